"""CPU pins of the oracle's trainer control flow against tables written from
the reference's text (src/full-train.c, src/budgeted-train.c).

Every expected value below was worked out by hand from the cited lines, not
computed by a second implementation: the working-set rebuild
(src/full-train.c:540-643) on a hand-built vector covering each (class,
state, violating) branch, the found flags, the exact thresholds, -0.0 and
values one ulp either side of C; its fill rule, including nSC == W - nSW
(the strict < sends that case through rpermute) and the number of rand()
draws; the inner regrouping (:264-300) and the initial groups (:131-155)
row by row, including the order of the tests; the inner loop rule (:158,
C precedence); the outer stop and best tracking (:514-533); IRWLSpar's
loop rule (src/budgeted-train.c:399) and M decay (:479)."""
import ctypes as C

import numpy as np
import pytest

from oracle import pyoracle as po

L = po.lib()
_libc = C.CDLL(None)
_libc.rand.restype = C.c_int
_libc.srand.argtypes = [C.c_uint]

ULP_BELOW_2 = np.nextafter(2.0, 0.0)
ULP_ABOVE_2 = np.nextafter(2.0, 3.0)

# ---- working-set rebuild, C = 2, threshold 0.001 ------------------------------
# (y, beta, e, where it goes and why; "SW" = direct pick)
WS_ROWS = [
    (+1, 2.0, -0.5, "SW"),        # 0  at bound (beta y == C), e y = -0.5 < -0.001: first +1 bound violator (found12)
    (+1, 2.0, -0.3, "SC"),        # 1  the same, found12 already set
    (+1, 2.0, -0.001, "SIN"),     # 2  e y = -0.001 is not < -0.001
    (+1, 2.0, 0.5, "SIN"),        # 3  at bound, not violating
    (-1, -2.0, 0.5, "SW"),        # 4  beta y = 2 == C, e y = -0.5: first -1 bound violator (found02)
    (-1, -2.0, 0.0011, "SC"),     # 5  e y = -0.0011, found02 set
    (+1, 0.0, 0.5, "SW"),         # 6  beta == 0, e y = 0.5 > 0.001: first +1 zero violator (found10)
    (+1, 0.0, 0.001, "SIN"),      # 7  e y = 0.001 is not > 0.001
    (+1, 0.0, -0.7, "SIN"),       # 8  zero, not violating
    (+1, -0.0, 0.2, "SC"),        # 9  -0.0 == 0.0: zero state, violating, found10 set
    (-1, 0.0, -0.2, "SW"),        # 10 e y = 0.2: first -1 zero violator (found00)
    (-1, 0.0, -0.2, "SC"),        # 11 found00 set
    (+1, 0.5, 0.2, "SW"),         # 12 free (beta y = 0.5), |e y| = 0.2: first +1 free violator (found11)
    (+1, 0.5, -0.2, "SC"),        # 13 |e y| = 0.2, found11 set
    (+1, 0.5, 0.0005, "SC"),      # 14 free and not violating: still a candidate, never SIN
    (-1, -1.0, 0.3, "SW"),        # 15 free (beta y = 1), |e y| = 0.3: first -1 free violator (found01)
    (-1, -1.0, 0.0, "SC"),        # 16 free, not violating
    (+1, ULP_BELOW_2, 0.0, "SC"),  # 17 beta y one ulp below C: free, not violating
    (+1, ULP_ABOVE_2, 0.5, "SC"),  # 18 one ulp above C: free (not == C), violating, found11 set
    (-1, 2.0, 0.3, "SC"),         # 19 beta y = -2: neither C nor 0, so free; |e y| = 0.3, found01 set
]
SW_DIRECT = [0, 4, 6, 10, 12, 15]
SC_ORDER = [1, 5, 9, 11, 13, 14, 16, 17, 18, 19]
SIN_ORDER = [2, 3, 7, 8]


def _rebuild(W, seed=None):
    y = np.array([r[0] for r in WS_ROWS], dtype=np.float64)
    b = np.array([r[1] for r in WS_ROWS], dtype=np.float64)
    e = np.array([r[2] for r in WS_ROWS], dtype=np.float64)
    n = len(WS_ROWS)
    SW, SIN, SC = (C.c_int * W)(), (C.c_int * n)(), (C.c_int * n)()
    nsw, nsin = C.c_int(), C.c_int()
    found = (C.c_int * 6)()
    if seed is not None:
        po.srand(seed)
    dp = C.POINTER(C.c_double)
    ncand = L.orc_ws_rebuild(n, b.ctypes.data_as(dp), e.ctypes.data_as(dp), y.ctypes.data_as(dp), 2.0, W,
                             SW, C.byref(nsw), SIN, C.byref(nsin), SC, found)
    return ncand, list(SW[:nsw.value]), list(SIN[:nsin.value]), list(found)


def test_table_rows_match_their_labels():
    assert [i for i, r in enumerate(WS_ROWS) if r[3] == "SW"] == SW_DIRECT
    assert [i for i, r in enumerate(WS_ROWS) if r[3] == "SC"] == SC_ORDER
    assert [i for i, r in enumerate(WS_ROWS) if r[3] == "SIN"] == SIN_ORDER


def test_ws_rebuild_every_branch_candidates_fit():
    """W = 17: space 11 > 10 candidates, appended in index order, no draw."""
    _libc.srand(5)
    nxt = _libc.rand()
    ncand, sw, sin, found = _rebuild(17, seed=5)
    assert ncand == len(SC_ORDER)
    assert sw == SW_DIRECT + SC_ORDER
    assert sin == SIN_ORDER
    assert found == [1, 1, 1, 1, 1, 1]   # [y==-1: zero, free, bound], [y==+1: zero, free, bound]
    assert _libc.rand() == nxt           # rand() untouched


@pytest.mark.parametrize("W", [16, 12])
def test_ws_rebuild_permuted_fill(W):
    """W = 16: space == 10 candidates, which the strict < sends through
    rpermute (nSC - 1 = 9 rand() draws) even though all fit; W = 12: six
    go to the working set in permuted order, the other four to SIN after
    its own members."""
    seed = 11
    _libc.srand(seed)
    stream = [_libc.rand() for _ in range(12)]
    ncand, sw, sin, _ = _rebuild(W, seed=seed)
    assert _libc.rand() == stream[9]     # exactly 9 draws consumed
    po.srand(seed)
    perm = po.take_ints(L.orc_rpermute(10), 10)
    space = W - len(SW_DIRECT)
    assert sw[:len(SW_DIRECT)] == SW_DIRECT
    assert sw[len(SW_DIRECT):] == [SC_ORDER[p] for p in perm[:space]]
    assert sin == SIN_ORDER + [SC_ORDER[p] for p in perm[space:]]
    assert sorted(sw[len(SW_DIRECT):] + sin[len(SIN_ORDER):]) == SC_ORDER


def test_rpermute_draw_rule():
    """rpermute(n): k = n-1 .. 1, j = rand() % (k+1), swap a[j] and a[k]
    (src/full-train.c:65-77), checked for n = 4 from the libc stream."""
    _libc.srand(3)
    r = [_libc.rand() for _ in range(3)]
    a = [0, 1, 2, 3]
    for k, v in zip([3, 2, 1], r):
        j = v % (k + 1)
        a[j], a[k] = a[k], a[j]
    po.srand(3)
    assert list(po.take_ints(L.orc_rpermute(4), 4)) == a


# ---- inner regrouping, C = 1 (src/full-train.c:264-300) ----------------------------
# (group before, y, e, beta new) -> (a after, group after)
REGROUP = [
    ((1, +1, 0.5, 0.5), (2.0, 1)),                    # a = y C / e
    ((1, +1, 0.5, 0.99), (2.0, 3)),                   # y beta = 0.99 C: >= 0.99 C -> group 3
    ((1, +1, 0.5, np.nextafter(0.99, 0)), (2.0, 1)),  # one ulp below 0.99 C
    ((1, -1, -0.25, -1.01), (4.0, 3)),                # y beta = 1.01 C: <= 1.01 C -> group 3
    ((1, -1, -0.25, np.nextafter(-1.01, -2)), (4.0, 1)),  # one ulp above 1.01 C
    ((1, +1, -0.1, 0.995), (0.0, 2)),                 # e y < 0 first: group 2, so the 0.99 C test never applies
    ((1, +1, 5e-5, 0.2), (1e4, 1)),                   # 0 <= e y < 1e-4: a = C 1e4
    ((1, +1, 0.0, 0.2), (1e4, 1)),                    # e y = 0: not < 0, so C 1e4
    ((1, -1, -0.0, 0.2), (1e4, 1)),                   # e y = (-0.0)(-1) = 0.0
    ((1, +1, 1e-4, 0.2), (1.0 / 1e-4, 1)),            # e y = 1e-4 is not < 1.0/10000: a = y C / e
    ((2, +1, 0.3, 0.0), (1.0 / 0.3, 1)),              # group 2 with a != 0 -> 1
    ((2, +1, -0.3, 0.0), (0.0, 2)),                   # stays 2
    ((3, +1, 0.3, 0.2), (1.0 / 0.3, 3)),              # group 3 stays 3 while e y >= 0, whatever beta
    ((3, +1, -0.3, 0.2), (0.0, 2)),                   # e y < 0: 3 -> 2, and a = 0 keeps it there
    ((2, +1, 0.3, 0.995), (1.0 / 0.3, 1)),            # 2 -> 1 comes after the 0.99 C test: ends in 1, not 3
    ((2, -1, 5e-5, 0.0), (0.0, 2)),                   # e y = -5e-5 < 0
]


def test_regroup_table():
    n = len(REGROUP)
    g0 = np.array([r[0][0] for r in REGROUP], dtype=np.int32)
    y = np.array([r[0][1] for r in REGROUP], dtype=np.float64)
    e = np.array([r[0][2] for r in REGROUP], dtype=np.float64)
    b = np.array([r[0][3] for r in REGROUP], dtype=np.float64)
    a = np.full(n, -7.0)
    grp = g0.copy()
    dp, ip = C.POINTER(C.c_double), C.POINTER(C.c_int)
    L.orc_regroup(n, e.ctypes.data_as(dp), y.ctypes.data_as(dp), b.ctypes.data_as(dp), 1.0,
                  a.ctypes.data_as(dp), grp.ctypes.data_as(ip))
    for k, (_, (wa, wg)) in enumerate(REGROUP):
        assert (a[k], grp[k]) == (wa, wg), (k, REGROUP[k], a[k], grp[k])


# ---- initial groups, C = 1 (src/full-train.c:131-155) --------------------------------
# (y, e, beta) -> (a, group); no 1e-4 clip at this point
INITIAL = [
    ((+1, -0.5, 0.0), (0.0, 2)),          # e y < 0: a = 0, group 2
    ((+1, 0.5, 1.0), (2.0, 3)),           # beta == y C: group 3
    ((+1, 0.5, 0.99), (2.0, 1)),          # otherwise group 1
    ((-1, -0.5, -1.0), (2.0, 3)),         # beta == y C = -1
    ((+1, 0.0, 0.0), (np.inf, 1)),        # e y = 0: a = y C / 0 = inf, nonzero
    ((+1, -0.0, 0.0), (-np.inf, 1)),      # e y = -0.0 is not < 0: a = 1 / -0.0 = -inf
]


def test_initial_groups_table():
    n = len(INITIAL)
    y = np.array([r[0][0] for r in INITIAL], dtype=np.float64)
    e = np.array([r[0][1] for r in INITIAL], dtype=np.float64)
    b = np.array([r[0][2] for r in INITIAL], dtype=np.float64)
    a = np.zeros(n)
    grp = np.zeros(n, dtype=np.int32)
    dp, ip = C.POINTER(C.c_double), C.POINTER(C.c_int)
    with np.errstate(divide="ignore"):
        L.orc_initial_groups(n, e.ctypes.data_as(dp), y.ctypes.data_as(dp), b.ctypes.data_as(dp), 1.0,
                             a.ctypes.data_as(dp), grp.ctypes.data_as(ip))
    for k, (_, (wa, wg)) in enumerate(INITIAL):
        assert (a[k], grp[k]) == (wa, wg), (k, a[k], grp[k])


# ---- inner loop rule, C = 1 (src/full-train.c:158) ---------------------------------
# (iter, minbeta, maxbeta, since, dW, nW) -> continue?
INNER = [
    ((0, 0.0, 0.5, 9, 0.0, 1.0), 1),        # iter < 5 runs whatever the rest says
    ((4, 0.0, 0.5, 9, 0.0, 0.0), 1),
    ((5, 0.0, 0.5, 0, 1.0, 1.0), 0),        # no bound violated
    ((5, 0.0, 1.0, 0, 1.0, 1.0), 0),        # maxbeta == C is not > C
    ((5, -0.1, 0.5, 0, 1.0, 1.0), 1),       # minbeta < 0
    ((5, 0.0, 1.0000001, 0, 1.0, 1.0), 1),  # maxbeta > C
    ((5, -0.1, 0.5, 5, 1.0, 1.0), 0),       # since best = 5
    ((5, -0.1, 0.5, 4, 1.0, 1.0), 1),
    ((999, -0.1, 0.5, 0, 1.0, 1.0), 1),
    ((1000, -0.1, 0.5, 0, 1.0, 1.0), 0),    # 1000 iterations
    ((5, -0.1, 0.5, 0, 1e-6, 1.0), 0),      # ratio 1e-6 is not > 1e-6
    ((5, -0.1, 0.5, 0, 1.000001e-6, 1.0), 1),
    ((5, -0.1, 0.5, 0, 0.0, 0.0), 0),       # 0/0 = NaN: the comparison is false
    ((5, -0.1, 0.5, 0, 1.0, 0.0), 1),       # 1/0 = inf > 1e-6
]


def test_inner_loop_rule_table():
    with np.errstate(divide="ignore", invalid="ignore"):
        for args, want in INNER:
            assert L.orc_inner_continue(args[0], args[1], args[2], 1.0, args[3], args[4], args[5]) == want, args


# ---- outer stop and best tracking (src/full-train.c:514-533) -------------------------
def test_outer_tracking_sequence():
    """done = dW/nW < Eta; a strictly smaller ratio resets the since-best
    counter, anything else (equal, larger, NaN, inf) increments it; the
    first outer iteration has nW = |beta_0|^2 = 0, so its ratio is inf."""
    best, since = C.c_double(1e20), C.c_int(0)
    seq = [  # (dW, nW) -> (done, best after, since after)
        ((1.0, 0.0), (0, 1e20, 1)),         # inf ratio: not below 1e20
        ((4.0, 2.0), (0, 2.0, 0)),
        ((4.0, 2.0), (0, 2.0, 1)),          # equal is not better
        ((8.0, 2.0), (0, 2.0, 2)),
        ((0.0, 0.0), (0, 2.0, 3)),          # NaN
        ((1.0, 4.0), (0, 0.25, 0)),
        ((1e-3, 1.0), (0, 1e-3, 0)),        # 1e-3 is not < Eta = 1e-3
        ((9e-4, 1.0), (1, 9e-4, 0)),        # done
    ]
    with np.errstate(divide="ignore", invalid="ignore"):
        for (dW, nW), (wd, wb, ws) in seq:
            done = L.orc_outer_track(dW, nW, 1e-3, C.byref(best), C.byref(since))
            assert (done, best.value, since.value) == (wd, wb, ws), (dW, nW)


def test_outer_runs_until_300_iterations_without_improvement():
    """The outer loop condition (endNorm == 0 && SinceBest < 300,
    src/full-train.c:431): 300 ratios no better than the best end it."""
    best, since = C.c_double(1e20), C.c_int(0)
    assert L.orc_outer_track(1.0, 1.0, 1e-3, C.byref(best), C.byref(since)) == 0
    for k in range(300):
        assert since.value < 300
        L.orc_outer_track(2.0, 1.0, 1e-3, C.byref(best), C.byref(since))
    assert since.value == 300


# ---- IRWLSpar (src/budgeted-train.c:399, 479-489) -----------------------------------
def test_irwlspar_loop_rule_table():
    rows = [((0, 1.0, 1.0, 0), 1), ((499, 1.0, 1.0, 4), 1), ((500, 1.0, 1.0, 0), 0), ((3, 1e-6, 1.0, 0), 0),
            ((3, 2e-6, 1.0, 0), 1), ((3, 1.0, 1.0, 5), 0), ((0, 1e9, 1.0, 0), 1)]
    for (it, dW, nW, since), want in rows:
        assert L.orc_par_continue(it, dW, nW, since) == want, (it, dW, nW, since)


def test_irwlspar_m_decay_and_best_tracking():
    """M /= 10 only when iter > 10 and the ratio exceeds 100 x the previous
    one (strictly); the previous ratio always becomes this one; a strictly
    smaller ratio is the new best."""
    M, old, best, since = C.c_double(1e4), C.c_double(0.25), C.c_double(1e9), C.c_int(0)
    steps = [  # (iter, ratio) -> (M, old, best, since, improved)
        ((10, 30.0), (1e4, 30.0, 30.0, 0, 1)),      # 30 > 100 x 0.25 but iter 10 is not > 10
        ((11, 3000.0), (1e4, 3000.0, 30.0, 1, 0)),  # 3000 == 100 x 30: not strictly larger
        ((12, 3e5 + 1e-9), (1e3, 3e5 + 1e-9, 30.0, 2, 0)),  # > 100 x 3000: M /= 10
        ((13, 10.0), (1e3, 10.0, 10.0, 0, 1)),
        ((14, 10.0), (1e3, 10.0, 10.0, 1, 0)),      # equal: not a new best
        ((15, 1001.0), (1e2, 1001.0, 10.0, 2, 0)),  # > 1000: decays again
    ]
    for (it, ratio), (wM, wo, wb, ws, wi) in steps:
        imp = L.orc_par_track(it, ratio, C.byref(M), C.byref(old), C.byref(best), C.byref(since))
        assert (M.value, old.value, best.value, since.value, imp) == (wM, wo, wb, ws, wi), (it, ratio)
