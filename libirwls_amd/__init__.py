"""MI355X-native IRWLS SVM training hot path (LIBIRWLS interface)."""
