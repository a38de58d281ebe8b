"""MI355X-native close-coupling S-matrix solver (multiscat drop-in)."""
