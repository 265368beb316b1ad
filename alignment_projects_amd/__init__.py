"""MI355X-native FFT tile-pair cross-correlation (billkarsh/Alignment_Projects' correlation path)."""
