"""MI355X-native Bloch simulation engine with MRiLab's scan interface."""
