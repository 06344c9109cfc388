"""MI355X-native density-fitted J/K path."""
