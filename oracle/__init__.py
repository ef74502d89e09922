"""FP64 CPU oracle of the split-step path (test infrastructure only)."""
