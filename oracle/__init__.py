"""CPU oracle of the exaStitch hot path (test infrastructure only)."""
