"""CPU checker for the grain-scoring path (test infrastructure only)."""
