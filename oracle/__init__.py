"""Oracle: CPU restatement of the reference parse path (test infrastructure only)."""
