"""C restatement of oracle arithmetic (test infrastructure only)."""
