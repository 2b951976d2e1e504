"""CPU oracle of the Monte-Carlo PDA hot path (test infrastructure only)."""
