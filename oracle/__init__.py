"""CPU oracle for the particle pipeline (test infrastructure, not product)."""
