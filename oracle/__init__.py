"""CPU oracle of the qubatron traversal (test infrastructure only)."""
