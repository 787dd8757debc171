"""CPU oracle of the reference path (test infrastructure only)."""
