"""CPU oracle of the tile processor hot path (test infrastructure only)."""
