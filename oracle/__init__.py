"""CPU oracle of the mechanical-equilibration path (test infrastructure only)."""
