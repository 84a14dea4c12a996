"""CPU oracle of the DEM step (test infrastructure only; see gdem_oracle.h)."""
