"""CPU oracle of the render path (test infrastructure, see pbr_oracle.c)."""
