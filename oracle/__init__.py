"""CPU oracle package (test infrastructure only; see rk_oracle.h)."""
