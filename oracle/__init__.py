"""CPU parity oracle (test infrastructure only; see hb_oracle.c)."""
