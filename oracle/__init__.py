"""CPU oracle of the PM loop (test infrastructure; see oracle/hgpm_oracle.c)."""
