"""CPU oracle of the genotype-probe path — test infrastructure only (see oracle.py)."""
