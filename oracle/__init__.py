"""CPU oracle of the D2Q9 core -- test infrastructure only (see oracle.py)."""
