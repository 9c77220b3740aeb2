"""CPU oracle of linon's compute pass — test infrastructure only (see oracle.py)."""
