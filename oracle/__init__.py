"""CPU oracle of the Wilson Dslash -- test infrastructure only (see oracle.py)."""
