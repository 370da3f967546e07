"""CPU oracle for the fCWT hot path -- test infrastructure only (see oracle.py)."""
