"""CPU oracle for the pyloki search hot path -- test infrastructure only (see oracle.py)."""
