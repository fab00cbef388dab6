"""CPU oracle for XBFS - test infrastructure only (see oracle.py)."""
