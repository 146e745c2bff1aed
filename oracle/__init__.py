"""FP64 CPU oracle of the VBD path -- test infrastructure only (see oracle.py)."""
