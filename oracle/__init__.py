"""The CPU oracle for the flock step (test infrastructure only; see flockref.h)."""
