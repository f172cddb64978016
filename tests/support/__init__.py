"""Shared helpers for the test suite: FCHK and Cube readers written for the
tests, and the ctypes binding of the CPU oracle (oracle/liboracle.so)."""
