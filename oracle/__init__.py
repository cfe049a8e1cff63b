"""Checker package: ctypes bindings for the CPU restatement (rz_oracle.c) and
for the reference's own sources compiled by oracle/Makefile (oracle/_ref).

TEST INFRASTRUCTURE ONLY: imported by tests/, the package's smoke() and the
cpu_baseline leg of bench.py as the checker; never by the product path.
"""
