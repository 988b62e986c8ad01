"""ORACLE — test infrastructure only.

CPU restatements of the reference path's arithmetic, used as the checker by tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg. Nothing in rascil_amd imports
or calls anything here.
"""
