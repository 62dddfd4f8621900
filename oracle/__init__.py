"""Test infrastructure: CPU oracle for the Bluebild hot path (numpy + a C restatement).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this package.
"""
