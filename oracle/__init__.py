"""The oracle: a float64 CPU restatement of the reference's imaging path.

Test infrastructure only -- imported by tests/, __graft_entry__.smoke() and
the cpu_baseline leg of bench.py, never by the lithobat_amd package.
"""
