"""CPU oracle for the AXHEL NTT / element-wise path.

Test infrastructure only: tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg load it as the checker.  The product package (axhel_amd)
never imports it.
"""
