"""TEST INFRASTRUCTURE ONLY: CPU oracle for the all-pairs NCC path.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use
this package, as the checker or the reported CPU baseline.  See
oracle/xcorr_oracle.c for what each function restates (reference file:line).
"""
