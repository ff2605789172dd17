"""Oracle for the OSHUN distribution-function update (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import
this package, and only as the checker or the timed CPU baseline.  The product
(oshun_amd and its HIP library) never imports it.
"""
