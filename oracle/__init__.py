"""CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import
anything from this package, and only as the test reference or the timed CPU baseline.
The product (mutual-polarization-slab_amd/) never imports it.
"""
