"""CPU oracle for the DRR path -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use
this package, and only to check results against; the product (deepdrr_amd)
never imports it.
"""
