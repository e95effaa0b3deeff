"""CPU oracle for the DRR path -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use
this package, as the checker; the product (deepdrr_amd) never imports it.
"""
