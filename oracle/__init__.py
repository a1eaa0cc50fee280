"""CPU oracle for the layer optical-depth path (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import
this package, and only to check results.  The product (clblm_amd) never uses it.
"""
