"""TEST INFRASTRUCTURE ONLY: CPU oracles for the rasterize-then-splat path.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import
this package.  Nothing in diffren_amd/ imports it.
"""
