"""FP64 CPU oracle of the DEM time step -- TEST INFRASTRUCTURE ONLY.

Loaded only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg, as the checker and CPU baseline; the product never imports it.
"""
