"""CPU oracle of gala's hot path -- TEST INFRASTRUCTURE ONLY.

Imported only by tests/, __graft_entry__.smoke() (as the checker) and
bench.py's cpu_baseline leg.  The product never imports it.
"""
