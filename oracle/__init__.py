"""CPU oracle of the ir_interpolate_at path -- TEST INFRASTRUCTURE ONLY.

Imported only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg, as the checker.  The product (brille_amd) never imports it.
"""
