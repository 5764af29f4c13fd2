"""CPU oracle of the D3Q19 step (NumPy: lbm_oracle.py; C: lbm_oracle.c via c_oracle.py).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and bench.py's CPU
baseline, as the checker.  The product (tack-lbm_amd/) never imports it.
"""
