"""CPU oracle of the reference's steady-state diffusion solve.

TEST INFRASTRUCTURE: only tests/, __graft_entry__.smoke() and the cpu_baseline
leg of bench.py import this package, as the checker of the HIP path.  The
product (taufactor_amd) never imports it.
"""
