"""CPU oracle of the reference's steady-state diffusion solve.

TEST INFRASTRUCTURE: only tests/, __graft_entry__.smoke() and the cpu_baseline
leg of bench.py import this package, to compare the HIP path's results against
it.  The product (taufactor_amd) never imports it.
"""
