"""CPU oracle for the DiaQ GPU hot path -- test infrastructure only (the checker and the CPU baseline instrument).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this package; the product (diaq_amd)
never does.
"""
