"""CPU oracle (test infrastructure only).

Plain-C FP64 restatements of the reference solvers, loaded through ctypes.
Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py use
this package; the product library never imports it.
"""
