"""CPU oracle for the parallel-A* hot path.

TEST INFRASTRUCTURE ONLY: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg may import it, and only as the checker or the CPU baseline.
The product library never links or calls it.
"""
