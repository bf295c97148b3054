"""FP64 CPU oracle of the pressure solve — test infrastructure only.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this package, and only as the checker. See pcg_pfmg.c for the
algorithm and the NGA2 lines it restates.
"""
