"""TEST INFRASTRUCTURE ONLY: the CPU oracle (see oracle/oracle.py).  Only
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import it."""
