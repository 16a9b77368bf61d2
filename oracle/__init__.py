"""ORACLE -- test infrastructure only: the FP64 CPU restatement of the
reference step (see oracle/orc.h).  Only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg import this package."""
