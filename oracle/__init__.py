"""ORACLE — test infrastructure only.

A CPU restatement of the FEDEM Part Reducer hot path (SAM numbering, Femlib
element matrices, assembly, Guyan/Craig-Bampton reduction, gravity), used by
tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py as the
checker.  The product library never imports, links or calls anything here.
"""
