"""Oracle (test infrastructure only): CPU restatement of the reference path.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this package. The product package never does.
"""
