"""Oracle package: CPU FP64 restatement of the reference hot path.

Test infrastructure only (tests/, __graft_entry__.smoke(), bench.py's
cpu_baseline leg). The product path must never import it.
"""
