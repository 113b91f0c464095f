"""ORACLE -- test infrastructure only.

CPU restatements of the reference's solve-and-sensitivity path that the tests,
`__graft_entry__.smoke()` and bench.py's `cpu_baseline` leg use as the checker.
Nothing in the product (`spacetime-topology-optimization_amd/`) imports this package.
"""
