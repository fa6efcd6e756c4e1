"""CPU oracle for the fixed-dt trap-kinetics Monte Carlo (test infrastructure).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
package, and only as the checker.  The product (mcluminescence_amd) never
imports it.
"""
