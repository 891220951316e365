"""ORACLE -- test infrastructure only (tests/, __graft_entry__.smoke(), bench.py cpu_baseline).

CPU checkers of the HIP path: the FP64 restatement of the reference time step
(fvm_oracle.cpp) and, where the reference checkout exists, the reference's own
geometry code compiled into _ref/ (build_ref.sh). The product never imports this.
"""
