"""Regenerates tests/golden/oracle_dq_single12.npz from the CPU oracle:
dq of the 12^3 single-block mixing layer after initialisation, and the
primitives and dt after three SSP-RK3 steps.  The reference ships no golden
vectors for this path, so this fixture guards the oracle against drift; the
oracle itself is pinned by tests/test_oracle_kernels.py and
tests/test_oracle_rhs.py."""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import test_oracle_rhs  # noqa: E402

if __name__ == "__main__":
    test_oracle_rhs.make_golden()
    print("wrote", test_oracle_rhs.GOLDEN)
