"""Regenerate the committed golden vectors of the reference example data set.

The expected outputs come from the FP64 oracle (oracle/mra_oracle.cpp), which
tests/test_oracle.py pins against the dense-covariance oracle; the inputs are
the reference's own example files copied under tests/golden/example/.
Run from the repository root:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle  # noqa: E402
from tests import cases  # noqa: E402


def main():
    tree, cov, prm = cases.example_case()
    D, mean, var = oracle.evaluate(tree, cov, prm)
    lev = oracle.evaluate_levels(tree, cov, prm)
    out = os.path.join(ROOT, "tests", "golden", "example_golden.npz")
    np.savez(out, D=D, mean=mean, var=var, levels=lev, pred_index=tree["pred_index"])
    print("wrote", out, "D =", D)


if __name__ == "__main__":
    main()
