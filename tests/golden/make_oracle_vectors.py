#!/usr/bin/env python3
"""Write oracle golden vectors for cases too slow to run the oracle inside
the test suite.

diffuse_dipole (512x1024, one 8760 h step, 36,205 RKG2 stages in 7 PTL
cycles) takes about four minutes in the CPU oracle; its history rows are
stored as tests/golden/oracle_diffuse_dipole.npz so the GPU test can compare
against the oracle bit for bit.  Run from the repository root after
building the oracle:  python tests/golden/make_oracle_vectors.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

from cases import case_cfg  # noqa: E402
from oracle.oracle import run_case  # noqa: E402


def main():
    sol, num, o = run_case(case_cfg("diffuse_dipole"))
    np.savez_compressed(os.path.join(HERE, "oracle_diffuse_dipole.npz"), sol=sol, num=num)
    print(num[0])


if __name__ == "__main__":
    main()
