"""Writes tests/golden/bp5_example.npz: the reference's BP5-QD example model
(examples/BP5-QD: parameter.txt, bp5t.msh, bp5tparam.dat) as read by
pyquake3d_amd.reference_io -- mesh, friction, stresses, loading and initial
slip rate per element -- so that GPU tests and bench.py can run that model
where the reference tree is absent. Run in a container that has the
reference at /root/reference:

    python tests/golden/make_bp5_fixture.py
"""
import os
import sys

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, REPO)

from pyquake3d_amd.problem import save_problem  # noqa: E402
from pyquake3d_amd.reference_io import problem_from_parameters  # noqa: E402

SRC = "/root/reference/examples/BP5-QD/parameter.txt"


def main():
    p, hs, para = problem_from_parameters(SRC)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "bp5_example.npz")
    p.name = "BP5-QD"
    save_problem(out, p, halfspace=hs)
    print(out, os.path.getsize(out), "bytes,", p.n, "elements")


if __name__ == "__main__":
    main()
