"""Writes the reference's example models as read by pyquake3d_amd.reference_io
(mesh, friction, stresses, loading and initial slip rate per element), so
that GPU tests and bench.py can run them where the reference tree is absent:

  bp5_example.npz   examples/BP5-QD (parameter.txt, bp5t.msh, bp5tparam.dat)
  lab_example.npz   examples/Lab-model (parameter.txt, lab.msh; Initlab)
  eafz_example.npz  examples/EAFZ-model (parameter.txt, turkey.msh)

(examples/HF-model holds only the script that generates its mesh and element
file, not the files themselves.) Run in a container that has the reference
at /root/reference:

    python tests/golden/make_model_fixtures.py
"""
import os
import sys

REPO = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, REPO)

from pyquake3d_amd.problem import save_problem  # noqa: E402
from pyquake3d_amd.reference_io import problem_from_parameters  # noqa: E402

SRC = "/root/reference/examples"
MODELS = (("BP5-QD", "bp5_example.npz"), ("Lab-model", "lab_example.npz"),
          ("EAFZ-model", "eafz_example.npz"))


def main():
    here = os.path.dirname(os.path.abspath(__file__))
    for d, name in MODELS:
        p, hs, para = problem_from_parameters(os.path.join(SRC, d, "parameter.txt"))
        p.name = d
        out = os.path.join(here, name)
        save_problem(out, p, halfspace=hs)
        print(out, os.path.getsize(out), "bytes,", p.n, "elements")


if __name__ == "__main__":
    main()
