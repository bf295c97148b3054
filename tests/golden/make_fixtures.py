"""Writes tests/golden/fixtures.npz: right-hand sides and reference solutions
of three small problems, computed independently of the oracle.

  periodic16  16^3 periodic, closed-form FFT solution (the fourier3d method,
              NGA2 src/solver/fourier3d_class.f90:1103-1196)
  channel16   16^3 stretched channel with walls, scipy sparse direct solve
  spheres24   24^3 periodic, 6 spheres, density ratio 1000, sparse direct solve

Solutions are mean-free (all three systems are singular). The reference's own
tests hold no golden vectors for this path, so these fixtures come from the
mathematics (closed form, direct factorisation), not from a reference run.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, os.path.dirname(HERE))

from conftest import assemble_csr, closed_form_periodic, direct_solve_singular  # noqa: E402
from nga2_amd import inputs  # noqa: E402
from oracle.orc import interior  # noqa: E402


def main():
    out = {}
    P = inputs.periodic_gaussian(n=16)
    b = interior(P.rhs, P.cfg)
    out["periodic16_rhs"] = b
    out["periodic16_sol"] = closed_form_periodic(b, 16, 1.0 / 16)
    P = inputs.channel_spectral(n=16)
    b = interior(P.rhs, P.cfg)
    out["channel16_rhs"] = b
    out["channel16_sol"] = direct_solve_singular(assemble_csr(P.cfg, P.opr), b)
    P = inputs.spheres_two_phase(n=24, nsph=6, rmin=0.1, rmax=0.25)
    b = interior(P.rhs, P.cfg)
    out["spheres24_rhs"] = b
    out["spheres24_sol"] = direct_solve_singular(assemble_csr(P.cfg, P.opr), b)
    np.savez_compressed(os.path.join(HERE, "fixtures.npz"), **out)
    for k, v in out.items():
        print(k, v.shape, float(np.abs(v).max()))


if __name__ == "__main__":
    main()
