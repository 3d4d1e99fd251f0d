"""Writes the regression fixtures of tests/test_golden.py from the CPU oracle.

    python tests/golden/make_golden.py

Each fixture holds seeded inputs, the parameters, the fields after a number
of steps and the evaluated free energy per area and mean densities.  The
reference ships no test vectors; these pin the oracle against later changes
and let the GPU tests compare with stored numbers.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, os.path.dirname(HERE))

from oracle_lib import Oracle  # noqa: E402
from pfc_amd import LATTICE_A, graphene_params, hetero_params  # noqa: E402
from pfc_amd.inputs import crystal, hetero_patchwork  # noqa: E402

CASES = {
    # name: (params factory, fields factory, Nx, Ny, dx, dy, steps)
    "graphene_64x64_nc": (lambda: graphene_params(c=0.0),
                          lambda: crystal(64, 64, LATTICE_A / 8, LATTICE_A / 8, 0.2, -0.2,
                                          LATTICE_A, sym=3, theta=0.3, noise=0.01, seed=1)[None],
                          64, 64, LATTICE_A / 8, LATTICE_A / 8, 30),
    "hetero_64x64_c": (lambda: hetero_params(c=1.0),
                       lambda: hetero_patchwork(64, 64, 0.7, 0.7, domain=24, seed=3),
                       64, 64, 0.7, 0.7, 30),
}


def compute(name):
    pf, ff, Nx, Ny, dx, dy, steps = CASES[name]
    p = pf()
    f0 = ff()
    o = Oracle(p, Nx, Ny, dx, dy)
    o.set_fields(f0)
    assert o.steps(0, steps) == 0
    fave, nave = o.evaluate()
    return dict(fields_in=f0, fields_out=o.get_fields(), fave=np.array(fave), nave=nave,
                Nx=Nx, Ny=Ny, dx=dx, dy=dy, steps=steps)


def main():
    for name in CASES:
        d = compute(name)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **d)
        print("wrote", name, "fave", float(d["fave"]))


if __name__ == "__main__":
    main()
