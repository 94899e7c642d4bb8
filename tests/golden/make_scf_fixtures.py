"""Write tests/golden/scf_fortran.npz from the SCF golden vectors that gala's
own tests hold (tests/potential/scf/data/: positions.dat.gz, the five
*.coeff coefficient files in Hernquist & Ostriker 1992 normalisation, and
the Fortran SCF code's *-accp.dat.gz outputs: acceleration x, y, z and
potential, 7 significant figures).  Every 4th of the 32,768 grid points is
kept (8,192 points).  Run in a container that has the reference tree; the
fixture is data only.
"""
import gzip
import os

import numpy as np

SRC = "/root/reference/tests/potential/scf/data"
NAMES = ["simple-hernquist", "multi-hernquist", "simple-nonsph", "random", "wang-zhao"]
STRIDE = 4


def main():
    out = {}
    with gzip.open(os.path.join(SRC, "positions.dat.gz"), "rt") as f:
        xyz = np.loadtxt(f, skiprows=1)
    out["xyz"] = xyz[::STRIDE]
    for name in NAMES:
        coeff = np.atleast_2d(np.loadtxt(os.path.join(SRC, f"{name}.coeff"), skiprows=1))
        with gzip.open(os.path.join(SRC, f"{name}-accp.dat.gz"), "rt") as f:
            accp = np.loadtxt(f)
        out[f"{name}/coeff"] = coeff
        out[f"{name}/accp"] = accp[::STRIDE]
    here = os.path.dirname(os.path.abspath(__file__))
    np.savez_compressed(os.path.join(here, "scf_fortran.npz"), **out)


if __name__ == "__main__":
    main()
