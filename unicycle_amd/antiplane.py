"""Drop-in driver of the reference's 2-D antiplane rate-and-state program
(fortran/2d/antiplane/src/ratestate.f90).

    python -m unicycle_amd.antiplane [-e EPS] [-m MAXSTEP] [-i MAXITER] [--dry-run] deck.in

Reads the 2-D deck (unicycle_amd.deck.parse_antiplane: shear modulus,
interval, vertical segments with friction, observed patches, observation
points), builds K and the displacement kernels on the GPU and runs the main
loop (ratestate.f90:254-276) on the device with the loop of
unicycle_amd.ratestate: time.txt, patch-%08d.dat (4 state entries and their
derivatives) and opts-<name>.dat (one displacement component).  The NetCDF
exports of the reference (--export-netcdf, --export-greens) are not written.
"""
from __future__ import annotations

import sys

from .ratestate import main as _main


def main(argv=None) -> int:
    return _main(argv, antiplane=True)


if __name__ == "__main__":
    sys.exit(main())
