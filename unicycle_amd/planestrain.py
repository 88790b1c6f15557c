"""Drop-in driver of the reference's 2-D plane-strain rate-and-state program
(fortran/2d/planestrain/src/ratestate.f90).

    python -m unicycle_amd.planestrain [-e EPS] [-m MAXSTEP] [-i MAXITER] [--dry-run] deck.in

Reads the plane-strain deck (unicycle_amd.deck.parse_planestrain: elastic
moduli mu and lambda, interval, dipping segments with friction rows that
begin with tau0, observed patches, observation points), builds K and the
displacement kernels on the GPU and runs the main loop (ratestate.f90:
285-305) on the device with the loop of unicycle_amd.ratestate: time.txt,
patch-%08d.dat (5 state entries and their derivatives, each after a blank)
and opts-<name>.dat (east and down displacement).  The NetCDF exports of the
reference are not written.
"""
from __future__ import annotations

import sys

from .ratestate import main as _main


def main(argv=None) -> int:
    return _main(argv, planestrain=True)


if __name__ == "__main__":
    sys.exit(main())
