"""Drop-in driver of the reference's 2-D plane-strain viscoelastic cycle
program (fortran/2d/planestrain/src/viscouscycles.f90).

    python -m unicycle_amd.planestrain_viscouscycles [-e EPS] [-m MAXSTEP] [-i MAXITER] [--dry-run] deck.in

Reads the deck (unicycle_amd.deck.parse_planestrain_viscous: the plane-strain
ratestate deck with strain volumes, their nonlinear Maxwell rows with the
initial stress s0, thermal rows and observed volumes) and runs the main loop
(viscouscycles.f90:347-378) on the device with the loop of
unicycle_amd.ratestate: time.txt, patch-%08d.dat, volume-%08d.dat (6 state
entries and their derivatives), opts-<name>.dat (d, df, dl: east and down
each) and eMax in the progress line.  The NetCDF exports of the reference
are not written.
"""
from __future__ import annotations

import sys

from .ratestate import main as _main


def main(argv=None) -> int:
    return _main(argv, viscous=True, planestrain=True)


if __name__ == "__main__":
    sys.exit(main())
