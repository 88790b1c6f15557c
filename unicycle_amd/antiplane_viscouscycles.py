"""Drop-in driver of the reference's 2-D antiplane viscoelastic cycle program
(fortran/2d/antiplane/src/viscouscycles.f90).

    python -m unicycle_amd.antiplane_viscouscycles [-e EPS] [-m MAXSTEP] [-i MAXITER] [--dry-run] deck.in

Reads the deck (unicycle_amd.deck.parse_antiplane_viscous: the antiplane
ratestate deck with vertical strain volumes, their Maxwell rows (Gm,
gammadot0, n, Q, R), thermal rows and observed volumes) and runs the main
loop on the device with the loop of unicycle_amd.ratestate: time.txt,
patch-%08d.dat, volume-%08d.dat (4 state entries and their derivatives),
opts-<name>.dat (d, df, dl) and eMax in the progress line.  The NetCDF
exports of the reference are not written.
"""
from __future__ import annotations

import sys

from .ratestate import main as _main


def main(argv=None) -> int:
    return _main(argv, viscous=True, antiplane=True)


if __name__ == "__main__":
    sys.exit(main())
