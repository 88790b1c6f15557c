"""Drop-in driver of the reference's 3-D viscoelastic cycle program.

    python -m unicycle_amd.viscouscycles [-e EPS] [-m MAXSTEP] [-i MAXITER] [--dry-run] deck.in

Reads the input deck of viscouscycles.f90 (unicycle_amd.deck.parse_viscous:
faults, cuboid strain volumes with nonlinear Maxwell and thermal properties,
observed patches and volumes, observation points, perturbations) and runs
the main loop of viscouscycles.f90:413-434 on the device with the loop of
unicycle_amd.ratestate, which writes time.txt, patch-%08d.dat,
volume-%08d.dat, opts-<name>.dat (d, df, dl), patch-rectangular.vtp and
volume.vtp in the reference's formats.
"""
from __future__ import annotations

import sys

from .ratestate import main as _main
from .ratestate import run  # noqa: F401  (the same loop runs a ViscousDeck)


def main(argv=None) -> int:
    return _main(argv, viscous=True)


if __name__ == "__main__":
    sys.exit(main())
