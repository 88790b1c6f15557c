"""Drop-in driver of the reference's 0-D rate-and-state program
(fortran/0d/src/ratestate.f90: a circular patch of radius r on a spring).

    python -m unicycle_amd.springslider [-e EPS] [-m MAXSTEP] [-i MAXITER] [--dry-run] deck.in

The 0-D odefun (ratestate.f90:230-277) is the 2-D antiplane one with one
patch and the circular-crack stiffness as K: t = -7 pi / 16 mu / r (V - Vpl)
(ratestate.f90:255), so the run is the device-resident main loop of that
model (ucy_set_antiplane with one element, ucy_set_kernel with K = [[k]]).
It writes what the program writes (ratestate.f90:100-151):

  <wdir>/time.txt    the header, (0, dt_next), then per iteration the time at
                     its start and the previous step (ES19.12E2);
  <wdir>/patch.dat   the program's header line, then per iteration the time
                     (ES19.12E2) and, each after one blank, the 4 state
                     entries and their derivatives (ES20.12E3);
  stdout             every 50 iterations: iteration, time, step, V.
"""
from __future__ import annotations

import argparse
import math
import os
import sys

import numpy as np

from . import deck as deckmod
from .engine import Engine

PI = 3.1415926535897932384626433832795028  # macros.f90:12 (0-D)


def stiffness(d: deckmod.PointDeck) -> float:
    """-7 pi / 16 mu / r in the reference's order of operations (:255)."""
    return -7.0 * PI / 16.0 * d.mu / d.radius


def setup(d: deckmod.PointDeck, device: int = 0) -> Engine:
    eng = Engine(device)
    eng.set_antiplane(d.mu, np.zeros((1, 3)), [d.radius], [math.pi / 2])
    eng.set_friction([d.Vpl], [0.0], [d.mu0], [d.sig], [d.a], [d.b], [d.L], [d.Vo], [d.damping])
    eng.set_kernel(np.array([[stiffness(d)]]))
    return eng


def run(d: deckmod.PointDeck, epsilon: float = 1.0e-6,
        max_step: float = 1.7976931348623158e308, max_iterations: int = 1000000,
        device: int = 0, chunk: int = 20000, out=sys.stdout, write_files: bool = True):
    """Run the deck; returns (iterations, final time, final state)."""
    fmt = deckmod.es_field
    eng = setup(d, device)
    y = eng.initStateVector()   # ratestate.f90:195-216 (0-D)
    time, dt_next, dt_done = 0.0, 1.0e-3, 0.0
    idx = np.arange(4, dtype=np.int64)
    if write_files:
        os.makedirs(d.wdir, exist_ok=True)
        ftime = open(os.path.join(d.wdir, "time.txt"), "w")
        fstate = open(os.path.join(d.wdir, "patch.dat"), "w")
        ftime.write("#               time                 dt\n")
        ftime.write(fmt(0.0, 19, 12, 2) + fmt(dt_next, 19, 12, 2) + "\n")
        fstate.write("# time, slip, stress, log10(state), log10(V), slip rate, stress rate, "
                     "stress rate / stress, a / V\n")
    out.write("#       n               time                 dt          v\n")
    out.write(f"{0:09d}" + fmt(time, 19, 12, 2) + fmt(dt_next, 19, 12, 2) + "\n")
    it = 0
    try:
        while it < max_iterations and time < d.interval:
            n = min(chunk, max_iterations - it)
            res = eng.integrate(y, d.interval, dt0=dt_next, max_steps=n, epsilon=epsilon,
                                max_step=max_step, log_cap=n, observe=idx, t0=time)
            for k in range(res.steps):
                i = it + k + 1
                t_start = res.t_log[k - 1] if k else time
                prev = res.dt_log[k - 1] if k else dt_done
                if write_files:
                    # export, ratestate.f90:155-184: (ES19.12E2, 4 x (X, ES20.12E3, X,
                    # ES20.12E3)) over y(1:4) then dydt(1:4)
                    vals = list(res.obs_y[k]) + list(res.obs_dydt[k])
                    fstate.write(fmt(t_start, 19, 12, 2)
                                 + "".join(" " + fmt(v, 20, 12, 3) for v in vals) + "\n")
                    ftime.write(fmt(t_start, 19, 12, 2) + fmt(prev, 19, 12, 2) + "\n")
                if i % 50 == 0:
                    out.write(f"{i:09d}" + fmt(t_start, 19, 12, 2) + fmt(prev, 19, 12, 2)
                              + fmt(res.vmax_log[k], 11, 4, 2) + "\n")
            if res.steps:
                dt_done = res.dt_log[res.steps - 1]
            it += res.steps
            time, y, dt_next = res.time, res.y, res.dt_next
            out.flush()
            if res.steps < n and time < d.interval:
                break
    finally:
        if write_files:
            ftime.close()
            fstate.close()
        eng.close()
    out.write(f"{it if time >= d.interval else it + 1:09d} time steps.\n")
    return it, time, y


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="unicycle_amd.springslider",
                                 description="quasi-dynamic slip evolution for a point source "
                                             "(GPU)")
    ap.add_argument("-e", "--epsilon", type=float, default=1.0e-6)
    ap.add_argument("-m", "--maximum-step", type=float, default=1.7976931348623158e308)
    ap.add_argument("-i", "--maximum-iterations", type=int, default=1000000)
    ap.add_argument("--dry-run", action="store_true")
    ap.add_argument("--version", action="store_true")
    ap.add_argument("deck", nargs="?", help="input deck (standard input if omitted)")
    a = ap.parse_args(argv)
    if a.version:
        print("unicycle_amd springslider (gfx950)")
        return 0
    text = open(a.deck).read() if a.deck else sys.stdin.read()
    try:
        d = deckmod.parse_point(text)
    except deckmod.DeckError as exc:
        print(f"input error: {exc}", file=sys.stderr)
        return 1
    print("# RATESTATE")
    print("# quasi-dynamic slip evolution for a point source")
    print(f"# numerical accuracy: {deckmod.es_field(a.epsilon, 11, 4, 2)}")
    print(f"# maximum iterations: {a.maximum_iterations:11d}")
    print(f"# output directory\n  {d.wdir}")
    if a.dry_run:
        print("dry run: abort calculation")
        return 0
    run(d, epsilon=a.epsilon, max_step=a.maximum_step, max_iterations=a.maximum_iterations)
    return 0


if __name__ == "__main__":
    sys.exit(main())
