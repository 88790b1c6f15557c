"""Drop-in driver of the reference's 3-D rate-and-state program.

    python -m unicycle_amd.ratestate [-e EPS] [-m MAXSTEP] [-i MAXITER] [--dry-run] deck.in

Reads the input deck of ratestate.f90 (unicycle_amd.deck), builds K on the
GPU, and runs the main loop of ratestate.f90:307-328 on the device in chunks
of iterations.  It writes what the reference writes each iteration
(ratestate.f90:253-290, 488-536):

  <wdir>/time.txt               "#               time                 dt", the
                                row (0, dt_next), then per iteration the time
                                at its start and the previous step (ES19.12E2);
  <wdir>/patch-%08d.dat         per observed patch and iteration: the time, the
                                patch's 7 state entries, then their
                                derivatives, at the iteration's start
                                (ES19.12E2, then ES20.12E3);
  <wdir>/opts-<name>.dat        per observation point and iteration: the time
                                and the displacement (north, east, down) at
                                the iteration's start, d = O u of exportPoints
                                (ratestate.f90:414-486), 4 x ES19.12E2;
  stdout                        every 50 iterations: iteration, time, step, vMax.

The VTP geometry files are not written.
"""
from __future__ import annotations

import argparse
import os
import sys

import numpy as np

from . import deck as deckmod
from .engine import PREC_F64, PREC_MIXED, PREC_SPLIT_BF16, Engine

DPATCH = 7   # state entries per patch (STATE_VECTOR_DGF of macros.f90)


def setup(d: deckmod.Deck, device: int = 0, precision: int = PREC_F64) -> Engine:
    eng = Engine(device)
    eng.set_elastic(d.lambda_, d.mu)
    if d.n_rect:
        eng.set_rectangles(d.rect_x, d.rect_length, d.rect_width, d.rect_strike, d.rect_dip)
    if d.n_tri:
        eng.set_triangles(d.tri_i1, d.tri_i2, d.tri_i3, d.vertices)
    eng.set_friction(*d.friction_arrays())
    eng.buildG(precision)
    if d.observation_points:   # buildO (greens.f90:302-462, ratestate.f90:208)
        eng.buildO([p[1:] for p in d.observation_points])
    return eng


def run(d: deckmod.Deck, epsilon: float = 1.0e-6, max_step: float = 1.7976931348623158e308,
        max_iterations: int = 1000000, device: int = 0, precision: int = PREC_F64,
        chunk: int = 20000, out=sys.stdout, write_files: bool = True):
    """Run the deck; returns (iterations, final time, final state)."""
    fmt = deckmod.es_field
    eng = setup(d, device, precision)
    y = eng.initStateVector()
    time, dt_next, dt_done = 0.0, 1.0e-3, 0.0     # ratestate.f90:166-167
    obs = [j - 1 for j in d.observation_states]   # 0-based patches
    idx = np.array([DPATCH * e + k for e in obs for k in range(DPATCH)], dtype=np.int64)
    files, pfiles = [], []
    if write_files:
        os.makedirs(d.wdir, exist_ok=True)
        ftime = open(os.path.join(d.wdir, "time.txt"), "w")
        ftime.write("#               time                 dt\n")
        ftime.write(fmt(0.0, 19, 12, 2) + fmt(dt_next, 19, 12, 2) + "\n")
        files = [open(os.path.join(d.wdir, f"patch-{j:08d}.dat"), "w") for j in d.observation_states]
        # ratestate.f90:291-304
        pfiles = [open(os.path.join(d.wdir, f"opts-{p[0].strip()}.dat"), "w")
                  for p in d.observation_points]
    out.write("#       n               time                 dt       vMax\n")
    out.write(f"{0:09d}" + fmt(time, 19, 12, 2) + fmt(dt_next, 19, 12, 2) + "\n")
    it = 0
    try:
        while it < max_iterations and time < d.interval:
            n = min(chunk, max_iterations - it)
            res = eng.integrate(y, d.interval, dt0=dt_next, max_steps=n, epsilon=epsilon,
                                max_step=max_step, log_cap=n, observe=idx if idx.size else None,
                                t0=time, displacement=bool(d.observation_points))
            for k in range(res.steps):
                i = it + k + 1
                t_start = res.t_log[k - 1] if k else time
                prev = res.dt_log[k - 1] if k else dt_done
                if write_files:
                    ftime.write(fmt(t_start, 19, 12, 2) + fmt(prev, 19, 12, 2) + "\n")
                    for j, f in enumerate(files):
                        ys = res.obs_y[k, DPATCH * j:DPATCH * (j + 1)]
                        ds = res.obs_dydt[k, DPATCH * j:DPATCH * (j + 1)]
                        f.write(fmt(t_start, 19, 12, 2)
                                + "".join(fmt(v, 20, 12, 3) for v in ys)
                                + "".join(fmt(v, 20, 12, 3) for v in ds) + "\n")
                    for j, f in enumerate(pfiles):
                        f.write(fmt(t_start, 19, 12, 2)
                                + "".join(fmt(v, 19, 12, 2) for v in res.disp[k, 3 * j:3 * j + 3])
                                + "\n")
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
            for f in files + pfiles:
                f.close()
        eng.close()
    # the reference's loop index after EXIT (or past the last iteration)
    out.write(f"{it if time >= d.interval else it + 1:09d} time steps.\n")
    return it, time, y


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="unicycle_amd.ratestate",
                                 description="quasi-dynamic earthquake simulation in three-"
                                             "dimensional media (GPU)")
    ap.add_argument("-e", "--epsilon", type=float, default=1.0e-6)
    ap.add_argument("-m", "--maximum-step", type=float, default=1.7976931348623158e308)
    ap.add_argument("-i", "--maximum-iterations", type=int, default=1000000)
    ap.add_argument("--dry-run", action="store_true")
    ap.add_argument("--version", action="store_true")
    ap.add_argument("--precision", choices=["f64", "mixed", "split_bf16"], default="f64",
                    help="K of the stage matvecs (reduced forms keep binary64 state)")
    ap.add_argument("deck", nargs="?", help="input deck (standard input if omitted)")
    a = ap.parse_args(argv)
    if a.version:
        print("unicycle_amd ratestate (gfx950)")
        return 0
    text = open(a.deck).read() if a.deck else sys.stdin.read()
    try:
        d = deckmod.parse(text)
    except deckmod.DeckError as exc:
        print(f"input error: {exc}", file=sys.stderr)
        return 1
    print("# RATESTATE")
    print("# quasi-dynamic earthquake simulation in three-dimensional media")
    print(f"# numerical accuracy: {deckmod.es_field(a.epsilon, 11, 4, 2)}")
    print(f"# maximum iterations: {a.maximum_iterations:11d}")
    print(f"# output directory\n  {d.wdir}")
    print(f"# {d.n_rect} rectangular and {d.n_tri} triangular patches, "
          f"{len(d.observation_states)} observed")
    print(f"# {len(d.observation_points)} observation points")
    if a.dry_run:
        print("dry run: abort calculation")
        return 0
    prec = {"f64": PREC_F64, "mixed": PREC_MIXED, "split_bf16": PREC_SPLIT_BF16}[a.precision]
    run(d, epsilon=a.epsilon, max_step=a.maximum_step, max_iterations=a.maximum_iterations,
        precision=prec)
    return 0


if __name__ == "__main__":
    sys.exit(main())
