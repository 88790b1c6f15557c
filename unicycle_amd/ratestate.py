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
  <wdir>/patch-rectangular.vtp the rectangles' geometry and friction
                                (exportGeometry, ratestate.f90:374-399);
  stdout                        every 50 iterations: iteration, time, step, vMax.

A 2-D antiplane deck (unicycle_amd.antiplane, fortran/2d/antiplane/src/
ratestate.f90:254-276) runs the same loop with 4 state entries per patch,
one displacement column in opts-<name>.dat and no VTP file; a 2-D
plane-strain deck (unicycle_amd.planestrain, fortran/2d/planestrain/src/
ratestate.f90:285-305) with 5 state entries per patch, each value of a
patch row after a blank (its format's X), two displacement columns (east,
down).  The 2-D viscouscycles decks (unicycle_amd.antiplane_viscouscycles,
unicycle_amd.planestrain_viscouscycles) add volume-%08d.dat (4 or 6 state
entries and their derivatives), d, df and dl in opts-<name>.dat and eMax
in the progress line, as the 3-D one.

A viscouscycles deck (unicycle_amd.viscouscycles, viscouscycles.f90:413-434)
runs the same loop with strain volumes and adds: volume-%08d.dat per observed
volume (12 state entries and their derivatives), the fault and flow parts of
the displacement in opts-<name>.dat (10 columns, viscouscycles.f90:612-629),
eMax in the progress line and <wdir>/volume.vtp.
"""
from __future__ import annotations

import argparse
import os
import sys

import numpy as np

from . import deck as deckmod
from . import vtp
from .engine import PREC_F64, PREC_MIXED, PREC_SPLIT_BF16, Engine

DPATCH = 7    # state entries per patch (STATE_VECTOR_DGF_PATCH of macros.f90)
DVOLUME = 12  # state entries per strain volume (STATE_VECTOR_DGF_VOLUME)


def setup(d, device: int = 0, precision: int = PREC_F64) -> Engine:
    """Mesh, friction (and for a viscouscycles deck the volumes and their
    rheology), buildG and buildO on the GPU."""
    eng = Engine(device)
    if isinstance(d, deckmod.AntiplaneDeck):   # fortran/2d
        if d.planestrain:
            eng.set_planestrain(d.lambda_, d.mu, d.x, d.width, d.dip)
            if np.any(d.tau0 > 0):
                eng.set_initial_traction(d.tau0)
            if d.viscous and d.n_vol:
                eng.set_planestrain_volumes(d.vol_x, d.vol_thickness, d.vol_width, d.vol_dip)
        else:
            eng.set_antiplane(d.mu, d.x, d.width, d.dip)
            if d.viscous and d.n_vol:
                eng.set_antiplane_volumes(d.vol_x, d.vol_thickness, d.vol_width, d.vol_dip)
        if d.viscous and d.n_vol:
            eng.set_rheology(*rheology(d))
            if d.planestrain and d.s0.size and np.any(d.s0 != 0.0):
                eng.set_initial_volume_stress(d.s0)
        if d.n:
            eng.set_friction(*d.friction_arrays())
        eng.buildG(precision)
        if d.observation_points:
            eng.buildO([p[1:] for p in d.observation_points])
        return eng
    eng.set_elastic(d.lambda_, d.mu)
    if d.n_rect:
        eng.set_rectangles(d.rect_x, d.rect_length, d.rect_width, d.rect_strike, d.rect_dip)
    if d.n_tri:
        eng.set_triangles(d.tri_i1, d.tri_i2, d.tri_i3, d.vertices)
    if isinstance(d, deckmod.ViscousDeck) and d.n_vol:
        eng.set_volumes(d.vol_x, d.vol_length, d.vol_thickness, d.vol_width, d.vol_strike)
        eng.set_rheology(*rheology(d))
    if d.n_rect + d.n_tri:
        eng.set_friction(*d.friction_arrays())
    eng.buildG(precision)
    if d.observation_points:   # buildO (greens.f90:302-462, ratestate.f90:208)
        eng.buildO([p[1:] for p in d.observation_points])
    return eng


def rheology(d):
    """Engine.set_rheology arguments of a viscouscycles deck (3-D or 2-D): the
    Maxwell rows (gammadot0, n, Gm, Q, R), the volumes' loading strain rates
    and the thermal To.  Without Maxwell rows the volumes do not flow (viscouscycles.f90:963:
    the strain rate stays zero); without thermal rows the Arrhenius factor
    needs Q = 0 (the reference would read an unset temperature)."""
    nv = d.n_vol
    if d.n_maxwell:
        g0, npow, Gm, Q, R = (np.asarray(d.maxwell[k], dtype=np.float64)
                              for k in ("gammadot0", "npower", "Gm", "Q", "R"))
    else:
        g0, npow, Gm, Q, R = np.zeros(nv), np.ones(nv), np.ones(nv), np.zeros(nv), np.ones(nv)
    if d.thermal:
        To = np.asarray(d.thermal["To"], dtype=np.float64)
    else:
        if np.any(Q != 0.0):
            raise deckmod.DeckError("thermal properties are needed when Q is not zero")
        To = np.ones(nv)
    return g0, npow, d.vol_ebg, Gm, Q, R, To


def rect_friction(d: deckmod.Deck):
    """(Vpl, rake, mu0, sig, a, b, L) of the rectangles, deck order."""
    fr = d.friction_arrays()
    return tuple(np.asarray(a)[:d.n_rect] for a in fr[:7])


def run(d: deckmod.Deck, epsilon: float = 1.0e-6, max_step: float = 1.7976931348623158e308,
        max_iterations: int = 1000000, device: int = 0, precision: int = PREC_F64,
        chunk: int = 20000, out=sys.stdout, write_files: bool = True):
    """Run the deck (ratestate, or viscouscycles for a ViscousDeck); returns
    (iterations, final time, final state)."""
    fmt = deckmod.es_field
    viscous3d = isinstance(d, deckmod.ViscousDeck)
    ap = isinstance(d, deckmod.AntiplaneDeck)          # the 2-D programs
    ps = ap and d.planestrain
    viscous = viscous3d or (ap and d.viscous)
    # STATE_VECTOR_DGF_PATCH / _VOLUME of the macros.f90 of each program
    dpatch = 5 if ps else 4 if ap else DPATCH
    dvol = 6 if ps else 4 if ap else DVOLUME
    eng = setup(d, device, precision)
    # initStateVector: ratestate.f90:548-610, viscouscycles.f90:710-820, 2-D
    # antiplane ratestate.f90:568-605, viscouscycles.f90:697-770, 2-D plane
    # strain ratestate.f90:585-627, viscouscycles.f90:714-837
    y = eng.initStateVector(1 if viscous else 0)
    npatch = d.n if ap else d.n_rect + d.n_tri
    nvol = d.n_vol if viscous else 0
    if viscous3d and nvol and not d.n_maxwell:
        # no Maxwell rows: the reference's initial stress reads unset
        # properties; here the volumes start unstressed
        for v in range(nvol):
            y[DPATCH * npatch + DVOLUME * v + 6:DPATCH * npatch + DVOLUME * (v + 1)] = 0.0
    time, dt_next, dt_done = 0.0, 1.0e-3, 0.0     # ratestate.f90:166-167
    obs_vol = d.observation_volumes if viscous else []
    idx = [dpatch * (j - 1) + k for j in d.observation_states for k in range(dpatch)]
    idx += [dpatch * npatch + dvol * (j - 1) + k for j in obs_vol for k in range(dvol)]
    ncol = 2 if ps else 1 if ap else 3   # displacement components per point
    # the plane-strain ratestate program writes each value of a patch row
    # after a blank ("X," in its format, ratestate.f90:543-547)
    sep = " " if ps and not viscous else ""
    idx = np.array(idx, dtype=np.int64)
    files, pfiles = [], []
    if write_files:
        os.makedirs(d.wdir, exist_ok=True)
        # exportGeometry (ratestate.f90:374-399, viscouscycles.f90:473-515;
        # the 2-D program writes none)
        if not ap:
            with open(os.path.join(d.wdir, "patch-rectangular.vtp"), "w") as f:
                f.write(vtp.rectangles_vtp(d.rect_x, d.rect_length, d.rect_width, d.rect_strike,
                                           d.rect_dip, rect_friction(d)))
        if viscous3d:
            with open(os.path.join(d.wdir, "volume.vtp"), "w") as f:
                f.write(vtp.volumes_vtp(d.vol_x, d.vol_length, d.vol_thickness, d.vol_width,
                                        d.vol_strike, d.vol_dip))
        ftime = open(os.path.join(d.wdir, "time.txt"), "w")
        ftime.write("#               time                 dt\n")
        ftime.write(fmt(0.0, 19, 12, 2) + fmt(dt_next, 19, 12, 2) + "\n")
        # observed elements, patches then volumes (ratestate.f90:273-288,
        # viscouscycles.f90:370-392)
        files = [(open(os.path.join(d.wdir, f"patch-{j:08d}.dat"), "w"), dpatch)
                 for j in d.observation_states]
        files += [(open(os.path.join(d.wdir, f"volume-{j:08d}.dat"), "w"), dvol)
                  for j in obs_vol]
        # ratestate.f90:291-304
        pfiles = [open(os.path.join(d.wdir, f"opts-{p[0].strip()}.dat"), "w")
                  for p in d.observation_points]
    out.write("#       n               time                 dt       vMax"
              + ("       eMax" if viscous else "") + "\n")
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
                    o = 0
                    for f, dgf in files:
                        ys = res.obs_y[k, o:o + dgf]
                        ds = res.obs_dydt[k, o:o + dgf]
                        o += dgf
                        f.write(fmt(t_start, 19, 12, 2)
                                + "".join(sep + fmt(v, 20, 12, 3) for v in ys)
                                + "".join(sep + fmt(v, 20, 12, 3) for v in ds) + "\n")
                    for j, f in enumerate(pfiles):
                        c0, c1 = ncol * j, ncol * (j + 1)
                        cols = [res.disp[k, c0:c1]]
                        if viscous:
                            cols += [res.disp_f[k, c0:c1], res.disp_l[k, c0:c1]]
                        f.write(fmt(t_start, 19, 12, 2)
                                + "".join(fmt(v, 19, 12, 2) for c in cols for v in c) + "\n")
                if i % 50 == 0:
                    line = (f"{i:09d}" + fmt(t_start, 19, 12, 2) + fmt(prev, 19, 12, 2)
                            + fmt(res.vmax_log[k], 11, 4, 2))
                    if viscous:
                        line += fmt(res.emax_log[k], 11, 4, 2)
                    out.write(line + "\n")
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
            for f in [f for f, _ in files] + pfiles:
                f.close()
        eng.close()
    # the reference's loop index after EXIT (or past the last iteration)
    out.write(f"{it if time >= d.interval else it + 1:09d} time steps.\n")
    return it, time, y


def main(argv=None, viscous: bool = False, antiplane: bool = False,
         planestrain: bool = False) -> int:
    two_d = antiplane or planestrain
    name = ("antiplane" if antiplane else "planestrain" if planestrain else "")
    name = (name + "_viscouscycles" if viscous else name) if two_d else (
        "viscouscycles" if viscous else "ratestate")
    media = ("antiplane strain" if antiplane else "plane strain" if planestrain
             else "three-dimensional media")
    if viscous:
        media = "viscoelastic media, " + media if two_d else "three-dimensional viscoelastic media"
    ap = argparse.ArgumentParser(prog=f"unicycle_amd.{name}",
                                 description=f"quasi-dynamic earthquake simulation in {media} "
                                             "(GPU)")
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
        print(f"unicycle_amd {name} (gfx950)")
        return 0
    text = open(a.deck).read() if a.deck else sys.stdin.read()
    try:
        if antiplane:
            d = (deckmod.parse_antiplane_viscous if viscous else deckmod.parse_antiplane)(text)
        elif planestrain:
            d = (deckmod.parse_planestrain_viscous if viscous else deckmod.parse_planestrain)(text)
        else:
            d = deckmod.parse_viscous(text) if viscous else deckmod.parse(text)
        if viscous and d.n_vol:
            rheology(d)
    except deckmod.DeckError as exc:
        print(f"input error: {exc}", file=sys.stderr)
        return 1
    print("# VISCOUSCYCLES" if viscous else "# RATESTATE")
    print(f"# quasi-dynamic earthquake simulation in {media}")
    print(f"# numerical accuracy: {deckmod.es_field(a.epsilon, 11, 4, 2)}")
    print(f"# maximum iterations: {a.maximum_iterations:11d}")
    print(f"# output directory\n  {d.wdir}")
    if two_d:
        print(f"# {d.n} patches, {len(d.observation_states)} observed")
    else:
        print(f"# {d.n_rect} rectangular and {d.n_tri} triangular patches, "
              f"{len(d.observation_states)} observed")
    if viscous:
        print(f"# {d.n_vol} strain volumes, {len(d.observation_volumes)} observed")
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
