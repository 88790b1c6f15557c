"""Instances of one geometry beyond 64: the library stores them as
consecutive groups of 64 instance columns ("slices") that share one K, and
the stage GEMM streams that K once for all the slices of the family (the
waves of a block take the slices of one row item and k-split).

- matvec: every slice against K V in binary64, within the bound of each
  precision (test_gpu_ensemble.py states the bounds), on every tiling
  variant the shared form admits (with bf16+i8 K, the wave-shaped
  direct-to-LDS variants too); the variants that stage one V per block are
  refused.
- four slices: the K-sharing variant with every wave busy, bit for bit
  against the wave-shaped variant.
- stepping: a family of two slices runs the same arithmetic as the same
  instances added as two separate groups (each with its own copy of K) on
  the same tiling, so every instance's state, time and counters agree bit
  for bit.
"""
import numpy as np
import pytest

import orc
from unicycle_amd import UcyError, configs
from unicycle_amd.ensemble import Ensemble, PREC_BF16_I8, PREC_F64, PREC_MIXED, PREC_SPLIT_BF16

pytestmark = pytest.mark.gpu


def _instances(model, n_inst, seed):
    draws = configs.ensemble_instances(model, n_inst, seed)
    zeros = np.zeros((3 * model.n, 2 * model.n))
    y0 = []
    for i in range(n_inst):
        s = orc.System(configs.instance_friction(model, draws, i), zeros)
        y0.append(configs.perturbed_initial_state(model, s.init_state(), int(draws["seeds"][i])))
    return draws, np.array(y0)


def _add(ens, model, d, y0, sl=slice(None)):
    return ens.add_group(model, d["a"][sl], d["b"][sl], d["sig"][sl], d["L"][sl], y0[sl])


@pytest.mark.parametrize("precision", [PREC_SPLIT_BF16, PREC_F64, PREC_MIXED, PREC_BF16_I8])
def test_shared_k_matvec(engine_factory, precision):
    eng = engine_factory()
    ma = configs.thrust_fault(nx=8, nz=8, ellipse=None)
    mb = configs.thrust_fault(nx=16, nz=8, ellipse=None)
    da, ya = _instances(ma, 150, seed=1)
    db, yb = _instances(mb, 37, seed=2)
    ens = Ensemble(eng, precision)
    assert _add(ens, ma, da, ya) == 0
    assert _add(ens, mb, db, yb) == 3
    assert ens.groups == [(ma.n, 64), (ma.n, 64), (ma.n, 22), (mb.n, 37)]
    assert ens.families == [[0, 1, 2], [3]]
    ens.build()
    Ka = orc.buildG(orc.Mesh(ma.lambda_, ma.mu, rect=ma.geometry()), nthreads=8)
    Kb = orc.buildG(orc.Mesh(mb.lambda_, mb.mu, rect=mb.geometry()), nthreads=8)
    Ks = [Ka, Ka, Ka, Kb]
    rng = np.random.default_rng(0)
    variants = {PREC_F64: [0], PREC_BF16_I8: [0, 1, 2, 5, 8, 9, 10]}.get(precision,
                                                                        [0, 1, 2, 5, 9, 10])
    Vs = [rng.standard_normal((K.shape[1], n)) * 10.0 ** rng.uniform(-12, 0, (K.shape[1], n))
          for K, (_, n) in zip(Ks, ens.groups)]
    got = {}
    for kb, v in [(kb, v) for v in variants for kb in (4, 64)]:
        ens.set_tiling(kb, v)
        for g, K in enumerate(Ks):
            V = Vs[g]
            T = ens.matvec(g, V)
            got[kb, v, g] = T
            want = K @ V
            scale = np.abs(K) @ np.abs(V)
            err = np.abs(T - want)
            bound = 1e-13 if precision == PREC_F64 else 2.0 ** -13
            assert np.all(err <= bound * scale), (kb, v, g, (err / scale).max())
    if precision != PREC_F64:
        # the K-sharing variant runs the same products in the same order as
        # the wave-shaped 128-row variant: equal bit for bit, every slice
        # (three of a family of 3 with a wave left over, and a family of 1)
        for kb in (4, 64):
            for g in range(len(Ks)):
                assert np.array_equal(got[kb, 10, g], got[kb, 5, g]), (kb, g)
    if precision != PREC_F64:
        for v in (3, 4) + ((6, 7) if precision == PREC_BF16_I8 else ()):
            with pytest.raises(UcyError):
                ens.set_tiling(64, v)


@pytest.mark.parametrize("precision", [PREC_F64, PREC_BF16_I8])
def test_shared_k_stepping_equals_separate_groups(engine_factory, precision):
    eng = engine_factory()
    model = configs.thrust_fault(nx=8, nz=4, ellipse=None)
    d, y0 = _instances(model, 100, seed=7)
    runs = []
    # (shared K, GEMM variant): the same split-K sums in every run; in
    # bf16+i8 also the K-sharing variant 10 on the family, whose products
    # run in the wave-shaped variants' order
    cases = [(True, 0), (False, 0)] + ([(True, 10)] if precision == PREC_BF16_I8 else [])
    for shared, variant in cases:
        ens = Ensemble(eng, precision)
        if shared:
            _add(ens, model, d, y0)
            assert ens.families == [[0, 1]]
        else:
            _add(ens, model, d, y0, slice(0, 64))
            _add(ens, model, d, y0, slice(64, 100))
            assert ens.families == [[0], [1]]
        ens.build()
        ens.set_tiling(4, variant)
        ens.reset()
        st = ens.run(1e30, max_attempts=1500)
        assert st.n_failed == 0
        runs.append([ens.state(g) for g in range(2)])
        ens.close()
    for other in runs[1:]:
        for g in range(2):
            for a, b in zip(runs[0][g], other[g]):
                assert np.array_equal(a, b)


@pytest.mark.parametrize("precision", [PREC_SPLIT_BF16, PREC_BF16_I8])
def test_kshare_four_slices(engine_factory, precision):
    """A full family of four slices (256 instances): every wave of the
    K-sharing variant multiplies (none idle, as with fewer slices), its
    product equals the wave-shaped 128-row variant's bit for bit, slice by
    slice, at one and at several k-splits (a 16 x 8 thrust: 16 k-blocks;
    384 rows, three row items, the last of two 32-row blocks), and K V
    within the precision's bound."""
    eng = engine_factory()
    model = configs.thrust_fault(nx=16, nz=8, ellipse=None)
    d, y0 = _instances(model, 256, seed=11)
    ens = Ensemble(eng, precision)
    assert _add(ens, model, d, y0) == 0
    assert ens.families == [[0, 1, 2, 3]]
    ens.build()
    K = orc.buildG(orc.Mesh(model.lambda_, model.mu, rect=model.geometry()), nthreads=8)
    rng = np.random.default_rng(3)
    V = [rng.standard_normal((K.shape[1], 64)) * 10.0 ** rng.uniform(-12, 0, (K.shape[1], 64))
         for _ in range(4)]
    for kb in (4, 64):
        got = {}
        for v in (5, 10):
            ens.set_tiling(kb, v)
            got[v] = [ens.matvec(g, V[g]) for g in range(4)]
        for g in range(4):
            assert np.array_equal(got[10][g], got[5][g]), (kb, g)
            err = np.abs(got[10][g] - K @ V[g])
            assert np.all(err <= 2.0 ** -13 * (np.abs(K) @ np.abs(V[g]))), (kb, g)
    ens.close()


def test_kshare_target_size(engine_factory):
    """The 256-instance target's shape: one 64 x 64 thrust (4096 patches, K of
    12288 x 8192) as a family of four slices in bf16+i8.  The K-sharing
    variant at the tuner's depth for it (104 k-blocks: 5 splits) and at 64
    (8 splits) equals the wave-shaped 128-row variant bit for bit, every
    slice, and each slice's product is within 2^-13 sum |K||V| per row of
    K V in binary64 with the K the plain path assembles from the same
    inputs."""
    eng = engine_factory()
    model = configs.thrust_fault(nx=64, nz=64, ellipse=None)
    plain = engine_factory()
    plain.set_elastic(model.lambda_, model.mu)
    plain.set_rectangles(*model.geometry())
    plain.set_friction(*model.friction())
    y_base = plain.initStateVector()
    plain.buildG()
    K = plain.get_kernel()
    d = configs.ensemble_instances(model, 256, seed=12)
    ens = Ensemble(eng, PREC_BF16_I8)
    assert ens.add_group(model, d["a"], d["b"], d["sig"], d["L"], np.tile(y_base, (256, 1))) == 0
    assert ens.families == [[0, 1, 2, 3]]
    ens.build()
    rng = np.random.default_rng(13)
    V = [rng.standard_normal((K.shape[1], 64)) * 10.0 ** rng.uniform(-12, 0, (K.shape[1], 64))
         for _ in range(4)]
    want = [K @ v for v in V]
    scale = [np.abs(K) @ np.abs(v) for v in V]
    for kb in (104, 64):
        got = {}
        for v in (5, 10):
            ens.set_tiling(kb, v)
            got[v] = [ens.matvec(g, V[g]) for g in range(4)]
        for g in range(4):
            assert np.array_equal(got[10][g], got[5][g]), (kb, g)
            assert np.all(np.abs(got[10][g] - want[g]) <= 2.0 ** -13 * scale[g]), (kb, g)
    ens.close()
