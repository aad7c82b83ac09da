"""Batched persistent launch (tp_batch_*) against per-variant launches and the oracle.

A batch must reproduce, variant by variant, what tp_track / tp_adjoint give for that
variant alone: the same per-lane code runs in both, so the parity build agrees bit for
bit (coordinates, loss records, dL/dx0); strength gradients are sums over a different
partition of the beam (per tile instead of per block) and agree to 1e-12 relative.  The
oracle and quad-precision finite differences pin the batched results directly as well.
"""

import numpy as np
import pytest

from oracle import pyoracle as O
from trackpad_hip import (Beam, Drift, ExactSBend, HIPBatch, HIPParticles, Lattice, Quadrupole, Sextupole,
                          TrackPadHIPError, compile_table, ringpass_grad, set_strengths, strength_positions, track_)
from trackpad_hip import workloads as W

pytestmark = pytest.mark.gpu


def _beam():
    return Beam(W.ENERGY)


def _variants(nvar, seed):
    rings, _, _ = W.variant_rings(nvar=nvar, seed=seed)
    return rings


def _coords(n, seed, nlost=0):
    c = W.dba_particles(n, seed=seed)
    c[:nlost, 1] = 0.9  # lost in the first drift
    return c


def _same(a, b, fast, what):
    """Parity build: bit for bit.  Fast build: FMA contraction is the compiler's choice per
    kernel, and the drifts round differently in the batch kernel (series form of 1/pz) and in
    a single-beam launch that leaves its waves alone on their SIMDs (refined rsqrt), so
    within 1e-9 of each plane's largest entry (the differences grow through the turns), plus
    1e-12 absolute: z adds NormL (beti + delta) - L beti at every drift, a difference of
    terms of the drift's length, so its rounding is absolute (~1e-16 m per drift) however
    small z itself is."""
    if not fast or a.size == 0:
        assert np.array_equal(a, b), (what, np.max(np.abs(a - b)) if a.size else 0)
    else:
        scale = np.max(np.abs(b), axis=0, keepdims=True)
        assert np.all(np.abs(a - b) <= 1e-9 * np.where(scale > 0, scale, 1.0) + 1e-12), what


def _cross_variant_close(g, r, v, fast):
    """Variant v of a batched adjoint against its own tp_adjoint call.  Parity build: loss
    within 1e-13, dx0 bit for bit, dk within 1e-12 of the largest.  Fast build: the kernels
    round differently (FMA contraction is the compiler's choice per kernel; the batch kernel's
    drifts take the series form of 1/pz, a small single-beam launch the refined rsqrt, and a
    batch holding a variant with more features runs that variant's kernel for all), and the
    differences grow through the turns, so loss, dx0 and dk agree within the adjoint's own
    accuracy, 1e-9."""
    tol = 1e-9 if fast else 1e-13
    assert abs(g.loss[v] - r.loss) <= tol * abs(r.loss), (v, abs(g.loss[v] - r.loss) / abs(r.loss))
    if fast:
        scale = np.max(np.abs(r.dx0), axis=0, keepdims=True)
        assert np.all(np.abs(g.dx0[v] - r.dx0) <= 1e-9 * scale), v
    else:
        assert np.array_equal(g.dx0[v], r.dx0), v
    scale = np.max(np.abs(r.dk))
    tol = 1e-9 if fast else 1e-12
    assert np.all(np.abs(g.dk[v] - r.dk) <= tol * scale), (v, np.max(np.abs(g.dk[v] - r.dk)) / scale)


def _per_variant_track(ctx, tabs, coords, flags, ts, t0s, fast):
    out = []
    for tab, c, lf, t, t0 in zip(tabs, coords, flags, ts, t0s):
        p = HIPParticles.from_coords(c, lf, ctx)
        track_(p, tab, nturns=int(t), turn=int(t0), fast=fast)
        out.append(p.download())
    return out


def _setup(ctx, beam):
    rings = _variants(4, seed=3)
    lats = rings + [W.fodo_lattice(), rings[0]]
    ns = [130, 64, 1, 0, 77, 200]
    ts = np.array([3, 5, 2, 4, 0, 1], np.int32)
    t0s = np.array([0, 10, 0, 0, 5, 7], np.int32)
    coords = [_coords(n, seed=40 + i, nlost=min(n, 3) if i in (0, 5) else 0) for i, n in enumerate(ns)]
    flags = []
    for i, n in enumerate(ns):
        lf = np.zeros(n, np.int64)
        if n > 4:
            lf[4] = 1  # arrives flagged: skipped, lost_turn -2
        flags.append(lf)
    tabs = [ctx.table(compile_table(l, beam)) for l in lats]
    return lats, tabs, coords, flags, ts, t0s


@pytest.mark.parametrize("fast", [False, True])
def test_batch_track_matches_per_variant(ctx, fast):
    beam = _beam()
    lats, tabs, coords, flags, ts, t0s = _setup(ctx, beam)
    parts = [HIPParticles.from_coords(c, lf, ctx) for c, lf in zip(coords, flags)]
    b = HIPBatch(tabs, parts)
    b.track(ts, t0s, fast=fast)
    ref = _per_variant_track(ctx, tabs, coords, flags, ts, t0s, fast)
    for v, (p, r) in enumerate(zip(parts, ref)):
        got = p.download()
        for k in (1, 2, 3):  # lost flags, turns, positions
            assert np.array_equal(got[k], r[k]), (v, k)
        _same(got[0], r[0], fast, v)
    # losses exercised, with the turn offset of their variant
    lt0 = parts[0].download()[2]
    assert np.all(lt0[:3] == t0s[0]) and lt0[4] == -2
    assert np.all(parts[5].download()[2][:3] == t0s[5])
    info = b.info()
    assert info["n_items"] == sum((n + 63) // 64 for n, t in zip([130, 64, 1, 0, 77, 200], ts) if t > 0)


def test_batch_track_against_oracle(ctx):
    """Parity build of the batch against the oracle directly (DBA variant, 20 turns)."""
    beam = _beam()
    rings = _variants(3, seed=5)
    tabs = [ctx.table(compile_table(l, beam)) for l in rings]
    coords = [_coords(96, seed=50 + i, nlost=2) for i in range(3)]
    parts = [HIPParticles.from_coords(c, None, ctx) for c in coords]
    b = HIPBatch(tabs, parts)
    ts = np.array([20, 7, 13], np.int32)
    b.track(ts)
    for lat, c, p, t in zip(rings, coords, parts, ts):
        oc, olf, olt, ole = O.ringpass(c, lat, beam, np.zeros(len(c), np.int64), int(t), nthreads=8)
        gc, glf, glt, gle = p.download()
        assert np.array_equal(glf, olf) and np.array_equal(glt, olt) and np.array_equal(gle, ole)
        scale = np.maximum(np.abs(oc), np.sqrt(np.mean(oc * oc, axis=0, keepdims=True)))
        assert np.max(np.abs(gc - oc) / np.where(scale > 0, scale, 1.0)) <= 1e-12


def test_batch_graph_replay_and_replan(ctx):
    """Replaying the captured graph continues tracking; a new turn vector re-plans."""
    beam = _beam()
    rings = _variants(3, seed=6)
    tabs = [ctx.table(compile_table(l, beam)) for l in rings]
    coords = [_coords(n, seed=60 + i) for i, n in enumerate((70, 129, 5))]
    parts = [HIPParticles.from_coords(c, None, ctx) for c in coords]
    b = HIPBatch(tabs, parts)
    ts = np.array([2, 3, 1], np.int32)
    b.track(ts)
    b.track(ts, ts)  # same shape except turn0: re-plan; then a pure replay
    b.track(ts, ts)
    ts2 = np.array([1, 1, 4], np.int32)
    b.track(ts2, 3 * ts)
    for lat, c, p, t, t2 in zip(rings, coords, parts, ts, ts2):
        ref = HIPParticles.from_coords(c, None, ctx)
        tab = ctx.table(compile_table(lat, beam))
        track_(ref, tab, nturns=int(t))
        track_(ref, tab, nturns=int(t), turn=int(t))
        track_(ref, tab, nturns=int(t), turn=int(t))
        track_(ref, tab, nturns=int(t2), turn=int(3 * t))
        r, g = ref.download(), p.download()
        assert np.array_equal(r[0], g[0]) and np.array_equal(r[2], g[2])


def test_batch_many_ragged_variants(ctx):
    """40 variants of random size (1..300) and turn count (1..6): the work queue hands out
    1..5 tiles per variant in cost order; every variant matches its own launch."""
    beam = _beam()
    rng = np.random.default_rng(8)
    base = _variants(4, seed=8)
    nv = 40
    lats = [base[i % 4] for i in range(nv)]
    ns = rng.integers(1, 301, nv)
    ts = rng.integers(1, 7, nv).astype(np.int32)
    coords = [_coords(int(n), seed=100 + i, nlost=int(i % 3)) for i, n in enumerate(ns)]
    tabs = [ctx.table(compile_table(l, beam)) for l in lats]
    parts = [HIPParticles.from_coords(c, None, ctx) for c in coords]
    b = HIPBatch(tabs, parts)
    b.set_launch(128, 5, 3)
    b.track(ts)
    ref = _per_variant_track(ctx, tabs, coords, [None] * nv, ts, np.zeros(nv, np.int32), False)
    for v, (p, r) in enumerate(zip(parts, ref)):
        g = p.download()
        assert np.array_equal(g[0], r[0]) and np.array_equal(g[2], r[2]) and np.array_equal(g[3], r[3]), v


@pytest.mark.parametrize("fast", [False, True])
def test_batch_adjoint_matches_per_variant(ctx, fast):
    beam = _beam()
    rings = _variants(3, seed=9)
    lats = rings + [W.fodo_lattice()]
    ns = [100, 64, 1, 50]
    ts = np.array([4, 7, 3, 5], np.int32)
    coords = [_coords(n, seed=70 + i, nlost=2 if i == 0 else 0) for i, n in enumerate(ns)]
    tabs = [ctx.table(compile_table(l, beam)) for l in lats]
    parts = [HIPParticles.from_coords(c, None, ctx) for c in coords]
    b = HIPBatch(tabs, parts)
    g = b.adjoint(ts, ckpt_stride=2, fast=fast)
    for v, (lat, c, t) in enumerate(zip(lats, coords, ts)):
        r = ringpass_grad(c, lat, beam, int(t), ckpt_stride=2, fast=fast, ctx=ctx)
        assert g.n_alive[v] == r.n_alive
        fin = parts[v].download()
        assert np.array_equal(fin[1], r.lost_flags)
        _same(np.ascontiguousarray(fin[0]), r.final, fast, v)
        _cross_variant_close(g, r, v, fast)
        assert np.all(g.dk[v][r.dk == 0.0] == 0.0)
    assert g.n_alive[0] == 98


@pytest.mark.parametrize("fast", [False, True])
def test_batch_adjoint_forest_variant(ctx, fast):
    """A batch holding a ring of the reference's ExactSBend (Forest yrot / fringe / wedge) next
    to a DBA variant runs the reverse kernel variant with the Forest maps for both; each
    variant's results equal its own tp_adjoint call as in the test above."""
    import lattices as LT
    beam = _beam()
    lats = [LT.forest_ring(W.ENERGY), _variants(1, seed=10)[0]]
    ns = [70, 64]
    ts = np.array([5, 3], np.int32)
    coords = [W.gaussian_particles(ns[0], 33, 2e-4, 2e-5, 2e-4, 2e-5, 1e-3), _coords(ns[1], seed=80)]
    tabs = [ctx.table(compile_table(l, beam)) for l in lats]
    parts = [HIPParticles.from_coords(c, None, ctx) for c in coords]
    b = HIPBatch(tabs, parts)
    g = b.adjoint(ts, ckpt_stride=2, fast=fast)
    for v, (lat, c, t) in enumerate(zip(lats, coords, ts)):
        r = ringpass_grad(c, lat, beam, int(t), ckpt_stride=2, fast=fast, ctx=ctx)
        assert g.n_alive[v] == r.n_alive == ns[v]
        _cross_variant_close(g, r, v, fast)


@pytest.mark.parametrize("fast", [False, True])
def test_batch_adjoint_every_family(ctx, fast):
    """A batch holding the ring of every family (rolled quadrupoles among them: the batch
    carries the rotation inverses at the concatenated pool offsets), a DBA variant and the
    Forest ring: each variant's gradients equal its own call."""
    import lattices as LT
    beam = _beam()
    lats = [_variants(1, seed=11)[0], LT.families_ring(), LT.forest_ring(W.ENERGY)]
    ns = [64, 70, 40]
    ts = np.array([3, 4, 2], np.int32)
    rng = np.random.default_rng(91)
    coords = [_coords(ns[0], seed=81),
              rng.normal(0, 1, (ns[1], 6)) * np.array([3e-4, 3e-5, 3e-4, 3e-5, 1e-3, 5e-4]),
              W.gaussian_particles(ns[2], 34, 2e-4, 2e-5, 2e-4, 2e-5, 1e-3)]
    tabs = [ctx.table(compile_table(l, beam)) for l in lats]
    parts = [HIPParticles.from_coords(c, None, ctx) for c in coords]
    b = HIPBatch(tabs, parts)
    g = b.adjoint(ts, ckpt_stride=2, fast=fast)
    for v, (lat, c, t) in enumerate(zip(lats, coords, ts)):
        r = ringpass_grad(c, lat, beam, int(t), ckpt_stride=2, fast=fast, ctx=ctx)
        assert g.n_alive[v] == r.n_alive == ns[v]
        _cross_variant_close(g, r, v, fast)


def test_batch_adjoint_vs_quad_precision_fd(ctx):
    """One variant of a batch (DBA, 8 particles, 30 turns, checkpoints every 7 turns):
    central differences with h = 1e-7 in quad precision agree to 1e-7 with the batched
    adjoint for the largest quad and sextupole gradients and two initial coordinates."""
    import mpmath as mp

    beam = _beam()
    rings = _variants(2, seed=10)
    n, T, h = 8, 30, 1e-7
    c0 = [_coords(n, seed=80), _coords(40, seed=81)]
    tabs = [ctx.table(compile_table(l, beam)) for l in rings]
    parts = [HIPParticles.from_coords(c, None, ctx) for c in c0]
    b = HIPBatch(tabs, parts)
    g = b.adjoint(np.array([T, 11], np.int32), ckpt_stride=7)
    lat, c = rings[0], c0[0]
    dk, dx0 = g.dk[0], g.dx0[0]

    def loss_q(lat_, c_):
        (hi, lo), na = O.loss_quad(c_, lat_, beam, T, nthreads=8)
        assert na == n
        return (mp.mpf(hi) + mp.mpf(lo)) / na

    pos = strength_positions(lat)
    quads = sorted((p for p in pos if isinstance(lat[p], Quadrupole)), key=lambda p: -abs(dk[p]))
    sexts = sorted((p for p in pos if isinstance(lat[p], Sextupole)), key=lambda p: -abs(dk[p]))
    for p in (quads[0], sexts[0]):
        k0 = lat[p].k1 if isinstance(lat[p], Quadrupole) else lat[p].k2
        fd = float((loss_q(set_strengths(lat, [p], [k0 + h]), c) - loss_q(set_strengths(lat, [p], [k0 - h]), c))
                   / (2 * h))
        assert abs(dk[p] - fd) <= 1e-7 * abs(fd), (p, dk[p], fd)
    for (i, j) in ((0, 0), (5, 2)):
        hh = h * 1e-3
        cp, cm = c.copy(), c.copy()
        cp[i, j] += hh
        cm[i, j] -= hh
        fd = float((loss_q(lat, cp) - loss_q(lat, cm)) / (2 * hh))
        assert abs(dx0[i, j] - fd) <= 1e-7 * abs(fd), (i, j, dx0[i, j], fd)


def test_batch_refusals_and_lifetime(ctx):
    beam = _beam()
    ring = _variants(1, seed=12)[0]
    other = Lattice([Drift(1.0), ExactSBend(1.0, 1e-8), Quadrupole(0.4, 1.0)])  # drift-limit bend: no adjoint
    c = _coords(20, seed=90)
    tabs = [ctx.table(compile_table(ring, beam)), ctx.table(compile_table(other, beam))]
    parts = [HIPParticles.from_coords(c, None, ctx) for _ in range(2)]
    b = HIPBatch(tabs, parts)
    b.track([2, 2])  # tracking is supported
    with pytest.raises(TrackPadHIPError, match="TP_EUNSUPPORTED"):
        b.adjoint([2, 2])
    with pytest.raises(TrackPadHIPError, match="TP_EINVAL"):
        b.track([-1, 2])
    mixed = [tabs[0], ctx.table(compile_table(ring, beam, exact_hamiltonian=False))]
    with pytest.raises(TrackPadHIPError, match="TP_EUNSUPPORTED"):
        HIPBatch(mixed, parts)
    # the batch keeps its particle sets alive after their owners release them
    q = [HIPParticles.from_coords(c, None, ctx) for _ in range(2)]
    b2 = HIPBatch([tabs[0], tabs[0]], q)
    b2.particles = []
    for p in q:
        p._fin()
    del q
    b2.track([1, 3])
    b2.close()
    # an empty batch is valid
    e = HIPBatch([], [], ctx=ctx)
    e.track([])
    assert e.adjoint([]).loss.size == 0
