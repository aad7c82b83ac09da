"""The launch-shape autotuner (tp_autotune, tp_batch_autotune).

Tuning must not change the caller's beam, must apply a valid shape, must not change any
tracking result (every shape runs the same per-particle code), and must reuse a cached
choice without timing.
"""

import numpy as np
import pytest

from trackpad_hip import HIPBatch, HIPContext, HIPParticles, Beam, compile_table, ringpass_grad, track_
from trackpad_hip import workloads as W

pytestmark = pytest.mark.gpu


def _dl(p):
    c, lf, lt, le = p.download()
    return np.ascontiguousarray(c), lf, lt, le


def test_autotune_applies_shape_and_keeps_results(tmp_path):
    ctx = HIPContext(0)
    beam = Beam(W.ENERGY)
    lat = W.dba_ring(beam=beam)
    tab = ctx.table(compile_table(lat, beam))
    c0 = W.dba_particles(5000, seed=31)
    c0[:4, 1] = 0.9
    p = HIPParticles.from_coords(c0, ctx=ctx)
    before = _dl(p)
    cache = tmp_path / "tune.txt"
    cfg = ctx.autotune(tab, p, probe_turns=2, fast=True, target_launch_ms=50.0, cache=cache)
    assert cfg["candidates"] == 12  # block {64,128,256} x (ppl 1 x budgets {0,5,6} + ppl 2)
    assert cfg["block"] in (64, 128, 256) and cfg["ppl"] in (1, 2) and cfg["fwd_waves"] in (0, 5, 6)
    assert cfg["turns_per_launch"] >= 1 and cfg["rate"] > 0
    after = _dl(p)
    for a, b in zip(before, after):
        assert np.array_equal(a, b)  # the caller's beam is untouched
    # results do not depend on the chosen shape (parity build: bit for bit)
    track_(p, tab, nturns=5)
    tuned = _dl(p)
    ref_ctx = HIPContext(0)
    q = HIPParticles.from_coords(c0, ctx=ref_ctx)
    track_(q, ref_ctx.table(compile_table(lat, beam)), nturns=5)
    ref = _dl(q)
    for a, b in zip(tuned, ref):
        assert np.array_equal(a, b)
    # a second call hits the cache: same choice, nothing timed
    lines = cache.read_text().splitlines()
    assert len(lines) == 1
    cfg2 = ctx.autotune(tab, p, probe_turns=2, fast=True, target_launch_ms=50.0, cache=cache)
    assert cfg2["candidates"] == 0
    for k in ("block", "ppl", "turns_per_launch", "fwd_waves"):
        assert cfg2[k] == cfg[k]
    assert len(cache.read_text().splitlines()) == 1
    # a different beam size is a different key
    p2 = HIPParticles.from_coords(c0[:1000], ctx=ctx)
    cfg3 = ctx.autotune(tab, p2, probe_turns=2, fast=True, target_launch_ms=50.0, cache=cache)
    assert cfg3["candidates"] == 12 and len(cache.read_text().splitlines()) == 2


def test_batch_autotune_restores_beams(tmp_path):
    ctx = HIPContext(0)
    beam = Beam(W.ENERGY)
    rings, _, _ = W.variant_rings(nvar=3, seed=32)
    tabs = [ctx.table(compile_table(r, beam)) for r in rings]
    coords = [W.dba_particles(n, seed=33 + i) for i, n in enumerate((300, 70, 1))]
    parts = [HIPParticles.from_coords(c, ctx=ctx) for c in coords]
    b = HIPBatch(tabs, parts)
    before = [_dl(p) for p in parts]
    cfg = b.autotune([2, 3, 1], ckpt_stride=2, adjoint=True, fast=True, cache=tmp_path / "b.txt")
    assert cfg["candidates"] == 9 + 3
    assert cfg["block"] in (64, 128, 256) and cfg["fwd_waves"] in (0, 5, 6) and cfg["back_waves"] in (0, 3, 4)
    for p, bf in zip(parts, before):
        for a, c in zip(_dl(p), bf):
            assert np.array_equal(a, c)
    g = b.adjoint([4, 3, 2], ckpt_stride=2)
    for v, (r, c) in enumerate(zip(rings, coords)):
        ref = ringpass_grad(c, r, beam, [4, 3, 2][v], ckpt_stride=2, ctx=ctx)
        assert np.array_equal(g.dx0[v], ref.dx0)
        assert g.loss[v] == pytest.approx(ref.loss, rel=1e-13)
    cfg2 = b.autotune([2, 3, 1], ckpt_stride=2, adjoint=True, fast=True, cache=tmp_path / "b.txt")
    assert cfg2["candidates"] == 0 and cfg2["back_waves"] == cfg["back_waves"]
