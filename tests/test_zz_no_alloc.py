"""tp_track_noalloc (the tracking entry point that allocates nothing: its survivor lists and
queue counter live in a static device array) against tp_track, through the C ABI: the same
bits, two calls in a row with the second smaller, and the refusal of a particle set larger
than the array holds.  Collected last (test_zz_*): it runs after the other GPU tests and also
alone in a fresh process."""

import ctypes as C

import numpy as np
import pytest

import trackpad_hip._lib as L
from trackpad_hip import Beam, HIPParticles, TrackPadHIPError, compile_table
from trackpad_hip import workloads as W

pytestmark = pytest.mark.gpu


def _run(ctx, fn, tab, c0, nturns, flags, tpl):
    ctx.set_launch(0, 0, tpl)
    try:
        p = HIPParticles.from_coords(c0, None, ctx)
        L.check(fn(ctx.handle, tab.handle, p.handle, 0, int(nturns), flags))
        c, lf, lt, le = p.download()
        return np.ascontiguousarray(c), lf, lt, le
    finally:
        ctx.set_launch(0, 0, 0)


def test_noalloc_matches_tp_track(ctx):
    lib = ctx._lib
    beam = Beam(W.ENERGY)
    # 1) a large FODO beam with aperture losses: persistent launches (queue counter in the
    #    scratch) and survivor compaction between one-turn launches (lists in the scratch)
    n1 = 400_000
    assert n1 <= L.TP_NOALLOC_MAX_N
    c1 = W.fodo_particles(n1, seed=21)
    c1[::50, 0] = 0.021
    fodo = ctx.table(compile_table(W.fodo_lattice(), beam, aperture=0.02))
    for flags in (0, L.TP_TRACK_FAST):
        ref = _run(ctx, lib.tp_track, fodo, c1, 4, flags, 1)
        got = _run(ctx, lib.tp_track_noalloc, fodo, c1, 4, flags, 1)
        assert ref[1].sum() >= n1 // 50
        for a, b in zip(ref, got):
            assert np.array_equal(a, b), flags
    # 2) then a smaller DBA beam losing a third of its particles, launches of a quarter turn
    #    and one turn: the scratch keeps the last call's contents and must not leak into this one
    n2 = 3000
    c2 = W.dba_particles(n2, seed=8)
    c2[::3, 1] = 0.3
    dba = ctx.table(compile_table(W.dba_ring(beam=beam), beam))
    for flags in (0, L.TP_TRACK_FAST):
        ref = _run(ctx, lib.tp_track, dba, c2, 3, flags, 1)
        got = _run(ctx, lib.tp_track_noalloc, dba, c2, 3, flags, 1)
        assert 0.2 < ref[1].mean() < 0.5
        for a, b in zip(ref, got):
            assert np.array_equal(a, b), flags
    # 3) a single launch without compaction, and the Python entry point
    from trackpad_hip import track_
    p = HIPParticles.from_coords(c2, None, ctx)
    track_(p, dba, nturns=2, fast=True, no_alloc=True)
    q = HIPParticles.from_coords(c2, None, ctx)
    track_(q, dba, nturns=2, fast=True)
    assert all(np.array_equal(a, b) for a, b in zip(p.download(), q.download()))


def test_noalloc_refuses_sets_beyond_the_scratch(ctx):
    beam = Beam(W.ENERGY)
    tab = ctx.table(compile_table(W.fodo_lattice(), beam))
    big = HIPParticles(L.TP_NOALLOC_MAX_N + 1, ctx)
    rc = ctx._lib.tp_track_noalloc(ctx.handle, tab.handle, big.handle, 0, 1, 0)
    assert rc == -1
    assert b"static scratch maximum" in ctx._lib.tp_last_error()
    with pytest.raises(TrackPadHIPError, match="TP_EINVAL"):
        L.check(ctx._lib.tp_track_noalloc(None, None, None, 0, 1, 0))
