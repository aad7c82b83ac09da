"""Device parity: libtrackpad_hip (through its C ABI) against the CPU oracle.

Tolerances (written where they are used):
  * parity build, maps with only + - * / sqrt on coordinates: bit-exact (array_equal);
  * parity build with per-particle transcendentals (exact bends, RF, tan edges): within
    the reference tests' own atol = 1e-15 for single passes of the reference's 10 x 6 input
    (test/verify_elements.jl:6), and, for rings, |gpu - oracle| <= rtol * max(|oracle|,
    rms of that coordinate plane) with rtol = 1e-12 after 1 turn and 1e-9 after 100 turns;
  * fast build: it rounds differently from the reference, and the reference formulas are
    ill-conditioned (z += NormL (beti + delta) - L beti and Forest's exact-body x cancel
    O(1) terms).  Here it is checked for accuracy parity: measured against the same
    algorithm run in long double (oracle ``extended=True``), its per-plane error
    max|fast - ld| / rms(ld) must not exceed 3x the reference path's own error (plus 1e-14
    for planes the reference gets exactly); tests/test_gpu_fast_tolerance.py holds it to the
    stated tolerances plane by plane against the oracle itself.  Single-element passes are
    held to the per-coordinate conditioning bounds that pin the oracle against 50-digit
    values (tests/test_oracle_pinning.py);
  * loss flags, lost turn and lost element: identical (the test inputs stay away from the
    aperture boundary; DA-grid particles within a boundary band are excluded).
"""

import math

import numpy as np
import pytest

from conftest import ENERGY_VAL, PARITY_ATOL, PARTICLES_INITIAL
from test_oracle_pinning import TOL_EXACT_BEND, TOL_PLAIN
import lattices as LT
from oracle import pyoracle as O
from trackpad_hip import (Beam, Drift, HIPParticles, Lattice, Quadrupole, compile_table, track_)
from trackpad_hip import workloads as W

pytestmark = pytest.mark.gpu


def gpu_track(ctx, lat, beam, coords, nturns, *, fast=False, aperture=None, lost_flags=None, turn=0,
              block=0, ppl=0, tpl=0):
    ctx.set_launch(block, ppl, tpl)
    try:
        tab = ctx.table(compile_table(lat, beam, aperture=aperture))
        p = HIPParticles.from_coords(coords, lost_flags, ctx)
        track_(p, tab, nturns=nturns, turn=turn, fast=fast)
        c, lf, lt, le = p.download()
        return np.ascontiguousarray(c), lf, lt, le
    finally:
        ctx.set_launch(0, 0, 0)


def ring_close(a, b, rtol):
    """max over entries of |a - b| / max(|b|, rms of b's coordinate plane)."""
    rms = np.sqrt(np.mean(b * b, axis=0, keepdims=True))
    scale = np.maximum(np.abs(b), rms)
    scale = np.where(scale > 0, scale, 1.0)
    err = np.max(np.abs(a - b) / scale)
    return err <= rtol, err


def accuracy_ratio_ok(fast, orc, ld, factor=3.0, floor=1e-14):
    """Per coordinate plane: max|fast - ld| <= factor * max|orc - ld| + floor * rms(ld)."""
    rms = np.sqrt(np.mean(ld * ld, axis=0))
    rms = np.where(rms > 0, rms, 1.0)
    ef = np.max(np.abs(fast - ld), axis=0) / rms
    eo = np.max(np.abs(orc - ld), axis=0) / rms
    return bool(np.all(ef <= factor * eo + floor)), ef, eo


def p10():
    return np.array(PARTICLES_INITIAL, dtype=np.float64)


# --------------------------------------------------------------------------------------
# Single passes of the reference's own test inputs
# --------------------------------------------------------------------------------------
@pytest.mark.parametrize("case", [c[0] for c in LT.element_cases()])
def test_element_single_pass(ctx, case):
    """Each hot-path element of test/verify_elements.jl on the reference's 10 x 6 input."""
    elem = dict(LT.element_cases())[case]
    beam = Beam(ENERGY_VAL)
    lat = Lattice([elem])
    oc, olf, _, _ = O.linepass(p10(), lat, beam, np.zeros(10, np.int64))
    gc, glf, _, _ = gpu_track(ctx, lat, beam, p10(), 1)
    assert np.array_equal(glf, olf)
    if case in LT.BITEXACT:
        assert np.array_equal(gc, oc), np.max(np.abs(gc - oc))
    else:
        assert np.max(np.abs(gc - oc)) < PARITY_ATOL
    fc, flf, _, _ = gpu_track(ctx, lat, beam, p10(), 1, fast=True)
    assert np.array_equal(flf, olf)
    tol = TOL_EXACT_BEND if case in ("ExactSBend", "ERBend") else TOL_PLAIN
    assert np.all(np.max(np.abs(fc - oc), axis=0) <= tol)


@pytest.mark.parametrize("name", ["fodo", "dba", "bend", "exact_bend"])
def test_reference_lines(ctx, name):
    """The lines of test/verify_{fodo,dba,bend,exact_bend}.jl, one pass (atol 1e-15)."""
    lat = {"fodo": LT.verify_fodo_line, "dba": LT.verify_dba_line, "bend": LT.verify_bend_line,
           "exact_bend": LT.verify_exact_bend_line}[name]()
    beam = Beam(ENERGY_VAL)
    oc, olf, _, _ = O.linepass(p10(), lat, beam, np.zeros(10, np.int64))
    gc, glf, _, _ = gpu_track(ctx, lat, beam, p10(), 1)
    assert np.array_equal(glf, olf)
    if name != "exact_bend":
        assert np.array_equal(gc, oc)
    assert np.max(np.abs(gc - oc)) < PARITY_ATOL
    fc, _, _, _ = gpu_track(ctx, lat, beam, p10(), 1, fast=True)
    tol = TOL_EXACT_BEND if name == "exact_bend" else TOL_PLAIN * len(lat)
    assert np.all(np.max(np.abs(fc - oc), axis=0) <= tol)


# --------------------------------------------------------------------------------------
# Multi-turn rings
# --------------------------------------------------------------------------------------
def test_fodo_bitexact_multiturn(ctx):
    """FODO workload (n=4 Yoshida, R = 2 cm aperture), 3,000 particles x 100 turns: the
    parity build reproduces the oracle bit for bit, coordinates and loss record."""
    beam = Beam(W.ENERGY)
    lat = W.fodo_lattice()
    c0 = W.fodo_particles(3000, seed=5)
    c0[:20, 0] = np.linspace(0.0195, 0.0205, 20)  # straddle the aperture at the start
    oc, olf, olt, ole = O.ringpass(c0, lat, beam, np.zeros(3000, np.int64), 100, aperture=0.02, nthreads=8)
    gc, glf, glt, gle = gpu_track(ctx, lat, beam, c0, 100, aperture=0.02)
    assert np.array_equal(glf, olf) and np.array_equal(glt, olt) and np.array_equal(gle, ole)
    assert np.array_equal(gc, oc), np.max(np.abs(gc - oc))
    assert olf.sum() > 0  # the aperture is exercised
    fc, flf, _, _ = gpu_track(ctx, lat, beam, c0, 100, aperture=0.02, fast=True)
    assert np.array_equal(flf, olf)
    alive = olf == 0
    ld = O.ringpass(c0, lat, beam, np.zeros(3000, np.int64), 100, aperture=0.02, nthreads=8, extended=True)[0]
    ok, ef, eo = accuracy_ratio_ok(fc[alive], oc[alive], ld[alive])
    assert ok, (ef, eo)


@pytest.mark.parametrize("nturns,rtol", [(1, 1e-12), (100, 1e-9)])
def test_dba_ring(ctx, nturns, rtol):
    """DBA workload ring (exact bends with linear edges, sextupoles, RF), 512 particles:
    parity build within rtol of the oracle; fast build at accuracy parity with it."""
    beam = Beam(W.ENERGY)
    lat = W.dba_ring(beam=beam)
    c0 = W.dba_particles(512, seed=3)
    oc, olf, olt, ole = O.ringpass(c0, lat, beam, np.zeros(512, np.int64), nturns, nthreads=8)
    gc, glf, glt, gle = gpu_track(ctx, lat, beam, c0, nturns)
    assert np.array_equal(glf, olf) and np.array_equal(glt, olt) and np.array_equal(gle, ole)
    ok, err = ring_close(gc, oc, rtol)
    assert ok, err
    fc, flf, _, _ = gpu_track(ctx, lat, beam, c0, nturns, fast=True)
    assert np.array_equal(flf, olf)
    ld = O.ringpass(c0, lat, beam, np.zeros(512, np.int64), nturns, nthreads=8, extended=True)[0]
    ok, ef, eo = accuracy_ratio_ok(fc, oc, ld)
    assert ok, (ef, eo)


def test_forest_exact_bend_ring(ctx):
    """Reference ExactSBend (yrot, Forest fringe, wedge) with e1, e2 != 0 in a ring of
    quads: 5 turns, tolerance 1e-12."""
    beam = Beam(W.ENERGY)
    from trackpad_hip import ExactSBend
    cell = [Drift(1.0), Quadrupole(0.4, 1.2), Drift(0.5), ExactSBend(1.0, 0.1, 0.03, 0.02, gk=0.01),
            Drift(0.5), Quadrupole(0.4, -1.2), Drift(1.0)]
    lat = Lattice(cell * 4)
    c0 = W.gaussian_particles(256, 9, 1e-4, 1e-5, 1e-4, 1e-5, 1e-3)
    oc, olf, _, _ = O.ringpass(c0, lat, beam, np.zeros(256, np.int64), 5)
    gc, glf, _, _ = gpu_track(ctx, lat, beam, c0, 5)
    assert np.array_equal(glf, olf)
    ok, err = ring_close(gc, oc, 1e-12)
    assert ok, err


def test_exact_bend_large_angles(ctx):
    """Exact bends at large sub-step angles and near-grazing momenta, where the fast build's
    single-arcsine form of the body does not apply and it takes the reference's two clamped
    arcsines: one pass, both builds, loss records identical with the oracle, coordinates
    within 1e-12 (absolute; the body's x is a difference of O(1) terms over irho)."""
    from trackpad_hip import ExactSBend, ExactSBendLinearEdge
    beam = Beam(W.ENERGY)
    c = np.zeros((8, 6))
    c[0] = [1e-3, 0.3, 0.0, 0.1, 0.0, 0.0]
    c[1] = [-2e-3, -0.6, 1e-3, 0.0, 0.0, 0.01]
    c[2] = [0.0, 0.85, 0.0, 0.0, 0.0, -0.02]
    c[3] = [0.0, 0.999, 0.0, 0.0, 0.0, 0.0]     # pz + pzn above d2/16 at small body angles
    c[4] = [0.0, 0.9999, 0.0, 0.0, 0.0, 0.0]    # below it: the two-arcsine form
    c[5] = [0.0, -0.9999, 0.0, 0.01, 0.0, 0.0]
    c[6] = [1e-4, 2e-4, -1e-4, 1e-4, 0.0, 1e-3]
    c[7] = [0.0, 0.5, 0.0, 0.5, 0.0, 0.0]
    for lat in (Lattice([ExactSBendLinearEdge(0.5, 0.8, 0.1, 0.1, num_int_steps=1)]),
                Lattice([ExactSBendLinearEdge(2.0, 0.02, num_int_steps=10)]),
                Lattice([ExactSBend(1.0, 0.6, 0.2, 0.1)])):
        oc, olf, olt, ole = O.ringpass(c, lat, beam, np.zeros(len(c), np.int64), 1)
        for fast in (False, True):
            gc, glf, glt, gle = gpu_track(ctx, lat, beam, c, 1, fast=fast)
            assert np.array_equal(glf, olf) and np.array_equal(gle, ole), (lat.elements[0], fast)
            alive = olf == 0
            assert np.max(np.abs(gc - oc)[alive]) <= 1e-12, (lat.elements[0], fast, np.abs(gc - oc).max(axis=0))
            np.testing.assert_array_equal(gc[~alive], oc[~alive])


def test_persistent_launch_large_beam(ctx):
    """A beam with more waves than the GPU holds at once runs as a persistent grid whose waves
    take 64-particle tiles from a queue: 400,000 particles of the FODO workload (with 2 % of
    them started outside the aperture), 3 turns in one launch and in three (queue plus
    survivor compaction), parity build bit for bit with the oracle, both PPL."""
    beam = Beam(W.ENERGY)
    lat = W.fodo_lattice()
    n = 400_000
    c0 = W.fodo_particles(n, seed=21)
    c0[::50, 0] = 0.021
    oc, olf, olt, ole = O.ringpass(c0, lat, beam, np.zeros(n, np.int64), 3, aperture=0.02, nthreads=8)
    assert olf.sum() >= n // 50
    for block, ppl, tpl in ((256, 1, 0), (64, 2, 1), (128, 1, 1)):
        gc, glf, glt, gle = gpu_track(ctx, lat, beam, c0, 3, aperture=0.02, block=block, ppl=ppl, tpl=tpl)
        assert np.array_equal(glf, olf) and np.array_equal(glt, olt) and np.array_equal(gle, ole), (block, ppl, tpl)
        assert np.array_equal(gc, oc), (block, ppl, tpl)


def test_survivor_compaction_between_launches(ctx):
    """Tracking in several launches compacts the survivors into a particle list between
    launches (lost particles stop occupying lanes); lanes map to particles one to one, so
    coordinates and loss records are the same bits as in one launch.  DA ring on a wide
    24 x 24 x 3 grid: most particles are lost within the 12 turns."""
    beam = Beam(W.ENERGY)
    lat = W.da_ring(ncell=80, beam=beam)
    grid = W.da_grid(24, 24, half=0.018)
    for fast in (False, True):
        one = gpu_track(ctx, lat, beam, grid, 12, fast=fast)
        for tpl, ppl in ((1, 1), (5, 2), (3, 1)):
            many = gpu_track(ctx, lat, beam, grid, 12, fast=fast, tpl=tpl, ppl=ppl)
            for a, b in zip(one, many):
                assert np.array_equal(a, b), (fast, tpl, ppl)
        lost = one[1].astype(bool)
        assert 0.3 < lost.mean() < 0.95, lost.mean()
        assert (one[2][lost] >= 0).all() and (one[2][lost] < 12).all()
    # the grid loses particles fast enough that launches cover quarter turns (more launches
    # than turns); the records above are the same bits all the same
    from trackpad_hip import HIPParticles, compile_table, track_
    ctx.set_launch(0, 0, 3)
    try:
        p = HIPParticles.from_coords(grid, None, ctx)
        track_(p, ctx.table(compile_table(lat, beam)), nturns=12, timed=True)
        assert ctx.timing()["launches"] > 12
        for a, b in zip(gpu_track(ctx, lat, beam, grid, 12), p.download()):
            assert np.array_equal(a, np.ascontiguousarray(b))
    finally:
        ctx.set_launch(0, 0, 0)
    # all particles lost early: the remaining launches are skipped, records unchanged
    wide = W.da_grid(4, 4, half=0.5)
    one = gpu_track(ctx, lat, beam, wide, 6)
    many = gpu_track(ctx, lat, beam, wide, 6, tpl=1)
    assert one[1].all() and all(np.array_equal(a, b) for a, b in zip(one, many))
    # a scan continued in chunks: each call compacts the particles lost in earlier calls out
    # before its first launch; the records and coordinates equal the one-call run's
    tab = ctx.table(compile_table(lat, beam))
    for fast in (False, True):
        one = gpu_track(ctx, lat, beam, grid, 12, fast=fast)
        ctx.set_launch(0, 0, 2)
        try:
            p = HIPParticles.from_coords(grid, None, ctx)
            for t0 in (0, 4, 8):
                track_(p, tab, nturns=4, turn=t0, fast=fast)
            c, lf, lt, le = p.download()
        finally:
            ctx.set_launch(0, 0, 0)
        for a, b in zip(one, (np.ascontiguousarray(c), lf, lt, le)):
            assert np.array_equal(a, b), fast
    # every particle already lost: a call launches nothing and changes nothing
    p = HIPParticles.from_coords(wide, None, ctx)
    track_(p, tab, nturns=6)
    before = p.download()
    ctx.set_launch(0, 0, 1)
    try:
        track_(p, tab, nturns=3, turn=6)
    finally:
        ctx.set_launch(0, 0, 0)
    assert all(np.array_equal(a, b) for a, b in zip(before, p.download()))


def test_da_grid_orders_to_10(ctx):
    """DA ring (order-10 normal/skew errors on every magnet) on a 16 x 16 x 3 grid, 10
    turns: loss records identical away from the boundary, survivors within 1e-9."""
    beam = Beam(W.ENERGY)
    lat = W.da_ring(ncell=80, beam=beam)
    grid = W.da_grid(16, 16)
    n = grid.shape[0]
    oc, olf, olt, ole = O.ringpass(grid, lat, beam, np.zeros(n, np.int64), 10, nthreads=8)
    gc, glf, glt, gle = gpu_track(ctx, lat, beam, grid, 10)
    # boundary band: grid points with a neighbour of different fate
    lost = olf.reshape(3, 16, 16).astype(bool)
    band = np.zeros_like(lost)
    for ax in (1, 2):
        for sh in (1, -1):
            band |= lost != np.roll(lost, sh, axis=ax)
    keep = ~band.reshape(-1)
    assert np.array_equal(glf[keep], olf[keep])
    assert np.array_equal(glt[keep], olt[keep]) and np.array_equal(gle[keep], ole[keep])
    # coordinates on regular orbits: initial |x|, |y| <= 5 mm (larger amplitudes of this
    # order-10 lattice are chaotic, where any rounding difference grows exponentially)
    regular = keep & (olf == 0) & (np.abs(grid[:, 0]) <= 0.005) & (np.abs(grid[:, 2]) <= 0.005)
    assert regular.sum() > 20
    ok, err = ring_close(gc[regular], oc[regular], 1e-9)
    assert ok, err


# --------------------------------------------------------------------------------------
# Loss semantics and edge cases
# --------------------------------------------------------------------------------------
def test_loss_semantics(ctx):
    """src/lattice.jl:269-284: flag-1 inputs are skipped; a particle lost at an element
    exit keeps its element-entry coordinates; NaN / pz^2 <= 0 / |x| > 1 all lose."""
    beam = Beam(W.ENERGY)
    lat = Lattice([Drift(1.0), Quadrupole(0.5, 1.0), Drift(1.0)])
    c = np.zeros((6, 6))
    c[0] = [0.999, 0.01, 0, 0, 0, 0]      # |x| > 1 after the first drift
    c[1] = [0.0, 0.9, 0, 0.9, 0, 0]       # pz^2 < 0 -> NaN -> lost at the first drift
    c[2] = [np.nan, 0, 0, 0, 0, 0]        # NaN input
    c[3] = [1e-3, 1e-4, 0, 0, 0, 0]       # survives
    c[4] = [5.0, 0, 0, 0, 0, 0]           # arrives flagged: skipped, untouched
    c[5] = [0.0, 0.0, 0.5, 0.6, 0, 0]     # |y| > 1 inside the quad-drift region
    lf = np.array([0, 0, 0, 0, 1, 0], np.int64)
    oc, olf, olt, ole = O.ringpass(c, lat, beam, lf, 2)
    gc, glf, glt, gle = gpu_track(ctx, lat, beam, c, 2, lost_flags=lf)
    assert list(olf) == [1, 1, 1, 0, 1, 1]
    assert np.array_equal(glf, olf) and np.array_equal(glt, olt) and np.array_equal(gle, ole)
    assert list(glt) == [0, 0, 0, -1, -2, 0]
    np.testing.assert_array_equal(gc[[0, 1, 3, 4, 5]], oc[[0, 1, 3, 4, 5]])
    assert np.isnan(gc[2, 0]) and np.array_equal(gc[2, 1:], oc[2, 1:])
    assert np.array_equal(gc[0], c[0]) and np.array_equal(gc[4], c[4])
    # the fast build: the same losses (its drift has no pz^2 test: the refined hardware
    # reciprocal square root is NaN for pz^2 <= 0) and the same kept entry coordinates
    fc, flf, flt, fle = gpu_track(ctx, lat, beam, c, 2, lost_flags=lf, fast=True)
    assert np.array_equal(flf, olf) and np.array_equal(flt, olt) and np.array_equal(fle, ole)
    np.testing.assert_array_equal(fc[[0, 1, 4]], oc[[0, 1, 4]])
    for i in (3, 5):
        assert np.max(np.abs(fc[i] - oc[i])) <= 1e-15
    # pz^2 exactly 0 (px = 1) and -1.1e-16 (0.6^2 + 0.8^2 rounds above 1) at the first
    # drift: lost there in both builds with the entry coordinates; px = py = 0.5 passes it
    e = np.zeros((3, 6))
    e[0] = [0.0, 1.0, 0.0, 0.0, 0.0, 0.0]
    e[1] = [0.0, 0.6, 0.0, 0.8, 0.0, 0.0]
    e[2] = [0.0, 0.5, 0.0, 0.5, 0.0, 0.0]
    oe = O.ringpass(e, lat, beam, np.zeros(3, np.int64), 1)
    assert list(oe[1][:2]) == [1, 1] and list(oe[3][:2]) == [0, 0]
    for fast in (False, True):
        ge = gpu_track(ctx, lat, beam, e, 1, fast=fast)
        assert np.array_equal(ge[1], oe[1]) and np.array_equal(ge[2], oe[2]) and np.array_equal(ge[3], oe[3])
        np.testing.assert_array_equal(ge[0][:2], e[:2])


def test_empty_and_ragged(ctx):
    beam = Beam(W.ENERGY)
    lat = W.fodo_lattice()
    # empty particle set
    gc, glf, _, _ = gpu_track(ctx, lat, beam, np.zeros((0, 6)), 3)
    assert gc.shape == (0, 6)
    # empty lattice: identity
    c0 = W.fodo_particles(77, seed=8)
    gc, glf, _, _ = gpu_track(ctx, Lattice([]), beam, c0, 5)
    assert np.array_equal(gc, c0) and glf.sum() == 0
    # ragged sizes, launch shapes and turn chunking all give the same bits
    for n in (1, 63, 65, 257, 1001):
        c0 = W.fodo_particles(n, seed=n)
        ref, rlf, _, _ = O.ringpass(c0, lat, beam, np.zeros(n, np.int64), 7, aperture=0.02)
        for block, ppl, tpl in ((256, 1, 0), (64, 2, 3), (128, 1, 1), (192, 2, 0)):
            gc, glf, _, _ = gpu_track(ctx, lat, beam, c0, 7, aperture=0.02, block=block, ppl=ppl, tpl=tpl)
            assert np.array_equal(gc, ref) and np.array_equal(glf, rlf), (n, block, ppl, tpl)


def test_turn_continuation(ctx):
    """ringpass! resumes from (coords, lost_flags, turn): 2 x 5 turns == 10 turns."""
    beam = Beam(W.ENERGY)
    lat = W.dba_ring(beam=beam)
    c0 = W.dba_particles(200, seed=4)
    a = gpu_track(ctx, lat, beam, c0, 10)
    b1 = gpu_track(ctx, lat, beam, c0, 5)
    b2 = gpu_track(ctx, lat, beam, b1[0], 5, lost_flags=b1[1], turn=5)
    assert np.array_equal(a[0], b2[0])


def test_host_entry_points(ctx):
    """linepass_/ringpass_ mutate host arrays in place like linepass!/ringpass!."""
    from trackpad_hip import linepass_, ringpass_

    beam = Beam(ENERGY_VAL)
    lat = LT.verify_dba_line()
    c = p10()
    lf = np.zeros(10, np.int64)
    linepass_(c, lat, beam, lf, ctx=ctx)
    oc, olf, _, _ = O.linepass(p10(), lat, beam, np.zeros(10, np.int64))
    assert np.array_equal(c, oc) and np.array_equal(lf, olf)
    c = np.asfortranarray(p10())
    lf = np.zeros(10, np.int64)
    ringpass_(c, lat, beam, lf, 3, ctx=ctx)
    oc, _, _, _ = O.ringpass(p10(), lat, beam, np.zeros(10, np.int64), 3)
    assert np.array_equal(c, oc)


@pytest.mark.parametrize("fbe,fbx", [(2, 2), (3, 3), (3, 2), (2, 3)])
def test_edge_methods_soleil_thomx(ctx, fbe, fbx):
    """Linear edge methods 2 (SOLEIL: the vertical lens divided by 1 + delta) and 3 (THOMX: the
    edge angle seen with the slope px / (1 + delta), + at the entrance, - at the exit),
    src/tracking.jl:174-178, 207-213, with fringe (fint gap != 0) on both faces: one pass of
    the reference's 10 x 6 input and of a wider beam (|px| to 5e-3, delta to 2 %), parity build
    within the reference tests' 1e-15 and the fast build within the oracle-pinning bounds; a
    ring of such bends with quads for 1 and 50 turns, parity build within 1e-12 / 1e-9 of the
    oracle."""
    from trackpad_hip import SBend

    beam = Beam(ENERGY_VAL)
    bend = SBend(1.5, 0.157, 0.02, 0.03, num_int_steps=10, fint1=0.5, fint2=0.4, gap=0.03,
                 fringe_bend_entrance=fbe, fringe_bend_exit=fbx)
    lat = Lattice([bend])
    rng = np.random.default_rng(fbe * 10 + fbx)
    wide = np.column_stack([rng.normal(0, 2e-3, 64), rng.normal(0, 5e-3, 64), rng.normal(0, 2e-3, 64),
                            rng.normal(0, 5e-3, 64), rng.normal(0, 1e-3, 64), rng.uniform(-0.02, 0.02, 64)])
    for c0 in (p10(), wide):
        n = len(c0)
        oc, olf, _, _ = O.linepass(c0, lat, beam, np.zeros(n, np.int64))
        gc, glf, _, _ = gpu_track(ctx, lat, beam, c0, 1)
        assert np.array_equal(glf, olf)
        assert np.max(np.abs(gc - oc)) < PARITY_ATOL * max(1.0, np.abs(c0).max() / 1e-3)
        fc, flf, _, _ = gpu_track(ctx, lat, beam, c0, 1, fast=True)
        assert np.array_equal(flf, olf)
        assert np.all(np.max(np.abs(fc - oc), axis=0) <= TOL_PLAIN * 10 * max(1.0, np.abs(c0).max() / 1e-3))
    # the methods differ from Brown's on this input (the cases exercise them)
    brown = Lattice([SBend(1.5, 0.157, 0.02, 0.03, num_int_steps=10, fint1=0.5, fint2=0.4, gap=0.03)])
    ob = O.linepass(wide, brown, beam, np.zeros(64, np.int64))[0]
    om = O.linepass(wide, lat, beam, np.zeros(64, np.int64))[0]
    assert np.max(np.abs(ob[:, 3] - om[:, 3])) > 1e-9
    cell = [Drift(1.0), Quadrupole(0.4, 1.1), Drift(0.5), bend, Drift(0.5), Quadrupole(0.4, -1.0), Drift(1.0)]
    ring = Lattice(cell * 6)
    c0 = W.gaussian_particles(256, 9, 1e-4, 1e-5, 1e-4, 1e-5, 1e-3)
    for nturns, rtol in ((1, 1e-12), (50, 1e-9)):
        oc, olf, olt, ole = O.ringpass(c0, ring, beam, np.zeros(256, np.int64), nturns, nthreads=8)
        gc, glf, glt, gle = gpu_track(ctx, ring, beam, c0, nturns)
        assert np.array_equal(glf, olf) and np.array_equal(glt, olt)
        ok, err = ring_close(gc, oc, rtol)
        assert ok, (nturns, err)
