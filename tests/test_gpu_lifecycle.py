"""Handle lifetimes through the C ABI: destroy order and clean process exit (GPU)."""

import os
import subprocess
import sys
import textwrap

import numpy as np
import pytest

from trackpad_hip import _lib as L

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_destroy_context_first():
    """Julia/Python finalizers run in any order: destroying the context before its table
    and particle set must be safe; the last release frees the stream."""
    import ctypes as C

    lib = L.lib()
    from trackpad_hip import Beam, compile_table
    from trackpad_hip import workloads as W

    blob = compile_table(W.fodo_lattice(), Beam(3e9))
    ctx, tab, par = C.c_void_p(), C.c_void_p(), C.c_void_p()
    L.check(lib.tp_ctx_create(0, C.byref(ctx)))
    buf = C.create_string_buffer(blob, len(blob))
    L.check(lib.tp_table_create(ctx, buf, len(blob), C.byref(tab)))
    L.check(lib.tp_particles_create(ctx, 100, C.byref(par)))
    c0 = np.asfortranarray(W.fodo_particles(100, seed=1))
    L.check(lib.tp_particles_upload(par, c0.ctypes.data_as(C.POINTER(C.c_double)), None, None, None))
    L.check(lib.tp_track(ctx, tab, par, 0, 3, 0))
    L.check(lib.tp_ctx_destroy(ctx))
    # no new work on a destroyed context
    assert lib.tp_track(ctx, tab, par, 0, 1, 0) == -1
    L.check(lib.tp_particles_destroy(par))
    L.check(lib.tp_table_destroy(tab))


def test_process_exit_is_clean():
    """A script that tracks and exits with live handles returns 0 (no teardown crash)."""
    code = textwrap.dedent("""
        import sys
        sys.path[:0] = [%r, %r]
        import numpy as np
        from trackpad_hip import Beam, HIPParticles, HIPContext, compile_table, track_
        from trackpad_hip import workloads as W
        ctx = HIPContext(0)
        tab = ctx.table(compile_table(W.fodo_lattice(), Beam(3e9)))
        p = HIPParticles.from_coords(W.fodo_particles(1000), ctx=ctx)
        track_(p, tab, nturns=5)
        q = HIPParticles(1000, ctx)
        q.copy_from(p)
        print("ok", p.download()[1].sum())
    """ % (ROOT, os.path.join(ROOT, "trackpad.jl_amd")))
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, (r.returncode, r.stdout[-2000:], r.stderr[-2000:])
    assert r.stdout.startswith("ok")
