"""The table writer (trackpad_hip.table.compile_table) against committed blobs, and the
table's field semantics (CPU only)."""

import math
import os
import struct
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_tables  # noqa: E402

from trackpad_hip import Beam, Drift, Lattice, Quadrupole, Sextupole, TableInfo, compile_table  # noqa: E402
from trackpad_hip.table import DRIFT1, DRIFT2, KICK1, KICK2  # noqa: E402
from trackpad_hip import workloads as W  # noqa: E402


@pytest.mark.parametrize("name,blob", list(make_tables.cases()))
def test_blob_unchanged(name, blob):
    with open(os.path.join(HERE, "golden", "tables", name + ".bin"), "rb") as f:
        assert f.read() == blob, name


def _parse(blob):
    magic, ver, n_seq, n_rec, pool_len, flags, _, _ = struct.unpack_from("<8I", blob, 0)
    beti, cl, al, ap2 = struct.unpack_from("<4d", blob, 32)
    off = 64
    seq = list(struct.unpack_from("<%di" % n_seq, blob, off))
    off += (n_seq * 4 + 7) // 8 * 8
    recs = [struct.unpack_from("<4iI3i", blob, off + 32 * i) for i in range(n_rec)]
    off += 32 * n_rec
    pool = np.frombuffer(blob, dtype="<f8", count=pool_len, offset=off)
    return dict(magic=magic, ver=ver, seq=seq, recs=recs, pool=pool, beti=beti, cl=cl, al=al, ap2=ap2, flags=flags)


def test_quad_record_constants():
    """Yoshida constants formed as in symplectic4_pass (src/tracking.jl:324-328); b1 <- k1."""
    beam = Beam(3e9)
    t = _parse(compile_table(Lattice([Quadrupole(0.4, 1.3, num_int_steps=10)]), beam))
    assert t["magic"] == 0x54504854 and t["ver"] == 1 and t["flags"] == 1
    tag, n, m, off, fl, moff, g, kick = t["recs"][0]
    assert (tag, n, m, g, moff) == (2, 10, 1, 1, -1)
    SL = 0.4 / 10
    bt = 1.0 / beam.beta
    assert list(t["pool"][off:off + 6]) == [SL * DRIFT1, SL * DRIFT2, SL * KICK1, SL * KICK2, SL * DRIFT1 * bt,
                                           SL * DRIFT2 * bt]
    assert list(t["pool"][off + 6:off + 10]) == [0.0, 0.0, 1.3, 0.0]  # (b0, a0, b1, a1)


def test_dedup_and_info():
    info = TableInfo()
    blob = compile_table(W.dba_ring(), Beam(W.ENERGY), info=info)
    assert info.n_seq == 501 and info.n_rec == 13
    t = _parse(blob)
    assert len(t["seq"]) == 501 and max(t["seq"]) == 12
    assert sum(1 for g in info.gorder if g >= 0) == 200
    assert sorted(set(info.gscale)) == [0.0, 0.5, 1.0]


def test_sextupole_and_aperture():
    t = _parse(compile_table(Lattice([Drift(1.0), Sextupole(0.2, 30.0)]), Beam(3e9), aperture=0.02))
    assert t["ap2"] == 0.02 * 0.02
    tag, n, m, off, fl, moff, g, kick = t["recs"][1]
    assert tag == 2 and m == 2 and g == 2 and t["pool"][off + 6 + 4] == 15.0
