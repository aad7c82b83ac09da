"""Write tests/golden/tables/*.bin: table blobs of the Python writer for a few lattices.

Run from the repository root:  python tests/golden/make_tables.py

The blobs pin the table format: tests/test_table_writer.py checks that the Python writer
still produces them byte for byte, and the Julia writer's test
(trackpad.jl_amd/julia/TrackPadHIP/test/runtests.jl) compares its output with them.
"""

import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "tests"), os.path.join(ROOT, "trackpad.jl_amd")]

import lattices as LT  # noqa: E402
from trackpad_hip import Beam, Lattice, compile_table  # noqa: E402
from trackpad_hip import workloads as W  # noqa: E402


def cases():
    b35 = Beam(3.5e9)
    yield "verify_fodo_line", compile_table(LT.verify_fodo_line(), b35)
    yield "verify_dba_line", compile_table(LT.verify_dba_line(), b35)
    yield "verify_exact_bend_line", compile_table(LT.verify_exact_bend_line(), b35)
    yield "verify_elements", compile_table(Lattice([e for _, e in LT.element_cases()]), b35)
    yield "fodo_workload", compile_table(W.fodo_lattice(), Beam(W.ENERGY), aperture=0.02)
    yield "dba_workload", compile_table(W.dba_ring(), Beam(W.ENERGY))


def main():
    os.makedirs(os.path.join(HERE, "tables"), exist_ok=True)
    for name, blob in cases():
        with open(os.path.join(HERE, "tables", name + ".bin"), "wb") as f:
            f.write(blob)
        print(name, len(blob))


if __name__ == "__main__":
    main()
