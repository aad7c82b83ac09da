"""The C ABI's host code (table parser, argument checks, error paths) and the CPU oracle under
AddressSanitizer and UndefinedBehaviorSanitizer (no GPU).

tests/sanitize/Makefile builds both with the sanitizers on the host pass only; this test runs
the ABI and oracle tests of this directory in a child Python with clang's shared ASan runtime
preloaded and the sanitized libraries in place of the ordinary ones (TRACKPAD_HIP_LIB,
TRACKPAD_ORACLE_LIB).  The mutation and truncation fuzzing of test_abi_fuzz then runs the
parser on every rejected and accepted blob with each heap read checked, and the oracle's
element maps, Jacobians and derivative passes run under the same checks.  A sanitizer report
aborts the child: the test fails with its output."""

import glob
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = os.path.join(ROOT, "tests", "sanitize")
OUT = os.path.join(SAN, "_build")
KOBJ = os.path.join(ROOT, "trackpad.jl_amd", "_build")

# tests whose native calls are host code only: the ABI without a device, and the oracle
SUITES = ["test_abi_cpu.py", "test_abi_fuzz.py", "test_oracle_pinning.py", "test_oracle_closed_forms.py",
          "test_oracle_elements_extra.py", "test_oracle_wiggler.py", "test_table_writer.py", "test_madx.py"]


def _asan_runtime():
    hits = sorted(glob.glob("/opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so"))
    return hits[-1] if hits else None


@pytest.fixture(scope="module")
def built():
    rt = _asan_runtime()
    if rt is None or shutil.which("make") is None or not os.path.exists("/opt/rocm/bin/hipcc"):
        pytest.skip("clang's ASan runtime or hipcc is not installed")
    if not all(os.path.exists(os.path.join(KOBJ, k)) for k in
               ("tp_track_exact.o", "tp_track_fast.o", "tp_adjoint_exact.o", "tp_adjoint_fast.o")):
        pytest.skip("the library's kernel objects are not built (make -C trackpad.jl_amd/csrc)")
    r = subprocess.run(["make", "-s", "-C", SAN, "-j4"], capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    return rt


def test_abi_and_oracle_under_asan_ubsan(built):
    env = dict(os.environ)
    env.update({
        "LD_PRELOAD": built,
        # Python's own allocations are not freed at exit; leaks are not what this run looks for
        "ASAN_OPTIONS": "detect_leaks=0:abort_on_error=1:halt_on_error=1:detect_stack_use_after_return=1",
        "UBSAN_OPTIONS": "print_stacktrace=1:halt_on_error=1",
        "TRACKPAD_HIP_LIB": os.path.join(OUT, "libtrackpad_hip_san.so"),
        "TRACKPAD_ORACLE_LIB": os.path.join(OUT, "liboracle_san.so"),
        "OMP_NUM_THREADS": "4",
        # every Python allocation through malloc, so that ctypes buffers carry ASan's redzones
        "PYTHONMALLOC": "malloc",
    })
    cmd = [sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider", "-p", "no:faulthandler", "-m", "not gpu",
           *[os.path.join(ROOT, "tests", s) for s in SUITES], "-k", "not test_abi_and_oracle_under_asan_ubsan"]
    r = subprocess.run(cmd, cwd=ROOT, env=env, capture_output=True, text=True, timeout=1200)
    out = r.stdout[-4000:] + r.stderr[-6000:]
    assert r.returncode == 0, out
    assert "ERROR: AddressSanitizer" not in out and "runtime error:" not in out, out
    assert " passed" in r.stdout, out


def test_the_sanitized_libraries_are_the_ones_loaded(built):
    """The child's loaders resolve to the sanitized builds, and those carry the ASan and UBSan
    instrumentation (their checks call into the preloaded runtime)."""
    code = ("import os, sys; sys.path[:0] = [os.path.join(%r, 'trackpad.jl_amd'), %r];"
            "from trackpad_hip import _lib as L; from oracle import pyoracle as O;"
            "L.lib(); O.lib(); print(L.LIB_PATH); print(O.LIB)" % (ROOT, ROOT))
    env = dict(os.environ, LD_PRELOAD=built, ASAN_OPTIONS="detect_leaks=0",
               TRACKPAD_HIP_LIB=os.path.join(OUT, "libtrackpad_hip_san.so"),
               TRACKPAD_ORACLE_LIB=os.path.join(OUT, "liboracle_san.so"))
    r = subprocess.run([sys.executable, "-c", code], cwd=ROOT, env=env, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-3000:]
    lines = r.stdout.split()
    assert lines[0].endswith("libtrackpad_hip_san.so") and lines[1].endswith("liboracle_san.so")
    for lib in ("libtrackpad_hip_san.so", "liboracle_san.so"):
        syms = subprocess.run(["nm", "-D", os.path.join(OUT, lib)], capture_output=True, text=True).stdout
        assert "__asan_report_load8" in syms and "__ubsan_handle" in syms, lib
