"""The host setup under sanitizers, on the CPU.

tests/cpp/layout_check drives the library's own host code (host_setup.cpp:
the ABMC schedule, the IC(0) factor with auto-shift, the transpose, the
SELL-64 and sweep layouts, the split SGS factors, and a batch of systems
planned side by side) and checks every layout against the host factor.  It
is built here with g++ twice -- AddressSanitizer + UndefinedBehaviorSanitizer
(any report fails the run), and ThreadSanitizer -- and run on systems large
enough for the setup to take its threads (parallel_chunks, the three layout
threads of build_plan, the batch's worker pool)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from sparsesolv_amd import problems as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SOURCES = [os.path.join(ROOT, "tests", "cpp", "layout_check.cpp"),
           os.path.join(ROOT, "ngsolve-sparsesolv_amd", "csrc", "host_setup.cpp")]
FLAGS = ["-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-ffp-contract=off", "-pthread",
         "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "ngsolve-sparsesolv_amd", "csrc")]
SANITIZERS = {
    "asan_ubsan": ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined"],
    "tsan": ["-fsanitize=thread"],
}
# a report ends the run with this status (any other failure has its own)
ENV = {
    "ASAN_OPTIONS": "verify_asan_link_order=0:detect_leaks=1:exitcode=86",
    "UBSAN_OPTIONS": "print_stacktrace=1:halt_on_error=1:exitcode=86",
    "TSAN_OPTIONS": "halt_on_error=1:exitcode=86",
}

pytestmark = pytest.mark.skipif(shutil.which("g++") is None, reason="no host C++ compiler")


@pytest.fixture(scope="module")
def binaries(tmp_path_factory):
    d = tmp_path_factory.mktemp("sanitize")
    out, procs = {}, {}
    for name, flags in SANITIZERS.items():     # both builds side by side
        exe = d / ("layout_check_" + name)
        procs[name] = subprocess.Popen(["g++", *FLAGS, *flags, *SOURCES, "-o", str(exe)],
                                       stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
        out[name] = str(exe)
    for name, p in procs.items():
        _, err = p.communicate(timeout=600)
        assert p.returncode == 0, err[-4000:]
    return out


def write_matrix(path, A):
    with open(path, "wb") as f:
        np.array([A.n], dtype=np.int32).tofile(f)
        A.row_ptr.astype(np.int32).tofile(f)
        A.col_idx.astype(np.int32).tofile(f)
        A.values.astype(np.complex128 if A.is_complex else np.float64).tofile(f)


CASES = [
    # (name, matrix, block, colours, use_abmc)
    ("cc24_threads", lambda: P.curlcurl(24).A, 4, 4, 1),
    ("eddy10_complex", lambda: P.eddy(10).A, 4, 8, 1),
    ("lap16_b16", lambda: P.laplace7(16), 16, 16, 1),
    ("lap20_level", lambda: P.laplace7(20), 4, 4, 0),
]


@pytest.mark.parametrize("sanitizer", list(SANITIZERS))
@pytest.mark.parametrize("name,builder,block,colours,abmc", CASES, ids=[c[0] for c in CASES])
def test_host_setup_under_sanitizer(binaries, tmp_path, sanitizer, name, builder, block, colours, abmc):
    A = builder()
    path = tmp_path / "A.bin"
    write_matrix(path, A)
    env = dict(os.environ, **ENV)
    r = subprocess.run([binaries[sanitizer], str(path), str(int(A.is_complex)), str(block),
                        str(colours), str(abmc)], capture_output=True, text=True, timeout=600, env=env)
    report = r.stderr[-6000:]
    assert r.returncode != 86, f"{sanitizer} report:\n{report}"
    assert r.returncode == 0 and r.stdout.startswith("ok"), r.stdout + report
    assert "runtime error" not in r.stderr and "WARNING: ThreadSanitizer" not in r.stderr, report
