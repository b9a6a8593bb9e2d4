"""AddressSanitizer + UndefinedBehaviorSanitizer build (host code only) of the
device TDE arithmetic (tests/host_build/tde_host.cpp: stress full and half
space, displacement) and of the plain-C oracle (oracle/tde_oracle.c), run over
every golden point record: no sanitizer report (any report aborts the driver),
and the sanitized device arithmetic still equals the reference C++ bit for
bit on the full-space records."""
import os
import subprocess
from pathlib import Path

import numpy as np
import pytest

REPO = Path(__file__).resolve().parents[1]
CLANG = Path("/opt/rocm/llvm/bin/clang")
CLANGXX = Path("/opt/rocm/llvm/bin/clang++")
HIPCC = Path("/opt/rocm/bin/hipcc")
SAN = ["address", "undefined"]

pytestmark = pytest.mark.skipif(not (CLANG.exists() and CLANGXX.exists() and HIPCC.exists()),
                                reason="ROCm clang / hipcc not installed")


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    d = tmp_path_factory.mktemp("sanitized")
    csrc = REPO / "pyquake3d_amd" / "csrc"
    common = ["-O1", "-g", "-fno-omit-frame-pointer", "-ffp-contract=off"]
    host_san = []
    for s in SAN:
        host_san += ["-Xarch_host", f"-fsanitize={s}"]
    host_san += ["-Xarch_host", "-fno-sanitize-recover=all"]
    steps = [
        [str(CLANG), *common, "-fopenmp", f"-fsanitize={','.join(SAN)}",
         "-fno-sanitize-recover=all", "-c", str(REPO / "oracle" / "tde_oracle.c"),
         "-o", str(d / "oracle.o")],
        [str(HIPCC), "-x", "hip", "--offload-host-only", *common, *host_san, f"-I{csrc}", "-c",
         str(REPO / "tests" / "host_build" / "tde_host.cpp"), "-o", str(d / "tde_host.o")],
        [str(HIPCC), "-x", "hip", "--offload-host-only", *common, *host_san, "-c",
         str(REPO / "tests" / "host_build" / "sanitize_driver.cpp"), "-o", str(d / "drv.o")],
        [str(CLANGXX), "-fopenmp", f"-fsanitize={','.join(SAN)}", str(d / "tde_host.o"),
         str(d / "drv.o"), str(d / "oracle.o"), "-o", str(d / "driver"), "-lm",
         "-Wl,-rpath,/opt/rocm/llvm/lib"],
    ]
    for cmd in steps:
        r = subprocess.run(cmd, capture_output=True, text=True)
        assert r.returncode == 0, r.stderr[-3000:]
    return d


def _run(driver, rec, hs):
    inp, out = driver / f"in{hs}.bin", driver / f"out{hs}.bin"
    with open(inp, "wb") as f:
        np.array([len(rec), hs], dtype=np.int64).tofile(f)
        np.ascontiguousarray(rec, dtype=np.float64).tofile(f)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([str(driver / "driver"), str(inp), str(out)], capture_output=True,
                       text=True, env=env, timeout=600)
    assert r.returncode == 0 and "runtime error" not in r.stderr, r.stderr[-3000:]
    return np.fromfile(out).reshape(len(rec), 15)


def test_full_space_records_clean_and_bitwise(driver, golden):
    g = golden["g1_points"]
    rec = np.concatenate([g["pts"], g["tri"], g["bv"], g["ml"]], axis=1)
    o = _run(driver, rec, 0)
    fin = np.isfinite(g["stress"]).all(1)
    assert np.array_equal(o[fin, :6], g["stress"][fin])
    scale = np.abs(g["stress"][fin]).max(1, keepdims=True)
    assert (np.abs(o[fin, 6:12] - g["stress"][fin]) <= 1e-10 * scale).all()


def test_half_space_records_clean(driver, golden):
    g = golden["g4_halfspace"]
    rec = np.concatenate([g["pts"], g["tri"], g["bv"], g["ml"]], axis=1)
    o = _run(driver, rec, 1)
    fin = np.isfinite(g["stress"]).all(1)
    scale = np.abs(g["stress"][fin]).max(1, keepdims=True)
    assert (np.abs(o[fin, :6] - g["stress"][fin]) <= 1e-10 * scale).all()
