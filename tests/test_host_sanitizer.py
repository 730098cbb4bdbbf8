"""The library's host code under AddressSanitizer and UndefinedBehaviorSanitizer.

tests/hostsan builds hsh_api.cpp, hsh_ingest.cpp and hsh_jit.cpp with g++
-fsanitize=address,undefined against a mock HIP runtime (device memory in
host memory; every kernel launch checks the pointers of its arguments against
the live allocations, with the extents its grid uses; captured graphs are
re-checked at every replay) and drives it through the C ABI with
tests/hostsan/scenarios.py: heterogeneous layouts and the autotuner, mixed
tile/pack layouts included (the sanitizer build sets HSH_AUTOTUNE_MIXED),
events, inputs, diffusion shells, calcium, givals, groups, symmetric
compartments, the adjoint and the isopotential kernel.  A scenario passes
when no sanitizer or pointer check fires and every device allocation and
module is released when its handles are destroyed.  No kernel runs: the
numbers are covered by the GPU parity tests.
"""
import os
import shutil
import subprocess
import sys

import pytest

SAN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "hostsan")
ALL = ["hetero", "layouts", "network", "shells", "calcium", "givals", "groups", "symmetric", "adjoint", "iso"]


def _libasan():
    if shutil.which("gcc") is None or shutil.which("g++") is None:
        return None
    p = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    return p if os.path.isabs(p) and os.path.exists(p) else None


@pytest.fixture(scope="module")
def libasan():
    asan = _libasan()
    if asan is None:
        pytest.skip("g++ with AddressSanitizer is not available")
    r = subprocess.run(["make", "-C", SAN, "-j8"], capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    return asan


def run(asan, names, **env):
    pre = os.environ.get("LD_PRELOAD", "")
    e = dict(os.environ, LD_PRELOAD=(asan + " " + pre).strip(), ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
             UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1", HSH_AUTOTUNE_LOG="1", **env)
    r = subprocess.run([sys.executable, os.path.join(SAN, "scenarios.py"), *names], env=e, capture_output=True,
                       text=True, timeout=600)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-5000:]
    for n in names:
        assert f"ok {n} " in out, n
    return out


def test_host_code_clean_with_specialised_kernels(libasan):
    run(libasan, ALL)


def test_host_code_clean_with_interpreter_kernels(libasan):
    """The run-time compiler fails every program: the handles fall back to
    the built-in interpreter kernels."""
    out = run(libasan, ALL, MOCKHIP_RTC_FAIL="1")
    assert "using the interpreter kernels" in out


@pytest.mark.parametrize("seed", [1, 2, 3, 4])
def test_autotuner_over_mixed_layouts(libasan, seed):
    """Candidate timings differ per seed, so the autotuner keeps a different
    layout each time; the sweep includes layouts of tiles and packs at once,
    and the tuned handle keeps stepping on what it kept."""
    out = run(libasan, ["hetero"], MOCKHIP_SEED=str(seed))
    mixed = [ln for ln in out.splitlines() if ln.startswith("hsh_autotune: tile_max_rows")
             and " 0 tiles" not in ln and " 0 packs" not in ln]
    assert mixed, out[-3000:]
