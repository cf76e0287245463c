"""The library's kernels run on the CPU under sanitizers (tests/hostemu: the
HIP sources compiled as C++ against a host stand-in for the HIP runtime,
workgroups as fibers that meet at every barrier and lane exchange).

The adjoint's kernels (cfd2d_adjoint, cfd2d_adjoint_geometry: LAX, HLLC and
exact-Godunov flux VJPs, Green-Gauss and least-squares gradient reverses,
checkpoint recomputation) run under AddressSanitizer and
UndefinedBehaviorSanitizer with array-bounds checks on the kernels' private
arrays, in four builds:
  locals initialised to a byte pattern, or to zero      (a read of a local
      before it is written gives different results)
  the primitive and gradient adjoints Pb [4][nc], Gb [8][nc] as component
      arrays (the library's layout), or as per-cell records [nc][4], [nc][8]
Every build must finish without a sanitizer report and all four must give
bitwise the same outputs.  The emulated kernels are also checked against
the tangent-linear model (1e-9 relative), so the agreement is of correct
results.
"""
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
EMU = os.path.join(HERE, "hostemu")
CLANG = "/opt/rocm/llvm/bin/clang++"
BUILDS = ["pattern", "zero", "records-pattern", "records-zero"]

pytestmark = pytest.mark.slow


def _asan_runtime():
    if not os.path.exists(CLANG):
        return None
    p = subprocess.run([CLANG, "-print-file-name=libclang_rt.asan-x86_64.so"], capture_output=True, text=True)
    path = p.stdout.strip()
    return path if os.path.isfile(path) else None


@pytest.fixture(scope="module")
def outputs(tmp_path_factory):
    asan = _asan_runtime()
    if asan is None or shutil.which("make") is None:
        pytest.skip("clang++ with the AddressSanitizer runtime is not installed")
    jobs = [subprocess.Popen(["make", "-s", "-j2", "-C", EMU] + extra, stdout=subprocess.PIPE,
                             stderr=subprocess.STDOUT, text=True) for extra in ([], ["LAYOUT=records"])]
    for j in jobs:
        out, _ = j.communicate()
        assert j.returncode == 0, out[-4000:]
    tmp = tmp_path_factory.mktemp("hostemu")
    env = dict(os.environ, LD_PRELOAD=asan, ASAN_OPTIONS="detect_leaks=0:halt_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    runs = {}
    for b in BUILDS:
        e = dict(env, CFD2D_LIB=os.path.join(EMU, "_build", b, "libcfd2d.so"))
        runs[b] = subprocess.Popen([sys.executable, os.path.join(EMU, "run_adjoint.py"), str(tmp / (b + ".npz"))],
                                   env=e, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    res = {}
    for b, p in runs.items():
        out, _ = p.communicate(timeout=900)
        assert p.returncode == 0, (b, out[-6000:])
        assert "runtime error" not in out and "ERROR: AddressSanitizer" not in out, (b, out[-6000:])
        res[b] = dict(np.load(tmp / (b + ".npz")))
    return res


def test_sanitized_builds_agree_bitwise(outputs):
    ref = outputs["zero"]
    assert len(ref) >= 30
    for b in BUILDS[1:]:
        diff = [k for k in ref if not np.array_equal(ref[k], outputs[b][k])]
        assert not diff, (b, diff)
    assert all(np.all(np.isfinite(v)) for v in ref.values())


@pytest.mark.parametrize("name", ["bump_HLLC_LSQ_3", "bump_GODUNOV_GREEN_GAUSS_2"])
def test_emulated_adjoint_matches_tangent(outputs, name):
    import oracle
    import tangent
    sys.path.insert(0, EMU)
    import run_adjoint as ra
    kind, flux, gradient, rk = [x for x in ra.CASES if "%s_%s_%s_%d" % x == name][0]
    c = ra.make(kind, flux, gradient, rk)
    s = oracle.OracleSolver(c)
    o = outputs["zero"]
    tau = float(o[name + "/tau"])
    eg = s.geometry["edge_group"]
    mask = ((eg == c.mesh.group_index("bump")) | (eg == c.mesh.group_index("wall_bottom"))).astype(np.int32)
    rng = np.random.default_rng(3)
    U0 = c.initial_state()
    v = rng.standard_normal(U0.shape) * np.abs(U0) * 1e-2
    m = tangent.Model(s)
    w = rng.standard_normal(m.bcpar0.shape) * np.abs(m.bcpar0) * 1e-2
    Jm, dJ = tangent.tangent(s, U0, v, w, ra.NSTEPS, tau, mask)
    assert abs(float(o[name + "/J"]) - Jm) <= 1e-12 * abs(Jm)
    ad = float(np.sum(o[name + "/gU"] * v)) + float(np.sum(o[name + "/gb"] * w))
    assert abs(ad - dJ) <= 1e-9 * abs(dJ), (ad, dJ)
    dxy = rng.standard_normal(c.mesh.xy.shape) * 1e-2
    _, dJx = tangent.tangent_geometry(s, c.mesh.xy, dxy, ra.NSTEPS, tau, mask)
    adx = float(np.sum(o[name + "/geo_gx"] * dxy))
    assert abs(adx - dJx) <= 1e-9 * abs(dJx), (adx, dJx)
