"""The library's host-side bookkeeping under AddressSanitizer, on the CPU: every HIP source compiled for the host only
(no device code) and linked against a HIP runtime stub (tests/hoststub/hip_stub.cpp) in which device memory is host
memory, kernels do not run, and streams, events, graphs and executables are heap objects with lifetime checks. The
driver (tests/hoststub/lifetime_driver.cpp) runs a level through more distinct pass counts than its graph cache
holds with no host synchronisation between runs, timing classes, layouts, the autotuner, the controller and a batch
of three levels, then destroys everything. It passes when ASan reports nothing and the stub finds no executable graph
destroyed while a launch of it may still run (the graph-cache eviction the engine synchronises for)."""
import os
import shutil
import subprocess
from concurrent.futures import ThreadPoolExecutor

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "autoplace_amd", "csrc")
INCLUDE = os.path.join(HERE, "..", "include")
STUB = os.path.join(HERE, "hoststub")
SOURCES = ["gp_kernels", "gp_wa", "gp_poisson", "gp_fft", "gp_engine", "gp_mlp", "gp_cluster", "gp_prep",
           "gp_devmem", "gp_renumber"]


def _hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    return None


def test_engine_bookkeeping_under_asan(tmp_path):
    hipcc = _hipcc()
    if hipcc is None or shutil.which("g++") is None or shutil.which("nm") is None:
        pytest.skip("hipcc, g++ or nm not found")
    flags = ["-fPIC", "-O1", "-g", "-std=c++17", "-x", "hip", "--cuda-host-only", "-Xarch_host", "-fsanitize=address",
             "-Xarch_host", "-fno-omit-frame-pointer", "-I", CSRC, "-I", INCLUDE]
    jobs = [(os.path.join(CSRC, s + ".hip"), str(tmp_path / (s + ".o"))) for s in SOURCES]
    jobs += [(os.path.join(STUB, "hip_stub.cpp"), str(tmp_path / "hip_stub.o")),
             (os.path.join(STUB, "lifetime_driver.cpp"), str(tmp_path / "driver.o"))]

    def compile_one(job):
        r = subprocess.run([hipcc] + flags + ["-c", job[0], "-o", job[1]], capture_output=True, text=True)
        assert r.returncode == 0, "%s:\n%s" % (job[0], r.stderr[-4000:])
        return job[1]

    with ThreadPoolExecutor(max_workers=4) as ex:
        objs = list(ex.map(compile_one, jobs))
    # the device-code symbols a host-only object refers to: dummies (no device code is ever loaded)
    r = subprocess.run(["nm", "-u"] + objs, capture_output=True, text=True, check=True)
    fatbins = sorted({w for w in r.stdout.split() if w.startswith("__hip_fatbin")})
    (tmp_path / "fatbins.cpp").write_text("".join('extern "C" { unsigned char %s[16]; }\n' % f for f in fatbins))
    subprocess.run(["g++", "-fPIC", "-c", str(tmp_path / "fatbins.cpp"), "-o", str(tmp_path / "fatbins.o")],
                   check=True)
    exe = str(tmp_path / "driver")
    r = subprocess.run(["g++", "-fsanitize=address", "-o", exe] + objs + [str(tmp_path / "fatbins.o"), "-lpthread"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300, env=env, cwd=str(tmp_path))
    assert r.returncode == 0 and "lifetime driver: ok" in r.stdout, (r.stdout[-2000:] + r.stderr[-6000:])
