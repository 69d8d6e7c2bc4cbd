"""Build libautoplace_gp.so (the HIP kernels and the C ABI) in-tree for gfx950.

    python -m autoplace_amd.build_lib            # or __graft_entry__.build()

hipcc compiles each csrc/*.hip for --offload-arch=gfx950 (packed-FP32 instructions disabled) and links one shared
library next to this file. No GPU is needed to build.
"""
import os
import shutil
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libautoplace_gp.so")
SOURCES = ["gp_kernels.hip", "gp_wa.hip", "gp_poisson.hip", "gp_fft.hip", "gp_engine.hip", "gp_mlp.hip",
           "gp_cluster.hip", "gp_prep.hip", "gp_devmem.hip",
           "gp_renumber.hip"]
ARCH = os.environ.get("AUTOPLACE_ARCH", "gfx950")
FLAGS = ["-O3", "-fPIC", "-std=c++17", "--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang",
         "-packed-fp32-ops", "-fvisibility=hidden", "-Wall", "-Wno-unused-function"]
# Per-source device flags. gp_wa.hip: its max / min reductions take finite positions (or the +-FLT_MAX of empty
# lanes), so the device code is compiled without NaN semantics: fmaxf / fminf then need no canonicalising
# instruction per operand and the DPP move folds into v_max_f32 / v_min_f32 (about 15 % fewer VALU instructions in
# the WA kernels, the same bits for every finite input; a NaN position still makes its net's WA sum NaN).
SOURCE_FLAGS = {"gp_wa.hip": ["-Xarch_device", "-fno-honor-nans"]}


def _hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found (ROCm is needed to build the HIP library)")


def build(verbose=False):
    """Compile every HIP source and link the library (always rebuilds). -> path of the library."""
    hipcc = _hipcc()
    missing = [s for s in SOURCES if not os.path.exists(os.path.join(CSRC, s))]
    if missing:
        raise RuntimeError("HIP sources missing from %s: %s" % (CSRC, ", ".join(missing)))
    srcs = list(SOURCES)
    with tempfile.TemporaryDirectory() as tmp:
        def compile_one(src):
            obj = os.path.join(tmp, src.replace(".hip", ".o"))
            cmd = [hipcc] + FLAGS + SOURCE_FLAGS.get(src, []) + ["-c", os.path.join(CSRC, src), "-o", obj]
            r = subprocess.run(cmd, capture_output=True, text=True)
            if r.returncode != 0:
                raise RuntimeError("hipcc failed on %s:\n%s" % (src, r.stderr))
            if verbose and r.stderr.strip():
                print(r.stderr, file=sys.stderr)
            return obj

        with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
            objs = list(ex.map(compile_one, srcs))
        out = LIB + ".tmp"
        cmd = [hipcc, "-shared", "-fPIC", "--offload-arch=" + ARCH] + objs + ["-o", out]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("link failed:\n%s" % r.stderr)
        os.replace(out, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
