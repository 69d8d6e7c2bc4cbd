"""The density maps' fixed point: fixed_point32 (autoplace_amd/csrc/gp_common.h), the integer form the tile kernels
use, equals the reference's (unsigned long long)((double)w * 2^32) (cpp_to_py/gp_cuda_kernel.cu:84-95) for every
float w in [0, 2^31) -- checked exhaustively on the host (1.3e9 values, a few seconds), through the header itself
built by hipcc for the host only."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "autoplace_amd", "csrc")

PROGRAM = r"""
#include "gp_common.h"
#include <cstdio>
int main() {
    unsigned long long bad = 0, first = 0;
    for (unsigned long long b = 0; b < 0x4F000000ull; b++) {           // every float in [0, 2^31)
        const float w = __builtin_bit_cast(float, (unsigned)b);
        if ((unsigned long long)((double)w * ap::kDensityScale) != ap::fixed_point32(w) && !bad++) first = b;
    }
    // zero, negative zero, negative, infinite and NaN inputs give 0
    const float odd[] = {-0.0f, -1.0f, -1e-30f, __builtin_inff(), -__builtin_inff(), __builtin_nanf("")};
    for (float w : odd) bad += ap::fixed_point32(w) != 0ull;
    printf("%llu %llx\n", bad, first);
    return bad != 0;
}
"""


def _hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    return None


def test_fixed_point32_equals_the_double_conversion(tmp_path):
    hipcc = _hipcc()
    if hipcc is None:
        pytest.skip("hipcc not found")
    src, exe = tmp_path / "fp.cpp", tmp_path / "fp"
    src.write_text(PROGRAM)
    r = subprocess.run([hipcc, "-O2", "-std=c++17", "-I", CSRC, "-x", "hip", "--cuda-host-only", str(src), "-o",
                        str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, "mismatches (count, first float bits): " + r.stdout
