"""Generated gfx950 code of every source that uses the matrix cores: no MFMA may write a
destination that partly overlaps its accumulator (C) operand.

Such an instruction (for example v_mfma_f32_16x16x4_f32 a[4:7], ..., a[6:9]) reads the C
registers over several passes while writing D, and gave wrong sums on gfx950 (registers 1 and
2 of one output tile of the whitening product). The compiler emitted it when a kernel ran short
of registers; the kernels are written to avoid that, and this test checks the assembly so a
later change cannot bring it back unnoticed. Runs on the CPU (hipcc cross-compiles).
"""
import os
import re
import shutil
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "pibronic_amd", "csrc")
SOURCES = ["pib_gemm.hip", "pib_flow.hip", "pib_flow_fused.hip", "pib_eval_wide.hip"]
MFMA = re.compile(r"\s+(v_mfma\S+)\s+([av])\[(\d+):(\d+)\],\s*[^,]+,\s*[^,]+,\s*([av])\[(\d+):(\d+)\]")


def partial_overlaps(asm_text):
    bad, total = [], 0
    for line in asm_text.splitlines():
        m = MFMA.match(line)
        if not m:
            continue
        total += 1
        fd, d0, d1, fc, c0, c1 = m.group(2), int(m.group(3)), int(m.group(4)), m.group(5), int(m.group(6)), int(m.group(7))
        if fd == fc and (d0, d1) != (c0, c1) and not (d1 < c0 or c1 < d0):
            bad.append(line.strip())
    return bad, total


def test_detector_flags_the_pattern():
    bad, total = partial_overlaps("\tv_mfma_f32_16x16x4_f32 a[4:7], v32, v27, a[6:9]\n"
                                  "\tv_mfma_f32_16x16x4_f32 a[0:3], v32, v27, a[0:3]\n"
                                  "\tv_mfma_f32_16x16x32_bf16 v[8:11], v[0:3], v[4:7], a[8:11]\n")
    assert total == 3 and len(bad) == 1


@pytest.mark.parametrize("src", SOURCES)
def test_no_partially_overlapping_mfma(src):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    from pibronic_amd import build
    with tempfile.TemporaryDirectory() as tmp:
        out = os.path.join(tmp, "k.s")
        flags = [f for f in build.COMMON_FLAGS if f != "-fPIC"]
        r = subprocess.run([hipcc, *flags, "--cuda-device-only", "-S", os.path.join(CSRC, src), "-o", out],
                           capture_output=True, text=True)
        assert r.returncode == 0, r.stderr[-2000:]
        bad, total = partial_overlaps(open(out).read())
    assert total > 0
    assert not bad, bad[:5]
