"""The kernels' few inline-asm instructions never read a register that a
transcendental instruction (v_sqrt, v_rcp, v_rsq, v_log, v_exp, v_sin, v_cos)
wrote within the four instructions before them.

On gfx950 a VALU instruction that reads the result of a transcendental needs a
wait state, which the compiler inserts for its own instructions but not in
front of an inline-asm statement; a read without it sees the old register
value. The device code is compiled to assembly with the library's own flags
(no GPU needed) and every asm statement is checked.
"""
import os
import re
import subprocess

import pytest

from quijibo_amd import build as B

TRANS = ("v_sqrt", "v_rcp", "v_rsq", "v_log", "v_exp", "v_sin", "v_cos")
SOURCES_WITH_ASM = ["csrc/qj_field.hip", "csrc/qj_mc.hip", "csrc/qj_grad.hip"]


def _regs(text):
    out = set()
    for m in re.finditer(r"v\[(\d+):(\d+)\]|\bv(\d+)\b", text):
        if m.group(3):
            out.add(int(m.group(3)))
        else:
            out.update(range(int(m.group(1)), int(m.group(2)) + 1))
    return out


def hazards(asm_lines, window=4):
    """(asm line, offending instruction) pairs: an instruction inside an
    ASMSTART/ASMEND block reading a VGPR written by a transcendental among the
    `window` instructions before the block."""
    found = []
    for i, line in enumerate(asm_lines):
        if "ASMSTART" not in line:
            continue
        k = i + 1
        while k < len(asm_lines) and "ASMEND" not in asm_lines[k]:
            parts = asm_lines[k].strip().split(None, 1)
            k += 1
            if len(parts) < 2 or "," not in parts[1]:
                continue
            srcs = _regs(parts[1].split(",", 1)[1])
            j, seen = i - 1, 0
            while j >= 0 and seen < window:
                prev = asm_lines[j].strip()
                j -= 1
                if not asm_lines[j + 1].startswith("\t") or prev.startswith((";", ".")) or not prev:
                    continue
                seen += 1
                if prev.startswith(TRANS) and _regs(prev.split(None, 1)[1].split(",", 1)[0]) & srcs:
                    found.append((asm_lines[k - 1].strip(), prev))
    return found


def test_scanner_flags_a_transcendental_read():
    lines = ["\tv_sqrt_f32_e32 v27, v26", "\t;;#ASMSTART", "\tv_max_f32_e64 v28, v27, |v29|", "\t;;#ASMEND"]
    assert len(hazards(lines)) == 1
    lines = ["\tv_sqrt_f32_e32 v27, v26", "\t;;#ASMSTART", "\tv_max_f32_e64 v28, v25, |v29|", "\t;;#ASMEND"]
    assert hazards(lines) == []


@pytest.mark.parametrize("src", SOURCES_WITH_ASM)
def test_no_inline_asm_reads_a_fresh_transcendental(src, tmp_path):
    if not os.path.exists(B.HIPCC):
        pytest.skip("hipcc not available")
    out = tmp_path / (os.path.basename(src) + ".s")
    cmd = ([B.HIPCC, "-x", "hip"] + B.COMMON + B.DEVICE_FLAGS + B.FILE_FLAGS.get(src, []) +
           ["--cuda-device-only", "-S", os.path.join(B.HERE, src), "-o", str(out)])
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    lines = out.read_text().splitlines()
    assert any("ASMSTART" in l for l in lines) or src != "csrc/qj_field.hip"
    assert hazards(lines) == []
