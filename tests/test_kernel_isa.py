"""Build-time check of the responsibilities kernels' gfx950 assembly (no GPU): the BF16 pair
conversion written as inline asm must never read the result of a transcendental instruction
(v_exp_f32 ...) on the very next instruction.  gfx950 needs a wait state there, which the compiler
inserts for the instructions it emits but not in front of inline asm; a build without it gave wrong
responsibilities whenever the scheduler put the conversion right after the exp."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

TRANS = re.compile(r"^v_(exp|log|rcp|rsq|sqrt|sin|cos)_(f16|f32|f64|legacy_f32)\w*$")


def vgprs(operand: str):
    """VGPR numbers named by one operand: v7, v[4:5]."""
    m = re.fullmatch(r"v(\d+)", operand)
    if m:
        return {int(m.group(1))}
    m = re.fullmatch(r"v\[(\d+):(\d+)\]", operand)
    if m:
        return set(range(int(m.group(1)), int(m.group(2)) + 1))
    return set()


def trans_then_read(asm_text: str, reader: str = "v_cvt_pk_bf16_f32"):
    """(trans line, reader line) pairs where `reader` reads a VGPR the instruction right before it
    (a transcendental) wrote."""
    bad, prev = [], None
    for line in asm_text.splitlines():
        t = line.split(";")[0].strip()
        if not t or t.startswith(".") or t.endswith(":"):
            continue
        op, _, rest = t.partition(" ")
        ops = [o.strip() for o in rest.split(",")] if rest else []
        if op == reader and prev is not None:
            pop, pops = prev
            if TRANS.match(pop) and pops and vgprs(pops[0]) & set().union(*[vgprs(o) for o in ops[1:]]):
                bad.append((f"{pop} {', '.join(pops)}", t))
        prev = (op, ops)
    return bad


def test_checker_finds_the_hazard():
    hazard = "\tv_exp_f32_e32 v2, v2\n\tv_cvt_pk_bf16_f32 v5, v2, v60\n"
    assert len(trans_then_read(hazard)) == 1
    assert trans_then_read("\tv_exp_f32_e32 v2, v2\n\ts_nop 0\n\tv_cvt_pk_bf16_f32 v5, v2, v60\n") == []
    assert trans_then_read("\tv_exp_f32_e32 v3, v2\n\tv_cvt_pk_bf16_f32 v5, v2, v60\n") == []
    assert len(trans_then_read("\tv_exp_f32_e32 v61, v2\n\tv_cvt_pk_bf16_f32 v5, v2, v61\n")) == 1


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
def test_softmax_kernels_wait_before_the_inline_conversion(tmp_path):
    src = os.path.join(ROOT, "emc_amd", "csrc", "hip", "softmax.hip")
    out = tmp_path / "softmax.s"
    cmd = [HIPCC, "--offload-arch=gfx950", "-O3", "-std=c++17", "-munsafe-fp-atomics", "-Xclang", "-target-feature",
           "-Xclang", "-packed-fp32-ops", "--cuda-device-only", "-S", "-o", str(out), src,
           "-I" + os.path.join(ROOT, "emc_amd", "csrc")]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    text = out.read_text()
    assert text.count("v_cvt_pk_bf16_f32") > 0
    assert trans_then_read(text) == []
