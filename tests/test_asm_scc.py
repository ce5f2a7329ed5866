"""The 64-lane encoder narrows exec around its emit store with an inline-asm
s_and_saveexec_b64, and advances its stage pointer with inline-asm
s_bcnt1 / s_lshl1_add; all of them write SCC.  If an asm did not declare the
SCC clobber, the compiler could keep a scalar compare (the stage-upkeep test)
live in SCC across the block and branch on the clobbered value.  This test
compiles the encoder for gfx950 and checks the generated assembly: no SCC
value defined before such a block may be read after it."""
import os
import re
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "mans_amd", "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SCC_DEF = re.compile(r"^\s*(s_cmp|s_bitcmp|s_and_b|s_or_b|s_xor_b|s_andn2_b|s_orn2_b|s_nand|s_nor|s_xnor|"
                     r"s_add_|s_sub_|s_addc|s_subb|s_lshl|s_lshr|s_ashr|s_bfe|s_min_|s_max_|s_abs|s_not|"
                     r"s_bcnt|s_ff|s_flbit|s_mul_hi|s_absdiff|s_cselect)")
SCC_USE = re.compile(r"^\s*(s_cbranch_scc|s_cselect|s_addc|s_subb|s_cmov)")


def scc_violations(asm: str) -> list[int]:
    bad = []
    clobbered = False
    inasm = False
    for i, line in enumerate(asm.split("\n")):
        if "ASMSTART" in line:
            inasm = True
            continue
        if "ASMEND" in line:
            inasm = False
            continue
        if SCC_USE.match(line) and clobbered:
            bad.append(i + 1)
        if inasm and ("s_and_saveexec" in line or SCC_DEF.match(line)):
            clobbered = True
        elif not inasm and SCC_DEF.match(line):
            clobbered = False
    return bad


def test_checker_flags_a_live_scc():
    asm = "\n".join([
        "\ts_cmpk_lt_u32 s33, 0x200",
        "\t;;#ASMSTART",
        "\ts_and_saveexec_b64 s[30:31], s[34:35]",
        "\tds_write_b16 v12, v103",
        "\ts_mov_b64 exec, s[30:31]",
        "\t;;#ASMEND",
        "\ts_cbranch_scc1 .LBB3_77",
    ])
    assert scc_violations(asm) == [7]
    fixed = asm.replace("\t;;#ASMEND", "\t;;#ASMEND\n\ts_cmpk_lt_u32 s33, 0x200")
    assert scc_violations(fixed) == []


@pytest.mark.skipif(not shutil.which(HIPCC) and not os.path.exists(HIPCC), reason="hipcc not available")
def test_encoder_keeps_no_scc_across_exec_asm(tmp_path):
    out = tmp_path / "encode.s"
    cmd = [HIPCC, "--offload-arch=gfx950", "-O3", "-std=c++17", "-fno-strict-aliasing",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", f"-I{CSRC}",
           "-S", "--cuda-device-only", os.path.join(CSRC, "encode.hip"), "-o", str(out)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    asm = out.read_text()
    assert "s_and_saveexec_b64" in asm
    assert scc_violations(asm) == []
