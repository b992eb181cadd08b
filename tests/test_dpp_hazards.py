"""Wait states before every DPP instruction of the built library.

The FP32 butterflies are inline assembly (v_fmac_f32 with a DPP source, in
csrc/fft_reg.hpp), and the compiler does not insert the wait states a DPP
instruction needs for registers or EXEC written just before it when the DPP is
inside inline assembly.  This test extracts the gfx950 code object from
libcpwgpu.so, disassembles it (llvm-objdump of the code built here) and checks,
for every DPP instruction: at least 2 wait states since the last VALU write of
its DPP source register, and at least 5 since the last write of EXEC, along
the straight-line code before it (a branch target inside that window counts as
a failure: the other predecessor is not checked)."""

import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "cpw2000_amd", "libcpwgpu.so")
LLVM = "/opt/rocm/llvm/bin"
LINE = re.compile(r"^\s+(\S.*?)\s*//\s*([0-9A-F]{12}):")
BRANCH = re.compile(r"^s_(cbranch_\w+|branch)\s+(-?\d+)")
VREG = re.compile(r"^v(\d+)$|^v\[(\d+):(\d+)\]$")


def _regs(op):
    m = VREG.match(op)
    if not m:
        return set()
    if m.group(1):
        return {int(m.group(1))}
    return set(range(int(m.group(2)), int(m.group(3)) + 1))


def _writes_exec(ins, ops):
    op = ins.split()[0]
    return ("saveexec" in op or op.startswith("v_cmpx")
            or (ops and ops[0] in ("exec", "exec_lo", "exec_hi")))


def _disassemble(tmp):
    fb, co = os.path.join(tmp, "fb.bin"), os.path.join(tmp, "co.o")
    subprocess.run([f"{LLVM}/llvm-objcopy", f"--dump-section=.hip_fatbin={fb}", LIB, os.path.join(tmp, "x.so")],
                   check=True)
    subprocess.run([f"{LLVM}/clang-offload-bundler", "--unbundle", "--type=o",
                    "--targets=hipv4-amdgcn-amd-amdhsa--gfx950", f"--input={fb}", f"--output={co}"], check=True)
    out = subprocess.run([f"{LLVM}/llvm-objdump", "-d", "--mcpu=gfx950", co], check=True,
                         capture_output=True, text=True).stdout
    code = []
    for line in out.split("\n"):
        m = LINE.match(line)
        if m:
            code.append((int(m.group(2), 16), m.group(1)))
    return code


def dpp_hazards(code):
    """code: [(address, instruction text)] in program order.  Returns the DPP
    instructions that lack wait states, and the counts of DPP and of
    v_fmac_f32_dpp instructions seen."""
    targets = set()
    for addr, ins in code:
        m = BRANCH.match(ins)
        if m:
            off = int(m.group(2))
            off = off - 65536 if off > 32767 else off
            targets.add(addr + 4 + 4 * off)
    bad, n_dpp, n_fmac = [], 0, 0
    for i, (addr, ins) in enumerate(code):
        if "quad_perm" not in ins and "row_" not in ins:
            continue
        n_dpp += 1
        n_fmac += ins.startswith("v_fmac_f32_dpp")
        ops = [o.strip() for o in ins.split(None, 1)[1].split(",")] if " " in ins else []
        src = _regs(ops[1].split()[0]) if len(ops) > 1 else set()
        ws, src_ws, exec_ws, j, cur = 0, None, None, i - 1, addr
        while j >= 0 and ws < 5:
            if cur in targets:  # another predecessor: unknown
                break
            a, prev = code[j]
            cur = a
            op = prev.split()[0]
            pops = [o.strip() for o in prev.split(None, 1)[1].split(",")] if " " in prev else []
            if exec_ws is None and _writes_exec(prev, pops):
                exec_ws = ws
            if src_ws is None and op.startswith("v_") and pops and _regs(pops[0]) & src:
                src_ws = ws
            ws += int(pops[0], 0) + 1 if op == "s_nop" else 1
            j -= 1
        if ws < 5 or (exec_ws is not None and exec_ws < 5) or (src_ws is not None and src_ws < 2):
            bad.append(f"{addr:#x} {ins} (exec {exec_ws}, src {src_ws}, window {ws})")
    return bad, n_dpp, n_fmac


def _seq(lines):
    return [(4 * i, t) for i, t in enumerate(lines)]


FILL = ["v_mul_f32_e32 v40, v41, v42"] * 6
BFLY = "v_fmac_f32_dpp v25, v25, v17 row_half_mirror row_mask:0xf bank_mask:0xf"


def test_checker_flags_planted_hazards():
    # the butterfly as the library emits it: s_nop 1 after the write of its source
    ok = _seq(FILL + ["v_mul_f32_e32 v25, v26, v27", "s_nop 1", BFLY])
    assert dpp_hazards(ok)[0] == []
    # no wait state between the write of the DPP source and the DPP read
    assert dpp_hazards(_seq(FILL + ["v_mul_f32_e32 v25, v26, v27", BFLY]))[0]
    # a register range that holds the source
    assert dpp_hazards(_seq(FILL + ["v_pk_mov_b32 v[24:25], v[2:3], v[2:3]", BFLY]))[0]
    # EXEC written three wait states before
    assert dpp_hazards(_seq(FILL + ["s_and_saveexec_b64 s[2:3], vcc", "s_nop 1", BFLY]))[0]
    assert dpp_hazards(_seq(FILL + ["s_mov_b64 exec, s[4:5]", "s_nop 1", BFLY]))[0]
    # five wait states after the EXEC write are enough
    assert dpp_hazards(_seq(FILL + ["s_mov_b64 exec, s[4:5]", "s_nop 4", BFLY]))[0] == []
    # a branch into the window: the other path is not known
    seq = _seq(FILL + ["s_nop 1", BFLY] + ["s_cbranch_scc1 -3"])  # to the s_nop
    assert dpp_hazards(seq)[0]


@pytest.mark.skipif(not os.path.exists(LIB) or not os.path.exists(f"{LLVM}/llvm-objdump"),
                    reason="libcpwgpu.so not built or LLVM tools absent")
def test_dpp_wait_states(tmp_path):
    code = _disassemble(str(tmp_path))
    assert code, "no device code found"
    bad, n_dpp, n_fmac = dpp_hazards(code)
    assert n_fmac > 0, "no one-instruction FP32 butterflies in the library"
    assert not bad, f"{len(bad)} of {n_dpp} DPP instructions lack wait states: " + "; ".join(bad[:5])
