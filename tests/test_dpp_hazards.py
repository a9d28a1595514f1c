"""Static check of the DPP kernels' compiled code (CPU; needs hipcc).

The small-solve and panel substitutions (small, chained, panel and
variable-size batch kernels) issue their row broadcasts as inline assembly (v_fmac_{f64,f32}_dpp, v_mov_{b64,b32}_dpp), for which the
compiler inserts no hazard padding.  On CDNA a DPP instruction must not read a
VGPR written by a VALU instruction within the two preceding wait states, nor
follow an EXEC write within five.  This test compiles those translation units
with the library's own flags to gfx950 assembly and scans each such kernel up
to its end label: s_nop N counts N+1 wait states, any other instruction one,
and a label counts as an unknown predecessor.  The LDS-DMA GEMM kernels,
which issue their tile loads as inline assembly, are checked the same way for
the M0 set-up each load needs.
"""
import os
import re
import subprocess
import concurrent.futures

import pytest

from rocblas_amd import build as rbuild

REG = re.compile(r"\bv\[(\d+):(\d+)\]|\bv(\d+)\b")
DPP_WAIT = 2
EXEC_WAIT = 5


def vgprs(text):
    out = set()
    for m in REG.finditer(text):
        if m.group(3) is not None:
            out.add(int(m.group(3)))
        else:
            out.update(range(int(m.group(1)), int(m.group(2)) + 1))
    return out


def scan(body):
    """Hazards in one kernel body: list of (line, message)."""
    hist = []  # (kind, vgprs written, writes exec, wait states)
    found = []
    for ln, raw in body:
        s = raw.split(";")[0].strip()
        if not s or s.startswith("."):
            continue
        if s.endswith(":"):
            hist.append(("label", set(), True, 0))
            continue
        op, _, args = s.partition(" ")
        operands = [a.strip() for a in args.split(",")]
        if "_dpp" in op or "row_newbcast" in args:
            reads = set()
            for o in operands:
                reads |= vgprs(o.split(" row_")[0])
            ws = 0
            for kind, wr, wex, w in reversed(hist):
                if kind == "label":
                    if ws < EXEC_WAIT:
                        found.append((ln, f"DPP {ws} wait states after a label: {s}"))
                    break
                if kind == "valu" and ws < DPP_WAIT and wr & reads:
                    found.append((ln, f"DPP reads v{sorted(wr & reads)} {ws} wait states "
                                      f"after a VALU write: {s}"))
                if wex and ws < EXEC_WAIT:
                    found.append((ln, f"DPP {ws} wait states after an EXEC write: {s}"))
                ws += w
                if ws >= EXEC_WAIT:
                    break
        if op == "s_nop":
            hist.append(("nop", set(), False, int(args, 0) + 1))
        elif op.startswith("v_"):
            writes_vgpr = not op.startswith(("v_cmp", "v_readlane", "v_readfirstlane"))
            wr = vgprs(operands[0]) if writes_vgpr and operands else set()
            wex = op.startswith("v_cmpx") or operands[0].startswith("exec")
            hist.append(("valu", wr, wex, 1))
        else:
            hist.append(("other", set(), bool(operands) and operands[0].startswith("exec"), 1))
        hist = hist[-64:]
    return found


def kernels(asm, pattern):
    cur, body = None, []
    for i, line in enumerate(asm.split("\n"), 1):
        m = re.match(r"^(_Z\S+):", line)
        if m:
            cur, body = m.group(1), []
            continue
        if cur:
            body.append((i, line))
            if line.startswith(".Lfunc_end"):
                if re.search(pattern, cur):
                    yield cur, body
                cur = None


def compile_asm(src, out):
    flags = [f for f in rbuild.CXXFLAGS if f != "-fPIC"]
    cmd = [rbuild.HIPCC] + flags + ["-S", "--cuda-device-only", src, "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    return out


@pytest.mark.skipif(not os.path.exists(rbuild.HIPCC), reason="hipcc not available")
def test_dpp_kernels_have_no_missing_wait_states(tmp_path):
    srcs = [os.path.join(rbuild.CSRC, f) for f in
            ("small_k8.hip", "small_k16.hip", "small_k32.hip", "chain_trsm.hip", "panel_trsm.hip",
             "vbatched.hip")]
    with concurrent.futures.ThreadPoolExecutor(6) as ex:
        outs = list(ex.map(lambda s: compile_asm(s, str(tmp_path / (os.path.basename(s) + ".s"))),
                           srcs))
    checked, problems = 0, []
    for out in outs:
        pattern = r"small_trsm_dpp_kernel|chain_trsm_kernel|panel_trsm_kernel|vb_panel_kernel"
        for name, body in kernels(open(out).read(), pattern):
            checked += 1
            problems += [f"{name[:70]} line {ln}: {msg}" for ln, msg in scan(body)]
    assert checked == 3 * 32 + 3 * 2, checked
    assert not problems, "\n".join(problems[:20])


@pytest.mark.skipif(not os.path.exists(rbuild.HIPCC), reason="hipcc not available")
def test_gemm_lds_dma_loads_are_set_up(tmp_path):
    """The DMA GEMM kernels issue global_load_lds_dwordx4 from inline assembly:
    every one must follow `s_mov_b32 m0, <scalar register>` (the LDS
    destination; a vector register there would be wrong) and one wait state
    (`s_nop 0`), and the kernels must not spill to scratch."""
    out = compile_asm(os.path.join(rbuild.CSRC, "gemm.hip"), str(tmp_path / "gemm.s"))
    asm = open(out).read()
    checked = 0
    for name, body in kernels(asm, r"gemm_dma"):
        ops = [l.split(";")[0].strip() for _, l in body]
        ops = [o for o in ops if o and not o.startswith(".") and not o.endswith(":")]
        for i, o in enumerate(ops):
            if o.startswith("global_load_lds"):
                checked += 1
                assert ops[i - 1] == "s_nop 0", (name[:60], ops[i - 2:i + 1])
                m = re.match(r"s_mov_b32 m0, (\S+)$", ops[i - 2])
                assert m and not re.match(r"v\d|v\[", m.group(1)), (name[:60], ops[i - 2])
    assert checked >= 16 * 8, checked
    blocks = re.findall(r"\.amdhsa_kernel (\S+)\n(.*?)\.end_amdhsa_kernel", asm, re.S)
    dma = [(n, b) for n, b in blocks if "gemm_dma" in n]
    # FP64: 2 depths x 4 transpose forms dense, 3 forms for each of the 4
    # kinds of triangular operand (8-deep only), the same 12 as the balanced
    # kernel and its 4 reduce kernels; FP32: 2 x 4
    assert len(dma) == 44, len(dma)
    for n, b in dma:
        size = re.search(r"\.amdhsa_private_segment_fixed_size (\d+)", b)
        assert size and int(size.group(1)) == 0, (n[:60], size and size.group(1))


def test_scanner_flags_a_planted_hazard():
    """The scanner itself: a DPP read right after the VALU write of its
    operand is flagged, and two wait states of padding clear it."""
    bad = [(1, "\tv_mul_f64 v[4:5], v[0:1], v[2:3]"),
           (2, "\tv_fmac_f64_dpp v[6:7], v[8:9], v[4:5] row_newbcast:3 row_mask:0xf bank_mask:0xf")]
    assert len(scan(bad)) == 1
    ok = [bad[0], (2, "\ts_nop 1"), (3, bad[1][1])]
    assert scan(ok) == []
    exec_bad = [(1, "\ts_mov_b64 exec, s[2:3]"), (2, "\ts_nop 2"), bad[1]]
    assert len(scan(exec_bad)) == 1
