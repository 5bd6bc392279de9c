"""The MFMA kernels' gfx950 assembly has no read of a matrix product's result that a branch brings
too close to the product.

A v_mfma writes its destination registers some cycles after it issues, and the compiler has to
put enough independent instructions (or s_nop) between the product and any non-MFMA instruction
that reads or overwrites those registers. In the cache training kernel it once placed a
`round == 0` branch right after a product, and on the taken path the store of the result followed
with no wait states: lanes whose results had not landed stored the previous partial sum, which
showed up as a few wrong weight-gradient entries from run to run. This test compiles the kernels
that use MFMA, walks every path from each product (fallthrough and branch targets) and fails if a
path that crosses a branch reaches a reader or a writer of the product's registers within 12 wait
states (more than the 16x16x32 and 32x32x16 products need on gfx950); straight-line code is left
to the compiler, whose wait-state insertion handles it. Runs on the CPU: hipcc cross-compiles.
"""
import os
import re
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "myrt_amd", "csrc")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
WAIT_STATES = 12
LOADS = ("global_load", "buffer_load", "ds_read", "scratch_load", "flat_load")
STORES = ("global_store", "buffer_store", "flat_store", "scratch_store", "ds_write", "ds_add", "ds_store")


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    return None


def registers(tok):
    m = re.match(r"([va])\[(\d+):(\d+)\]$", tok)
    if m:
        return {(m.group(1), i) for i in range(int(m.group(2)), int(m.group(3)) + 1)}
    m = re.match(r"([va])(\d+)$", tok)
    return {(m.group(1), int(m.group(2)))} if m else set()


def functions(asm):
    """{kernel: [("label", name) | ("ins", text)]} from an assembly listing."""
    funcs, cur = {}, None
    for line in asm.splitlines():
        text = line.split(";")[0].rstrip()
        if not text.strip():
            continue
        m = re.match(r"^([A-Za-z_.$][\w.$]*):", text)
        if m:
            if not m.group(1).startswith(".L"):
                cur = m.group(1)
                funcs[cur] = []
            if cur is not None:
                funcs[cur].append(("label", m.group(1)))
            continue
        t = text.strip()
        if cur is not None and not t.startswith("."):
            funcs[cur].append(("ins", t))
    return funcs


def close_accesses(body):
    """(product index, access index, wait states, text) for each access of a product's registers
    reached within WAIT_STATES along a path that crosses a branch."""
    labels = {b[1]: i for i, b in enumerate(body) if b[0] == "label"}
    found = []
    for i, (kind, text) in enumerate(body):
        if kind != "ins" or not text.startswith("v_mfma"):
            continue
        dst = registers(text.split(None, 1)[1].split(",")[0].strip())
        stack, seen = [(i + 1, 0, False)], set()
        while stack:
            j, ws, crossed = stack.pop()
            if j >= len(body) or ws >= WAIT_STATES or (j, ws, crossed) in seen:
                continue
            seen.add((j, ws, crossed))
            kind, t = body[j]
            if kind == "label":
                stack.append((j + 1, ws, crossed))
                continue
            op = t.split()[0]
            args = t.split(None, 1)[1] if " " in t else ""
            toks = [a.strip() for a in args.split(",")]
            if op.startswith("v_mfma"):
                src = set().union(*[registers(x) for x in toks[1:3]]) if len(toks) >= 3 else set()
                written = set()  # a product accumulating into the same registers is forwarded
            elif op.startswith(STORES):
                src, written = set().union(*[registers(x) for x in toks]), set()
            elif op.startswith("s_"):
                src, written = set(), set()
            else:
                src = set().union(*[registers(x) for x in toks[1:]]) if toks else set()
                written = registers(toks[0]) if toks and not op.startswith(LOADS) else set()
            if crossed and (src & dst or written & dst):
                found.append((i, j, ws, t))
            m = re.match(r"s_nop (\d+)", t)
            ws2 = ws + (int(m.group(1)) + 1 if m else 1)
            if op == "s_endpgm":
                continue
            if op.startswith("s_branch"):
                stack.append((labels.get(args.strip(), len(body)), ws2, True))
                continue
            if op.startswith("s_cbranch"):
                stack.append((labels.get(args.strip(), len(body)), ws2, True))
                stack.append((j + 1, ws2, True))
                continue
            stack.append((j + 1, ws2, crossed))
    return found


def test_scanner_finds_a_store_right_after_a_branch():
    body = functions("""kern:
\tv_mfma_f32_16x16x32_bf16 v[0:3], v[8:11], v[12:15], v[0:3]
\ts_cbranch_vccnz .LBB0_2
\tglobal_load_dwordx4 v[20:23], v[4:5], off
\ts_waitcnt vmcnt(0)
\ts_nop 7
\ts_nop 7
\tv_add_f32_e32 v0, v0, v20
.LBB0_2:
\tglobal_store_dwordx4 v[4:5], v[0:3], off
\ts_endpgm
""")["kern"]
    found = close_accesses(body)
    assert any("global_store" in t and ws == 1 for _, _, ws, t in found)  # only the branch between
    assert not any("v_add" in t for _, _, _, t in found)  # 19 wait states on that path


@pytest.mark.parametrize("source", ["mlp.hip", "unet.hip"])
def test_no_product_result_is_read_across_a_branch_too_early(source, tmp_path):
    cc = hipcc()
    if cc is None:
        pytest.skip("hipcc not found")
    out = str(tmp_path / (source + ".s"))
    cmd = [cc, "-S", "--offload-device-only", os.path.join(CSRC, source), "-o", out, "-O3", "-std=c++17",
           "-ffp-contract=off", "-I", CSRC, "-I", INCLUDE, "--offload-arch=gfx950", "-Xclang", "-target-feature",
           "-Xclang", "-packed-fp32-ops"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    funcs = functions(open(out).read())
    products = sum(t.startswith("v_mfma") for body in funcs.values() for k, t in body if k == "ins")
    assert products > 0
    bad = {name: close_accesses(body) for name, body in funcs.items()}
    bad = {name: f for name, f in bad.items() if f}
    assert not bad, {name: [(ws, t) for _, _, ws, t in f[:4]] for name, f in bad.items()}
