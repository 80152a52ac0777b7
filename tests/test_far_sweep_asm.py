"""The far kernel's generated asm sweep (bit_amd/csrc/far_sweep_asm.h), checked on
the CPU by emulating its instructions.

* The committed header is what gen_far_sweep.py writes (no stale edit).
* Every register read of a loaded value comes after an s_waitcnt that covers
  its ds_read_b128, no instruction writes a register with a load in flight,
  every load is consumed before the statement ends, and only the declared
  clobbers are written.
* The emulated result equals the C++ sweep's operation order (qstep in
  kernels.hip: the Horner chains over l2, l1, l0 and t) bit for bit, in FP32
  with fused multiply-adds, on random tiles and per-pair values for every
  nterms; a planted mistake (one multiplier index moved) changes it.
"""
import importlib.util
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "bit_amd", "csrc")
HEADER = os.path.join(CSRC, "far_sweep_asm.h")


def _gen():
    spec = importlib.util.spec_from_file_location("gen_far_sweep",
                                                  os.path.join(CSRC, "gen_far_sweep.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def f32(x):
    return np.float32(x)


def fma32(a, b, c):
    # the product of two FP32 values is exact in FP64; the FP64 sum then rounds
    # once more, so this is FP32's fused multiply-add up to rare double rounding
    # (identical for two emulations that perform the same operations)
    return np.float32(np.float64(a) * np.float64(b) + np.float64(c))


def parse(text):
    """{nt: (instruction lines, input expressions, clobbered registers)}"""
    out = {}
    for m in re.finditer(r"qsweep32<(\d+)>\(const QPair<\d+, float>& v,\s*unsigned addr\) \{"
                         r"(.*?)return r;", text, re.S):
        nt = int(m.group(1))
        body = m.group(2)
        lines = re.findall(r'"([^"]*)\\n"', body)
        ins = re.findall(r'"v"\(([^)]*)\)', body)
        clob = re.findall(r'"v(\d+)"', body.split(": \"=&v\"(r)")[1])
        out[nt] = (lines, ins, set(int(c) for c in clob))
    return out


def run_asm(lines, ins, clob, Q, vals):
    """Emulate the statement; vals maps an input expression to its value."""
    ops = {1: None}
    for i, e in enumerate(ins[1:]):
        ops[i + 2] = vals[e]
    regs, pending, issued = {}, {}, 0
    result = None

    def rd(tok):
        if tok.startswith("%"):
            return ops[int(tok[1:])]
        r = int(tok[1:])
        assert r not in pending, f"v{r} read with its load in flight"
        return regs[r]

    def wr(tok, v):
        r = int(tok[1:])
        assert r in clob, f"v{r} written but not clobbered"
        assert r not in pending, f"v{r} written with a load in flight"
        regs[r] = v

    for ln in lines:
        op, _, rest = ln.partition(" ")
        args = [a.strip() for a in rest.split(",")]
        if op == "ds_read_b128":
            m = re.match(r"v\[(\d+):(\d+)\], %1 offset:(\d+)", rest)
            r0, r1, off = int(m.group(1)), int(m.group(2)), int(m.group(3))
            assert r1 == r0 + 3 and r0 % 4 == 0 and off % 16 == 0
            for j in range(4):
                assert r0 + j in clob and r0 + j not in pending
                regs[r0 + j] = Q[off // 4 + j]
                pending[r0 + j] = issued
            issued += 1
        elif op == "s_waitcnt":
            n = int(re.match(r"lgkmcnt\((\d+)\)", rest).group(1))
            for r in [r for r, k in pending.items() if k < issued - n]:
                del pending[r]
        elif op == "v_fmac_f32":
            d, a, b = args
            wr(d, fma32(rd(a), rd(b), rd(d)))
        elif op == "v_fma_f32":
            d, a, b, c = args
            wr(d, fma32(rd(a), rd(b), rd(c)))
        elif op == "v_mul_f32":
            d, a, b = args
            wr(d, f32(np.float64(rd(a)) * np.float64(rd(b))))
        elif op == "v_mov_b32":
            d, s = args
            if d == "%0":
                result = rd(s)
            else:
                wr(d, rd(s))
        else:
            raise AssertionError(f"unexpected instruction {ln}")
    assert not pending, "loads still in flight at the end of the statement"
    return result


def run_ref(nt, Q, mult, g, z, plant=False):
    """The C++ sweep's operations in its order (qstep)."""
    p = nt - 1
    b = a = phi = s = None
    for c, (l0, l1, l2, t) in enumerate(_gen().qtab(nt)):
        L = p - 2 * t
        q = Q[c]
        if l2 == L - l0 - l1:
            b = q
        else:
            i = l0 + l1 + l2 + t
            if plant and c == 7:
                i = (i + 1) % max(p, 1)
            b = fma32(b, mult[2][i], q)
        if l2 == 0:
            a = b if l1 == L - l0 else fma32(a, mult[1][l0 + l1 + t], b)
            if l1 == 0:
                phi = a if l0 == L else fma32(phi, mult[0][l0 + t], a)
                if l0 == 0:
                    tmp = f32(np.float64(g[t]) * np.float64(phi))
                    s = tmp if t == p // 2 else fma32(s, z, tmp)
    return s


def test_header_is_generated():
    assert open(HEADER).read() == _gen().generate()


@pytest.mark.parametrize("nt", range(1, 13))
def test_asm_sweep_equals_cpp_order(nt):
    specs = parse(open(HEADER).read())
    lines, ins, clob = specs[nt]
    assert ins[0] == "addr"
    nq = len(_gen().qtab(nt))
    rng = np.random.default_rng(nt)
    for trial in range(3):
        Q = np.zeros((nq + 3) // 4 * 4, np.float32)
        Q[:nq] = rng.uniform(-1, 1, nq).astype(np.float32)
        mult = rng.uniform(-1.5, 1.5, (3, max(nt - 1, 1))).astype(np.float32)
        g = rng.uniform(0.2, 1.0, nt // 2 + 1).astype(np.float32)
        z = np.float32(rng.uniform(0.0, 0.5))
        vals = {}
        for e in ins[1:]:
            m = re.match(r"v\.mult\[(\d)\]\[(\d+)\]", e)
            if m:
                vals[e] = mult[int(m.group(1))][int(m.group(2))]
            elif e.startswith("v.g["):
                vals[e] = g[int(e[4:-1])]
            else:
                assert e == "v.z"
                vals[e] = z
        got = run_asm(lines, ins, clob, Q, vals)
        want = run_ref(nt, Q, mult, g, z)
        assert got.tobytes() == want.tobytes(), (nt, trial, got, want)
        if nt >= 4:
            assert run_ref(nt, Q, mult, g, z, plant=True).tobytes() != want.tobytes()


def test_far_order_window_matches_tests():
    """The pairing-order window the GPU tests restate is the kernel's."""
    import re as _re
    src = open(os.path.join(CSRC, "kernels.hip")).read()
    m = _re.search(r"#define BIT_FAR_ORDER_WINDOW (\d+)", src)
    spec = importlib.util.spec_from_file_location(
        "tgp", os.path.join(ROOT, "tests", "test_gpu_parity.py"))
    text = open(spec.origin).read()
    w = int(_re.search(r"FAR_ORDER_WINDOW = (\d+)", text).group(1))
    assert m and int(m.group(1)) == w
