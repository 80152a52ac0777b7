"""Generate far_sweep_asm.h: the FP32 far kernel's Horner sweep as inline asm.

    python bit_amd/csrc/gen_far_sweep.py        # rewrites far_sweep_asm.h beside this file

The sweep evaluates, for one (target, cluster) pair, the nested Horner form of
the radial-derivative Taylor sum over the cluster's Q tile in LDS, in the
tile's consumption order (t descending, then l0, l1, l2 descending).  Per
entry (l0, l1, l2, t) it performs exactly the operations of the C++ sweep
(`qstep` in kernels.hip), in the same order and with the same operands:

    b   = l2 chain start ? q : fma(b, mult2[l0+l1+l2+t], q)
    a   = l1 chain start ? b : fma(a, mult1[l0+l1+t], b)          (l2 == 0)
    phi = l0 chain start ? a : fma(phi, mult0[l0+t], a)           (l1 == 0)
    s   = t == tmax ? g[t]*phi : fma(s, z, g[t]*phi)              (l0 == 0)

so its results are bit-identical.  What differs is the encoding: the value
being accumulated moves into the register of the Q value it adds (v_fmac,
4 bytes, destination = addend), which the compiler does not do for values
loaded four at a time by ds_read_b128 (it ties only the tuple's last
register and emits the 8-byte three-VGPR v_fma_f32 for the others).

Registers: a ring of RING four-register tuples (RING_NT for some nterms) at
the top of the wave's 128-VGPR budget (v[128 - 4 RING, 127]) receives the tile, ds_read_b128 at a
time, at most INFLIGHT tuples ahead of use; phi and s live in the two
registers below the ring.  Every load is consumed, with its s_waitcnt, inside
the same asm statement (LDS returns in order; scalar loads still in flight
from before the statement only make a wait stricter).
"""
from __future__ import annotations

import os
import sys

RING = 7          # ring tuples (28 VGPRs)
RING_NT = {}      # per-nterms ring sizes where the default does not fit
INFLIGHT = 5      # tuples loaded ahead of the one in use
TOP = 128         # the kernel runs at four waves per SIMD: 128 VGPRs
MAX_NT = 12      # nterms 13 keeps the compiler's sweep: with the ring its
                 # other live values no longer fit in 128 VGPRs


def qtab(nt: int):
    """Entries (l0, l1, l2, t) in consumption order (QTab in kernels.hip)."""
    p = nt - 1
    out = []
    for t in range(p // 2, -1, -1):
        L = p - 2 * t
        for l0 in range(L, -1, -1):
            for l1 in range(L - l0, -1, -1):
                for l2 in range(L - l0 - l1, -1, -1):
                    out.append((l0, l1, l2, t))
    return out


def gen_one(nt: int) -> str:
    ring = RING_NT.get(nt, RING)
    p = nt - 1
    ents = qtab(nt)
    nq = len(ents)
    ntup = (nq + 3) // 4
    ring0 = TOP - 4 * ring
    phi, s = ring0 - 2, ring0 - 1
    tmax = p // 2

    # operands: %0 = result, %1 = LDS byte address, then the per-pair values
    ops = {}
    order = []

    def op(name):
        if name not in ops:
            ops[name] = len(order) + 2
            order.append(name)
        return f"%{ops[name]}"

    busy = [False] * (4 * ring)
    slot_of = {}
    issued = 0
    waited = -1
    lines = []

    def reg(r):
        return f"v{ring0 + r}"

    def free(r):
        assert busy[r]
        busy[r] = False

    def try_issue(k_now):
        nonlocal issued
        while issued < ntup and issued - k_now < INFLIGHT:
            used = set(slot_of[k] for k in range(k_now, issued))
            cand = [j for j in range(ring)
                    if j not in used and not any(busy[4 * j:4 * j + 4])]
            if not cand:
                break
            j = cand[0]
            lines.append(f"ds_read_b128 v[{ring0 + 4 * j}:{ring0 + 4 * j + 3}], %1 "
                         f"offset:{16 * issued}")
            for r in range(4 * j, 4 * j + 4):
                busy[r] = True
            slot_of[issued] = j
            issued += 1

    b = a = None
    for c, (l0, l1, l2, t) in enumerate(ents):
        k = c // 4
        try_issue(k)
        if k >= issued:
            raise RuntimeError(f"NT={nt}: ring too small at entry {c}")
        if waited < k:
            lines.append(f"s_waitcnt lgkmcnt({issued - k - 1})")
            waited = k
        q = 4 * slot_of[k] + c % 4
        L = p - 2 * t
        if l2 == L - l0 - l1:
            b = q
        else:
            lines.append(f"v_fmac_f32 {reg(q)}, {reg(b)}, {op(('m', 2, l0 + l1 + l2 + t))}")
            free(b)
            b = q
        if l2 != 0:
            continue
        if l1 == L - l0:
            a = b
        else:
            lines.append(f"v_fmac_f32 {reg(b)}, {reg(a)}, {op(('m', 1, l0 + l1 + t))}")
            free(a)
            a = b
        b = None
        if l1 != 0:
            continue
        if l0 == L:
            lines.append(f"v_mov_b32 v{phi}, {reg(a)}")
        else:
            lines.append(f"v_fma_f32 v{phi}, v{phi}, {op(('m', 0, l0 + t))}, {reg(a)}")
        free(a)
        a = None
        if l0 != 0:
            continue
        if t == tmax:
            lines.append(f"v_mul_f32 v{s}, {op(('g', t))}, v{phi}")
        else:
            lines.append(f"v_mul_f32 v{phi}, {op(('g', t))}, v{phi}")
            lines.append(f"v_fmac_f32 v{phi}, v{s}, {op(('z',))}")
            phi, s = s, phi
        # the padded tail of the last tuple is never consumed
    for c in range(nq, 4 * ntup):
        busy[4 * slot_of[ntup - 1] + c % 4] = False
    assert issued == ntup and not any(busy)
    lines.append(f"v_mov_b32 %0, v{s}")

    def cexpr(name):
        if name[0] == "m":
            return f"v.mult[{name[1]}][{name[2]}]"
        if name[0] == "g":
            return f"v.g[{name[1]}]"
        return "v.z"

    ins = ", ".join(['"v"(addr)'] + [f'"v"({cexpr(n)})' for n in order])
    clob = ", ".join(f'"v{r}"' for r in range(ring0 - 2, TOP))
    body = "\n".join(f'      "{ln}\\n"' for ln in lines)
    return (f"template <>\n"
            f"__device__ __forceinline__ float qsweep32<{nt}>(const QPair<{nt}, float>& v,\n"
            f"                                                unsigned addr) {{\n"
            f"  float r;\n"
            f"  asm volatile(\n{body}\n"
            f"      : \"=&v\"(r)\n"
            f"      : {ins}\n"
            f"      : {clob}, \"memory\");\n"
            f"  return r;\n"
            f"}}\n")


def generate() -> str:
    head = ("// Generated by gen_far_sweep.py: the FP32 far sweep as inline asm, one\n"
            "// specialisation per nterms.  Do not edit; rerun the generator.\n"
            "#pragma once\n\n"
            "// Sweep over a cluster's Q tile at LDS byte address `addr` for one\n"
            "// target; bit-identical to qall<NT, float>.\n"
            f"constexpr int kAsmSweepMaxNt = {MAX_NT};\n\n"
            "template <int NT>\n"
            "__device__ float qsweep32(const QPair<NT, float>& v, unsigned addr);\n\n")
    return head + "\n".join(gen_one(nt) for nt in range(1, MAX_NT + 1))


def main():
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "far_sweep_asm.h")
    text = generate()
    old = open(out).read() if os.path.exists(out) else None
    if text != old:
        with open(out, "w") as f:
            f.write(text)
    return 0


if __name__ == "__main__":
    sys.exit(main())
