"""numpy wrappers over the checker libraries (test infrastructure only).

``Oracle`` wraps ``oracle/libbit_oracle.so`` (the FP64 C restatement of the
reference treecode, ``bit_oracle.c``); ``Reference`` wraps
``oracle/_ref/libbitref.so`` (the reference's own ``src/tree3d.cpp`` compiled
with ``ref_shim.cpp``), which exists only where the reference checkout was
present at build time.  Only tests, ``__graft_entry__.smoke`` and the CPU
baseline leg of ``bench.py`` import this module.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "libbit_oracle.so")
REF_SO = os.path.join(HERE, "_ref", "libbitref.so")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")


class OraParams(C.Structure):
    _fields_ = [("nterms", C.c_int), ("acc", C.c_double), ("kappa", C.c_double),
                ("delta2", C.c_double), ("mac_sub", C.c_double), ("kernel_order", C.c_int),
                ("root", C.c_int)]


@dataclass
class Tree:
    numtree: int
    boxes: np.ndarray      # [numtree, 13]
    children: np.ndarray   # [numtree, 8]
    nummemb: np.ndarray    # [numtree]
    members: np.ndarray    # concatenated member lists
    offsets: np.ndarray    # [numtree+1] into members

    def node_members(self, i: int) -> np.ndarray:
        return self.members[self.offsets[i]:self.offsets[i + 1]]


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


class Oracle:
    """The FP64 restatement; each method names the reference function it follows."""

    def __init__(self, path: str = ORACLE_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run `make -C oracle`")
        L = C.CDLL(path)
        L.ora_tree_build.restype = C.c_void_p
        L.ora_tree_build.argtypes = [C.c_int, _dp, _dp, _dp, C.c_int]
        L.ora_tree_free.argtypes = [C.c_void_p]
        L.ora_tree_numtree.restype = C.c_int
        L.ora_tree_numtree.argtypes = [C.c_void_p]
        L.ora_tree_quirk_hits.restype = C.c_int
        L.ora_tree_quirk_hits.argtypes = [C.c_void_p]
        L.ora_tree_total_members.restype = C.c_int64
        L.ora_tree_total_members.argtypes = [C.c_void_p]
        L.ora_tree_export.argtypes = [C.c_void_p, _dp, _ip, _ip, _ip]
        L.ora_moments.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, C.c_int, _dp]
        L.ora_taylor_coeffs.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double,
                                        C.c_double, C.c_double, _dp]
        L.ora_potential.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _dp, OraParams, C.c_int,
                                    _ip, _dp, _lp, _lp, C.c_int]
        L.ora_lists.argtypes = [C.c_void_p, _dp, _dp, _dp, OraParams, C.c_int, _ip,
                                C.POINTER(C.c_int), _ip, C.POINTER(C.c_int), C.c_int]
        L.ora_direct.argtypes = [C.c_int, _dp, _dp, _dp, _dp, C.c_double, C.c_double,
                                 C.c_int, _ip, _dp, C.c_int]
        L.ora_treeforce.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _dp, OraParams, C.c_int,
                                    _ip, _dp, _dp, _dp, C.c_int]
        L.ora_potential_points.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _dp, OraParams,
                                           C.c_int, _dp, _dp, _dp, _dp, C.c_int]
        L.ora_treeforce_points.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _dp, OraParams,
                                           C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, C.c_int]
        L.ora_direct_force.argtypes = [C.c_int, _dp, _dp, _dp, _dp, C.c_double, C.c_int,
                                       C.c_int, _ip, _dp, _dp, _dp, C.c_int]
        L.ora_powinv_coeffs.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                        C.c_double, _dp]
        self.L = L

    # treemake (tree3d.cpp:1345)
    def tree(self, x, y, z, maxmemb: int):
        x, y, z = _f64(x), _f64(y), _f64(z)
        h = self.L.ora_tree_build(len(x), x, y, z, int(maxmemb))
        if not h:
            raise RuntimeError("oracle tree build failed")
        return OracleTree(self, h, x, y, z)

    def taylor_coeffs(self, nterms, dx, dy, dz, R2, kappa):
        S = nterms + 1
        a = np.zeros(S * S * S)
        self.L.ora_taylor_coeffs(nterms, dx, dy, dz, R2, kappa, a)
        return a.reshape(S, S, S)

    # shielded_pot_particle_ds (tree3d.cpp:1966), +phi convention
    def direct(self, x, y, z, q, kappa, delta2, tidx=None, nthreads=0):
        x, y, z, q = _f64(x), _f64(y), _f64(z), _f64(q)
        tidx = _i32(np.arange(len(x)) if tidx is None else tidx)
        out = np.zeros(len(tidx))
        self.L.ora_direct(len(x), x, y, z, q, kappa, delta2, len(tidx), tidx, out, nthreads)
        return out

    def direct_force(self, x, y, z, q, delta2, tidx=None, nthreads=0, kernel_order=0):
        x, y, z, q = _f64(x), _f64(y), _f64(z), _f64(q)
        tidx = _i32(np.arange(len(x)) if tidx is None else tidx)
        e = [np.zeros(len(tidx)) for _ in range(3)]
        self.L.ora_direct_force(len(x), x, y, z, q, delta2, kernel_order, len(tidx), tidx, *e,
                                nthreads)
        return np.stack(e, axis=1)

    def powinv_coeffs(self, nterms, l, dx, dy, dz, R2):
        S = nterms + 1
        a = np.zeros(S * S * S)
        self.L.ora_powinv_coeffs(nterms, l, dx, dy, dz, R2, a)
        return a.reshape(S, S, S)


class OracleTree:
    def __init__(self, ora: Oracle, h, x, y, z):
        self.ora, self.h, self.x, self.y, self.z = ora, h, x, y, z
        self.numtree = ora.L.ora_tree_numtree(h)
        self.quirk_hits = ora.L.ora_tree_quirk_hits(h)

    def __del__(self):
        try:
            self.ora.L.ora_tree_free(self.h)
        except Exception:
            pass

    def export(self) -> Tree:
        nt = self.numtree
        tot = self.ora.L.ora_tree_total_members(self.h)
        boxes = np.zeros(nt * 13)
        ch = np.zeros(nt * 8, dtype=np.int32)
        nm = np.zeros(nt, dtype=np.int32)
        mem = np.zeros(max(tot, 1), dtype=np.int32)
        self.ora.L.ora_tree_export(self.h, boxes, ch, nm, mem)
        off = np.zeros(nt + 1, dtype=np.int64)
        off[1:] = np.cumsum(nm)
        return Tree(nt, boxes.reshape(nt, 13), ch.reshape(nt, 8), nm, mem[:tot], off)

    # treemoments (tree3d.cpp:2631): [numtree, nterms, nterms, nterms]
    def moments(self, q, nterms):
        out = np.zeros(self.numtree * nterms ** 3)
        self.ora.L.ora_moments(self.h, self.x, self.y, self.z, _f64(q), nterms, out)
        return out.reshape(self.numtree, nterms, nterms, nterms)

    def _root(self, root):
        if not 0 <= int(root) < self.numtree:
            raise ValueError(f"root {root} is not a node (numtree {self.numtree})")
        return int(root)

    # root: the node the traversal starts at (the reference's `ind`), 0 = whole tree
    def potential(self, q, moments, nterms, acc, kappa, delta2, mac_sub, tidx=None,
                  nthreads=0, counts=False, root=0):
        q = _f64(q)
        tidx = _i32(np.arange(len(q)) if tidx is None else tidx)
        nt = len(tidx)
        pot = np.zeros(nt)
        nn = np.zeros(nt, dtype=np.int64)
        nf = np.zeros(nt, dtype=np.int64)
        prm = OraParams(nterms, acc, kappa, delta2, mac_sub, 0, self._root(root))
        self.ora.L.ora_potential(self.h, self.x, self.y, self.z, q, _f64(moments).ravel(),
                                 prm, nt, tidx, pot, nn, nf, nthreads)
        return (pot, nn, nf) if counts else pot

    def potential_points(self, q, moments, nterms, acc, kappa, delta2, mac_sub, px, py, pz,
                         nthreads=0, root=0):
        px, py, pz = _f64(px), _f64(py), _f64(pz)
        out = np.zeros(len(px))
        prm = OraParams(nterms, acc, kappa, delta2, mac_sub, 0, self._root(root))
        self.ora.L.ora_potential_points(self.h, self.x, self.y, self.z, _f64(q),
                                        _f64(moments).ravel(), prm, len(px), px, py, pz, out,
                                        nthreads)
        return out

    def treeforce_points(self, q, moments, nterms, acc, delta2, mac_sub, px, py, pz, nthreads=0,
                         kernel_order=0, root=0):
        px, py, pz = _f64(px), _f64(py), _f64(pz)
        e = [np.zeros(len(px)) for _ in range(3)]
        prm = OraParams(nterms, acc, 0.0, delta2, mac_sub, kernel_order, self._root(root))
        self.ora.L.ora_treeforce_points(self.h, self.x, self.y, self.z, _f64(q),
                                        _f64(moments).ravel(), prm, len(px), px, py, pz, *e,
                                        nthreads)
        return np.stack(e, axis=1)

    def lists(self, ti, nterms, acc, delta2, mac_sub, cap=4096, root=0):
        far = np.zeros(cap, dtype=np.int32)
        near = np.zeros(cap, dtype=np.int32)
        nf, nn = C.c_int(0), C.c_int(0)
        prm = OraParams(nterms, acc, 0.0, delta2, mac_sub, 0, self._root(root))
        self.ora.L.ora_lists(self.h, self.x, self.y, self.z, prm, int(ti), far, C.byref(nf),
                             near, C.byref(nn), cap)
        return far[:nf.value].copy(), near[:nn.value].copy()

    def treeforce(self, q, moments, nterms, acc, delta2, mac_sub, tidx=None, nthreads=0,
                  kernel_order=0, root=0):
        q = _f64(q)
        tidx = _i32(np.arange(len(q)) if tidx is None else tidx)
        e = [np.zeros(len(tidx)) for _ in range(3)]
        prm = OraParams(nterms, acc, 0.0, delta2, mac_sub, kernel_order, self._root(root))
        self.ora.L.ora_treeforce(self.h, self.x, self.y, self.z, q, _f64(moments).ravel(), prm,
                                 len(tidx), tidx, *e, nthreads)
        return np.stack(e, axis=1)


class Reference:
    """The reference library itself (built from its sources by oracle/Makefile)."""

    DEL = 0.01  # problem.h:12, compiled in

    def __init__(self, path: str = REF_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run `make -C oracle ref`")
        L = C.CDLL(path)
        L.ref_tree.restype = C.c_int
        L.ref_tree.argtypes = [C.c_int, _dp, _dp, _dp, _dp, C.c_int, C.c_int, C.c_int,
                               C.c_int64, _dp, _ip, _ip, _ip, C.c_void_p]
        L.ref_shielded_treepot.argtypes = [C.c_int, _dp, _dp, _dp, _dp, C.c_int, C.c_int,
                                           C.c_double, C.c_double, C.c_int, _ip, _dp]
        L.ref_treepotential_point.argtypes = [C.c_int, _dp, _dp, _dp, _dp, C.c_int, C.c_int,
                                              C.c_double, C.c_int, _ip, _dp]
        L.ref_treeforce.argtypes = [C.c_int, _dp, _dp, _dp, _dp, C.c_int, C.c_int, C.c_double,
                                    C.c_int, _ip, _dp, _dp, _dp]
        L.ref_shielded_pot_ds.argtypes = [C.c_int, _dp, _dp, _dp, _dp, C.c_double, C.c_int,
                                          _ip, _dp]
        L.ref_force_ds_reg.argtypes = [C.c_int, _dp, _dp, _dp, _dp, C.c_int, _ip, _dp, _dp,
                                       _dp]
        L.ref_subtree_calls.argtypes = [C.c_int, _dp, _dp, _dp, _dp, C.c_int, C.c_int,
                                        C.c_double, C.c_double, C.c_int, _ip, _ip, _dp, _dp,
                                        _dp, _dp, _dp]
        L.ref_boxparticles.argtypes = [C.c_int, _dp, _dp, _dp, _dp]
        L.ref_self_intG.restype = C.c_double
        self.L = L
        self.intG = L.ref_self_intG()

    def tree(self, x, y, z, q, maxmemb, nterms, with_moments=True):
        x, y, z, q = _f64(x), _f64(y), _f64(z), _f64(q)
        n = len(x)
        cap_nodes = 4 * n + 64
        cap_mem = n * 64
        boxes = np.zeros(cap_nodes * 13)
        ch = np.zeros(cap_nodes * 8, dtype=np.int32)
        nm = np.zeros(cap_nodes, dtype=np.int32)
        mem = np.zeros(cap_mem, dtype=np.int32)
        mom = np.zeros(cap_nodes * nterms ** 3) if with_moments else None
        nt = self.L.ref_tree(n, x, y, z, q, maxmemb, nterms, cap_nodes, cap_mem, boxes, ch, nm,
                             mem, mom.ctypes.data if with_moments else None)
        if nt < 0:
            raise RuntimeError("reference tree export overflow")
        off = np.zeros(nt + 1, dtype=np.int64)
        off[1:] = np.cumsum(nm[:nt])
        t = Tree(nt, boxes[:nt * 13].reshape(nt, 13).copy(), ch[:nt * 8].reshape(nt, 8).copy(),
                 nm[:nt].copy(), mem[:off[-1]].copy(), off)
        if with_moments:
            return t, mom[:nt * nterms ** 3].reshape(nt, nterms, nterms, nterms).copy()
        return t

    def shielded_treepot_raw(self, x, y, z, q, maxmemb, nterms, acc, kappa, tidx):
        tidx = _i32(tidx)
        out = np.zeros(len(tidx))
        self.L.ref_shielded_treepot(len(x), _f64(x), _f64(y), _f64(z), _f64(q), maxmemb, nterms,
                                    acc, kappa, len(tidx), tidx, out)
        return out

    def shielded_phi(self, x, y, z, q, maxmemb, nterms, acc, kappa, tidx):
        """+phi (delta=0, MAC shift DEL) from shielded_treepot: undo its sign and self term."""
        raw = self.shielded_treepot_raw(x, y, z, q, maxmemb, nterms, acc, kappa, tidx)
        qi = _f64(q)[_i32(tidx)]
        return -(raw - qi * self.intG)

    def coulomb_reg_phi(self, x, y, z, q, maxmemb, nterms, acc, tidx):
        """+phi with delta^2 = DEL from treepotential(p=-1) at the particle positions,
        minus the target's own regularized self term q_i/sqrt(DEL)."""
        tidx = _i32(tidx)
        out = np.zeros(len(tidx))
        self.L.ref_treepotential_point(len(x), _f64(x), _f64(y), _f64(z), _f64(q), maxmemb,
                                       nterms, acc, len(tidx), tidx, out)
        return out - _f64(q)[tidx] / np.sqrt(self.DEL)

    def treeforce(self, x, y, z, q, maxmemb, nterms, acc, tidx):
        tidx = _i32(tidx)
        e = [np.zeros(len(tidx)) for _ in range(3)]
        self.L.ref_treeforce(len(x), _f64(x), _f64(y), _f64(z), _f64(q), maxmemb, nterms, acc,
                             len(tidx), tidx, *e)
        return np.stack(e, axis=1)

    def subtree_calls(self, x, y, z, q, maxmemb, nterms, acc, kappa, inds, tidx):
        """Raw reference values of the calls that start at node inds[t] for
        particle tidx[t]: shielded_treepot(ind, i), treepotential(ind, -1, x_i)
        and treeforce(ind, i) -> (shielded, pot_point, field [npairs, 3])."""
        inds, tidx = _i32(inds), _i32(tidx)
        m = len(tidx)
        sh, pp = np.zeros(m), np.zeros(m)
        e = [np.zeros(m) for _ in range(3)]
        self.L.ref_subtree_calls(len(x), _f64(x), _f64(y), _f64(z), _f64(q), maxmemb, nterms,
                                 acc, kappa, m, inds, tidx, sh, pp, *e)
        return sh, pp, np.stack(e, axis=1)

    def shielded_phi_ds(self, x, y, z, q, kappa, tidx):
        tidx = _i32(tidx)
        out = np.zeros(len(tidx))
        self.L.ref_shielded_pot_ds(len(x), _f64(x), _f64(y), _f64(z), _f64(q), kappa, len(tidx),
                                   tidx, out)
        return -out

    def force_ds_reg(self, x, y, z, q, tidx):
        tidx = _i32(tidx)
        e = [np.zeros(len(tidx)) for _ in range(3)]
        self.L.ref_force_ds_reg(len(x), _f64(x), _f64(y), _f64(z), _f64(q), len(tidx), tidx, *e)
        return np.stack(e, axis=1)

    def boxparticles(self, n):
        x, y, z, q = (np.zeros(n) for _ in range(4))
        self.L.ref_boxparticles(n, x, y, z, q)
        return x, y, z, q
