"""bit_amd -- MI355X particle-cluster treecode (host mirror over the C-ABI).

The native library ``bit_amd/libbit_amd.so`` (include/bit_amd.h) holds the
whole evaluation path: the host tree build (bit-exact with the reference
``treemake``, tree3d.cpp:1345) and the gfx950 HIP kernels (moments, batch
interaction lists, far field, near field).  This module loads it with ctypes
and mirrors the reference's operations on numpy / torch arrays:

    ===========================  ==========================================
    reference (src/tree3d.h)     here
    ===========================  ==========================================
    treemake                     System(ctx, x, y, z, q, maxmemb)
    treemoments                  System.moments(nterms) / moments_cube()
    shielded_treepot loop        System.potential(nterms, acc, kappa, ...)
    treepotential loop           System.potential(kappa=0, delta2=DEL, ...)
    treeforce loop               System.field(nterms, acc, delta2, ...)
    shielded_pot_particle_ds     System.direct(kappa, delta2, tidx)
    ===========================  ==========================================

There is no CPU fallback: if the library is missing, importing this module
raises ImportError (build it with ``python -m bit_amd.build``).
"""
from __future__ import annotations

import ctypes as C
import os
import weakref
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libbit_amd.so")

BIT_OK = 0
BIT_E_INVALID = -1
BIT_E_OOM = -2
BIT_E_HIP = -3
BIT_E_TREE = -4
BIT_E_STATE = -5
BIT_FP64 = 0
BIT_FP32 = 1
MAX_NTERMS = 13

# Reference compile-time constants (problem.h:11-13, constants.h:10).
ACC = 0.5
DEL = 0.01
MAXTERM = 12


class BitError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"bit_amd error {code}: {msg}")
        self.code = code


class Params(C.Structure):
    _fields_ = [("nterms", C.c_int), ("acc", C.c_double), ("kappa", C.c_double),
                ("delta2", C.c_double), ("mac_sub", C.c_double), ("precision", C.c_int),
                ("kernel_order", C.c_int), ("root", C.c_int)]


class Stats(C.Structure):
    _fields_ = [("near_pairs", C.c_int64), ("far_pairs", C.c_int64), ("far_items", C.c_int64),
                ("near_items", C.c_int64), ("lane_far", C.c_int64),
                ("lane_near_pairs", C.c_int64), ("ms_tree", C.c_double),
                ("ms_upload", C.c_double), ("ms_moments", C.c_double), ("ms_lists", C.c_double),
                ("ms_far", C.c_double), ("ms_near", C.c_double), ("list_passes", C.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


# Every symbol include/bit_amd.h declares.
EXPORTS = [
    "bit_params_default", "bit_last_error", "bit_ctx_create", "bit_ctx_destroy",
    "bit_ctx_set_stream", "bit_ctx_sync", "bit_ctx_bytes_in_use",
    "bit_ctx_inject_alloc_failures", "bit_system_create", "bit_system_create_device",
    "bit_system_destroy", "bit_system_numtree", "bit_system_maxdepth",
    "bit_system_export_tree", "bit_moments", "bit_moments_download", "bit_eval_potential",
    "bit_eval_field", "bit_system_stats", "bit_direct_potential", "bit_tree_build",
    "bit_tree_destroy", "bit_tree_numtree", "bit_tree_maxdepth", "bit_tree_export",
    "bit_system_export_lists", "bit_device_alloc", "bit_device_free", "bit_copy_to_device",
    "bit_copy_to_host", "bit_system_set_charges", "bit_solve_second_kind", "bit_gmres_dense",
    "bit_system_prepare", "bit_eval_potential_multi", "bit_gmres", "bit_ctx_set_host_threads",
    "bit_copy_on_device", "bit_eval_potential_points", "bit_eval_field_points",
    "bit_direct_field", "bit_direct_field_points", "bit_eval_potential_noalloc",
    "bit_krylov_create", "bit_krylov_destroy", "bit_krylov_status", "bit_krylov_solution",
    "bit_serve_round", "bit_valu_ceiling",
]


# void (*)(void *user, const double *in, double *out)
MATVEC_FN = C.CFUNCTYPE(None, C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double))


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built: run `python -m bit_amd.build`")
    # torch (ROCm build) ships its own HIP runtime under the same soname.  If
    # torch is imported first, this library binds to that runtime (one HIP
    # runtime per process).  Importing torch after this library has loaded the
    # image's runtime gives the process two runtimes and torch then finds no
    # device: import torch first when both are used.
    L = C.CDLL(LIB_PATH)
    vp, dp, ip = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int)
    L.bit_last_error.restype = C.c_char_p
    L.bit_params_default.argtypes = [C.POINTER(Params)]
    L.bit_ctx_create.argtypes = [C.c_int, C.c_size_t, C.POINTER(vp)]
    L.bit_ctx_destroy.argtypes = [vp]
    L.bit_ctx_set_stream.argtypes = [vp, vp]
    L.bit_ctx_sync.argtypes = [vp]
    L.bit_ctx_bytes_in_use.argtypes = [vp]
    L.bit_ctx_bytes_in_use.restype = C.c_size_t
    L.bit_ctx_inject_alloc_failures.argtypes = [vp, C.c_int]
    for f in ("bit_system_create", "bit_system_create_device"):
        getattr(L, f).argtypes = [vp, C.c_int, vp, vp, vp, vp, C.c_int, C.POINTER(vp)]
    L.bit_system_destroy.argtypes = [vp]
    L.bit_system_numtree.argtypes = [vp]
    L.bit_system_maxdepth.argtypes = [vp]
    L.bit_system_export_tree.argtypes = [vp, vp, vp, vp, vp, vp]
    L.bit_moments.argtypes = [vp, C.c_int]
    L.bit_moments_download.argtypes = [vp, vp]
    L.bit_eval_potential.argtypes = [vp, C.POINTER(Params), vp, C.c_int]
    L.bit_eval_field.argtypes = [vp, C.POINTER(Params), vp, C.c_int]
    L.bit_eval_potential_noalloc.argtypes = [vp, C.POINTER(Params), vp]
    L.bit_system_stats.argtypes = [vp, C.POINTER(Stats)]
    L.bit_direct_potential.argtypes = [vp, C.c_double, C.c_double, C.c_int, vp, vp]
    L.bit_direct_field.argtypes = [vp, C.c_double, C.c_int, C.c_int, vp, vp]
    L.bit_direct_field_points.argtypes = [vp, C.c_double, C.c_int, C.c_int, vp, vp, vp, vp]
    L.bit_system_export_lists.argtypes = [vp, C.POINTER(Params), vp, vp, vp, vp, vp, vp, vp,
                                          vp, vp]
    L.bit_device_alloc.argtypes = [vp, C.c_size_t, C.POINTER(vp)]
    L.bit_device_free.argtypes = [vp, vp]
    L.bit_copy_to_device.argtypes = [vp, vp, vp, C.c_size_t]
    L.bit_copy_to_host.argtypes = [vp, vp, vp, C.c_size_t]
    L.bit_copy_on_device.argtypes = [vp, vp, vp, C.c_size_t]
    L.bit_system_set_charges.argtypes = [vp, vp, C.c_int]
    L.bit_solve_second_kind.argtypes = [vp, C.POINTER(Params), C.c_double, C.c_double, vp, vp,
                                        vp, C.c_int, C.c_double, C.c_int, vp, vp]
    L.bit_gmres_dense.argtypes = [C.c_int, vp, vp, vp, C.c_int, C.c_double, C.c_int, vp, vp]
    L.bit_system_prepare.argtypes = [vp, C.POINTER(Params)]
    L.bit_eval_potential_multi.argtypes = [vp, C.c_int, vp, vp, vp, C.c_int, vp]
    L.bit_gmres.argtypes = [C.c_int, MATVEC_FN, vp, vp, vp, C.c_int, C.c_double, C.c_int, vp,
                            vp]
    L.bit_ctx_set_host_threads.argtypes = [vp, C.c_int]
    L.bit_krylov_create.argtypes = [vp, C.POINTER(Params), C.c_double, C.c_double, vp, vp, vp,
                                    C.c_int, C.c_double, C.c_int, C.POINTER(vp)]
    L.bit_krylov_destroy.argtypes = [vp]
    L.bit_krylov_status.argtypes = [vp, vp, vp, vp]
    L.bit_krylov_solution.argtypes = [vp, vp]
    L.bit_serve_round.argtypes = [vp, C.c_int, vp, vp, vp, C.c_int, vp, vp]
    L.bit_valu_ceiling.argtypes = [vp, dp, dp]
    for f in ("bit_eval_potential_points", "bit_eval_field_points"):
        getattr(L, f).argtypes = [vp, C.POINTER(Params), C.c_int, vp, vp, vp, vp]
    L.bit_tree_build.argtypes = [C.c_int, vp, vp, vp, C.c_int, C.POINTER(vp)]
    L.bit_tree_destroy.argtypes = [vp]
    L.bit_tree_numtree.argtypes = [vp]
    L.bit_tree_maxdepth.argtypes = [vp]
    L.bit_tree_export.argtypes = [vp, vp, vp, vp, vp, vp]
    for name in EXPORTS:
        if name not in ("bit_last_error", "bit_params_default", "bit_ctx_bytes_in_use"):
            getattr(L, name).restype = C.c_int
    return L


class _Lib:
    """The native library, loaded on first use (ImportError if it is not built;
    there is no fallback).  Loading lazily lets `python -m bit_amd.build`
    import this package before the library exists."""

    _L = None

    def __getattr__(self, name):
        if _Lib._L is None:
            _Lib._L = _load()
        return getattr(_Lib._L, name)


lib = _Lib()


def _check(rc: int):
    if rc != BIT_OK:
        raise BitError(rc, lib.bit_last_error().decode(errors="replace"))


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def make_params(nterms=6, acc=ACC, kappa=0.0, delta2=0.0, mac_sub=None, precision="fp64",
                kernel_order=0, root=0):
    """Evaluation parameters; mac_sub defaults to delta2 (the MAC on the true
    squared distance), the reference's treepotential/treeforce convention.
    root: the node the traversal starts at (the reference's `ind`); 0 = the
    whole tree, another node gives that subtree's contribution."""
    p = Params()
    p.nterms = int(nterms)
    p.acc = float(acc)
    p.kappa = float(kappa)
    p.delta2 = float(delta2)
    p.mac_sub = float(delta2 if mac_sub is None else mac_sub)
    p.precision = BIT_FP32 if precision in ("fp32", BIT_FP32) else BIT_FP64
    p.kernel_order = int(kernel_order)
    p.root = int(root)
    return p


class Context:
    """One device + one stream + an HBM budget (bytes, 0 = unlimited)."""

    def __init__(self, device: int = 0, hbm_budget: int = 0):
        self.h = None
        self._systems = weakref.WeakSet()
        self._arrays = weakref.WeakSet()
        h = C.c_void_p()
        _check(lib.bit_ctx_create(int(device), int(hbm_budget), C.byref(h)))
        self.h = h

    def close(self):
        """Closes the context's systems and device arrays, then the context."""
        if self.h:
            for s in list(self._systems):
                s.close()
            for a in list(self._arrays):
                a.ptr = None  # freed with the context
            _check(lib.bit_ctx_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_stream(self, stream_ptr: int | None):
        _check(lib.bit_ctx_set_stream(self.h, C.c_void_p(stream_ptr or 0)))

    def sync(self):
        _check(lib.bit_ctx_sync(self.h))

    def valu_ceiling(self) -> tuple[float, float]:
        """(FP32, FP64) lane TFLOP/s of independent FMA chains on this device:
        the issue ceiling of the library's instruction mix (no packed FP32)."""
        a, b = C.c_double(), C.c_double()
        _check(lib.bit_valu_ceiling(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    @property
    def bytes_in_use(self) -> int:
        return int(lib.bit_ctx_bytes_in_use(self.h))

    def inject_alloc_failures(self, k: int):
        _check(lib.bit_ctx_inject_alloc_failures(self.h, int(k)))

    def set_host_threads(self, k: int):
        _check(lib.bit_ctx_set_host_threads(self.h, int(k)))


class DeviceArray:
    """A float64 (or int32) array in HBM owned by a Context (bit_device_alloc)."""

    def __init__(self, ctx: Context, n: int, dtype=np.float64):
        self.ctx, self.n, self.dtype = ctx, int(n), np.dtype(dtype)
        h = C.c_void_p()
        _check(lib.bit_device_alloc(ctx.h, self.n * self.dtype.itemsize, C.byref(h)))
        self.ptr = h.value
        ctx._arrays.add(self)

    @classmethod
    def from_numpy(cls, ctx: Context, a) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(ctx, a.size, a.dtype)
        _check(lib.bit_copy_to_device(ctx.h, C.c_void_p(d.ptr), _ptr(a), a.nbytes))
        return d

    def numpy(self) -> np.ndarray:
        out = np.empty(self.n, dtype=self.dtype)
        _check(lib.bit_copy_to_host(self.ctx.h, _ptr(out), C.c_void_p(self.ptr), out.nbytes))
        return out

    def data_ptr(self) -> int:
        return self.ptr

    def copy_from(self, other: "DeviceArray"):
        """Device-to-device copy of another array of the same size."""
        _check(lib.bit_copy_on_device(self.ctx.h, C.c_void_p(self.ptr), C.c_void_p(other.ptr),
                                      self.n * self.dtype.itemsize))

    def free(self):
        if self.ptr and self.ctx.h:
            _check(lib.bit_device_free(self.ctx.h, C.c_void_p(self.ptr)))
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _dev_ptr(a):
    """Device address of a DeviceArray or a torch device tensor, else None."""
    if isinstance(a, DeviceArray):
        return a.ptr
    if hasattr(a, "is_cuda") and a.is_cuda:
        return a.data_ptr()
    return None


@dataclass
class TreeExport:
    numtree: int
    boxes: np.ndarray     # [numtree, 13] as TREE (tree3d.h:51-74)
    children: np.ndarray  # [numtree, 8]
    begin: np.ndarray
    end: np.ndarray
    perm: np.ndarray      # sorted slot -> original index

    def members(self, i: int) -> np.ndarray:
        """Original indices of node i in slot order (inside a leaf: Morton order)."""
        return self.perm[self.begin[i]:self.end[i]]

    def member_list(self, i: int) -> np.ndarray:
        """TREE::members of node i: ascending original indices."""
        return np.sort(self.members(i))


def build_tree(x, y, z, maxmemb: int) -> TreeExport:
    """Host-only tree build (treemake, tree3d.cpp:1345); needs no device."""
    x, y, z = _f64(x), _f64(y), _f64(z)
    n = len(x)
    h = C.c_void_p()
    _check(lib.bit_tree_build(n, _ptr(x), _ptr(y), _ptr(z), int(maxmemb), C.byref(h)))
    try:
        nt = lib.bit_tree_numtree(h)
        boxes = np.zeros(nt * 13)
        ch = np.zeros(nt * 8, dtype=np.int32)
        b = np.zeros(nt, dtype=np.int32)
        e = np.zeros(nt, dtype=np.int32)
        p = np.zeros(n, dtype=np.int32)
        _check(lib.bit_tree_export(h, _ptr(boxes), _ptr(ch), _ptr(b), _ptr(e), _ptr(p)))
        t = TreeExport(nt, boxes.reshape(nt, 13), ch.reshape(nt, 8), b, e, p)
        t.maxdepth = lib.bit_tree_maxdepth(h)
        return t
    finally:
        lib.bit_tree_destroy(h)


def gmres_dense(A, b, x0=None, restart=50, tol=1e-10, maxit=1000):
    """Host-only FP64 restarted GMRES on a dense matrix (the solver core)."""
    A = np.ascontiguousarray(A, dtype=np.float64)
    b = _f64(b)
    x = np.zeros(len(b)) if x0 is None else _f64(x0).copy()
    it, rr = C.c_int(), C.c_double()
    _check(lib.bit_gmres_dense(len(b), _ptr(A), _ptr(b), _ptr(x), int(restart), float(tol),
                               int(maxit), C.byref(it), C.byref(rr)))
    return x, it.value, rr.value


def gmres(n, matvec, b, x0=None, restart=50, tol=1e-6, maxit=500):
    """FP64 restarted GMRES with a Python operator matvec(x) -> A x (numpy)."""
    b = _f64(b)
    x = np.zeros(n) if x0 is None else _f64(x0).copy()
    err = []

    def cb(_user, pin, pout):
        try:
            v = np.ctypeslib.as_array(pin, shape=(n,)).copy()
            np.ctypeslib.as_array(pout, shape=(n,))[:] = matvec(v)
        except Exception as e:  # keep the C loop alive; report after it returns
            err.append(e)
            np.ctypeslib.as_array(pout, shape=(n,))[:] = 0.0

    fn = MATVEC_FN(cb)
    it, rr = C.c_int(), C.c_double()
    _check(lib.bit_gmres(int(n), fn, None, _ptr(b), _ptr(x), int(restart), float(tol),
                         int(maxit), C.byref(it), C.byref(rr)))
    if err:
        raise err[0]
    return x, it.value, rr.value


def eval_multi(ctx: "Context", systems, params, precision_out=None):
    """Potentials of several systems in one persistent launch per precision
    (bit_eval_potential_multi); returns (list of numpy arrays, GPU ms of the
    launches)."""
    k = len(systems)
    outs = [np.zeros(s.n) for s in systems]
    sys_arr = (C.c_void_p * k)(*[s.h.value for s in systems])
    prm_arr = (Params * k)(*params)
    out_arr = (C.c_void_p * k)(*[o.ctypes.data for o in outs])
    ms = C.c_double()
    _check(lib.bit_eval_potential_multi(ctx.h, k, sys_arr, prm_arr, out_arr, 0, C.byref(ms)))
    return outs, ms.value


class Krylov:
    """A second-kind GMRES solve on a System with its Krylov state resident in
    HBM (bit_krylov_*), advanced one matrix-vector product per serve_round."""

    def __init__(self, system: "System", w, f, diag, self_coef=0.0, sigma0=None, restart=50,
                 tol=1e-6, maxit=500, nterms=7, acc=0.6, kappa=0.0, delta2=0.0, mac_sub=None,
                 precision="fp64"):
        self.system = system
        prm = make_params(nterms, acc, kappa, delta2, mac_sub, precision)
        w, f = _f64(w), _f64(f)
        s0 = None if sigma0 is None else _f64(sigma0)
        h = C.c_void_p()
        _check(lib.bit_krylov_create(system.h, C.byref(prm), float(diag), float(self_coef),
                                     _ptr(w), _ptr(f), None if s0 is None else _ptr(s0),
                                     int(restart), float(tol), int(maxit), C.byref(h)))
        self.h = h

    def status(self):
        """(done, iterations, relative residual); done 0 running, 1 converged,
        2 stopped at maxit."""
        d, it, rr = C.c_int(), C.c_int(), C.c_double()
        _check(lib.bit_krylov_status(self.h, C.byref(d), C.byref(it), C.byref(rr)))
        return d.value, it.value, rr.value

    @property
    def done(self) -> bool:
        return self.status()[0] != 0

    def solution(self) -> np.ndarray:
        out = np.zeros(self.system.n)
        _check(lib.bit_krylov_solution(self.h, _ptr(out)))
        return out

    def close(self):
        if getattr(self, "h", None):
            if self.system.ctx.h and self.system.h:
                lib.bit_krylov_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def serve_round(ctx: "Context", systems=(), params=(), solves=()):
    """One serving round: potentials of `systems` (host outputs) and one step of
    every running solve in `solves`; returns (outputs, GPU ms of the round)."""
    k, m = len(systems), len(solves)
    outs = [np.zeros(s.n) for s in systems]
    sys_arr = (C.c_void_p * max(k, 1))(*[s.h.value for s in systems])
    prm_arr = (Params * max(k, 1))(*params)
    out_arr = (C.c_void_p * max(k, 1))(*[o.ctypes.data for o in outs])
    kry_arr = (C.c_void_p * max(m, 1))(*[q.h.value for q in solves])
    ms = C.c_double()
    _check(lib.bit_serve_round(ctx.h, k, sys_arr, prm_arr, out_arr, m, kry_arr, C.byref(ms)))
    return outs, ms.value


class System:
    """One particle system resident on the device (tree + sorted sources)."""

    def __init__(self, ctx: Context, x, y, z, q, maxmemb: int):
        self.ctx = ctx
        h = C.c_void_p()
        if _dev_ptr(x) is not None:  # device-resident float64 arrays
            n = x.n if isinstance(x, DeviceArray) else x.numel()
            ptrs = [C.c_void_p(_dev_ptr(a)) for a in (x, y, z, q)]
            _check(lib.bit_system_create_device(ctx.h, n, *ptrs, int(maxmemb), C.byref(h)))
        else:
            x, y, z, q = _f64(x), _f64(y), _f64(z), _f64(q)
            n = len(x)
            _check(lib.bit_system_create(ctx.h, n, _ptr(x), _ptr(y), _ptr(z), _ptr(q),
                                         int(maxmemb), C.byref(h)))
        self.h = h
        ctx._systems.add(self)
        self.n = n
        self.numtree = lib.bit_system_numtree(h)
        self.maxdepth = lib.bit_system_maxdepth(h)

    def close(self):
        if getattr(self, "h", None):
            if self.ctx.h:  # a closed context already destroyed its systems
                lib.bit_system_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def export_tree(self) -> TreeExport:
        nt, n = self.numtree, self.n
        boxes = np.zeros(nt * 13)
        ch = np.zeros(nt * 8, dtype=np.int32)
        b = np.zeros(nt, dtype=np.int32)
        e = np.zeros(nt, dtype=np.int32)
        p = np.zeros(n, dtype=np.int32)
        _check(lib.bit_system_export_tree(self.h, _ptr(boxes), _ptr(ch), _ptr(b), _ptr(e),
                                          _ptr(p)))
        return TreeExport(nt, boxes.reshape(nt, 13), ch.reshape(nt, 8), b, e, p)

    def moments(self, nterms: int):
        _check(lib.bit_moments(self.h, int(nterms)))
        self._mom_nterms = int(nterms)

    def moments_cube(self) -> np.ndarray:
        p = self._mom_nterms
        out = np.zeros(self.numtree * p ** 3)
        _check(lib.bit_moments_download(self.h, _ptr(out)))
        return out.reshape(self.numtree, p, p, p)

    def potential(self, nterms=6, acc=ACC, kappa=0.0, delta2=0.0, mac_sub=None,
                  precision="fp64", out=None, root=0):
        """+phi at every particle (self excluded), original order.  `out` may be
        a float64 DeviceArray / torch device tensor of n elements (the result
        then stays in HBM).  root: start the traversal at that node (the
        subtree's contribution, the reference's `ind`)."""
        prm = make_params(nterms, acc, kappa, delta2, mac_sub, precision, 0, root)
        if out is not None and _dev_ptr(out) is not None:
            _check(lib.bit_eval_potential(self.h, C.byref(prm), C.c_void_p(_dev_ptr(out)), 1))
            return out
        res = np.zeros(self.n)
        _check(lib.bit_eval_potential(self.h, C.byref(prm), _ptr(res), 0))
        return res

    def potential_noalloc(self, out, nterms=6, acc=ACC, kappa=0.0, delta2=0.0, mac_sub=None,
                          precision="fp64"):
        """+phi into `out` (a float64 DeviceArray / torch device tensor of n
        elements) with no device allocation; needs prepare() with the same
        parameters.  Synchronises the context before returning."""
        prm = make_params(nterms, acc, kappa, delta2, mac_sub, precision)
        ptr = _dev_ptr(out)
        if ptr is None:
            raise ValueError("potential_noalloc needs a device output array")
        _check(lib.bit_eval_potential_noalloc(self.h, C.byref(prm), C.c_void_p(ptr)))
        self.ctx.sync()
        return out

    def field(self, nterms=6, acc=ACC, delta2=DEL, mac_sub=None, kernel_order=0,
              precision="fp64", root=0):
        """E = sum q (x_i - x_j)/R^3 (regularized Coulomb), [n, 3], original order;
        kernel_order >= 1 adds the regularized family's higher terms."""
        prm = make_params(nterms, acc, 0.0, delta2, mac_sub, precision, kernel_order, root)
        res = np.zeros(3 * self.n)
        _check(lib.bit_eval_field(self.h, C.byref(prm), _ptr(res), 0))
        return res.reshape(self.n, 3)

    def potential_at(self, px, py, pz, nterms=6, acc=ACC, kappa=0.0, delta2=0.0, mac_sub=None,
                     precision="fp64", root=0):
        """+phi at arbitrary points (all sources of the subtree at `root`)."""
        prm = make_params(nterms, acc, kappa, delta2, mac_sub, precision, 0, root)
        px, py, pz = _f64(px), _f64(py), _f64(pz)
        out = np.zeros(len(px))
        _check(lib.bit_eval_potential_points(self.h, C.byref(prm), len(px), _ptr(px), _ptr(py),
                                             _ptr(pz), _ptr(out)))
        return out

    def field_at(self, px, py, pz, nterms=6, acc=ACC, delta2=DEL, mac_sub=None, kernel_order=0,
                 precision="fp64", root=0):
        """Regularized Coulomb field at arbitrary points, [m, 3]."""
        prm = make_params(nterms, acc, 0.0, delta2, mac_sub, precision, kernel_order, root)
        px, py, pz = _f64(px), _f64(py), _f64(pz)
        out = np.zeros(3 * len(px))
        _check(lib.bit_eval_field_points(self.h, C.byref(prm), len(px), _ptr(px), _ptr(py),
                                         _ptr(pz), _ptr(out)))
        return out.reshape(-1, 3)

    def direct(self, kappa, delta2, tidx):
        """FP64 GPU direct sum at particle targets tidx (original indices)."""
        t = np.ascontiguousarray(tidx, dtype=np.int32)
        out = np.zeros(len(t))
        _check(lib.bit_direct_potential(self.h, float(kappa), float(delta2), len(t), _ptr(t),
                                        _ptr(out)))
        return out

    def direct_field(self, delta2, tidx, kernel_order=0):
        """FP64 GPU direct-sum field at particle targets tidx (own particle
        skipped), [len(tidx), 3]."""
        t = np.ascontiguousarray(tidx, dtype=np.int32)
        out = np.zeros(3 * len(t))
        _check(lib.bit_direct_field(self.h, float(delta2), int(kernel_order), len(t), _ptr(t),
                                    _ptr(out)))
        return out.reshape(-1, 3)

    def direct_field_at(self, px, py, pz, delta2, kernel_order=0):
        """FP64 GPU direct-sum field at arbitrary points (every source), [m, 3]."""
        px, py, pz = _f64(px), _f64(py), _f64(pz)
        out = np.zeros(3 * len(px))
        _check(lib.bit_direct_field_points(self.h, float(delta2), int(kernel_order), len(px),
                                           _ptr(px), _ptr(py), _ptr(pz), _ptr(out)))
        return out.reshape(-1, 3)

    def lists(self, nterms=6, acc=ACC, delta2=0.0, mac_sub=None, root=0):
        """Batch interaction lists for the MAC parameters: dict of far/near
        offsets [nbatch+1], masks (uint64) and nodes."""
        prm = make_params(nterms, acc, 0.0, delta2, mac_sub, "fp64", 0, root)
        nb, nf, nn = C.c_int(), C.c_int(), C.c_int()
        _check(lib.bit_system_export_lists(self.h, C.byref(prm), C.byref(nb), C.byref(nf),
                                           C.byref(nn), None, None, None, None, None, None))
        fo = np.zeros(nb.value + 1, dtype=np.int32)
        no = np.zeros(nb.value + 1, dtype=np.int32)
        fm = np.zeros(max(nf.value, 1), dtype=np.uint64)
        fn = np.zeros(max(nf.value, 1), dtype=np.int32)
        nm = np.zeros(max(nn.value, 1), dtype=np.uint64)
        nd = np.zeros(max(nn.value, 1), dtype=np.int32)
        _check(lib.bit_system_export_lists(self.h, C.byref(prm), C.byref(nb), C.byref(nf),
                                           C.byref(nn), _ptr(fo), _ptr(fm), _ptr(fn), _ptr(no),
                                           _ptr(nm), _ptr(nd)))
        return dict(far_off=fo, far_mask=fm[:nf.value], far_node=fn[:nf.value], near_off=no,
                    near_mask=nm[:nn.value], near_node=nd[:nn.value])

    def set_charges(self, q):
        """New charges for the same positions (tree and lists are kept)."""
        if _dev_ptr(q) is not None:
            _check(lib.bit_system_set_charges(self.h, C.c_void_p(_dev_ptr(q)), 1))
        else:
            q = _f64(q)
            _check(lib.bit_system_set_charges(self.h, _ptr(q), 0))

    def solve_second_kind(self, w, f, diag, self_coef=0.0, sigma0=None, restart=50, tol=1e-6,
                          maxit=500, nterms=7, acc=0.6, kappa=0.0, delta2=0.0, mac_sub=None,
                          precision="fp64"):
        """GMRES solve of diag*s + S_w s (+ self*w*s) = f with treecode
        matvecs; returns (sigma, iterations, relative residual)."""
        prm = make_params(nterms, acc, kappa, delta2, mac_sub, precision)
        w, f = _f64(w), _f64(f)
        sig = np.zeros(self.n) if sigma0 is None else _f64(sigma0).copy()
        it, rr = C.c_int(), C.c_double()
        _check(lib.bit_solve_second_kind(self.h, C.byref(prm), float(diag), float(self_coef),
                                         _ptr(w), _ptr(f), _ptr(sig), int(restart), float(tol),
                                         int(maxit), C.byref(it), C.byref(rr)))
        return sig, it.value, rr.value

    def prepare(self, nterms=6, acc=ACC, kappa=0.0, delta2=0.0, mac_sub=None,
                precision="fp64"):
        """Moments + interaction lists for these parameters (no evaluation)."""
        prm = make_params(nterms, acc, kappa, delta2, mac_sub, precision)
        _check(lib.bit_system_prepare(self.h, C.byref(prm)))

    def stats(self) -> dict:
        s = Stats()
        _check(lib.bit_system_stats(self.h, C.byref(s)))
        return s.as_dict()
