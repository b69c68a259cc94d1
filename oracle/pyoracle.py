"""Python binding of the CPU oracle (liboracle.so, built from pbf_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as
the checker and by bench.py's cpu_baseline leg.  The product path never
imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import sys

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
_PKG_DIR = os.path.join(os.path.dirname(_HERE), "pbd-simulation-framework-cuda_amd")
if _PKG_DIR not in sys.path:
    sys.path.insert(0, _PKG_DIR)

from pbdfluid.capi import PbdParams  # noqa: E402  (struct layout only)

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_up = C.POINTER(C.c_uint32)


class OrcDerived(C.Structure):
    _fields_ = [
        ("variant", C.c_int), ("cap", C.c_uint32), ("kw", C.c_uint32),
        ("diameter", C.c_double), ("h_spiky", C.c_double), ("h_poly", C.c_double), ("dq", C.c_double),
        ("w_dq", C.c_double), ("rho0", C.c_double), ("k_scorr", C.c_double), ("n_scorr", C.c_int),
        ("eps", C.c_double), ("stiffness", C.c_double), ("c_visc", C.c_double), ("e_vort", C.c_double),
        ("dt", C.c_double), ("dt_f", C.c_float), ("gravity_f", C.c_float),
        ("bmin_f", C.c_float * 3), ("bmax_f", C.c_float * 3), ("vdamp_f", C.c_float), ("pdamp_f", C.c_float),
        ("cell_f", C.c_float), ("inv_cell_f", C.c_float), ("origin_f", C.c_float * 3),
        ("dims", C.c_uint32 * 3), ("table_size", C.c_uint32), ("collision_type", C.c_int),
        ("solver_iters", C.c_uint32), ("stab_iters", C.c_uint32), ("max_batches", C.c_uint32),
        ("g_apply_forces", C.c_int), ("g_find_contacts", C.c_int), ("g_find_neighbours", C.c_int),
        ("g_lambda", C.c_int), ("g_delta", C.c_int), ("g_apply_delta", C.c_int), ("g_update", C.c_int),
        ("g_omega", C.c_int), ("g_vorticity", C.c_int), ("g_viscosity", C.c_int),
        ("mu_s", C.c_double), ("mu_k", C.c_double),
    ]


class OrcIo(C.Structure):
    _fields_ = [
        ("n", C.c_uint32),
        ("x4", _fp), ("p4", _fp), ("v4", _fp), ("c4", _fp), ("f4", _fp),
        ("lam", _dp), ("dp4", _dp),
        ("out_x4", _dp), ("out_p4", _dp), ("out_v4", _dp), ("out_f4", _dp), ("out_density", _dp),
        ("out_omega3", _dp),
        ("out_keys_unsorted", _up), ("out_keys", _up), ("out_ids", _up), ("out_starts", _up), ("out_ends", _up),
        ("out_nbr_count", _up), ("out_contact_count", _up),
        ("min_cutoff_margin", C.c_double),
        ("phase", _up),
    ]


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle not built: {LIB_PATH} missing (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        L.orc_derive.argtypes = [C.POINTER(PbdParams), C.POINTER(OrcDerived)]
        L.orc_derive.restype = C.c_int
        L.orc_expand_bits.argtypes = [C.c_uint32]
        L.orc_expand_bits.restype = C.c_uint32
        L.orc_morton.argtypes = [C.c_uint32] * 3
        L.orc_morton.restype = C.c_uint32
        L.orc_cell_key.argtypes = [C.POINTER(OrcDerived), _fp, _up]
        L.orc_cell_key.restype = C.c_uint32
        L.orc_w.argtypes = [C.POINTER(OrcDerived), C.c_double]
        L.orc_w.restype = C.c_double
        L.orc_grad.argtypes = [C.POINTER(OrcDerived), _dp, _dp]
        L.orc_grad.restype = None
        L.orc_lattice_rest_density.argtypes = [C.POINTER(OrcDerived), C.c_double]
        L.orc_lattice_rest_density.restype = C.c_double
        L.orc_predict.argtypes = [C.POINTER(OrcDerived), C.c_uint32, _fp, _fp, _fp, _fp]
        L.orc_predict.restype = None
        L.orc_sort.argtypes = [C.c_uint32, _up, _up, _up]
        L.orc_sort.restype = None
        L.orc_cell_ranges.argtypes = [C.c_uint32, _up, C.c_uint32, _up, _up]
        L.orc_cell_ranges.restype = None
        L.orc_neighbours.argtypes = [C.POINTER(OrcDerived), C.c_uint32, _up, _up, _up, C.c_uint32, _up, _up]
        L.orc_neighbours.restype = C.c_uint32
        L.orc_step.argtypes = [C.POINTER(PbdParams), C.POINTER(OrcIo)]
        L.orc_step.restype = C.c_int
        L.orc64_forward.argtypes = [C.POINTER(PbdParams), C.c_double, C.c_double, C.c_uint32, C.c_uint32,
                                    _dp, _dp, _dp, _up, _up]
        L.orc64_forward.restype = C.c_int
        L.orc64_gradient.argtypes = [C.POINTER(PbdParams), C.c_double, C.c_double, C.c_uint32, C.c_uint32,
                                     _dp, _dp, _dp, _up, _dp, _dp, _dp, _dp, _dp, _dp]
        L.orc64_gradient.restype = C.c_int
        L.orc_threads.argtypes = []
        L.orc_threads.restype = C.c_int
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_set_threads.restype = None
        _lib = L
    return _lib


def _ptr(a, t):
    return a.ctypes.data_as(t) if a is not None else None


def derive(p: PbdParams) -> OrcDerived:
    d = OrcDerived()
    rc = lib().orc_derive(C.byref(p), C.byref(d))
    if rc != 0:
        raise ValueError(f"orc_derive failed ({rc})")
    return d


def w(d: OrcDerived, r: float) -> float:
    return lib().orc_w(C.byref(d), float(r))


def grad(d: OrcDerived, r) -> np.ndarray:
    rv = np.ascontiguousarray(r, np.float64)
    g = np.zeros(3, np.float64)
    lib().orc_grad(C.byref(d), _ptr(rv, _dp), _ptr(g, _dp))
    return g


def cell_key(d: OrcDerived, pos) -> tuple[int, tuple[int, int, int]]:
    p = np.ascontiguousarray(np.asarray(pos, np.float32)[:3])
    ijk = np.zeros(3, np.uint32)
    k = lib().orc_cell_key(C.byref(d), _ptr(p, _fp), _ptr(ijk, _up))
    return k, tuple(int(v) for v in ijk)


def sort(keys: np.ndarray):
    keys = np.ascontiguousarray(keys, np.uint32)
    n = keys.shape[0]
    ko = np.zeros(n, np.uint32)
    io = np.zeros(n, np.uint32)
    lib().orc_sort(n, _ptr(keys, _up), _ptr(ko, _up), _ptr(io, _up))
    return ko, io


def cell_ranges(keys_sorted: np.ndarray, table_size: int):
    keys_sorted = np.ascontiguousarray(keys_sorted, np.uint32)
    st = np.zeros(table_size, np.uint32)
    en = np.zeros(table_size, np.uint32)
    lib().orc_cell_ranges(keys_sorted.shape[0], _ptr(keys_sorted, _up), table_size, _ptr(st, _up), _ptr(en, _up))
    return st, en


def neighbours(d: OrcDerived, keys_sorted, starts, ends, i: int):
    out = np.zeros(d.cap, np.uint32)
    nc = C.c_uint32(0)
    cnt = lib().orc_neighbours(C.byref(d), keys_sorted.shape[0], _ptr(keys_sorted, _up), _ptr(starts, _up),
                               _ptr(ends, _up), i, _ptr(out, _up), C.byref(nc))
    return out[:cnt].copy(), nc.value


class State:
    """FP32 particle state of the oracle (the reference's surfaces + persistent arrays)."""

    def __init__(self, x4, v4=None, c4=None, f4=None, p4=None, phase=None):
        x4 = np.ascontiguousarray(x4, np.float32).reshape(-1, 4)
        n = x4.shape[0]
        self.n = n
        self.x4 = x4.copy()
        self.p4 = (x4.copy() if p4 is None else np.ascontiguousarray(p4, np.float32).reshape(-1, 4).copy())
        self.v4 = (np.zeros((n, 4), np.float32) if v4 is None
                   else np.ascontiguousarray(v4, np.float32).reshape(-1, 4).copy())
        self.c4 = (np.ones((n, 4), np.float32) if c4 is None
                   else np.ascontiguousarray(c4, np.float32).reshape(-1, 4).copy())
        self.f4 = (np.zeros((n, 4), np.float32) if f4 is None
                   else np.ascontiguousarray(f4, np.float32).reshape(-1, 4).copy())
        self.lam = np.zeros(n, np.float64)
        self.dp4 = np.zeros((n, 4), np.float64)
        self.phase = None if phase is None else np.ascontiguousarray(phase, np.uint32).copy()


def step(p: PbdParams, st: State, want_outputs: bool = True) -> dict:
    """One oracle step (Fluid::compute); updates `st` in place, returns the FP64 / integer outputs."""
    d = derive(p)
    n = st.n
    io = OrcIo()
    io.n = n
    io.x4, io.p4, io.v4 = _ptr(st.x4, _fp), _ptr(st.p4, _fp), _ptr(st.v4, _fp)
    io.c4, io.f4 = _ptr(st.c4, _fp), _ptr(st.f4, _fp)
    io.lam, io.dp4 = _ptr(st.lam, _dp), _ptr(st.dp4, _dp)
    io.phase = _ptr(st.phase, _up)
    out = {}
    if want_outputs:
        out = dict(
            x4=np.zeros((n, 4)), p4=np.zeros((n, 4)), v4=np.zeros((n, 4)), f4=np.zeros((n, 4)),
            density=np.zeros(n), omega3=np.zeros((n, 3)),
            keys_unsorted=np.zeros(n, np.uint32), keys=np.zeros(n, np.uint32), ids=np.zeros(n, np.uint32),
            starts=np.zeros(d.table_size, np.uint32), ends=np.zeros(d.table_size, np.uint32),
            nbr_count=np.zeros(n, np.uint32), contact_count=np.zeros(n, np.uint32),
        )
        io.out_x4, io.out_p4 = _ptr(out["x4"], _dp), _ptr(out["p4"], _dp)
        io.out_v4, io.out_f4 = _ptr(out["v4"], _dp), _ptr(out["f4"], _dp)
        io.out_density, io.out_omega3 = _ptr(out["density"], _dp), _ptr(out["omega3"], _dp)
        io.out_keys_unsorted = _ptr(out["keys_unsorted"], _up)
        io.out_keys, io.out_ids = _ptr(out["keys"], _up), _ptr(out["ids"], _up)
        io.out_starts, io.out_ends = _ptr(out["starts"], _up), _ptr(out["ends"], _up)
        io.out_nbr_count = _ptr(out["nbr_count"], _up)
        io.out_contact_count = _ptr(out["contact_count"], _up)
    rc = lib().orc_step(C.byref(p), C.byref(io))
    if rc != 0:
        raise RuntimeError(f"orc_step failed ({rc})")
    out["lambda"] = st.lam.copy()
    out["dp4"] = st.dp4.copy()
    out["min_cutoff_margin"] = io.min_cutoff_margin
    return out


def _state64(a, n):
    if a is None:
        return np.zeros((n, 4), np.float64)
    a = np.asarray(a, np.float64).reshape(n, -1)
    out = np.zeros((n, 4), np.float64)
    out[:, :3] = a[:, :3]
    return out


def forward64(p: PbdParams, x, v=None, f=None, nsub: int = 1, rho0: float = 0.0, eps: float = -1.0, phase=None):
    """FP64 forward of nsub steps (textbook, collisionType 0) from state x, v, f
    (and initial phases, None: all fluid).

    Returns (x, v, f, orig): the final state in its sorted order and, per final
    slot, the index of the particle in the initial state.  rho0 <= 0 / eps < 0
    select the values derived from p."""
    x = np.asarray(x)
    n = x.shape[0]
    X, V, F = _state64(x, n), _state64(v, n), _state64(f, n)
    orig = np.arange(n, dtype=np.uint32)
    ph = None if phase is None else np.ascontiguousarray(phase, dtype=np.uint32).copy()
    rc = lib().orc64_forward(C.byref(p), float(rho0), float(eps), n, int(nsub), _ptr(X, _dp), _ptr(V, _dp),
                             _ptr(F, _dp), _ptr(orig, _up), None if ph is None else _ptr(ph, _up))
    if rc != 0:
        raise RuntimeError(f"orc64_forward failed ({rc})")
    return X[:, :3].copy(), V[:, :3].copy(), F[:, :3].copy(), orig


def gradient64(p: PbdParams, x, v, gx, gv, f=None, nsub: int = 1, rho0: float = 0.0, eps: float = -1.0,
               phase=None):
    """Reverse mode of forward64: given dL/dx, dL/dv of the final state (final
    order), returns dict(x0, v0: (n, 3) in the initial order, rho0, eps)."""
    x = np.asarray(x)
    n = x.shape[0]
    X, V, F = _state64(x, n), _state64(v, n), _state64(f, n)
    GX, GV = _state64(gx, n), _state64(gv, n)
    ph = None if phase is None else np.ascontiguousarray(phase, dtype=np.uint32)
    gx0 = np.zeros((n, 4))
    gv0 = np.zeros((n, 4))
    gr = C.c_double(0.0)
    ge = C.c_double(0.0)
    rc = lib().orc64_gradient(C.byref(p), float(rho0), float(eps), n, int(nsub), _ptr(X, _dp), _ptr(V, _dp),
                              _ptr(F, _dp), None if ph is None else _ptr(ph, _up), _ptr(GX, _dp), _ptr(GV, _dp),
                              _ptr(gx0, _dp), _ptr(gv0, _dp), C.byref(gr), C.byref(ge))
    if rc != 0:
        raise RuntimeError(f"orc64_gradient failed ({rc})")
    return dict(x0=gx0[:, :3].copy(), v0=gv0[:, :3].copy(), rho0=gr.value, eps=ge.value)


def set_threads(n: int) -> None:
    lib().orc_set_threads(int(n))


def threads() -> int:
    return lib().orc_threads()
