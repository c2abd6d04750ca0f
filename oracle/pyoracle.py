"""ctypes loader of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use
this module, and only as the checker.  It loads oracle/build/liboracle.so
(the restatement in mhd_oracle.hpp) and, where built, the reference-header
wrapper oracle/_ref/libmhdref_{fp64,fp32}.so.  Arrays are AoS (n, 9).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

ORACLE_DIR = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(ORACLE_DIR, "build", "liboracle.so")
REF_DIR = os.path.join(ORACLE_DIR, "_ref")
ORACLE_BIN = {p: os.path.join(ORACLE_DIR, "build", f"mhd2d_glm_oracle_{p}") for p in ("fp64", "fp32")}
REF_BIN = {p: os.path.join(REF_DIR, f"mhd2d_glm_cpu_{p}") for p in ("fp64", "fp32")}

_SUF = {np.float64: "f64", np.float32: "f32"}
_CT = {np.float64: ctypes.c_double, np.float32: ctypes.c_float}


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


class Oracle:
    def __init__(self, path: str = LIB):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run `make -C {ORACLE_DIR} oracle`")
        self.lib = ctypes.CDLL(path)
        P, I, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        for dt, suf in _SUF.items():
            R = _CT[dt]
            f = getattr(self.lib, f"orc_glm_flux_batch_{suf}")
            f.argtypes = [I, I, P, P, R, I, D, P, P]
            f.restype = I
            for name in ("cons_to_prim_batch", "prim_to_cons_batch"):
                g = getattr(self.lib, f"orc_{name}_{suf}")
                g.argtypes = [I, P, P]
                g.restype = None
            g = getattr(self.lib, f"orc_fast_speed_batch_{suf}")
            g.argtypes = [I, P, I, P]
            g.restype = None
            g = getattr(self.lib, f"orc_reconstruct_batch_{suf}")
            g.argtypes = [I, P, P, P, P, P]
            g.restype = None
            g = getattr(self.lib, f"orc_init_{suf}")
            g.argtypes = [ctypes.c_char_p, I, I, P]
            g.restype = None
            g = getattr(self.lib, f"orc_init_params_{suf}")
            g.argtypes = [ctypes.c_char_p, P, I, I, P]
            g.restype = None
            g = getattr(self.lib, f"orc_max_signal_speed_{suf}")
            g.argtypes = [I, I, P]
            g.restype = D
            g = getattr(self.lib, f"orc_compute_rhs_{suf}")
            g.argtypes = [I, I, P, R, R, I, I, D, P, P]
            g.restype = None
            g = getattr(self.lib, f"orc_advance_{suf}")
            g.argtypes = [I, I, P, I, R, R, R, I, I, D, P, P, P, P, P]
            g.restype = I

    def _fn(self, name, dtype):
        return getattr(self.lib, f"orc_{name}_{_SUF[np.dtype(dtype).type]}")

    def glm_flux(self, UL, UR, ch, direction=0, hlld=True, dtype=np.float64, tol=1e-4):
        a = np.ascontiguousarray(UL, dtype=dtype)
        b = np.ascontiguousarray(UR, dtype=dtype)
        n = a.shape[0]
        F = np.empty((n, 9), dtype=dtype)
        st = np.empty(n, dtype=np.int32)
        self._fn("glm_flux_batch", dtype)(n, direction, _p(a), _p(b), ch, int(hlld), tol, _p(F), _p(st))
        return F, st

    def cons_to_prim(self, U, dtype=np.float64):
        a = np.ascontiguousarray(U, dtype=dtype)
        W = np.empty_like(a)
        self._fn("cons_to_prim_batch", dtype)(a.shape[0], _p(a), _p(W))
        return W

    def prim_to_cons(self, W, dtype=np.float64):
        a = np.ascontiguousarray(W, dtype=dtype)
        U = np.empty_like(a)
        self._fn("prim_to_cons_batch", dtype)(a.shape[0], _p(a), _p(U))
        return U

    def fast_speed(self, W, direction=0, dtype=np.float64):
        a = np.ascontiguousarray(W, dtype=dtype)
        out = np.empty(a.shape[0], dtype=dtype)
        self._fn("fast_speed_batch", dtype)(a.shape[0], _p(a), direction, _p(out))
        return out

    def reconstruct(self, Pm, P0, Pp, dtype=np.float64):
        a, b, c = (np.ascontiguousarray(x, dtype=dtype) for x in (Pm, P0, Pp))
        Pl, Pr = np.empty_like(a), np.empty_like(a)
        self._fn("reconstruct_batch", dtype)(a.shape[0], _p(a), _p(b), _p(c), _p(Pl), _p(Pr))
        return Pl, Pr

    def init(self, problem, nx, ny, dtype=np.float64, params=None):
        """Initial state as AoS (ny, nx, 9); `params` for "otv" (5 values)."""
        u = np.empty((ny, nx, 9), dtype=dtype)
        if params is None:
            self._fn("init", dtype)(problem.encode(), nx, ny, _p(u))
        else:
            p = np.ascontiguousarray(params, dtype=np.float64)
            assert problem == "otv" and p.size == 5
            self._fn("init_params", dtype)(problem.encode(), _p(p), nx, ny, _p(u))
        return u

    def max_signal_speed(self, u_aos):
        ny, nx, _ = u_aos.shape
        a = np.ascontiguousarray(u_aos)
        return self._fn("max_signal_speed", a.dtype)(nx, ny, _p(a))

    def compute_rhs(self, u_aos, ch, alpha=0.18, muscl=True, hlld=True, tol=1e-4):
        ny, nx, _ = u_aos.shape
        a = np.ascontiguousarray(u_aos)
        rhs = np.empty_like(a)
        cnt = np.zeros(4, dtype=np.uint64)
        self._fn("compute_rhs", a.dtype)(nx, ny, _p(a), ch, alpha, int(muscl), int(hlld), tol,
                                         _p(rhs), _p(cnt))
        return rhs, cnt

    def advance(self, u_aos, max_steps, t_end=0.5, cfl=0.25, alpha=0.18, muscl=True,
                hlld=True, tol=1e-4, t0=0.0):
        """Advances a copy of u_aos; returns (u, info dict)."""
        ny, nx, _ = u_aos.shape
        a = np.array(u_aos, copy=True, order="C")
        dt = a.dtype.type
        t = np.array([t0], dtype=dt)
        steps = np.zeros(1, dtype=np.int32)
        cnt = np.zeros(4, dtype=np.uint64)
        bad = np.full(2, -1, dtype=np.int32)
        ch = np.zeros(1, dtype=dt)
        code = self._fn("advance", dt)(nx, ny, _p(a), max_steps, t_end, cfl, alpha, int(muscl),
                                       int(hlld), tol, _p(t), _p(steps), _p(cnt), _p(bad), _p(ch))
        return a, {"code": code, "t": t[0], "steps": int(steps[0]), "counters": cnt,
                   "bad": (int(bad[0]), int(bad[1])), "ch": ch[0]}


    def refmixed_f32(self, name, *arrays, direction=None):
        """FP32 stage function of the reference-mixed restatement over a
        batch: cons_to_prim, prim_to_cons, fast_speed (direction) or recon
        (Pm, P0, Pp -> Pl, Pr)."""
        P, I = ctypes.c_void_p, ctypes.c_int
        a = [np.ascontiguousarray(x, dtype=np.float32) for x in arrays]
        n = a[0].shape[0]
        f = getattr(self.lib, f"orc_refmixed_{name}_batch")
        if name == "fast_speed":
            out = np.empty(n, dtype=np.float32)
            f.argtypes = [I, P, I, P]
            f(n, _p(a[0]), int(direction), _p(out))
            return out
        if name == "recon":
            Pl, Pr = np.empty_like(a[0]), np.empty_like(a[0])
            f.argtypes = [I, P, P, P, P, P]
            f(n, _p(a[0]), _p(a[1]), _p(a[2]), _p(Pl), _p(Pr))
            return Pl, Pr
        out = np.empty_like(a[0])
        f.argtypes = [I, P, P]
        f(n, _p(a[0]), _p(out))
        return out

    REGIONS = {"recovery": 1, "cfl": 2, "recon": 4, "rhs": 8, "glm": 16}

    def refmixed_advance(self, u_aos, max_steps, t_end=0.5, cfl=0.25, alpha=0.18, hlld=True,
                         tol=1e-4, t0=0.0, regions=None):
        """The reference's stage-level mixed build (MUSCL), FP64 state:
        advances a copy of u_aos; returns (u, info dict) as advance().
        `regions` (names of REGIONS) puts only those regions in float; None:
        all five (the mixed build), (): none (the FP64 step)."""
        mask = 31 if regions is None else sum(self.REGIONS[r] for r in regions)
        ny, nx, _ = u_aos.shape
        a = np.array(u_aos, dtype=np.float64, copy=True, order="C")
        t = np.array([t0], dtype=np.float64)
        steps = np.zeros(1, dtype=np.int32)
        cnt = np.zeros(4, dtype=np.uint64)
        bad = np.full(2, -1, dtype=np.int32)
        ch = np.zeros(1, dtype=np.float64)
        f = self.lib.orc_refmixed_advance_regions
        D, I, P = ctypes.c_double, ctypes.c_int, ctypes.c_void_p
        f.argtypes = [I, I, P, I, D, D, D, I, D, P, P, P, P, P, I]
        f.restype = I
        code = f(nx, ny, _p(a), max_steps, t_end, cfl, alpha, int(hlld), tol, _p(t), _p(steps),
                 _p(cnt), _p(bad), _p(ch), mask)
        return a, {"code": code, "t": t[0], "steps": int(steps[0]), "counters": cnt,
                   "bad": (int(bad[0]), int(bad[1])), "ch": ch[0]}


class RefHeaders:
    """The reference's own physics functions (oracle/_ref, when built)."""

    def __init__(self, precision="fp64"):
        path = os.path.join(REF_DIR, f"libmhdref_{precision}.so")
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        self.lib = ctypes.CDLL(path)
        self.dtype = np.float64 if precision == "fp64" else np.float32
        R = _CT[self.dtype]
        P, I = ctypes.c_void_p, ctypes.c_int
        self.lib.ref_glm_flux_batch.argtypes = [I, I, P, P, R, I, P, P]
        self.lib.ref_cons_to_prim_batch.argtypes = [I, P, P]
        self.lib.ref_prim_to_cons_batch.argtypes = [I, P, P]
        self.lib.ref_fast_speed_batch.argtypes = [I, P, I, P]
        assert self.lib.ref_real_size() == np.dtype(self.dtype).itemsize

    def glm_flux(self, UL, UR, ch, direction=0, hlld=True):
        a = np.ascontiguousarray(UL, dtype=self.dtype)
        b = np.ascontiguousarray(UR, dtype=self.dtype)
        n = a.shape[0]
        F = np.empty((n, 9), dtype=self.dtype)
        st = np.empty(n, dtype=np.int32)
        self.lib.ref_glm_flux_batch(n, direction, _p(a), _p(b), ch, int(hlld), _p(F), _p(st))
        return F, st

    def cons_to_prim(self, U):
        a = np.ascontiguousarray(U, dtype=self.dtype)
        W = np.empty_like(a)
        self.lib.ref_cons_to_prim_batch(a.shape[0], _p(a), _p(W))
        return W

    def prim_to_cons(self, W):
        a = np.ascontiguousarray(W, dtype=self.dtype)
        U = np.empty_like(a)
        self.lib.ref_prim_to_cons_batch(a.shape[0], _p(a), _p(U))
        return U

    def fast_speed(self, W, direction=0):
        a = np.ascontiguousarray(W, dtype=self.dtype)
        out = np.empty(a.shape[0], dtype=self.dtype)
        self.lib.ref_fast_speed_batch(a.shape[0], _p(a), direction, _p(out))
        return out


def soa_to_aos(u_soa):
    """(9, ny, nx) -> (ny, nx, 9)."""
    return np.ascontiguousarray(np.moveaxis(u_soa, 0, -1))


def aos_to_soa(u_aos):
    """(ny, nx, 9) -> (9, ny, nx)."""
    return np.ascontiguousarray(np.moveaxis(u_aos, -1, 0))
