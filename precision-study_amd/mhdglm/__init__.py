"""Python binding of the MI355X MHD-GLM solver (ctypes over include/mhd_glm.h).

Mirrors the reference's interfaces for the 2D MHD-GLM path:
  * :class:`Solver` -- the time loop of mhd2d_glm_cpu.cpp:513-570 behind the
    C ABI (create / set_state / advance / get_state);
  * :func:`glm_flux`, :func:`cons_to_prim`, :func:`fast_speed` -- the
    by-array physics functions of include/mhd/mhd_glm_common.hpp and
    include/mhd/mhd_common.hpp, batched on the device;
  * :func:`run_cli` -- the command-line drop-in (argv, OUTDIR, .dat).

The library is loaded from this package's ``lib/`` directory (built by
``make -C precision-study_amd``).  There is no CPU fallback: a missing
library or device raises.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass

import numpy as np

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# The in-tree build; MHDGLM_LIB names another build of the same library.
LIB_PATH = os.environ.get("MHDGLM_LIB") or os.path.join(PKG_ROOT, "lib", "libmhdglm.so")
CLI_PATH = os.path.join(PKG_ROOT, "bin", "mhd2d_glm_amd")

PREC_FP64 = 0
PREC_FP32 = 1
PREC_MIXED = 2
PREC_MIXED_KAHAN = 3
PREC_FP64_FMA = 4
PREC_REF_MIXED = 5
PREC_MIXED_RECON64 = 6
PREC_MIXED_FLUX64 = 7
PREC_MIXED_FLUX64_NARROW = 8
PRECISIONS = {"fp64": PREC_FP64, "fp32": PREC_FP32, "mixed": PREC_MIXED,
              "mixed_kahan": PREC_MIXED_KAHAN, "fp64_fma": PREC_FP64_FMA,
              "ref_mixed": PREC_REF_MIXED, "mixed_recon64": PREC_MIXED_RECON64,
              "mixed_flux64": PREC_MIXED_FLUX64, "mixed_flux64_narrow": PREC_MIXED_FLUX64_NARROW}
# Element type of host state buffers.
DTYPES = {p: (np.float32 if p == PREC_FP32 else np.float64) for p in PRECISIONS.values()}

# Public symbols of include/mhd_glm.h.
EXPORTED = (
    "mhd_create", "mhd_destroy", "mhd_last_error", "mhd_problem_state",
    "mhd_set_state", "mhd_init_problem", "mhd_advance", "mhd_get_status",
    "mhd_get_state", "mhd_device_state", "mhd_stream", "mhd_enqueue_steps",
    "mhd_time_kernels", "mhd_checkpoint_save", "mhd_checkpoint_load", "mhd_glm_flux_batch", "mhd_cons_to_prim_batch",
    "mhd_fast_speed_batch", "mhd_problem_state_ex",
    "mhd_ensemble_create", "mhd_ensemble_destroy", "mhd_ensemble_last_error",
    "mhd_ensemble_set_state", "mhd_ensemble_init_problem", "mhd_ensemble_advance",
    "mhd_ensemble_get_status", "mhd_ensemble_get_state", "mhd_ensemble_enqueue_steps",
    "mhd_ensemble_time_kernels", "mhd_autotuned_tile", "mhd_autotuned_fuse",
    "mhd_selector_configure", "mhd_selector_stats", "mhd_selector_tiles", "mhd_tile_features",
    "mhd_pair_tile_errors", "mhd_mlp_param_count", "mhd_mlp_forward", "mhd_mlp_gradients",
    "mhd_mlp_train", "mhd_run_noalloc",
)


class MhdError(RuntimeError):
    """A non-zero return of the C ABI (code = reference exit code)."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"mhd error {code}: {msg}")
        self.code = code


class Cfg(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int),
        ("cfl", ctypes.c_double), ("alpha", ctypes.c_double),
        ("muscl", ctypes.c_int), ("hlld", ctypes.c_int),
        ("precision", ctypes.c_int), ("hlld_tol", ctypes.c_double),
        ("tile", ctypes.c_int), ("batch_steps", ctypes.c_int),
        ("lowbeta_ratio", ctypes.c_double), ("fallback", ctypes.c_int),
        ("fuse", ctypes.c_int), ("rollback_steps", ctypes.c_int),
    ]


class Status(ctypes.Structure):
    _fields_ = [
        ("code", ctypes.c_int), ("done", ctypes.c_int), ("steps", ctypes.c_int),
        ("t", ctypes.c_double), ("dt", ctypes.c_double), ("ch_last", ctypes.c_double),
        ("x_evaluations", ctypes.c_ulonglong), ("y_evaluations", ctypes.c_ulonglong),
        ("degenerate", ctypes.c_ulonglong), ("nonphysical", ctypes.c_ulonglong),
        ("bad_i", ctypes.c_int), ("bad_j", ctypes.c_int), ("fail_step", ctypes.c_int),
        ("fallback_tiles", ctypes.c_ulonglong), ("fused_redo_steps", ctypes.c_ulonglong),
        ("exact_steps", ctypes.c_ulonglong), ("rollbacks", ctypes.c_int),
        ("restarts", ctypes.c_int),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


_LIB = None


def load(path: str = LIB_PATH) -> ctypes.CDLL:
    """Loads the solver library (raises if it was not built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(path):
        raise FileNotFoundError(
            f"{path} is missing: build it with `make -C {PKG_ROOT}` "
            "(or __graft_entry__.build())")
    lib = ctypes.CDLL(path)
    P, I, D, S = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_size_t
    lib.mhd_create.argtypes = [ctypes.POINTER(Cfg), ctypes.POINTER(P)]
    lib.mhd_create.restype = I
    lib.mhd_destroy.argtypes = [P]
    lib.mhd_destroy.restype = None
    lib.mhd_last_error.argtypes = [P]
    lib.mhd_last_error.restype = ctypes.c_char_p
    lib.mhd_problem_state.argtypes = [ctypes.c_char_p, I, I, I, P, S]
    lib.mhd_problem_state.restype = I
    lib.mhd_set_state.argtypes = [P, P, S]
    lib.mhd_set_state.restype = I
    lib.mhd_init_problem.argtypes = [P, ctypes.c_char_p]
    lib.mhd_init_problem.restype = I
    lib.mhd_advance.argtypes = [P, D, ctypes.c_longlong, ctypes.POINTER(Status)]
    lib.mhd_advance.restype = I
    lib.mhd_get_status.argtypes = [P, ctypes.POINTER(Status)]
    lib.mhd_get_status.restype = I
    lib.mhd_get_state.argtypes = [P, I, P, S]
    lib.mhd_get_state.restype = I
    lib.mhd_device_state.argtypes = [P, ctypes.POINTER(P), ctypes.POINTER(S)]
    lib.mhd_device_state.restype = I
    lib.mhd_stream.argtypes = [P]
    lib.mhd_stream.restype = P
    lib.mhd_enqueue_steps.argtypes = [P, D, I]
    lib.mhd_enqueue_steps.restype = I
    lib.mhd_checkpoint_save.argtypes = [P, ctypes.c_char_p]
    lib.mhd_checkpoint_save.restype = I
    lib.mhd_checkpoint_load.argtypes = [P, ctypes.c_char_p]
    lib.mhd_checkpoint_load.restype = I
    lib.mhd_time_kernels.argtypes = [P, D, I, ctypes.POINTER(ctypes.c_double)]
    lib.mhd_time_kernels.restype = I
    lib.mhd_run_noalloc.argtypes = [ctypes.POINTER(Cfg), P, S, D, ctypes.c_longlong, P,
                                    ctypes.POINTER(Status)]
    lib.mhd_run_noalloc.restype = I
    lib.mhd_glm_flux_batch.argtypes = [I, I, I, P, P, D, I, D, P, P]
    lib.mhd_glm_flux_batch.restype = I
    lib.mhd_cons_to_prim_batch.argtypes = [I, I, P, P]
    lib.mhd_cons_to_prim_batch.restype = I
    lib.mhd_fast_speed_batch.argtypes = [I, I, I, P, P]
    lib.mhd_fast_speed_batch.restype = I
    F = ctypes.c_float
    lib.mhd_selector_configure.argtypes = [P, P, P, P, F, F, F]
    lib.mhd_selector_configure.restype = I
    lib.mhd_selector_stats.argtypes = [P, P]
    lib.mhd_selector_stats.restype = I
    lib.mhd_selector_tiles.argtypes = [P]
    lib.mhd_selector_tiles.restype = I
    lib.mhd_tile_features.argtypes = [P, P]
    lib.mhd_tile_features.restype = I
    lib.mhd_pair_tile_errors.argtypes = [P, P]
    lib.mhd_pair_tile_errors.restype = I
    lib.mhd_mlp_param_count.argtypes = []
    lib.mhd_mlp_param_count.restype = I
    lib.mhd_mlp_forward.argtypes = [P, P, I, P]
    lib.mhd_mlp_forward.restype = I
    lib.mhd_mlp_gradients.argtypes = [P, P, P, P, I, P, P]
    lib.mhd_mlp_gradients.restype = I
    lib.mhd_mlp_train.argtypes = [P, P, P, P, I, I, I, F, ctypes.c_uint, P]
    lib.mhd_mlp_train.restype = I
    lib.mhd_autotuned_tile.argtypes = [I, I, I]
    lib.mhd_autotuned_tile.restype = I
    lib.mhd_autotuned_fuse.argtypes = [I, I, I]
    lib.mhd_autotuned_fuse.restype = I
    lib.mhd_problem_state_ex.argtypes = [ctypes.c_char_p, P, I, I, I, I, P, S]
    lib.mhd_problem_state_ex.restype = I
    lib.mhd_ensemble_create.argtypes = [ctypes.POINTER(Cfg), I, P, P, ctypes.POINTER(P)]
    lib.mhd_ensemble_create.restype = I
    lib.mhd_ensemble_destroy.argtypes = [P]
    lib.mhd_ensemble_destroy.restype = None
    lib.mhd_ensemble_last_error.argtypes = [P]
    lib.mhd_ensemble_last_error.restype = ctypes.c_char_p
    lib.mhd_ensemble_set_state.argtypes = [P, I, P, S]
    lib.mhd_ensemble_set_state.restype = I
    lib.mhd_ensemble_init_problem.argtypes = [P, I, ctypes.c_char_p, P, I]
    lib.mhd_ensemble_init_problem.restype = I
    lib.mhd_ensemble_advance.argtypes = [P, D, ctypes.c_longlong, P]
    lib.mhd_ensemble_advance.restype = I
    lib.mhd_ensemble_get_status.argtypes = [P, I, ctypes.POINTER(Status)]
    lib.mhd_ensemble_get_status.restype = I
    lib.mhd_ensemble_get_state.argtypes = [P, I, I, P, S]
    lib.mhd_ensemble_get_state.restype = I
    lib.mhd_ensemble_enqueue_steps.argtypes = [P, D, I]
    lib.mhd_ensemble_enqueue_steps.restype = I
    lib.mhd_ensemble_time_kernels.argtypes = [P, D, I, ctypes.POINTER(ctypes.c_double)]
    lib.mhd_ensemble_time_kernels.restype = I
    _LIB = lib
    return lib


def _ptr(a: np.ndarray) -> ctypes.c_void_p:
    return ctypes.c_void_p(a.ctypes.data)


def problem_state(problem: str, nx: int, ny: int, precision: int = PREC_FP64) -> np.ndarray:
    """Host-evaluated initial state, shape (9, ny, nx) (no device needed)."""
    lib = load()
    out = np.empty((9, ny, nx), dtype=DTYPES[precision])
    rc = lib.mhd_problem_state(problem.encode(), nx, ny, precision, _ptr(out), nx * ny)
    if rc:
        raise MhdError(rc, f"unknown problem {problem!r} or bad size")
    return out


def problem_state_ex(problem: str, params, nx: int, ny: int,
                     precision: int = PREC_FP64) -> np.ndarray:
    """Initial state of a problem with parameters ("otv": B amplitude and
    the phases of the vx, vy, Bx, By sine terms), shape (9, ny, nx)."""
    lib = load()
    out = np.empty((9, ny, nx), dtype=DTYPES[precision])
    p = np.ascontiguousarray(params, dtype=np.float64)
    rc = lib.mhd_problem_state_ex(problem.encode(), _ptr(p), p.size, nx, ny, precision,
                                  _ptr(out), nx * ny)
    if rc:
        raise MhdError(rc, f"bad problem {problem!r} or parameters")
    return out


@dataclass
class SolverConfig:
    nx: int
    ny: int
    cfl: float = 0.25
    alpha: float = 0.18
    muscl: bool = True
    hlld: bool = True
    precision: int = PREC_FP64
    hlld_tol: float = 1.0e-4
    tile: int = -1
    batch_steps: int = 0
    lowbeta_ratio: float = 0.0
    fallback: int = 0  # MIXED: 0 FP64 fallback on, 1 off, 2 every tile (check mode)
    fuse: int = 0  # stages per launch: 1, 2 (one launch per step); 0 autotuned (tile -1) else 1
    rollback_steps: int = 0  # MIXED: FP64 rollback window/snapshot interval; 0 = 256, < 0 off


class Solver:
    """One solver context (device state + stream) behind the C ABI."""

    def __init__(self, cfg: SolverConfig):
        self.lib = load()
        self.cfg = cfg
        c = _cfg(cfg)
        h = ctypes.c_void_p()
        rc = self.lib.mhd_create(ctypes.byref(c), ctypes.byref(h))
        if rc:
            raise MhdError(rc, "mhd_create failed")
        self.h = h
        self.dtype = DTYPES[cfg.precision]

    def close(self):
        if getattr(self, "h", None):
            self.lib.mhd_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _check(self, rc: int, allow=(0,)) -> int:
        if rc not in allow:
            raise MhdError(rc, self.lib.mhd_last_error(self.h).decode())
        return rc

    def set_state(self, u: np.ndarray):
        u = np.ascontiguousarray(u, dtype=self.dtype).reshape(9, -1)
        assert u.shape[1] == self.cfg.nx * self.cfg.ny
        self._check(self.lib.mhd_set_state(self.h, _ptr(u), u.shape[1]))

    def init_problem(self, problem: str = "ot"):
        self._check(self.lib.mhd_init_problem(self.h, problem.encode()))

    def advance(self, t_end: float, max_steps: int = -1) -> Status:
        st = Status()
        self._check(self.lib.mhd_advance(self.h, t_end, max_steps, ctypes.byref(st)),
                    allow=(0, 2, 3, 4))
        return st

    def status(self) -> Status:
        st = Status()
        self._check(self.lib.mhd_get_status(self.h, ctypes.byref(st)), allow=(0, 2, 3, 4))
        return st

    def get_state(self, which: int = 0) -> np.ndarray:
        out = np.empty((9, self.cfg.ny, self.cfg.nx), dtype=self.dtype)
        self._check(self.lib.mhd_get_state(self.h, which, _ptr(out), self.cfg.nx * self.cfg.ny))
        return out

    def enqueue_steps(self, t_end: float, nsteps: int):
        self._check(self.lib.mhd_enqueue_steps(self.h, t_end, nsteps))

    def device_state(self):
        """(device pointer of state plane 0, elements between planes)."""
        ptr, ld = ctypes.c_void_p(), ctypes.c_size_t()
        self._check(self.lib.mhd_device_state(self.h, ctypes.byref(ptr), ctypes.byref(ld)))
        return ptr.value, ld.value

    def save(self, path: str):
        """Binary checkpoint (control block + raw state planes)."""
        self._check(self.lib.mhd_checkpoint_save(self.h, os.fsencode(path)))

    def load(self, path: str):
        """Restores a checkpoint written by a context of the same shape/form."""
        self._check(self.lib.mhd_checkpoint_load(self.h, os.fsencode(path)))

    # --- learned precision selector (MIXED form) -------------------------
    def selector_tiles(self) -> int:
        return self.lib.mhd_selector_tiles(self.h)

    def tile_features(self) -> np.ndarray:
        out = np.empty((self.selector_tiles(), 16), dtype=np.float32)
        self._check(self.lib.mhd_tile_features(self.h, _ptr(out)))
        return out

    def pair_tile_errors(self) -> np.ndarray:
        out = np.empty(self.selector_tiles(), dtype=np.float32)
        self._check(self.lib.mhd_pair_tile_errors(self.h, _ptr(out)))
        return out

    def configure_selector(self, model, threshold=None, audit_frac=0.01, audit_budget=None):
        """model: dict params/mean/std/threshold/budget[/audit_budget]
        (mhdglm.selector); None: off.  The audit budget defaults to the
        model's audit_budget, else 10 x its label budget."""
        if model is None:
            self._check(self.lib.mhd_selector_configure(self.h, None, None, None, 0.0, 0.0, 0.0))
            return
        p = np.ascontiguousarray(model["params"], dtype=np.float32)
        m = np.ascontiguousarray(model["mean"], dtype=np.float32)
        sd = np.ascontiguousarray(model["std"], dtype=np.float32)
        th = float(model["threshold"] if threshold is None else threshold)
        if audit_budget is None:
            audit_budget = model.get("audit_budget", 10.0 * float(model["budget"]))
        bud = float(audit_budget)
        self._check(self.lib.mhd_selector_configure(self.h, _ptr(p), _ptr(m), _ptr(sd), th,
                                                    float(audit_frac), bud))

    def selector_stats(self) -> dict:
        out = np.zeros(4, dtype=np.uint64)
        self._check(self.lib.mhd_selector_stats(self.h, _ptr(out)))
        return {"audits": int(out[0]), "misses": int(out[1]), "fp64_tile_steps": int(out[2]),
                "steps": int(out[3])}

    def time_kernels(self, t_end: float, nsteps: int):
        """Average ms of (stage-1, stage-2, finish) kernels over nsteps steps."""
        out = (ctypes.c_double * 3)()
        self._check(self.lib.mhd_time_kernels(self.h, t_end, nsteps, out))
        return tuple(out)

    def stream(self) -> int:
        return self.lib.mhd_stream(self.h)


def run_noalloc(cfg: SolverConfig, state_ptr: int, ld: int, t_end: float,
                max_steps: int = -1, stream: int = 0) -> Status:
    """mhd_run_noalloc: advances a device-resident state (pointer to 9 planes
    `ld` elements apart) in place from t=0, with no device allocation."""
    st = Status()
    rc = load().mhd_run_noalloc(ctypes.byref(_cfg(cfg)), ctypes.c_void_p(state_ptr), ld, t_end,
                                max_steps, ctypes.c_void_p(stream), ctypes.byref(st))
    if rc not in (0, 2, 3, 4):
        raise MhdError(rc, "mhd_run_noalloc failed (see stderr)")
    return st


def mlp_param_count() -> int:
    return load().mhd_mlp_param_count()


def mlp_forward(params, x) -> np.ndarray:
    """Device MLP forward: x (n, 16) standardised features -> probabilities."""
    lib = load()
    p = np.ascontiguousarray(params, dtype=np.float32)
    xs = np.ascontiguousarray(x, dtype=np.float32)
    out = np.empty(xs.shape[0], dtype=np.float32)
    rc = lib.mhd_mlp_forward(_ptr(p), _ptr(xs), xs.shape[0], _ptr(out))
    if rc:
        raise MhdError(rc, "mhd_mlp_forward failed")
    return out


def mlp_gradients(params, x, y, w):
    """Device gradients of the weighted BCE: (grad[kParams], loss)."""
    lib = load()
    p = np.ascontiguousarray(params, dtype=np.float32)
    xs = np.ascontiguousarray(x, dtype=np.float32)
    ys = np.ascontiguousarray(y, dtype=np.float32)
    ws = np.ascontiguousarray(w, dtype=np.float32)
    g = np.empty(p.size, dtype=np.float32)
    loss = np.zeros(1, dtype=np.float32)
    rc = lib.mhd_mlp_gradients(_ptr(p), _ptr(xs), _ptr(ys), _ptr(ws), xs.shape[0], _ptr(g),
                               _ptr(loss))
    if rc:
        raise MhdError(rc, "mhd_mlp_gradients failed")
    return g, float(loss[0])


def mlp_train(params, x, y, w, epochs=20, batch=1024, lr=1e-3, seed=0):
    """Adam training on the device; returns (params, per-epoch loss)."""
    lib = load()
    p = np.array(params, dtype=np.float32, copy=True)
    xs = np.ascontiguousarray(x, dtype=np.float32)
    ys = np.ascontiguousarray(y, dtype=np.float32)
    ws = np.ascontiguousarray(w, dtype=np.float32)
    losses = np.zeros(epochs, dtype=np.float32)
    rc = lib.mhd_mlp_train(_ptr(p), _ptr(xs), _ptr(ys), _ptr(ws), xs.shape[0], epochs, batch, lr,
                           seed, _ptr(losses))
    if rc:
        raise MhdError(rc, "mhd_mlp_train failed")
    return p, losses


def _cfg(cfg: "SolverConfig", nx: int = 0, ny: int = 0) -> Cfg:
    return Cfg(nx or cfg.nx, ny or cfg.ny, cfg.cfl, cfg.alpha, int(cfg.muscl), int(cfg.hlld),
               cfg.precision, cfg.hlld_tol, cfg.tile, cfg.batch_steps, cfg.lowbeta_ratio,
               cfg.fallback, cfg.fuse, cfg.rollback_steps)


class Ensemble:
    """Many instances of uneven sizes advanced together (one kernel per
    stage over a device list of the running instances' tiles, per-instance
    dt) behind the C ABI."""

    def __init__(self, cfg: SolverConfig, sizes):
        self.lib = load()
        self.cfg = cfg
        self.sizes = [(int(a), int(b)) for a, b in sizes]
        nx = np.array([a for a, _ in self.sizes], dtype=np.int32)
        ny = np.array([b for _, b in self.sizes], dtype=np.int32)
        c = _cfg(cfg, 2, 2)
        h = ctypes.c_void_p()
        rc = self.lib.mhd_ensemble_create(ctypes.byref(c), len(self.sizes), _ptr(nx), _ptr(ny),
                                          ctypes.byref(h))
        if rc:
            raise MhdError(rc, "mhd_ensemble_create failed")
        self.h = h
        self.dtype = DTYPES[cfg.precision]

    def close(self):
        if getattr(self, "h", None):
            self.lib.mhd_ensemble_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _check(self, rc, allow=(0,)):
        if rc not in allow:
            raise MhdError(rc, self.lib.mhd_ensemble_last_error(self.h).decode())
        return rc

    def set_state(self, i: int, u: np.ndarray):
        u = np.ascontiguousarray(u, dtype=self.dtype).reshape(9, -1)
        self._check(self.lib.mhd_ensemble_set_state(self.h, i, _ptr(u), u.shape[1]))

    def init_problem(self, i: int, problem: str = "ot", params=()):
        p = np.ascontiguousarray(params, dtype=np.float64)
        self._check(self.lib.mhd_ensemble_init_problem(self.h, i, problem.encode(),
                                                       _ptr(p) if p.size else None, p.size))

    def advance(self, t_end: float, max_steps: int = -1):
        st = (Status * len(self.sizes))()
        self._check(self.lib.mhd_ensemble_advance(self.h, t_end, max_steps, st),
                    allow=(0, 2, 3, 4))
        return list(st)

    def status(self, i: int) -> Status:
        st = Status()
        self._check(self.lib.mhd_ensemble_get_status(self.h, i, ctypes.byref(st)),
                    allow=(0, 2, 3, 4))
        return st

    def get_state(self, i: int, which: int = 0) -> np.ndarray:
        nx, ny = self.sizes[i]
        out = np.empty((9, ny, nx), dtype=self.dtype)
        self._check(self.lib.mhd_ensemble_get_state(self.h, i, which, _ptr(out), nx * ny))
        return out

    def enqueue_steps(self, t_end: float, nsteps: int):
        self._check(self.lib.mhd_ensemble_enqueue_steps(self.h, t_end, nsteps))

    def time_kernels(self, t_end: float, nsteps: int):
        out = (ctypes.c_double * 3)()
        self._check(self.lib.mhd_ensemble_time_kernels(self.h, t_end, nsteps, out))
        return tuple(out)


def _aos(a, dtype) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=dtype)
    assert a.ndim == 2 and a.shape[1] == 9
    return a


def glm_flux(UL, UR, ch: float, direction: int = 0, hlld: bool = True,
             precision: int = PREC_FP64, hlld_tol: float = 1.0e-4):
    """Batched mhdGlmFluxX/Y on the device: (F[n,9], status[n])."""
    lib = load()
    dt = DTYPES[precision]
    a, b = _aos(UL, dt), _aos(UR, dt)
    n = a.shape[0]
    F = np.empty((n, 9), dtype=dt)
    st = np.empty(n, dtype=np.int32)
    rc = lib.mhd_glm_flux_batch(precision, direction, n, _ptr(a), _ptr(b), ch, int(hlld),
                                hlld_tol, _ptr(F), _ptr(st))
    if rc:
        raise MhdError(rc, "mhd_glm_flux_batch failed")
    return F, st


def cons_to_prim(U, precision: int = PREC_FP64) -> np.ndarray:
    """Batched mhdGlmConsToPrim on the device."""
    lib = load()
    dt = DTYPES[precision]
    a = _aos(U, dt)
    W = np.empty_like(a)
    rc = lib.mhd_cons_to_prim_batch(precision, a.shape[0], _ptr(a), _ptr(W))
    if rc:
        raise MhdError(rc, "mhd_cons_to_prim_batch failed")
    return W


def fast_speed(W, direction: int = 0, precision: int = PREC_FP64) -> np.ndarray:
    """Batched mhdFastSpeedX/Y on the device."""
    lib = load()
    dt = DTYPES[precision]
    a = _aos(W, dt)
    out = np.empty(a.shape[0], dtype=dt)
    rc = lib.mhd_fast_speed_batch(precision, direction, a.shape[0], _ptr(a), _ptr(out))
    if rc:
        raise MhdError(rc, "mhd_fast_speed_batch failed")
    return out


def run_cli(args, outdir: str, precision: str = "fp64", env=None, timeout=600):
    """Runs the drop-in executable; returns the CompletedProcess."""
    e = dict(os.environ if env is None else env)
    e["OUTDIR"] = outdir
    e["MHD_PRECISION"] = precision
    return subprocess.run([CLI_PATH] + [str(x) for x in args], env=e, capture_output=True,
                          text=True, timeout=timeout)
