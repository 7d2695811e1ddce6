"""MI355X (gfx950) per-timestep liquid solve of the vdropimpact drop-impact code.

The compute path is the HIP library ``_lib/libvdi_hip.so`` behind the C ABI
of include/vdi.h.  This module is the Python mirror of the reference's C++
entry points for that path:

  Fluid2D.step_forward(dt)          fluid_2d::step_forward (fluid_2d.cc:93)
  Fluid2D.init_fields()             fluid_2d::init_fields (fluid_2d_sio.cc:14)
  Fluid2D.gas_layer_update(t, dt)   gas_layer::update (gas_layer.cc:36)
  Fluid2D.mg_solve_v_cycle(k, t)    mgs_mac/mgs_fem::solve_v_cycle (mgs_mac.hh:54, mgs_fem.hh:61)
  Fluid2D.visco_setup_and_solve()   visco_impl::setup_and_solve (visco_impl.cc:21)
  MGSystem                          the standalone systems of mg_test.cc

There is no CPU fallback: if the library is missing, loading fails loudly.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from .params import PhysicalConfig, VdiParams, drop_impact_params, h0_nd_for_hmin

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "_lib", "libvdi_hip.so")

# status codes (include/vdi.h)
OK, ERR_MG_NOCONV, ERR_NEWTON_NOCONV, ERR_TRIDIAG, ERR_HIP, ERR_ARG = range(6)
# solver kinds
MAC, FEM, VX, VY = 0, 1, 2, 3
# arithmetic forms
FORM_FP64, FORM_MIXED = 0, 1
# passes for Fluid2D.run_pass
PASS_GODUNOV, PASS_CORRECT_ADVECT = 0, 1
# kernel classes for Fluid2D.time_kernels
KCLASS_MG_DOWN0, KCLASS_MG_UP0, KCLASS_MG_MID, KCLASS_MG_COARSE, KCLASS_GAS, KCLASS_SIMPSON = range(6)
KCLASS_CC_DOWN, KCLASS_CC_UP, KCLASS_FEM_DOWN, KCLASS_FEM_UP, KCLASS_XSOLVE = range(6, 11)
KCLASS_SCC_DOWN, KCLASS_SCC_UP, KCLASS_SFEM_DOWN, KCLASS_SFEM_UP = range(11, 15)
KCLASS_NAMES = {0: "k_cc_down/k_fem_down (level 0)", 1: "k_cc_up/k_fem_up (level 0)", 2: "level kernels 1..lc-1",
                3: "k_coarse", 4: "k_gas", 5: "k_simpson"}
# the same launches by kernel function, as rocprofv3 names them
KFUNC_NAMES = {6: "vdi::k_down<double, false>", 7: "vdi::k_up<double, false>", 8: "vdi::k_down<double, true>",
               9: "vdi::k_up<double, true>", 3: "vdi::k_coarse<double>", 10: "vdi::k_xsolve<double>",
               4: "vdi::k_gas", 5: "vdi::k_simpson", 11: "vdi::k_sdown<double, nu>", 12: "vdi::k_sup<double, nu>",
               13: "vdi::k_sndown<double, nu>", 14: "vdi::k_snup<double, nu>"}
# array ids
(U, V, P, UL, VL, UD, VD, C0, C1, SRC, GP, GTAU, GP_OLD, GTAU_OLD, H, HX, HT, FBD, PG, PG_OLD,
 Z_MAC, Z_FEM, Z_VX, Z_VY) = range(24)

# Exported C ABI: name -> (restype, argtypes).  Kept in step with include/vdi.h.
_vp = ctypes.c_void_p
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_i, _d, _l = ctypes.c_int, ctypes.c_double, ctypes.c_long


class VdiStats(ctypes.Structure):
    _fields_ = [
        ("last_cycles", ctypes.c_int * 4),
        ("total_cycles", ctypes.c_long * 4),
        ("solves", ctypes.c_long * 4),
        ("last_newton", ctypes.c_int),
        ("steps", ctypes.c_long),
        ("time", ctypes.c_double),
        ("last_height_passes", ctypes.c_int),
    ]


ABI = {
    "vdi_status_string": (ctypes.c_char_p, [_i]),
    "vdi_last_error": (ctypes.c_char_p, []),
    "vdi_create": (_i, [ctypes.POINTER(VdiParams), ctypes.POINTER(_vp)]),
    "vdi_destroy": (_i, [_vp]),
    "vdi_set_solver": (_i, [_vp, _i, _i, _d, _i]),
    "vdi_set_schedule": (_i, [_vp, _i, _i, _i]),
    "vdi_get_schedule": (_i, [_vp, _i, _ip, _ip]),
    "vdi_set_schedule_levels": (_i, [_vp, _i, _i, _i, _i, _i]),
    "vdi_get_schedule_levels": (_i, [_vp, _i, _ip]),
    "vdi_set_form": (_i, [_vp, _i]),
    "vdi_set_graphs": (_i, [_vp, _i]),
    "vdi_set_concurrency": (_i, [_vp, _i]),
    "vdi_get_counters": (_i, [_vp, ctypes.POINTER(_l), ctypes.POINTER(_l)]),
    "vdi_get_budgets": (_i, [_vp, _ip]),
    "vdi_init_fields": (_i, [_vp]),
    "vdi_step_forward": (_i, [_vp, _d]),
    "vdi_gas_layer_update": (_i, [_vp, _d, _d]),
    "vdi_mg_solve_v_cycle": (_i, [_vp, _i, _d, _ip]),
    "vdi_visco_setup_and_solve": (_i, [_vp, _i, _d, _d, _d, _ip]),
    "vdi_run_pass": (_i, [_vp, _i, _d, _i, _dp]),
    "vdi_set_tracers": (_i, [_vp, _dp, _i]),
    "vdi_get_tracers": (_i, [_vp, _dp, _i]),
    "vdi_array_len": (_l, [_vp, _i]),
    "vdi_get_array": (_i, [_vp, _i, _dp, _l]),
    "vdi_set_array": (_i, [_vp, _i, _dp, _l]),
    "vdi_get_time": (_i, [_vp, _dp]),
    "vdi_set_time": (_i, [_vp, _d]),
    "vdi_get_stats": (_i, [_vp, ctypes.POINTER(VdiStats)]),
    "vdi_reset_stats": (_i, [_vp]),
    "vdi_synchronize": (_i, [_vp]),
    "vdi_time_kernels": (_i, [_vp, _i, _i]),
    "vdi_kernel_times": (_i, [_vp, ctypes.POINTER(_l), _dp, _dp]),
    "vdi_copy_probe": (_i, [_l, _i, _dp]),
    "vdi_device_info": (_i, [_ip, _i]),
    "vdi_state_len": (_l, [_vp]),
    "vdi_download_state": (_i, [_vp, _dp, _l]),
    "vdi_upload_state": (_i, [_vp, _dp, _l]),
    "vdi_mg_create": (_i, [_i, _i, _i, _d, _d, _i, ctypes.POINTER(_vp)]),
    "vdi_mg_destroy": (_i, [_vp]),
    "vdi_mg_set_schedule": (_i, [_vp, _i, _i]),
    "vdi_mg_set_form": (_i, [_vp, _i]),
    "vdi_mg_levels": (_i, [_vp]),
    "vdi_mg_set_b": (_i, [_vp, _dp, _l]),
    "vdi_mg_set_z": (_i, [_vp, _dp, _l]),
    "vdi_mg_get_z": (_i, [_vp, _dp, _l]),
    "vdi_mg_apply": (_i, [_vp, _dp, _dp, _l]),
    "vdi_mg_solve": (_i, [_vp, _d, _i, _ip]),
    "vdi_mg_vcycles": (_i, [_vp, _i]),
    "vdi_mg_solve_fixed": (_i, [_vp, _d, _i, _ip, _ip]),
    "vdi_mg_device_arrays": (_i, [_vp, ctypes.POINTER(_vp), ctypes.POINTER(_vp)]),
    "vdi_mg_solve_noalloc": (_i, [_i, _i, _i, _d, _d, _i, _vp, _vp, _d, _i, _ip]),
    "vdi_mg_residual": (_i, [_vp, _dp, _dp]),
    "vdi_mg_synchronize": (_i, [_vp]),
    "vdi_mg_cycle_bytes": (_i, [_vp, _dp]),
    "vdi_tile_shape": (_i, [_i, _ip, _ip]),
    "vdi_tile_stream": (_i, [_i]),
    "vdi_mg_set_tile": (_i, [_vp, _i, _i, _i, _i]),
    "vdi_mg_autotune": (_i, [_vp, _i]),
    "vdi_mg_time_kernels": (_i, [_vp, _i, _i]),
    "vdi_mg_kernel_times": (_i, [_vp, ctypes.POINTER(_l), _dp, _dp]),
    "vdi_autotune": (_i, [_vp, _i]),
    "vdi_get_tiles": (_i, [_vp, _i, _ip, _i]),
    "vdi_batch_create": (_i, [ctypes.POINTER(VdiParams), _i, _i, ctypes.POINTER(_vp)]),
    "vdi_batch_destroy": (_i, [_vp]),
    "vdi_batch_size": (_i, [_vp]),
    "vdi_batch_sim": (_i, [_vp, _i, ctypes.POINTER(_vp)]),
    "vdi_batch_init_fields": (_i, [_vp]),
    "vdi_batch_step": (_i, [_vp]),
    "vdi_batch_autotune": (_i, [_vp, _i]),
    "vdi_batch_step_times": (_i, [_vp, _dp, _dp, _i]),
    "vdi_batch_member_spans": (_i, [_vp, _dp, _i]),
    "vdi_batch_member_status": (_i, [_vp, _ip, _i]),
    "vdi_batch_restart_member": (_i, [_vp, _i]),
    "vdi_batch_counters": (_i, [_vp, ctypes.POINTER(_l), ctypes.POINTER(_l)]),
    "vdi_batch_tiles": (_i, [_vp, _ip, _i]),
    "vdi_guess_attach": (_i, [_vp, _i, ctypes.c_ulonglong]),
    "vdi_guess_mode": (_i, [_vp, _i]),
    "vdi_guess_train": (_i, [_vp, _i, _i, _d, ctypes.c_ulonglong, _dp]),
    "vdi_guess_stats": (_i, [_vp, ctypes.POINTER(_l), ctypes.POINTER(_l), ctypes.POINTER(_l), _dp]),
    "vdi_guess_copy": (_i, [_vp, _vp]),
    "vdi_nn_create": (_i, [_i, _i, ctypes.c_ulonglong, ctypes.POINTER(_vp)]),
    "vdi_nn_destroy": (_i, [_vp]),
    "vdi_nn_nparams": (_l, [_vp]),
    "vdi_nn_get_params": (_i, [_vp, ctypes.POINTER(ctypes.c_float), _l]),
    "vdi_nn_set_params": (_i, [_vp, ctypes.POINTER(ctypes.c_float), _l]),
    "vdi_nn_forward": (_i, [_vp, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)]),
    "vdi_nn_gradient": (_i, [_vp, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float),
                             ctypes.POINTER(ctypes.c_float), _dp]),
}

_lib = None


class VdiError(RuntimeError):
    def __init__(self, status: int, where: str):
        lib = _lib
        msg = lib.vdi_status_string(status).decode() if lib else str(status)
        detail = lib.vdi_last_error().decode() if (lib and status == ERR_HIP) else ""
        super().__init__(f"{where}: {msg}" + (f" [{detail}]" if detail else ""))
        self.status = status


def load_library(path: str = LIB_PATH) -> ctypes.CDLL:
    """Loads libvdi_hip.so and binds the C ABI.  Raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"libvdi_hip.so not built ({path}); run python -m vdropimpact_amd.build")
    lib = ctypes.CDLL(path)
    for name, (res, args) in ABI.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _check(status: int, where: str):
    if status != OK:
        raise VdiError(status, where)


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_dp)


class Fluid2D:
    """GPU-resident drop-impact liquid solve (the reference's fluid_2d)."""

    def __init__(self, params: VdiParams, nu1: int = 0, nu2: int = 0, rel_tol: float = 1e-10,
                 max_cycles: int = 200, _handle=None):
        self.lib = load_library()
        self.params = params
        self.m, self.n = params.m, params.n
        self._owned = _handle is None
        if _handle is None:
            h = _vp()
            _check(self.lib.vdi_create(ctypes.byref(params), ctypes.byref(h)), "vdi_create")
            self.h = h
            self.set_solver(nu1, nu2, rel_tol, max_cycles)
        else:
            self.h = _handle

    def close(self):
        if getattr(self, "h", None):
            if self._owned:
                self.lib.vdi_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_solver(self, nu1=0, nu2=0, rel_tol=1e-10, max_cycles=200):
        """Tolerance, V-cycle cap and (unless both are 0) one V(nu1, nu2) for every solver."""
        _check(self.lib.vdi_set_solver(self.h, nu1, nu2, rel_tol, max_cycles), "vdi_set_solver")

    def set_schedule(self, which: int, nu1: int, nu2: int, nu1s=None, nu2s=None):
        """Smoothing sweeps of one solver (MAC, FEM, VX or VY): V(nu1, nu2)
        on the finest level and V(nu1s, nu2s) on the restricted levels
        (the same pair when these are omitted)."""
        nu1s = nu1 if nu1s is None else nu1s
        nu2s = nu2 if nu2s is None else nu2s
        _check(self.lib.vdi_set_schedule_levels(self.h, which, nu1, nu2, nu1s, nu2s), "vdi_set_schedule_levels")

    def schedule(self, which: int) -> tuple:
        """(nu1, nu2, nu1s, nu2s) of one solver."""
        out = (ctypes.c_int * 4)()
        _check(self.lib.vdi_get_schedule_levels(self.h, which, out), "vdi_get_schedule_levels")
        return tuple(out)

    def set_form(self, form: int):
        _check(self.lib.vdi_set_form(self.h, form), "vdi_set_form")

    def set_graphs(self, enable: bool):
        _check(self.lib.vdi_set_graphs(self.h, int(enable)), "vdi_set_graphs")

    def set_concurrency(self, enable: bool):
        """Parallel graph branches for the two viscous solves (default on)."""
        _check(self.lib.vdi_set_concurrency(self.h, int(enable)), "vdi_set_concurrency")

    def counters(self):
        """(steps run as captured graphs, of those redone on the direct path)."""
        a, b = ctypes.c_long(), ctypes.c_long()
        _check(self.lib.vdi_get_counters(self.h, ctypes.byref(a), ctypes.byref(b)), "vdi_get_counters")
        return a.value, b.value

    def budgets(self) -> list:
        """V-cycle budgets [MAC, VX, VY, FEM] of the next captured step."""
        arr = (ctypes.c_int * 4)()
        _check(self.lib.vdi_get_budgets(self.h, arr), "vdi_get_budgets")
        return list(arr)

    def init_fields(self):
        _check(self.lib.vdi_init_fields(self.h), "init_fields")

    def step_forward(self, dt: float):
        _check(self.lib.vdi_step_forward(self.h, dt), "step_forward")

    def gas_layer_update(self, t: float, dt: float):
        _check(self.lib.vdi_gas_layer_update(self.h, t, dt), "gas_layer::update")

    def mg_solve_v_cycle(self, which: int, t: float) -> int:
        c = ctypes.c_int(0)
        _check(self.lib.vdi_mg_solve_v_cycle(self.h, which, t, ctypes.byref(c)), "solve_v_cycle")
        return c.value

    def visco_setup_and_solve(self, which: int, vfx: float, vfy: float, t: float) -> int:
        c = ctypes.c_int(0)
        _check(self.lib.vdi_visco_setup_and_solve(self.h, which, vfx, vfy, t, ctypes.byref(c)),
               "visco_impl::setup_and_solve")
        return c.value

    def synchronize(self):
        _check(self.lib.vdi_synchronize(self.h), "synchronize")

    def set_tracers(self, pos):
        """Tracer positions, shape (ntrace, 2) or flat (x0, y0, x1, ...)."""
        a = np.ascontiguousarray(pos, np.float64).ravel()
        self.ntrace = a.size // 2
        _check(self.lib.vdi_set_tracers(self.h, _ptr(a), self.ntrace), "vdi_set_tracers")

    def tracers(self) -> np.ndarray:
        nt = getattr(self, "ntrace", 0)
        out = np.empty(2 * nt, np.float64)
        _check(self.lib.vdi_get_tracers(self.h, _ptr(out), nt), "vdi_get_tracers")
        return out.reshape(nt, 2)

    def run_pass(self, which: int, dt: float, reps: int = 1) -> float:
        """The advection passes alone (PASS_GODUNOV, PASS_CORRECT_ADVECT);
        returns the average launch time in ms."""
        ms = ctypes.c_double(0)
        _check(self.lib.vdi_run_pass(self.h, which, dt, reps, ctypes.byref(ms)), "vdi_run_pass")
        return ms.value

    @property
    def time(self) -> float:
        t = ctypes.c_double()
        _check(self.lib.vdi_get_time(self.h, ctypes.byref(t)), "get_time")
        return t.value

    @time.setter
    def time(self, t: float):
        _check(self.lib.vdi_set_time(self.h, t), "set_time")

    def get(self, aid: int) -> np.ndarray:
        n = self.lib.vdi_array_len(self.h, aid)
        if n <= 0:
            raise ValueError(f"array {aid} not available")
        out = np.empty(n, np.float64)
        _check(self.lib.vdi_get_array(self.h, aid, _ptr(out), n), "get_array")
        return out

    def set(self, aid: int, arr):
        a = np.ascontiguousarray(arr, np.float64).ravel()
        _check(self.lib.vdi_set_array(self.h, aid, _ptr(a), a.size), "set_array")

    def padded(self, aid: int) -> np.ndarray:
        return self.get(aid).reshape(self.n + 4, self.m + 4)

    def stats(self) -> VdiStats:
        s = VdiStats()
        _check(self.lib.vdi_get_stats(self.h, ctypes.byref(s)), "get_stats")
        return s

    def reset_stats(self):
        _check(self.lib.vdi_reset_stats(self.h), "reset_stats")

    def time_kernels(self, kclass: int, max_launches: int = 8192):
        """Times every launch of one kernel class (KCLASS_*) with HIP events."""
        _check(self.lib.vdi_time_kernels(self.h, kclass, max_launches), "time_kernels")

    def kernel_times(self):
        """(launches, total ms, algorithmic bytes) of the timed launches."""
        n, t, b = ctypes.c_long(), ctypes.c_double(), ctypes.c_double()
        _check(self.lib.vdi_kernel_times(self.h, ctypes.byref(n), ctypes.byref(t), ctypes.byref(b)), "kernel_times")
        return n.value, t.value, b.value

    def download_state(self) -> np.ndarray:
        n = self.lib.vdi_state_len(self.h)
        out = np.empty(n, np.float64)
        _check(self.lib.vdi_download_state(self.h, _ptr(out), n), "download_state")
        return out

    def upload_state(self, state):
        a = np.ascontiguousarray(state, np.float64).ravel()
        _check(self.lib.vdi_upload_state(self.h, _ptr(a), a.size), "upload_state")

    def autotune(self, reps: int = 5):
        """Picks the fastest tile shape of every multigrid level kernel."""
        _check(self.lib.vdi_autotune(self.h, reps), "autotune")

    def tile_indices(self, which: int) -> list:
        """Chosen tile-shape indices of solver `which`: [prec][down/up][level] then the mixed level-0 up."""
        arr = (ctypes.c_int * 256)()
        n = self.lib.vdi_get_tiles(self.h, which, arr, 256)
        if n < 0:
            raise VdiError(-n, "get_tiles")
        return list(arr)[:n]

    def tiles(self, which: int) -> list:
        """Chosen tile shapes (tx, ty) of solver `which`: [prec][down/up][level] then the mixed level-0 up."""
        return [tile_shape(v) for v in self.tile_indices(which)]


    # ---- learned initial guess of the MAC solve ----
    def guess_attach(self, capacity: int = 0, seed: int = 1):
        _check(self.lib.vdi_guess_attach(self.h, capacity, seed), "guess_attach")

    def guess_mode(self, mode: int):
        """0 off, 1 record training pairs, 2 use the network."""
        _check(self.lib.vdi_guess_mode(self.h, mode), "guess_mode")

    def guess_train(self, epochs: int, batch: int = 4, lr: float = 1e-3, seed: int = 7) -> float:
        loss = ctypes.c_double()
        _check(self.lib.vdi_guess_train(self.h, epochs, batch, lr, seed, ctypes.byref(loss)), "guess_train")
        return loss.value

    def guess_stats(self) -> dict:
        a, b, c, k = ctypes.c_long(), ctypes.c_long(), ctypes.c_long(), ctypes.c_double()
        _check(self.lib.vdi_guess_stats(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c), ctypes.byref(k)),
               "guess_stats")
        return {"pairs": a.value, "accepted": b.value, "rejected": c.value, "kappa": k.value}

    def guess_copy_from(self, other: "Fluid2D"):
        _check(self.lib.vdi_guess_copy(self.h, other.h), "guess_copy")


_fp = ctypes.POINTER(ctypes.c_float)
NN_INPUTS = 3  # network input planes: residual, extrapolated guess, previous correction


def _fptr(a: np.ndarray):
    return a.ctypes.data_as(_fp)


class UNet:
    """The learned-guess network alone (H x W pixels), for gradient checks."""

    def __init__(self, H: int, W: int, seed: int = 1):
        self.lib = load_library()
        self.H, self.W = H, W
        h = _vp()
        _check(self.lib.vdi_nn_create(H, W, seed, ctypes.byref(h)), "vdi_nn_create")
        self.h = h
        self.nparams = self.lib.vdi_nn_nparams(self.h)

    def close(self):
        if getattr(self, "h", None):
            self.lib.vdi_nn_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self) -> np.ndarray:
        out = np.empty(self.nparams, np.float32)
        _check(self.lib.vdi_nn_get_params(self.h, _fptr(out), out.size), "vdi_nn_get_params")
        return out

    def set_params(self, p):
        a = np.ascontiguousarray(p, np.float32)
        _check(self.lib.vdi_nn_set_params(self.h, _fptr(a), a.size), "vdi_nn_set_params")

    def forward(self, x2) -> np.ndarray:
        """x2: the NN_INPUTS input planes, (NN_INPUTS, H*W)."""
        x = np.ascontiguousarray(x2, np.float32).reshape(NN_INPUTS, -1)
        y = np.empty(self.H * self.W, np.float32)
        _check(self.lib.vdi_nn_forward(self.h, _fptr(x), _fptr(y)), "vdi_nn_forward")
        return y

    def gradient(self, x2, t):
        x = np.ascontiguousarray(x2, np.float32).reshape(NN_INPUTS, -1)
        tt = np.ascontiguousarray(t, np.float32).ravel()
        g = np.empty(self.nparams, np.float32)
        loss = ctypes.c_double()
        _check(self.lib.vdi_nn_gradient(self.h, _fptr(x), _fptr(tt), _fptr(g), ctypes.byref(loss)), "vdi_nn_gradient")
        return loss.value, g


def tile_shape(index: int):
    lib = load_library()
    tx, ty = ctypes.c_int(), ctypes.c_int()
    _check(lib.vdi_tile_shape(index, ctypes.byref(tx), ctypes.byref(ty)), "tile_shape")
    return tx.value, ty.value


def tile_stream(index: int) -> bool:
    """True for a streaming-launch shape (one wavefront per strip and segment)."""
    return load_library().vdi_tile_stream(index) == 1


def tile_label(index: int) -> str:
    """'64x32' for a tile, 'stream/16' for a streaming launch of 16-row segments."""
    tx, ty = tile_shape(index)
    return f"stream/{ty}" if tile_stream(index) else f"{tx}x{ty}"


class Batch:
    """Many independent drop simulations advanced by one captured graph of
    persistent kernels per step (uneven grids and parameters)."""

    def __init__(self, params_list, form: int = FORM_FP64):
        self.lib = load_library()
        self.params = list(params_list)
        arr = (VdiParams * len(self.params))(*self.params)
        h = _vp()
        _check(self.lib.vdi_batch_create(arr, len(self.params), form, ctypes.byref(h)), "vdi_batch_create")
        self.h = h
        self.size = len(self.params)

    def close(self):
        if getattr(self, "h", None):
            self.lib.vdi_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sim(self, k: int) -> Fluid2D:
        """Borrowed view of member k (valid while the batch lives)."""
        h = _vp()
        _check(self.lib.vdi_batch_sim(self.h, k, ctypes.byref(h)), "vdi_batch_sim")
        return Fluid2D(self.params[k], _handle=h)

    def init_fields(self):
        _check(self.lib.vdi_batch_init_fields(self.h), "vdi_batch_init_fields")

    def step(self):
        _check(self.lib.vdi_batch_step(self.h), "vdi_batch_step")

    def autotune(self, reps: int = 5):
        _check(self.lib.vdi_batch_autotune(self.h, reps), "vdi_batch_autotune")

    def step_times(self):
        """(graph ms of the last batched step, per-member busy ms array)."""
        g = ctypes.c_double()
        busy = np.zeros(self.size)
        _check(self.lib.vdi_batch_step_times(self.h, ctypes.byref(g), _ptr(busy), self.size), "step_times")
        return g.value, busy

    def member_spans(self) -> np.ndarray:
        """Per-member step time (ms) of the last batched step: first work item
        start to last work item end, from the device clock."""
        out = np.zeros(self.size)
        _check(self.lib.vdi_batch_member_spans(self.h, _ptr(out), self.size), "member_spans")
        return out

    def member_status(self) -> np.ndarray:
        """Status of each member's last step (OK, or why its run ended)."""
        out = (ctypes.c_int * self.size)()
        _check(self.lib.vdi_batch_member_status(self.h, out, self.size), "member_status")
        return np.array(out[:], dtype=int)

    def restart(self, k: int):
        """Starts member k over as a new simulation with the same parameters."""
        _check(self.lib.vdi_batch_restart_member(self.h, k), "vdi_batch_restart_member")

    def counters(self):
        a, b = ctypes.c_long(), ctypes.c_long()
        _check(self.lib.vdi_batch_counters(self.h, ctypes.byref(a), ctypes.byref(b)), "counters")
        return a.value, b.value

    def tiles(self) -> list:
        arr = (ctypes.c_int * 4096)()
        n = self.lib.vdi_batch_tiles(self.h, arr, 4096)
        return [tile_shape(v) for v in list(arr)[:n]]


def copy_probe(n: int, reps: int = 10) -> float:
    """Average ms of a plain device copy of n doubles (counter calibration)."""
    lib = load_library()
    ms = ctypes.c_double()
    _check(lib.vdi_copy_probe(n, reps, ctypes.byref(ms)), "copy_probe")
    return ms.value


def device_info() -> dict:
    """LDS per workgroup, CU count and other facts of device 0."""
    lib = load_library()
    arr = (ctypes.c_int * 9)()
    _check(lib.vdi_device_info(arr, 9), "device_info")
    keys = ["lds_per_block", "cus", "threads_per_block", "wavefront", "clock_khz", "mem_clock_khz",
            "mem_bus_bits", "lds_per_cu", "l2_bytes"]
    return dict(zip(keys, list(arr)))


class MGSystem:
    """A standalone linear system (mgs_mac / mgs_fem / visco_impl test
    constructors, mg_test.cc) solved on the GPU."""

    def __init__(self, kind: int, m: int, n: int, cx: float, cy: float, x_sym: int = 1,
                 nu1: int = 2, nu2: int = 2):
        self.lib = load_library()
        self.kind, self.m, self.n = kind, m, n
        h = _vp()
        _check(self.lib.vdi_mg_create(kind, m, n, cx, cy, x_sym, ctypes.byref(h)), "vdi_mg_create")
        self.h = h
        _check(self.lib.vdi_mg_set_schedule(self.h, nu1, nu2), "vdi_mg_set_schedule")

    def close(self):
        if getattr(self, "h", None):
            self.lib.vdi_mg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def levels(self) -> int:
        return self.lib.vdi_mg_levels(self.h)

    def set_form(self, form: int):
        _check(self.lib.vdi_mg_set_form(self.h, form), "vdi_mg_set_form")

    def set_b(self, b):
        a = np.ascontiguousarray(b, np.float64).ravel()
        _check(self.lib.vdi_mg_set_b(self.h, _ptr(a), a.size), "vdi_mg_set_b")

    def set_z(self, z):
        a = np.ascontiguousarray(z, np.float64).ravel()
        _check(self.lib.vdi_mg_set_z(self.h, _ptr(a), a.size), "vdi_mg_set_z")

    def get_z(self) -> np.ndarray:
        out = np.empty(self.m * self.n, np.float64)
        _check(self.lib.vdi_mg_get_z(self.h, _ptr(out), out.size), "vdi_mg_get_z")
        return out

    def apply(self, z) -> np.ndarray:
        a = np.ascontiguousarray(z, np.float64).ravel()
        out = np.empty_like(a)
        _check(self.lib.vdi_mg_apply(self.h, _ptr(a), _ptr(out), a.size), "vdi_mg_apply")
        return out

    def solve(self, rel_tol: float = 1e-10, max_cycles: int = 200) -> int:
        c = ctypes.c_int(0)
        _check(self.lib.vdi_mg_solve(self.h, rel_tol, max_cycles, ctypes.byref(c)), "vdi_mg_solve")
        return c.value

    def vcycles(self, k: int):
        _check(self.lib.vdi_mg_vcycles(self.h, k), "vdi_mg_vcycles")

    def device_arrays(self):
        """(z, b) device addresses of the level-0 arrays (as ctypes void pointers)."""
        z, b = _vp(), _vp()
        _check(self.lib.vdi_mg_device_arrays(self.h, ctypes.byref(z), ctypes.byref(b)), "vdi_mg_device_arrays")
        return z, b

    def solve_fixed(self, rel_tol: float, budget: int):
        """Fixed V-cycle budget without host reads in between (a captured
        step's form); returns (done flag, V-cycles run)."""
        d, c = ctypes.c_int(0), ctypes.c_int(0)
        _check(self.lib.vdi_mg_solve_fixed(self.h, rel_tol, budget, ctypes.byref(d), ctypes.byref(c)),
               "vdi_mg_solve_fixed")
        return d.value, c.value

    def residual(self):
        r, b = ctypes.c_double(), ctypes.c_double()
        _check(self.lib.vdi_mg_residual(self.h, ctypes.byref(r), ctypes.byref(b)), "vdi_mg_residual")
        return r.value, b.value

    def synchronize(self):
        _check(self.lib.vdi_mg_synchronize(self.h), "vdi_mg_synchronize")

    def cycle_bytes(self) -> float:
        b = ctypes.c_double()
        _check(self.lib.vdi_mg_cycle_bytes(self.h, ctypes.byref(b)), "vdi_mg_cycle_bytes")
        return b.value

    def set_tile(self, prec: int, which: int, level: int, shape: int):
        _check(self.lib.vdi_mg_set_tile(self.h, prec, which, level, shape), "vdi_mg_set_tile")

    def autotune(self, reps: int = 5):
        _check(self.lib.vdi_mg_autotune(self.h, reps), "vdi_mg_autotune")

    def time_kernels(self, kclass: int, max_launches: int = 1024):
        """Stamps every launch of one kernel class (KCLASS_*; -1 stops) with
        the device clock."""
        _check(self.lib.vdi_mg_time_kernels(self.h, kclass, max_launches), "vdi_mg_time_kernels")

    def kernel_times(self):
        """(launches, total ms, algorithmic bytes) of the stamped launches."""
        n, t, b = ctypes.c_long(), ctypes.c_double(), ctypes.c_double()
        _check(self.lib.vdi_mg_kernel_times(self.h, ctypes.byref(n), ctypes.byref(t), ctypes.byref(b)),
               "vdi_mg_kernel_times")
        return n.value, t.value, b.value


NUM_TILE_SHAPES = 22

__all__ = [
    "Fluid2D", "MGSystem", "Batch", "UNet", "tile_shape", "tile_label", "tile_stream", "NUM_TILE_SHAPES", "VdiParams", "VdiStats", "VdiError", "PhysicalConfig",
    "drop_impact_params", "h0_nd_for_hmin", "load_library", "LIB_PATH", "ABI",
]
