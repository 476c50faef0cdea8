"""Python host mirror of the MI355X SQP tracking-MPC library (libmpcpcg.so).

A thin ctypes layer over the C-ABI declared in include/mpcpcg.h.  Device
memory and streams come from PyTorch (ROCm); every compute call goes to the
HIP library.  There is no CPU fallback: importing works without a GPU, but
creating a Solver requires the built library and a visible GPU, and fails
loudly otherwise.

Names follow the reference call surface (EmreAdabag/gpu-pcg-tracking):
  Solver.form_kkt         <- gato_form_kkt            (schur.cuh:391-524)
  Solver.form_schur       <- form_schur               (schur.cuh:545-559)
  Solver.pcg              <- GPU-PCG pcg<T,14,N>      (toplevel.cuh:143-157, :292)
  Solver.compute_dz       <- compute_dz               (schur.cuh:562-566)
  Solver.compute_merit    <- compute_merit            (merit.cuh:99-145)
  Solver.ls_compute_merit <- ls_gato_compute_merit x8 (merit.cuh:17-94)
  Solver.sqp_solve        <- sqpSolve                 (toplevel.cuh:24-503)
  Solver.fk               <- grid::end_effector_positions_kernel
  Solver.dynamics_gradient<- gato_plant::forwardDynamicsAndGradient
"""
from __future__ import annotations

import ctypes as C
import os

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "libmpcpcg.so")

NQ, NX, NU, NXU = 7, 14, 7, 21
MAX_SQP_ITERS = 64

# status codes / bits (include/mpcpcg.h)
OK, ERR_ARG, ERR_HIP, ERR_NOT_READY, ERR_NONFINITE, ERR_NO_DEVICE, ERR_PCG_FAULT = range(7)
ST_ACTIVE, ST_PCG_MAXITER, ST_RHO_ABORT, ST_NONFINITE, ST_PCG_FAULT, ST_TIME_BUDGET = 1, 2, 4, 8, 16, 32
FLAG_LAST_KNOT_OWN_STATE = 1
FLAG_PCG_STREAM = 2
FLAG_DIRECT_SOLVE = 4
FLAG_PCG_STAIR = 8
FLAG_PCG_CAP_BY_N = 16

# every symbol the header declares (checked by the CPU test suite)
EXPORTS = (
    "mpc_default_config", "mpc_create", "mpc_destroy", "mpc_get_config", "mpc_set_config",
    "mpc_status_string", "mpc_fk", "mpc_dynamics_gradient", "mpc_form_kkt", "mpc_form_schur",
    "mpc_pcg_solve", "mpc_compute_dz", "mpc_compute_merit", "mpc_line_search_merits",
    "mpc_sqp_solve", "mpc_batch_setup", "mpc_batch_sqp", "mpc_batch_set_timing",
    "mpc_batch_phase_times", "mpc_batch_pcg_layout", "mpc_batch_pcg_launch_map", "mpc_batch_set_graph", "mpc_batch_get_plan",
    "mpc_batch_set_plan", "mpc_batch_autotune", "mpc_set_cost_weights", "mpc_get_cost_weights", "mpc_batch_qp_layer",
    "mpc_batch_qp_loss", "mpc_batch_qp_adjoint", "mpc_synth_instance", "mpc_synth_horizons",
    "mpc_inverse_dynamics", "mpc_default_closed_loop", "mpc_closed_loop", "mpc_synth_track", "mpc_direct_solve",
    "mpc_bw_copy", "mpc_bw_lds", "mpc_direct_solve_static", "mpc_set_fault_injection", "mpc_batch_pcg_exit",
    "mpc_batch_direct_resolve",
)


# iterative-refinement steps of the QP layer and adjoint solves: the fp32
# Schur / PCG solve leaves relative errors of 1e-3..1e-2 on ill-conditioned
# instances (rho = 1e-3, control weights near 1e-4); each step with an fp64
# KKT residual removes most of it
REFINE_STEPS = 2


class MpcError(RuntimeError):
    pass


class Config(C.Structure):
    _fields_ = [
        ("max_knots", C.c_uint32), ("max_instances", C.c_uint32), ("max_total_knots", C.c_uint32),
        ("dt", C.c_float), ("q_ee", C.c_float), ("qd_cost", C.c_float), ("r_cost", C.c_float),
        ("mu", C.c_float), ("rho_factor", C.c_float), ("rho_max", C.c_float), ("rho_min", C.c_float),
        ("rho_reset", C.c_float), ("pcg_exit_tol", C.c_float), ("pcg_max_iter", C.c_uint32),
        ("sqp_max_iter", C.c_uint32), ("flags", C.c_uint32), ("direct_refine", C.c_uint32),
        ("direct_kkt_refine", C.c_uint32), ("sqp_max_time_us", C.c_float),
    ]


class SqpStats(C.Structure):
    _fields_ = [
        ("sqp_iters", C.c_uint32), ("not_aborted", C.c_uint32), ("sqp_us", C.c_double),
        ("pcg_iters", C.c_uint32 * MAX_SQP_ITERS), ("pcg_maxiter", C.c_uint32 * MAX_SQP_ITERS),
        ("linsys_us", C.c_double * MAX_SQP_ITERS), ("ls_step", C.c_int32 * MAX_SQP_ITERS),
        ("time_exit", C.c_uint32), ("reserved", C.c_uint32),
    ]


class ClosedLoopConfig(C.Structure):
    _fields_ = [
        ("steps", C.c_uint32), ("sqp_iters", C.c_uint32), ("substeps", C.c_uint32), ("sim_dt", C.c_float),
        ("noise_sigma", C.c_float), ("use_graph", C.c_uint32), ("noise_seed", C.c_uint64),
        ("timing", C.c_uint32), ("reserved", C.c_uint32),
    ]


class ClosedLoopStats(C.Structure):
    _fields_ = [
        ("total_us", C.c_double), ("solve_us", C.c_double), ("linsys_us", C.c_double), ("plant_us", C.c_double),
        ("graph", C.c_uint32), ("reserved", C.c_uint32),
    ]


class PcgPlan(C.Structure):
    """mpc_pcg_plan: launch plan of the batched PCG."""
    _fields_ = [
        ("small_nt", C.c_uint32), ("max_items", C.c_uint32), ("mixed", C.c_uint32), ("teams", C.c_uint32),
        ("us", C.c_float), ("split", C.c_uint32),
    ]


_PLAN_KEYS = ("small_nt", "max_items", "mixed", "teams", "split")


_lib = None


def lib():
    """Load libmpcpcg.so (built by __graft_entry__.build()); raise if absent."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise MpcError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        vp, u32, f32 = C.c_void_p, C.c_uint32, C.c_float
        sigs = {
            "mpc_create": [C.POINTER(Config), C.POINTER(vp)],
            "mpc_destroy": [vp],
            "mpc_get_config": [vp, C.POINTER(Config)],
            "mpc_set_config": [vp, C.POINTER(Config)],
            "mpc_fk": [vp, u32, vp, u32, vp, vp],
            "mpc_dynamics_gradient": [vp, u32, vp, vp, vp, vp],
            "mpc_form_kkt": [vp, u32, vp, vp, vp, vp, vp, vp, vp, vp],
            "mpc_form_schur": [vp, u32, vp, vp, vp, vp, vp, vp, vp, f32, vp],
            "mpc_pcg_solve": [vp, u32, vp, vp, vp, vp, u32, f32, vp, vp, vp],
            "mpc_compute_dz": [vp, u32, vp, vp, vp, vp, vp, vp],
            "mpc_compute_merit": [vp, u32, vp, vp, vp, vp],
            "mpc_line_search_merits": [vp, u32, vp, vp, vp, vp, vp, vp],
            "mpc_sqp_solve": [vp, u32, vp, vp, vp, C.POINTER(f32), f32, C.POINTER(SqpStats), vp],
            "mpc_batch_setup": [vp, u32, C.POINTER(u32)],
            "mpc_batch_sqp": [vp, vp, vp, vp, vp, u32, vp, vp, vp],
            "mpc_batch_set_timing": [vp, C.c_int],
            "mpc_batch_set_graph": [vp, C.c_int],
            "mpc_batch_get_plan": [vp, C.POINTER(PcgPlan)],
            "mpc_batch_set_plan": [vp, C.POINTER(PcgPlan)],
            "mpc_batch_autotune": [vp, u32, u32, C.POINTER(PcgPlan), u32, C.POINTER(u32), vp],
            "mpc_batch_phase_times": [vp, C.POINTER(C.c_double)],
            "mpc_batch_pcg_layout": [vp, C.POINTER(C.c_uint32)],
            "mpc_batch_pcg_launch_map": [vp, vp],
            "mpc_set_cost_weights": [vp, vp, vp],
            "mpc_get_cost_weights": [vp, vp, vp],
            "mpc_batch_qp_layer": [vp, vp, vp, vp, vp, f32, u32, f32, vp, vp, vp],
            "mpc_batch_qp_loss": [vp, vp, vp, vp, vp, vp, vp],
            "mpc_batch_qp_adjoint": [vp, vp, vp, vp, vp, vp, f32, u32, f32, vp, vp, vp],
            "mpc_inverse_dynamics": [vp, u32, vp, vp, vp],
            "mpc_bw_copy": [vp, vp, C.c_size_t, vp],
            "mpc_bw_lds": [vp, u32, u32, vp],
            "mpc_direct_solve": [vp, u32, vp, vp, vp, u32, vp, vp],
            "mpc_direct_solve_static": [u32, vp, vp, vp, u32, vp],
            "mpc_set_fault_injection": [vp, u32],
            "mpc_batch_pcg_exit": [vp, vp, vp, vp, vp],
            "mpc_batch_direct_resolve": [vp, vp, u32, vp],
            "mpc_closed_loop": [vp, C.POINTER(ClosedLoopConfig), u32, vp, vp, vp, vp, vp, vp, vp, vp,
                                C.POINTER(ClosedLoopStats), vp],
        }
        for name, argt in sigs.items():
            fn = getattr(L, name)
            fn.argtypes = argt
            fn.restype = C.c_int
        L.mpc_default_config.argtypes = [C.POINTER(Config)]
        L.mpc_default_config.restype = None
        L.mpc_status_string.argtypes = [C.c_int]
        L.mpc_status_string.restype = C.c_char_p
        L.mpc_synth_instance.argtypes = [C.c_uint64, u32, f32, vp, vp]
        L.mpc_synth_instance.restype = None
        L.mpc_synth_horizons.argtypes = [C.c_uint64, u32, C.POINTER(u32), u32, C.POINTER(u32)]
        L.mpc_synth_horizons.restype = None
        L.mpc_synth_track.argtypes = [C.c_uint64, u32, f32, vp, vp]
        L.mpc_synth_track.restype = None
        L.mpc_default_closed_loop.argtypes = [C.POINTER(ClosedLoopConfig)]
        L.mpc_default_closed_loop.restype = None
        _lib = L
    return _lib


def _check(rc, what):
    if rc != OK:
        msg = lib().mpc_status_string(rc).decode()
        raise MpcError(f"{what} failed: {msg} ({rc})")


def default_config(**kw) -> Config:
    cfg = Config()
    lib().mpc_default_config(C.byref(cfg))
    for k, v in kw.items():
        if not hasattr(cfg, k):
            raise KeyError(k)
        setattr(cfg, k, v)
    return cfg


# ---- synthetic workload (host side, no GPU) ---------------------------------------------------

def synth_instance(seed: int, N: int, dt: float = 0.01):
    """(x0 (14,), q_ref (N, 7)) float32 numpy arrays of one seeded instance."""
    import numpy as np
    x0 = np.zeros(NX, np.float32)
    qref = np.zeros((N, NQ), np.float32)
    lib().mpc_synth_instance(C.c_uint64(seed), N, dt, x0.ctypes.data_as(C.c_void_p),
                             qref.ctypes.data_as(C.c_void_p))
    return x0, qref


def synth_track(seed: int, K: int, dt: float = 0.01):
    """(x0 (14,), [q, qd, qdd] (K, 21)) float32: the seeded joint path and its derivatives."""
    import numpy as np
    x0 = np.zeros(NX, np.float32)
    out = np.zeros((K, 21), np.float32)
    lib().mpc_synth_track(C.c_uint64(seed), K, dt, x0.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p))
    return x0, out


def default_closed_loop(**kw) -> ClosedLoopConfig:
    cl = ClosedLoopConfig()
    lib().mpc_default_closed_loop(C.byref(cl))
    for k, v in kw.items():
        if not hasattr(cl, k):
            raise KeyError(k)
        setattr(cl, k, v)
    return cl


def synth_horizons(seed: int, n: int, choices=(32, 64, 128, 256)):
    import numpy as np
    ch = (C.c_uint32 * len(choices))(*choices)
    out = (C.c_uint32 * n)()
    lib().mpc_synth_horizons(C.c_uint64(seed), n, ch, len(choices), out)
    return np.frombuffer(out, dtype=np.uint32).copy()


# ---- solver -------------------------------------------------------------------------------------

def _ptr(t, n=0, kind="f32", what="buffer"):
    """Device address of a CUDA tensor after checking what the kernels
    assume of it: contiguous, on the current device, float32 (kind "f32")
    or 32-bit integer (kind "u32"), at least n elements.  None passes as a
    null pointer (optional outputs)."""
    if t is None:
        return None
    import torch
    if not (t.is_cuda and t.is_contiguous()):
        raise MpcError(f"{what}: expected a contiguous CUDA tensor")
    if t.device.index != torch.cuda.current_device():
        raise MpcError(f"{what}: tensor on cuda:{t.device.index}, current device is cuda:{torch.cuda.current_device()}")
    ok = t.dtype == torch.float32 if kind == "f32" else t.dtype in (torch.int32, getattr(torch, "uint32", torch.int32))
    if not ok:
        raise MpcError(f"{what}: expected {'float32' if kind == 'f32' else 'int32 / uint32'}, got {t.dtype}")
    if t.numel() < n:
        raise MpcError(f"{what}: {t.numel()} elements, needs {n}")
    return C.c_void_p(t.data_ptr())


def bw_copy(src, dst):
    """dst = src (float32 CUDA tensors, numel % 4 == 0, 16-byte aligned) by
    the HBM bandwidth probe of the library; 8 * numel bytes moved."""
    if src.numel() != dst.numel():
        raise MpcError("bw_copy: sizes differ")
    _check(lib().mpc_bw_copy(_ptr(src, what="src"), _ptr(dst, what="dst"), src.numel(), _stream()), "mpc_bw_copy")


def bw_lds(out, n_wg: int, reps: int):
    """The LDS bandwidth probe: n_wg workgroups of 256 lanes, 64 * 256 * reps
    bytes of LDS reads each; out (int32 CUDA tensor, n_wg * 256) gets each
    lane's folded sum."""
    if out.numel() < 256 * n_wg:
        raise MpcError("bw_lds: out too small")
    _check(lib().mpc_bw_lds(_ptr(out, 256 * n_wg, "u32", "out"), n_wg, reps, _stream()), "mpc_bw_lds")


# largest horizon of direct_solve_static (MPC_DIRECT_STATIC_MAX_KNOTS)
DIRECT_STATIC_MAX_KNOTS = 512


def direct_solve_static(N, S, gamma, lam, refine=2):
    """lambda = S^-1 gamma as Solver.direct_solve, through the entry point
    that allocates nothing (no handle, factor workspace in a library-owned
    device array for N <= DIRECT_STATIC_MAX_KNOTS; calls overlapping in
    time on one GPU race on it)."""
    _check(lib().mpc_direct_solve_static(N, _ptr(S, 588 * N, what="S"), _ptr(gamma, 14 * N, what="gamma"),
                                         _ptr(lam, 14 * N, what="lambda"), int(refine), _stream()),
           "mpc_direct_solve_static")


def _stream():
    import torch
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


class Solver:
    """Owns an mpc_handle (workspace sized at construction)."""

    def __init__(self, cfg: Config | None = None, **kw):
        import torch
        if not torch.cuda.is_available():
            raise MpcError("no GPU visible: the HIP library has no CPU path")
        self.cfg = cfg if cfg is not None else default_config(**kw)
        h = C.c_void_p()
        _check(lib().mpc_create(C.byref(self.cfg), C.byref(h)), "mpc_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().mpc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_config(self, **kw):
        for k, v in kw.items():
            setattr(self.cfg, k, v)
        _check(lib().mpc_set_config(self.h, C.byref(self.cfg)), "mpc_set_config")

    # -- robot model
    def fk(self, q, out=None):
        import torch
        n = q.shape[0]
        out = out if out is not None else torch.empty((n, 6), device=q.device, dtype=torch.float32)
        _check(lib().mpc_fk(self.h, n, _ptr(q, n * q.shape[1], what="q"), q.shape[1], _ptr(out, 6 * n, what="ee"),
                            _stream()), "mpc_fk")
        return out

    def dynamics_gradient(self, qqdu):
        import torch
        n = qqdu.shape[0]
        qdd = torch.empty((n, NQ), device=qqdu.device, dtype=torch.float32)
        df = torch.empty((n, 147), device=qqdu.device, dtype=torch.float32)
        _check(lib().mpc_dynamics_gradient(self.h, n, _ptr(qqdu, 21 * n, what="qqdu"), _ptr(qdd), _ptr(df),
                                           _stream()), "mpc_dynamics_gradient")
        return qdd, df

    # -- stages (one instance, reference layouts)
    def form_kkt(self, N, xu, ee_ref, xs, G, Cm, g, c):
        _check(lib().mpc_form_kkt(self.h, N, _ptr(xu, 21 * N - 7, what="xu"), _ptr(ee_ref, 6 * N, what="ee_ref"),
                                  _ptr(xs, NX, what="xs"), _ptr(G, 245 * N - 49, what="G"),
                                  _ptr(Cm, 294 * (N - 1), what="C"), _ptr(g, 21 * N - 7, what="g"),
                                  _ptr(c, 14 * N, what="c"), _stream()), "mpc_form_kkt")

    def form_schur(self, N, G, Cm, g, c, S, Pinv, gamma, rho):
        _check(lib().mpc_form_schur(self.h, N, _ptr(G, 245 * N - 49, what="G"), _ptr(Cm, 294 * (N - 1), what="C"),
                                    _ptr(g, 21 * N - 7, what="g"), _ptr(c, 14 * N, what="c"),
                                    _ptr(S, 588 * N, what="S"), _ptr(Pinv, 588 * N, what="Pinv"),
                                    _ptr(gamma, 14 * N, what="gamma"), float(rho), _stream()), "mpc_form_schur")

    def pcg(self, N, S, Pinv, gamma, lam, max_iter, exit_tol, iters=None, flag=None):
        _check(lib().mpc_pcg_solve(self.h, N, _ptr(S, 588 * N, what="S"), _ptr(Pinv, 588 * N, what="Pinv"),
                                   _ptr(gamma, 14 * N, what="gamma"), _ptr(lam, 14 * N, what="lambda"), max_iter,
                                   float(exit_tol), _ptr(iters, 1, "u32", "iters"), _ptr(flag, 1, "u32", "flag"),
                                   _stream()), "mpc_pcg_solve")

    def direct_solve(self, N, S, gamma, lam, refine=2, work=None):
        """lambda = S^-1 gamma by block LDL^T on the GPU with `refine` steps of
        iterative refinement (work: 588 N floats)."""
        import torch
        work = torch.empty(588 * N, device="cuda") if work is None else work
        _check(lib().mpc_direct_solve(self.h, N, _ptr(S, 588 * N, what="S"), _ptr(gamma, 14 * N, what="gamma"),
                                      _ptr(lam, 14 * N, what="lambda"), int(refine), _ptr(work, 588 * N, what="work"),
                                      _stream()), "mpc_direct_solve")

    def compute_dz(self, N, Ginv, Cm, g, lam, dz):
        _check(lib().mpc_compute_dz(self.h, N, _ptr(Ginv, 245 * N - 49, what="Ginv"), _ptr(Cm, 294 * (N - 1), what="C"),
                                    _ptr(g, 21 * N - 7, what="g"), _ptr(lam, 14 * N, what="lambda"),
                                    _ptr(dz, 21 * N - 7, what="dz"), _stream()), "mpc_compute_dz")

    def compute_merit(self, N, xu, ee_ref, out):
        _check(lib().mpc_compute_merit(self.h, N, _ptr(xu, 21 * N - 7, what="xu"), _ptr(ee_ref, 6 * N, what="ee_ref"),
                                       _ptr(out, 1, what="merit"), _stream()), "mpc_compute_merit")

    def ls_compute_merit(self, N, xs, xu, ee_ref, dz, out):
        _check(lib().mpc_line_search_merits(self.h, N, _ptr(xs, NX, what="xs"), _ptr(xu, 21 * N - 7, what="xu"),
                                            _ptr(ee_ref, 6 * N, what="ee_ref"), _ptr(dz, 21 * N - 7, what="dz"),
                                            _ptr(out, 8, what="merits"), _stream()), "mpc_line_search_merits")

    def sqp_solve(self, N, ee_ref, lam, xu, rho, rho_reset=1e-3, stats=None):
        """Drop-in for sqpSolve: updates xu and lam in place; returns
        (pcg_iters, linsys_us, sqp_us, sqp_iters, not_aborted, pcg_maxiter, rho,
        ls_steps); `stats` (an SqpStats) receives the raw statistics."""
        r = C.c_float(rho)
        st = SqpStats() if stats is None else stats
        _check(lib().mpc_sqp_solve(self.h, N, _ptr(ee_ref, 6 * N, what="ee_ref"), _ptr(lam, 14 * N, what="lambda"),
                                   _ptr(xu, 21 * N - 7, what="xu"), C.byref(r), float(rho_reset), C.byref(st),
                                   _stream()), "mpc_sqp_solve")
        k = st.sqp_iters
        return (list(st.pcg_iters[:k]), list(st.linsys_us[:k]), st.sqp_us, k, bool(st.not_aborted),
                [bool(x) for x in st.pcg_maxiter[:k]], r.value, list(st.ls_step[:k]))

    # -- batched
    def batch_setup(self, horizons):
        arr = (C.c_uint32 * len(horizons))(*[int(x) for x in horizons])
        _check(lib().mpc_batch_setup(self.h, len(horizons), arr), "mpc_batch_setup")
        self._n_inst = len(horizons)
        self._T = int(sum(int(x) for x in horizons))

    def _layout(self):
        if getattr(self, "_n_inst", None) is None:
            raise MpcError("no batch layout (call batch_setup first)")
        return self._n_inst, self._T

    def batch_sqp(self, xu, lam, ee_ref, rho, sqp_iters, status, pcg_iters=None):
        n, T = self._layout()
        _check(lib().mpc_batch_sqp(self.h, _ptr(xu, 21 * T - 7 * n, what="xu"), _ptr(lam, 14 * T, what="lambda"),
                                   _ptr(ee_ref, 6 * T, what="ee_ref"), _ptr(rho, n, what="rho"), sqp_iters,
                                   _ptr(status, n, "u32", "status"), _ptr(pcg_iters, n * sqp_iters, "u32", "pcg_iters"),
                                   _stream()), "mpc_batch_sqp")

    def batch_set_timing(self, on=True):
        _check(lib().mpc_batch_set_timing(self.h, int(on)), "mpc_batch_set_timing")

    def batch_set_graph(self, on=True):
        """Run batch_sqp as a captured HIP graph, replayed while its buffers,
        iteration count, configuration and layout stay the same."""
        _check(lib().mpc_batch_set_graph(self.h, int(on)), "mpc_batch_set_graph")

    def set_cost_weights(self, wx, wu):
        import numpy as np
        wx = np.ascontiguousarray(np.asarray(wx, np.float32))
        wu = np.ascontiguousarray(np.asarray(wu, np.float32))
        if wx.shape != (14,) or wu.shape != (7,):
            raise MpcError("weights: wx has 14 entries, wu has 7")
        _check(lib().mpc_set_cost_weights(self.h, wx.ctypes.data, wu.ctypes.data), "mpc_set_cost_weights")

    def set_fault_injection(self, part_plus_one):
        """Test hook: part part_plus_one - 1 of every PCG team posts under a
        wrong epoch tag, so team solves end with flag 2 (0 turns it off)."""
        _check(lib().mpc_set_fault_injection(self.h, int(part_plus_one)), "mpc_set_fault_injection")

    def cost_weights(self):
        import numpy as np
        wx, wu = np.zeros(14, np.float32), np.zeros(7, np.float32)
        _check(lib().mpc_get_cost_weights(self.h, wx.ctypes.data, wu.ctypes.data), "mpc_get_cost_weights")
        return wx, wu

    def qp_layer(self, xu, ee_ref, rho, lam, dz, xstar, pcg_tol=0.0, refine=REFINE_STEPS, refine_tol=0.0):
        n, T = self._layout()
        L = 21 * T - 7 * n
        _check(lib().mpc_batch_qp_layer(self.h, _ptr(xu, L, what="xu"), _ptr(ee_ref, 6 * T, what="ee_ref"),
                                        _ptr(rho, n, what="rho"), _ptr(lam, 14 * T, what="lambda"), float(pcg_tol),
                                        int(refine), float(refine_tol), _ptr(dz, L, what="dz"),
                                        _ptr(xstar, L, what="xstar"), _stream()), "mpc_batch_qp_layer")

    def qp_loss(self, xu, ee_ref, loss, dLdx=None, dLdref=None):
        n, T = self._layout()
        L = 21 * T - 7 * n
        _check(lib().mpc_batch_qp_loss(self.h, _ptr(xu, L, what="xu"), _ptr(ee_ref, 6 * T, what="ee_ref"),
                                       _ptr(loss, n, what="loss"), _ptr(dLdx, L, what="dLdx"),
                                       _ptr(dLdref, 6 * T, what="dLdref"), _stream()), "mpc_batch_qp_loss")

    def qp_adjoint(self, xu, ee_ref, rho, dz, a, grad_w, grad_ref, pcg_tol=0.0, refine=REFINE_STEPS,
                   refine_tol=0.0):
        n, T = self._layout()
        L = 21 * T - 7 * n
        _check(lib().mpc_batch_qp_adjoint(self.h, _ptr(xu, L, what="xu"), _ptr(ee_ref, 6 * T, what="ee_ref"),
                                          _ptr(rho, n, what="rho"), _ptr(dz, L, what="dz"), _ptr(a, L, what="a"),
                                          float(pcg_tol), int(refine), float(refine_tol),
                                          _ptr(grad_w, 21 * n, what="grad_w"), _ptr(grad_ref, 6 * T, what="grad_ref"),
                                          _stream()), "mpc_batch_qp_adjoint")

    def inverse_dynamics(self, qqdqdd, out=None):
        """tau (n, 7) of n points [q qd qdd] (n, 21) on the device."""
        import torch
        n = qqdqdd.shape[0]
        out = torch.empty(n, NQ, device=qqdqdd.device) if out is None else out
        _check(lib().mpc_inverse_dynamics(self.h, n, _ptr(qqdqdd, 21 * n, what="qqdqdd"), _ptr(out, NQ * n, what="tau"),
                                          _stream()), "mpc_inverse_dynamics")
        return out

    def track_inputs(self, seed, steps, N, dt=0.01):
        """Device inputs of a closed-loop run on the seeded path: x0 (14), the
        reference window path ee (6 (steps + N)), and the precomputed trajectory
        (q_ref, qd_ref and the inverse-dynamics torques, xu layout over
        steps + N knots)."""
        import torch
        K = steps + N
        x0, qqdqdd = synth_track(seed, K, dt)
        pts = torch.from_numpy(qqdqdd).cuda()
        ee = self.fk(pts[:, :NQ].contiguous()).reshape(-1).contiguous()
        tau = self.inverse_dynamics(pts)
        traj = torch.zeros(21 * K, device="cuda")
        tv = traj.view(K, 21)
        tv[:, :NX] = pts[:, :NX]
        tv[:, NX:] = tau
        return torch.from_numpy(x0).cuda(), ee, traj[:21 * K - 7].contiguous()

    def closed_loop(self, N, x0, ee_path, xu, lam, cl=None, traj=None, states=None, ee_err=None, pcg_iters=None):
        """Closed-loop tracking on the device (see mpc_closed_loop); returns
        (states (steps+1, 14), ee_err (steps, 2), pcg_iters (steps, K) or None, stats dict)."""
        import torch
        cl = default_closed_loop() if cl is None else cl
        steps, K = cl.steps, cl.sqp_iters
        states = torch.empty(steps + 1, NX, device="cuda") if states is None else states
        ee_err = torch.empty(steps, 2, device="cuda") if ee_err is None else ee_err
        st = ClosedLoopStats()
        K_all = steps + N
        _check(lib().mpc_closed_loop(self.h, C.byref(cl), N, _ptr(x0, NX, what="x0"),
                                     _ptr(ee_path, 6 * K_all, what="ee_path"),
                                     _ptr(traj, 21 * K_all - 7, what="traj"), _ptr(xu, 21 * N - 7, what="xu"),
                                     _ptr(lam, 14 * N, what="lambda"), _ptr(states, NX * (steps + 1), what="states"),
                                     _ptr(ee_err, 2 * steps, what="ee_err"),
                                     _ptr(pcg_iters, steps * K, "u32", "pcg_iters"), C.byref(st), _stream()),
               "mpc_closed_loop")
        self._n_inst, self._T = 1, N  # the run sets the batch layout to its one instance
        stats = {k: getattr(st, k) for k, _ in ClosedLoopStats._fields_ if k != "reserved"}
        return states, ee_err, pcg_iters, stats

    def batch_pcg_layout(self):
        """{bins, teams, streamed, team_launches} of the current batch layout."""
        out = (C.c_uint32 * 4)()
        _check(lib().mpc_batch_pcg_layout(self.h, out), "mpc_batch_pcg_layout")
        return dict(zip(("bins", "teams", "streamed", "team_launches"), list(out)))

    def batch_pcg_launch_map(self):
        """Per instance, the PCG launch of a phase that solves it: 0 .. L-1
        the team launches (the last also holds the first resident bins), then
        the streaming launch (if any), then the resident launch."""
        import numpy as np
        n = getattr(self, "_n_inst", None)
        if n is None:
            raise MpcError("batch_pcg_launch_map: no batch layout (call batch_setup first)")
        out = np.zeros(n, np.uint32)
        _check(lib().mpc_batch_pcg_launch_map(self.h, out.ctypes.data), "mpc_batch_pcg_launch_map")
        return out

    def batch_get_plan(self):
        """The current PCG launch plan {small_nt, max_items, mixed, teams, split}."""
        p = PcgPlan()
        _check(lib().mpc_batch_get_plan(self.h, C.byref(p)), "mpc_batch_get_plan")
        return {k: int(getattr(p, k)) for k in _PLAN_KEYS}

    def batch_set_plan(self, **kw):
        """Set fields of the PCG launch plan (the others kept) and re-pack."""
        cur = self.batch_get_plan()
        unknown = set(kw) - set(_PLAN_KEYS)
        if unknown:
            raise MpcError(f"batch_set_plan: unknown fields {sorted(unknown)}")
        cur.update({k: int(v) for k, v in kw.items()})
        p = PcgPlan(**cur)
        _check(lib().mpc_batch_set_plan(self.h, C.byref(p)), "mpc_batch_set_plan")

    def batch_autotune(self, iters=16, reps=3):
        """Time every launch plan that differs for the current layout and keep
        the fastest; returns the timed plans (dicts with "us") in order."""
        res = (PcgPlan * 64)()
        n = C.c_uint32(0)
        _check(lib().mpc_batch_autotune(self.h, int(iters), int(reps), res, 64, C.byref(n), _stream()),
               "mpc_batch_autotune")
        return [dict({k: int(getattr(r, k)) for k in _PLAN_KEYS}, us=float(r.us)) for r in res[:n.value]]

    def batch_pcg_exit(self, iters=None, flags=None, eta=None):
        """Per instance, the last PCG solve's iterations, max-iter flag and
        |r^T Pinv r| at its last exit test, into the given CUDA tensors."""
        n, _ = self._layout()
        _check(lib().mpc_batch_pcg_exit(self.h, _ptr(iters, n, "u32", "iters"), _ptr(flags, n, "u32", "flags"),
                                        _ptr(eta, n, what="eta"), _stream()), "mpc_batch_pcg_exit")

    def batch_direct_resolve(self, lam_out, refine=2):
        """The last batched Schur systems solved directly into lam_out (the
        batch's lambda layout); overwrites the preconditioner workspace."""
        _, T = self._layout()
        _check(lib().mpc_batch_direct_resolve(self.h, _ptr(lam_out, 14 * T, what="lambda"), int(refine), _stream()),
               "mpc_batch_direct_resolve")

    def batch_phase_times(self):
        out = (C.c_double * 7)()
        _check(lib().mpc_batch_phase_times(self.h, out), "mpc_batch_phase_times")
        return list(out)
