"""ctypes mirror of ``include/simplediffeq_hip.h`` (ABI version 1).

Only plain ctypes: the structs below must match the C header field for field.
``tests/test_abi.py`` checks the sizes against the compiled library.
"""
import ctypes
import os

ABI_VERSION = 5

# sdeh_status
OK = 0
E_INVALID = -1
E_HIP = -2
E_COMPILE = -3
E_UNSUPPORTED = -4
E_NOMEM = -5
E_DTMIN = -6
E_SAVE_OVERFLOW = -7
STATUS_NAMES = {0: "OK", -1: "INVALID", -2: "HIP", -3: "COMPILE", -4: "UNSUPPORTED",
                -5: "NOMEM", -6: "DTMIN", -7: "SAVE_OVERFLOW"}

# sdeh_alg
ALG_TSIT5, ALG_ATSIT5, ALG_VERN7, ALG_AVERN7, ALG_VERN9, ALG_AVERN9, ALG_RK4, ALG_EULER = range(8)
FIXED_GRID_ONLY_ALGS = (ALG_RK4, ALG_EULER)  # no interpolant: every grid point is returned
ADAPTIVE_ALGS = (ALG_ATSIT5, ALG_AVERN7, ALG_AVERN9)

# sdeh_dtype
F64, F32 = 0, 1

# sdeh_retcode
RET_DEFAULT, RET_DTLESSTHANMIN, RET_NAN, RET_MAXITERS, RET_SAVE_OVERFLOW = range(5)
RETCODE_NAMES = {0: "Default", 1: "DtLessThanMin", 2: "NaN", 3: "MaxIters", 4: "SaveOverflow"}

# sdeh_builtin_rhs: id -> (name, n_state, n_param, n_shared)
RHS_LORENZ, RHS_GLV16, RHS_EXPDECAY3, RHS_AFFINE1, RHS_LINEAR2 = 1, 2, 3, 4, 5
BUILTIN_RHS = {
    RHS_LORENZ: ("lorenz", 3, 3, 0),
    RHS_GLV16: ("glv16", 16, 16, 256),
    RHS_EXPDECAY3: ("expdecay3", 3, 1, 0),
    RHS_AFFINE1: ("affine1", 1, 2, 0),
    RHS_LINEAR2: ("linear2", 2, 4, 0),
}

# sdeh_schedule / precision / memory
SAVE_POINT_MAJOR, SAVE_TRAJ_MAJOR, SAVE_PACKED = 0, 1, 2
SCHED_AUTO, SCHED_ONE_LAUNCH, SCHED_PERSISTENT, SCHED_STAGED = 0, 1, 2, 3
PREC_PLAIN, PREC_KAHAN = 0, 1
MEM_HOST, MEM_DEVICE = 0, 1

MLP_NIN, MLP_NH, MLP_NOUT = 6, 64, 2
MLP_NPARAM = MLP_NH * MLP_NIN + MLP_NH + MLP_NH * MLP_NH + MLP_NH + MLP_NOUT * MLP_NH + MLP_NOUT


class SolveDesc(ctypes.Structure):
    _fields_ = [
        ("alg", ctypes.c_int32),
        ("dtype", ctypes.c_int32),
        ("rhs", ctypes.c_int64),
        ("n_traj", ctypes.c_int64),
        ("t0", ctypes.c_double),
        ("tf", ctypes.c_double),
        ("dt", ctypes.c_double),
        ("abstol", ctypes.c_double),
        ("reltol", ctypes.c_double),
        ("save_everystep", ctypes.c_int32),
        ("max_saves", ctypes.c_int32),
        ("saveat", ctypes.POINTER(ctypes.c_double)),
        ("n_saveat", ctypes.c_int64),
        ("maxiters", ctypes.c_int64),
        ("quirks", ctypes.c_int32),
        ("schedule", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("block", ctypes.c_int32),
        ("grid", ctypes.c_int32),
        ("chunk", ctypes.c_int32),
        ("lanes", ctypes.c_int32),
        ("save_layout", ctypes.c_int32),
        ("span_order", ctypes.c_int32),
        ("waves", ctypes.c_int32),
        ("per_lane", ctypes.c_int32),
        ("refill", ctypes.c_int32),
    ]


class SolveIO(ctypes.Structure):
    _fields_ = [
        ("u0", ctypes.c_void_p),
        ("p", ctypes.c_void_p),
        ("p_shared", ctypes.c_void_p),
        ("n_shared", ctypes.c_int64),
        ("tspan", ctypes.c_void_p),
        ("dt0", ctypes.c_void_p),
        ("order", ctypes.c_void_p),
        ("u_end", ctypes.c_void_p),
        ("t_end", ctypes.c_void_p),
        ("retcode", ctypes.c_void_p),
        ("naccept", ctypes.c_void_p),
        ("nreject", ctypes.c_void_p),
        ("u_saveat", ctypes.c_void_p),
        ("u_steps", ctypes.c_void_p),
        ("t_steps", ctypes.c_void_p),
        ("n_steps", ctypes.c_void_p),
        ("steps_offset", ctypes.c_void_p),
        ("steps_capacity", ctypes.c_int64),
        ("visits", ctypes.c_void_p),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("kernel_ms", ctypes.c_double),
        ("total_ms", ctypes.c_double),
        ("attempts", ctypes.c_int64),
        ("accepts", ctypes.c_int64),
        ("block", ctypes.c_int32),
        ("grid", ctypes.c_int32),
        ("chunk", ctypes.c_int32),
        ("schedule", ctypes.c_int32),
        ("lanes", ctypes.c_int32),
        ("waves", ctypes.c_int32),
        ("steps_needed", ctypes.c_int64),
        ("per_lane", ctypes.c_int32),
        ("refill", ctypes.c_int32),
    ]


def default_desc():
    """The reference solve defaults (gpuatsit5.jl:205-212): dt = 0.1f0,
    abstol = 1f-6, reltol = 1f-3, save_everystep = true."""
    d = SolveDesc()
    d.alg = ALG_ATSIT5
    d.dtype = F64
    d.rhs = RHS_LORENZ
    d.n_traj = 0
    d.t0, d.tf = 0.0, 1.0
    d.dt = float(_f32(0.1))
    d.abstol = float(_f32(1e-6))
    d.reltol = float(_f32(1e-3))
    d.save_everystep = 1
    d.max_saves = 0
    d.saveat = None
    d.n_saveat = 0
    d.maxiters = 0
    d.quirks = 1
    d.schedule = SCHED_AUTO
    d.precision = PREC_PLAIN
    d.block = d.grid = d.chunk = 0
    return d


def _f32(x):
    import numpy as np
    return np.float32(x)


# Exported symbols of libsimplediffeq_hip (every function include/simplediffeq_hip.h declares).
EXPORTS = [
    "sdeh_desc_init", "sdeh_abi_version", "sdeh_create", "sdeh_destroy", "sdeh_last_error",
    "sdeh_rhs_info", "sdeh_compile_rhs", "sdeh_solve", "sdeh_solve_static", "sdeh_get_stats",
    "sdeh_graph_capture",
    "sdeh_graph_launch", "sdeh_graph_destroy", "sdeh_autotune", "sdeh_synchronize",
    "sdeh_mlp_forward", "sdeh_mlp_plan_lorenz", "sdeh_plan_order", "sdeh_mlp_loss_grad",
    "sdeh_mlp_adam", "sdeh_fixed_grid",
]

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_ROOT, "lib", "libsimplediffeq_hip.so")

_lib = None


def load_library(path=None):
    """Load libsimplediffeq_hip.so and declare its signatures.  Raises if the
    library is missing: there is no fallback path."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError("libsimplediffeq_hip.so not built (expected at %s); run "
                           "__graft_entry__.build()" % p)
    # One HIP runtime per process: torch ships its own libamdhip64.so.7 and
    # loads it by file name, so if this library pulled in the system copy first
    # the process would hold two runtimes and torch would find no GPU.  Loading
    # torch first makes this library bind to the copy torch already mapped.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(p)
    vp, i32, i64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64
    lib.sdeh_desc_init.argtypes = [ctypes.POINTER(SolveDesc)]
    lib.sdeh_desc_init.restype = None
    lib.sdeh_abi_version.restype = ctypes.c_int
    lib.sdeh_create.argtypes = [ctypes.POINTER(vp), ctypes.c_int]
    lib.sdeh_create.restype = ctypes.c_int
    lib.sdeh_destroy.argtypes = [vp]
    lib.sdeh_destroy.restype = None
    lib.sdeh_last_error.argtypes = [vp]
    lib.sdeh_last_error.restype = ctypes.c_char_p
    lib.sdeh_rhs_info.argtypes = [vp, i64, ctypes.POINTER(i32), ctypes.POINTER(i32),
                                  ctypes.POINTER(i64)]
    lib.sdeh_rhs_info.restype = ctypes.c_int
    lib.sdeh_compile_rhs.argtypes = [vp, ctypes.c_char_p, ctypes.c_char_p, i32, i32, i64,
                                     ctypes.POINTER(i64)]
    lib.sdeh_compile_rhs.restype = ctypes.c_int
    lib.sdeh_solve.argtypes = [vp, ctypes.POINTER(SolveDesc), ctypes.POINTER(SolveIO), i32, vp,
                               i32]
    lib.sdeh_solve.restype = ctypes.c_int
    lib.sdeh_solve_static.argtypes = [vp, ctypes.POINTER(SolveDesc), ctypes.POINTER(SolveIO),
                                      vp, i32]
    lib.sdeh_solve_static.restype = ctypes.c_int
    lib.sdeh_get_stats.argtypes = [vp, ctypes.POINTER(Stats)]
    lib.sdeh_get_stats.restype = ctypes.c_int
    lib.sdeh_graph_capture.argtypes = [vp, ctypes.POINTER(SolveDesc), ctypes.POINTER(SolveIO),
                                       ctypes.POINTER(i64)]
    lib.sdeh_graph_capture.restype = ctypes.c_int
    lib.sdeh_graph_launch.argtypes = [vp, i64, vp]
    lib.sdeh_graph_launch.restype = ctypes.c_int
    lib.sdeh_graph_destroy.argtypes = [vp, i64]
    lib.sdeh_graph_destroy.restype = ctypes.c_int
    lib.sdeh_autotune.argtypes = [vp, ctypes.POINTER(SolveDesc), ctypes.POINTER(SolveIO),
                                  ctypes.POINTER(Stats)]
    lib.sdeh_autotune.restype = ctypes.c_int
    lib.sdeh_synchronize.argtypes = [vp]
    lib.sdeh_synchronize.restype = ctypes.c_int
    lib.sdeh_mlp_forward.argtypes = [vp, vp, vp, vp, i64, vp]
    lib.sdeh_mlp_forward.restype = ctypes.c_int
    lib.sdeh_mlp_plan_lorenz.argtypes = [vp, vp, vp, vp, vp, i64, vp, vp, vp, vp, vp, vp]
    lib.sdeh_mlp_plan_lorenz.restype = ctypes.c_int
    lib.sdeh_plan_order.argtypes = [vp, vp, i64, vp, vp]
    lib.sdeh_plan_order.restype = ctypes.c_int
    lib.sdeh_mlp_loss_grad.argtypes = [vp, vp, vp, vp, i64, vp, vp, vp]
    lib.sdeh_mlp_loss_grad.restype = ctypes.c_int
    lib.sdeh_mlp_adam.argtypes = [vp, vp, vp, vp, vp, ctypes.c_float, ctypes.c_float,
                                  ctypes.c_float, ctypes.c_float, i32, vp]
    lib.sdeh_mlp_adam.restype = ctypes.c_int
    if lib.sdeh_abi_version() != ABI_VERSION:
        raise RuntimeError("libsimplediffeq_hip ABI version mismatch")
    if path is None:
        _lib = lib
    return lib
