"""ORACLE — TEST INFRASTRUCTURE ONLY.

ctypes loader of the CPU restatement (oracle/liboracle.so, built from
oracle_capi.cpp by oracle/Makefile).  Only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg may use it, as the checker or the timed CPU
baseline; the product library never calls it.
"""
import ctypes
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "liboracle.so")

ARITH_FUSED, ARITH_PLAIN = 0, 1
POW_PORTABLE, POW_LIBM = 0, 1

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("oracle not built: run `make -C oracle`")
    lib = ctypes.CDLL(LIB_PATH)
    lib.oracle_solve.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_int]
    lib.oracle_solve.restype = ctypes.c_int
    lib.oracle_pow_ctrl.argtypes = [ctypes.c_double, ctypes.c_double]
    lib.oracle_pow_ctrl.restype = ctypes.c_double
    lib.oracle_log_dd.argtypes = [ctypes.c_double, ctypes.c_void_p]
    lib.oracle_log_dd.restype = None
    lib.oracle_exp_dd.argtypes = [ctypes.c_double, ctypes.c_double]
    lib.oracle_exp_dd.restype = ctypes.c_double
    lib.oracle_pow_ctrl_domain.argtypes = [ctypes.c_double, ctypes.c_double]
    lib.oracle_pow_ctrl_domain.restype = ctypes.c_double
    lib.oracle_minmax.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double]
    lib.oracle_minmax.restype = ctypes.c_double
    lib.oracle_range.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                 ctypes.c_void_p, ctypes.c_int64, ctypes.POINTER(ctypes.c_int32)]
    lib.oracle_range.restype = ctypes.c_int64
    lib.oracle_btheta.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_void_p]
    lib.oracle_btheta.restype = ctypes.c_int
    _lib = lib
    return lib


def solve(desc, bufs, arith=ARITH_FUSED, pow_kind=POW_PORTABLE, nthreads=1):
    """Run the restatement on host EnsembleBuffers (numpy backend).  Returns the
    sdeh_status the library would return."""
    if bufs.backend != "numpy":
        raise ValueError("the oracle runs on host buffers")
    lib = load()
    return lib.oracle_solve(ctypes.addressof(desc), ctypes.addressof(bufs.io), int(arith),
                            int(pow_kind), int(nthreads))


def julia_range(start, step, stop, cap=100000):
    import numpy as np
    lib = load()
    out = np.empty(cap, dtype=np.float64)
    rat = ctypes.c_int32()
    n = lib.oracle_range(float(start), float(step), float(stop), out.ctypes.data, cap,
                         ctypes.byref(rat))
    return n, out[:min(n, cap)], bool(rat.value)


def btheta(which, theta):
    import numpy as np
    out = np.empty(19, dtype=np.float64)
    k = load().oracle_btheta(int(which), float(theta), out.ctypes.data)
    return out[:k]
