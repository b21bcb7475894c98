"""The C ABI: the library loads, exports every function the header declares,
and the ctypes structs match the C layout.  No GPU needed."""
import ctypes
import os
import re
import subprocess
import tempfile

import pytest

from simplediffeq_hip import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "simplediffeq_hip.h")


def declared_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|void|int64_t|const char\*)\s+(sdeh_\w+)\s*\(", txt, re.M)))


def test_header_declarations_are_all_exported():
    lib = _abi.load_library()
    decl = declared_functions()
    assert set(decl) == set(_abi.EXPORTS)
    for name in decl:
        assert hasattr(lib, name), name
    assert lib.sdeh_abi_version() == _abi.ABI_VERSION


def test_struct_layout_matches_c():
    src = r'''
#include <stddef.h>
#include <stdio.h>
#include "simplediffeq_hip.h"
int main(void) {
  printf("%zu %zu %zu\n", sizeof(sdeh_solve_desc), sizeof(sdeh_io), sizeof(sdeh_stats));
  printf("%zu %zu %zu %zu\n", offsetof(sdeh_solve_desc, saveat), offsetof(sdeh_solve_desc, maxiters),
         offsetof(sdeh_solve_desc, chunk), offsetof(sdeh_io, n_steps));
  return 0;
}
'''
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.c")
        open(c, "w").write(src)
        exe = os.path.join(d, "t")
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()
    sizes = [int(x) for x in out]
    assert sizes[:3] == [ctypes.sizeof(_abi.SolveDesc), ctypes.sizeof(_abi.SolveIO),
                         ctypes.sizeof(_abi.Stats)]
    assert sizes[3:] == [_abi.SolveDesc.saveat.offset, _abi.SolveDesc.maxiters.offset,
                         _abi.SolveDesc.chunk.offset, _abi.SolveIO.n_steps.offset]


def test_desc_defaults_mirror_reference_keywords():
    lib = _abi.load_library()
    d = _abi.SolveDesc()
    lib.sdeh_desc_init(ctypes.byref(d))
    # gpuatsit5.jl:208-210: dt = 0.1f0, abstol = 1f-6, reltol = 1f-3, save_everystep = true
    import numpy as np
    assert d.dt == float(np.float32(0.1)) and d.abstol == float(np.float32(1e-6))
    assert d.reltol == float(np.float32(1e-3)) and d.save_everystep == 1 and d.quirks == 1
    py = _abi.default_desc()
    for f, _ in _abi.SolveDesc._fields_:
        if f not in ("saveat", "reserved"):
            assert getattr(py, f) == getattr(d, f), f


def test_create_without_gpu_fails_cleanly():
    try:
        import torch
        if torch.cuda.device_count() > 0:
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    lib = _abi.load_library()
    h = ctypes.c_void_p()
    st = lib.sdeh_create(ctypes.byref(h), 0)
    assert st != 0 and not h.value
