"""pdx_node_stresses_noalloc: the nodal stress recovery with its scratch in the library's fixed
device array (no allocation in the call), against pdx_node_stresses on the same inputs (same
kernels, so the same bits) and against the oracle (1e-12 of the largest component), on two calls
in a row through the C-ABI, the second on a smaller model, so it runs on the scratch the first
call left behind.  Device buffers for the caller's inputs and outputs come from the HIP runtime
the library itself is linked with."""
import ctypes as C
import os

import numpy as np
import pytest

from helpers import load_oracle, load_pdx, workloads

pytestmark = pytest.mark.gpu

O = load_oracle()
W = workloads()
pdx = load_pdx()


def _loaded_hip_path():
    """The file of the HIP runtime mapped into this process: the one libpdx.so resolved, which is
    torch's copy when torch was imported first and ROCm's otherwise."""
    with open("/proc/self/maps") as f:
        for line in f:
            path = line.split()[-1] if len(line.split()) >= 6 else ""
            if os.path.basename(path).startswith("libamdhip64.so"):
                return path
    raise RuntimeError("no HIP runtime mapped after loading libpdx.so")


def _hip():
    pdx.lib()  # the library first: its HIP runtime is the one these buffers must come from
    h = C.CDLL(_loaded_hip_path())
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


H2D, D2H = 1, 2  # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


def _noalloc(cx, u):
    h = _hip()
    n = cx.model.nn
    u = np.ascontiguousarray(u, np.float64)
    out = np.full(n * 6, np.nan)
    du, ds = C.c_void_p(), C.c_void_p()
    assert h.hipMalloc(C.byref(du), u.nbytes) == 0 and h.hipMalloc(C.byref(ds), out.nbytes) == 0
    try:
        assert h.hipMemcpy(du, u.ctypes.data, u.nbytes, H2D) == 0
        assert h.hipMemcpy(ds, out.ctypes.data, out.nbytes, H2D) == 0  # NaN: every output must be written
        rc = pdx.lib().pdx_node_stresses_noalloc(cx.h, du, ds)
        assert h.hipDeviceSynchronize() == 0  # the launches are asynchronous on the context's stream
        assert h.hipMemcpy(out.ctypes.data, ds, out.nbytes, D2H) == 0
    finally:
        h.hipFree(du)
        h.hipFree(ds)
    return rc, out.reshape(n, 6)


def test_noalloc_stresses_match_the_allocating_entry_point():
    cases = (W.box3d(n=14, pd_block=6), W.crack2d(seed=3, nx=31, ny=17, dx=1e-3))
    rng = np.random.default_rng(7)
    for mesh in cases:  # the second call is the smaller model
        mp = pdx.Model(mesh)
        cx = pdx.Context(mp)
        u = rng.uniform(-1e-6, 1e-6, mp.nn * mp.dim)
        want = cx.node_stresses(u)  # assembles K, allocating path
        rc, got = _noalloc(cx, u)
        assert rc == 0, pdx.lib().pdx_last_error().decode()
        assert np.array_equal(got, want)
        ref = O.Model(mesh).node_stresses(u)
        assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()


def test_noalloc_stresses_need_an_assembled_K():
    cx = pdx.Context(pdx.Model(W.box3d(n=6, pd_block=2)))
    rc, _ = _noalloc(cx, np.zeros(cx.model.nn * 3))
    assert rc == -1  # PDX_ERR_ARG, nothing launched
    assert "not assembled" in pdx.lib().pdx_last_error().decode()
