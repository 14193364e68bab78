"""Krylov SpMVs over mirrored storage (pdx_solve_opts.mirrored): each lower block that no family
stores in either orientation is read as its upper partner's transpose.  The blocks are the same
values, so the products and their order are the same: solves and Newmark runs must agree with the
full-storage runs bit for bit, iteration counts included.  The device check that enables the
storage must pass for the stiffness of 2D and 3D models and for the 3D Newmark operator
beta dt^2 K + M (hex8 consistent mass is symmetric term by term)."""
import numpy as np
import pytest

from helpers import load_pdx, workloads

W = workloads()
pdx = load_pdx()

STATIC = [
    ("box3d_fp64_block_jacobi", lambda: W.box3d(n=14, pd_block=6), pdx.PDX_PREC_BLOCK_JACOBI, pdx.PDX_FP64),
    ("box3d_fp32_two_level", lambda: W.box3d(n=14, pd_block=6), pdx.PDX_PREC_TWO_LEVEL, pdx.PDX_FP32),
    ("crack2d_fp64_jacobi", lambda: W.crack2d(seed=3, nx=81, ny=41, dx=1e-3), pdx.PDX_PREC_JACOBI, pdx.PDX_FP64),
    ("crack2d_fp32_ilu0", lambda: W.crack2d(seed=3, nx=81, ny=41, dx=1e-3), pdx.PDX_PREC_ILU0, pdx.PDX_FP32),
]


@pytest.mark.gpu
@pytest.mark.parametrize("name,make,precond,precision", STATIC, ids=[c[0] for c in STATIC])
def test_static_mirrored_is_bitwise_full(name, make, precond, precision):
    mesh = make()
    kw = dict(precond=precond, precision=precision, tol=1e-8, max_iter=20000, coarse_agg=4)
    u0, s0 = pdx.Context(pdx.Model(mesh)).static_solve(pdx.solve_opts(**kw))
    cx = pdx.Context(pdx.Model(mesh))
    for rep in range(2):  # the second solve reuses the checked copies (and replays the graphs)
        u1, s1 = cx.static_solve(pdx.solve_opts(mirrored=True, **kw))
        print(f"{name} rep {rep}: iterations {s1['iterations']} / {s0['iterations']}, "
              f"mirrored {s1['mirrored_used']}", flush=True)
        assert s0["mirrored_used"] == 0 and s1["mirrored_used"] == 1
        assert s1["iterations"] == s0["iterations"] and s1["refinements"] == s0["refinements"]
        assert np.array_equal(u1, u0)
    # a new generation of K is checked and gathered again
    cx.assemble(pdx.PDX_MODE_STATIC)
    u2, s2 = cx.static_solve(pdx.solve_opts(mirrored=True, **kw))
    assert s2["mirrored_used"] == 1 and np.array_equal(u2, u0)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [pdx.PDX_FP64, pdx.PDX_FP32], ids=["fp64", "fp32"])
def test_newmark_mirrored_is_bitwise_full_3d(precision):
    mesh = W.penny3d(n=16, pd_block=6)
    kw = dict(precision=precision, precond=pdx.PDX_PREC_BLOCK_JACOBI, tol=1e-9)
    a = pdx.Context(pdx.Model(mesh)).newmark_run(4, pdx.solve_opts(**kw))
    b = pdx.Context(pdx.Model(mesh)).newmark_run(4, pdx.solve_opts(mirrored=True, **kw))
    print(f"penny3d newmark: iterations {b[3]['iterations']} / {a[3]['iterations']}, "
          f"mirrored {b[3]['mirrored_used']}", flush=True)
    assert b[3]["mirrored_used"] == 1
    assert a[3]["iterations"] == b[3]["iterations"]
    for x, y in zip(a[:3], b[:3]):
        assert np.array_equal(x, y)


@pytest.mark.gpu
def test_newmark_mirrored_2d_same_result_either_way():
    # quad4 consistent mass is formed as (N_a rho) N_b detJ, which need not be symmetric in the
    # last bit: the device check decides, and the result is the full-storage one in both cases
    mesh = W.plate2d(nx=41, ny=21, dx=1e-3, pd_x=(0.015, 0.025))
    mesh["ebc"][2].update(val=0.0, vel=1e-2, vary=True)
    mesh["dt"] = 1e-6
    kw = dict(precision=pdx.PDX_FP32, precond=pdx.PDX_PREC_BLOCK_JACOBI, tol=1e-10)
    a = pdx.Context(pdx.Model(mesh)).newmark_run(4, pdx.solve_opts(**kw))
    b = pdx.Context(pdx.Model(mesh)).newmark_run(4, pdx.solve_opts(mirrored=True, **kw))
    print(f"plate2d newmark: mirrored {b[3]['mirrored_used']}", flush=True)
    assert a[3]["iterations"] == b[3]["iterations"]
    for x, y in zip(a[:3], b[:3]):
        assert np.array_equal(x, y)


@pytest.mark.gpu
def test_csr_solve_ignores_mirrored():
    # the PARDISO seam's 1 x 1-block operator has no mirrored storage
    import scipy.sparse as sp
    n = 200
    A = sp.diags([-np.ones(n - 1), 4.0 * np.ones(n), -np.ones(n - 1)], [-1, 0, 1], format="csr")
    b = np.linspace(1.0, 2.0, n)
    x0, s0 = pdx.solve_csr(A.indptr, A.indices, A.data, b, pdx.solve_opts())
    x1, s1 = pdx.solve_csr(A.indptr, A.indices, A.data, b, pdx.solve_opts(mirrored=True))
    assert s1["mirrored_used"] == 0 and np.array_equal(x0, x1)
