"""Krylov iterations replayed from captured hipGraphs (chunks of 10 iterations)
give the same iterates, bit for bit, as launching every kernel one by one: the
reductions are fixed-order, so a replay must reproduce the eager run exactly.
Each case solves twice on one context (the second solve replays the graphs the
first captured, with the host's residual checks and copies between replays),
re-assembles K in between where stated (new operator generation: re-capture),
and compares with a no-graph run on a fresh context."""
import numpy as np
import pytest

from helpers import load_oracle, load_pdx, workloads

O = load_oracle()
W = workloads()
pdx = load_pdx()

CASES = [
    ("bicgstab_jacobi_fp64", pdx.PDX_KRYLOV_BICGSTAB, pdx.PDX_PREC_JACOBI, pdx.PDX_FP64),
    ("bicgstab_two_level_fp64", pdx.PDX_KRYLOV_BICGSTAB, pdx.PDX_PREC_TWO_LEVEL, pdx.PDX_FP64),
    ("bicgstab_two_level_fp32", pdx.PDX_KRYLOV_BICGSTAB, pdx.PDX_PREC_TWO_LEVEL, pdx.PDX_FP32),
    ("bicgstab_ilu0_fp64", pdx.PDX_KRYLOV_BICGSTAB, pdx.PDX_PREC_ILU0, pdx.PDX_FP64),
    ("bicgstab_ilu0_fp32", pdx.PDX_KRYLOV_BICGSTAB, pdx.PDX_PREC_ILU0, pdx.PDX_FP32),
]


def _mesh():
    return W.crack2d(seed=5, nx=121, ny=61, dx=1e-3)


@pytest.mark.gpu
@pytest.mark.parametrize("name,krylov,precond,precision", CASES, ids=[c[0] for c in CASES])
def test_graph_replay_is_bitwise_eager(name, krylov, precond, precision):
    mesh = _mesh()
    kw = dict(krylov=krylov, precond=precond, precision=precision, tol=1e-8, max_iter=20000)
    ref = pdx.Context(pdx.Model(mesh))
    u0, s0, rc0 = ref.static_solve(pdx.solve_opts(no_graph=True, **kw), check=False)
    cx = pdx.Context(pdx.Model(mesh))
    for rep in range(2):
        u1, s1, rc1 = cx.static_solve(pdx.solve_opts(**kw), check=False)
        print(f"{name} replay {rep}: rc {rc1} iterations {s1['iterations']} relres {s1['true_relres']:.3e}", flush=True)
        assert rc1 == rc0 and s1["iterations"] == s0["iterations"] and s1["refinements"] == s0["refinements"]
        assert np.array_equal(u1, u0)
        assert s1["iterations"] >= 20  # at least one replay of each captured chunk
    # a new operator generation (re-assembly) re-captures and still matches
    cx.assemble(pdx.PDX_MODE_STATIC)
    u2, s2, rc2 = cx.static_solve(pdx.solve_opts(**kw), check=False)
    assert rc2 == rc0 and np.array_equal(u2, u0)


@pytest.mark.gpu
def test_graph_replay_cg_symmetric_plate():
    mesh = W.plate2d(nx=61, ny=31, dx=1e-3, pd_x=(0.0, 0.0))
    kw = dict(krylov=pdx.PDX_KRYLOV_CG, precond=pdx.PDX_PREC_BLOCK_JACOBI, precision=pdx.PDX_FP64, tol=1e-8)
    u0, s0 = pdx.Context(pdx.Model(mesh)).static_solve(pdx.solve_opts(no_graph=True, **kw))
    cx = pdx.Context(pdx.Model(mesh))
    for _ in range(2):
        u1, s1 = cx.static_solve(pdx.solve_opts(**kw))
        assert s1["iterations"] == s0["iterations"] and np.array_equal(u1, u0)


@pytest.mark.gpu
def test_graph_replay_newmark():
    mesh = W.plate2d(nx=41, ny=21, dx=1e-3, pd_x=(0.015, 0.025))
    mesh["ebc"][2].update(val=0.0, vel=1e-2, vary=True)
    mesh["dt"] = 1e-6
    kw = dict(precision=pdx.PDX_FP32, precond=pdx.PDX_PREC_TWO_LEVEL, tol=1e-10, coarse_agg=6)
    a = pdx.Context(pdx.Model(mesh)).newmark_run(6, pdx.solve_opts(no_graph=True, **kw))
    b = pdx.Context(pdx.Model(mesh)).newmark_run(6, pdx.solve_opts(**kw))
    for x, y in zip(a[:3], b[:3]):
        assert np.array_equal(x, y)
    assert a[3]["iterations"] == b[3]["iterations"]
