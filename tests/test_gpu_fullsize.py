"""Properties at the benchmark sizes, where the oracle would take too long: the 60^3 explicit box
(644,397 equations) and the 100^3 penny-crack Newmark model (2.97M equations).

  rigid translations   K_ff t = 0 for t a unit translation in x, y or z, on every row whose node
                       has no prescribed neighbour (every family, element and boundary operator
                       of PDLSM-FEM annihilates rigid translations): max |(K_ff t)_i| / ||row_i||_1
                       <= 1e-12; a dropped term or a wrong sign in any assembly kernel breaks it
  linearity            zero initial state, twice the traction: every state of the explicit step
                       is exactly twice the first run's in FP64 (scaling by 2 is exact in every
                       operation), bit for bit; in FP32 force to 1e-12 (FP32 flushes the
                       displacement differences at the wave front's far tail, below its smallest
                       normal number, so there the scaling is not exact)
  precision            FP32 force / FP64 state against FP64 after 200 steps: <= 1e-6 relative
  Newmark              FP32 BiCGStab with FP64 refinement and the two-level preconditioner against
                       plain FP64 block-Jacobi BiCGStab, 3 steps of the 100^3 model: displacements
                       <= 1e-7 relative, true relative residual <= 1e-8 in both
  set-up on the GPU    the 100^3 model's families and node graph built on the GPU equal the host
                       set-up's bit for bit
"""
import numpy as np
import pytest
import scipy.sparse as sp

from helpers import load_pdx, rel, workloads

pytestmark = pytest.mark.gpu

W = workloads()
pdx = load_pdx()


def _box_dt(cx):
    return 0.8 * 2.0 / np.sqrt(cx.max_eig(200))


def test_box60_rigid_translations():
    mp = pdx.Model(W.box3d())
    cx = pdx.Context(mp)
    K, _ = cx.assemble(pdx.PDX_MODE_STATIC)
    ia, ja = mp.csr()
    A = sp.csr_matrix((K, ja, ia), shape=(mp.neq, mp.neq))
    eq = mp.nodes()["eq"]
    off, adj = mp.adjacency()
    node_of = np.zeros(mp.neq, np.int64)
    for d in range(3):
        free = eq[:, d] >= 0
        node_of[eq[free, d]] = np.nonzero(free)[0]
    # nodes with a prescribed neighbour (or prescribed themselves) in direction d
    rows_abs = np.abs(A).sum(axis=1).A1
    for d in range(3):
        presc = np.nonzero(eq[:, d] < 0)[0]
        near = np.zeros(mp.nn, bool)
        near[presc] = True
        for k in presc:
            near[adj[off[k]:off[k + 1]]] = True
        t = np.zeros(mp.neq)
        t[eq[eq[:, d] >= 0, d]] = 1.0
        r = A @ t
        keep = ~near[node_of]
        assert keep.sum() > 0.9 * mp.neq
        worst = np.max(np.abs(r[keep]) / rows_abs[keep])
        assert worst <= 1e-12, (d, worst)
        # the excluded rows do see the removed columns
        assert np.max(np.abs(r[~keep]) / rows_abs[~keep]) > 1e-6


@pytest.mark.parametrize("prec", [0, 1])
def test_box60_explicit_linear_in_the_load(prec):
    c1 = pdx.Context(pdx.Model(W.box3d(traction=1e6)))
    dt = _box_dt(c1)
    c2 = pdx.Context(pdx.Model(W.box3d(traction=2e6)))
    for c in (c1, c2):
        c.set_dt(dt)
        c.explicit_set_form(prec, 21)  # one form (one order of the row sums) for both runs
    u1, up1 = c1.explicit_run(30, precision=prec)
    u2, up2 = c2.explicit_run(30, precision=prec)
    assert np.abs(u1).max() > 0
    if prec == 0:
        assert np.array_equal(u2, 2.0 * u1) and np.array_equal(up2, 2.0 * up1)
    else:
        assert rel(u2, 2.0 * u1) <= 1e-12 and rel(up2, 2.0 * up1) <= 1e-12


def test_box60_fp32_force_within_1e6():
    mp = pdx.Model(W.box3d())
    c64 = pdx.Context(mp)
    dt = _box_dt(c64)
    c64.set_dt(dt)
    c32 = pdx.Context(mp)
    c32.set_dt(dt)
    u64, _ = c64.explicit_run(200, precision=0)
    u32, _ = c32.explicit_run(200, precision=1)
    assert np.abs(u64).max() > 0
    assert rel(u32, u64) <= 1e-6


def test_penny100_newmark_fp32_refined_matches_fp64():
    m0 = W.penny3d()
    c0 = pdx.Context(pdx.Model(m0))
    dt = 10.0 * 2.0 / np.sqrt(c0.max_eig(200))
    del c0
    mp = pdx.Model(W.penny3d(dt=dt, ramp_steps=50))
    runs = {}
    for name, o in (("fp64", pdx.solve_opts(precision=0, precond=pdx.PDX_PREC_BLOCK_JACOBI, tol=1e-8,
                                            max_iter=100000)),
                    ("fp32_ir", pdx.solve_opts(precision=1, precond=pdx.PDX_PREC_TWO_LEVEL, tol=1e-8,
                                               max_iter=100000, coarse_agg=10))):
        cx = pdx.Context(mp)
        u, v, a, st = cx.newmark_run(3, o)
        assert st["converged"] and st["true_relres"] <= 1e-8, (name, st)  # the last step's solve
        runs[name] = u
        del cx
    assert np.abs(runs["fp64"]).max() > 0
    assert rel(runs["fp32_ir"], runs["fp64"]) <= 1e-7


def test_penny100_gpu_setup_matches_host():
    mesh = W.penny3d()
    h = pdx.Model(mesh)
    pdx.set_setup_device(0)
    try:
        g = pdx.Model(mesh)
    finally:
        pdx.set_setup_device(-1)
    fh, fg = h.families(), g.families()
    for key in ("node", "off", "mem", "bond"):
        assert np.array_equal(fh[key], fg[key]), key
    oh, ah = h.adjacency()
    og, ag = g.adjacency()
    assert np.array_equal(oh, og) and np.array_equal(ah, ag)
    assert fh["mem"].size > 3_000_000
