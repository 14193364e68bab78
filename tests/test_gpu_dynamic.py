"""GPU parity of the dynamic drivers against the oracle, through the C-ABI: the consistent-mass
explicit path (LUMPMASS OFF, the reference's default), bond failure inside the explicit and
Newmark loops (FC 2), and Newmark in 3D with a ramped traction.

Tolerances:
  consistent mass on its pattern      <= 1e-14 relative (values), pattern bit-exact
  explicit, consistent mass, FP64     ||u_gpu - u_oracle|| / ||u_oracle|| <= 1e-10 after 60 steps
                                      (the mass solve runs to a 1e-13 relative residual)
  explicit, consistent mass, FP32 force <= 1e-6
  dynamic FC 2                        bonds kept / broken per step and the final bond states
                                      exact; u <= 1e-8
  Newmark 3D (penny crack, ramped traction) FP64 and FP32-refined: u <= 1e-8
"""
import numpy as np
import pytest

from helpers import load_oracle, load_pdx, rel, workloads

pytestmark = pytest.mark.gpu

O = load_oracle()
W = workloads()
pdx = load_pdx()


def _consistent_case():
    m = W.plate2d(nx=41, ny=21, dx=1e-3, pd_x=(0.015, 0.025))
    m["ebc"][2].update(val=0.0, vel=0.05, vary=True)
    top = np.arange(20 * 41 + 1, 21 * 41 + 1, dtype=np.int32)
    m["nbc"] = [dict(con=np.stack([top[:-1], top[1:]], axis=1), val=1e6, vel=2e11, vary=True)]
    m["lumped_mass"] = False
    m["dt"] = 5e-8
    return m


@pytest.mark.parametrize("name", ["plate2d", "box3d"])
def test_consistent_mass_matches_oracle(name):
    mesh = _consistent_case() if name == "plate2d" else dict(W.box3d(n=10, pd_block=4), lumped_mass=False)
    mo = O.Model(mesh)
    ia0, ja0, v0 = mo.mass_csr()
    mp = pdx.Model(mesh)
    ia1, ja1 = mp.mass_csr()
    assert np.array_equal(ia0, ia1) and np.array_equal(ja0, ja1)
    v1 = pdx.Context(mp).mass_csr()
    assert rel(v1, v0) <= 1e-14


@pytest.mark.parametrize("name", ["plate2d", "box3d"])
@pytest.mark.parametrize("prec,tol", [(0, 1e-10), (1, 1e-6)])
def test_explicit_consistent_mass_trajectory(name, prec, tol):
    """pdfemDynamicSolver_CSRformat with LUMPMASS OFF: M a = F every step (the oracle factorises
    M with SuperLU, the GPU runs PCG to 1e-13); 60 steps with ramped EBC and traction."""
    if name == "plate2d":
        mesh = _consistent_case()
    else:
        mesh = W.box3d(n=12, pd_block=6)
        mesh["ebc"].append(dict(nid=np.array([12 * 12 * 11 + 1], np.int32), dof=0, val=0.0, vel=0.05, vary=True))
        mesh["nbc"][0].update(vel=1e9, vary=True)
        mesh.update(lumped_mass=False, dt=5e-8)
    u0, up0, _ = O.Model(mesh).dynamic_run(60)
    cx = pdx.Context(pdx.Model(mesh))
    u1, up1 = cx.explicit_run(60, precision=prec)
    assert np.abs(u0).max() > 0
    assert rel(u1, u0) <= tol and rel(up1, up0) <= tol, (rel(u1, u0), rel(up1, up0))


def _fc2_case(lumped):
    """Edge-cracked 41 x 21 plate pulled apart at the top and bottom (the shape of the reference's
    plateCrack dynamic examples: VEBC -1 / +1 on the two loaded edges, FC 2 0).  Every node is
    jittered by up to 0.15 dx: on an exact grid an FC 2 crack segment (through a bond's midpoint,
    normal to the principal direction) runs almost along a row of nodes and bonds, where the
    reference's segment test decides on the last bits of the stresses, which the GPU and the
    oracle share only to 1e-12."""
    m = W.crack2d(seed=1, nx=41, ny=21, dx=1e-3)
    x = m["x"].copy()
    x[:, :2] += np.random.default_rng(11).uniform(-0.15e-3, 0.15e-3, (len(x), 2))
    m["x"] = x
    m["ebc"][0].update(val=0.0, vel=-0.5, vary=True)
    m["ebc"][2].update(val=0.0, vel=0.5, vary=True)
    m.update(lumped_mass=lumped, fc=2, topk=0, dt=5e-8)
    return m


@pytest.mark.parametrize("lumped", [True, False])
def test_explicit_fc2_matches_oracle(lumped):
    """Bond failure after every stored explicit step: the bonds selected and broken per step and
    the final bond states equal the oracle's; u within 1e-8.  The run covers the first breaks up
    to the onset of the cascade, where hundreds of bonds sit near the strength every step and a
    stress difference in the last bits (the GPU and the oracle agree to 1e-12) can decide one of
    them either way."""
    m = _fc2_case(lumped)
    N = 32
    _, _, probe = O.Model(m).dynamic_run(N, sigma_ult=1e30)
    s = probe[:, 2]
    m["sigma_ult"] = float(0.5 * (s[N // 2 - 1] + s[N // 2]))  # bonds start to break half-way
    mo = O.Model(m)
    u0, up0, info0 = mo.dynamic_run(N)
    assert info0[:, 1].sum() > 0
    cx = pdx.Context(pdx.Model(m))
    u1, up1 = cx.explicit_run(N)
    info1 = cx.step_log()
    assert info1.shape == info0.shape
    assert np.array_equal(info1[:, :2], info0[:, :2]), np.nonzero(info1[:, :2] != info0[:, :2])
    assert rel(info1[:, 2], info0[:, 2]) <= 1e-8
    assert np.array_equal(cx.bonds(), mo.families()["bond"])
    assert rel(u1, u0) <= 1e-8 and rel(up1, up0) <= 1e-8


def test_newmark_fc2_matches_oracle():
    """Bond failure after every Newmark step (at the advanced BC values): per-step bonds and the
    final bond states exact, u within 1e-8."""
    m = _fc2_case(False)
    m["dt"] = 2e-7
    N = 12
    probe = []
    O.newmark_run(O.Model(m), N, fc=2, topk=0, sigma_ult=1e30, info=probe)
    s = np.array(probe)[:, 2]
    sig = float(0.5 * (s[N // 2 - 1] + s[N // 2]))
    m["sigma_ult"] = sig
    mo = O.Model(m)
    info0 = []
    u0, v0, a0 = O.newmark_run(mo, N, fc=2, topk=0, sigma_ult=sig, info=info0)
    info0 = np.array(info0)
    assert info0[:, 1].sum() > 0
    cx = pdx.Context(pdx.Model(m))
    u1, v1, a1, st = cx.newmark_run(N, pdx.solve_opts(precision=0, precond=2, tol=1e-12, max_iter=100000))
    info1 = cx.step_log()
    assert np.array_equal(info1[:, :2], info0[:, :2])
    assert np.array_equal(cx.bonds(), mo.families()["bond"])
    assert rel(u1, u0) <= 1e-8


@pytest.mark.parametrize("precision", [0, 1])
def test_newmark_penny3d_matches_oracle(precision):
    """3D Newmark on the penny-crack box (20^3 nodes, 10^3-node PD region, pre-broken bonds across
    the crack) with the traction ramped from zero: FP64 and FP32-refined (FP32 inner BiCGStab,
    FP64 refinement to the same 1e-12 true residual) against the oracle's direct solves."""
    mesh = W.penny3d(n=20, pd_block=10, ramp_steps=6, dt=1e-6)
    N = 6
    mo = O.Model(mesh)
    u0, v0, a0 = O.newmark_run(mo, N)
    cx = pdx.Context(pdx.Model(mesh))
    u1, v1, a1, st = cx.newmark_run(N, pdx.solve_opts(precision=precision, precond=2, tol=1e-12, max_iter=200000,
                                                      inner_tol=1e-5))
    assert np.abs(u0).max() > 0
    assert rel(u1, u0) <= 1e-8 and rel(v1, v0) <= 1e-7, (rel(u1, u0), rel(v1, v0))


def test_newmark_ramped_ebc_and_nbc_matches_oracle():
    """2D Newmark with a ramped EBC and a ramped Line2 traction (the predictor's prescribed
    increment and the lagged traction, pinned in the oracle by a one-DOF recurrence)."""
    mesh = _consistent_case()
    mesh["dt"] = 1e-6
    mo = O.Model(mesh)
    u0, v0, a0 = O.newmark_run(mo, 5)
    cx = pdx.Context(pdx.Model(mesh))
    u1, v1, a1, st = cx.newmark_run(5, pdx.solve_opts(precision=0, precond=2, tol=1e-12, max_iter=100000))
    assert rel(u1, u0) <= 1e-8 and rel(v1, v0) <= 1e-7


def test_dynamic_drivers_reject_unsupported_failure_criteria():
    """FC 1 (J-integral crack growth) is not built: both dynamic drivers refuse it instead of
    running without failure."""
    m = _fc2_case(True)
    m["fc"] = 1
    cx = pdx.Context(pdx.Model(m))
    with pytest.raises(pdx.PdxError, match="FC 1"):
        cx.explicit_run(2)
    with pytest.raises(pdx.PdxError, match="FC 1"):
        cx.newmark_run(2)


def _run_fc2(driver, m, incremental, n):
    cx = pdx.Context(pdx.Model(m))
    cx.set_incremental_bonds(incremental)
    if driver == "explicit":
        # one form (one order of the row sums) for both runs: the tuner picks by timing, and forms
        # with different lane counts round the row sums differently
        cx.explicit_set_form(0, 0)
        u, up = cx.explicit_run(n)
        out, info = (u, up), cx.step_log()
    elif driver == "newmark":
        u, v, a, st = cx.newmark_run(n, pdx.solve_opts(precision=0, precond=2, tol=1e-12, max_iter=100000))
        out, info = (u, v, a), cx.step_log()
    else:
        u, info, st = cx.quasi_static_run(n, pdx.solve_opts(precision=0, precond=2, tol=1e-12, max_iter=100000),
                                          fc=2, topk=2, sigma_ult=3.56e7)
        out = (u,)
    return out, np.asarray(info), cx.bonds()


@pytest.mark.parametrize("driver", ["explicit", "newmark", "quasi_static"])
def test_fc2_incremental_update_equals_whole_assembly(driver):
    """Bond failure applied to K by recomputing only the changed families' rows (the default)
    gives, bit for bit, the run that assembles the whole K again after every change: the DOF
    values, the bonds selected and broken per step and the final bond states, into the cascade
    where bonds break in most steps."""
    if driver == "quasi_static":
        m = W.crack2d(seed=1, nx=41, ny=21, dx=1e-3)
        m["ebc"][2] = dict(nid=m["ebc"][2]["nid"], dof=1, val=0.0, vary=True, vel=2e-6)
        n = 10
    else:
        m = _fc2_case(driver == "explicit")
        n = 40 if driver == "explicit" else 12
        if driver == "newmark":
            m["dt"] = 2e-7
        m["sigma_ult"] = 1e30
        _, probe, _ = _run_fc2(driver, m, True, n)
        m["sigma_ult"] = float(probe[n // 4, 2])  # bonds break from a quarter of the run on
    a_out, a_info, a_bond = _run_fc2(driver, m, True, n)
    b_out, b_info, b_bond = _run_fc2(driver, m, False, n)
    assert a_info[:, 1].sum() > 0 and (a_info[:, 1] > 0).sum() >= 2
    for x, y in zip(a_out, b_out):
        assert np.array_equal(x, y)
    assert np.array_equal(a_info, b_info) and np.array_equal(a_bond, b_bond)
