"""Nodal stresses (calGlobalNodeStresses) on the GPU against the oracle.

For a seeded random DOF field, every node's stress agrees with the oracle's to
1e-12 of the largest stress component (PD nodes through the cached bond strain
weights, FE nodes through Gauss-point stresses and the extrapolation matrix L,
the averaging and the plane-strain sigma_zz); the explicit-state path gives the
same as passing the state's DOF values.
"""
import numpy as np
import pytest

from helpers import load_oracle, load_pdx, tet_box_mesh, tri_tet_mesh, workloads

pytestmark = pytest.mark.gpu

O = load_oracle()
W = workloads()
pdx = load_pdx()


def _plane_strain():
    m = W.plate2d(nx=41, ny=21, dx=1e-3, pd_x=(0.015, 0.025))
    m["proType"] = 2
    return m


CASES = {
    "plate2d": lambda: W.plate2d(nx=61, ny=31, dx=1e-3, pd_x=(0.02, 0.04)),
    "plate2d_strain": _plane_strain,
    "crack2d": lambda: W.crack2d(seed=5, nx=121, ny=61, dx=1e-3),
    "box3d": lambda: W.box3d(n=14, pd_block=8),
    "penny3d": lambda: W.penny3d(n=20, pd_block=10),
    "tri_tet": tri_tet_mesh,
    "tet_box": tet_box_mesh,
}


@pytest.mark.parametrize("name", list(CASES))
def test_node_stresses_match_oracle(name):
    mesh = CASES[name]()
    mo = O.Model(mesh)
    cx = pdx.Context(pdx.Model(mesh))
    rng = np.random.default_rng(7)
    u = rng.normal(size=mo.nn * mesh["dim"]) * 1e-6
    s0 = mo.node_stresses(u)
    s1 = cx.node_stresses(u)
    assert np.abs(s1 - s0).max() <= 1e-12 * np.abs(s0).max(), np.abs(s1 - s0).max() / np.abs(s0).max()


def test_node_stresses_of_explicit_state():
    mesh = W.box3d(n=12, pd_block=6)
    mesh["nbc"][0]["val"] = 1e6
    mesh["dt"] = 2e-8
    mp = pdx.Model(mesh)
    cx = pdx.Context(mp)
    cx.explicit_run(5)
    s_state = cx.node_stresses()
    s_vals = cx.node_stresses(mp_values(cx, mp))
    assert np.array_equal(s_state, s_vals)


def mp_values(cx, mp):
    """DOF values [nn*dim] of the explicit state u_{n+1} (prescribed ones included)."""
    u, _ = cx.explicit_run(0)
    vals = mp.values().copy()  # prescribed values at set-up
    eq = mp.nodes()["eq"][:, :mp.dim].ravel()
    free = eq >= 0
    vals[free] = u[eq[free]]
    return vals
