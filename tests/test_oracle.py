"""Pins of the CPU oracle (oracle/pdlsm_oracle.c) -- what fixes each function.

  a1-a5 inflFunc / shapTens / vec_gd / matG / matH  -> Navier operator of a
        quadratic field on jittered 2D and 3D families (closed form), and the
        reference's golden static result examples/box (end to end)
  a6    matC                       -> exact strain of a quadratic field
  a7    PD assembly (-H cd_v, free/fixed split)  -> box golden; pattern counts
  a8    PD-boundary terms 2D/3D    -> traction sigma.n closed form on the PD
        region boundary; box golden (3D); a planted drop is detected
  a9    CSR pattern                -> plate2d N_eq / nnz measured with the
        reference binary (40,299 / 1,043,589); box N_eq 102,500
  a12   element stiffness          -> textbook quad4 K11, strain energy of a
        constant-strain field (quad4, tri3, hex8, tet4), rigid-body null space
  a14   natural-BC force           -> total force = t * area
  a15   lumped / consistent mass   -> totals = rho * V
  a17   explicit driver            -> closed form of the central-difference
        recurrence of a one-DOF model, with and without a ramped EBC (pins the
        start-up step and the advance-before-force order)
  a19   Newmark driver             -> energy conservation of the average
        acceleration rule on a one-DOF model from rest (pins beta, gamma, the
        predictor and the update)
  S9    nodal stresses             -> sigma = D eps of a uniform strain at every
        node (PD, FE, averaging, plane-strain sigma_zz); corner values of a
        linear stress field built from the reference's Gauss-point order and L
"""
import os

import numpy as np
import pytest

from helpers import box_fixture, load_oracle, one_quad, tet_box_mesh, tri_tet_mesh, workloads

O = load_oracle()
W = workloads()


# ---------------------------------------------------------------------------
# golden: the reference's examples/box static result
# ---------------------------------------------------------------------------

def _box_correction(skip=0):
    import scipy.sparse.linalg as spla
    mesh, U = box_fixture()
    M = O.Model(mesh)
    O.set_skip(skip)
    try:
        K, F = M.assemble(1)
    finally:
        O.set_skip(0)
    eq = M.nodes()["eq"]
    u = np.zeros(M.neq)
    u[eq[eq >= 0]] = U.astype(np.float64)[eq >= 0]
    A = M.sparse(K)
    r = F - A @ u
    d = A.diagonal()
    P = spla.LinearOperator(A.shape, matvec=lambda v: v / d)
    delta, _ = spla.bicgstab(A, r, M=P, rtol=1e-3, maxiter=400)
    return M, np.linalg.norm(delta) / np.linalg.norm(u), np.linalg.norm(r) / np.linalg.norm(F)


@pytest.mark.slow
def test_box_golden_static_solution():
    """The oracle's exact static solution equals the reference's committed
    boxVpd4_static.vtk U to float32 rounding (correction ||delta||/||u||)."""
    M, corr, res = _box_correction()
    assert M.neq == 102500
    assert corr < 2e-7, corr
    assert res < 1e-4, res


@pytest.mark.slow
def test_box_golden_detects_dropped_boundary_terms():
    """Planted mistake in the oracle: without the PD-boundary terms the VTK
    solution is far from the oracle's."""
    _, corr, _ = _box_correction(skip=1)
    assert corr > 1e-3, corr


def test_box_golden_node_stresses():
    """calGlobalNodeStresses against the reference's own output: the oracle's
    nodal stresses of the boxVpd4_static.vtk displacements equal that file's
    stress tensor to the float32 rounding of its U (SURVEY a6 / §8f stress row)."""
    mesh, U = box_fixture()
    ref = np.load(os.path.join(os.path.dirname(__file__), "golden", "box", "box_vtk_fields.npz"))["sigma"]
    ref = ref.astype(np.float64)
    M = O.Model(mesh)
    S = M.node_stresses(U.astype(np.float64).reshape(-1))
    assert np.linalg.norm(S - ref) <= 5e-6 * np.linalg.norm(ref)
    assert np.abs(S - ref).max() <= 1e-5 * np.abs(ref).max()
    # a planted mistake (strain from u scaled by 1 + 1e-4) is far outside the bound
    S2 = M.node_stresses(U.astype(np.float64).reshape(-1) * (1 + 1e-4))
    assert np.linalg.norm(S2 - ref) > 5e-5 * np.linalg.norm(ref)


def test_plate2d_pattern_counts_match_reference_binary():
    M = O.Model(W.plate2d())
    assert (M.neq, M.nnz, M.nfam) == (40299, 1043589, 4141)
    ia, ja = M.csr()
    for r in range(0, M.neq, 997):
        row = ja[ia[r]: ia[r + 1]]
        assert np.all(np.diff(row) > 0)


# ---------------------------------------------------------------------------
# PDLSM family: closed forms
# ---------------------------------------------------------------------------

def _quad_field(dim, seed):
    rng = np.random.default_rng(seed)
    c0 = rng.normal(size=dim)
    B = rng.normal(size=(dim, dim))
    Q = rng.normal(size=(dim, dim, dim))
    Q = 0.5 * (Q + Q.transpose(0, 2, 1))  # u_i = c_i + B_ij x_j + Q_ijk x_j x_k

    def u(x):
        x = np.asarray(x)[:, :dim]
        return c0 + x @ B.T + np.einsum("ijk,nj,nk->ni", Q, x, x)

    def grad(x):
        x = np.asarray(x)[:dim]
        return B + 2 * np.einsum("ijk,k->ij", Q, x)

    def navier(lam, mu):
        # L(u)_i = (lam+mu) d_i(div u) + mu lap u_i ; second derivatives d_jk u_i = 2 Q_ijk
        H = 2 * Q
        div_grad = np.einsum("jji->i", H)
        lap = np.einsum("ijj->i", H)
        return (lam + mu) * div_grad + mu * lap

    return u, grad, navier


def _interior_families(M, dim, n):
    fam = M.families()
    x = M.mesh["x"]
    lo, hi = x.min(0), x.max(0)
    out = []
    for f in range(M.nfam):
        k = fam["node"][f]
        d = fam["delta"][f]
        if np.all(x[k, :dim] - 1.2 * d > lo[:dim]) and np.all(x[k, :dim] + 1.2 * d < hi[:dim]):
            out.append(f)
        if len(out) >= n:
            break
    return fam, out


@pytest.mark.parametrize("dim", [2, 3])
def test_family_H_reproduces_navier_operator(dim):
    """sum_m G_m V_m (u_m - u_k) = (lam+mu) grad div u + mu lap u for any
    quadratic u: pins inflFunc, the shape tensor with its 1/2 factors, the per-bond
    solve and matG/matH."""
    if dim == 2:
        mesh = W.plate2d(nx=41, ny=41, pd_x=(0.0, 1.0))
        mesh["x"][:, :2] += np.random.default_rng(1).uniform(-0.15, 0.15, (len(mesh["x"]), 2)) * 5e-4
    else:
        mesh = W.box3d(n=14, pd_block=14, jitter=0.2)
    M = O.Model(mesh)
    g = M.geom()
    fam, fams = _interior_families(M, dim, 6)
    assert fams
    u, _, navier = _quad_field(dim, 7)
    L = navier(g["lam"], g["mu"])
    x = mesh["x"]
    for f in fams:
        mem = fam["mem"][fam["off"][f]: fam["off"][f + 1]] - 1
        H = M.family_H(f)
        val = H @ u(x[mem]).ravel()
        assert np.allclose(val, L, rtol=1e-6, atol=1e-6 * np.abs(L).max()), (val, L)


@pytest.mark.parametrize("dim", [2, 3])
def test_family_C_gives_exact_strain(dim):
    if dim == 2:
        mesh = W.plate2d(nx=31, ny=31, pd_x=(0.0, 1.0))
        mesh["x"][:, :2] += np.random.default_rng(2).uniform(-0.15, 0.15, (len(mesh["x"]), 2)) * 5e-4
    else:
        mesh = W.box3d(n=12, pd_block=12, jitter=0.2, seed=3)
    M = O.Model(mesh)
    fam, fams = _interior_families(M, dim, 4)
    u, grad, _ = _quad_field(dim, 11)
    x = mesh["x"]
    for f in fams:
        mem = fam["mem"][fam["off"][f]: fam["off"][f + 1]] - 1
        k = mem[0]
        eps = M.family_C(f) @ u(x[mem]).ravel()
        G = grad(x[k])
        if dim == 2:
            ex = np.array([G[0, 0], G[1, 1], G[0, 1] + G[1, 0]])
        else:
            ex = np.array([G[0, 0], G[1, 1], G[2, 2], G[0, 1] + G[1, 0], G[1, 2] + G[2, 1], G[0, 2] + G[2, 0]])
        assert np.allclose(eps, ex, rtol=1e-7, atol=1e-7 * np.abs(ex).max())


def _linear_u(M, dim, B):
    x = M.mesh["x"]
    return (x[:, :dim] @ B.T).ravel()


def test_pdbe_2d_traction_closed_form():
    """PD-boundary terms alone, linear u (constant stress sigma): at a node on
    the straight left edge of the PD strip the rows receive -dx * sigma.n with
    the outward normal -x, i.e. F = dx * (sxx, sxy)."""
    dx = 5e-4
    mesh = W.plate2d(nx=41, ny=21, dx=dx, pd_x=(0.005, 0.015))
    mesh["ebc"] = []
    M = O.Model(mesh)
    B = np.array([[1e-3, 2e-4], [-5e-4, 3e-4]])
    M.store_u(_linear_u(M, 2, B))
    O.set_skip(2 | 4)
    try:
        _, F = M.assemble(2, with_ext=False)
    finally:
        O.set_skip(0)
    g = M.geom()
    lam, mu = g["lam"], g["mu"]
    eps = np.array([B[0, 0], B[1, 1], B[0, 1] + B[1, 0]])
    E, nu = mesh["E"], mesh["nu"]
    D = E / (1 - nu ** 2) * np.array([[1, nu, 0], [nu, 1, 0], [0, 0, (1 - nu) / 2]])
    s = D @ eps
    eq = M.nodes()["eq"]
    i0 = 10  # node column x = 10 dx = 0.005: left edge of the PD strip
    for j in range(3, 18, 4):
        k = j * 41 + i0
        got = F[eq[k, :2]]
        assert np.allclose(got, dx * np.array([s[0], s[2]]), rtol=1e-9), (got, dx * s)


def test_pdbe_3d_traction_closed_form():
    dx = 1e-3
    mesh = W.box3d(n=14, dx=dx, jitter=0.0, pd_block=8)
    mesh["ebc"] = []
    M = O.Model(mesh)
    B = np.array([[1e-3, 2e-4, -1e-4], [-5e-4, 3e-4, 2e-4], [1e-4, -2e-4, 4e-4]])
    M.store_u(_linear_u(M, 3, B))
    O.set_skip(2 | 4)
    try:
        _, F = M.assemble(2, with_ext=False)
    finally:
        O.set_skip(0)
    E, nu = mesh["E"], mesh["nu"]
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    eps = 0.5 * (B + B.T)
    sig = lam * np.trace(eps) * np.eye(3) + 2 * mu * eps
    eq = M.nodes()["eq"]
    lo = (14 - 8) // 2  # PD node block [3, 10]; face x = 3 dx has outward normal -x
    k = 6 * 14 * 14 + 6 * 14 + lo
    got = F[eq[k]]
    # F = -sum_e w A_e sigma.n over the four faces around the node, n = -x, sum w A = dx^2
    assert np.allclose(got, dx * dx * sig[:, 0], rtol=1e-9), (got, dx * dx * sig[:, 0])


# ---------------------------------------------------------------------------
# elements
# ---------------------------------------------------------------------------

def _single(dim, nen, x, E=2.0e3, nu=0.3):
    econ = np.zeros((1, 8), np.int32)
    econ[0, :nen] = np.arange(1, nen + 1)
    return O.Model(dict(dim=dim, proType=1 if dim == 2 else -1, E=E, nu=nu, rho=3.0, x=x, econ=econ,
                        nen=np.array([nen], np.int32), algo=np.array([2], np.int32), ebc=[], nbc=[],
                        cracks=np.zeros((0, 3, 3)), fac=3.0, nlf=0.33))


def _D(dim, E, nu):
    if dim == 2:
        return E / (1 - nu ** 2) * np.array([[1, nu, 0], [nu, 1, 0], [0, 0, (1 - nu) / 2]])
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    D = np.zeros((6, 6))
    D[:3, :3] = lam
    D[np.arange(3), np.arange(3)] += 2 * mu
    D[3:, 3:] = mu * np.eye(3)
    return D


def test_quad4_textbook_diagonal():
    x = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], float)
    E, nu = 2.0e3, 0.3
    K = _single(2, 4, x, E, nu).element_K(0)
    assert np.isclose(K[0, 0], E / (1 - nu ** 2) * (0.5 - nu / 6), rtol=1e-13)
    assert np.allclose(K, K.T, rtol=1e-13, atol=1e-12 * abs(K).max())


@pytest.mark.parametrize("kind", ["quad4", "tri3", "hex8", "tet4"])
def test_element_constant_strain_energy_and_rigid_modes(kind):
    rng = np.random.default_rng(5)
    if kind == "quad4":
        dim, x = 2, np.array([[0, 0, 0], [1.1, 0.1, 0], [1.0, 0.9, 0], [-0.1, 1.2, 0]], float)
    elif kind == "tri3":
        dim, x = 2, np.array([[0, 0, 0], [1.2, 0.1, 0], [0.3, 0.9, 0]], float)
    elif kind == "hex8":
        dim = 3
        x = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)
        x += rng.uniform(-0.1, 0.1, x.shape)
    else:
        dim, x = 3, np.array([[0, 0, 0], [1, 0.1, 0], [0.2, 1, 0.1], [0.1, 0.2, 1.1]], float)
    nen = len(x)
    E, nu = 2.0e3, 0.3
    M = _single(dim, nen, x, E, nu)
    K = M.element_K(0)
    V = M.nodes()["ev"][0]
    B = rng.normal(size=(dim, dim))
    u = (x[:, :dim] @ B.T).ravel()
    eps = 0.5 * (B + B.T)
    ev = np.array([eps[0, 0], eps[1, 1], 2 * eps[0, 1]]) if dim == 2 else \
        np.array([eps[0, 0], eps[1, 1], eps[2, 2], 2 * eps[0, 1], 2 * eps[1, 2], 2 * eps[0, 2]])
    assert np.isclose(u @ K @ u, ev @ _D(dim, E, nu) @ ev * V, rtol=1e-11)
    W_ = np.zeros((dim, dim))
    W_[0, 1], W_[1, 0] = 1.0, -1.0  # rigid rotation + translation
    r = (x[:, :dim] @ W_.T + 0.7).ravel()
    assert np.abs(K @ r).max() < 1e-10 * np.abs(K).max()
    Me = M.element_M(0)
    assert np.isclose(Me.sum(), dim * 3.0 * V, rtol=1e-12)


def test_natural_bc_total_force_and_lumped_mass():
    mesh = W.box3d(n=8, dx=1e-3, jitter=0.2, pd_block=4, traction=2.5e6)
    mesh["ebc"] = []
    M = O.Model(mesh)
    M.store_u(np.zeros(M.nn * 3))
    O.set_skip(1 | 2 | 4)
    try:
        _, F = M.assemble(2, with_ext=True)
    finally:
        O.set_skip(0)
    L = 7e-3
    Fz = F.reshape(-1, 3)
    assert np.isclose(Fz[:, 2].sum(), 2.5e6 * L * L, rtol=1e-12)
    assert abs(Fz[:, 0].sum()) < 1e-9 * 2.5e6 * L * L
    m = M.lumped_mass()
    assert np.isclose(m.sum(), 3 * mesh["rho"] * M.nodes()["ev"].sum(), rtol=1e-12)


# ---------------------------------------------------------------------------
# drivers
# ---------------------------------------------------------------------------

def _one_dof_terms(mesh):
    M = O.Model(mesh)
    K, F = M.assemble(1)
    m = M.lumped_mass()
    return M, K, F, m


def test_explicit_one_dof_closed_form():
    """u_n = (f/k)(1 - cos n theta), 2 - 2 cos theta = dt^2 k/m, for the
    reference's start-up step u1 = u0 + dt^2 a0 / 2 and the recurrence."""
    mesh = one_quad()
    M, K, F, m = _one_dof_terms(mesh)
    assert M.neq == 1
    k, f, mm, dt = K[0], F[0], m[0], mesh["dt"]
    th = np.arccos(1 - dt * dt * k / (2 * mm))
    for n in (0, 1, 5, 40):
        u, _ = O.Model(mesh).explicit_run(n)
        want = f / k * (1 - np.cos((n + 1) * th))
        assert np.isclose(u[0], want, rtol=1e-9, atol=1e-15), (n, u[0], want)


def test_explicit_ramped_bc_order():
    """Ramped EBC on a coupled DOF: the force of loop step n uses the BC value
    advanced n+1 times (advance before force).  Exact discrete solution of
    u_{n+1} - 2u_n + u_{n-1} = dt^2/m (f - k u_n - c v_{n+1})."""
    mesh = one_quad(vebc=True)
    M0 = O.Model(one_quad(vebc=False))
    K, F = M0.assemble(1)
    k, f = K[0], F[0]
    mm = M0.lumped_mass()[0]
    # coupling c = K(u3x, u2x): force per unit prescribed value
    Mq = O.Model(dict(one_quad(), ebc=one_quad()["ebc"][:2]))
    Kq, _ = Mq.assemble(1)
    Sq = Mq.sparse(Kq).toarray()
    eqq = Mq.nodes()["eq"]
    c = Sq[eqq[2, 0], eqq[1, 0]]
    dt, vel = mesh["dt"], 1.0
    v = lambda j: j * dt * vel
    # reference recurrence, computed independently of the oracle driver
    u_prev, u = 0.0, 0.0
    a0 = (f - c * v(1)) / mm
    u_next = u + 0.5 * dt * dt * a0
    hist = [u_next]
    for n in range(1, 30):
        u_prev, u = u, u_next
        u_next = 2 * u + dt * dt * (f - k * u - c * v(n + 1)) / mm - u_prev
        hist.append(u_next)
    got, _ = O.Model(mesh).explicit_run(29)
    assert np.isclose(got[0], hist[-1], rtol=1e-9)
    # the off-by-one order (force with v_n) gives a measurably different answer
    u_prev, u = 0.0, 0.0
    u_next = 0.5 * dt * dt * (f - c * v(0)) / mm
    for n in range(1, 30):
        u_prev, u = u, u_next
        u_next = 2 * u + dt * dt * (f - k * u - c * v(n)) / mm - u_prev
    assert abs(got[0] - u_next) > 1e-3 * abs(got[0])


def test_newmark_average_acceleration_conserves_energy():
    """From rest (a0 = 0 as in the reference), after the first step
    E = m v^2/2 + k (u - f/k)^2/2 is constant for beta = 1/4, gamma = 1/2."""
    mesh = one_quad(dt=2e-4)
    M = O.Model(mesh)
    K, F = M.assemble(1)
    k, f, mm = K[0], F[0], M.mass_on_K()[0]
    Es = []
    for n in (1, 2, 7, 20):
        u, v, a = O.newmark_run(O.Model(mesh), n)
        Es.append(0.5 * mm * v[0] ** 2 + 0.5 * k * (u[0] - f / k) ** 2)
        assert np.isclose(mm * a[0] + k * u[0], f, rtol=1e-9)
    assert np.allclose(Es, Es[0], rtol=1e-9)


# ---------------------------------------------------------------------------
# nodal stresses (calGlobalNodeStresses)
# ---------------------------------------------------------------------------

def _voigt_sigma(dim, E, nu, G, plane_strain=False):
    """sigma of the linear field u = G x in the oracle's 6-slot layout."""
    eps = 0.5 * (G + G.T)
    if dim == 2:
        if plane_strain:
            lam = E * nu / ((1 + nu) * (1 - 2 * nu)); mu = E / (2 * (1 + nu))
            D = np.array([[lam + 2 * mu, lam, 0], [lam, lam + 2 * mu, 0], [0, 0, mu]])
        else:
            D = _D(2, E, nu)
        s = D @ np.array([eps[0, 0], eps[1, 1], 2 * eps[0, 1]])
        out = np.array([s[0], s[1], 0.0, s[2], 0.0, 0.0])
        if plane_strain:
            out[2] = nu * (s[0] + s[1])
        return out
    return _D(3, E, nu) @ np.array([eps[0, 0], eps[1, 1], eps[2, 2], 2 * eps[0, 1], 2 * eps[1, 2], 2 * eps[0, 2]])


@pytest.mark.parametrize("case", ["plate2d_stress", "plate2d_strain", "box3d", "tri_tet", "tet_box"])
def test_node_stresses_of_uniform_strain(case):
    """A linear displacement gives sigma = D eps at every node: PD nodes through
    matC (polynomial reproduction), FE nodes through B at the Gauss points and
    the extrapolation L (unit row sums), all through the averaging."""
    rng = np.random.default_rng(3)
    if case.startswith("plate2d"):
        m = W.plate2d(nx=41, ny=21, dx=1e-3, pd_x=(0.015, 0.025))
        if case == "plate2d_strain":
            m["proType"] = 2
        dim = 2
    elif case == "tri_tet":
        m, dim = tri_tet_mesh(), 2
    elif case == "tet_box":
        m, dim = tet_box_mesh(), 3
    else:
        m = W.box3d(n=12, pd_block=6)
        dim = 3
    M = O.Model(m)
    G = rng.normal(size=(dim, dim)) * 1e-4
    x = np.asarray(m["x"])[:, :dim]
    u = (x @ G.T).ravel()
    sig = M.node_stresses(u)
    want = _voigt_sigma(dim, m["E"], m["nu"], G, plane_strain=(case == "plate2d_strain"))
    assert np.abs(sig - want[None, :]).max() <= 1e-9 * np.abs(want).max()


@pytest.mark.parametrize("kind", ["quad4", "hex8"])
def test_node_stresses_extrapolate_linear_stress(kind):
    """On one square/cube element, u_x = x*y gives stresses linear in x and y.
    Expected corner values are built here from the reference's written
    algorithm, independently of the oracle's code: Gauss-point stresses in the
    reference's point order, then its matrix L (pdsolve.cpp:9176-9196, :9254-9277).
    hex8: points and L follow the node order, so the corners are exact.
    quad4: the points are stored p-fastest, (-,-), (+,-), (-,+), (+,+), while L is
    written for the node order (-,-), (+,-), (+,+), (-,+); the reference therefore
    extrapolates fields varying along x with points 2 and 3 exchanged, and the
    oracle must do the same."""
    E, nu = 2.0e3, 0.3
    g = 1 / np.sqrt(3)
    if kind == "quad4":
        dim, x = 2, np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], float)
        gp = [(-g, -g), (g, -g), (-g, g), (g, g)]  # q outer, p inner (pdfemEleQuad4N.cpp:93-110)
        a, b, c = 1 + 0.5 * np.sqrt(3), -0.5, 1 - 0.5 * np.sqrt(3)
        L = np.array([[a, b, c, b], [b, a, b, c], [c, b, a, b], [b, c, b, a]])
    else:
        dim = 3
        x = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)
        pI, qI, rI = [-1, 1, 1, -1, -1, 1, 1, -1], [-1, -1, 1, 1, -1, -1, 1, 1], [-1, -1, -1, -1, 1, 1, 1, 1]
        gp = [(g * pI[i], g * qI[i], g * rI[i]) for i in range(8)]
        L = np.array([[(1 + pI[j] * np.sqrt(3) * pI[i]) * (1 + qI[j] * np.sqrt(3) * qI[i]) *
                       (1 + rI[j] * np.sqrt(3) * rI[i]) / 8 for j in range(8)] for i in range(8)])
    M = _single(dim, len(x), x, E, nu)
    u = np.zeros((len(x), dim))
    u[:, 0] = x[:, 0] * x[:, 1]
    sig = M.node_stresses(u.ravel())
    D = _D(dim, E, nu)
    sgp = []
    for pt in gp:
        X, Y = 0.5 * (pt[0] + 1), 0.5 * (pt[1] + 1)  # unit element: natural -> physical
        eps = np.array([Y, 0.0, X]) if dim == 2 else np.array([Y, 0.0, 0.0, X, 0.0, 0.0])
        sgp.append(D @ eps)
    nodal = L @ np.array(sgp)
    for i in range(len(x)):
        want = np.array([nodal[i, 0], nodal[i, 1], 0, nodal[i, 2], 0, 0]) if dim == 2 else nodal[i]
        assert np.abs(sig[i] - want).max() <= 1e-10 * np.abs(nodal).max(), (i, sig[i], want)
    if kind == "hex8":  # node-ordered points: exact corner values
        for i in range(len(x)):
            X, Y = x[i, 0], x[i, 1]
            assert np.allclose(sig[i], D @ np.array([Y, 0, 0, X, 0, 0]), rtol=0, atol=1e-9 * np.abs(D).max())


# ---------------------------------------------------------------------------
# quasi-static load steps with the maximum-principal-stress bond failure (2D, FC 2)
# ---------------------------------------------------------------------------

def _quasi_case(seed=1):
    m = W.crack2d(seed=seed, nx=41, ny=21, dx=1e-3)
    m["ebc"][2] = dict(nid=m["ebc"][2]["nid"], dof=1, val=0.0, vary=True, vel=2e-6)
    return m


def test_quasi_static_first_load_step_is_the_static_solution():
    """Step 1 of pdfemQuasiStaticSolver carries no load (the ramped EBC starts at 0) and
    advances the EBC by vel; step 2 solves K_ff du = -K u at the new prescribed values,
    which for a linear model is the static solution with that EBC (known answer)."""
    import scipy.sparse.linalg as spla
    m = _quasi_case()
    M = O.Model(m)
    u, info = M.quasi_static_run(2, fc=2, topk=2, sigma_ult=1e9)
    assert info[0, 3] == 0 and info[0, 2] == 1.0 and info[1, 1] == 0
    ms = _quasi_case()
    ms["ebc"][2] = dict(nid=ms["ebc"][2]["nid"], dof=1, val=2e-6)
    S = O.Model(ms)
    K, F = S.assemble(1)
    us = spla.spsolve(S.sparse(K).tocsc(), F)
    eq = S.nodes()["eq"][:, :2].ravel()
    free = eq >= 0  # the prescribed values have already advanced to the next load level
    assert np.linalg.norm(u[free] - us[eq[free]]) <= 1e-12 * np.linalg.norm(us)


def test_quasi_static_fc2_breaks_bond_pairs_along_cracks():
    """The FC 2 update puts each crack segment through a bond's midpoint; the bond and its
    reverse (same midpoint, same stress mean) are cut together, so broken bonds come in
    pairs; bonds break, the load is held while they do, and the damage is the broken
    fraction of each family."""
    M = O.Model(_quasi_case())
    u, info = M.quasi_static_run(6, fc=2, topk=2, sigma_ult=3.56e7)
    assert info[:, 1].sum() > 0 and np.all(info[:, 0] <= 2)
    fam = M.families()
    off, mem, bond = fam["off"], fam["mem"], fam["bond"]
    state = {}
    for f in range(len(off) - 1):
        k = mem[off[f]]
        for q in range(off[f] + 1, off[f + 1]):
            state[(k, mem[q])] = bond[q]
    broken = [(a, b) for (a, b), v in state.items() if v == 0]
    assert broken and all(state.get((b, a), 0) == 0 for a, b in broken)
    phi = M.local_damage()
    for f in range(len(off) - 1):
        n = off[f + 1] - off[f] - 1
        assert phi[mem[off[f]] - 1] == 1.0 - bond[off[f] + 1: off[f + 1]].sum() / n


def _seg_cross(a1, a2, b1, b2):
    """pdsolve::intersection (pdsolve.cpp:9058-9079), brute-force copy for the check."""
    if (max(a1[0], a2[0]) < min(b1[0], b2[0]) or max(a1[1], a2[1]) < min(b1[1], b2[1]) or
            max(b1[0], b2[0]) < min(a1[0], a2[0]) or max(b1[1], b2[1]) < min(a1[1], a2[1])):
        return False
    s1 = (((a1[0] - b1[0]) * (b2[1] - b1[1]) - (a1[1] - b1[1]) * (b2[0] - b1[0])) *
          ((a2[0] - b1[0]) * (b2[1] - b1[1]) - (a2[1] - b1[1]) * (b2[0] - b1[0])))
    s2 = (((b1[0] - a1[0]) * (a2[1] - a1[1]) - (b1[1] - a1[1]) * (a2[0] - a1[0])) *
          ((b2[0] - a1[0]) * (a2[1] - a1[1]) - (b2[1] - a1[1]) * (a2[0] - a1[0])))
    return not (s1 > 0 or s2 > 0)


@pytest.mark.parametrize("field", ["syy", "sxx", "sxy"])
def test_quasi_static_fc2_cracks_are_normal_to_the_principal_stress(field):
    """One FC 2 update on an all-PD plate under a uniform stress (sigma_yy, sigma_xx or
    pure shear): every bond ties, so the topk bonds kept are the first ones in family /
    member order; each puts a crack segment of length minDelta through its midpoint,
    normal to the principal direction (y, x, or the diagonal of the shear); the bonds cut
    are exactly those a brute-force segment test over all bonds finds (a crack through a
    bond's midpoint on a node row also cuts the collinear bonds of that row, as the
    reference's segment test counts a collinear overlap as a crossing)."""
    m = W.crack2d(seed=1, nx=15, ny=11, dx=1e-3)
    m["cracks"] = np.zeros((0, 3, 3))
    m["algo"] = np.ones_like(m["algo"])
    M = O.Model(m)
    sig = np.zeros((M.nn, 6))
    sig[:, {"syy": 1, "sxx": 0, "sxy": 3}[field]] = 1e6
    kept, broke, smax = M.fc2_step(sig, 5e5, 3)
    assert kept == 3 and smax == 1e6 and broke > 0
    fam = M.families()
    off, mem, bond = fam["off"], fam["mem"], fam["bond"]
    x = m["x"]
    e = {"syy": (0.0, 1.0), "sxx": (1.0, 0.0), "sxy": (2 ** -0.5, 2 ** -0.5)}[field]
    half = 0.5 * M.geom()["minDelta"]
    firsts = [(f, q) for f in range(len(off) - 1) for q in range(off[f] + 1, off[f + 1])][:3]
    segs = []
    for f, q in firsts:
        xc = 0.5 * (x[mem[off[f]] - 1] + x[mem[q] - 1])
        segs.append(((xc[0] - half * e[1], xc[1] + half * e[0]), (xc[0] + half * e[1], xc[1] - half * e[0])))
    expect = np.ones_like(bond)
    for f in range(len(off) - 1):
        xk = x[mem[off[f]] - 1]
        for q in range(off[f] + 1, off[f + 1]):
            xm = x[mem[q] - 1]
            if any(_seg_cross(xk, xm, s0, s1) for s0, s1 in segs):
                expect[q] = 0
    assert np.array_equal(bond, expect)


def test_threaded_explicit_force_matches_serial():
    """The CPU baseline's threaded explicit force (the reference's rank slices as threads,
    F summed in rank order) gives the serial trajectory to rounding, for thread counts that
    do and do not divide the family, boundary-entity and element counts."""
    mesh = W.box3d(n=10, pd_block=4)
    mesh["dt"] = 2e-8
    u1, _ = O.Model(mesh).explicit_run(6)
    for p in (3, 8):
        O.lib().om_set_threads(p)
        try:
            up, _ = O.Model(mesh).explicit_run(6)
        finally:
            O.lib().om_set_threads(1)
        assert np.linalg.norm(up - u1) <= 1e-13 * np.linalg.norm(u1)


# ---------------------------------------------------------------------------
# consistent mass of the explicit driver (LUMPMASS OFF, the reference's default)
# ---------------------------------------------------------------------------

def test_consistent_mass_textbook_quad4_and_hex8():
    """Rectangle a x b (unit thickness) and box a x b x c: the exactly integrated bilinear /
    trilinear consistent mass is rho V / 36 (quad4) or rho V / 216 (hex8) times 2^s, s = number
    of natural coordinates two corner nodes share (4 2 1 / 8 4 2 1), on each displacement
    component, with no coupling between components."""
    a, b, c, rho = 2e-3, 1e-3, 3e-3, 2700.0
    q = one_quad(rho=rho)
    q["x"] = np.array([[0, 0, 0], [a, 0, 0], [a, b, 0], [0, b, 0]], float)
    Me = O.Model(q).element_M(0)
    nat = np.array([[-1, -1], [1, -1], [1, 1], [-1, 1]])
    want = np.zeros((8, 8))
    for i in range(4):
        for j in range(4):
            s = int((nat[i] == nat[j]).sum())
            for d in range(2):
                want[2 * i + d, 2 * j + d] = rho * a * b / 36 * 2 ** s
    assert np.allclose(Me, want, rtol=1e-13, atol=0)
    # hex8 via a one-element box
    sgn = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1], [-1, -1, 1], [1, -1, 1], [1, 1, 1],
                    [-1, 1, 1]])
    x = (sgn + 1) / 2 * np.array([a, b, c])
    h = dict(dim=3, proType=-1, E=70e9, nu=0.33, rho=rho, x=x, econ=np.arange(1, 9, dtype=np.int32)[None, :],
             nen=np.array([8], np.int32), algo=np.array([2], np.int32),
             ebc=[dict(nid=np.array([1], np.int32), dof=d, val=0.0) for d in range(3)], nbc=[],
             cracks=np.zeros((0, 3, 3)), fac=3.0, nlf=0.33, me=0.0, dt=1.0, beta=0.25, gamma=0.5)
    Me = O.Model(h).element_M(0)
    want = np.zeros((24, 24))
    for i in range(8):
        for j in range(8):
            s = int((sgn[i] == sgn[j]).sum())
            for d in range(3):
                want[3 * i + d, 3 * j + d] = rho * a * b * c / 216 * 2 ** s
    assert np.allclose(Me, want, rtol=1e-13, atol=0)


def _mass_pattern_brute(M):
    """setCSRIndexes_gloMassMat by definition, with sets: a free row of node k holds the free
    DOFs of k itself when k has a family, and of every node sharing an FE element with k;
    rows of pure PD nodes (type 2) only their diagonal."""
    nd = M.nodes()
    eq, ntype, algo = nd["eq"], nd["ntype"], nd["algo"]
    d = M.dim
    fam_nodes = set(M.families()["node"].tolist())
    econ, nen = M.mesh["econ"], M.mesh["nen"]
    near = [set() for _ in range(M.nn)]
    for k in fam_nodes:
        near[k].add(k)
    for e in range(len(econ)):
        if algo[e] == 2:
            ns = [int(v) - 1 for v in econ[e, : nen[e]]]
            for p in ns:
                near[p].update(ns)
    rows = {}
    for k in range(M.nn):
        for i in range(d):
            r = eq[k, i]
            if r < 0:
                continue
            if ntype[k] == 2:
                rows[r] = [r]
            else:
                rows[r] = sorted(eq[q, j] for q in near[k] for j in range(d) if eq[q, j] >= 0)
    ia = np.zeros(M.neq + 1, np.int64)
    for r in range(M.neq):
        ia[r + 1] = ia[r] + len(rows[r])
    return ia, np.concatenate([np.array(rows[r], np.int64) for r in range(M.neq)])


@pytest.mark.parametrize("case", ["plate_pd_strip", "crack2d", "box3d"])
def test_mass_pattern_brute_force(case):
    """The explicit driver's consistent-mass pattern against its definition (sets over a tiny
    mesh); its values hold rho cd_v on the diagonal of pure PD rows and sum to the FE mass plus
    the PD share over free rows and columns of a mesh without prescribed DOFs."""
    if case == "plate_pd_strip":
        mesh = W.plate2d(nx=11, ny=6, dx=1e-3, pd_x=(0.003, 0.006))
    elif case == "crack2d":
        mesh = W.crack2d(seed=3, nx=21, ny=11, dx=1e-3)
    else:
        mesh = W.box3d(n=7, pd_block=4)
    M = O.Model(mesh)
    ia, ja, v = M.mass_csr()
    ib, jb = _mass_pattern_brute(M)
    assert np.array_equal(ia, ib) and np.array_equal(ja, jb)
    nd = M.nodes()
    fam = M.families()
    rho = mesh["rho"]
    for f, k in enumerate(fam["node"]):
        if nd["ntype"][k] != 2:
            continue
        for i in range(M.dim):
            r = nd["eq"][k, i]
            if r >= 0:
                assert ia[r + 1] - ia[r] == 1 and v[ia[r]] == rho * fam["cd_v"][f]
    free = dict(mesh, ebc=[])
    F = O.Model(free)
    ia, ja, v = F.mass_csr()
    ev, algo = F.nodes()["ev"], F.nodes()["algo"]
    want = F.dim * rho * (ev[algo == 2].sum() + F.families()["cd_v"].sum())
    assert np.isclose(v.sum(), want, rtol=1e-12)


def test_explicit_consistent_mass_two_dof_closed_form():
    """Two free DOFs (nodes 2 and 3, UX) of the unit quad with consistent mass: the central
    difference with the start-up step u1 = u0 + dt^2/2 a0 and M a = F is, mode by mode of
    M^-1 K = V diag(lam) V^-1, u_n = K^-1 f - V diag(cos n theta) V^-1 K^-1 f with
    2 - 2 cos theta = dt^2 lam (from rest).  The lumped mass gives a measurably different answer."""
    mesh = dict(one_quad(), ebc=one_quad()["ebc"][:2], lumped_mass=False)
    M = O.Model(mesh)
    assert M.neq == 2
    K, F = M.assemble(1)
    Kd = M.sparse(K).toarray()
    ia, ja, mv = M.mass_csr()
    import scipy.sparse as sp
    Md = sp.csr_matrix((mv, ja, ia), shape=(2, 2)).toarray()
    assert Md[0, 1] != 0  # coupled: nodes 2 and 3 share an edge
    dt = mesh["dt"]
    lam, V = np.linalg.eig(np.linalg.solve(Md, Kd))
    th = np.arccos(1 - dt * dt * lam.real / 2)
    us = np.linalg.solve(Kd, F)
    c0 = np.linalg.solve(V.real, us)
    for n in (0, 3, 40):
        u, _, _ = O.Model(mesh).dynamic_run(n)
        want = us - V.real @ (np.cos((n + 1) * th) * c0)
        assert np.allclose(u, want, rtol=1e-9, atol=1e-12 * np.abs(us).max()), (n, u, want)
    ul, _, _ = O.Model(dict(mesh, lumped_mass=True)).dynamic_run(40)
    assert np.linalg.norm(ul - u) > 1e-3 * np.linalg.norm(u)


def test_dynamic_run_lumped_equals_c_driver():
    """The Python dynamic driver with lumped mass and FC 0 is the C driver's loop bit for bit."""
    mesh = W.box3d(n=8, pd_block=4)
    mesh["dt"] = 2e-8
    mesh["nbc"][0].update(vel=1e9, vary=True)
    u0, up0 = O.Model(mesh).explicit_run(7)
    u1, up1, _ = O.Model(mesh).dynamic_run(7, lumped=True, fc=0)
    assert np.array_equal(u0, u1) and np.array_equal(up0, up1)


# ---------------------------------------------------------------------------
# Newmark with ramped essential and natural BCs
# ---------------------------------------------------------------------------

def test_newmark_ramped_ebc_and_nbc_order():
    """One free DOF u (node 3 UX) coupled to a ramped prescribed DOF p (node 2 UX, velocity w)
    and loaded by a ramped traction t (velocity s): step n (1-based) solves
    (m + beta dt^2 k) a' = f1 t_{n-1} - k U_N - c p_n, U_N = u + dt v + dt^2 (1/2 - beta) a, with
    the predictor's prescribed value already advanced (p_n = n dt w, pdsolve.cpp:6350-6366) and
    the traction not yet advanced (t_{n-1} = t0 + (n-1) dt s: setVaryNaturalBC runs after the
    update, :6252).  Either other order gives a measurably different answer."""
    t0, s, w = 1e5, 4e9, 0.5
    base = one_quad(dt=2e-4)
    M0 = O.Model(base)
    K, F = M0.assemble(1)
    k, f1 = K[0], F[0] / t0
    m = M0.mass_on_K()[0]
    Mq = O.Model(dict(one_quad(), ebc=one_quad()["ebc"][:2]))
    Kq, _ = Mq.assemble(1)
    eqq = Mq.nodes()["eq"]
    c = Mq.sparse(Kq).toarray()[eqq[2, 0], eqq[1, 0]]
    mesh = one_quad(vebc=True, dt=2e-4)
    mesh["ebc"][2]["vel"] = w
    mesh["nbc"][0].update(val=t0, vel=s, vary=True)
    dt, beta, gamma = mesh["dt"], mesh["beta"], mesh["gamma"]

    def run(nsteps, p_of, t_of):
        u = v = a = 0.0
        for n in range(1, nsteps + 1):
            UN = u + dt * v + dt * dt * (0.5 - beta) * a
            a1 = (f1 * t_of(n) - k * UN - c * p_of(n)) / (m + beta * dt * dt * k)
            u, v, a = (u + dt * v + dt * dt * ((0.5 - beta) * a + beta * a1),
                       v + dt * ((1 - gamma) * a + gamma * a1), a1)
        return u, v, a

    N = 25
    want = run(N, lambda n: n * dt * w, lambda n: t0 + (n - 1) * dt * s)
    u, v, a = O.newmark_run(O.Model(mesh), N)
    assert np.allclose([u[0], v[0], a[0]], want, rtol=1e-9)
    for alt in (run(N, lambda n: (n - 1) * dt * w, lambda n: t0 + (n - 1) * dt * s),
                run(N, lambda n: n * dt * w, lambda n: t0 + n * dt * s)):
        assert abs(alt[0] - u[0]) > 1e-4 * abs(u[0])


# ---------------------------------------------------------------------------
# quasi-static load ratio (setVaryEssentialBC with rati, failureProcess's 1.001 sigma_ult / s1)
# ---------------------------------------------------------------------------

def test_quasi_static_load_ratio_rule():
    """A linear model loaded by one ramped EBC (velocity w, dt = 1): the largest principal bond
    stress s1 is proportional to the EBC value.  With sigma_ult = 2.6 s1(w): step 1 has
    (rati - 1) w > w, so the value grows by w (s1 doubles at step 2); step 2 has
    (rati - 1) 2w < w, so the value becomes rati * 2w and step 3's s1 is 1.001 sigma_ult
    (pdsolve.cpp:5274-5296, :6616-6627), where bonds are then selected."""
    m = _quasi_case()
    _, info0 = O.Model(m).quasi_static_run(2, fc=2, topk=2, sigma_ult=1e30)
    s1 = info0[1, 3]
    assert s1 > 0 and info0[1, 0] == 0
    sig = 2.6 * s1
    _, info = O.Model(m).quasi_static_run(4, fc=2, topk=2, sigma_ult=sig)
    assert info[0, 3] == 0 and info[0, 2] == 1.0
    assert np.isclose(info[1, 3], s1, rtol=1e-12) and info[1, 0] == 0
    assert np.isclose(info[1, 2], 1.001 * sig / s1, rtol=1e-12)
    assert np.isclose(info[2, 3], 2 * s1, rtol=1e-10) and info[2, 0] == 0
    assert np.isclose(info[3, 3], 1.001 * sig, rtol=1e-9)
    assert info[3, 0] > 0


# ---------------------------------------------------------------------------
# bond failure inside the dynamic drivers (FC 2, 2D)
# ---------------------------------------------------------------------------

def _dyn_fail_case(lumped):
    m = W.crack2d(seed=1, nx=31, ny=16, dx=1e-3)
    m["ebc"][2] = dict(nid=m["ebc"][2]["nid"], dof=1, val=0.0, vary=True, vel=0.2)
    m.update(lumped_mass=lumped, dt=2e-8, fc=2, topk=0)
    return m


@pytest.mark.parametrize("lumped", [True, False])
def test_dynamic_fc2_after_each_stored_step(lumped):
    """failureProcess runs after every stored step on that step's DOF values (free u_{n+1},
    prescribed values of the step's force): with a strength between the largest bond stresses
    of steps j-1 and j the FC 2 trajectory equals the FC 0 one through step j, the bonds broken at
    step j are those one FC 2 update breaks on the FC 0 state of step j, and the next step
    differs."""
    m = _dyn_fail_case(lumped)
    j = 6  # info row j belongs to the stored step u_{j+1}
    _, _, info_ref = O.Model(m).dynamic_run(j + 1, sigma_ult=1e30)
    s = info_ref[:, 2]
    assert np.all(np.diff(s[: j + 1]) > 0) and np.all(info_ref[:, 1] == 0)
    sig_ult = 0.5 * (s[j - 1] + s[j])
    Mj = O.Model(m)
    uj, _, _ = Mj.dynamic_run(j, sigma_ult=1e30)
    S = O.Model(m)
    kept, broke, smax = S.fc2_step(S.node_stresses(Mj.values()), sig_ult, 0)
    assert kept > 0 and broke > 0 and smax == s[j]
    F = O.Model(m)
    uf, _, info = F.dynamic_run(j, sigma_ult=sig_ult)
    assert np.array_equal(uf, uj) and np.all(info[:j, 1] == 0)
    assert info[j, 0] == kept and info[j, 1] == broke
    assert np.array_equal(F.families()["bond"], S.families()["bond"])
    ug, _, _ = O.Model(m).dynamic_run(j + 1, sigma_ult=sig_ult)
    uh, _, _ = O.Model(m).dynamic_run(j + 1, sigma_ult=1e30)
    assert not np.array_equal(ug, uh)
