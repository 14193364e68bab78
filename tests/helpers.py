"""Shared test helpers: loaders for the package, the oracle and the fixtures."""
import importlib.util
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "pdlsm-fem_mpi_amd")


def load_pdx():
    name = "pdlsm_fem_mpi_amd"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def load_oracle():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    oracle.build()
    return oracle


def load_gnn_oracle():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import gnn_oracle
    return gnn_oracle


def workloads():
    sys.path.insert(0, PKG_DIR)
    import workloads as W
    return W


def box_fixture():
    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    import make_box_fixture
    return make_box_fixture.load_box()


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def row_block_rel(Ka, Kb, ia):
    """max over rows of ||row_a - row_b|| / ||row_b|| (norm-wise per CSR row)."""
    d = (np.asarray(Ka) - np.asarray(Kb)) ** 2
    b = np.asarray(Kb) ** 2
    rd = np.sqrt(np.add.reduceat(d, ia[:-1])) if len(d) else np.zeros(0)
    rb = np.sqrt(np.add.reduceat(b, ia[:-1])) if len(b) else np.zeros(0)
    nz = rb > 0
    return float(np.max(rd[nz] / rb[nz])) if nz.any() else 0.0


def one_quad(vebc=False, E=1e9, nu=0.25, rho=1000.0, t=1e5, dt=1e-5):
    """Unit-square quad4 with a single free DOF (node 3 UX); traction t on edge 2-3."""
    x = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], float)
    econ = np.zeros((1, 8), np.int32)
    econ[0, :4] = [1, 2, 3, 4]
    ebc = [dict(nid=np.array([1, 4], np.int32), dof=0, val=0.0), dict(nid=np.array([1, 2, 3, 4], np.int32), dof=1, val=0.0),
           dict(nid=np.array([2], np.int32), dof=0, val=0.0, vel=(1.0 if vebc else 0.0), vary=vebc)]
    nbc = [dict(con=np.array([[2, 3]], np.int32), val=t)]
    return dict(dim=2, proType=1, E=E, nu=nu, rho=rho, x=x, econ=econ, nen=np.array([4], np.int32),
                algo=np.array([2], np.int32), ebc=ebc, nbc=nbc, cracks=np.zeros((0, 3, 3)), fac=3.0, nlf=0.33,
                me=0.0, dt=dt, beta=0.25, gamma=0.5)


def tri_tet_mesh():
    """Mixed 2D mesh: quads in the left half, each right-half quad split into two triangles."""
    W = workloads()
    m = W.plate2d(nx=31, ny=17, dx=1e-3, pd_x=(0.008, 0.02))
    econ, nen, algo = [], [], []
    for e in range(len(m["econ"])):
        c = m["econ"][e]
        cx = m["x"][c[:4] - 1, 0].mean()
        if cx > 0.015:
            for tri in ([c[0], c[1], c[2]], [c[0], c[2], c[3]]):
                row = np.zeros(8, np.int32)
                row[:3] = tri
                econ.append(row); nen.append(3); algo.append(m["algo"][e])
        else:
            econ.append(c); nen.append(4); algo.append(m["algo"][e])
    m.update(econ=np.array(econ), nen=np.array(nen, np.int32), algo=np.array(algo, np.int32))
    return m


def tet_box_mesh(n=6, dx=1e-3, seed=0):
    """All-FE 3D box of (n-1)^3 cubes, each split into 6 tetrahedra around the main
    diagonal; bottom fixed, top traction."""
    rng = np.random.default_rng(seed)
    idx = lambda i, j, k: i + n * (j + n * k)
    x = np.array([[i, j, k] for k in range(n) for j in range(n) for i in range(n)], float) * dx
    inner = (x > 0).all(1) & (x < (n - 1) * dx).all(1)
    x[inner] += rng.uniform(-0.15, 0.15, (inner.sum(), 3)) * dx
    tets = []
    for k in range(n - 1):
        for j in range(n - 1):
            for i in range(n - 1):
                v = [idx(i, j, k), idx(i + 1, j, k), idx(i + 1, j + 1, k), idx(i, j + 1, k),
                     idx(i, j, k + 1), idx(i + 1, j, k + 1), idx(i + 1, j + 1, k + 1), idx(i, j + 1, k + 1)]
                for a, b in ((1, 2), (2, 3), (3, 7), (7, 4), (4, 5), (5, 1)):
                    t = [v[0], v[a], v[b], v[6]]
                    p = x[t]
                    if np.linalg.det(p[1:] - p[0]) < 0:
                        t[1], t[2] = t[2], t[1]
                    tets.append(t)
    econ = np.zeros((len(tets), 8), np.int32)
    econ[:, :4] = np.array(tets) + 1
    bottom = np.array([idx(i, j, 0) + 1 for j in range(n) for i in range(n)], np.int32)
    ebc = [dict(nid=bottom, dof=d, val=0.0) for d in range(3)]
    return dict(dim=3, proType=-1, E=70e9, nu=0.33, rho=2700.0, x=x, econ=econ,
                nen=np.full(len(tets), 4, np.int32), algo=np.full(len(tets), 2, np.int32), ebc=ebc, nbc=[],
                cracks=np.zeros((0, 3, 3)), fac=3.015, nlf=0.33, me=0.0, dt=1.0, beta=0.25, gamma=0.5)
