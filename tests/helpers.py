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
