"""Python binding of libpdx.so, the MI355X PDLSM-FEM hot-path library.

The package directory name contains a hyphen, so import it by path:

    import importlib.util, sys
    spec = importlib.util.spec_from_file_location(
        "pdlsm_fem_mpi_amd", "pdlsm-fem_mpi_amd/__init__.py",
        submodule_search_locations=["pdlsm-fem_mpi_amd"])

Everything here is a thin ctypes layer over the C-ABI declared in
include/pdx.h; all compute runs in the HIP library.  A missing library raises
at import of the first object instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpdx.so")

PDX_MODE_EXPLICIT, PDX_MODE_STATIC, PDX_MODE_QUASI, PDX_MODE_NEWMARK = 0, 1, 2, 3
PDX_FP64, PDX_FP32 = 0, 1
PDX_KRYLOV_BICGSTAB, PDX_KRYLOV_CG = 0, 1
PDX_PREC_NONE, PDX_PREC_JACOBI, PDX_PREC_BLOCK_JACOBI, PDX_PREC_LEARNED, PDX_PREC_TWO_LEVEL = 0, 1, 2, 3, 4
PDX_PREC_ILU0 = 5
FB_VARIANTS = 7  # family-batch launch forms (PDX_FB_VARIANTS)
EXPLICIT_FORMS = 24  # explicit-step kernel forms (PDX_EXPLICIT_FORMS)

_P, _I, _D, _LL = C.c_void_p, C.c_int, C.c_double, C.c_longlong


class MeshDesc(C.Structure):
    _fields_ = [("dim", _I), ("pro_type", _I), ("E", _D), ("nu", _D), ("rho", _D),
                ("n_node", _I), ("xyz", _P), ("n_ele", _I), ("ele_nodes", _P), ("ele_nen", _P),
                ("ele_algo", _P), ("n_ebc", _I), ("ebc_off", _P), ("ebc_nid", _P), ("ebc_dof", _P),
                ("ebc_val", _P), ("ebc_vel", _P), ("ebc_vary", _P), ("n_nbc", _I), ("nbc_off", _P),
                ("nbc_con", _P), ("nbc_val", _P), ("nbc_vel", _P), ("nbc_vary", _P), ("n_crack", _I),
                ("crack", _P), ("horizon_factor", _D), ("nlf", _D), ("me", _D), ("dt", _D), ("beta", _D),
                ("gamma", _D), ("sigma_ult", _D), ("kic", _D), ("fc", _I), ("topk", _I),
                ("lumped_mass", _I), ("newmark", _I), ("save_freq", _I), ("vtk_ascii", _I)]


class Settings(C.Structure):
    _fields_ = [("solver_flag", _I), ("newmark", _I), ("lumped_mass", _I), ("num_steps", _I), ("save_freq", _I),
                ("vtk_binary", _I), ("fc", _I), ("topk", _I), ("n_rf", _I), ("rf_ids", _I * 16), ("dt", _D),
                ("beta", _D), ("gamma", _D), ("horizon_factor", _D), ("nlf", _D), ("sigma_ult", _D), ("kic", _D)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_ if k != "rf_ids"}
        d["rf_ids"] = [self.rf_ids[i] for i in range(min(self.n_rf, 16))]
        return d


class SolveOpts(C.Structure):
    _fields_ = [("krylov", _I), ("precond", _I), ("precision", _I), ("tol", _D), ("max_iter", _I),
                ("inner_tol", _D), ("warm_start", _I), ("coarse_agg", _I), ("fallback", _I),
                ("no_graph", _I), ("mirrored", _I)]


class SolveStats(C.Structure):
    _fields_ = [("iterations", _I), ("refinements", _I), ("true_relres", _D), ("seconds", _D),
                ("converged", _I), ("precond_used", _I), ("precision_used", _I), ("attempts", _I),
                ("mirrored_used", _I)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None

EXPORTS = {
    "pdx_last_error": (C.c_char_p, []),
    "pdx_version": (C.c_char_p, []),
    "pdx_model_create": (_I, [_P, _P]),
    "pdx_model_read": (_I, [C.c_char_p, _P, _P, _P]),
    "pdx_write_vtk": (_I, [_P, C.c_char_p, C.c_char_p, _P, _P, _P, _I]),
    "pdx_write_vtk_part": (_I, [_P, C.c_char_p, C.c_char_p, _P, _P, _P, _I]),
    "pdx_write_node_table": (_I, [_P, C.c_char_p, _P, _P, _I]),
    "pdx_quasi_static_run": (_I, [_P, _I, _P, _I, _I, _D, _P, _P, _P]),
    "pdx_ctx_bonds": (_I, [_P, _P]),
    "pdx_set_incremental_bonds": (_I, [_P, _I]),
    "pdx_local_damage": (_I, [_P, _P]),
    "pdx_model_destroy": (None, [_P]),
    "pdx_model_settings": (_I, [_P, _P]),
    "pdx_model_mass_csr": (_I, [_P, _P, _P, _P]),
    "pdx_mass_csr": (_I, [_P, _P]),
    "pdx_step_log": (_I, [_P, _P, _P]),
    "pdx_model_sizes": (_I, [_P, _P]),
    "pdx_model_csr": (_I, [_P, _P, _P]),
    "pdx_model_adjacency": (_I, [_P, _P, _P]),
    "pdx_model_families": (_I, [_P, _P, _P, _P, _P, _P, _P]),
    "pdx_model_nodes": (_I, [_P, _P, _P, _P, _P, _P, _P]),
    "pdx_model_values": (_I, [_P, _P]),
    "pdx_ctx_create": (_I, [_P, _I, _P]),
    "pdx_ctx_destroy": (None, [_P]),
    "pdx_ctx_stream": (_P, [_P]),
    "pdx_assemble": (_I, [_P, _I, _P, _P, _P, _I]),
    "pdx_lumped_mass": (_I, [_P, _P]),
    "pdx_assemble_timed": (_I, [_P, _P]),
    "pdx_mass_on_K": (_I, [_P, _P]),
    "pdx_set_dof_values": (_I, [_P, _P]),
    "pdx_explicit_run": (_I, [_P, _I, _I, _P, _P]),
    "pdx_explicit_steps": (_I, [_P, _I, _I, _I]),
    "pdx_max_eig": (_I, [_P, _I, _P]),
    "pdx_set_dt": (_I, [_P, _D]),
    "pdx_static_solve": (_I, [_P, _P, _P, _P]),
    "pdx_ilu0": (_I, [_P, _P, _P, _P, _P, _P, _P]),
    "pdx_newmark_run": (_I, [_P, _I, _P, _P, _P, _P, _P]),
    "pdx_solve_csr": (_I, [_I, _LL, _P, _P, _P, _P, _P, _P, _P]),
    "pdx_last_timing": (_I, [_P, _P, _P]),
    "pdx_explicit_tuning": (_I, [_P, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "pdx_explicit_set_form": (_I, [_P, _I, _I]),
    "pdx_node_stresses": (_I, [_P, _P, _P]),
    "pdx_node_stresses_timed": (_I, [_P, _P, _P]),
    "pdx_node_stresses_noalloc": (_I, [_P, _P, _P]),
    "pdx_set_setup_device": (_I, [_I]),
    "pdx_gnn_operator": (_I, [_P, _I]),
    "pdx_gnn_attach": (_I, [_P, _P]),
    "pdx_gnn_create": (_I, [_I, _I, C.c_ulonglong, _P]),
    "pdx_gnn_destroy": (None, [_P]),
    "pdx_gnn_num_params": (_LL, [_P]),
    "pdx_gnn_get_params": (_I, [_P, _P]),
    "pdx_gnn_set_params": (_I, [_P, _P]),
    "pdx_gnn_train": (_I, [_P, _P, _I, C.c_ulonglong, _I, _I, C.c_float, _P]),
    "pdx_gnn_loss_grad": (_I, [_P, _P, _I, C.c_ulonglong, _I, _P, _P]),
    "pdx_gnn_inputs": (_I, [_P, _I, C.c_ulonglong, _I, _P, _P, _P, _P, _P, _P]),
    "pdx_gnn_apply": (_I, [_P, _P]),
    "pdx_gnn_infer": (_I, [_P, _P, _I, _P, _P]),
    "pdx_fb_create": (_I, [_I, _I, _LL, _P, _P, _P, _D, _D, _D, _P]),
    "pdx_fb_destroy": (None, [_P]),
    "pdx_fb_sizes": (_I, [_P, _P, _P, _P]),
    "pdx_fb_families": (_I, [_P, _P, _P]),
    "pdx_fb_run": (_I, [_P, _I, _P]),
    "pdx_fb_tuning": (_I, [_P, _P, _P]),
    "pdx_fb_results": (_I, [_P, _LL, _LL, _P, _P, _P]),
}


def lib():
    """Load libpdx.so; raises if it was not built (no fallback path exists)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run `python pdlsm-fem_mpi_amd/build.py` first")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in EXPORTS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class PdxError(RuntimeError):
    pass


def _chk(rc):
    if rc != 0:
        raise PdxError(f"pdx error {rc}: {lib().pdx_last_error().decode()}")


def set_setup_device(device: int) -> None:
    """Run the family search and the node graph of later Model(...) set-ups on GPU `device`
    (-1: on the host, the default); both give the same families and pattern."""
    _chk(lib().pdx_set_setup_device(int(device)))


def mesh_desc(mesh: dict):
    """Build a pdx_mesh_desc from a raw-mesh dict (see workloads.py); returns
    (desc, keepalive list)."""
    keep = []

    def arr(x, dt):
        a = np.ascontiguousarray(np.asarray(x, dtype=dt))
        keep.append(a)
        return a

    dim = int(mesh["dim"])
    x = arr(mesh["x"], np.float64).reshape(-1, 3)
    econ = arr(mesh["econ"], np.int32).reshape(-1, 8)
    nen = arr(mesh["nen"], np.int32)
    algo = arr(mesh["algo"], np.int32)
    ebc = mesh.get("ebc", [])
    nbc = mesh.get("nbc", [])
    tn = 4 if dim == 3 else 2
    ebc_off = arr(np.concatenate([[0], np.cumsum([len(s["nid"]) for s in ebc])]), np.int32)
    ebc_nid = arr(np.concatenate([np.asarray(s["nid"]) for s in ebc]) if ebc else np.zeros(1), np.int32)
    nbc_off = arr(np.concatenate([[0], np.cumsum([len(s["con"]) for s in nbc])]), np.int32)
    nbc_con = arr(np.concatenate([np.asarray(s["con"]).reshape(-1, tn) for s in nbc]).ravel()
                  if nbc else np.zeros(1), np.int32)
    cracks = arr(np.asarray(mesh.get("cracks", np.zeros((0, 3, 3)))).reshape(-1, 9), np.float64)
    d = MeshDesc()
    d.dim, d.pro_type = dim, int(mesh.get("proType", 1))
    d.E, d.nu, d.rho = float(mesh["E"]), float(mesh["nu"]), float(mesh["rho"])
    d.n_node, d.xyz = len(x), _p(x)
    d.n_ele, d.ele_nodes, d.ele_nen, d.ele_algo = len(econ), _p(econ), _p(nen), _p(algo)
    d.n_ebc, d.ebc_off, d.ebc_nid = len(ebc), _p(ebc_off), _p(ebc_nid)
    d.ebc_dof = _p(arr([s["dof"] for s in ebc] or [0], np.int32))
    d.ebc_val = _p(arr([s["val"] for s in ebc] or [0.0], np.float64))
    d.ebc_vel = _p(arr([s.get("vel", 0.0) for s in ebc] or [0.0], np.float64))
    d.ebc_vary = _p(arr([int(s.get("vary", False)) for s in ebc] or [0], np.int32))
    d.n_nbc, d.nbc_off, d.nbc_con = len(nbc), _p(nbc_off), _p(nbc_con)
    d.nbc_val = _p(arr([s["val"] for s in nbc] or [0.0], np.float64))
    d.nbc_vel = _p(arr([s.get("vel", 0.0) for s in nbc] or [0.0], np.float64))
    d.nbc_vary = _p(arr([int(s.get("vary", False)) for s in nbc] or [0], np.int32))
    d.n_crack, d.crack = len(cracks), _p(cracks)
    d.horizon_factor, d.nlf, d.me = float(mesh["fac"]), float(mesh["nlf"]), float(mesh.get("me", 0.0))
    d.dt, d.beta, d.gamma = float(mesh.get("dt", 1.0)), float(mesh.get("beta", 0.25)), float(mesh.get("gamma", 0.5))
    d.sigma_ult, d.kic = float(mesh.get("sigma_ult", 0.0)), float(mesh.get("kic", 0.0))
    d.fc, d.topk = int(mesh.get("fc", 0)), int(mesh.get("topk", 0))
    # run settings: a dict without them takes the reference's defaults (consistent mass, no Newmark)
    d.lumped_mass, d.newmark = int(bool(mesh.get("lumped_mass", False))), int(bool(mesh.get("newmark", False)))
    d.save_freq, d.vtk_ascii = int(mesh.get("save_freq", 1)), int(not mesh.get("vtk_binary", True))
    return d, keep


class Model:
    """Host data model (set-up of pdsolve::setDatModel and the CSR pattern)."""

    def __init__(self, mesh: dict | None = None, cmd_file: str | None = None):
        L = lib()
        h = C.c_void_p()
        if cmd_file is not None:
            sf, ns = C.c_int(-1), C.c_int(1)
            _chk(L.pdx_model_read(cmd_file.encode(), C.byref(h), C.byref(sf), C.byref(ns)))
            self.solver_flag, self.num_steps = sf.value, ns.value
        else:
            desc, keep = mesh_desc(mesh)
            _chk(L.pdx_model_create(C.byref(desc), C.byref(h)))
        self.h = h
        s = np.zeros(16, np.int64)
        _chk(L.pdx_model_sizes(self.h, _p(s)))
        (self.nn, self.ne, self.dim, self.neq, self.nnz, self.nfam, self.nmem, self.npdbe, self.npde, self.nfe,
         self.nadj, self.npresc, nb0, nb1, nb2, self.max_family) = (int(v) for v in s)
        self.nblocks = (nb0, nb1, nb2)

    def __del__(self):
        if getattr(self, "h", None):
            lib().pdx_model_destroy(self.h)
            self.h = None

    def write_vtk(self, path, u=None, sigma=None, phi=None, title=None, binary=True):
        """Write the whole mesh with nodal displacements u ([nn, dim] or [nn, 3], or
        flat), stresses sigma [nn, 6] (xx, yy, zz, xy, yz, zx) and damage phi [nn]
        in the reference's legacy VTK layout, byte for byte
        (writeReslutsTOTAL_vtk_Binary / _ASCII, fioFile.cpp:562-769); None writes
        zeros, title None takes the mesh file's title."""
        nn = self.nn
        u3 = None
        if u is not None:
            u = np.asarray(u, np.float64).reshape(nn, -1)
            u3 = np.zeros((nn, 3))
            u3[:, : u.shape[1]] = u
        sg = None if sigma is None else np.ascontiguousarray(np.asarray(sigma, np.float64).reshape(nn, 6))
        ph = None if phi is None else np.ascontiguousarray(np.asarray(phi, np.float64).reshape(nn))
        _chk(lib().pdx_write_vtk(self.h, str(path).encode(), None if title is None else title.encode(),
                                 None if u3 is None else _p(u3), None if sg is None else _p(sg),
                                 None if ph is None else _p(ph), int(bool(binary))))

    def write_vtk_part(self, path, part, u=None, sigma=None, phi=None, title=None):
        """The reference's PD-only (part 0) or FE-only (part 1) ASCII result file
        (writeResultsPD_vtk / writeResultsFEM_vtk, fioFile.cpp:771-1001)."""
        u3, sg, ph = self._result_arrays(u, sigma, phi)
        _chk(lib().pdx_write_vtk_part(self.h, str(path).encode(), None if title is None else title.encode(),
                                      None if u3 is None else _p(u3), None if sg is None else _p(sg),
                                      None if ph is None else _p(ph), int(part)))

    def write_node_table(self, path, kind, u=None, sigma=None):
        """The .disp (kind 0) or .sig (kind 1) node table of an ASCII result (fioFile.cpp:486-538)."""
        u3, sg, _ = self._result_arrays(u, sigma, None)
        _chk(lib().pdx_write_node_table(self.h, str(path).encode(), None if u3 is None else _p(u3),
                                        None if sg is None else _p(sg), int(kind)))

    def _result_arrays(self, u, sigma, phi):
        nn = self.nn
        u3 = None
        if u is not None:
            u = np.asarray(u, np.float64).reshape(nn, -1)
            u3 = np.zeros((nn, 3))
            u3[:, : u.shape[1]] = u
        sg = None if sigma is None else np.ascontiguousarray(np.asarray(sigma, np.float64).reshape(nn, 6))
        ph = None if phi is None else np.ascontiguousarray(np.asarray(phi, np.float64).reshape(nn))
        return u3, sg, ph

    def adjacency(self):
        off = np.zeros(self.nn + 1, np.int32)
        adj = np.zeros(max(self.nadj, 1), np.int32)
        _chk(lib().pdx_model_adjacency(self.h, _p(off), _p(adj)))
        return off, adj[: self.nadj]

    def csr(self):
        ia = np.zeros(self.neq + 1, np.int64)
        ja = np.zeros(max(self.nnz, 1), np.int64)
        _chk(lib().pdx_model_csr(self.h, _p(ia), _p(ja)))
        return ia, ja[: self.nnz]

    def families(self):
        fn = np.zeros(max(self.nfam, 1), np.int32)
        fo = np.zeros(self.nfam + 1, np.int32)
        fm = np.zeros(max(self.nmem, 1), np.int32)
        bd = np.zeros(max(self.nmem, 1), np.int32)
        fv = np.zeros(max(self.nfam, 1))
        fd = np.zeros(max(self.nfam, 1))
        _chk(lib().pdx_model_families(self.h, _p(fn), _p(fo), _p(fm), _p(bd), _p(fv), _p(fd)))
        n, m = self.nfam, self.nmem
        return dict(node=fn[:n], off=fo, mem=fm[:m], bond=bd[:m], cd_v=fv[:n], delta=fd[:n])

    def nodes(self):
        eq = np.zeros(self.nn * 3, np.int32)
        nt = np.zeros(self.nn, np.int32)
        nv = np.zeros(self.nn)
        ev = np.zeros(self.ne)
        pb = np.zeros(max(self.npdbe, 1) * 4, np.int32)
        al = np.zeros(self.ne, np.int32)
        _chk(lib().pdx_model_nodes(self.h, _p(eq), _p(nt), _p(nv), _p(ev), _p(pb), _p(al)))
        return dict(eq=eq.reshape(-1, 3), ntype=nt, nv=nv, ev=ev, pdbe=pb[: self.npdbe * 4].reshape(-1, 4), algo=al)

    def values(self):
        u = np.zeros(self.nn * self.dim)
        _chk(lib().pdx_model_values(self.h, _p(u)))
        return u

    def settings(self) -> dict:
        """Run settings of the command file (solver flag, NEWMARK, LUMPMASS, SETSOLVING, FC, RF,
        VTKFORMAT) or of the mesh dict."""
        st = Settings()
        _chk(lib().pdx_model_settings(self.h, C.byref(st)))
        return st.as_dict()

    def mass_csr(self):
        """(ia, ja) of the explicit driver's consistent-mass pattern (setCSRIndexes_gloMassMat)."""
        nz = C.c_longlong(0)
        _chk(lib().pdx_model_mass_csr(self.h, None, None, C.byref(nz)))
        ia = np.zeros(self.neq + 1, np.int64)
        ja = np.zeros(max(nz.value, 1), np.int64)
        _chk(lib().pdx_model_mass_csr(self.h, _p(ia), _p(ja), C.byref(nz)))
        return ia, ja[: nz.value]


def solve_opts(krylov=PDX_KRYLOV_BICGSTAB, precond=PDX_PREC_JACOBI, precision=PDX_FP64, tol=1e-8, max_iter=20000,
               inner_tol=1e-4, warm_start=False, coarse_agg=0, fallback=False, no_graph=False, mirrored=False):
    o = SolveOpts()
    o.krylov, o.precond, o.precision, o.tol, o.max_iter = krylov, precond, precision, tol, max_iter
    o.inner_tol, o.warm_start, o.coarse_agg, o.fallback = inner_tol, int(warm_start), int(coarse_agg), int(fallback)
    o.no_graph, o.mirrored = int(no_graph), int(mirrored)
    return o


class Context:
    """Device-resident state of one model on one GPU."""

    def __init__(self, model: Model, device: int = 0):
        self.model = model
        h = C.c_void_p()
        _chk(lib().pdx_ctx_create(model.h, device, C.byref(h)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            lib().pdx_ctx_destroy(self.h)
            self.h = None

    def stream(self):
        return lib().pdx_ctx_stream(self.h)

    def set_dof_values(self, u_all):
        _chk(lib().pdx_set_dof_values(self.h, _p(np.ascontiguousarray(u_all, np.float64))))

    def assemble(self, mode=PDX_MODE_STATIC, UN=None, with_ext=True, want_K=True):
        m = self.model
        F = np.zeros(max(m.neq, 1))
        K = np.zeros(max(m.nnz, 1)) if (want_K and mode != PDX_MODE_EXPLICIT) else None
        un = None if UN is None else np.ascontiguousarray(UN, np.float64)
        _chk(lib().pdx_assemble(self.h, mode, _p(un), _p(K), _p(F), int(with_ext)))
        return (None if K is None else K[: m.nnz]), F[: m.neq]

    def assemble_timed(self):
        """Re-assemble K on the device only; returns the kernel time in ms."""
        ms = C.c_double(0)
        _chk(lib().pdx_assemble_timed(self.h, C.byref(ms)))
        return ms.value

    def lumped_mass(self):
        out = np.zeros(max(self.model.neq, 1))
        _chk(lib().pdx_lumped_mass(self.h, _p(out)))
        return out[: self.model.neq]

    def mass_on_K(self):
        out = np.zeros(max(self.model.nnz, 1))
        _chk(lib().pdx_mass_on_K(self.h, _p(out)))
        return out[: self.model.nnz]

    def mass_csr(self):
        """Values of the consistent mass on the model's mass pattern (Model.mass_csr)."""
        ia, ja = self.model.mass_csr()
        out = np.zeros(max(len(ja), 1))
        _chk(lib().pdx_mass_csr(self.h, _p(out)))
        return out[: len(ja)]

    def step_log(self):
        """Per stored step of the dynamic drivers with FC 2: [n, 3] bonds kept, bonds broken,
        largest principal bond stress."""
        n = C.c_int(0)
        _chk(lib().pdx_step_log(self.h, None, C.byref(n)))
        out = np.zeros(max(n.value, 1) * 3)
        _chk(lib().pdx_step_log(self.h, _p(out), C.byref(n)))
        return out[: n.value * 3].reshape(-1, 3)

    def explicit_run(self, nsteps, precision=PDX_FP64):
        u = np.zeros(max(self.model.neq, 1))
        up = np.zeros(max(self.model.neq, 1))
        _chk(lib().pdx_explicit_run(self.h, nsteps, precision, _p(u), _p(up)))
        return u[: self.model.neq], up[: self.model.neq]

    def explicit_steps(self, nsteps, precision=PDX_FP64, use_graph=True):
        _chk(lib().pdx_explicit_steps(self.h, nsteps, precision, int(use_graph)))

    def last_timing(self):
        t, s = C.c_double(0), C.c_double(0)
        _chk(lib().pdx_last_timing(self.h, C.byref(t), C.byref(s)))
        return t.value, s.value

    def node_stresses(self, u_all=None):
        """Nodal stresses [nn, 6] of the DOF values u_all [nn*dim] (None: explicit state)."""
        sig = np.zeros(self.model.nn * 6)
        if u_all is None:
            _chk(lib().pdx_node_stresses(self.h, None, _p(sig)))
        else:
            u = np.ascontiguousarray(u_all, np.float64)
            _chk(lib().pdx_node_stresses(self.h, _p(u), _p(sig)))
        return sig.reshape(-1, 6)

    def node_stresses_timed(self, u_all) -> float:
        u = np.ascontiguousarray(u_all, np.float64)
        ms = C.c_double()
        _chk(lib().pdx_node_stresses_timed(self.h, _p(u), C.byref(ms)))
        return ms.value

    def gnn_operator(self, mode: int):
        """Operator of the network's inputs, training and apply: 0 K, 1 the Newmark K_eff."""
        _chk(lib().pdx_gnn_operator(self.h, int(mode)))

    def explicit_set_form(self, precision, form):
        """Step with kernel form `form` (-1: tune again at the next step)."""
        _chk(lib().pdx_explicit_set_form(self.h, precision, form))

    def explicit_tuning(self, precision):
        """The explicit-step form the tuner chose: (lanes per short row, non-temporal K
        loads, XCD-contiguous block order, row-local K value order, column pattern
        classes, mirrored lower blocks, ms per form)."""
        return self.explicit_tuning_full(precision)[:7]

    def explicit_tuning_full(self, precision):
        """explicit_tuning plus (alternating sweep direction, form index)."""
        lanes, nt, xm, soa, rel, sym, alt, form = (C.c_int(0) for _ in range(8))
        ms = np.zeros(EXPLICIT_FORMS, np.float32)
        _chk(lib().pdx_explicit_tuning(self.h, precision, C.byref(lanes), C.byref(nt), C.byref(xm), C.byref(soa),
                                       C.byref(rel), C.byref(sym), C.byref(alt), C.byref(form), _p(ms)))
        return (lanes.value, bool(nt.value), bool(xm.value), bool(soa.value), bool(rel.value), bool(sym.value),
                ms, bool(alt.value), form.value)

    def max_eig(self, iters=100):
        v = C.c_double(0)
        _chk(lib().pdx_max_eig(self.h, iters, C.byref(v)))
        return v.value

    def set_dt(self, dt):
        _chk(lib().pdx_set_dt(self.h, float(dt)))

    def ilu0(self, r=None):
        """Block ILU(0) factor of the assembled K (pdx_ilu0): dict with colors [nn], F [blocks, d, d]
        (strict lower blocks L, the rest U, in adjacency order), Dinv [nn, d, d], ncolors, and z =
        (L U)^-1 r when r [nn*d] is given."""
        m = self.model
        d = m.dim
        nblk = m.nadj
        col = np.zeros(max(m.nn, 1), np.int32)
        F = np.zeros(max(nblk, 1) * d * d)
        Dinv = np.zeros(max(m.nn, 1) * d * d)
        nc = C.c_int(0)
        rr = None if r is None else np.ascontiguousarray(r, np.float64)
        z = None if r is None else np.zeros(m.nn * d)
        _chk(lib().pdx_ilu0(self.h, _p(rr), _p(z), _p(col), _p(F), _p(Dinv), C.byref(nc)))
        out = dict(colors=col[: m.nn], F=F[: nblk * d * d].reshape(nblk, d, d), Dinv=Dinv[: m.nn * d * d].reshape(m.nn, d, d),
                   ncolors=nc.value)
        if z is not None:
            out["z"] = z
        return out

    def static_solve(self, opts=None, u0=None, check=True):
        """u, stats; with check=False also the status code instead of raising on failure."""
        opts = opts or solve_opts()
        u = np.zeros(max(self.model.neq, 1)) if u0 is None else np.array(u0, np.float64)
        st = SolveStats()
        rc = lib().pdx_static_solve(self.h, C.byref(opts), _p(u), C.byref(st))
        if not check:
            return u[: self.model.neq], st.as_dict(), rc
        _chk(rc)
        return u[: self.model.neq], st.as_dict()

    def quasi_static_run(self, nsteps, opts=None, fc=-1, topk=-1, sigma_ult=0.0):
        """Quasi-static load steps with the maximum-principal-stress bond failure (fc=2,
        2D); returns (u_all [nn*dim], step_info [nsteps, 4]: bonds selected, bonds broken,
        load ratio, largest principal bond stress, stats)."""
        opts = opts or solve_opts()
        u = np.zeros(self.model.nn * self.model.dim)
        info = np.zeros((max(nsteps, 1), 4))
        st = SolveStats()
        _chk(lib().pdx_quasi_static_run(self.h, nsteps, C.byref(opts), fc, topk, sigma_ult, _p(u), _p(info),
                                        C.byref(st)))
        return u, info[:nsteps], st.as_dict()

    def set_incremental_bonds(self, on=True):
        """On (the default): bond failure recomputes only the changed families' rows of K (the same
        K bit for bit); off: every change assembles the whole K again."""
        _chk(lib().pdx_set_incremental_bonds(self.h, int(bool(on))))

    def bonds(self):
        """Bond states of this context (1 intact, 0 broken), laid out like Model.families()['bond']."""
        b = np.zeros(max(self.model.nmem, 1), np.int32)
        _chk(lib().pdx_ctx_bonds(self.h, _p(b)))
        return b[: self.model.nmem]

    def local_damage(self):
        phi = np.zeros(self.model.nn)
        _chk(lib().pdx_local_damage(self.h, _p(phi)))
        return phi

    def newmark_run(self, nsteps, opts=None):
        opts = opts or solve_opts()
        n = max(self.model.neq, 1)
        u, v, a = np.zeros(n), np.zeros(n), np.zeros(n)
        st = SolveStats()
        _chk(lib().pdx_newmark_run(self.h, nsteps, C.byref(opts), _p(u), _p(v), _p(a), C.byref(st)))
        k = self.model.neq
        return u[:k], v[:k], a[:k], st.as_dict()


def solve_csr(ia, ja, a, b, opts=None, device=0, x0=None):
    """PARDISO seam: solve A x = b for a zero-based int64 CSR matrix."""
    opts = opts or solve_opts()
    ia = np.ascontiguousarray(ia, np.int64)
    ja = np.ascontiguousarray(ja, np.int64)
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    x = np.zeros(len(b)) if x0 is None else np.array(x0, np.float64)
    st = SolveStats()
    _chk(lib().pdx_solve_csr(device, len(b), _p(ia), _p(ja), _p(a), _p(b), _p(x), C.byref(opts), C.byref(st)))
    return x, st.as_dict()


class GNN:
    """Graph network predicting the learned smoother of PDX_PREC_LEARNED."""

    def __init__(self, dim: int, device: int = 0, seed: int = 0):
        h = C.c_void_p()
        _chk(lib().pdx_gnn_create(dim, device, seed, C.byref(h)))
        self.h, self.dim = h, dim

    def __del__(self):
        if getattr(self, "h", None):
            lib().pdx_gnn_destroy(self.h)
            self.h = None

    def num_params(self) -> int:
        return int(lib().pdx_gnn_num_params(self.h))

    def params(self) -> np.ndarray:
        w = np.zeros(self.num_params(), np.float32)
        _chk(lib().pdx_gnn_get_params(self.h, _p(w)))
        return w

    def set_params(self, w):
        w = np.ascontiguousarray(w, np.float32)
        assert w.size == self.num_params()
        _chk(lib().pdx_gnn_set_params(self.h, _p(w)))

    def train(self, cx: "Context", steps: int, n_probes: int = 8, probe_seed: int = 0, lr: float = 1e-3,
              coarse_agg: int = 0) -> np.ndarray:
        loss = np.zeros(steps + 1)
        _chk(lib().pdx_gnn_train(cx.h, self.h, n_probes, probe_seed, coarse_agg, steps, lr, _p(loss)))
        return loss

    def loss_grad(self, cx: "Context", n_probes: int = 4, probe_seed: int = 0, coarse_agg: int = 0):
        g = np.zeros(self.num_params(), np.float32)
        loss = C.c_double()
        _chk(lib().pdx_gnn_loss_grad(cx.h, self.h, n_probes, probe_seed, coarse_agg, C.byref(loss), _p(g)))
        return loss.value, g

    def apply(self, cx: "Context"):
        _chk(lib().pdx_gnn_apply(cx.h, self.h))

    def attach(self, cx: "Context"):
        """Make this the network cx's learned Newmark runs predict their smoothers with (the
        context keeps a reference to it)."""
        _chk(lib().pdx_gnn_attach(cx.h, self.h))
        cx._gnn = self

    def infer(self, cx: "Context", bf16: bool = False):
        """Smoother [nn, d, d] predicted for cx's K (FP32, or BF16 MFMA inference) and the
        inference time in ms; it becomes cx's learned smoother."""
        m = cx.model
        M = np.zeros(m.nn * m.dim * m.dim)
        ms = C.c_double()
        _chk(lib().pdx_gnn_infer(cx.h, self.h, int(bool(bf16)), _p(M), C.byref(ms)))
        return M.reshape(m.nn, m.dim, m.dim), ms.value


def gnn_inputs(cx: "Context", n_probes: int = 4, probe_seed: int = 0, coarse_agg: int = 0) -> dict:
    """The exact network inputs the library uses for the context's current K."""
    m = cx.model
    nn, dim = m.nn, m.dim
    adj_off, adj = m.adjacency()
    nblk = int(adj_off[-1])
    X = np.zeros((nn, 8), np.float32)
    cw = np.zeros(nblk, np.float32)
    Aaa = np.zeros((nn, dim * dim), np.float32)
    B = np.zeros((nn, dim * dim), np.float32)
    f = np.zeros((n_probes, nn * dim), np.float32)
    y = np.zeros((n_probes, nn * dim), np.float32)
    _chk(lib().pdx_gnn_inputs(cx.h, n_probes, probe_seed, coarse_agg, _p(X), _p(cw), _p(Aaa), _p(B), _p(f), _p(y)))
    return dict(X=X, cw=cw, Aaa=Aaa, B=B, f=f, y=y, adj_off=adj_off, adj=adj, eq=m.nodes()["eq"][:, :dim])


class FamilyBatch:
    """Device-resident batch of PD families over a node cloud (every node a family)."""

    def __init__(self, x, vol, delta, nlf=0.33, lam=0.0, mu=0.0, dim=3, device=0):
        x = np.ascontiguousarray(x, np.float64)
        vol = np.ascontiguousarray(vol, np.float64)
        delta = np.ascontiguousarray(delta, np.float64)
        h = C.c_void_p()
        _chk(lib().pdx_fb_create(device, dim, len(vol), _p(x), _p(vol), _p(delta), nlf, lam, mu, C.byref(h)))
        self.h, self.dim = h, dim
        nf, nb, mf = C.c_longlong(), C.c_longlong(), C.c_int()
        _chk(lib().pdx_fb_sizes(h, C.byref(nf), C.byref(nb), C.byref(mf)))
        self.n_fam, self.n_bond, self.max_family = nf.value, nb.value, mf.value

    def __del__(self):
        if getattr(self, "h", None):
            lib().pdx_fb_destroy(self.h)
            self.h = None

    def families(self):
        off = np.zeros(self.n_fam + 1, np.int64)
        mem = np.zeros(max(self.n_bond, 1), np.int32)
        _chk(lib().pdx_fb_families(self.h, _p(off), _p(mem)))
        return off, mem[: self.n_bond]

    def run(self, variant=-1) -> float:
        ms = C.c_float()
        _chk(lib().pdx_fb_run(self.h, variant, C.byref(ms)))
        return ms.value

    def tuning(self):
        best = C.c_int()
        ms = np.zeros(FB_VARIANTS, np.float32)
        _chk(lib().pdx_fb_tuning(self.h, C.byref(best), _p(ms)))
        return best.value, ms

    def results(self, first, count, off=None):
        p = 9 if self.dim == 3 else 5
        dd = self.dim * self.dim
        if off is None:
            off, _ = self.families()
        nb = int(off[first + count] - off[first])
        A = np.zeros((count, p, p))
        L = np.zeros((count, p, p))
        H = np.zeros((max(nb, 1), dd), np.float32)
        _chk(lib().pdx_fb_results(self.h, first, count, _p(A), _p(L), _p(H)))
        return A, L, H[:nb]
