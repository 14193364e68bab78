"""CPU oracle for the PDLSM-FEM hot path -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this module.  It wraps oracle/pdlsm_oracle.c (a plain-C restatement of
SourceFiles/Solver/pdsolve.cpp and the element files of QibangLiu/PDLSM-FEM_mpi)
and restates the reference's mesh/command readers (datModel.cpp:58-326,
fioFile.cpp:69-462) in Python.

Sparse direct solves, which the reference hands to MKL PARDISO
(calMatrixOperations.cpp:378-498), are done with SciPy SuperLU here.

Parity status of each function is listed in the module docstring of
tests/test_oracle.py (what pins it).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

OM_ERRORS = {-1: "bad argument", -2: "singular matrix", -3: "CSR slot missing", -4: "bad mesh"}


def build(force: bool = False) -> str:
    """Compile the C restatement with gcc (no FMA contraction, like the
    reference's x86-64 -O2 build which has no FMA instructions; OpenMP only for the
    opt-in threaded explicit force of the CPU baseline)."""
    src = os.path.join(_HERE, "pdlsm_oracle.c")
    if force or not os.path.exists(_LIB_PATH) or os.path.getmtime(_LIB_PATH) < os.path.getmtime(src):
        subprocess.check_call(["gcc", "-O2", "-ffp-contract=off", "-fopenmp", "-fPIC", "-shared", "-o", _LIB_PATH,
                               src, "-lm"])
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        P, I, D = C.c_void_p, C.c_int, C.c_double
        L.om_create.restype = P
        L.om_create.argtypes = [I, I, D, D, D, I, P, I, P, P, P, I, P, P, P, P, P, P, I, P, P, P, P, P, I, P,
                                D, D, D, D, D, D, C.POINTER(I)]
        L.om_destroy.argtypes = [P]
        for name in ("om_sizes", "om_get_csr", "om_get_families", "om_get_nodes", "om_get_geom",
                     "om_lumped_mass", "om_store_u", "om_get_values", "om_newmark_UN", "om_advance_bc",
                     "om_element_M"):
            getattr(L, name).restype = None
        L.om_sizes.argtypes = [P, P]
        L.om_get_csr.argtypes = [P, P, P]
        L.om_get_families.argtypes = [P, P, P, P, P, P, P]
        L.om_get_nodes.argtypes = [P, P, P, P, P, P, P]
        L.om_get_geom.argtypes = [P, P]
        L.om_assemble.restype = I
        L.om_assemble.argtypes = [P, I, I, P, P, P, I]
        L.om_family_H.argtypes = [P, I, P]
        L.om_family_C.argtypes = [P, I, P]
        L.om_family_A.argtypes = [P, I, P]
        L.om_element_K.argtypes = [P, I, P, P]
        L.om_element_M.argtypes = [P, I, P, P]
        L.om_lumped_mass.argtypes = [P, P]
        L.om_mass_on_K.restype = I
        L.om_mass_on_K.argtypes = [P, P]
        L.om_mass_csr.restype = I
        L.om_mass_csr.argtypes = [P, P, P, P, P]
        L.om_advance_bc.argtypes = [P]
        L.om_store_u.argtypes = [P, P]
        L.om_get_values.argtypes = [P, P]
        L.om_explicit_run.restype = I
        L.om_explicit_run.argtypes = [P, I, P, P]
        L.om_newmark_UN.argtypes = [P, P, P, P]
        L.om_set_skip.restype = None
        L.om_set_skip.argtypes = [I]
        L.om_node_stresses.restype = I
        L.om_node_stresses.argtypes = [P, P, P]
        L.om_fb_members.restype = I
        L.om_fb_members.argtypes = [I, I, P, P, P, P, I]
        L.om_fb_family.restype = I
        L.om_fb_family.argtypes = [I, P, P, P, D, D, D, I, P, P, P]
        for name in ("om_set_values", "om_set_bonds", "om_advance_bc_ratio", "om_local_damage"):
            getattr(L, name).restype = None
        L.om_set_values.argtypes = [P, P]
        L.om_set_bonds.argtypes = [P, P]
        L.om_advance_bc_ratio.argtypes = [P, D]
        L.om_local_damage.argtypes = [P, P]
        L.om_fc2_step.restype = I
        L.om_fc2_step.argtypes = [P, P, D, I, P]
        L.om_set_threads.restype = None
        L.om_set_threads.argtypes = [I]
        L.om_greedy_colors.restype = I
        L.om_greedy_colors.argtypes = [I, P, P, P]
        L.om_bilu0.restype = I
        L.om_bilu0.argtypes = [I, I, P, P, P, P, P, P]
        L.om_bilu0_apply.restype = None
        L.om_bilu0_apply.argtypes = [I, I, P, P, P, P, P, P, P]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def _chk(rc):
    if rc != 0:
        raise RuntimeError(f"oracle error {rc}: {OM_ERRORS.get(rc, '?')}")


class Model:
    """A set-up reference model (pdsolve::pdsolve + setDatModel, pdsolve.cpp:20-151)."""

    def __init__(self, mesh: dict):
        L = lib()
        self.mesh = mesh
        self.dim = int(mesh["dim"])
        self.mesh_sigma_ult = float(mesh.get("sigma_ult", 0.0))
        self._keep = []

        def arr(x, dt):
            a = np.ascontiguousarray(np.asarray(x, dtype=dt))
            self._keep.append(a)
            return a

        x = arr(mesh["x"], np.float64).reshape(-1, 3)
        econ = arr(mesh["econ"], np.int32).reshape(-1, 8)
        nen = arr(mesh["nen"], np.int32)
        algo = arr(mesh["algo"], np.int32)
        ebc = mesh.get("ebc", [])
        ebc_off = arr(np.concatenate([[0], np.cumsum([len(s["nid"]) for s in ebc])]).astype(np.int32), np.int32)
        ebc_nid = arr(np.concatenate([np.asarray(s["nid"]) for s in ebc]) if ebc else np.zeros(0), np.int32)
        ebc_dof = arr([s["dof"] for s in ebc], np.int32)
        ebc_val = arr([s["val"] for s in ebc], np.float64)
        ebc_vel = arr([s.get("vel", 0.0) for s in ebc], np.float64)
        ebc_vary = arr([int(s.get("vary", False)) for s in ebc], np.int32)
        nbc = mesh.get("nbc", [])
        tn = 4 if self.dim == 3 else 2
        nbc_off = arr(np.concatenate([[0], np.cumsum([len(s["con"]) for s in nbc])]).astype(np.int32), np.int32)
        nbc_con = arr(np.concatenate([np.asarray(s["con"]).reshape(-1, tn) for s in nbc]).ravel()
                      if nbc else np.zeros(0), np.int32)
        nbc_val = arr([s["val"] for s in nbc], np.float64)
        nbc_vel = arr([s.get("vel", 0.0) for s in nbc], np.float64)
        nbc_vary = arr([int(s.get("vary", False)) for s in nbc], np.int32)
        cracks = arr(np.asarray(mesh.get("cracks", np.zeros((0, 3, 3)))).reshape(-1, 9), np.float64)
        err = C.c_int(0)
        self.h = L.om_create(self.dim, int(mesh.get("proType", 1)), float(mesh["E"]), float(mesh["nu"]),
                             float(mesh["rho"]), len(x), _p(x), len(econ), _p(econ), _p(nen), _p(algo),
                             len(ebc), _p(ebc_off), _p(ebc_nid), _p(ebc_dof), _p(ebc_val), _p(ebc_vel),
                             _p(ebc_vary), len(nbc), _p(nbc_off), _p(nbc_con), _p(nbc_val), _p(nbc_vel),
                             _p(nbc_vary), len(cracks), _p(cracks), float(mesh["fac"]), float(mesh["nlf"]),
                             float(mesh.get("me", 0.0)), float(mesh.get("dt", 1.0)),
                             float(mesh.get("beta", 0.25)), float(mesh.get("gamma", 0.5)), C.byref(err))
        if not self.h:
            _chk(err.value or -1)
        s = np.zeros(10, np.int64)
        L.om_sizes(self.h, _p(s))
        (self.neq, self.nnz, self.nfam, self.nmem, self.npdbe, self.npde, self.nfe) = (int(v) for v in s[:7])
        self.nblocks = tuple(int(v) for v in s[7:10])
        self.nn = len(x)
        self.ne = len(econ)

    def __del__(self):
        if getattr(self, "h", None):
            lib().om_destroy(self.h)
            self.h = None

    # ---- setup outputs ---------------------------------------------------
    def csr(self):
        ia = np.zeros(self.neq + 1, np.int64)
        ja = np.zeros(self.nnz, np.int64)
        lib().om_get_csr(self.h, _p(ia), _p(ja))
        return ia, ja

    def families(self):
        fn = np.zeros(self.nfam, np.int32)
        fo = np.zeros(self.nfam + 1, np.int32)
        fm = np.zeros(self.nmem, np.int32)
        bd = np.zeros(self.nmem, np.int32)
        fv = np.zeros(self.nfam, np.float64)
        fd = np.zeros(self.nfam, np.float64)
        lib().om_get_families(self.h, _p(fn), _p(fo), _p(fm), _p(bd), _p(fv), _p(fd))
        return dict(node=fn, off=fo, mem=fm, bond=bd, cd_v=fv, delta=fd)

    def nodes(self):
        eq = np.zeros(self.nn * 3, np.int32)
        nt = np.zeros(self.nn, np.int32)
        nv = np.zeros(self.nn, np.float64)
        ev = np.zeros(self.ne, np.float64)
        pb = np.zeros(max(self.npdbe, 1) * 4, np.int32)
        al = np.zeros(self.ne, np.int32)
        lib().om_get_nodes(self.h, _p(eq), _p(nt), _p(nv), _p(ev), _p(pb), _p(al))
        return dict(eq=eq.reshape(-1, 3), ntype=nt, nv=nv, ev=ev, pdbe=pb[: self.npdbe * 4].reshape(-1, 4), algo=al)

    def geom(self):
        g = np.zeros(11)
        lib().om_get_geom(self.h, _p(g))
        return dict(minDelta=g[0], maxDelta=g[1], bsize=g[2], lbc=g[3:6].copy(), rtc=g[6:9].copy(),
                    lam=g[9], mu=g[10])

    # ---- kernels ---------------------------------------------------------
    def family_H(self, f):
        fam = self.families()
        n = fam["off"][f + 1] - fam["off"][f]
        H = np.zeros((self.dim, self.dim * n))
        _chk(lib().om_family_H(self.h, f, _p(H)))
        return H

    def family_C(self, f):
        fam = self.families()
        n = fam["off"][f + 1] - fam["off"][f]
        nv = 3 if self.dim == 2 else 6
        Cm = np.zeros((nv, self.dim * n))
        _chk(lib().om_family_C(self.h, f, _p(Cm)))
        return Cm

    def family_A(self, f):
        p = 5 if self.dim == 2 else 9
        A = np.zeros((p, p))
        _chk(lib().om_family_A(self.h, f, _p(A)))
        return A

    def node_stresses(self, u_all):
        """calGlobalNodeStresses for DOF values u_all [nn*dim] -> sigma [nn, 6]."""
        u = np.ascontiguousarray(u_all, np.float64)
        sig = np.zeros(self.nn * 6)
        _chk(lib().om_node_stresses(self.h, _p(u), _p(sig)))
        return sig.reshape(self.nn, 6)

    def element_K(self, e):
        out = np.zeros(576)
        nd = C.c_int(0)
        _chk(lib().om_element_K(self.h, e, _p(out), C.byref(nd)))
        return out[: nd.value * nd.value].reshape(nd.value, nd.value)

    def element_M(self, e):
        out = np.zeros(576)
        nd = C.c_int(0)
        lib().om_element_M(self.h, e, _p(out), C.byref(nd))
        return out[: nd.value * nd.value].reshape(nd.value, nd.value)

    def assemble(self, mode=1, newmark=False, UN=None, with_ext=True):
        """K (CSR values on self.csr()) and F; mode 1 static, 2 quasi-static, 0 dynamic."""
        F = np.zeros(self.neq)
        K = None if (mode == 0 and not newmark) else np.zeros(self.nnz)
        un = None if UN is None else np.ascontiguousarray(UN, np.float64)
        _chk(lib().om_assemble(self.h, mode, int(newmark), _p(un), _p(K), _p(F), int(with_ext)))
        return K, F

    def lumped_mass(self):
        m = np.zeros(self.neq)
        lib().om_lumped_mass(self.h, _p(m))
        return m

    def mass_on_K(self):
        Mv = np.zeros(self.nnz)
        _chk(lib().om_mass_on_K(self.h, _p(Mv)))
        return Mv

    def mass_csr(self):
        """(ia, ja, values) of the explicit driver's consistent mass on its own pattern
        (setCSRIndexes_gloMassMat pdsolve.cpp:5674-5817, assembleElemassMatFEM_CSRformat
        :5154-5210, assembleMassMatPD_CSRformat :4568-4595)."""
        nz = C.c_longlong(0)
        _chk(lib().om_mass_csr(self.h, None, None, None, C.byref(nz)))
        ia = np.zeros(self.neq + 1, np.int64)
        ja = np.zeros(max(nz.value, 1), np.int64)
        vals = np.zeros(max(nz.value, 1))
        _chk(lib().om_mass_csr(self.h, _p(ia), _p(ja), _p(vals), C.byref(nz)))
        return ia, ja[: nz.value], vals[: nz.value]

    def advance_bc(self):
        lib().om_advance_bc(self.h)

    def store_u(self, U):
        lib().om_store_u(self.h, _p(np.ascontiguousarray(U, np.float64)))

    def values(self):
        out = np.zeros(self.nn * self.dim)
        lib().om_get_values(self.h, _p(out))
        return out

    def set_values(self, u_all):
        lib().om_set_values(self.h, _p(np.ascontiguousarray(u_all, np.float64)))

    def set_bonds(self, bond):
        lib().om_set_bonds(self.h, _p(np.ascontiguousarray(bond, np.int32)))

    def advance_bc_ratio(self, rati):
        lib().om_advance_bc_ratio(self.h, float(rati))

    def fc2_step(self, sigma, sigma_ult, topk):
        """(bonds kept, bonds broken, largest principal bond stress) of one FC 2 update."""
        out = np.zeros(3)
        _chk(lib().om_fc2_step(self.h, _p(np.ascontiguousarray(sigma, np.float64)), float(sigma_ult), int(topk),
                               _p(out)))
        return int(out[0]), int(out[1]), float(out[2])

    def local_damage(self):
        phi = np.zeros(self.nn)
        lib().om_local_damage(self.h, _p(phi))
        return phi

    def quasi_static_run(self, nsteps, fc=2, topk=0, sigma_ult=None):
        """pdfemQuasiStaticSolver_CSRformat (pdsolve.cpp:6464-6580) with failureProcess
        (:6597-6641), single rank: per step K_ff du = F_ext - K u (quasi-static assembly,
        SuperLU in place of PARDISO), u += du, nodal stresses, the FC 2 bond update (2D),
        then the BCs advance with the load ratio unless a bond was selected.  Returns the DOF
        values [nn*dim] and per step (bonds kept, bonds broken, ratio, largest s1)."""
        import scipy.sparse.linalg as spla
        sigma_ult = self.mesh_sigma_ult if sigma_ult is None else sigma_ult
        info = np.zeros((nsteps, 4))
        add_load = True
        for n in range(nsteps):
            K, F = self.assemble(2)
            du = spla.spsolve(self.sparse(K).tocsc(), F)
            u = self.values()
            eq = self.nodes()["eq"][:, : self.dim].ravel()
            u[eq >= 0] += du[eq[eq >= 0]]
            self.set_values(u)
            rati, kept, broke, smax = 1.0, 0, 0, 0.0
            if fc == 2 and self.dim == 2:
                sig = self.node_stresses(u)
                kept, broke, smax = self.fc2_step(sig, sigma_ult, topk)
                add_load = kept == 0
                rati = 1.0 if abs(smax) < 1e-20 * sigma_ult else 1.001 * sigma_ult / smax
            info[n] = (kept, broke, rati, smax)
            if add_load:
                self.advance_bc_ratio(rati)
        return self.values(), info

    def explicit_run(self, nsteps):
        """u_{nsteps+1} and u_{nsteps} of pdfemDynamicSolver_CSRformat with lumped mass."""
        u = np.zeros(self.neq)
        up = np.zeros(self.neq)
        _chk(lib().om_explicit_run(self.h, nsteps, _p(u), _p(up)))
        return u, up

    def _fc2_after_store(self, fc, topk, sigma_ult):
        """calGlobalNodeStresses + failureProcess as the dynamic drivers call them after storing
        the step's displacement (pdsolve.cpp:6060-6064, :6258-6261): FC 2 in 2D only (3D
        returns at once, :6594-6597).  Returns (bonds kept, bonds broken, largest s1)."""
        if fc != 2 or self.dim != 2:
            return 0, 0, 0.0
        sig = self.node_stresses(self.values())
        return self.fc2_step(sig, sigma_ult, topk)

    def dynamic_run(self, nsteps, lumped=None, fc=None, topk=None, sigma_ult=None):
        """pdfemDynamicSolver_CSRformat (pdsolve.cpp:5842-6090), single rank: the start-up step
        u1 = u0 + dt^2/2 a0 (:5979-6017) and nsteps loop steps, each advancing the BCs before the
        force (setVaryEssentialBC / setVaryNaturalBC), then
          lumped mass:     u_{n+1} = 2 u_n + dt^2 F/m - u_{n-1}    (timeIntegration_LM :5113-5132)
          consistent mass: M a = F (PARDISO in the reference, SuperLU here), then
                           u_{n+1} = 2 u_n + dt^2 a - u_{n-1}       (timeIntegration_NLM :5134-5152)
        and after every stored step the nodal stresses and the FC 2 bond update (2D).
        lumped / fc / topk / sigma_ult default to the model's LUMPMASS, FC, TOPK and strength.
        Returns u_{nsteps+1}, u_{nsteps} (free DOFs) and per step (kept, broken, largest s1)."""
        import scipy.sparse as sp
        import scipy.sparse.linalg as spla
        lumped = self.mesh.get("lumped_mass", False) if lumped is None else lumped
        fc = self.mesh.get("fc", 0) if fc is None else fc
        topk = self.mesh.get("topk", 0) if topk is None else topk
        sigma_ult = self.mesh_sigma_ult if sigma_ult is None else sigma_ult
        dt = self.mesh.get("dt", 1.0)
        eq = self.nodes()["eq"][:, : self.dim].ravel()
        vals = self.values()
        un = np.zeros(self.neq)
        un[eq[eq >= 0]] = vals[eq >= 0]
        if lumped:
            m = self.lumped_mass()
            accel = lambda F: F / m
        else:
            ia, ja, mv = self.mass_csr()
            lu = spla.splu(sp.csr_matrix((mv, ja, ia), shape=(self.neq, self.neq)).tocsc())
            accel = lu.solve
        info = np.zeros((nsteps + 1, 3))
        self.advance_bc()
        _, F = self.assemble(0)
        if lumped:
            unp1 = F / m * (0.5 * dt * dt) + un
        else:
            unp1 = accel(F) * (0.5 * dt * dt) + un
        self.store_u(unp1)
        info[0] = self._fc2_after_store(fc, topk, sigma_ult)
        unm1 = un
        for n in range(1, nsteps + 1):
            unm1, un = un, unp1
            self.advance_bc()
            _, F = self.assemble(0)
            if lumped:
                tA = dt * dt * F / m
            else:
                tA = dt * dt * accel(F)
            unp1 = un * 2.0 + tA - unm1
            self.store_u(unp1)
            info[n] = self._fc2_after_store(fc, topk, sigma_ult)
        return unp1, un, info

    def newmark_UN(self, v, a):
        UN = np.zeros(self.nn * self.dim)
        lib().om_newmark_UN(self.h, _p(np.ascontiguousarray(v)), _p(np.ascontiguousarray(a)), _p(UN))
        return UN

    def sparse(self, vals):
        import scipy.sparse as sp
        ia, ja = self.csr()
        return sp.csr_matrix((vals, ja, ia), shape=(self.neq, self.neq))


def set_skip(flags: int):
    """Test hook: 1 drops PD-boundary terms, 2 FE terms, 4 PD terms (planted mistakes)."""
    lib().om_set_skip(int(flags))


def static_solve(model: Model):
    """pdfemStaticSolver_CSRformat (pdsolve.cpp:5557-5654) with SuperLU in place of PARDISO."""
    import scipy.sparse.linalg as spla
    K, F = model.assemble(1)
    A = model.sparse(K).tocsc()
    u = spla.spsolve(A, F)
    return u, K, F


def newmark_run(model: Model, nsteps: int, fc=0, topk=0, sigma_ult=None, info=None):
    """pdfemDynamicNewmarkSolver_CSRformat (pdsolve.cpp:6092-6290) with SuperLU.
    Returns the free-DOF displacement, velocity and acceleration after nsteps.  With fc == 2
    (2D) every step ends with the nodal stresses and the FC 2 bond update at the advanced BC
    values (:6250-6261); info (or None) receives per step (kept, broken, largest s1)."""
    import scipy.sparse.linalg as spla
    neq, dt, beta, gamma = model.neq, model.mesh.get("dt", 1.0), model.mesh.get("beta", 0.25), model.mesh.get("gamma", 0.5)
    ia, ja = model.csr()
    eq = model.nodes()["eq"][:, : model.dim].ravel()
    vals = model.values()
    u = np.zeros(neq)
    u[eq[eq >= 0]] = vals[eq >= 0]
    v = np.zeros(neq)
    a = np.zeros(neq)
    Mv = model.mass_on_K()
    for n in range(nsteps):
        UN = model.newmark_UN(v, a)
        K, F = model.assemble(0, newmark=True, UN=UN)
        Keff = K * beta * dt * dt + Mv
        anp1 = spla.spsolve(model.sparse(Keff).tocsc(), F)
        t1 = u + dt * v + dt * dt * ((0.5 - beta) * a + beta * anp1)
        t2 = v + dt * ((1 - gamma) * a + gamma * anp1)
        u, v = t1, t2
        model.store_u(u)
        model.advance_bc()
        a = anp1
        r = model._fc2_after_store(fc, topk, model.mesh_sigma_ult if sigma_ult is None else sigma_ult)
        if info is not None:
            info.append(r)
    return u, v, a


# ---------------------------------------------------------------------------
# mesh + command file readers (datModel::readdata datModel.cpp:58-326,
# fioFiles::excuteCMD fioFile.cpp:69-462)
# ---------------------------------------------------------------------------

def read_lis(path: str) -> dict:
    with open(path) as f:
        text = f.read()
    lines = text.split("\n")
    title = lines[0].split()[0] if lines[0].split() else ""
    toks = " ".join(lines[2:]).split()
    pos = 0

    def nxt():
        nonlocal pos
        t = toks[pos]
        pos += 1
        return t

    sdim = nxt()
    pro = int(nxt())
    E, nu, rho = float(nxt()), float(nxt()), float(nxt())
    kic, sigma_ult = float(nxt()), float(nxt())
    dim = 3 if sdim == "3D" else 2
    nn, ne = int(nxt()), int(nxt())
    x = np.zeros((nn, 3))
    for i in range(nn):
        nxt()
        x[i] = [float(nxt()), float(nxt()), float(nxt())]
    econ = np.zeros((ne, 8), np.int32)
    nen = np.zeros(ne, np.int32)
    algo = np.zeros(ne, np.int32)
    for i in range(ne):
        nxt()
        algo[i] = int(nxt())
        if dim == 3:
            c = [int(nxt()) for _ in range(8)]
            if c[2] != c[3] and c[4] != c[5]:
                econ[i] = c
                nen[i] = 8
            else:
                econ[i, :4] = [c[0], c[1], c[2], c[4]]
                nen[i] = 4
        else:
            c = [int(nxt()) for _ in range(4)]
            econ[i, :4] = c
            nen[i] = 4 if c[3] != c[2] else 3
    # the reader skips one header line before each section (getline x2 after a token)
    rest = _sections_after_elements(lines)
    return dict(title=title, dim=dim, proType=pro, E=E, nu=nu, rho=rho, x=x, econ=econ, nen=nen, algo=algo,
                kic=kic, sigma_ult=sigma_ult, **rest)


def _sections_after_elements(lines):
    # locate section headers "====" and parse the following integers/floats
    idx = [i for i, l in enumerate(lines) if l.strip().startswith("=")]
    secs = []
    for k, i in enumerate(idx):
        j = idx[k + 1] if k + 1 < len(idx) else len(lines)
        secs.append(" ".join(lines[i + 1: j]).split())
    ebc_t, nbc_t, nof_t, crk_t = secs[0], secs[1], secs[2], secs[3]
    return dict(_ebc_tokens=ebc_t, _nbc_tokens=nbc_t, _crack_tokens=crk_t)


def _finish_lis(m: dict) -> dict:
    dim = m["dim"]
    t = m.pop("_ebc_tokens")
    n = int(t[0])
    sets, p = [], 1
    for _ in range(n):
        sets.append(dict(count=int(t[p]), dof={"UX": 0, "UY": 1, "UZ": 2}[t[p + 1]], val=float(t[p + 2])))
        p += 3
    for s in sets:
        s["nid"] = np.array([int(v) for v in t[p: p + s["count"]]], np.int32)
        p += s["count"]
        del s["count"]
    m["ebc"] = sets
    t = m.pop("_nbc_tokens")
    n = int(t[0])
    tn = 4 if dim == 3 else 2
    nsets, p = [], 1
    for _ in range(n):
        nsets.append(dict(count=int(t[p]), val=float(t[p + 1])))
        p += 2
    for s in nsets:
        s["con"] = np.array([int(v) for v in t[p: p + s["count"] * tn]], np.int32).reshape(-1, tn)
        p += s["count"] * tn
        del s["count"]
    m["nbc"] = nsets
    t = m.pop("_crack_tokens")
    n = int(t[0])
    npts = 3 if dim == 3 else 2
    cr = np.zeros((n, 3, 3))
    p = 1
    for i in range(n):
        for j in range(npts):
            cr[i, j] = [float(v) for v in t[p: p + 3]]
            p += 3
    m["cracks"] = cr
    return m


def read_case(cmd_path: str) -> dict:
    """Read a reference command file (.in) and the mesh it names."""
    base = os.path.dirname(os.path.abspath(cmd_path))
    m = None
    opts = dict(fac=3.0, nlf=0.33, me=0.0, dt=1.0, nsteps=1, save=1, gamma=0.5, beta=0.25, solver=-1,
                newmark=False, lumped=False, fc=0, topk=0, vtk_binary=True, rf=[])
    pend = []
    with open(cmd_path) as f:
        for line in f:
            tk = line.split()
            if not tk or tk[0].startswith("#"):
                continue
            if tk[0] == "MSHFILE":
                m = _finish_lis(read_lis(os.path.join(base, tk[1])))
            else:
                pend.append(tk)
    for tk in pend:
        k = tk[0]
        if k == "SOLVER":
            opts["solver"] = {"DYNAMIC": 0, "STATIC": 1, "QUASI-STATIC": 2}[tk[1]]
        elif k == "SETSOLVING" and len(tk) >= 6:
            opts["dt"] = 1.0 if opts["solver"] == 2 else float(tk[1])
            opts["nsteps"], opts["save"] = int(tk[2]), int(tk[3])
            opts["fac"], opts["nlf"] = float(tk[4]), float(tk[5])
            if len(tk) > 6:
                opts["me"] = float(tk[6])
        elif k == "NEWMARK" and len(tk) > 1:
            if tk[1] in ("ON", "OFF"):  # anything else is ignored (fioFile.cpp:151-181)
                opts["newmark"] = tk[1] == "ON"
            if len(tk) > 3:
                opts["gamma"], opts["beta"] = float(tk[2]), float(tk[3])
        elif k == "LUMPMASS" and len(tk) > 1:
            if tk[1] in ("ON", "OFF"):  # fioFile.cpp:208-230
                opts["lumped"] = tk[1] == "ON"
        elif k == "VTKFORMAT" and len(tk) > 1:
            if tk[1] in ("ASCII", "BINARY"):  # fioFile.cpp:422-444
                opts["vtk_binary"] = tk[1] == "BINARY"
        elif k == "RF" and len(tk) > 1:
            opts["rf"] += [int(v) for v in tk[1:]]
        elif k == "TOPK" and len(tk) > 1:
            opts["topk"] = int(tk[1])
        elif k == "EBC":
            m["ebc"][int(tk[1])]["val"] = float(tk[2])
        elif k == "VEBC":
            m["ebc"][int(tk[1])]["vary"] = True
            m["ebc"][int(tk[1])]["vel"] = float(tk[2])
        elif k == "NBC":
            m["nbc"][int(tk[1])]["val"] = float(tk[2])
        elif k == "VNBC":
            m["nbc"][int(tk[1])]["vary"] = True
            m["nbc"][int(tk[1])]["vel"] = float(tk[2])
        elif k == "FC":
            opts["fc"] = int(tk[1])
            if opts["fc"] == 2 and len(tk) > 2:
                opts["topk"] = int(tk[2])
    m.update(fac=opts["fac"], nlf=opts["nlf"], me=opts["me"], dt=opts["dt"], beta=opts["beta"],
             gamma=opts["gamma"], lumped_mass=opts["lumped"], newmark=opts["newmark"], fc=opts["fc"],
             topk=opts["topk"], save_freq=opts["save"], vtk_binary=opts["vtk_binary"])
    m["opts"] = opts
    return m


# ---- PD family batch ------------------------------------------------------
def fb_members(x, delta, cap=None):
    """Brute-force members of every node's family (initialPDN_Fami rule)."""
    x = np.ascontiguousarray(x, np.float64)
    delta = np.ascontiguousarray(delta, np.float64)
    n = len(delta)
    cap = cap or n * 200
    off = np.zeros(n + 1, np.int32)
    mem = np.zeros(cap, np.int32)
    c = lib().om_fb_members(3, n, _p(x), _p(delta), _p(off), _p(mem), cap)
    if c < 0:
        raise RuntimeError("member capacity exceeded")
    return off, mem[:c]


def fb_family(dim, x, vol, delta, nlf, lam, mu, mem):
    """Shape tensor A and H blocks [n, d*d] (self block first) of one family."""
    x = np.ascontiguousarray(x, np.float64)
    vol = np.ascontiguousarray(vol, np.float64)
    delta = np.ascontiguousarray(delta, np.float64)
    mem = np.ascontiguousarray(mem, np.int32)
    p = 5 if dim == 2 else 9
    A = np.zeros(p * p)
    H = np.zeros(len(mem) * dim * dim)
    _chk(lib().om_fb_family(dim, _p(x), _p(vol), _p(delta), nlf, lam, mu, len(mem), _p(mem), _p(A), _p(H)))
    return A.reshape(p, p), H.reshape(len(mem), dim * dim)


# ---- block ILU(0) (textbook restatement; the reference solves with PARDISO) ----
def greedy_colors(off, adj):
    """Greedy colouring of the node graph in node order (om_greedy_colors)."""
    off = np.ascontiguousarray(off, np.int32)
    adj = np.ascontiguousarray(adj, np.int32)
    col = np.zeros(len(off) - 1, np.int32)
    nc = lib().om_greedy_colors(len(off) - 1, _p(off), _p(adj), _p(col))
    return col, nc


def colour_rank(col):
    """Elimination rank: by colour, node order within a colour."""
    order = np.lexsort((np.arange(len(col)), col))
    rank = np.zeros(len(col), np.int32)
    rank[order] = np.arange(len(col), dtype=np.int32)
    return rank


def bilu0(off, adj, Kb, rank, d):
    """Block ILU(0) (om_bilu0): F [blocks, d, d] (strict lower L, rest U), Dinv [nn, d, d]."""
    off = np.ascontiguousarray(off, np.int32)
    adj = np.ascontiguousarray(adj, np.int32)
    Kb = np.ascontiguousarray(Kb, np.float64).reshape(-1)
    rank = np.ascontiguousarray(rank, np.int32)
    nn = len(off) - 1
    F = np.zeros_like(Kb)
    Dinv = np.zeros(nn * d * d)
    _chk(lib().om_bilu0(nn, d, _p(off), _p(adj), _p(Kb), _p(rank), _p(F), _p(Dinv)))
    return F.reshape(-1, d, d), Dinv.reshape(nn, d, d)


def bilu0_apply(off, adj, F, Dinv, rank, r, d):
    """z = (L U)^-1 r with the factors of bilu0 (om_bilu0_apply)."""
    off = np.ascontiguousarray(off, np.int32)
    adj = np.ascontiguousarray(adj, np.int32)
    nn = len(off) - 1
    z = np.zeros(nn * d)
    lib().om_bilu0_apply(nn, d, _p(off), _p(adj), _p(np.ascontiguousarray(F, np.float64).reshape(-1)),
                         _p(np.ascontiguousarray(Dinv, np.float64).reshape(-1)),
                         _p(np.ascontiguousarray(rank, np.int32)), _p(np.ascontiguousarray(r, np.float64)), _p(z))
    return z


def masked_blocks(off, adj, eq, d, ia, ja, K):
    """Node-block form [blocks, d, d] of a free-DOF CSR matrix (ia, ja, K) on the node adjacency
    (off, adj); rows and columns of prescribed DOFs (eq < 0) zero with 1 on their diagonal."""
    off = np.asarray(off, np.int64)
    adj = np.asarray(adj, np.int64)
    eq = np.asarray(eq)[:, :d]
    nn = len(off) - 1
    node = np.full(int(eq.max()) + 1, -1, np.int64)
    comp = np.full(int(eq.max()) + 1, -1, np.int64)
    a_idx, c_idx = np.nonzero(eq >= 0)
    node[eq[a_idx, c_idx]] = a_idx
    comp[eq[a_idx, c_idx]] = c_idx
    rows = np.repeat(np.arange(len(ia) - 1), np.diff(ia))
    keys = np.repeat(np.arange(nn, dtype=np.int64), np.diff(off)) * nn + adj
    ra, cb = node[rows], node[np.asarray(ja)]
    slot = np.searchsorted(keys, ra * nn + cb)
    assert np.array_equal(keys[slot], ra * nn + cb)
    Kb = np.zeros((len(adj), d, d))
    np.add.at(Kb, (slot, comp[rows], comp[np.asarray(ja)]), np.asarray(K))
    diag = np.searchsorted(keys, np.arange(nn, dtype=np.int64) * (nn + 1))
    pa, pc = np.nonzero(eq < 0)
    Kb[diag[pa], pc, pc] = 1.0
    return Kb
