"""Host side of the library (no GPU): the set-up that feeds the kernels
(DOF numbering, volumes, PD/FE split, boundary entities, families, bond
states, CSR pattern) against the oracle, the .lis/.in reader, error codes, and
the exported C-ABI."""
import ctypes
import os
import re

import numpy as np
import pytest

from helpers import tri_tet_mesh, ROOT, box_fixture, load_oracle, load_pdx, workloads

O = load_oracle()
W = workloads()
pdx = load_pdx()


def _compare(mesh):
    mo, mp = O.Model(mesh), pdx.Model(mesh)
    d = mo.dim
    ia0, ja0 = mo.csr()
    ia1, ja1 = mp.csr()
    assert (mo.neq, mo.nnz, mo.nfam) == (mp.neq, mp.nnz, mp.nfam)
    assert np.array_equal(ia0, ia1) and np.array_equal(ja0, ja1)
    fo, fp = mo.families(), mp.families()
    for k in ("node", "off", "mem", "bond"):
        assert np.array_equal(fo[k], fp[k]), k
    for k in ("cd_v", "delta"):
        assert np.array_equal(fo[k], fp[k]), k  # bit-exact
    no, npd = mo.nodes(), mp.nodes()
    assert np.array_equal(no["eq"][:, :d], npd["eq"][:, :d])
    for k in ("ntype", "nv", "ev", "algo"):
        assert np.array_equal(no[k], npd[k]), k
    w = 2 if d == 2 else 4
    assert np.array_equal(no["pdbe"][:, :w], npd["pdbe"][:, :w])
    return mo, mp


@pytest.mark.parametrize("name", ["plate2d", "crack2d", "box3d", "penny3d", "tri_tet"])
def test_setup_matches_oracle(name):
    if name == "plate2d":
        mesh = W.plate2d()
    elif name == "crack2d":
        mesh = W.crack2d(seed=5, nx=121, ny=61, dx=1e-3)
    elif name == "box3d":
        mesh = W.box3d(n=18, pd_block=8)
    elif name == "penny3d":
        mesh = W.penny3d(n=22, pd_block=10)
    else:
        mesh = tri_tet_mesh()
    mo, mp = _compare(mesh)
    if name == "crack2d":
        assert (mo.families()["bond"] == 0).sum() > 0
    if name == "penny3d":
        assert (mp.families()["bond"] == 0).sum() > 0


def test_box_example_setup_matches_oracle():
    mesh, _ = box_fixture()
    mo, mp = _compare(mesh)
    assert mp.neq == 102500


def write_lis(path, mesh, title="t"):
    """Write a mesh dict in the reference .lis layout (datModel.cpp:58-326)."""
    dim = mesh["dim"]
    with open(path, "w") as f:
        f.write(f"{title}\n***generated***\n{dim}D\t{mesh['proType']}\n")
        f.write(f"{mesh['E']:.17g}\t{mesh['nu']:.17g}\t{mesh['rho']:.17g}\t1e6\t1e6\n")
        f.write(f"{len(mesh['x'])}\t{len(mesh['econ'])}\n")
        for i, p in enumerate(mesh["x"]):
            f.write(f"{i + 1}\t{p[0]:.17g}\t{p[1]:.17g}\t{p[2]:.17g}\n")
        for e, c in enumerate(mesh["econ"]):
            if dim == 2:
                cc = list(c[:4]) if mesh["nen"][e] == 4 else [c[0], c[1], c[2], c[2]]
            else:
                cc = list(c[:8]) if mesh["nen"][e] == 8 else [c[0], c[1], c[2], c[2], c[3], c[3], c[3], c[3]]
            f.write(f"{e + 1}\t{mesh['algo'][e]}\t" + " ".join(str(int(v)) for v in cc) + "\n")
        f.write("=====Essential BCs=====\n")
        f.write(f"{len(mesh['ebc'])}\n")
        for s in mesh["ebc"]:
            f.write(f"{len(s['nid'])}\t{['UX', 'UY', 'UZ'][s['dof']]}\t{s['val']:.17g}\n")
        for s in mesh["ebc"]:
            f.write(" ".join(str(int(v)) for v in s["nid"]) + "\n")
        f.write("=====Natural BCs=====\n")
        f.write(f"{len(mesh['nbc'])}\n")
        for s in mesh["nbc"]:
            f.write(f"{len(s['con'])}\t{s['val']:.17g}\n")
        for s in mesh["nbc"]:
            for c in s["con"]:
                f.write(" ".join(str(int(v)) for v in c) + "\n")
        f.write("=====NO FAIL region=====\n0\n=====pre-exist crack=====\n")
        cr = mesh.get("cracks", np.zeros((0, 3, 3)))
        f.write(f"{len(cr)}\n")
        for c in cr:
            for j in range(3 if dim == 3 else 2):
                f.write(" ".join(f"{v:.17g}" for v in c[j]) + "\n")


@pytest.mark.parametrize("name", ["box3d", "crack2d"])
def test_lis_reader_matches_arrays(tmp_path, name):
    mesh = W.box3d(n=10, pd_block=4) if name == "box3d" else W.crack2d(seed=2, nx=61, ny=31, dx=1e-3)
    write_lis(tmp_path / "m.lis", mesh)
    (tmp_path / "c.in").write_text(f"MSHFILE m.lis\nSOLVER STATIC\nSETSOLVING 1 3 1 {mesh['fac']!r} {mesh['nlf']!r}\n")
    a, b = pdx.Model(mesh), pdx.Model(cmd_file=str(tmp_path / "c.in"))
    assert b.solver_flag == 1 and b.num_steps == 3
    assert np.array_equal(a.csr()[1], b.csr()[1])
    fa, fb = a.families(), b.families()
    assert np.array_equal(fa["mem"], fb["mem"]) and np.array_equal(fa["bond"], fb["bond"])
    # the oracle's Python reader agrees too
    mo = O.Model(O.read_case(str(tmp_path / "c.in")))
    assert np.array_equal(mo.csr()[1], b.csr()[1])


def test_error_codes():
    mesh = W.box3d(n=6, pd_block=2)
    mesh["ebc"] = mesh["ebc"] + [dict(nid=mesh["ebc"][0]["nid"][:1], dof=2, val=0.0)]
    with pytest.raises(pdx.PdxError, match="repeatly defined"):
        pdx.Model(mesh)
    mesh = W.box3d(n=6, pd_block=2)
    mesh["econ"] = mesh["econ"].copy()
    mesh["econ"][0, 0] = 10 ** 6
    with pytest.raises(pdx.PdxError, match="out of range"):
        pdx.Model(mesh)
    with pytest.raises(pdx.PdxError, match="is not exist"):
        pdx.Model(cmd_file="/nonexistent/x.in")


def test_c_abi_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "pdx.h")).read()
    names = sorted(set(re.findall(r"\b(pdx_[a-z0-9_]+)\s*\(", hdr)))
    assert len(names) > 20
    L = ctypes.CDLL(pdx.LIB_PATH)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert pdx.lib().pdx_version().decode().startswith("pdx")


def test_context_entry_points_reject_a_null_context():
    """Every context entry point returns PDX_ERR_ARG for a null context (before any device
    call) and leaves a message in pdx_last_error."""
    L = pdx.lib()
    null = None
    buf = (ctypes.c_double * 8)()
    ibuf = (ctypes.c_int * 8)()
    calls = {
        "pdx_explicit_steps": (null, 1, 0, 1),
        "pdx_explicit_run": (null, 1, 0, buf, buf),
        "pdx_newmark_run": (null, 1, null, buf, buf, buf, null),
        "pdx_node_stresses": (null, buf, buf),
        "pdx_node_stresses_noalloc": (null, buf, buf),
        "pdx_assemble": (null, 0, null, buf, buf, 1),
        "pdx_static_solve": (null, null, buf, null),
        "pdx_quasi_static_run": (null, 1, null, 2, 1, 1.0, buf, buf, null),
        "pdx_last_timing": (null, buf, buf),
        "pdx_step_log": (null, buf, ibuf),
        "pdx_set_incremental_bonds": (null, 1),
        "pdx_ctx_bonds": (null, ibuf),
        "pdx_local_damage": (null, buf),
        "pdx_explicit_set_form": (null, 0, 0),
        "pdx_max_eig": (null, 10, buf),
        "pdx_set_dt": (null, 1e-7),
    }
    for name, args in calls.items():
        rc = getattr(L, name)(*args)
        assert rc == -1, (name, rc)
        assert L.pdx_last_error(), name


def test_python_constants_match_the_header():
    hdr = open(os.path.join(ROOT, "include", "pdx.h")).read()
    defs = dict(re.findall(r"#define\s+(PDX_[A-Z0-9_]+)\s+(-?\d+)", hdr))
    for name in ("PDX_FP64", "PDX_FP32", "PDX_KRYLOV_BICGSTAB", "PDX_KRYLOV_CG", "PDX_PREC_NONE",
                 "PDX_PREC_JACOBI", "PDX_PREC_BLOCK_JACOBI", "PDX_PREC_LEARNED", "PDX_PREC_TWO_LEVEL", "PDX_PREC_ILU0"):
        if name in defs:
            assert getattr(pdx, name) == int(defs[name]), name
    assert pdx.FB_VARIANTS == int(defs["PDX_FB_VARIANTS"])
    assert pdx.EXPLICIT_FORMS == int(defs["PDX_EXPLICIT_FORMS"])


def _box_vtk_inputs():
    from helpers import box_fixture
    mesh, U = box_fixture()
    z = np.load(os.path.join(ROOT, "tests", "golden", "box", "box_vtk_fields.npz"))
    sha, size, _ = open(os.path.join(ROOT, "tests", "golden", "box", "box_vtk.sha256")).read().split()
    return mesh, U, z["sigma"], z["phi"], sha, int(size)


def test_vtk_binary_is_byte_identical_to_the_reference(tmp_path):
    """The reference's examples/box result file (boxVpd4_static.vtk, 3.5 MB): the
    writer, given the mesh and the file's own U, stress tensor and damage, reproduces
    it byte for byte (SHA-256); one changed stress value changes the hash."""
    import hashlib
    mesh, U, sig, phi, sha, size = _box_vtk_inputs()
    m = pdx.Model(mesh)
    path = tmp_path / "box.vtk"
    m.write_vtk(path, u=U.astype(np.float64), sigma=sig.astype(np.float64), phi=phi.astype(np.float64),
                title="boxVpd4", binary=True)
    raw = path.read_bytes()
    assert len(raw) == size and hashlib.sha256(raw).hexdigest() == sha
    sig2 = sig.astype(np.float64).copy()
    sig2[100, 3] *= 1.5
    m.write_vtk(path, u=U.astype(np.float64), sigma=sig2, phi=phi.astype(np.float64), title="boxVpd4")
    assert hashlib.sha256(path.read_bytes()).hexdigest() != sha


def test_vtk_ascii_layout(tmp_path):
    """The ASCII form (VTKFORMAT ASCII): sections, counts and 4-digit scientific
    values read back to the arrays."""
    mesh, U, sig, phi, _, _ = _box_vtk_inputs()
    m = pdx.Model(mesh)
    path = tmp_path / "box_a.vtk"
    m.write_vtk(path, u=U.astype(np.float64), sigma=sig.astype(np.float64), phi=phi.astype(np.float64),
                title="boxVpd4", binary=False)
    lines = path.read_text().split("\n")
    assert lines[:5] == ["# vtk DataFile Version 3.0", "boxVpd4", "ASCII", "DATASET UNSTRUCTURED_GRID",
                         f"POINTS {m.nn} float"]
    pts = np.array([[float(v) for v in l.split()] for l in lines[5: 5 + m.nn]])
    assert np.allclose(pts, mesh["x"], rtol=1e-4, atol=0)
    i = lines.index(f"CELLS {m.ne} {m.ne * 9}")
    cell = [int(v) for v in lines[i + 1].split()]
    assert cell == [8] + [int(v) - 1 for v in mesh["econ"][0][:8]]
    i = lines.index("VECTORS U float")
    u = np.array([[float(v) for v in l.split()] for l in lines[i + 1: i + 1 + m.nn]])
    assert np.allclose(u, U, rtol=1e-4, atol=1e-30)
    i = lines.index("TENSORS sigma float")
    t = np.array([[float(v) for v in l.split()] for l in lines[i + 1: i + 4]])
    s0 = sig[0].astype(np.float64)
    full = np.array([[s0[0], s0[3], s0[5]], [s0[3], s0[1], s0[4]], [s0[5], s0[4], s0[2]]])
    assert np.allclose(t, full, rtol=1e-4, atol=1e-30) and lines[i + 4] == ""


def test_command_settings_are_parsed(tmp_path):
    """NEWMARK ON|OFF (2 tokens, optional gamma beta), LUMPMASS, VTKFORMAT, SETSOLVING's save
    frequency, FC 2 TopK, TOPK and RF reach the model's settings (fioFile.cpp:69-462), the
    reference's defaults otherwise (datModel.cpp:17-56: no Newmark, consistent mass, binary)."""
    mesh = W.crack2d(seed=2, nx=21, ny=11, dx=1e-3)
    write_lis(tmp_path / "m.lis", mesh)
    (tmp_path / "a.in").write_text("MSHFILE m.lis\nSOLVER DYNAMIC\nSETSOLVING 5e-7 110 20 4 1.0\nNEWMARK ON\n"
                                   "LUMPMASS ON\nFC 2 3\nRF 1 0\nVTKFORMAT ASCII\n")
    (tmp_path / "b.in").write_text("MSHFILE m.lis\nSOLVER DYNAMIC\nSETSOLVING 5e-7 7 3 4 1.0\n"
                                   "NEWMARK OFF 0.6 0.3\nLUMPMASS MAYBE\nTOPK 5\nVTKFORMAT BINARY\n")
    (tmp_path / "c.in").write_text("MSHFILE m.lis\nSOLVER QUASI-STATIC\nSETSOLVING 5e-7 2 0 4 1.0\n")
    a = pdx.Model(cmd_file=str(tmp_path / "a.in")).settings()
    assert (a["solver_flag"], a["newmark"], a["lumped_mass"], a["num_steps"], a["save_freq"]) == (0, 1, 1, 110, 20)
    assert (a["fc"], a["topk"], a["vtk_binary"], a["rf_ids"]) == (2, 3, 0, [1, 0])
    assert a["dt"] == 5e-7 and a["horizon_factor"] == 4.0 and a["nlf"] == 1.0 and a["sigma_ult"] == 1e6
    b = pdx.Model(cmd_file=str(tmp_path / "b.in")).settings()
    assert (b["newmark"], b["lumped_mass"], b["topk"], b["vtk_binary"], b["save_freq"]) == (0, 0, 5, 1, 3)
    assert (b["gamma"], b["beta"]) == (0.6, 0.3)
    c = pdx.Model(cmd_file=str(tmp_path / "c.in")).settings()
    assert (c["solver_flag"], c["dt"], c["save_freq"], c["newmark"], c["lumped_mass"]) == (2, 1.0, 1, 0, 0)
    d = pdx.Model(dict(mesh, lumped_mass=True)).settings()
    assert (d["solver_flag"], d["lumped_mass"], d["vtk_binary"]) == (-1, 1, 1)
    o = O.read_case(str(tmp_path / "a.in"))
    assert o["lumped_mass"] and o["newmark"] and o["fc"] == 2 and o["topk"] == 3 and not o["vtk_binary"]


@pytest.mark.parametrize("name", ["plate2d", "crack2d", "box3d", "tri_tet"])
def test_mass_pattern_matches_oracle(name):
    """setCSRIndexes_gloMassMat: the library's consistent-mass pattern equals the oracle's bit for
    bit, and lies inside the stiffness pattern."""
    if name == "plate2d":
        mesh = W.plate2d(nx=61, ny=31, dx=1e-3, pd_x=(0.02, 0.035))
    elif name == "crack2d":
        mesh = W.crack2d(seed=5, nx=61, ny=31, dx=1e-3)
    elif name == "box3d":
        mesh = W.box3d(n=12, pd_block=6)
    else:
        mesh = tri_tet_mesh()
    mo, mp = O.Model(mesh), pdx.Model(mesh)
    ia0, ja0, _ = mo.mass_csr()
    ia1, ja1 = mp.mass_csr()
    assert np.array_equal(ia0, ia1) and np.array_equal(ja0, ja1)
    kia, kja = mp.csr()
    for r in range(0, mp.neq, max(1, mp.neq // 200)):
        assert set(ja1[ia1[r]: ia1[r + 1]]) <= set(kja[kia[r]: kia[r + 1]])


def test_vtk_ascii_quads_share_one_line(tmp_path):
    """pdfemEleQuad4N::print_vtk ends no line (pdfemEleQuad4N.cpp:201-215): in the ASCII result of
    a 2D quad mesh every cell follows on one line; triangles end theirs (pdfemEleTri3N.cpp:105-121)."""
    mesh = tri_tet_mesh()
    m = pdx.Model(mesh)
    path = tmp_path / "q.vtk"
    m.write_vtk(path, binary=False, title="q")
    lines = path.read_text().split("\n")
    i = next(k for k, l in enumerate(lines) if l.startswith("CELLS "))
    nq = int((np.asarray(mesh["nen"]) == 4).sum())
    want = []
    for e in range(len(mesh["econ"])):
        n = int(mesh["nen"][e])
        want.append(f"{n} " + "".join(f"{int(v) - 1} " for v in mesh["econ"][e][:n]))
    body = "\n".join(lines[i + 1: lines.index("", i + 1) + 1])
    expect = ""
    for e, w in enumerate(want):
        expect += w + ("" if mesh["nen"][e] == 4 else "\n")
    assert body.rstrip("\n") == expect.rstrip("\n") and nq > 0


def test_vtk_pd_and_fem_parts(tmp_path):
    """writeResultsPD_vtk / writeResultsFEM_vtk (fioFile.cpp:771-1001): the PD file holds the
    family nodes and the PD elements with U, sigma and phi; the FEM file the nodes that are not pure
    PD and the FE elements with U and sigma; cells index the part's own points."""
    mesh = W.crack2d(seed=2, nx=21, ny=11, dx=1e-3)
    m = pdx.Model(mesh)
    nd = m.nodes()
    rng = np.random.default_rng(1)
    U = rng.normal(size=(m.nn, 3))
    S = rng.normal(size=(m.nn, 6))
    phi = rng.uniform(size=m.nn)
    for part, keep_n, algo in ((0, nd["ntype"] != 0, 1), (1, nd["ntype"] != 2, 2)):
        path = tmp_path / f"p{part}.vtk"
        m.write_vtk_part(path, part, u=U, sigma=S, phi=phi, title="t")
        lines = path.read_text().split("\n")
        np_ = int(keep_n.sum())
        assert lines[:5] == ["# vtk DataFile Version 3.0", "t", "ASCII", "DATASET UNSTRUCTURED_GRID",
                             f"POINTS {np_} float"]
        pts = np.array([[float(v) for v in l.split()] for l in lines[5: 5 + np_]])
        assert np.allclose(pts, mesh["x"][keep_n], rtol=1e-4)
        ce = np.nonzero(nd["algo"] == algo)[0]
        i = lines.index(f"CELLS {len(ce)} {5 * len(ce)}")
        ids = [int(v) for v in lines[i + 1].split()]  # quads share one line
        mp = -np.ones(m.nn, int)
        mp[keep_n] = np.arange(np_)
        want = []
        for e in ce:
            want += [4] + [int(mp[v - 1]) for v in mesh["econ"][e][:4]]
        assert ids == want and min(want) >= 0
        assert lines[i + 2] == f"CELL_TYPES {len(ce)}"  # the endl after the cells ends their one line
        i = lines.index("VECTORS U float")
        u = np.array([[float(v) for v in l.split()] for l in lines[i + 1: i + 1 + np_]])
        assert np.allclose(u, U[keep_n], rtol=1e-4)
        assert ("SCALARS phi float 1" in lines) == (part == 0)


def test_node_tables(tmp_path):
    """writeUofNode / writeSigofNode (fioFile.cpp:486-538): header lines, then per node the id,
    the coordinates and the values, tab-separated, scientific with 4 (.disp) or 5 (.sig) digits."""
    mesh = W.box3d(n=5, pd_block=2)
    m = pdx.Model(mesh)
    rng = np.random.default_rng(2)
    U, S = rng.normal(size=(m.nn, 3)), rng.normal(size=(m.nn, 6))
    m.write_node_table(tmp_path / "a.disp", 0, u=U)
    m.write_node_table(tmp_path / "a.sig", 1, sigma=S)
    d = (tmp_path / "a.disp").read_text().split("\n")
    assert d[:2] == ["*****The final displacement of  point*****", "ID\tx\ty\tz\tUx\tUy\tUz"]
    row = d[2].split("\t")
    assert row[0] == "1" and row[4] == f"{U[0, 0]:.4e}" and len(row) == 7
    s = (tmp_path / "a.sig").read_text().split("\n")
    assert s[1] == "NID\tx\ty\tz\tSxx\tSyy\tSzz\tSxy\tSyz\tSzx"
    row = s[3].split("\t")
    assert row[0] == "2" and row[1] == f"{mesh['x'][1, 0]:.5e}" and row[9] == f"{S[1, 5]:.5e}"
