"""Write the reference's box example as small fixtures (run in the build
container, where the reference checkout is available):

    python tests/golden/make_box_fixture.py

Inputs (data files of the reference, read as text/binary, nothing executed):
  examples/box/boxVpd4_static.in, examples/box/boxVpd4.lis  (mesh + commands)
  examples/box/boxVpd4_static.vtk  (the reference's committed static result)

Outputs:
  tests/golden/box/box_mesh.npz  raw mesh arrays + solver settings
  tests/golden/box/box_U.npz     U (float32, big-endian in the VTK) per node
  tests/golden/box/box_vtk_fields.npz  the VTK's nodal stress tensor (6 components,
                                 float32) and damage phi (float32), decoded from the file
  tests/golden/box/box_vtk.sha256      SHA-256 and size of boxVpd4_static.vtk
"""
import hashlib
import json
import os
import sys

import numpy as np

REF = "/root/reference/examples/box"
HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "box")


def read_vtk_vectors(path, name, n):
    raw = open(path, "rb").read()
    key = f"VECTORS {name} float\n".encode()
    i = raw.index(key) + len(key)
    return np.frombuffer(raw[i: i + n * 12], dtype=">f4").reshape(n, 3).astype(np.float32)


def read_vtk_section(raw, key, count):
    """count big-endian float32 values after the line `key`."""
    i = raw.index(key) + len(key)
    return np.frombuffer(raw[i: i + count * 4], dtype=">f4").astype(np.float32)


def main():
    sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
    import oracle  # the Python restatement of datModel::readdata / excuteCMD

    m = oracle.read_case(os.path.join(REF, "boxVpd4_static.in"))
    os.makedirs(OUT, exist_ok=True)
    ebc = m["ebc"]
    nbc = m["nbc"]
    np.savez_compressed(
        os.path.join(OUT, "box_mesh.npz"),
        x=m["x"], econ=m["econ"], nen=m["nen"], algo=m["algo"],
        ebc_nid=np.concatenate([s["nid"] for s in ebc]), ebc_off=np.cumsum([0] + [len(s["nid"]) for s in ebc]),
        ebc_dof=np.array([s["dof"] for s in ebc]), ebc_val=np.array([s["val"] for s in ebc]),
        nbc_con=np.concatenate([s["con"] for s in nbc]), nbc_off=np.cumsum([0] + [len(s["con"]) for s in nbc]),
        nbc_val=np.array([s["val"] for s in nbc]),
        settings=np.array(json.dumps(dict(dim=m["dim"], proType=m["proType"], E=m["E"], nu=m["nu"], rho=m["rho"],
                                          fac=m["fac"], nlf=m["nlf"], me=m["me"], dt=m["dt"]))))
    U = read_vtk_vectors(os.path.join(REF, "boxVpd4_static.vtk"), "U", len(m["x"]))
    np.savez_compressed(os.path.join(OUT, "box_U.npz"), U=U)
    raw = open(os.path.join(REF, "boxVpd4_static.vtk"), "rb").read()
    n = len(m["x"])
    t = read_vtk_section(raw, b"TENSORS sigma float\n", 9 * n).reshape(n, 9)
    sig = t[:, [0, 4, 8, 1, 5, 2]]  # xx, yy, zz, xy, yz, zx from the written rows
    assert np.array_equal(t, sig[:, [0, 3, 5, 3, 1, 4, 5, 4, 2]])
    phi = read_vtk_section(raw, b"LOOKUP_TABLE  default  \n", n)
    np.savez_compressed(os.path.join(OUT, "box_vtk_fields.npz"), sigma=sig, phi=phi)
    with open(os.path.join(OUT, "box_vtk.sha256"), "w") as f:
        f.write(f"{hashlib.sha256(raw).hexdigest()} {len(raw)} boxVpd4_static.vtk\n")
    print("wrote", OUT, os.path.getsize(os.path.join(OUT, "box_mesh.npz")), os.path.getsize(os.path.join(OUT, "box_U.npz")))


def load_box():
    """Rebuild the raw-mesh dict from the fixture."""
    z = np.load(os.path.join(OUT, "box_mesh.npz"))
    s = json.loads(str(z["settings"]))
    eo, no = z["ebc_off"], z["nbc_off"]
    ebc = [dict(nid=z["ebc_nid"][eo[i]: eo[i + 1]], dof=int(z["ebc_dof"][i]), val=float(z["ebc_val"][i]))
           for i in range(len(eo) - 1)]
    nbc = [dict(con=z["nbc_con"][no[i]: no[i + 1]], val=float(z["nbc_val"][i])) for i in range(len(no) - 1)]
    m = dict(x=z["x"], econ=z["econ"], nen=z["nen"], algo=z["algo"], ebc=ebc, nbc=nbc,
             cracks=np.zeros((0, 3, 3)), beta=0.25, gamma=0.5, **s)
    U = np.load(os.path.join(OUT, "box_U.npz"))["U"]
    return m, U


if __name__ == "__main__":
    main()
