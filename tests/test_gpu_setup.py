"""Model set-up on the GPU (pdx_set_setup_device): the cell-list family search and the node graph
of the stiffness pattern give the host set-up's families (members in the reference's scan order)
and pattern bit for bit, on 2D and 3D, quad / hex / tri / tet, cracked and PD-to-the-surface
meshes; the batch's member search (always on the GPU) is checked against the oracle in
test_gpu_family_batch.py."""
import numpy as np
import pytest

from helpers import load_pdx, tri_tet_mesh, tet_box_mesh, workloads

pytestmark = pytest.mark.gpu

pdx = load_pdx()
W = workloads()

CASES = {
    "plate2d": lambda: W.plate2d(),
    "crack2d": lambda: W.crack2d(seed=5, nx=121, ny=61, dx=1e-3),
    "box3d": lambda: W.box3d(n=16, pd_block=8),
    "penny3d": lambda: W.penny3d(n=20, pd_block=10),
    "box3d_allpd": lambda: W.box3d(n=10, pd_block=10, jitter=0.3),
    "tri_tet": tri_tet_mesh,
    "tet_box": tet_box_mesh,
}


def _model(mesh, device):
    pdx.set_setup_device(device)
    try:
        return pdx.Model(mesh)
    finally:
        pdx.set_setup_device(-1)


@pytest.mark.parametrize("name", list(CASES))
def test_gpu_setup_matches_host(name):
    mesh = CASES[name]()
    h = _model(mesh, -1)
    g = _model(mesh, 0)
    fh, fg = h.families(), g.families()
    for key in ("node", "off", "mem", "bond", "cd_v", "delta"):
        assert np.array_equal(fh[key], fg[key]), key
    assert h.max_family == g.max_family
    oh, ah = h.adjacency()
    og, ag = g.adjacency()
    assert np.array_equal(oh, og) and np.array_equal(ah, ag)
    assert h.nnz == g.nnz and h.neq == g.neq
    assert ah.size > 0 and (fh["mem"].size > 0 or name == "tet_box")  # the tet box is FE only


def test_gpu_setup_device_checked():
    with pytest.raises(pdx.PdxError):
        pdx.set_setup_device(4096)
    with pytest.raises(pdx.PdxError):
        pdx.set_setup_device(-2)
