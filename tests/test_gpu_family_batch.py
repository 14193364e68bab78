"""Batched PD family kernel (node cloud, per-node horizons) against the oracle.

  members    the family lists equal the brute-force member sets of the reference
             rule (initialPDN_Fami), self first
  A          shape tensor within 1e-12 relative of the oracle's (the oracle sums
             in the reference's expression order, the kernel in a symmetric form)
  L          L L^T diag(s) reproduces A within 1e-13 relative (S = L L^T, the
             symmetric part of A; s = 1/2 on the squared terms)
  H          bond blocks (FP32) within 2e-6 relative of the oracle's FP64 blocks,
             per family; the self block is minus the sum of the others
  variants   every launch form gives the same A, L and H (bitwise)
"""
import numpy as np
import pytest

from helpers import load_oracle, load_pdx, workloads

pytestmark = pytest.mark.gpu

O = load_oracle()
W = workloads()
pdx = load_pdx()

E, NU = 70e9, 0.33
LAM = E * NU / ((1 + NU) * (1 - 2 * NU))
MU = E / (2 * (1 + NU))


def _cloud(n, dim=3, seed=0):
    c = W.family_batch(n=n, seed=seed)
    if dim == 2:
        x = c["x"].reshape(n, n, n, 3)[0].reshape(-1, 3).copy()
        x[:, 2] = 0.0
        return dict(x=x, delta=c["delta"][: n * n], vol=np.full(n * n, c["dx"] ** 2), dx=c["dx"])
    return c


def _rebuild(L, dim):
    s = np.ones(L.shape[0])
    s[[3, 4, 5] if dim == 3 else [2, 3]] = 0.5
    return (L @ L.T) * s[None, :]


@pytest.mark.parametrize("dim", [3, 2])
def test_family_batch_matches_oracle(dim):
    c = _cloud(11 if dim == 3 else 30, dim)
    fb = pdx.FamilyBatch(c["x"], c["vol"], c["delta"], nlf=0.33, lam=LAM, mu=MU, dim=dim)
    off, mem = fb.families()
    boff, bmem = O.fb_members(c["x"], c["delta"])
    assert np.array_equal(off, boff.astype(np.int64))
    for k in range(fb.n_fam):
        a, b = mem[off[k]:off[k + 1]], bmem[boff[k]:boff[k + 1]]
        assert a[0] == k and b[0] == k
        assert np.array_equal(np.sort(a[1:]), np.sort(b[1:])), k
    fb.run(0)
    A, L, H = fb.results(0, fb.n_fam, off)
    for k in range(fb.n_fam):
        A0, H0 = O.fb_family(dim, c["x"], c["vol"], c["delta"], 0.33, LAM, MU, mem[off[k]:off[k + 1]])
        assert np.linalg.norm(A[k] - A0) <= 1e-12 * np.linalg.norm(A0), k
        assert np.linalg.norm(_rebuild(L[k], dim) - A[k]) <= 1e-13 * np.linalg.norm(A[k]), k
        Hk = H[off[k]:off[k + 1]].astype(np.float64)
        assert np.linalg.norm(Hk - H0) <= 2e-6 * np.linalg.norm(H0), (k, np.linalg.norm(Hk - H0) / np.linalg.norm(H0))
        assert np.linalg.norm(Hk[0] + Hk[1:].sum(0)) <= 1e-5 * np.linalg.norm(Hk[1:]), k


def test_family_batch_variants_agree():
    c = _cloud(14)
    fb = pdx.FamilyBatch(c["x"], c["vol"], c["delta"], nlf=0.33, lam=LAM, mu=MU)
    off, _ = fb.families()
    fb.run(0)
    ref = fb.results(0, fb.n_fam, off)
    for v in range(1, pdx.FB_VARIANTS):
        try:
            fb.run(v)
        except pdx.PdxError:  # a form whose LDS slots exceed the device limit
            continue
        got = fb.results(0, fb.n_fam, off)
        for a, b in zip(got, ref):
            assert np.array_equal(a, b), v
    ms = fb.run(-1)
    best, times = fb.tuning()
    # the tuner picks the fastest of every form it could run, one family per block included
    assert 0 <= best < pdx.FB_VARIANTS and ms > 0 and times[best] > 0
    assert times[best] == min(t for t in times if t > 0)


def test_family_batch_full_cloud_sampled():
    """The 2,000,376-node cloud (126^3, jitter 0.3 dx, horizons U(2.5, 3.5) dx):
    sampled families against the oracle after the autotuned run."""
    c = W.family_batch()
    fb = pdx.FamilyBatch(c["x"], c["vol"], c["delta"], nlf=0.33, lam=LAM, mu=MU)
    assert fb.n_fam == 126 ** 3
    off, mem = fb.families()
    fb.run(-1)
    rng = np.random.default_rng(0)
    for k in rng.choice(fb.n_fam, 64, replace=False):
        A, L, H = fb.results(int(k), 1, off)
        A0, H0 = O.fb_family(3, c["x"], c["vol"], c["delta"], 0.33, LAM, MU, mem[off[k]:off[k + 1]])
        assert np.linalg.norm(A[0] - A0) <= 1e-12 * np.linalg.norm(A0)
        assert np.linalg.norm(H.astype(np.float64) - H0) <= 2e-6 * np.linalg.norm(H0)
