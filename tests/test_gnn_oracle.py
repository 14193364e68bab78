"""The CPU oracle of the learned smoother's network, pinned without the GPU:
finite differences of its loss against its autograd gradient, and the
zero-head case, which must reduce to node-block Jacobi computed with numpy."""
import importlib.util
import os

import numpy as np
import torch

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_spec = importlib.util.spec_from_file_location("gnn_oracle", os.path.join(ROOT, "oracle", "gnn_oracle.py"))
G = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(G)


def _tiny(dim=2, nn=12, P=3, seed=0):
    rng = np.random.default_rng(seed)
    adj_off = [0]
    adj = []
    for a in range(nn):
        nb = sorted(set([a] + list(rng.choice(nn, 4, replace=False))))
        adj += nb
        adj_off.append(len(adj))
    adj_off = np.array(adj_off, np.int32)
    adj = np.array(adj, np.int32)
    cw = rng.random(len(adj)).astype(np.float32)
    cw[np.repeat(np.arange(nn), np.diff(adj_off)) == adj] = 0
    eq = np.arange(nn * dim).reshape(nn, dim)
    eq[3, 0] = -1  # one prescribed DOF
    B = rng.normal(size=(nn, dim, dim)) * 0.1 + np.eye(dim)
    return dict(X=rng.normal(size=(nn, 8)).astype(np.float32), cw=cw, adj_off=adj_off, adj=adj,
                Aaa=rng.normal(size=(nn, dim * dim)).astype(np.float32),
                B=B.reshape(nn, dim * dim).astype(np.float32),
                f=rng.normal(size=(P, nn * dim)).astype(np.float32),
                y=rng.normal(size=(P, nn * dim)).astype(np.float32), eq=eq)


def test_layout_size():
    assert G.num_params(2) == 32 * 8 + 32 + 4 * (32 * 64 + 32) + 4 * 36 + 4
    assert G.num_params(3) == 32 * 8 + 32 + 4 * (32 * 64 + 32) + 9 * 41 + 9


def test_oracle_gradient_matches_finite_differences():
    for dim in (2, 3):
        inp = _tiny(dim)
        rng = np.random.default_rng(1)
        w = rng.normal(size=G.num_params(dim)) * 0.2
        L, g = G.loss_and_grad(w, inp, dim)
        idx = rng.choice(len(w), 40, replace=False)
        lay = G.layout(dim)
        idx = np.r_[idx, lay["w_o"][0], lay["b_o"][0], lay["w_in"][0] + 3, lay["w_0"][0] + 70]
        h = 1e-6
        for i in idx:
            wp, wm = w.copy(), w.copy()
            wp[i] += h
            wm[i] -= h
            fd = (G.loss_and_grad(wp, inp, dim)[0] - G.loss_and_grad(wm, inp, dim)[0]) / (2 * h)
            assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])) + 1e-8, (dim, i, fd, g[i])


def test_zero_head_is_block_jacobi():
    dim = 2
    inp = _tiny(dim)
    w = np.random.default_rng(2).normal(size=G.num_params(dim)) * 0.2
    lay = G.layout(dim)
    o, s = lay["w_o"]
    w[o:o + int(np.prod(s)) + dim * dim] = 0.0  # W_o and b_o
    nn = inp["X"].shape[0]
    B = inp["B"].astype(np.float64).reshape(nn, dim, dim).copy()
    free = inp["eq"] >= 0
    B *= free[:, :, None] * free[:, None, :]
    f = inp["f"].astype(np.float64).reshape(-1, nn, dim)
    y = inp["y"].astype(np.float64).reshape(-1, nn, dim)
    R = f - np.einsum("aij,paj->pai", B, y)
    want = (R ** 2).sum() / (f ** 2).sum()
    L, _ = G.loss_and_grad(w, inp, dim)
    assert abs(L - want) <= 1e-12 * want
    np.testing.assert_allclose(G.smoother(w, inp, dim), B, rtol=0, atol=1e-12)
