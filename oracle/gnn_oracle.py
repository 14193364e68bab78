"""CPU oracle of the learned smoother's network and loss -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline may use this
module, and only as the reference the HIP library (libpdx.so) is tested against.

The learned smoother has no counterpart in the reference: QibangLiu/PDLSM-FEM_mpi
solves its systems with MKL PARDISO (SourceFiles/Solver/calMatrixOperations.cpp:378-498)
and has no preconditioner or network.  This module restates, in float64 torch on the
CPU, the network and loss that libpdx.so trains (its definition: a message-passing
network on the node graph of K predicting M_a = (I + 0.5 tanh(o_a)) B_a, loss
sum_p ||f_p - M y_p||^2 / sum_p ||f_p||^2), so that torch.autograd supplies a
gradient computed independently of the library's hand-written backward kernels.

Parameter layout (flat float32 vector, the library's order):
  W_in [32 x 8], b_in [32], then per layer l < 4: W_l [32 x 64], b_l [32],
  then W_o [d*d x (32 + d*d)], b_o [d*d].
"""
from __future__ import annotations

import numpy as np
import torch

GF, GIN, GL = 32, 8, 4


def layout(dim: int) -> dict:
    dd = dim * dim
    off, o = {}, 0
    for name, shape in [("w_in", (GF, GIN)), ("b_in", (GF,))] + \
            [item for l in range(GL) for item in ((f"w_{l}", (GF, 2 * GF)), (f"b_{l}", (GF,)))] + \
            [("w_o", (dd, GF + dd)), ("b_o", (dd,))]:
        n = int(np.prod(shape))
        off[name] = (o, shape)
        o += n
    off["total"] = o
    return off


def num_params(dim: int) -> int:
    return layout(dim)["total"]


def unflatten(w: torch.Tensor, dim: int) -> dict:
    lay = layout(dim)
    out = {}
    for k, v in lay.items():
        if k != "total":
            o, s = v
            out[k] = w[o:o + int(np.prod(s))].reshape(s)
    return out


def forward(w: torch.Tensor, inp: dict, dim: int):
    """Smoother blocks M [nn, d, d] and head outputs O [nn, d*d]."""
    p = unflatten(w, dim)
    X = torch.as_tensor(inp["X"], dtype=torch.float64)
    cw = torch.as_tensor(inp["cw"], dtype=torch.float64)
    adj_off = np.asarray(inp["adj_off"])
    adj = torch.as_tensor(np.asarray(inp["adj"]), dtype=torch.long)
    nn = X.shape[0]
    erow = torch.as_tensor(np.repeat(np.arange(nn), np.diff(adj_off)), dtype=torch.long)
    h = torch.relu(X @ p["w_in"].T + p["b_in"])
    for l in range(GL):
        m = torch.zeros_like(h).index_add_(0, erow, cw[:, None] * h[adj])
        h = torch.relu(torch.cat([h, m], 1) @ p[f"w_{l}"].T + p[f"b_{l}"]) + h
    Aaa = torch.as_tensor(inp["Aaa"], dtype=torch.float64)
    O = torch.cat([h, Aaa], 1) @ p["w_o"].T + p["b_o"]
    T = 0.5 * torch.tanh(O).reshape(nn, dim, dim)
    B = torch.as_tensor(inp["B"], dtype=torch.float64).reshape(nn, dim, dim)
    M = (torch.eye(dim, dtype=torch.float64) + T) @ B
    free = torch.as_tensor(np.asarray(inp["eq"])[:, :dim] >= 0, dtype=torch.float64)
    M = M * free[:, :, None] * free[:, None, :]
    return M, O


def loss(w: torch.Tensor, inp: dict, dim: int) -> torch.Tensor:
    M, _ = forward(w, inp, dim)
    nn = M.shape[0]
    f = torch.as_tensor(inp["f"], dtype=torch.float64).reshape(-1, nn, dim)
    y = torch.as_tensor(inp["y"], dtype=torch.float64).reshape(-1, nn, dim)
    R = f - torch.einsum("aij,paj->pai", M, y)
    return (R * R).sum() / (f * f).sum()


def loss_and_grad(w_np: np.ndarray, inp: dict, dim: int):
    w = torch.tensor(np.asarray(w_np, np.float64), requires_grad=True)
    L = loss(w, inp, dim)
    L.backward()
    return float(L.detach()), w.grad.numpy().copy()


def smoother(w_np: np.ndarray, inp: dict, dim: int) -> np.ndarray:
    with torch.no_grad():
        M, _ = forward(torch.tensor(np.asarray(w_np, np.float64)), inp, dim)
    return M.numpy()
