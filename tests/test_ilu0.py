"""Block ILU(0) preconditioner (PDX_PREC_ILU0) and the solver fallback chain.

The reference solves with MKL PARDISO and has no incomplete factorisation, so
the oracle here is a textbook restatement (oracle/pdlsm_oracle.c om_bilu0:
IKJ ILU(0), Saad Alg. 10.4, over d x d node blocks).  The CPU tests pin that
restatement by what the mathematics fixes:
  * on a full pattern ILU(0) is the exact LU, so its solve is A^-1;
  * on any pattern the product L U equals A on every block of the pattern;
  * the block factor equals the scalar ILU(0) of the expanded DOF pattern
    (same incomplete product L U);
  * with a block-tridiagonal pattern in natural order nothing is dropped;
  * the colouring is proper and greedy.
The GPU tests compare the library's factor with the restatement bit for bit
and its solves with the true residual."""
import ctypes

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as sla

from helpers import load_oracle, load_pdx, workloads

O = load_oracle()
W = workloads()
pdx = load_pdx()




def _factor_product(off, adj, F, rank, d):
    nn = len(off) - 1
    L = np.eye(nn * d)
    U = np.zeros((nn * d, nn * d))
    for a in range(nn):
        for s in range(off[a], off[a + 1]):
            b = adj[s]
            blk = F[s]
            if rank[b] < rank[a]:
                L[a * d:(a + 1) * d, b * d:(b + 1) * d] = blk
            else:
                U[a * d:(a + 1) * d, b * d:(b + 1) * d] = blk
    return L @ U


def _dense(off, adj, Kb, d):
    nn = len(off) - 1
    A = np.zeros((nn * d, nn * d))
    for a in range(nn):
        for s in range(off[a], off[a + 1]):
            b = adj[s]
            A[a * d:(a + 1) * d, b * d:(b + 1) * d] = Kb[s]
    return A


def _random_blocks(off, adj, d, seed):
    rng = np.random.default_rng(seed)
    nn = len(off) - 1
    Kb = rng.uniform(-1, 1, (len(adj), d, d))
    for a in range(nn):
        for s in range(off[a], off[a + 1]):
            if adj[s] == a:
                Kb[s] += np.eye(d) * (3.0 * (off[a + 1] - off[a]))
    return Kb


def _small_model(name):
    if name == "crack2d":
        mesh = W.crack2d(seed=3, nx=41, ny=21, dx=2.5e-3)
    else:
        mesh = W.box3d(n=7, pd_block=3)
    mo = O.Model(mesh)
    mp = pdx.Model(mesh)
    K, F = mo.assemble(1)
    ia, ja = mo.csr()
    off, adj = mp.adjacency()
    d = mp.dim
    Kb = O.masked_blocks(off, adj, mp.nodes()["eq"], d, ia, ja, K)
    return mesh, mo, mp, off, adj, d, Kb, K, F


@pytest.mark.parametrize("d", [1, 2, 3])
def test_oracle_full_pattern_is_exact_lu(d):
    nn = 6
    off = np.arange(nn + 1, dtype=np.int32) * nn
    adj = np.tile(np.arange(nn, dtype=np.int32), nn)
    Kb = _random_blocks(off, adj, d, seed=d)
    A = _dense(off, adj, Kb, d)
    r = np.random.default_rng(7).standard_normal(nn * d)
    for rank in (np.arange(nn, dtype=np.int32), np.array([3, 0, 5, 1, 4, 2], np.int32)):
        F, Dinv = O.bilu0(off, adj, Kb, rank, d)
        z = O.bilu0_apply(off, adj, F, Dinv, rank, r, d)
        assert np.linalg.norm(z - np.linalg.solve(A, r)) <= 1e-13 * np.linalg.norm(z)
        assert np.abs(_factor_product(off, adj, F, rank, d) - A).max() <= 1e-13 * np.abs(A).max()


@pytest.mark.parametrize("name", ["crack2d", "box3d"])
def test_oracle_product_matches_A_on_the_pattern(name):
    _, _, _, off, adj, d, Kb, _, _ = _small_model(name)
    col, nc = O.greedy_colors(off, adj)
    rank = O.colour_rank(col)
    F, Dinv = O.bilu0(off, adj, Kb, rank, d)
    A = _dense(off, adj, Kb, d)
    M = _factor_product(off, adj, F, rank, d)
    mask = _dense(off, adj, np.ones_like(Kb), d) != 0
    scale = np.abs(A).max()
    assert np.abs((M - A)[mask]).max() <= 1e-12 * scale
    # something was dropped off the pattern: the factor is incomplete
    assert np.abs(M[~mask]).max() > 1e-6 * scale
    # pivot inverses are the inverses of U's diagonal blocks
    for a in range(0, len(off) - 1, 17):
        s = off[a] + int(np.nonzero(adj[off[a]:off[a + 1]] == a)[0][0])
        assert np.allclose(Dinv[a] @ F[s], np.eye(d), atol=1e-12)


def test_oracle_block_factor_equals_scalar_ilu0():
    _, _, _, off, adj, d, Kb, _, _ = _small_model("crack2d")
    col, _ = O.greedy_colors(off, adj)
    rank = O.colour_rank(col)
    F, Dinv = O.bilu0(off, adj, Kb, rank, d)
    nn = len(off) - 1
    # expand to the scalar DOF pattern: DOF a*d+i couples with every b*d+j of a neighbour b
    soff = [0]
    sadj = []
    sval = []
    for a in range(nn):
        for i in range(d):
            for s in range(off[a], off[a + 1]):
                for j in range(d):
                    sadj.append(adj[s] * d + j)
                    sval.append(Kb[s, i, j])
            soff.append(len(sadj))
    soff = np.array(soff, np.int32)
    sadj = np.array(sadj, np.int32)
    srank = (np.repeat(rank, d) * d + np.tile(np.arange(d), nn)).astype(np.int32)
    Fs, Ds = O.bilu0(soff, sadj, np.array(sval).reshape(-1, 1, 1), srank, 1)
    r = np.random.default_rng(1).standard_normal(nn * d)
    zb = O.bilu0_apply(off, adj, F, Dinv, rank, r, d)
    zs = O.bilu0_apply(soff, sadj, Fs, Ds, srank, r, 1)
    assert np.linalg.norm(zb - zs) <= 1e-11 * np.linalg.norm(zb)


@pytest.mark.parametrize("d", [2, 3])
def test_oracle_block_tridiagonal_natural_order_is_exact(d):
    nn = 9
    off = [0]
    adj = []
    for a in range(nn):
        adj += [b for b in (a - 1, a, a + 1) if 0 <= b < nn]
        off.append(len(adj))
    off = np.array(off, np.int32)
    adj = np.array(adj, np.int32)
    Kb = _random_blocks(off, adj, d, seed=11)
    rank = np.arange(nn, dtype=np.int32)
    F, Dinv = O.bilu0(off, adj, Kb, rank, d)
    r = np.random.default_rng(2).standard_normal(nn * d)
    z = O.bilu0_apply(off, adj, F, Dinv, rank, r, d)
    A = _dense(off, adj, Kb, d)
    assert np.linalg.norm(z - np.linalg.solve(A, r)) <= 1e-13 * np.linalg.norm(z)
    # red-black order drops fill on the same matrix
    col, nc = O.greedy_colors(off, adj)
    assert nc == 2
    rank2 = O.colour_rank(col)
    F2, D2 = O.bilu0(off, adj, Kb, rank2, d)
    z2 = O.bilu0_apply(off, adj, F2, D2, rank2, r, d)
    assert np.linalg.norm(z2 - np.linalg.solve(A, r)) > 1e-6 * np.linalg.norm(z)


@pytest.mark.parametrize("name", ["crack2d", "box3d"])
def test_oracle_colouring_is_proper_and_greedy(name):
    _, _, _, off, adj, _, _, _, _ = _small_model(name)
    col, nc = O.greedy_colors(off, adj)
    assert col.max() + 1 == nc
    nn = len(off) - 1
    rows = np.repeat(np.arange(nn), np.diff(off))
    S = sp.csr_matrix((np.ones(len(adj)), (rows, adj)), shape=(nn, nn))
    S = (S + S.T).tocsr()  # the colouring is of the symmetrised graph (K's node pattern need not be)
    assert (S != S.T).nnz == 0
    for a in range(nn):
        nb = S.indices[S.indptr[a]:S.indptr[a + 1]]
        nb = nb[nb != a]
        assert not np.any(col[nb] == col[a])
        lower = set(col[nb[nb < a]].tolist())
        assert set(range(col[a])) <= lower


def test_oracle_ilu0_cuts_iterations():
    """BiCGStab (SciPy) on the small cracked plate: colour-order ILU(0) needs fewer iterations
    than Jacobi."""
    _, mo, mp, off, adj, d, Kb, K, F = _small_model("crack2d")
    ia, ja = mo.csr()
    A = sp.csr_matrix((K, ja, ia), shape=(mo.neq, mo.neq))
    eq = mp.nodes()["eq"][:, :d]
    m = eq >= 0
    col, _ = O.greedy_colors(off, adj)
    rank = O.colour_rank(col)
    Fa, Dinv = O.bilu0(off, adj, Kb, rank, d)

    def ilu(v):
        full = np.zeros((len(off) - 1, d))
        full[m] = v[eq[m]]
        return O.bilu0_apply(off, adj, Fa, Dinv, rank, full.reshape(-1), d).reshape(-1, d)[m][np.argsort(eq[m])]

    def count(M):
        it = [0]
        x, info = sla.bicgstab(A, F, rtol=1e-8, maxiter=5000, M=M, callback=lambda _: it.__setitem__(0, it[0] + 1))
        assert info == 0 and np.linalg.norm(F - A @ x) <= 1e-8 * np.linalg.norm(F) * 1.01
        return it[0]

    dg = A.diagonal()
    n = mo.neq
    it_j = count(sla.LinearOperator((n, n), matvec=lambda v: v / dg))
    it_i = count(sla.LinearOperator((n, n), matvec=ilu))
    assert it_i < 0.6 * it_j, (it_i, it_j)


# ---------------------------------------------------------------- GPU --------
@pytest.mark.gpu
@pytest.mark.parametrize("name", ["crack2d", "box3d"])
def test_gpu_ilu0_factor_matches_oracle_bitwise(name):
    mesh, mo, mp, off, adj, d, _, _, _ = _small_model(name)
    cx = pdx.Context(mp)
    Kg, _ = cx.assemble(pdx.PDX_MODE_STATIC)
    ia, ja = mp.csr()
    Kb = O.masked_blocks(off, adj, mp.nodes()["eq"], d, ia, ja, Kg)
    r = np.random.default_rng(5).standard_normal(mp.nn * d)
    r.reshape(-1, d)[mp.nodes()["eq"][:, :d] < 0] = 0.0
    g = cx.ilu0(r)
    col, nc = O.greedy_colors(off, adj)
    print(f"ilu0 {name}: nodes {mp.nn} colours {nc}", flush=True)
    assert g["ncolors"] == nc and np.array_equal(g["colors"], col)
    rank = O.colour_rank(col)
    F, Dinv = O.bilu0(off, adj, Kb, rank, d)
    assert np.array_equal(g["F"], F)
    assert np.array_equal(g["Dinv"], Dinv)
    z = O.bilu0_apply(off, adj, F, Dinv, rank, r, d)
    assert np.linalg.norm(g["z"] - z) <= 1e-12 * np.linalg.norm(z)


def _true_relres(mo, K, F, u):
    A = mo.sparse(K)
    return np.linalg.norm(F - A @ u) / np.linalg.norm(F)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [pdx.PDX_FP64, pdx.PDX_FP32])
def test_gpu_static_solve_ilu0(precision):
    mesh = W.crack2d(seed=5, nx=121, ny=61, dx=1e-3)
    mo, mp = O.Model(mesh), pdx.Model(mesh)
    cx = pdx.Context(mp)
    K, F = mo.assemble(1)
    u, st = cx.static_solve(pdx.solve_opts(precision=precision, precond=pdx.PDX_PREC_ILU0, tol=1e-8))
    print(f"ilu0 solve precision {precision}: {st}", flush=True)
    assert st["converged"] and st["precond_used"] == pdx.PDX_PREC_ILU0 and st["attempts"] == 1
    assert _true_relres(mo, K, F, u) <= 1.2e-8
    _, sj = cx.static_solve(pdx.solve_opts(precision=precision, precond=pdx.PDX_PREC_JACOBI, tol=1e-8))
    assert st["iterations"] < 0.6 * sj["iterations"], (st["iterations"], sj["iterations"])


@pytest.mark.gpu
def test_gpu_fallback_chain():
    mesh = W.crack2d(seed=5, nx=61, ny=31, dx=2e-3)
    mo, mp = O.Model(mesh), pdx.Model(mesh)
    cx = pdx.Context(mp)
    K, F = mo.assemble(1)
    # no smoother has been predicted for this K: without the fallback the learned solve refuses
    with pytest.raises(pdx.PdxError, match="no learned smoother"):
        cx.static_solve(pdx.solve_opts(precond=pdx.PDX_PREC_LEARNED))
    u, st = cx.static_solve(pdx.solve_opts(precond=pdx.PDX_PREC_LEARNED, fallback=True))
    assert st["converged"] and st["attempts"] == 2 and st["precond_used"] == pdx.PDX_PREC_ILU0
    assert _true_relres(mo, K, F, u) <= 1.2e-8
    # an iteration budget too small for every method walks the whole chain and reports it
    u, st, rc = cx.static_solve(pdx.solve_opts(precond=pdx.PDX_PREC_TWO_LEVEL, precision=pdx.PDX_FP32, max_iter=2,
                                               fallback=True), check=False)
    assert rc == -7 and not st["converged"]
    assert st["attempts"] == 4 and st["precond_used"] == pdx.PDX_PREC_BLOCK_JACOBI
    assert st["precision_used"] == pdx.PDX_FP64
    # CG on the unsymmetric K with no preconditioner, rescued by BiCGStab + node-block Jacobi
    u, st, rc = cx.static_solve(pdx.solve_opts(krylov=pdx.PDX_KRYLOV_CG, precond=pdx.PDX_PREC_NONE, max_iter=300,
                                               fallback=True), check=False)
    assert rc == 0 and st["converged"]
    assert _true_relres(mo, K, F, u) <= 1.2e-8
