"""Benchmark of the PDLSM-FEM hot path on one MI355X.

    python bench.py [--gpus 1] [--steps 2000] [--warmup 30]

Workload (the explicit configuration the headline metric is quoted on): 3D box
of 60^3 nodes (dx = 1 mm, 205,379 hex8), interior nodes jittered
U(-0.2, 0.2) dx (seed 0), central 20^3-node PDLSM cube, horizon 3.015 dx,
E = 70 GPa, nu = 0.33, rho = 2700, bottom face UZ = 0 plus a corner pin,
1 MPa step traction on the top face, lumped mass, explicit central difference
at 0.8 x the critical dt (power iteration on M^-1 K).

A step is one explicit central-difference step (internal force K u_n over all
rows, external force, update of the free and prescribed DOFs) over the whole
mesh; state is resident in HBM when the timed region starts and the steps run
as replays of a captured 6-step hipGraph, captured before the timed region.
Value = DOF-steps/s in FP64 (the reference's precision); the FP32 internal-force /
FP64-state variant and its drift from the FP64 trajectory are reported beside it.

Also reported: K assembly time (family + element kernels), the static solve
time to a 1e-8 true relative residual on the edge-cracked plate (501 x 251
nodes), and the CPU oracle timed on a bounded sample of the same explicit
workload on this host.
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "pdlsm-fem_mpi_amd")
METRIC = "DOF-steps/s explicit; PCG s-to-1e-8 (learned vs plain) on 1 MI355X"
HBM_PEAK = 8.0e12


def load_pdx():
    spec = importlib.util.spec_from_file_location("pdlsm_fem_mpi_amd", os.path.join(PKG, "__init__.py"),
                                                  submodule_search_locations=[PKG])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["pdlsm_fem_mpi_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def stress_bytes(m):
    """Algorithmic bytes of one nodal-stress recovery: DOF values and coordinates read once,
    the element connectivity, the bond strain weights and member ids of the families, the
    node-to-family and node-to-element-corner maps, and the nodal stresses written.  The
    element corner stresses staged between the element and node kernels are excluded."""
    fam = m.families()
    nmem, nfam = len(fam["mem"]), m.nfam
    nfe = int((m.nodes()["algo"] == 2).sum())
    nen = 8 if m.dim == 3 else 4
    return (m.nn * m.dim * 8 + m.nn * 3 * 8 + nfe * nen * 4 + nmem * (m.dim * 8 + 4) + (nfam + 1) * 4
            + m.nn * 4 + (m.nn + 1) * 4 + nfe * nen * 4 + m.nn * 6 * 8)


def column_classes(m):
    """(classes, offsets in all classes) of the node rows grouped by the offsets of their
    block columns from the row index."""
    off, adj = m.adjacency()
    ids = {}
    for r in range(m.nn):
        key = (adj[off[r]: off[r + 1]] - r).tobytes()
        if key not in ids:
            ids[key] = off[r + 1] - off[r]
    return len(ids), int(sum(ids.values()))


def step_bytes(m, nblk, prec, col_classes=None, n_own=None):
    """Algorithmic HBM bytes of one explicit step in the node-block layout:
    block values (d*d per block) + one 4-byte column per block, row offsets,
    u_n gathered once (FP64), and per DOF: read u_n, u_{n-1}, m, unit traction
    force, write u_{n+1} (FP64).  With column pattern classes (col_classes =
    (classes, offsets)) the columns per block are replaced by one class per row
    and the classes' offset lists.  With mirrored storage (n_own own blocks) only
    the own blocks' values are read, plus one 4-byte source index per block."""
    d = m.dim
    sv = 4 if prec == 1 else 8
    cols = nblk * 4 if col_classes is None else m.nn * 4 + (col_classes[0] + 1) * 4 + col_classes[1] * 4
    vals = nblk * d * d * sv if n_own is None else n_own * d * d * sv + nblk * 4
    return vals + cols + (m.nn + 1) * 4 + m.nn * d * 8 + m.nn * d * (8 * 4 + 8)


def own_blocks(m):
    """Own blocks of the mirrored K storage: all blocks but the lower blocks (r, c) whose two
    orientations no family stores (those equal their partner's transpose)."""
    off, adj = m.adjacency()
    fam = m.families()
    nn = m.nn
    rows = np.repeat(np.arange(nn, dtype=np.int64), np.diff(off))
    fn = fam["node"].astype(np.int64)
    pdkeys = np.sort(np.repeat(fn, np.diff(fam["off"])) * nn + fam["mem"].astype(np.int64))

    def stored(keys):
        if len(pdkeys) == 0:
            return np.zeros(len(keys), bool)
        i = np.minimum(np.searchsorted(pdkeys, keys), len(pdkeys) - 1)
        return pdkeys[i] == keys

    adj64 = adj.astype(np.int64)
    mirrored = (adj64 < rows) & ~stored(rows * nn + adj64) & ~stored(adj64 * nn + rows)
    return int(len(adj) - mirrored.sum())


def scalar_csr_step_bytes(m, prec):
    """The scalar-CSR figure: nnz (s_v + 4) + (N_eq + 1) 4 + N_dof s_x + N_eq (8*4 + 8)."""
    sv = 4 if prec == 1 else 8
    return m.nnz * (sv + 4) + (m.neq + 1) * 4 + m.nn * m.dim * sv + m.neq * 40


def cpu_baseline(budget_s=15.0):
    """The oracle's explicit loop (per-step rebuild of every family and element, per-bond LU,
    binary-search scatter, as the reference runs it) on a bounded sample of the same workload,
    on all host cores (the reference's rank slices of families, boundary entities and
    elements as threads, F summed in rank order) and on one core."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    sys.path.insert(0, PKG)
    import oracle
    import workloads as W
    oracle.build()
    mesh = W.box3d(n=30, pd_block=10)
    mesh["dt"] = 1e-8
    cores = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)

    def run(threads, budget):
        oracle.lib().om_set_threads(threads)
        try:
            M = oracle.Model(mesh)
            t0 = time.time()
            M.explicit_run(0)
            one = time.time() - t0
            steps = max(1, min(200, int(budget / max(one, 1e-3)) - 1))
            M = oracle.Model(mesh)
            t0 = time.time()
            M.explicit_run(steps)
            return M.neq * (steps + 1) / (time.time() - t0), steps, time.time() - t0, M.neq
        finally:
            oracle.lib().om_set_threads(1)

    v1, s1, d1, neq = run(1, budget_s * 0.6)
    vn, sn, dn, _ = run(cores, budget_s * 0.4) if cores > 1 else (v1, s1, d1, neq)
    what = (f"oracle (per-step rebuild of every family and element, per-bond LU, binary-search scatter) "
            f"on the same workload at 30^3 nodes with a 10^3-node PD cube ({neq} DOF)")
    return {"value": vn, "unit": "DOF-steps/s", "cores": cores, "kind": "port",
            "sample": f"{what}, {cores} threads (the reference's rank slices), start-up + {sn} steps, {dn:.1f} s",
            "single_core": {"value": v1, "cores": 1, "sample": f"{what}, 1 thread, start-up + {s1} steps, {d1:.1f} s"}}


def _mkl_rt():
    """The MKL single dynamic library, if this host has one (None otherwise)."""
    import ctypes
    import ctypes.util
    names = [ctypes.util.find_library("mkl_rt")]
    for pre in (sys.prefix, "/opt/conda"):
        names += [os.path.join(pre, "lib", "libmkl_rt.so.1"), os.path.join(pre, "lib", "libmkl_rt.so")]
    for nm in names:
        if nm:
            try:
                return ctypes.CDLL(nm), nm
            except OSError:
                pass
    return None, None


def direct_solve(path):
    """CPU sparse direct solve of the system saved at path (ia, ja, a, b; zero-based CSR), as the
    reference solves it: MKL PARDISO, real unsymmetric (mtype 11), phases 11 / 22 / 33, with the
    reference's iparm (calMatrixOperations.cpp:277-301: METIS ordering, 2 refinement steps,
    1e-13 pivot perturbation, MPS scaling and weighted matching, zero-based indices) and the
    sequential threading layer it links (one core).  Without an MKL runtime: SciPy SuperLU
    (COLAMD).  Prints one JSON line.  Runs in a child process of the bench, on the CPU only."""
    import ctypes as C
    z = np.load(path)
    ia, ja = np.ascontiguousarray(z["ia"], np.int64), np.ascontiguousarray(z["ja"], np.int64)
    a, b = np.ascontiguousarray(z["a"], np.float64), np.ascontiguousarray(z["b"], np.float64)
    n = len(b)
    x = np.zeros(n)
    lib, name = _mkl_rt()
    res = {"neq": n, "nnz": int(len(a))}
    t0 = time.time()
    if lib is not None:
        pt = (C.c_void_p * 64)()
        iparm = np.zeros(64, np.int64)
        for k, v in {0: 1, 1: 2, 7: 2, 9: 13, 10: 1, 12: 1, 17: -1, 18: -1, 34: 1}.items():
            iparm[k] = v
        L = C.c_longlong
        one, mtype, msglvl, idum, err, nn = L(1), L(11), L(0), L(0), L(0), L(n)
        ptr = lambda v: v.ctypes.data_as(C.c_void_p)
        phases = {}
        for ph in (11, 22, 33, -1):
            t1 = time.time()
            lib.pardiso_64(pt, C.byref(one), C.byref(one), C.byref(mtype), C.byref(L(ph)), C.byref(nn), ptr(a),
                           ptr(ia), ptr(ja), C.byref(idum), C.byref(one), ptr(iparm), C.byref(msglvl), ptr(b),
                           ptr(x), C.byref(err))
            phases[str(ph)] = time.time() - t1
            if err.value != 0:
                res["error"] = f"PARDISO phase {ph} error {err.value}"
                break
        res.update(solver="MKL PARDISO (pardiso_64, sequential)", library=name, factor_nnz=int(iparm[17]),
                   seconds_per_phase={k: phases[k] for k in ("11", "22", "33") if k in phases})
        seconds = sum(phases.get(k, 0.0) for k in ("11", "22", "33"))
    else:
        import scipy.sparse as sp
        import scipy.sparse.linalg as spla
        A = sp.csr_matrix((a, ja, ia), shape=(n, n)).tocsc()
        lu = spla.splu(A, permc_spec="COLAMD")
        x = lu.solve(b)
        seconds = time.time() - t0
        res.update(solver="SciPy SuperLU (COLAMD); no MKL runtime on this host", factor_nnz=int(lu.nnz))
    Ax = np.zeros(n)
    for r in range(0, n, 65536):  # b - A x in CSR without SciPy
        e = min(n, r + 65536)
        rows = np.repeat(np.arange(r, e), np.diff(ia[r: e + 1]))
        np.add.at(Ax, rows, a[ia[r]: ia[e]] * x[ja[ia[r]: ia[e]]])
    res.update(seconds=seconds, cores=1, true_relres=float(np.linalg.norm(b - Ax) / np.linalg.norm(b)))
    print(json.dumps(res), flush=True)


def direct_solve_child(ia, ja, a, b, timeout=300):
    """direct_solve in a child process (a fault in the library cannot end the bench)."""
    import subprocess
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "system.npz")
        np.savez(path, ia=ia, ja=ja, a=a, b=b)
        env = dict(os.environ, MKL_THREADING_LAYER="SEQUENTIAL", MKL_NUM_THREADS="1")
        try:
            p = subprocess.run([sys.executable, os.path.abspath(__file__), "--direct-solve", path], env=env,
                               capture_output=True, text=True, timeout=timeout)
        except subprocess.TimeoutExpired:
            return {"error": f"no result within {timeout} s"}
        lines = [ln for ln in p.stdout.splitlines() if ln.startswith("{")]
        if p.returncode != 0 or not lines:
            return {"error": f"exit {p.returncode}: {p.stderr[-300:]}"}
        return json.loads(lines[-1])


def copy_probe(mib):
    """Five device-to-device copies of a known size: under rocprofv3 --pmc the
    FETCH_SIZE / WRITE_SIZE of this copy kernel against mib MiB read and written
    gives the scale of those counters on this GPU."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": n * 4, "copies": 5}), flush=True)


def hbm_probe(mib=2048, reps=10):
    """Practical HBM bandwidth of plain streaming kernels on this GPU, timed with
    events: a read-only reduction over mib MiB, and a device copy of mib / 2 MiB."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    half = a[: n // 2]
    b = torch.empty_like(half)
    out = {}
    for name, fn, nbytes in (("read_sum", lambda: a.sum(), n * 4),
                             ("copy", lambda: b.copy_(half), n * 4)):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        out[name] = {"bytes": nbytes, "ms": ms, "GBps": nbytes / (ms * 1e-3) / 1e9}
    print(json.dumps({"hbm_probe": out}), flush=True)


def measured_traffic(kernel, alt=False):
    """HBM bytes per launch of the explicit kernel on the 60^3 box, from the
    rocprofv3 FETCH_SIZE / WRITE_SIZE runs summarised in
    profiles/k_explicit_box3d_60_traffic.md (copy-calibrated); None if that
    kernel form was not profiled.  Forms that reverse the sweep every other step launch
    the same kernel: their figures are keyed '<kernel> alternating'."""
    path = os.path.join(ROOT, "profiles", "k_explicit_box3d_60_traffic.json")
    try:
        with open(path) as f:
            k = json.load(f)["kernels"].get(kernel + (" alternating" if alt else ""))
    except (OSError, ValueError, KeyError):
        return None
    return k["traffic_bytes"] if k else None


KRYLOV_NO_GRAPH = False  # --krylov-no-graph: launch Krylov iterations one by one (profiling runs)


def solve_opts(pdx, **kw):
    return pdx.solve_opts(no_graph=KRYLOV_NO_GRAPH, **kw)


def gnn_bench(pdx, cx, model):
    """Inference of the learned smoother's graph network (4 message-passing layers, width 32) on
    the node graph of a model, FP32 and BF16-activation MFMA, timed with HIP events (layers,
    head and smoother; the network inputs are built once per K).  Algorithmic bytes: per layer
    the activations read and written once, the edge weights and adjacency once; input layer,
    head and smoother likewise.  FLOPs: per layer 2 N F (2F) + 2 E F, plus input and head."""
    net = pdx.GNN(model.dim, seed=5)
    w = net.params()
    rng = np.random.default_rng(0)
    w[:] = rng.normal(size=w.size).astype(np.float32) * 0.1
    net.set_params(w)
    N, E, F, L, DD, GIN = model.nn, model.nadj, 32, 4, model.dim ** 2, 8
    flops = L * (2 * N * F * 2 * F + 2 * E * F) + 2 * N * GIN * F + 2 * N * (F + DD) * DD
    out = {"workload": f"node graph of the {model.nn}-node model, {E} edges, 4 layers of width 32",
           "flops": flops}
    for name, bf16, sa, peak_flops in (("fp32", False, 4, 157.3e12), ("bf16_mfma", True, 2, 2.5e15)):
        net.infer(cx, bf16=bf16)
        ms = min(net.infer(cx, bf16=bf16)[1] for _ in range(5))
        nbytes = (L * (2 * N * F * sa + E * 8 + N * 4) + N * GIN * 4 + N * F * sa
                  + N * F * sa + N * DD * 4 * 2 + N * DD * (4 + 4 + 8))
        out[name] = {"ms": ms, "alg_bytes": nbytes, "GBps": nbytes / (ms * 1e-3) / 1e9,
                     "hbm_frac": nbytes / (ms * 1e-3) / HBM_PEAK, "TFLOPs": flops / (ms * 1e-3) / 1e12,
                     "compute_frac": flops / (ms * 1e-3) / peak_flops}
    return out


def learned_bench(pdx, W, warm_seeds=range(8), train_seeds=range(8, 40), held_seeds=tuple(range(448, 512)),
                  plain_seeds=4):
    """The metric's learned-vs-plain comparison on the edge-cracked plate variants (501 x 251
    nodes, crack length and angle drawn per seed): a network trained on the GPU (the library's
    forward, backward and Adam kernels) -- 4 epochs x 10 Adam steps on 121 x 61 plates of
    warm_seeds, then one pass of 5 steps over 501 x 251 plates of train_seeds -- is applied with
    BF16 MFMA inference to the 64 held-out seeds 448-511 and solved to a 1e-8 true residual in FP64
    and in FP32 with FP64 refinement (inner solves to 1e-5: on these plates fewer refinement
    sweeps pay for the tighter inner tolerance; the Newmark runs keep the default 1e-4), beside
    two-level without the network and plain FP64 Jacobi (on the first plain_seeds held-out
    plates).  Each solve is timed on its second call (the first builds the coarse space).  The
    summary gives the mean iterations and solve times over the held-out plates and what the
    network gains over the same two-level solve without it."""
    t0 = time.time()
    net = pdx.GNN(2, seed=3)
    for epoch in range(4):
        for k, s in enumerate(warm_seeds):
            cx = pdx.Context(pdx.Model(W.crack2d(seed=s, nx=121, ny=61, dx=0.1 / 120)))
            cx.assemble(pdx.PDX_MODE_STATIC)
            net.train(cx, steps=10, n_probes=8, probe_seed=100 * epoch + k, lr=3e-3, coarse_agg=8)
    t_warm = time.time() - t0
    for k, s in enumerate(train_seeds):
        cx = pdx.Context(pdx.Model(W.crack2d(seed=s)))
        cx.assemble(pdx.PDX_MODE_STATIC)
        net.train(cx, steps=5, n_probes=4, probe_seed=1000 + k, lr=1e-3)
    out = {"train": f"4 epochs x 10 Adam steps on {len(warm_seeds)} seeds at 121x61 ({t_warm:.1f} s), then 5 steps on "
                    f"each of {len(train_seeds)} seeds ({min(train_seeds)}-{max(train_seeds)}) at 501x251",
           "train_s": time.time() - t0, "held_out_seeds": f"{min(held_seeds)}-{max(held_seeds)}"}
    cfgs = (("plain_fp64_jacobi", dict(precision=0, precond=1, tol=1e-8, max_iter=100000)),
            ("fp64_two_level", dict(precision=0, precond=pdx.PDX_PREC_TWO_LEVEL, tol=1e-8)),
            ("fp32_ir_two_level", dict(precision=1, precond=pdx.PDX_PREC_TWO_LEVEL, tol=1e-8, inner_tol=1e-5)),
            ("fp64_learned", dict(precision=0, precond=pdx.PDX_PREC_LEARNED, tol=1e-8)),
            ("fp32_ir_learned", dict(precision=1, precond=pdx.PDX_PREC_LEARNED, tol=1e-8, inner_tol=1e-5)))
    per = {name: {"iterations": [], "ms": []} for name, _ in cfgs}
    per["bf16_inference_ms"] = []
    errors = {}
    for i, s in enumerate(held_seeds):
        cx = pdx.Context(pdx.Model(W.crack2d(seed=s)))
        cx.assemble(pdx.PDX_MODE_STATIC)
        _, infer_ms = net.infer(cx, bf16=True)
        per["bf16_inference_ms"].append(infer_ms)
        for name, kw in cfgs:
            if name == "plain_fp64_jacobi" and i >= plain_seeds:
                continue
            try:
                cx.static_solve(solve_opts(pdx, **kw))
                _, st = cx.static_solve(solve_opts(pdx, **kw))
                if st["true_relres"] > 1e-8 * 1.2:
                    raise RuntimeError(f"true residual {st['true_relres']:.2e}")
                per[name]["iterations"].append(int(st["iterations"]))
                per[name]["ms"].append(st["seconds"] * 1e3)
            except Exception as e:  # reported, not hidden
                errors[f"seed {s} {name}"] = str(e)
    summ = {}
    for name, _ in cfgs:
        if per[name]["ms"]:
            summ[name] = {"plates": len(per[name]["ms"]), "mean_iterations": float(np.mean(per[name]["iterations"])),
                          "mean_ms": float(np.mean(per[name]["ms"]))}
    for prec in ("fp64", "fp32_ir"):
        a, b = summ.get(f"{prec}_two_level"), summ.get(f"{prec}_learned")
        if a and b:
            summ[f"{prec}_network_gain"] = {"iterations": a["mean_iterations"] / b["mean_iterations"],
                                            "time": a["mean_ms"] / b["mean_ms"]}
            ia, ib = np.array(per[f"{prec}_two_level"]["iterations"]), np.array(per[f"{prec}_learned"]["iterations"])
            if len(ia) == len(ib):
                summ[f"{prec}_network_gain"]["plates_fewer_iterations"] = int((ib < ia).sum())
                summ[f"{prec}_network_gain"]["plates_more_iterations"] = int((ib > ia).sum())
    if "plain_fp64_jacobi" in summ and "fp32_ir_learned" in summ:
        summ["fp32_ir_learned_vs_plain_time"] = summ["plain_fp64_jacobi"]["mean_ms"] / summ["fp32_ir_learned"]["mean_ms"]
        summ["fp64_learned_vs_plain_time"] = summ["plain_fp64_jacobi"]["mean_ms"] / summ["fp64_learned"]["mean_ms"]
    summ["mean_bf16_inference_ms"] = float(np.mean(per["bf16_inference_ms"]))
    out["summary"] = summ
    out["per_plate_iterations"] = {name: per[name]["iterations"] for name, _ in cfgs}
    if errors:
        out["errors"] = errors
    return out


def static_plate_bench(pdx, W):
    """The reference's CPU-runnable static case: a 2D plane-stress plate 100 x 50 mm
    (201 x 101 nodes, 40,299 eq) with a central PDLSM strip under displacement-controlled
    tension, model set-up on the host, K and F assembled on the GPU and solved to a 1e-8
    true residual (FP64 two-level BiCGStab, and plain FP64 Jacobi BiCGStab beside it).
    The reference spends 0.36 s assembling and 8.2 s in PARDISO on one core for it."""
    t0 = time.time()
    mp = pdx.Model(W.plate2d())
    setup_s = time.time() - t0
    cx = pdx.Context(mp)
    asm_ms = min(cx.assemble_timed() for _ in range(3))
    res = {"workload": f"plate2d 201x101 nodes, PD strip, neq {mp.neq}, nnz {mp.nnz}", "host_setup_s": setup_s,
           "assembly_ms": asm_ms}
    for name, o in (("fp64_two_level_bicgstab", solve_opts(pdx, precision=0, precond=pdx.PDX_PREC_TWO_LEVEL, tol=1e-8)),
                    ("plain_fp64_jacobi_bicgstab", solve_opts(pdx, precision=0, precond=1, tol=1e-8, max_iter=100000))):
        try:
            u, st0 = cx.static_solve(o)
            u, st = cx.static_solve(o)
            res[name] = {"seconds": st["seconds"], "seconds_first_call": st0["seconds"], "iterations": st["iterations"],
                         "true_relres": st["true_relres"]}
        except Exception as e:  # reported, not hidden
            res[name] = {"error": str(e)}
    return res


def newmark_bench(pdx, W, n=100, steps=50, coarse_agg=10):
    """The penny-crack Newmark configuration: 100^3 nodes (dx = 1 mm), penny
    crack of radius 0.2 L between two node layers, PD region the central 30^3
    nodes, bottom fixed, top traction ramped linearly to 5 MPa over the 50 steps,
    beta = 1/4, gamma = 1/2, dt = 10x the explicit critical step, a_{n+1}
    warm-started; each step solves (M + beta dt^2 K) a = F - K U_N to a 1e-8 true
    residual: plain (FP64 node-block-Jacobi BiCGStab), FP32 BiCGStab with FP64
    refinement and node-block Jacobi, the same with the two-level hybrid
    preconditioner (aggregates of coarse_agg^3 nodes, built once per run), the
    same with the learned smoother (a 3D network trained on the Newmark operator
    of smaller penny boxes, predicting the smoother of this run's K_eff), and
    the two block-Jacobi runs again with the SpMVs on mirrored storage (same
    results bit for bit).  The two-level and learned runs read mirrored storage
    (the faster of the two for the 3-DOF SpMV, same bits).  The other runs' displacements are compared with the
    plain run's."""
    m0 = W.penny3d(n=n)
    cx0 = pdx.Context(pdx.Model(m0))
    lmax = cx0.max_eig(200)
    dt = 10.0 * 2.0 / np.sqrt(lmax)
    del cx0
    m = W.penny3d(n=n, dt=dt, ramp_steps=steps)
    t0 = time.time()
    mp = pdx.Model(m)  # the reference's set-up: volumes, cells, PD/FE split, families, node graph, CSR rows
    setup_s = time.time() - t0
    # the same set-up with the family search and the node graph on the GPU
    pdx.set_setup_device(0)
    try:
        t0 = time.time()
        mg = pdx.Model(m)
        gpu_setup_s = time.time() - t0
    finally:
        pdx.set_setup_device(-1)
    fh, fg = mp.families(), mg.families()
    same = bool(np.array_equal(fh["off"], fg["off"]) and np.array_equal(fh["mem"], fg["mem"])
                and all(np.array_equal(a, b) for a, b in zip(mp.adjacency(), mg.adjacency())))
    del mg, fh, fg
    t0 = time.time()
    cx = pdx.Context(mp)  # device structures: uploads, block maps, column classes, mirrored index
    ctx_s = time.time() - t0
    del cx
    out = {"workload": f"penny3d {n}^3 nodes, PD 30^3, dt = 10 dt_crit = {dt:.3e} s, {steps} Newmark steps",
           "neq": mp.neq, "host_setup_s": setup_s, "gpu_setup_s": gpu_setup_s, "gpu_setup_identical": same,
           "context_setup_s": ctx_s}
    runs = {"plain_fp64_blockjacobi": solve_opts(pdx, precision=0, precond=pdx.PDX_PREC_BLOCK_JACOBI, tol=1e-8,
                                                     max_iter=100000),
            "fp32_ir_blockjacobi": solve_opts(pdx, precision=1, precond=pdx.PDX_PREC_BLOCK_JACOBI, tol=1e-8,
                                                  max_iter=100000),
            "fp32_ir_two_level": solve_opts(pdx, precision=1, precond=pdx.PDX_PREC_TWO_LEVEL, tol=1e-8,
                                                max_iter=100000, coarse_agg=coarse_agg, mirrored=True),
            "fp32_ir_learned": solve_opts(pdx, precision=1, precond=pdx.PDX_PREC_LEARNED, tol=1e-8,
                                              max_iter=100000, coarse_agg=coarse_agg, mirrored=True),
            "plain_fp64_blockjacobi_mirrored": solve_opts(pdx, precision=0, precond=pdx.PDX_PREC_BLOCK_JACOBI,
                                                              tol=1e-8, max_iter=100000, mirrored=True),
            "fp32_ir_blockjacobi_mirrored": solve_opts(pdx, precision=1, precond=pdx.PDX_PREC_BLOCK_JACOBI,
                                                           tol=1e-8, max_iter=100000, mirrored=True)}
    # the learned run's network: trained on the Newmark operator K_eff of smaller penny boxes at the
    # same spacing and time step, then attached to the run's context, which predicts the smoother
    # of its K_eff with it (inference inside the timed run)
    t1 = time.time()
    net = pdx.GNN(3, seed=5)
    for epoch in range(3):
        for k, nt in enumerate((24, 32, 40)):
            tr = pdx.Context(pdx.Model(W.penny3d(n=nt, pd_block=nt // 3, dt=dt, ramp_steps=steps)))
            tr.gnn_operator(1)
            net.train(tr, steps=10, n_probes=4, probe_seed=10 * epoch + k, lr=3e-3, coarse_agg=coarse_agg)
            del tr
    out["learned_train"] = {"models": "penny3d 24^3, 32^3, 40^3 (PD third), Newmark operator", "epochs": 3,
                            "adam_steps_per_model": 10, "seconds": time.time() - t1}
    us = {}
    for name, opts in runs.items():
        cx = pdx.Context(mp)
        if name == "fp32_ir_learned":
            net.attach(cx)
        t1 = time.time()
        try:
            u, v, a, st = cx.newmark_run(steps, opts)
        except Exception as e:  # reported, not hidden
            out[name] = {"error": str(e)}
            continue
        wall = time.time() - t1
        us[name] = u
        out[name] = {"seconds": wall, "ms_per_step": wall / steps * 1e3, "krylov_iterations": st["iterations"],
                     "refinements": st["refinements"], "last_true_relres": st["true_relres"],
                     "dof_steps_per_s": mp.neq * steps / wall, "u_max": float(np.abs(u).max())}
        if st["mirrored_used"]:
            out[name]["mirrored_storage"] = True
        del cx
    ref = us.get("plain_fp64_blockjacobi")
    for name in us:
        if name != "plain_fp64_blockjacobi" and ref is not None:
            out[name]["u_rel_diff_vs_plain"] = float(np.linalg.norm(us[name] - ref) / max(np.linalg.norm(ref), 1e-300))
            out[name]["speedup_vs_plain"] = out["plain_fp64_blockjacobi"]["seconds"] / out[name]["seconds"]
    return out


def family_batch_bench(pdx, W):
    """The 2,000,376-node PD cloud (126^3 jittered grid, per-node horizons
    U(2.5, 3.5) dx, members by the OR-of-horizons rule): shape tensor + Cholesky factor (FP64)
    and bond blocks H (FP32) of every family, one family per block vs the
    autotuned persistent launch.  Algorithmic bytes count every array element
    once (coordinates, volumes, horizons, offsets, member ids in; A, L, H out).
    Algorithmic FP64 flops: 3 per member per symmetric shape-tensor entry (45
    in 3D), p^3 / 3 per Cholesky, 2 p^3 for the p columns of S^-1, 2 p d^2 per
    bond block."""
    E, nu = 70e9, 0.33
    lam, mu = E * nu / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))
    t0 = time.time()
    c = W.family_batch()
    fb = pdx.FamilyBatch(c["x"], c["vol"], c["delta"], nlf=0.33, lam=lam, mu=mu)
    setup_s = time.time() - t0
    N, nb, P = fb.n_fam, fb.n_bond, 9
    alg_bytes = N * (24 + 8 + 8 + 8) + nb * 4 + N * (2 * P * P * 8) + nb * 9 * 4
    flops = 45 * 3 * (nb - N) + (P ** 3 // 3 + 2 * P ** 3) * N + 2 * P * 9 * (nb - N)
    fb.run(0)
    block_ms = min(fb.run(0) for _ in range(3))
    fb.run(-1)  # times every launch form (one family per block and the persistent shapes), keeps the fastest
    best, tune = fb.tuning()
    tuned_ms = min(fb.run(-1) for _ in range(3))
    pers = [(float(t), k) for k, t in enumerate(tune) if k > 0 and t > 0]
    best_pers = min(pers)[1] if pers else None
    pers_ms = min(fb.run(best_pers) for _ in range(3)) if best_pers is not None else None
    return {"workload": "126^3 jittered node cloud (2,000,376 nodes), horizons U(2.5,3.5) dx, OR-rule families",
            "n_family": N, "n_bond": nb, "mean_family": nb / N, "max_family": fb.max_family,
            "host_setup_s": setup_s, "one_family_per_block_ms": block_ms, "autotuned_ms": tuned_ms,
            "autotuned_variant": best, "best_persistent_ms": pers_ms, "best_persistent_variant": best_pers,
            "tuning_ms": [float(t) for t in tune],
            "speedup_autotuned_vs_one_per_block": block_ms / tuned_ms, "families_per_s": N / (tuned_ms * 1e-3),
            "algorithmic_GBps": alg_bytes / (tuned_ms * 1e-3) / 1e9, "fp64_TFLOPs": flops / (tuned_ms * 1e-3) / 1e12,
            "alg_bytes": alg_bytes, "flops": flops}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--n", type=int, default=60)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-solve", action="store_true")
    ap.add_argument("--no-fb", action="store_true", help="skip the 2M-node family batch")
    ap.add_argument("--no-newmark", action="store_true", help="skip the 100^3 penny-crack Newmark run")
    ap.add_argument("--fb-only", action="store_true",
                    help="only run the 2M-node family batch and print its figures (for kernel profiling)")
    ap.add_argument("--krylov-no-graph", action="store_true",
                    help="launch the Krylov iterations' kernels one by one instead of replaying captured graphs "
                         "(same results; for kernel-trace profiling)")
    ap.add_argument("--stress-only", action="store_true",
                    help="only time the nodal stress recovery of the box (for kernel profiling)")
    ap.add_argument("--hbm-probe", action="store_true",
                    help="only time a streaming read and a copy (practical HBM bandwidth) and exit")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only copy a MIB-MiB device buffer 5 times (HBM counter calibration) and exit")
    ap.add_argument("--direct-solve", default="", metavar="NPZ",
                    help="only solve the CSR system in NPZ on the CPU (PARDISO, else SuperLU) and exit")
    args = ap.parse_args()
    if args.direct_solve:
        direct_solve(args.direct_solve)
        return
    global KRYLOV_NO_GRAPH
    KRYLOV_NO_GRAPH = args.krylov_no_graph
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.hbm_probe:
        hbm_probe()
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: the whole mesh lives in one HBM and the library has no "
              "multi-GPU path; run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    pdx = load_pdx()
    sys.path.insert(0, PKG)
    import workloads as W

    if args.stress_only:
        model = pdx.Model(W.box3d(n=args.n))
        cx = pdx.Context(model)
        asm = [cx.assemble_timed() for _ in range(6)]  # the first also allocates the staging
        uall = np.random.default_rng(0).uniform(-1e-6, 1e-6, model.nn * model.dim)
        ms = [cx.node_stresses_timed(uall) for _ in range(max(args.steps, 1))]
        sb = stress_bytes(model)
        print(json.dumps({"assembly_ms": min(asm[1:]), "stress_recovery": {"ms": min(ms), "alg_bytes": sb,
                                                                           "GBps": sb / (min(ms) * 1e-3) / 1e9}}),
              flush=True)
        return
    if args.fb_only:
        print(json.dumps({"family_batch": family_batch_bench(pdx, W)}), flush=True)
        return

    t0 = time.time()
    mesh = W.box3d(n=args.n)
    model = pdx.Model(mesh)
    log(f"model: nodes {model.nn} eq {model.neq} nnz {model.nnz} families {model.nfam} "
        f"max family {model.max_family} ({time.time() - t0:.1f} s)")
    ctxs = {}
    for prec in (0, 1):
        ctxs[prec] = pdx.Context(model)
    cx = ctxs[0]
    asm_ms = [cx.assemble_timed() for _ in range(3)]
    lmax = cx.max_eig(300)
    dt = 0.8 * 2.0 / np.sqrt(lmax)
    log(f"assembly {min(asm_ms):.2f} ms; lambda_max {lmax:.4e}; dt {dt:.4e}")
    res = {}
    for prec in (0, 1):
        c = ctxs[prec]
        c.set_dt(dt)
        c.explicit_steps(args.warmup, precision=prec)
        c.explicit_steps(args.steps, precision=prec)  # timed with HIP events on the context stream
        total_ms, step_ms = c.last_timing()
        res[prec] = dict(total_ms=total_ms, step_ms=step_ms, dofs=model.neq * args.steps / (total_ms * 1e-3))
        log(f"precision {prec}: {step_ms * 1e3:.2f} us/step, {res[prec]['dofs']:.4e} DOF-steps/s")
    u64, _ = ctxs[0].explicit_run(0)
    u32, _ = ctxs[1].explicit_run(0, precision=1)
    drift = float(np.linalg.norm(u32 - u64) / max(np.linalg.norm(u64), 1e-300))
    # per-step nodal stress recovery (the reference recovers stresses after every explicit step,
    # pdsolve.cpp:6062), timed on its own: PD strains from the bond weights K's assembly left,
    # FE Gauss-point stresses extrapolated to the corners, node averages
    uall = np.random.default_rng(0).uniform(-1e-6, 1e-6, model.nn * model.dim)
    st_ms = min(cx.node_stresses_timed(uall) for _ in range(20))
    sb = stress_bytes(model)
    stress = {"workload": "nodal stresses of the same box (synthetic DOF values)", "ms": st_ms,
              "alg_bytes": sb, "GBps": sb / (st_ms * 1e-3) / 1e9, "roofline_frac": sb / (st_ms * 1e-3) / HBM_PEAK,
              "staged_bytes": int((model.nodes()["algo"] == 2).sum()) * 8 * 6 * 8 * 2}
    log(f"stress recovery {st_ms * 1e3:.1f} us ({stress['GBps']:.0f} GB/s algorithmic)")
    nblk = model.nadj
    classes = column_classes(model)
    n_own = own_blocks(model)
    tf = lambda b: "true" if b else "false"
    roof = {}
    for prec in (0, 1):
        lanes, nt, xm, soa, rel, sym, tune_ms, alt, form = ctxs[prec].explicit_tuning_full(prec)
        bytes_step = step_bytes(model, nblk, prec, classes if rel else None, n_own if sym else None)
        achieved = bytes_step / (res[prec]["step_ms"] * 1e-3)
        ty = "float" if prec else "double"
        kname = (f"pdx::k_explicit<{model.dim}, {ty}, {lanes}, {tf(nt)}, {tf(xm)}, {tf(soa)}, {tf(rel)}, "
                 f"{tf(sym)}, false>")
        roof[prec] = {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                      "frac": achieved / HBM_PEAK,
                      "traffic": measured_traffic(kname, alt) if args.n == 60 else None,
                      "kernel": kname + f" (node-block SpMV + central-difference update, {'FP32' if prec else 'FP64'} "
                                        f"values{', sweep reversed every other step' if alt else ''})",
                      "form": form,
                      "tuner_ms_per_form": [float(t) for t in tune_ms],
                      "bytes_per_launch": bytes_step,
                      "column_classes": {"classes": classes[0], "offsets": classes[1]},
                      "mirrored_storage": {"own_blocks": n_own, "blocks": nblk},
                      "scalar_csr_bytes_per_step": scalar_csr_step_bytes(model, prec)}
    # the headline is the reference's precision (FP64 values, FP64 accumulation, FP64 state);
    # the FP32-force / FP64-state figure and its drift against FP64 follow as secondary keys
    head = {"metric": METRIC, "value": res[0]["dofs"], "unit": "DOF-steps/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": res[0]["step_ms"], "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": f"box3d {args.n}^3 nodes hex8, central 20^3-node PDLSM cube, jitter 0.2dx seed 0, "
                                   f"m=3.015, E=70GPa nu=0.33 rho=2700, 1 MPa top traction, lumped mass, "
                                   f"explicit central difference dt=0.8 dt_crit, FP64 K values and state",
                       "neq": model.neq, "nnz": model.nnz, "dt": dt, "parallelism": "single GPU"},
            "roofline": roof[0]}
    fp32 = {"value": res[1]["dofs"], "unit": "DOF-steps/s", "ms_per_step": res[1]["step_ms"],
            "dtype": "f32 force, f64 state", "roofline": roof[1], "rel_diff_vs_fp64": None}
    out = {"stress_recovery": stress, "assembly_ms": min(asm_ms)}
    if not args.no_solve:
        sm = W.crack2d(seed=0)
        smod = pdx.Model(sm)
        scx = pdx.Context(smod)
        solves = {}
        cfgs = (("plain_fp64_jacobi_bicgstab", solve_opts(pdx, precision=0, precond=1, tol=1e-8, max_iter=100000)),
                ("fp32_ir_blockjacobi_bicgstab", solve_opts(pdx, precision=1, precond=2, tol=1e-8, max_iter=100000)),
                ("fp64_two_level_bicgstab", solve_opts(pdx, precision=0, precond=pdx.PDX_PREC_TWO_LEVEL, tol=1e-8)),
                ("fp32_ir_two_level_bicgstab", solve_opts(pdx, precision=1, precond=pdx.PDX_PREC_TWO_LEVEL, tol=1e-8,
                                                          inner_tol=1e-5)),
                ("fp64_ilu0_bicgstab", solve_opts(pdx, precision=0, precond=pdx.PDX_PREC_ILU0, tol=1e-8)),
                ("fp32_ir_ilu0_bicgstab", solve_opts(pdx, precision=1, precond=pdx.PDX_PREC_ILU0, tol=1e-8)),
                ("plain_fp64_jacobi_bicgstab_mirrored", solve_opts(pdx, precision=0, precond=1, tol=1e-8,
                                                                   max_iter=100000, mirrored=True)),
                ("fp64_two_level_bicgstab_mirrored", solve_opts(pdx, precision=0, precond=pdx.PDX_PREC_TWO_LEVEL,
                                                                tol=1e-8, mirrored=True)))
        for name, o in cfgs:
            try:
                # first call builds the coarse space of the two-level kinds (the
                # factor of ILU(0)); the second reuses it (repeated solves with one operator)
                u, st0 = scx.static_solve(o)
                u, st = scx.static_solve(o)
                solves[name] = {"seconds": st["seconds"], "seconds_first_call": st0["seconds"],
                                "iterations": st["iterations"], "refinements": st["refinements"],
                                "true_relres": st["true_relres"]}
                if st["mirrored_used"]:
                    solves[name]["mirrored_storage"] = True
            except Exception as e:  # reported, not hidden
                solves[name] = {"error": str(e)}
            log(name, solves[name])
        out["solve_to_1e-8"] = {"workload": f"crack2d 501x251 seed 0, neq {smod.neq}", **solves}
        if not args.no_cpu:
            # the same system solved on one host core as the reference solves it (sparse direct)
            K, F = scx.assemble(pdx.PDX_MODE_STATIC)
            ia, ja = smod.csr()
            cpu = direct_solve_child(ia, ja, K, F)
            gpu = solves.get("fp64_two_level_bicgstab", {}).get("seconds")
            if gpu and "seconds" in cpu:
                cpu["gpu_fp64_two_level_speedup"] = cpu["seconds"] / gpu
            out["solve_to_1e-8"]["cpu_direct_solve"] = cpu
            log("cpu direct solve", cpu)
        try:
            out["gnn_inference"] = gnn_bench(pdx, scx, smod)
        except Exception as e:  # reported, not hidden
            out["gnn_inference"] = {"error": str(e)}
        log("gnn inference", out["gnn_inference"])
        try:
            out["learned_vs_plain"] = learned_bench(pdx, W)
        except Exception as e:  # reported, not hidden
            out["learned_vs_plain"] = {"error": str(e)}
        log("learned vs plain", out["learned_vs_plain"])
    if not args.no_solve:
        out["static_plate2d"] = static_plate_bench(pdx, W)
        log("static plate2d", out["static_plate2d"])
    if not args.no_newmark:
        out["newmark_penny"] = newmark_bench(pdx, W)
        log("newmark", out["newmark_penny"])
    if not args.no_fb:
        out["family_batch"] = family_batch_bench(pdx, W)
        log("family batch", out["family_batch"])
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline()
        log("cpu baseline", out["cpu_baseline"])
    # the FP32 trajectory's drift from the FP64 one after the same warm-up + timed steps
    fp32["rel_diff_vs_fp64"] = drift
    # detail blocks first, the headline keys last (the end of the line is what a tail shows)
    out.update(head)
    out["fp32_force"] = {k: v for k, v in fp32.items() if k != "roofline"}
    out["fp32_force"]["roofline_frac"] = fp32["roofline"]["frac"]
    out["fp32_force"]["traffic"] = fp32["roofline"]["traffic"]
    out["fp32_force_roofline"] = fp32["roofline"]
    if "cpu_baseline" in out:
        out["cpu_baseline"] = out.pop("cpu_baseline")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
