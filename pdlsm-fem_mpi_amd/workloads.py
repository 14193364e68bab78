"""Synthetic PDLSM-FEM workloads in the reference's raw-mesh form.

Each generator returns a dict with the fields datModel::readdata
(datModel.cpp:58-326) fills from a .lis file plus the SETSOLVING /
EBC / NBC / VEBC / VNBC command values (fioFile.cpp:69-462):

  dim, proType, E, nu, rho, x[nn,3], econ[ne,8] (1-based node ids),
  nen[ne], algo[ne] (1 = PD element, 2 = FE element),
  ebc: [{nid, dof, val, vel, vary}], nbc: [{con[nele, 2|4], val, vel, vary}],
  cracks[nc,3,3], fac (horizon factor m), nlf (non-local factor a), me, dt,
  beta, gamma; optional run settings lumped_mass (LUMPMASS ON, default
  off as in the reference), newmark, fc, topk, sigma_ult.

Workloads (sizes as named):
  plate2d   2D plane-stress plate 100 x 50 mm, dx = 0.5 mm (201 x 101 nodes,
            20,000 quad4), PD strip of elements with centre x in (40, 60) mm,
            m = 3.015, E = 200 GPa, nu = 0.3; left edge fixed, right edge
            UX = 0.01 mm.  Static.
  box3d     3D box of n^3 nodes (default 60^3, dx = 1 mm, 205,379 hex8),
            interior nodes jittered U(-0.2, 0.2) dx, PD elements whose nodes lie
            in the central 20^3-node block, m = 3.015, E = 70 GPa, nu = 0.33,
            rho = 2700; bottom face UZ = 0 plus a corner pin, 1 MPa traction on
            the top face (quad4 faces).  Explicit central difference, lumped
            mass (LUMPMASS ON).
  crack2d   2D edge-cracked plate 501 x 251 nodes (dx = 0.2 mm), crack from
            the left edge at mid-height, length U[0.1, 0.4] x width, angle
            U[-30, 30] deg per seed; PD band of elements within 3 horizons of
            the crack; bottom UY = 0 + UX pin, top UY = 0.01 mm.  Static.
  penny3d   3D box of n^3 nodes with a penny crack (32-triangle fan) of radius
            0.2 L in the plane z = L/2, PD region = central 30^3 nodes, bottom
            fixed, top traction ramped to 5 MPa over 50 steps.  Newmark.
"""
from __future__ import annotations

import numpy as np


def _quad_grid(nx, ny, dx):
    """Nodes (i, j) numbered x-fastest, 1-based; quad (i,j),(i+1,j),(i+1,j+1),(i,j+1)."""
    ii, jj = np.meshgrid(np.arange(nx), np.arange(ny), indexing="xy")
    x = np.zeros((nx * ny, 3))
    x[:, 0] = (ii.ravel() * dx)
    x[:, 1] = (jj.ravel() * dx)
    nid = lambda i, j: j * nx + i + 1
    ei, ej = np.meshgrid(np.arange(nx - 1), np.arange(ny - 1), indexing="xy")
    ei, ej = ei.ravel(), ej.ravel()
    econ = np.zeros((len(ei), 8), np.int32)
    econ[:, 0] = nid(ei, ej)
    econ[:, 1] = nid(ei + 1, ej)
    econ[:, 2] = nid(ei + 1, ej + 1)
    econ[:, 3] = nid(ei, ej + 1)
    return x, econ, ei, ej


def plate2d(nx=201, ny=101, dx=5e-4, pd_x=(0.04, 0.06), m=3.015, nlf=0.33, ux_right=1e-5,
            E=200e9, nu=0.3, rho=7850.0):
    x, econ, ei, ej = _quad_grid(nx, ny, dx)
    ne = len(econ)
    cx = (ei + 0.5) * dx
    algo = np.where((cx > pd_x[0]) & (cx < pd_x[1]), 1, 2).astype(np.int32)
    left = np.array([j * nx + 1 for j in range(ny)], np.int32)
    right = np.array([j * nx + nx for j in range(ny)], np.int32)
    ebc = [dict(nid=left, dof=0, val=0.0), dict(nid=left, dof=1, val=0.0), dict(nid=right, dof=0, val=ux_right)]
    return dict(dim=2, proType=1, E=E, nu=nu, rho=rho, x=x, econ=econ, nen=np.full(ne, 4, np.int32), algo=algo,
                ebc=ebc, nbc=[], cracks=np.zeros((0, 3, 3)), fac=m, nlf=nlf, me=0.0, dt=1.0, beta=0.25, gamma=0.5)


def _hex_grid(n, dx, jitter, seed):
    rng = np.random.default_rng(seed)
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()
    x = np.stack([i * dx, j * dx, k * dx], axis=1).astype(np.float64)
    if jitter:
        interior = (i > 0) & (i < n - 1) & (j > 0) & (j < n - 1) & (k > 0) & (k < n - 1)
        d = rng.uniform(-jitter, jitter, size=(int(interior.sum()), 3)) * dx
        x[interior] += d
    nid = lambda a, b, c: c * n * n + b * n + a + 1
    ek, ej, ei = np.meshgrid(np.arange(n - 1), np.arange(n - 1), np.arange(n - 1), indexing="ij")
    ei, ej, ek = ei.ravel(), ej.ravel(), ek.ravel()
    econ = np.stack([nid(ei, ej, ek), nid(ei + 1, ej, ek), nid(ei + 1, ej + 1, ek), nid(ei, ej + 1, ek),
                     nid(ei, ej, ek + 1), nid(ei + 1, ej, ek + 1), nid(ei + 1, ej + 1, ek + 1),
                     nid(ei, ej + 1, ek + 1)], axis=1).astype(np.int32)
    return x, econ, (ei, ej, ek), nid


def box3d(n=60, dx=1e-3, jitter=0.2, seed=0, pd_block=20, m=3.015, nlf=0.33, E=70e9, nu=0.33, rho=2700.0,
          traction=1e6, dt=1.0):
    x, econ, (ei, ej, ek), nid = _hex_grid(n, dx, jitter, seed)
    lo = (n - pd_block) // 2
    hi = lo + pd_block - 1  # node index range [lo, hi]
    inside = (ei >= lo) & (ei + 1 <= hi) & (ej >= lo) & (ej + 1 <= hi) & (ek >= lo) & (ek + 1 <= hi)
    algo = np.where(inside, 1, 2).astype(np.int32)
    bottom = np.array([nid(a, b, 0) for b in range(n) for a in range(n)], np.int32)
    ebc = [dict(nid=bottom, dof=2, val=0.0),
           dict(nid=np.array([nid(0, 0, 0)], np.int32), dof=0, val=0.0),
           dict(nid=np.array([nid(0, 0, 0), nid(n - 1, 0, 0)], np.int32), dof=1, val=0.0)]
    a, b = np.meshgrid(np.arange(n - 1), np.arange(n - 1), indexing="xy")
    a, b = a.ravel(), b.ravel()
    top = np.stack([nid(a, b, n - 1), nid(a + 1, b, n - 1), nid(a + 1, b + 1, n - 1), nid(a, b + 1, n - 1)],
                   axis=1).astype(np.int32)
    nbc = [dict(con=top, val=traction)]
    return dict(dim=3, proType=-1, E=E, nu=nu, rho=rho, x=x, econ=econ, nen=np.full(len(econ), 8, np.int32),
                algo=algo, ebc=ebc, nbc=nbc, cracks=np.zeros((0, 3, 3)), fac=m, nlf=nlf, me=0.0, dt=dt,
                beta=0.25, gamma=0.5, lumped_mass=True)


def crack2d(seed=0, nx=501, ny=251, dx=2e-4, m=3.015, nlf=0.33, uy_top=1e-5, E=200e9, nu=0.3, rho=7850.0,
            band=3.0):
    rng = np.random.default_rng(seed)
    W = (nx - 1) * dx
    H = (ny - 1) * dx
    length = rng.uniform(0.1, 0.4) * W
    ang = np.deg2rad(rng.uniform(-30.0, 30.0))
    y0 = 0.5 * H + 0.5 * dx  # between node rows so bonds cross cleanly
    p0 = np.array([-1e-9, y0, 0.0])
    p1 = np.array([length * np.cos(ang), y0 + length * np.sin(ang), 0.0])
    x, econ, ei, ej = _quad_grid(nx, ny, dx)
    cx = (ei + 0.5) * dx
    cy = (ej + 0.5) * dx
    # element-centre distance to the crack segment
    d = p1[:2] - p0[:2]
    t = np.clip(((cx - p0[0]) * d[0] + (cy - p0[1]) * d[1]) / (d @ d), 0.0, 1.0)
    dist = np.hypot(cx - (p0[0] + t * d[0]), cy - (p0[1] + t * d[1]))
    algo = np.where(dist < band * m * dx, 1, 2).astype(np.int32)
    bottom = np.array([i + 1 for i in range(nx)], np.int32)
    top = np.array([(ny - 1) * nx + i + 1 for i in range(nx)], np.int32)
    ebc = [dict(nid=bottom, dof=1, val=0.0), dict(nid=np.array([1], np.int32), dof=0, val=0.0),
           dict(nid=top, dof=1, val=uy_top)]
    cracks = np.zeros((1, 3, 3))
    cracks[0, 0] = p0
    cracks[0, 1] = p1
    return dict(dim=2, proType=1, E=E, nu=nu, rho=rho, x=x, econ=econ, nen=np.full(len(econ), 4, np.int32),
                algo=algo, ebc=ebc, nbc=[], cracks=cracks, fac=m, nlf=nlf, me=0.0, dt=1.0, beta=0.25,
                gamma=0.5, crack_len=length, crack_angle=ang)


def penny3d(n=100, dx=1e-3, pd_block=30, radius_frac=0.2, nfan=32, m=3.015, nlf=0.33, E=70e9, nu=0.33,
            rho=2700.0, t_final=5e6, ramp_steps=50, dt=1e-7):
    x, econ, (ei, ej, ek), nid = _hex_grid(n, dx, 0.0, 0)
    L = (n - 1) * dx
    lo = (n - pd_block) // 2
    hi = lo + pd_block - 1
    inside = (ei >= lo) & (ei + 1 <= hi) & (ej >= lo) & (ej + 1 <= hi) & (ek >= lo) & (ek + 1 <= hi)
    algo = np.where(inside, 1, 2).astype(np.int32)
    bottom = np.array([nid(a, b, 0) for b in range(n) for a in range(n)], np.int32)
    ebc = [dict(nid=bottom, dof=d, val=0.0) for d in range(3)]
    a, b = np.meshgrid(np.arange(n - 1), np.arange(n - 1), indexing="xy")
    a, b = a.ravel(), b.ravel()
    top = np.stack([nid(a, b, n - 1), nid(a + 1, b, n - 1), nid(a + 1, b + 1, n - 1), nid(a, b + 1, n - 1)],
                   axis=1).astype(np.int32)
    nbc = [dict(con=top, val=0.0, vel=t_final / (ramp_steps * dt), vary=True)]
    # crack plane halfway between two node layers (a plane through nodes is degenerate)
    c = np.array([0.5 * L, 0.5 * L, (n // 2 - 0.5) * dx])
    R = radius_frac * L
    cracks = np.zeros((nfan, 3, 3))
    for q in range(nfan):
        a0, a1 = 2 * np.pi * q / nfan, 2 * np.pi * (q + 1) / nfan
        cracks[q, 0] = c
        cracks[q, 1] = c + R * np.array([np.cos(a0), np.sin(a0), 0.0])
        cracks[q, 2] = c + R * np.array([np.cos(a1), np.sin(a1), 0.0])
    return dict(dim=3, proType=-1, E=E, nu=nu, rho=rho, x=x, econ=econ, nen=np.full(len(econ), 8, np.int32),
                algo=algo, ebc=ebc, nbc=nbc, cracks=cracks, fac=m, nlf=nlf, me=0.0, dt=dt, beta=0.25,
                gamma=0.5)


def family_batch(n=126, jitter=0.3, s_range=(2.5, 3.5), seed=0):
    """2M-node family batch in [0,1]^3: jittered grid, per-node horizon
    delta_k = s_k * dx with s_k ~ U(2.5, 3.5)."""
    rng = np.random.default_rng(seed)
    dx = 1.0 / (n - 1)
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    x = np.stack([i.ravel(), j.ravel(), k.ravel()], axis=1).astype(np.float64) * dx
    x += rng.uniform(-jitter, jitter, size=x.shape) * dx
    s = rng.uniform(s_range[0], s_range[1], size=len(x))
    return dict(x=x, delta=s * dx, vol=np.full(len(x), dx ** 3), dx=dx)
