"""Grids and seeded synthetic inputs for the benchmark configurations.

Grids follow the reference's mesh generator (src/class_mesh/class_mesh.f90:112-335):
uniform, the live asin/cos stretching (:181-200) and the tanh stretching
present there as a comment (:203).  Random numbers come from SplitMix64
(published constants) -> 53-bit uniforms -> Box-Muller normals, so any
implementation of the same definition reproduces the inputs bit for bit.
"""
from __future__ import annotations

import numpy as np


def grid_uniform(n: int, a: float = 0.0, b: float = 1.0) -> np.ndarray:
    return a + np.arange(n, dtype=np.float64) * ((b - a) / (n - 1))


def grid_tanh(n: int, a: float, b: float, alpha: float) -> np.ndarray:
    s = np.arange(n, dtype=np.float64) / (n - 1)
    return a + (b - a) * (1.0 + np.tanh(alpha * (2.0 * s - 1.0)) / np.tanh(alpha)) / 2.0


def grid_asin(n: int, a: float, b: float, beta: float = 33.3) -> np.ndarray:
    alpha = 1.0 - beta / float(n * n)
    xi = a + np.arange(n, dtype=np.float64) * ((b - a) / (n - 1))
    return a + (b - a) * (np.arcsin(-alpha * np.cos(np.pi * (xi - a) / (b - a)))
                          / np.arcsin(alpha) + 1.0) * 0.5


def splitmix64(seed: int, count: int, offset: int = 0) -> np.ndarray:
    with np.errstate(over="ignore"):
        idx = np.arange(offset + 1, offset + count + 1, dtype=np.uint64)
        z = np.uint64(seed) + idx * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed: int, count: int, offset: int = 0) -> np.ndarray:
    return (splitmix64(seed, count, offset) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def normal(seed: int, count: int, offset: int = 0) -> np.ndarray:
    m = (count + 1) // 2
    u = uniform01(seed, 2 * m, offset)
    r = np.sqrt(-2.0 * np.log(1.0 - u[0::2]))
    z = np.empty(2 * m)
    z[0::2] = r * np.cos(2.0 * np.pi * u[1::2])
    z[1::2] = r * np.sin(2.0 * np.pi * u[1::2])
    return z[:count]


def subdomain_layout(ndom=(4, 4, 2), sizes=(48, 64, 96), seed: int = 1):
    """Influence-matrix decomposition of [0,1]^3 into ndom slabs per axis with
    point counts drawn per slab from `sizes` (so faces conform).  Each
    subdomain grid is the reference's asin/cos stretching on its slab
    (class_mesh.f90:156-200: per-subdomain split of [xa,xb], then stretch).

    Returns (slab_sizes, subdomains) with subdomains a list of
    dict(index=(a,b,c), grids=(gx,gy,gz), interfaces=[face ids]).
    """
    u = uniform01(seed, sum(ndom))
    draws = [sizes[int(v * len(sizes))] for v in u]
    slab = [draws[0:ndom[0]], draws[ndom[0]:ndom[0] + ndom[1]], draws[ndom[0] + ndom[1]:]]
    subs = []
    for c in range(ndom[2]):
        for b in range(ndom[1]):
            for a in range(ndom[0]):
                idx = (a, b, c)
                grids = []
                faces = []
                for ax in range(3):
                    k = idx[ax]
                    lo, hi = k / ndom[ax], (k + 1) / ndom[ax]
                    grids.append(grid_asin(slab[ax][k], lo, hi))
                    if k > 0:
                        faces.append(2 * ax)
                    if k < ndom[ax] - 1:
                        faces.append(2 * ax + 1)
                subs.append({"index": idx, "grids": tuple(grids), "interfaces": faces})
    return slab, subs


def helmholtz_256_problem(n: int = 256, stretch: float = 2.0, seed: int = 2, modes: int = 8,
                          kmax: int = 16, noise: float = 1e-3):
    """Single-domain n^3 box [0,1]^3, tanh-stretched in y, uniform x and z;
    RHS = sum of `modes` seeded random-phase cos modes with wavenumbers
    1..kmax per direction plus `noise` white noise; homogeneous Dirichlet.

    Returns (gx, gy, gz, g) with g a Fortran-ordered (n,n,n) array.
    """
    gx = grid_uniform(n, 0.0, 1.0)
    gy = grid_tanh(n, 0.0, 1.0, stretch)
    gz = grid_uniform(n, 0.0, 1.0)
    u = uniform01(seed, 8 * modes)
    g = np.zeros((n, n, n), order="F")
    for m in range(modes):
        k = 1 + np.floor(kmax * u[8 * m:8 * m + 3]).astype(int)
        ph = 2.0 * np.pi * u[8 * m + 3:8 * m + 6]
        amp = 0.5 + u[8 * m + 6]
        fx = np.cos(np.pi * k[0] * gx + ph[0])
        fy = np.cos(np.pi * k[1] * gy + ph[1])
        fz = np.cos(np.pi * k[2] * gz + ph[2])
        g += amp * fx[:, None, None] * fy[None, :, None] * fz[None, None, :]
    g += noise * normal(seed, n ** 3, offset=8 * modes).reshape((n, n, n), order="F")
    return gx, gy, gz, np.asfortranarray(g)


def channel_problem(nx: int = 192, ny: int = 256, nz: int = 192, stretch: float = 2.5,
                    re: float = 3000.0, amp: float = 0.1, kmax: int = 8, modes: int = 8,
                    seed: int = 4, lx: float = 2.0 * np.pi, lz: float = np.pi):
    """Wall-bounded channel: x in [0,lx] and z in [0,lz] uniform, y in [-1,1]
    tanh-stretched (walls at y = +-1, Re on the half-height).

    Velocity = laminar parabola (1 - y^2, 0, 0) + a divergence-free
    perturbation u' = curl(psi), psi_c = E(x,y,z) sum_m a cos(k.x + phase),
    with E = sin^2(pi x/lx) (1-y^2)^2 sin^2(pi z/lz) so that u' and its
    potential vanish on every face; wavenumbers 1..kmax per direction,
    phases and amplitudes from the seeded SplitMix64 stream, scaled to
    max|u'| = amp.  Boundary values: the parabola on every face (0 on the
    walls).  Returns (gx, gy, gz, u, v, w, bc) with bc = {c: (bcx, bcy, bcz)}
    and fields Fortran-ordered (nx, ny, nz).
    """
    gx = grid_uniform(nx, 0.0, lx)
    gy = grid_tanh(ny, -1.0, 1.0, stretch)
    gz = grid_uniform(nz, 0.0, lz)
    X, Y, Z = np.meshgrid(gx, gy, gz, indexing="ij")
    ex, dex = np.sin(np.pi * X / lx) ** 2, np.pi / lx * np.sin(2.0 * np.pi * X / lx)
    ey, dey = (1.0 - Y * Y) ** 2, -4.0 * Y * (1.0 - Y * Y)
    ez, dez = np.sin(np.pi * Z / lz) ** 2, np.pi / lz * np.sin(2.0 * np.pi * Z / lz)
    E = ex * ey * ez
    dE = (dex * ey * ez, ex * dey * ez, ex * ey * dez)
    r = uniform01(seed, 3 * modes * 7)
    grad = []  # grad[c][a] = d psi_c / d x_a
    for c in range(3):
        S = np.zeros_like(X)
        dS = [np.zeros_like(X) for _ in range(3)]
        for m in range(modes):
            q = r[7 * (c * modes + m):7 * (c * modes + m) + 7]
            n = 1 + np.floor(kmax * q[:3]).astype(int)
            k = (2.0 * np.pi * n[0] / lx, np.pi * n[1] / 2.0, 2.0 * np.pi * n[2] / lz)
            a = q[6] - 0.5
            arg = k[0] * X + k[1] * Y + k[2] * Z + 2.0 * np.pi * q[3]
            S += a * np.cos(arg)
            s = -a * np.sin(arg)
            for ax in range(3):
                dS[ax] += k[ax] * s
        grad.append([dE[ax] * S + E * dS[ax] for ax in range(3)])
    up = grad[2][1] - grad[1][2]
    vp = grad[0][2] - grad[2][0]
    wp = grad[1][0] - grad[0][1]
    scale = amp / max(np.abs(up).max(), np.abs(vp).max(), np.abs(wp).max())
    U = 1.0 - Y * Y
    u = np.asfortranarray(U + scale * up)
    v = np.asfortranarray(scale * vp)
    w = np.asfortranarray(scale * wp)

    def faces(f):
        return (np.asfortranarray(np.stack([f[0, 1:-1, 1:-1], f[-1, 1:-1, 1:-1]], -1)),
                np.asfortranarray(np.stack([f[1:-1, 0, 1:-1], f[1:-1, -1, 1:-1]], -1)),
                np.asfortranarray(np.stack([f[1:-1, 1:-1, 0], f[1:-1, 1:-1, -1]], -1)))

    zero = np.zeros_like(U)
    bc = {0: faces(U), 1: faces(zero), 2: faces(zero)}
    return gx, gy, gz, u, v, w, bc
