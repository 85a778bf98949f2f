"""Spline filter, LES eddy viscosities and the learned SGS closure: CPU restatement.

TEST INFRASTRUCTURE ONLY: imported by tests/ (and nothing in the product).

* lsq_filter_matrix: LSQUnivariateSpline (src/class_filter/class_filter.f90:
  130-174): a least-squares spline of degree k with the knots of curfit's
  iopt = -1 call (k+1 copies of each end point, the interior nodes xn(2:nn-1))
  fitted to the input line (weights wi) and evaluated at the output points
  (splev) or differentiated there (splder, the reference's ypo).  The fit is
  linear in the data, so the filter of one axis is the matrix
  F = E (B^T W^2 B)^-1 B^T W^2 (B, E: B-spline bases at the input and output
  points).  Pinned against scipy.interpolate.LSQUnivariateSpline, which wraps
  the same FITPACK routines, and by polynomial reproduction.
* les_nu: navier_LES (src/class_navier/class_navier_3D.f90:1225-1449): the
  Smagorinsky (les_type 1), WALE (2) and sigma (3, les_sigma :1453-1497)
  viscosities from the velocity gradients du(i,j) = d u_i / d x_j, times the
  filter width (dx^2+dy^2+dz^2)/3 of one-sided / centred spacings, zero on the
  box edges.  Pinned by closed forms (Smagorinsky on simple shear) and by the
  models' defining properties (WALE vanishes for pure shear, sigma for
  two-dimensional gradients and solid rotation).
* sgs_targets / sgs_features: the closure's training data from a fine field
  (the filter module's purpose, class_filter.f90): filtered velocity and
  tau_ij = F(u_i u_j) - F(u_i) F(u_j); inputs are the gradients du_ij (edges
  zeroed, as navier_LES zeroes them) and the filter width.  Pinned by the
  filter's polynomial reproduction (tau = 0 for a linear field when k >= 2,
  and for any field when the filter is the identity).
* MLP: the learned closure 10 -> 64 -> 64 -> 6 (tanh hidden layers) on
  features (9 velocity gradients, filter width) to the 6 SGS stresses, with
  bf16-rounded weights as the device kernels use them.  Its gradients are
  pinned by central finite differences.
"""
from __future__ import annotations

import numpy as np


# ----------------------------------------------------------------------------
# LSQ spline filter
# ----------------------------------------------------------------------------

def knots(xb: float, xe: float, xn, k: int) -> np.ndarray:
    """curfit iopt = -1 knots (class_filter.f90:149-154)."""
    xn = np.asarray(xn, dtype=np.float64)
    return np.concatenate([[xb] * (k + 1), xn[1:-1], [xe] * (k + 1)])


def bspline_basis(t: np.ndarray, k: int, x, deriv: int = 0) -> np.ndarray:
    """Matrix (len(x), len(t) - k - 1) of the B-splines of degree k on knots t
    (Cox-de Boor), or of their first derivatives (deriv=1); the right end
    point belongs to the last interval."""
    x = np.atleast_1d(np.asarray(x, dtype=np.float64))
    nc = len(t) - k - 1
    out = np.zeros((len(x), nc))
    last = nc - 1  # last span with t[l] < t[l+1]
    while t[last] >= t[last + 1]:
        last -= 1
    deg = k - deriv
    for r, xv in enumerate(x):
        l = int(np.searchsorted(t, xv, side="right") - 1)
        l = min(max(l, k), last)
        # de Boor's triangular recurrence for the deg+1 nonzero functions of
        # degree deg on span l (indices l-deg .. l)
        N = np.zeros(deg + 1)
        N[0] = 1.0
        left = np.zeros(deg + 1)
        right = np.zeros(deg + 1)
        for j in range(1, deg + 1):
            left[j] = xv - t[l + 1 - j]
            right[j] = t[l + j] - xv
            saved = 0.0
            for q in range(j):
                tmp = N[q] / (right[q + 1] + left[j - q])
                N[q] = saved + right[q + 1] * tmp
                saved = left[j - q] * tmp
            N[j] = saved
        if deriv == 0:
            out[r, l - k:l + 1] = N
        else:
            # N'_{i,k} = k (N_{i,k-1} / (t_{i+k} - t_i) - N_{i+1,k-1} / (t_{i+k+1} - t_{i+1}))
            for a in range(k + 1):
                i = l - k + a
                lo = N[a - 1] if a >= 1 else 0.0
                hi = N[a] if a <= k - 1 else 0.0
                d1 = t[i + k] - t[i]
                d2 = t[i + k + 1] - t[i + 1]
                out[r, i] = k * ((lo / d1 if d1 > 0 else 0.0) - (hi / d2 if d2 > 0 else 0.0))
    return out


def lsq_filter_matrix(xi, xo, xn, k: int = 3, wi=None, deriv: int = 0) -> np.ndarray:
    """F (len(xo) x len(xi)): the LSQ spline fitted to values on xi with
    weights wi (curfit minimises sum (wi (yi - s(xi)))^2), at xo (splev), or
    its first derivative there (deriv=1, splder)."""
    xi = np.asarray(xi, dtype=np.float64)
    w = np.ones_like(xi) if wi is None else np.asarray(wi, dtype=np.float64)
    t = knots(xi[0], xi[-1], xn, k)
    B = bspline_basis(t, k, xi) * w[:, None]
    E = bspline_basis(t, k, xo, deriv)
    q, r = np.linalg.qr(B)
    return E @ np.linalg.solve(r, q.T) * w[None, :]


# ----------------------------------------------------------------------------
# Eddy viscosities
# ----------------------------------------------------------------------------

def _spacing(g):
    g = np.asarray(g, dtype=np.float64)
    d = np.empty_like(g)
    d[0] = g[1] - g[0]
    d[-1] = g[-1] - g[-2]
    d[1:-1] = 0.5 * (g[2:] - g[:-2])
    return d


def zero_edges(f):
    f = f.copy()
    for a, b in ((0, 0), (0, -1), (-1, 0), (-1, -1)):
        f[a, b, :] = 0.0
        f[a, :, b] = 0.0
        f[:, a, b] = 0.0
    return f


def les_sigma(g2):
    """les_sigma (class_navier_3D.f90:1453-1497) on arrays of 3x3 gradients
    g2[..., l, m]: the singular values from the invariants of g2 g2^T."""
    g = np.einsum("...ik,...jk->...ij", g2, g2)
    i1 = g[..., 0, 0] + g[..., 1, 1] + g[..., 2, 2]
    i2 = (g[..., 0, 0] * g[..., 1, 1] + g[..., 0, 0] * g[..., 2, 2] + g[..., 1, 1] * g[..., 2, 2]
          - g[..., 0, 1] * g[..., 1, 0] - g[..., 0, 2] * g[..., 2, 0] - g[..., 1, 2] * g[..., 2, 1])
    i3 = np.linalg.det(g)
    a1 = i1 ** 2 / 9.0 - i2 / 3.0
    a2 = i1 ** 3 / 27.0 - i1 * i2 / 6.0 + i3 / 2.0
    with np.errstate(invalid="ignore", divide="ignore"):
        a3 = np.where(a1 < 1e-10, 0.0,
                      np.arccos(np.clip(a2 / np.maximum(a1, 1e-300) ** 1.5, -1.0, 1.0)) / 3.0)
    s1 = i1 / 3.0 + 2.0 * np.sqrt(np.maximum(a1, 0.0)) * np.cos(a3)
    s2 = i1 / 3.0 - 2.0 * np.sqrt(np.maximum(a1, 0.0)) * np.cos(a3 + np.pi / 3.0)
    s3 = i1 - s1 - s2
    return (np.sqrt(np.maximum(s1, 0.0)), np.sqrt(np.maximum(s2, 0.0)),
            np.sqrt(np.maximum(s3, 0.0)))


def les_nu(du, gx, gy, gz, les_type: int, c: float):
    """du[i][j] (full fields) -> les_nu (navier_LES :1269-1449)."""
    du = [[np.asarray(du[i][j], dtype=np.float64) for j in range(3)] for i in range(3)]
    du = [[zero_edges(du[i][j]) for j in range(3)] for i in range(3)]
    shape = du[0][0].shape
    nu = np.zeros(shape)
    if les_type in (1, 2):
        ss = np.zeros(shape)
        for i in range(3):
            ss += du[i][i] ** 2
            for j in range(i + 1, 3):
                ss += 0.5 * (du[i][j] + du[j][i]) ** 2
        if les_type == 1:
            nu = c * np.sqrt(2.0 * ss)
        else:
            sd = np.zeros(shape)
            tr = sum(du[m][k] * du[k][m] for m in range(3) for k in range(3))
            for i in range(3):
                for j in range(3):
                    a = np.zeros(shape)
                    for k in range(3):
                        a += 0.5 * (du[i][k] * du[k][j] + du[j][k] * du[k][i])
                    if i == j:
                        a -= tr / 3.0
                    sd += a ** 2
            with np.errstate(invalid="ignore", divide="ignore"):
                nu = np.where(sd < 1e-10, 0.0, c * sd ** 1.5 / (ss ** 2.5 + sd ** 1.25))
    elif les_type == 3:
        g2 = np.stack([np.stack([du[l][m] for m in range(3)], -1) for l in range(3)], -2)
        s1, s2, s3 = les_sigma(g2)
        with np.errstate(invalid="ignore", divide="ignore"):
            nu = np.where(s1 < 1e-10, 0.0, c * s3 * (s1 - s2) * (s2 - s3) / np.maximum(s1, 1e-300) ** 2)
    dx, dy, dz = (_spacing(g) ** 2 for g in (gx, gy, gz))
    nu = nu * (dx[:, None, None] + dy[None, :, None] + dz[None, None, :]) / 3.0
    return zero_edges(nu)


# ----------------------------------------------------------------------------
# Closure data
# ----------------------------------------------------------------------------

PAIRS = ((0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2))  # tau columns


def pair_col(i: int, j: int) -> int:
    return PAIRS.index((min(i, j), max(i, j)))


def filter3d(Fx, Fy, Fz, f):
    return np.einsum("ai,bj,ck,ijk->abc", Fx, Fy, Fz, f, optimize=True)


def sgs_targets(Fx, Fy, Fz, u, v, w, tscale: float = 1.0):
    """(ubar, vbar, wbar) and tau (n x 6, rows in Fortran point order) / tscale."""
    q = [np.asarray(a, dtype=np.float64) for a in (u, v, w)]
    qb = [filter3d(Fx, Fy, Fz, a) for a in q]
    tau = np.stack([(filter3d(Fx, Fy, Fz, q[i] * q[j]) - qb[i] * qb[j]).ravel(order="F")
                    for i, j in PAIRS], 1) / tscale
    return qb, tau


def eddy_targets(X, c: float):
    """Smagorinsky stress -2 c^2 Delta^2 |S| S (n x 6: xx yy zz xy xz yz) of
    unscaled feature rows (nine gradients row-major ij, then Delta)."""
    X = np.asarray(X, dtype=np.float64)
    g = [[X[:, 3 * i + j] for j in range(3)] for i in range(3)]
    S = [[0.5 * (g[i][j] + g[j][i]) for j in range(3)] for i in range(3)]
    SS = sum(S[i][j] ** 2 for i in range(3) for j in range(3))
    f = -2.0 * c * c * X[:, 9] ** 2 * np.sqrt(2.0 * SS)
    return np.stack([f * S[0][0], f * S[1][1], f * S[2][2], f * S[0][1], f * S[0][2], f * S[1][2]], 1)


def sgs_features(du, gx, gy, gz, gscale: float = 1.0, lscale: float = 1.0):
    """X (n x 10): gscale du_ij (row-major ij, edges zero), lscale Delta."""
    cols = [gscale * zero_edges(np.asarray(du[i][j], dtype=np.float64)).ravel(order="F")
            for i in range(3) for j in range(3)]
    dx, dy, dz = (_spacing(g) ** 2 for g in (gx, gy, gz))
    d = np.sqrt((dx[:, None, None] + dy[None, :, None] + dz[None, None, :]) / 3.0)
    cols.append(lscale * d.ravel(order="F"))
    return np.stack(cols, 1)


# ----------------------------------------------------------------------------
# Learned closure
# ----------------------------------------------------------------------------

SIZES = (10, 64, 64, 6)


def bf16(a):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    a = np.asarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32)


def unpack(p):
    """Flat parameter vector (W1, b1, W2, b2, W3, b3; W row-major out x in)."""
    out, o = [], 0
    for a, b in zip(SIZES[:-1], SIZES[1:]):
        out.append(p[o:o + a * b].reshape(b, a)); o += a * b
        out.append(p[o:o + b]); o += b
    return out


def forward(p, X, rounding: bool = True):
    """Y and the hidden activations, in float64 with bf16-rounded weights
    (rounding=False: the weights as given, for finite-difference checks)."""
    W1, b1, W2, b2, W3, b3 = unpack(np.asarray(p, dtype=np.float64 if not rounding else np.float32))
    if rounding:
        W1, W2, W3 = (bf16(w).astype(np.float64) for w in (W1, W2, W3))
    X = np.asarray(X, dtype=np.float64)
    h1 = np.tanh(X @ W1.T + b1)
    h2 = np.tanh(h1 @ W2.T + b2)
    return h2 @ W3.T + b3, h1, h2


def loss_and_grad(p, X, T, rounding: bool = True):
    """Mean over samples of sum((Y - T)^2) and its gradient w.r.t. p (the
    weight gradients taken at the bf16-rounded weights)."""
    W1, b1, W2, b2, W3, b3 = unpack(np.asarray(p, dtype=np.float32 if rounding else np.float64))
    if rounding:
        W2r, W3r = bf16(W2).astype(np.float64), bf16(W3).astype(np.float64)
    else:
        W2r, W3r = W2, W3
    Y, h1, h2 = forward(p, X, rounding)
    n = len(X)
    d = Y - np.asarray(T, dtype=np.float64)
    loss = float((d ** 2).sum() / n)
    dY = 2.0 * d / n
    gW3 = dY.T @ h2
    gb3 = dY.sum(0)
    dh2 = (dY @ W3r) * (1 - h2 ** 2)
    gW2 = dh2.T @ h1
    gb2 = dh2.sum(0)
    dh1 = (dh2 @ W2r) * (1 - h1 ** 2)
    gW1 = dh1.T @ np.asarray(X, dtype=np.float64)
    gb1 = dh1.sum(0)
    g = np.concatenate([a.ravel() for a in (gW1, gb1, gW2, gb2, gW3, gb3)])
    return loss, g


def adam_step(p, m, v, g, t: int, lr: float, beta1: float = 0.9, beta2: float = 0.999,
              eps: float = 1e-8):
    """One bias-corrected Adam update (Kingma & Ba 2015, Algorithm 1) in FP64:
    returns the new (p, m, v) after step t >= 1 with gradient g."""
    m = beta1 * m + (1.0 - beta1) * g
    v = beta2 * v + (1.0 - beta2) * g * g
    mh = m / (1.0 - beta1 ** t)
    vh = v / (1.0 - beta2 ** t)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v


def init_params(seed: int):
    """Uniform(+-1/sqrt(fan_in)) weights, zero biases (float32)."""
    rng = np.random.default_rng(seed)
    parts = []
    for a, b in zip(SIZES[:-1], SIZES[1:]):
        parts.append(rng.uniform(-1, 1, (b, a)).ravel() / np.sqrt(a))
        parts.append(np.zeros(b))
    return np.concatenate(parts).astype(np.float32)


NPARAM = sum(a * b + b for a, b in zip(SIZES[:-1], SIZES[1:]))
