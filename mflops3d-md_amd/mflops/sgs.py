"""Learned SGS closure workflow: fine run -> filtered training data -> MLP ->
coarse LES runs compared with the filtered fine run (a-posteriori error).

Every field operation runs in the library (NS steps, spline filter, SGS
targets and features, MLP training and inference); this module only
orchestrates the calls and keeps the normalisation scales.  The coarse
grid's points are the filter's output points and its nodes, the way the
reference's filter maps a fine field onto a coarser one (flt_init,
class_filter.f90:34-73, keeps the input, output and node grids;
LSQUnivariateSpline :130-174 fits on the nodes and evaluates on the output
grid).
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

from . import synthetic

MLP_SIZES = (10, 64, 64, 6)


def mlp_init(seed: int) -> np.ndarray:
    """Uniform(+-1/sqrt(fan_in)) weights, zero biases, in the parameter order
    of mf_mlp_create (W1, b1, W2, b2, W3, b3; W row-major out x in)."""
    rng = np.random.default_rng(seed)
    parts = []
    for a, b in zip(MLP_SIZES[:-1], MLP_SIZES[1:]):
        parts.append(rng.uniform(-1, 1, (b, a)).ravel() / np.sqrt(a))
        parts.append(np.zeros(b))
    return np.concatenate(parts).astype(np.float32)


@dataclass
class Scales:
    gscale: float
    lscale: float
    tscale: float


@dataclass
class Result:
    errors: dict = field(default_factory=dict)   # model -> (err_u, err_v, err_w)
    loss: tuple = (0.0, 0.0)                     # training loss, first and last step
    samples: int = 0
    timings: dict = field(default_factory=dict)  # phase -> seconds
    fine_shape: tuple = ()
    coarse_shape: tuple = ()


def training_data(mf, fs, coarse, snaps):
    """Features X and targets T (unscaled) of filtered snapshots, box edges dropped."""
    import torch
    dcs = [mf.DerivativeCoefficients(g, solver=True) for g in coarse]
    Xs, Ts = [], []
    for u, v, w in snaps:
        bars, tau = mf.sgs_targets(*fs, u, v, w)
        X = mf.sgs_features(dcs, *bars, coarse)
        keep = X[:, :9].abs().sum(1) > 0
        Xs.append(X[keep])
        Ts.append(tau[keep])
    return torch.cat(Xs).contiguous(), torch.cat(Ts).contiguous()


def fit_scales(X, T) -> Scales:
    return Scales(gscale=1.0 / float(X[:, :9].std()), lscale=1.0 / float(X[:, 9].max()),
                  tscale=float(T.std()))


def _ns(mf, grids, fields, bc, re, dt):
    s = mf.NavierStokes(*grids, re, dt)
    for c in range(3):
        s.set_bc(c, mf.BoundaryCondition(*[mf.to_device(b) for b in bc[c]]))
    s.set_state(*fields)
    s.set_flow_rate(True, 2.0 / re)
    return s


def _sync():
    import torch
    torch.cuda.synchronize()
    return time.perf_counter()


def run(mf, fine_shape=(64, 72, 48), coarse_shape=(24, 28, 18), re=5000.0, dt=1e-3,
        steps=60, snap_every=20, train_steps=300, lr=3e-3, amp=0.3, kmax=8, seed=4,
        les_c=0.165, params_seed=3, log=print) -> Result:
    """The workflow on the channel of mflops.synthetic (walls at y = +-1,
    the parabola on the other faces, a divergence-free perturbation of size
    amp with wavenumbers up to kmax per direction, which the coarse grid
    must resolve: the scheme's first projection of an unresolved field is an
    O(1) correction).  Returns the relative L2 error of u, v, w (interior points) of each
    coarse model ('none', 'smagorinsky', 'learned') against the filtered fine
    run after `steps` steps."""
    import torch
    res = Result(fine_shape=tuple(fine_shape), coarse_shape=tuple(coarse_shape))
    gx, gy, gz, u, v, w, bc = synthetic.channel_problem(*fine_shape, re=re, amp=amp, kmax=kmax,
                                                        seed=seed)
    cgx, cgy, cgz, _, _, _, cbc = synthetic.channel_problem(*coarse_shape, re=re, amp=0.0)
    fine, coarse = (gx, gy, gz), (cgx, cgy, cgz)
    fs = [mf.SplineFilter(fine[a], coarse[a], coarse[a], 3) for a in range(3)]
    ic = [mf.to_device(f) for f in (u, v, w)]

    # fine run with snapshots
    t0 = _sync()
    s = _ns(mf, fine, ic, bc, re, dt)
    snaps = []
    done = 0
    while done < steps:
        k = min(snap_every, steps - done)
        s.step(k, mf.MF_FP64)
        done += k
        snaps.append(s.get()[:3])
    t1 = _sync()
    res.timings["fine_run"] = t1 - t0
    log("fine run %s: %d steps in %.2f s" % (fine_shape, steps, t1 - t0))
    ref = [mf.filter3d(*fs, q) for q in snaps[-1]]

    # training data and the closure
    X, T = training_data(mf, fs, coarse, snaps)
    sc = fit_scales(X, T)
    res.samples = int(X.shape[0])
    scale = torch.tensor([sc.gscale] * 9 + [sc.lscale], device=X.device, dtype=torch.float32)
    Xs, Ts = (X * scale).contiguous(), (T / sc.tscale).contiguous()
    net = mf.SGSNet(mlp_init(params_seed))
    t2 = _sync()
    first = last = net.adam_step(Xs, Ts, lr)
    for _ in range(train_steps - 1):
        last = net.adam_step(Xs, Ts, lr)
    t3 = _sync()
    res.loss = (first, last)
    res.timings["training"] = t3 - t2
    log("training: %d samples, %d Adam steps in %.2f s, loss %.4e -> %.4e"
        % (res.samples, train_steps, t3 - t2, first, last))

    # coarse runs from the filtered initial field
    cic = [mf.filter3d(*fs, q) for q in ic]
    inner = (slice(1, -1),) * 3
    for model in ("none", "smagorinsky", "learned"):
        c = _ns(mf, coarse, cic, cbc, re, dt)
        if model == "smagorinsky":
            c.set_les(1, les_c)
        elif model == "learned":
            c.set_sgs(net, sc.gscale, sc.lscale, sc.tscale)
        t4 = _sync()
        c.step(steps, mf.MF_FP64)
        t5 = _sync()
        got = c.get()[:3]
        errs = []
        for a, b in zip(got, ref):
            d = (a - b)[inner]
            errs.append(float(torch.linalg.vector_norm(d) / torch.linalg.vector_norm(b[inner])))
        res.errors[model] = tuple(errs)
        res.timings["coarse_" + model] = t5 - t4
        log("coarse %-11s %s: error u %.4e v %.4e w %.4e (%.2f s)"
            % (model, coarse_shape, *errs, t5 - t4))
    return res


@dataclass
class CalibratedResult:
    c_fit: float = 0.0                              # Smagorinsky constant fitted on the training flow
    fit_errors: dict = field(default_factory=dict)  # candidate c -> error on the training flow
    errors: dict = field(default_factory=dict)      # model -> error on the test flow
    loss: tuple = (0.0, 0.0)
    samples: int = 0
    backscatter: float = 0.0                        # <-tau:S> / <|tau| |S|> of the filter's own stress
    timings: dict = field(default_factory=dict)


def _flow(mf, fine_shape, coarse_shape, re, dt, steps, snap_every, amp, kmax, modes, seed):
    """Fine run of one flow realisation: filters, coarse grids and boundary
    values, filtered snapshots and the filtered initial field."""
    gx, gy, gz, u, v, w, bc = synthetic.channel_problem(*fine_shape, re=re, amp=amp, kmax=kmax,
                                                        modes=modes, seed=seed)
    cgx, cgy, cgz, _, _, _, cbc = synthetic.channel_problem(*coarse_shape, re=re, amp=0.0)
    fine, coarse = (gx, gy, gz), (cgx, cgy, cgz)
    fs = [mf.SplineFilter(fine[a], coarse[a], coarse[a], 3) for a in range(3)]
    ic = [mf.to_device(f) for f in (u, v, w)]
    s = _ns(mf, fine, ic, bc, re, dt)
    snaps, done = [], 0
    while done < steps:
        k = min(snap_every, steps - done)
        s.step(k, mf.MF_FP64)
        done += k
        snaps.append(s.get()[:3])
    return dict(coarse=coarse, fs=fs, cbc=cbc, snaps=snaps,
                refs=[[mf.filter3d(*fs, q) for q in sn] for sn in snaps],
                start=[mf.filter3d(*fs, q) for q in ic])


def _coarse_error(mf, fl, re, dt, steps, model=None, net=None, scales=None):
    """Mean over u, v, w of the relative L2 error (interior points) of a
    coarse run from the filtered initial field against the filtered fine run."""
    import torch
    c = _ns(mf, fl["coarse"], fl["start"], fl["cbc"], re, dt)
    if model == "smagorinsky":
        c.set_les(1, scales)
    elif model == "learned":
        c.set_sgs(net, *scales)
        c.set_sgs_coupling(1)
    c.step(steps, mf.MF_FP64)
    inner = (slice(1, -1),) * 3
    errs = [float(torch.linalg.vector_norm((a - b)[inner]) / torch.linalg.vector_norm(b[inner]))
            for a, b in zip(c.get()[:3], fl["refs"][-1])]
    return float(np.mean(errs))


def run_calibrated(mf, fine_shape=(128, 128, 96), coarse_shape=(32, 32, 24), re=1e5, dt=1e-3,
                   steps=200, snap_every=20, amp=1.0, kmax=6, modes=64, train_seed=4,
                   test_seed=5, candidates=(0.05, 0.1, 0.15, 0.2, 0.25, 0.3, 0.4),
                   train_steps=1000, lr=3e-3, les_c=0.165, params_seed=3,
                   log=print) -> CalibratedResult:
    """The learned closure with a-posteriori-consistent targets.

    The filter's own stress tau = F(uu) - F(u)F(u) of the spline filter is
    not dissipative on average (its backscatter ratio is reported), and a
    closure trained on it injects energy in coarse runs.  Here the targets
    are the eddy-viscosity stress -2 c^2 Delta^2 |S| S of the filtered
    training snapshots with c fitted so that coarse runs of the training
    flow follow its filtered fine run; the network learns that stress from
    the filtered gradients and enters the coarse runs as an eddy viscosity.
    It is then judged on another flow realisation (test_seed) against no
    model and the reference's Smagorinsky model with constant les_c, by the
    relative error against that flow's filtered fine run."""
    import torch
    res = CalibratedResult()
    t0 = _sync()
    tr = _flow(mf, fine_shape, coarse_shape, re, dt, steps, snap_every, amp, kmax, modes, train_seed)
    t1 = _sync()
    res.timings["fine_run_train"] = t1 - t0
    # backscatter of the filter's own stress on the last training snapshot
    dcs = [mf.DerivativeCoefficients(g, solver=True) for g in tr["coarse"]]
    bars, tau = mf.sgs_targets(*tr["fs"], *tr["snaps"][-1])
    Xb = mf.sgs_features(dcs, *bars, tr["coarse"]).double()
    g = [[Xb[:, 3 * i + j] for j in range(3)] for i in range(3)]
    S = [[0.5 * (g[i][j] + g[j][i]) for j in range(3)] for i in range(3)]
    pair = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]
    tS = sum((1.0 if i == j else 2.0) * tau[:, k].double() * S[i][j] for k, (i, j) in enumerate(pair))
    nt = torch.sqrt(sum((1.0 if i == j else 2.0) * tau[:, k].double() ** 2 for k, (i, j) in enumerate(pair)))
    ns_ = torch.sqrt(sum(S[i][j] ** 2 for i in range(3) for j in range(3)))
    res.backscatter = float(-tS.sum() / (nt * ns_).sum())
    # the Smagorinsky constant with which coarse runs follow the training flow
    res.fit_errors = {c: _coarse_error(mf, tr, re, dt, steps, "smagorinsky", scales=c) for c in candidates}
    res.fit_errors[0.0] = _coarse_error(mf, tr, re, dt, steps)
    res.c_fit = min(res.fit_errors, key=res.fit_errors.get)
    t2 = _sync()
    res.timings["calibration"] = t2 - t1
    log("backscatter ratio of the filter's stress %.3f; fitted c = %.3f (%s)"
        % (res.backscatter, res.c_fit, ", ".join("%.2f: %.4f" % kv for kv in sorted(res.fit_errors.items()))))
    if res.c_fit <= 0.0:
        raise RuntimeError("no eddy viscosity improves the training flow; nothing to learn")
    # training data: filtered gradients -> eddy-viscosity stress of c_fit
    Xs, Ts = [], []
    for sn in tr["snaps"]:
        bars = [mf.filter3d(*tr["fs"], q) for q in sn]
        X = mf.sgs_features(dcs, *bars, tr["coarse"])
        X = X[X[:, :9].abs().sum(1) > 0].contiguous()
        Xs.append(X)
        Ts.append(mf.sgs_eddy_targets(X, res.c_fit))
    X, T = torch.cat(Xs).contiguous(), torch.cat(Ts).contiguous()
    sc = fit_scales(X, T)
    res.samples = int(X.shape[0])
    scale = torch.tensor([sc.gscale] * 9 + [sc.lscale], device=X.device, dtype=torch.float32)
    Xn, Tn = (X * scale).contiguous(), (T / sc.tscale).contiguous()
    net = mf.SGSNet(mlp_init(params_seed))
    first = last = net.adam_step(Xn, Tn, lr)
    for _ in range(train_steps - 1):
        last = net.adam_step(Xn, Tn, lr)
    res.loss = (first, last)
    t3 = _sync()
    res.timings["training"] = t3 - t2
    log("training: %d samples, %d Adam steps, loss %.4e -> %.4e" % (res.samples, train_steps, first, last))
    # judged on another realisation
    te = _flow(mf, fine_shape, coarse_shape, re, dt, steps, snap_every, amp, kmax, modes, test_seed)
    res.errors["none"] = _coarse_error(mf, te, re, dt, steps)
    res.errors["smagorinsky"] = _coarse_error(mf, te, re, dt, steps, "smagorinsky", scales=les_c)
    res.errors["learned"] = _coarse_error(mf, te, re, dt, steps, "learned", net,
                                          (sc.gscale, sc.lscale, sc.tscale))
    res.timings["test"] = _sync() - t3
    log("test flow: error none %.4f, smagorinsky %.4f, learned %.4f"
        % (res.errors["none"], res.errors["smagorinsky"], res.errors["learned"]))
    return res
