"""The learned initial guess end to end: solver-generated training data,
training on the device, and the guess a solve starts from.

    data = generate(seeds, n=32, maxk=3)          # cube problems, one batched solve per chunk
    g = LearnedGuess(); g.fit(data)               # Adam on the device
    st = g.solve(mesh, albedo, ang)               # guess + mixed solve, plain fallback

The target of vertex v is U_v / U_inf,v with U_inf = 4 pi I_p / (1 - albedo)
(I_p = 1 in the cube problems), so the network's sigmoid output maps to
f[v][00] = y_v I_p / (1 - albedo)."""
import numpy as np

from . import problems
from .api import MLP, Angular, Batch, HipContext, mlp_features, TR_SOLVER_MIXED

FOUR_PI = 4.0 * np.pi


def _pool():
    import concurrent.futures as cf
    import os
    return cf.ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1))


def cube_problems(seeds, n=32, maxk=3):
    """The cube problems of `seeds`, built on the host cores (the mesh and
    field generators run in the C library, outside the interpreter lock)."""
    with _pool() as ex:
        return list(ex.map(lambda s: problems.cube(n, maxk=maxk, seed=int(s)), seeds))


def generate(seeds, n=32, maxk=3, chunk=64, inner_tol=1e-4, progress=None):
    """Features and targets of the cube problems of `seeds`, solved to a 1e-8
    FP64 residual with the batched mixed solver.  Returns (feat [N][8],
    target [N], per-problem vertex counts)."""
    feats, targs, counts = [], [], []
    for c0 in range(0, len(seeds), chunk):
        P = cube_problems(seeds[c0:c0 + chunk], n, maxk)
        with _pool() as ex:
            F = list(ex.map(lambda p: mlp_features(p.mesh, p.albedo), P))
        B = Batch([p.mesh for p in P], Angular(maxk), [p.albedo for p in P])
        B.autotune(0)
        B.autotune(1)
        d_x = B.allocVector()
        st, per = B.solve_batch(d_x, solver=TR_SOLVER_MIXED, use_prec=2, tol_rel=1e-8, inner_tol=inner_tol, maxit=30)
        U, _ = B.moments(d_x)
        B.deviceFree(d_x)
        for k, p in enumerate(P):
            lo, hi = int(B.voff[k]), int(B.voff[k + 1])
            feats.append(F[k])
            targs.append(np.clip(U[lo:hi] * (1.0 - p.albedo) / FOUR_PI, 0.0, 1.0).astype(np.float32))
            counts.append(hi - lo)
        if progress:
            progress(min(c0 + chunk, len(seeds)))
        del B
    return np.concatenate(feats), np.concatenate(targs), counts


def generate_single(probs, inner_tol=1e-4, progress=None):
    """generate() for problems solved one at a time: scattering a batch
    cannot carry (Henyey-Greenstein), each problem with its own angular
    data.  I_p = 1 is assumed, as in the workloads."""
    feats, targs, counts = [], [], []
    for k, p in enumerate(probs):
        ctx = HipContext(p.mesh, p.ang)
        d_x = ctx.allocVector()
        st, _ = ctx.solve(d_x, solver=TR_SOLVER_MIXED, use_prec=2, tol_abs=0.0, tol_rel=1e-8, inner_tol=inner_tol,
                          maxit=30, inner_maxit=20000)
        U, _ = ctx.moments(d_x)
        ctx.close()
        feats.append(mlp_features(p.mesh, p.albedo))
        targs.append(np.clip(U * (1.0 - p.albedo) / FOUR_PI, 0.0, 1.0).astype(np.float32))
        counts.append(len(U))
        if progress:
            progress(k + 1)
    return np.concatenate(feats), np.concatenate(targs), counts


class LearnedGuess:
    def __init__(self, seed=1, params=None):
        self.net = MLP(params=params, seed=seed)
        self.mu = np.zeros(8, np.float32)
        self.sd = np.ones(8, np.float32)

    def fit(self, feat, target, steps=2000, batch=16384, lr=2e-3, seed=1):
        self.mu = feat.mean(0).astype(np.float32)
        # a feature constant over the training set carries no information and
        # its first-layer weights get no gradient: a huge scale makes the
        # network ignore it wherever it is applied
        sd = feat.std(0)
        self.sd = np.where(sd < 1e-6, 1e30, np.maximum(sd, 1e-3)).astype(np.float32)
        self.net.set_norm(self.mu, self.sd)
        d_f, d_t = self.net.upload(feat), self.net.upload(target)
        return self.net.train(d_f, d_t, len(target), steps, batch, lr, seed)

    def predict(self, feat):
        d_f = self.net.upload(feat)
        d_y = self.net.alloc(len(feat))
        self.net.predict(d_f, len(feat), d_y)
        return self.net.download(d_y, len(feat))

    def write_guess(self, ctx, mesh, albedo, Ip=1.0):
        """ctx's device vector holding the learned guess, built on the device
        (features, prediction and guess: tr_ctx_learned_guess).  Ip: the
        source scale I_p of U_inf = 4 pi I_p / (1 - albedo), or None for each
        vertex's own (feature 7 holds log I_p per vertex)."""
        f = ctx.allocVector()
        ctx.learned_guess(self.net, albedo, -1.0 if Ip is None else Ip, f)
        return f

    def write_guess_host(self, ctx, mesh, albedo, Ip=1.0):
        """write_guess with the features computed on the host
        (tr_mlp_features) and the prediction copied through the host."""
        feat = mlp_features(mesh, albedo)
        y = self.predict(feat)
        if Ip is None:
            y = (y * np.exp(feat[:, 7])).astype(np.float32)
            Ip = 1.0
        d_y = ctx.deviceAlloc(4 * len(y))
        ctx.copyToDev(d_y, y)
        f = ctx.allocVector()
        ctx.scalar_guess(d_y, Ip / (1.0 - albedo), f)
        ctx.deviceFree(d_y)
        return f
