"""Learned importance-sampling proposal for the three-phonon linewidth sum.

The deterministic linewidth of mode (q0, s0) is a sum over q1 of per-pair terms
t(q1, s0) = sum_{s1,s2} W(s0,s1,s2; q0,q1) G(s0,s1,s2) (CalculateLinewidth,
src/Kappa/Src/ThirdOrder_m.f90:490-541, with G the occupation-weighted Gaussians).  A sampled
estimate evaluates K of the Nq pair terms, drawn with probability p(q1), and returns
(1/K) sum_k t(q1_k)/p(q1_k) (unbiased for any p > 0).  The proposal is

    p(q1) = (1 - lam) S(q1) / sum S + lam / Nq,   S(q1) = sum_{(s0,s1,s2) resonant} exp(f(x)) G

where f is an MLP (spring_amd.lib.Mlp, trained on the GPU) that predicts log W from
x = (q0, q1, w0, w1, w2, s0, s1, s2), trained only on W values the method itself produced
(spr_v3_w_pairs), and G is known exactly.  The uniform floor lam keeps the weights bounded.
`physics_only=True` replaces exp(f) by 1 (the proposal without the network), so the
network's own gain can be measured against it and against uniform sampling.

The pair-level proposal (PairProposal) scores what the draw selects: a (q0, q1) pair carries
every branch.  Its target is the pair's measured share of the linewidths,
T(q1) = sum_s0 |t(q1, s0)| / sum_q1' |t(q1', s0)| (the s0 with a nonzero linewidth), learned from
the pair terms the method computes for a set of training q0; its inputs are the pair's
resonance features (spr_pair_scores: the G-only weight S = sum_s0 g(q1, s0) / sum_q1 g(q1, s0)
with g = sum |G| / (w1 w2) over the resonant triplets, and S-weighted frequency moments).  The
network-free proposal of the same form uses S itself.
"""
from __future__ import annotations

import numpy as np

from . import constants as C
from .lib import Mlp

NF = 12


def features(qm, ph, i0g, q1, s0, s1, s2, wmax):
    """Feature rows (float32) for triplets; arrays broadcast against each other."""
    q1 = np.asarray(q1)
    q2 = qm.index(-qm.q_int[i0g][None, :].astype(np.int64) - qm.q_int[q1].astype(np.int64))
    nd = ph.omega.shape[1]
    k0 = np.broadcast_to(qm.q_int[i0g] / qm.mesh, q1.shape + (3,))
    k1 = qm.q_int[q1] / qm.mesh
    cols = [k0[..., 0], k0[..., 1], k0[..., 2], k1[..., 0], k1[..., 1], k1[..., 2],
            ph.omega[i0g][s0] / wmax, ph.omega[q1, s1] / wmax, ph.omega[q2, s2] / wmax,
            s0 / (nd - 1.0), s1 / (nd - 1.0), s2 / (nd - 1.0)]
    return np.stack([np.broadcast_to(c, np.broadcast(*cols).shape) for c in cols], axis=-1).astype(np.float32)


def training_set(ctx, qm, ph, n_pairs: int, per_pair: int, seed: int = 0, chunk: int = 128):
    """(X, y) with y = log W of random branch triplets of random (q0, q1) pairs."""
    rng = np.random.default_rng(seed)
    nd = ph.omega.shape[1]
    wmax = float(np.abs(ph.omega).max())
    q0i = rng.integers(0, qm.nq0, n_pairs)
    q1 = rng.integers(0, qm.nq, n_pairs)
    Xs, ys = [], []
    for lo in range(0, n_pairs, chunk):
        W = ctx.v3_w_pairs(q0i[lo:lo + chunk], q1[lo:lo + chunk])
        for k in range(W.shape[0]):
            flat = W[k].ravel()
            nz = np.nonzero(flat > 0)[0]
            if nz.size == 0:
                continue
            pick = rng.choice(nz, size=min(per_pair, nz.size), replace=False)
            s0, s1, s2 = np.unravel_index(pick, (nd, nd, nd))
            ig = int(qm.irr_idx[q0i[lo + k]])
            Xs.append(features(qm, ph, ig, np.full(pick.size, q1[lo + k]), s0, s1, s2, wmax))
            ys.append(np.log(flat[pick]))
    X = np.concatenate(Xs)
    y = np.concatenate(ys).astype(np.float32)
    perm = rng.permutation(y.size)
    return X[perm], y[perm]


class Proposal:
    """Trained network + normalisation."""

    def __init__(self, mlp: Mlp, y_mean: float, y_std: float, wmax: float):
        self.mlp, self.mu, self.sd, self.wmax = mlp, y_mean, y_std, wmax

    @classmethod
    def train(cls, X, y, wmax, epochs=8, batch=4096, lr=2e-3, seed=0):
        mu, sd = float(y.mean()), float(y.std() + 1e-12)
        mlp = Mlp(NF, seed=seed)
        losses = mlp.train(X, (y - mu) / sd, epochs=epochs, batch=batch, lr=lr)
        p = cls(mlp, mu, sd, wmax)
        p.losses = losses
        return p

    def log_w(self, X, prec="bf16"):
        return self.mlp.predict(X, prec) * self.sd + self.mu


def gauss_weights(ph, i0g, q1, q2, sigma):
    """G(s0,s1,s2) for pairs (q1, q2): array (n, nd, nd, nd)."""
    w0 = ph.omega[i0g][None, :, None, None]
    w1 = ph.omega[q1][:, None, :, None]
    w2 = ph.omega[q2][:, None, None, :]
    n1 = ph.nbe[q1][:, None, :, None]
    n2 = ph.nbe[q2][:, None, None, :]
    c0 = C.ONE_BY_SQRT_2PI / sigma
    h = 0.5 / sigma ** 2
    d = lambda x: c0 * np.exp(-h * x * x)
    return (n1 + n2 + 1.0) * d(w0 - w1 - w2) + (n1 - n2) * (d(w0 + w1 - w2) - d(w0 - w1 + w2))


def proposal_scores(prop: Proposal | None, qm, ph, i0g, sigma, resonant=1e-3, prec="bf16", chunk=256):
    """S(q1) over the whole mesh for q0 (grid index i0g); prop=None gives the G-only proposal.

    Returns (S, n_evals): n_evals = number of network evaluations used."""
    nd = ph.omega.shape[1]
    q1_all = np.arange(qm.nq)
    q2_all = qm.index(-qm.q_int[i0g][None, :].astype(np.int64) - qm.q_int.astype(np.int64))
    S = np.zeros(qm.nq)
    n_eval = 0
    iw0 = ph.inv_omega[i0g]
    for lo in range(0, qm.nq, chunk):
        q1 = q1_all[lo:lo + chunk]
        q2 = q2_all[lo:lo + chunk]
        Graw = np.abs(gauss_weights(ph, i0g, q1, q2, sigma))
        # W = |V|^2 (1/w0)(1/w1)(1/w2) f: select triplets by G times the 1/omega factors
        scale = np.abs(iw0[None, :, None, None] * ph.inv_omega[q1][:, None, :, None] *
                       ph.inv_omega[q2][:, None, None, :])
        Gs = Graw * scale
        thr = resonant * Gs.max() if Gs.size else 0.0
        sel = np.nonzero(Gs > thr)
        if sel[0].size == 0:
            continue
        if prop is None:
            np.add.at(S, lo + sel[0], Gs[sel])
            continue
        X = features(qm, ph, i0g, q1[sel[0]], sel[1], sel[2], sel[3], prop.wmax)
        lw = prop.log_w(X, prec)
        n_eval += X.shape[0]
        np.add.at(S, lo + sel[0], np.exp(np.minimum(lw - prop.mu, 80.0)) * Graw[sel])
    return S, n_eval


def device_scores(prop: Proposal | None, ctx, q0_index, sigma, window=3.0):
    """S(q1) of every q1 for each irreducible q0 index, computed on the GPU (spr_mlp_score):
    triplets with min |w0 -+ w1 -+ w2| < window sigma, weight exp(sd f(x)) |G|.  prop=None gives
    the G-only proposal (no network).  Returns (S (n, nq), n_evals)."""
    from . import lib
    if prop is None:
        return lib.score(ctx, None, q0_index, sigma, window)
    return lib.score(ctx, prop.mlp, q0_index, sigma, window, prop.sd, prop.wmax)


def mixture(S, lam=0.1):
    nq = S.size
    tot = S.sum()
    base = S / tot if tot > 0 else np.full(nq, 1.0 / nq)
    return (1.0 - lam) * base + lam / nq


def head_tail_draw(S, K, head_frac, rng, lam=0.1):
    """K pair evaluations for one q0 split into a deterministic head and a sampled tail: the
    H = round(head_frac K) q1 of highest score (any H of them when scores tie) are evaluated
    exactly (weight 1), the other
    K - H are drawn with replacement from the mixture proposal restricted to the remaining q1
    (weight 1 / ((K - H) p)).  sum(weight * term) is an unbiased estimate of the full sum; the
    tail only has to cover what the head leaves, so a score that ranks the large terms first
    cuts the variance.  Returns (q1 indices, weights)."""
    nq = S.size
    H = min(nq, int(round(head_frac * K)))
    if 0 < H < nq:   # the H highest scores (any H of them when scores tie), in linear time
        part = np.argpartition(-S, H - 1)
        head, rest = part[:H], np.sort(part[H:])
    else:
        order = np.arange(nq)
        head, rest = order[:H], order[H:]
    M = K - H
    if M <= 0 or rest.size == 0:
        return head, np.ones(head.size)
    pr = mixture(S[rest], lam)
    pick = rng.choice(rest.size, size=M, replace=True, p=pr)
    return np.concatenate([head, rest[pick]]), np.concatenate([np.ones(H), 1.0 / (M * pr[pick])])


def is_rms_error(terms, p, K):
    """Relative RMS error of the K-sample importance estimate of sum(terms) (per s0),
    exact expectation: Var = (sum t^2/p - (sum t)^2) / K."""
    tot = terms.sum(axis=0)
    var = ((terms ** 2) / p[:, None]).sum(axis=0) - tot ** 2
    var = np.maximum(var, 0.0) / K
    with np.errstate(divide="ignore", invalid="ignore"):
        return np.sqrt(var) / np.abs(tot)


def sample_estimate(ctx, qi, p, K, sigma, prec, rng):
    """Draw K q1 from p and return the importance estimate (1/K) sum t(q1)/p(q1) (Ndof), the
    weighted sum formed on the GPU (spr_lw3_sampled)."""
    q1 = rng.choice(p.size, size=K, replace=True, p=p)
    return ctx.lw3_sampled(np.full(K, qi), q1, 1.0 / (K * p[q1]), sigma, prec)[qi]


PAIR_NF = 12


def pair_features(ctx, q0_index, sigma, wmax, window=3.0):
    """(S, X): the G-only pair weights (n, nq) and the pair feature rows (n, nq, 12)."""
    from . import lib
    return lib.pair_scores(ctx, None, q0_index, sigma, window, wmax, want_x=True)


def pair_targets(t):
    """log of each q1's share of the linewidths of one q0, normalised like the features:
    t (nq, Ndof) pair terms -> log(T Nq / n_s0), T = sum_s0 |t| / sum_q1 |t| over s0 with a
    nonzero sum; floored at 1e-12."""
    a = np.abs(t)
    L = a.sum(axis=0)
    act = L > 1e-12 * max(L.max(), 1e-300)
    if not act.any():
        return np.full(t.shape[0], np.log(1e-12), np.float32)
    T = (a[:, act] / L[act]).sum(axis=1)
    return np.log(np.maximum(T * t.shape[0] / act.sum(), 1e-12)).astype(np.float32)


def pair_training_set(ctx, qm, ph, q0_train, sigma, prec="fp32", window=3.0, chunk=64):
    """(X, y) over every q1 of the training q0: X the pair features, y = pair_targets of the
    pair terms the method computes (spr_lw3_pairs in `prec`)."""
    wmax = float(np.abs(ph.omega).max())
    Xs, ys = [], []
    q0_train = np.asarray(q0_train)
    for lo in range(0, q0_train.size, chunk):
        qs = q0_train[lo:lo + chunk]
        _, X = pair_features(ctx, qs, sigma, wmax, window)
        q0s = np.repeat(qs, qm.nq)
        q1s = np.tile(np.arange(qm.nq), qs.size)
        t = ctx.lw3_pairs(q0s, q1s, sigma, prec).reshape(qs.size, qm.nq, -1)
        for k in range(qs.size):
            Xs.append(X[k])
            ys.append(pair_targets(t[k]))
    return np.concatenate(Xs), np.concatenate(ys)


class PairProposal:
    """Pair-level network: features of (q0, q1) -> log of the pair's share of the linewidths."""

    def __init__(self, mlp, y_mean, y_std, wmax):
        self.mlp, self.mu, self.sd, self.wmax = mlp, y_mean, y_std, wmax

    @classmethod
    def train(cls, X, y, wmax, epochs=10, batch=4096, lr=2e-3, seed=0):
        mu, sd = float(y.mean()), float(y.std() + 1e-12)
        mlp = Mlp(PAIR_NF, seed=seed)
        rng = np.random.default_rng(seed)
        perm = rng.permutation(y.size)
        losses = mlp.train(X[perm], ((y - mu) / sd)[perm], epochs=epochs, batch=batch, lr=lr)
        p = cls(mlp, mu, sd, wmax)
        p.losses = losses
        return p


def pair_proposal_scores(prop, ctx, q0_index, sigma, wmax, window=3.0):
    """Scores S(q1) (n, nq) of the pair-level proposal: exp of the network's log-share prediction
    (prop given) or the G-only weight S (prop None); both computed on the device."""
    from . import lib
    if prop is None:
        return lib.pair_scores(ctx, None, q0_index, sigma, window, wmax)
    f = lib.pair_scores(ctx, prop.mlp, q0_index, sigma, window, wmax)
    return np.exp(np.minimum(f * prop.sd, 60.0))
