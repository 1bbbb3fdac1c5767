"""Device MLP (training gradients against the FP64 oracle, training, BF16 inference) and the
learned importance-sampling proposal against uniform sampling at a 5 % budget."""
import numpy as np
import pytest

from tests.conftest import problem

pytestmark = pytest.mark.gpu


def test_mlp_gradient_matches_oracle():
    from spring_amd.lib import Mlp
    from oracle import mlp_oracle as M
    rng = np.random.default_rng(0)
    F = 12
    mlp = Mlp(F, seed=3)
    p = mlp.params().astype(np.float64)
    X = rng.normal(size=(200, F)).astype(np.float32)
    y = rng.normal(size=200).astype(np.float32)
    g, loss = mlp.grad(X, y)
    loss_ref, g_ref = M.loss_grad(p, X.astype(np.float64), y.astype(np.float64))
    assert abs(loss - loss_ref) < 1e-4 * abs(loss_ref)
    assert np.max(np.abs(g - g_ref)) < 1e-4 * np.max(np.abs(g_ref))


def test_mlp_trains_and_bf16_inference():
    from spring_amd.lib import Mlp
    rng = np.random.default_rng(1)
    X = rng.uniform(-1, 1, size=(20000, 4)).astype(np.float32)
    y = (np.sin(2 * X[:, 0]) + X[:, 1] * X[:, 2] - 0.5 * X[:, 3] ** 2).astype(np.float32)
    mlp = Mlp(4, seed=0)
    losses = mlp.train(X, y, epochs=20, batch=1024, lr=3e-3)
    print("losses", losses[0], losses[-1], flush=True)
    assert losses[-1] < 0.05 * losses[0]
    p32 = mlp.predict(X[:3000], "fp32")
    pbf = mlp.predict(X[:3000], "bf16")
    assert np.mean((p32 - y[:3000]) ** 2) < 0.05 * np.var(y)
    assert np.max(np.abs(pbf - p32)) < 0.1 * np.std(y)


def test_pair_terms_sum_to_linewidth(rocksalt):
    from spring_amd.lib import GpuContext
    qm, ph = problem(rocksalt, (6, 6, 6))
    ctx = GpuContext(rocksalt, qm, ph)
    lw = ctx.lw3_all(0.1, "fp64")
    for qi in (0, 3, qm.nq0 - 1):
        t = ctx.lw3_pairs(np.full(qm.nq, qi), np.arange(qm.nq), 0.1, "fp64")
        assert np.allclose(t.sum(axis=0), lw[qi], rtol=1e-10, atol=1e-14)


def test_learned_proposal_beats_uniform(rocksalt):
    from spring_amd import learned as Lrn
    from spring_amd.lib import GpuContext
    qm, ph = problem(rocksalt, (12, 12, 12))
    ctx = GpuContext(rocksalt, qm, ph)
    sigma = 0.1
    X, y = Lrn.training_set(ctx, qm, ph, n_pairs=600, per_pair=216, seed=0)
    prop = Lrn.Proposal.train(X, y, wmax=float(np.abs(ph.omega).max()), epochs=15, batch=2048, lr=2e-3)
    print("train loss", prop.losses[0], prop.losses[-1], flush=True)
    K = int(0.05 * qm.nq)
    errs = {"uniform": [], "physics": [], "learned": []}
    for qi in (5, 17, 40, qm.nq0 - 3):
        ig = int(qm.irr_idx[qi])
        t = ctx.lw3_pairs(np.full(qm.nq, qi), np.arange(qm.nq), sigma, "fp64")
        sig = np.abs(t.sum(axis=0)) > 1e-6 * np.abs(t.sum(axis=0)).max()
        pu = np.full(qm.nq, 1.0 / qm.nq)
        pp = Lrn.mixture(Lrn.device_scores(None, ctx, [qi], sigma)[0][0])
        pl = Lrn.mixture(Lrn.device_scores(prop, ctx, [qi], sigma)[0][0])
        for name, p in (("uniform", pu), ("physics", pp), ("learned", pl)):
            errs[name].append(Lrn.is_rms_error(t, p, K)[sig])
    e = {k: float(np.mean(np.concatenate(v))) for k, v in errs.items()}
    print("mean relative RMS error at 5% budget:", e, flush=True)
    assert e["learned"] < e["uniform"]
    assert e["physics"] < e["uniform"]


def test_device_scores_match_host_definition(rocksalt):
    """spr_mlp_score (features, resonance window, BF16 network, exp(sd f) |G| sums on the
    device) against the same definition evaluated on the host with the FP32 network."""
    from spring_amd import learned as Lrn
    from spring_amd.lib import GpuContext, Mlp
    qm, ph = problem(rocksalt, (6, 6, 6))
    ctx = GpuContext(rocksalt, qm, ph)
    mlp = Mlp(Lrn.NF, seed=5)
    prop = Lrn.Proposal(mlp, 0.3, 1.7, float(np.abs(ph.omega).max()))
    sigma, window = 0.3, 3.0
    qis = [1, 4, qm.nq0 - 1]
    S, n_eval = Lrn.device_scores(prop, ctx, qis, sigma, window)
    n_ref = 0
    for k, qi in enumerate(qis):
        ig = int(qm.irr_idx[qi])
        q1 = np.arange(qm.nq)
        q2 = qm.index(-qm.q_int[ig][None, :].astype(np.int64) - qm.q_int.astype(np.int64))
        w0 = ph.omega[ig][None, :, None, None]
        w1 = ph.omega[q1][:, None, :, None]
        w2 = ph.omega[q2][:, None, None, :]
        d = np.minimum(np.abs(w0 - w1 - w2), np.minimum(np.abs(w0 + w1 - w2), np.abs(w0 - w1 + w2)))
        sel = np.nonzero(d < window * sigma)
        G = np.abs(Lrn.gauss_weights(ph, ig, q1, q2, sigma))
        X = Lrn.features(qm, ph, ig, q1[sel[0]], sel[1], sel[2], sel[3], prop.wmax)
        f = mlp.predict(X, "fp32")
        ref = np.zeros(qm.nq)
        np.add.at(ref, sel[0], np.exp(prop.sd * f) * G[sel])
        n_ref += X.shape[0]
        big = ref > 1e-3 * ref.max()
        assert np.max(np.abs(S[k] - ref)[big] / ref[big]) < 5e-2
        assert abs(S[k].sum() - ref.sum()) < 2e-2 * ref.sum()
        assert np.all(S[k][ref == 0] == 0)
    assert n_eval == n_ref


def _host_pair_features(qm, ph, ig, sigma, window, wmax):
    """spr_pair_scores' definition evaluated on the host (FP64) for one q0: (S, X)."""
    from spring_amd import learned as Lrn
    nd = ph.omega.shape[1]
    q1 = np.arange(qm.nq)
    q2 = qm.index(-qm.q_int[ig][None, :].astype(np.int64) - qm.q_int.astype(np.int64))
    w0 = ph.omega[ig][None, :, None, None]
    w1 = ph.omega[q1][:, None, :, None]
    w2 = ph.omega[q2][:, None, None, :]
    d = np.minimum(np.abs(w0 - w1 - w2), np.minimum(np.abs(w0 + w1 - w2), np.abs(w0 - w1 + w2)))
    live = (ph.inv_omega[ig] != 0.0)[None, :, None, None]
    inw = (d < window * sigma) & live
    wgt = np.abs(Lrn.gauss_weights(ph, ig, q1, q2, sigma)) * ph.inv_omega[q1][:, None, :, None] * \
        ph.inv_omega[q2][:, None, None, :] * inw
    g = wgt.sum(axis=(2, 3))
    h1 = (wgt * w1).sum(axis=(2, 3))
    h2 = (wgt * w2).sum(axis=(2, 3))
    ac = np.arange(nd) < 3
    a = (wgt * (ac[None, None, :, None] | ac[None, None, None, :])).sum(axis=(2, 3))
    cnt = inw.sum(axis=(1, 2, 3))
    gt = g.sum(axis=0)
    act = gt > 0
    S = (g[:, act] / gt[act]).sum(axis=1)
    fold = lambda k, n: np.where(k > n // 2, k - n, k) / n
    X = np.zeros((qm.nq, 12))
    k0 = np.array([ig % qm.mesh[0], (ig // qm.mesh[0]) % qm.mesh[1], ig // (qm.mesh[0] * qm.mesh[1])])
    X[:, 0:3] = fold(k0, qm.mesh)[None, :]
    X[:, 3:6] = fold(np.stack([q1 % qm.mesh[0], (q1 // qm.mesh[0]) % qm.mesh[1], q1 // (qm.mesh[0] * qm.mesh[1])], 1),
                     qm.mesh[None, :])
    X[:, 6] = np.log(np.maximum(S * qm.nq / max(act.sum(), 1), 1e-12))
    with np.errstate(divide="ignore", invalid="ignore"):
        X[:, 7] = np.where(S > 0, (h1[:, act] / gt[act]).sum(1) / S / wmax, 0.0)
        X[:, 8] = np.where(S > 0, (h2[:, act] / gt[act]).sum(1) / S / wmax, 0.0)
        X[:, 9] = np.where(S > 0, (g[:, act] / gt[act] * ph.omega[ig][act]).sum(1) / S / wmax, 0.0)
        X[:, 10] = np.where(S > 0, (a[:, act] / gt[act]).sum(1) / S, 0.0)
    X[:, 11] = np.log1p(cnt) / np.log(nd ** 3)
    return S, X


def test_pair_features_match_host_definition(rocksalt):
    """spr_pair_scores' resonance weights and feature rows against the host definition."""
    from spring_amd import learned as Lrn
    from spring_amd.lib import GpuContext
    qm, ph = problem(rocksalt, (5, 5, 5))
    ctx = GpuContext(rocksalt, qm, ph)
    wmax = float(np.abs(ph.omega).max())
    qis = [0, 3, qm.nq0 - 1]
    S, X = Lrn.pair_features(ctx, qis, 0.3, wmax, window=3.0)
    assert S[1].sum() > 0 and S[2].sum() > 0
    for k, qi in enumerate(qis):
        Sr, Xr = _host_pair_features(qm, ph, int(qm.irr_idx[qi]), 0.3, 3.0, wmax)
        assert np.allclose(S[k], Sr, rtol=1e-5, atol=1e-7 * Sr.max()), qi
        assert abs(S[k].sum() - Sr.sum()) <= 1e-5 * Sr.sum()
        assert np.allclose(X[k], Xr, rtol=1e-4, atol=1e-4), qi
    ctx.close()


def test_pair_proposal_beats_g_only_10atom(rocksalt10):
    """The pair-level learned proposal on the 10-atom cell (10^3 mesh with time reversal, 504
    irreducible q0 -- the largest mesh this test's budget allows; the 20^3 form is the bench's
    learned10 line): trained on every (q0, q1) pair term of 120 random q0 (120,000 samples the
    method computes), it is judged on 40 other q0 by the exact relative RMS error of the 5 %
    sampled estimate, (sum t^2/p - (sum t)^2)/K, against the G-only proposal and uniform
    sampling at the same budget."""
    from spring_amd import learned as Lrn
    from spring_amd.lib import GpuContext
    qm, ph = problem(rocksalt10, (10, 10, 10), time_rev=1)
    ctx = GpuContext(rocksalt10, qm, ph)
    sigma = 0.1
    wmax = float(np.abs(ph.omega).max())
    perm = np.random.default_rng(0).permutation(qm.nq0)
    train, held = perm[:120], perm[120:160]
    X, y = Lrn.pair_training_set(ctx, qm, ph, train, sigma, prec="fp32")
    prop = Lrn.PairProposal.train(X, y, wmax, epochs=10)
    print("pair proposal: samples", y.size, "loss", prop.losses[0], prop.losses[-1], flush=True)
    K = int(0.05 * qm.nq)
    Sg = Lrn.pair_proposal_scores(None, ctx, held, sigma, wmax)
    Sl = Lrn.pair_proposal_scores(prop, ctx, held, sigma, wmax)
    errs = {"uniform": [], "g_only": [], "learned": []}
    for k, qi in enumerate(held):
        t = ctx.lw3_pairs(np.full(qm.nq, qi), np.arange(qm.nq), sigma, "fp64")
        L = np.abs(t.sum(axis=0))
        sig = L > 1e-6 * L.max()
        for name, p in (("uniform", np.full(qm.nq, 1.0 / qm.nq)), ("g_only", Lrn.mixture(Sg[k])),
                        ("learned", Lrn.mixture(Sl[k]))):
            errs[name].append(Lrn.is_rms_error(t, p, K)[sig])
    e = {k: float(np.mean(np.concatenate(v))) for k, v in errs.items()}
    print("mean relative RMS error at 5% budget:", e, flush=True)
    assert e["learned"] < 0.9 * e["g_only"]
    assert e["learned"] < 0.9 * e["uniform"]
    ctx.close()
