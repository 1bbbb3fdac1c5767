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
