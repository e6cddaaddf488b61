"""Bermudan swaptions on the GPU against the oracle on shared host draws: the Longstaff-Schwartz
policy (coefficients, exercise years, pathwise cashflows, prices), its forward application, the
polynomial judge, and the BF16 continuation network (forward, gradient, Adam step) against the FP64
restatement with the same BF16 rounding points. Then training on the GPU: the loss falls, the
network's exposure agrees with the judge, and on single-exercise contracts, whose continuation value
is known in closed form, the network is compared with the polynomial path by path."""
import numpy as np
import pytest

import oracle
from oracle import mlp as omlp
from oracle import swaption
from engine_quant_amd.berm import BermudanBook, MlpConfig
from engine_quant_amd.portfolios import bermudan_book

pytestmark = pytest.mark.gpu

MODEL = (0.10, 0.03, 0.01, 0.02)
SPY, HORIZON = 4, 12


def _book(n=6, seed=3):
    rng = np.random.default_rng(seed)
    first = rng.integers(1, 5, n).astype(np.int32)
    n_ex = rng.integers(1, 7, n).astype(np.int32)
    n_ex = np.minimum(n_ex, HORIZON - first).astype(np.int32)
    strike = 0.028 + rng.normal(0, 0.004, n)
    notional = np.exp(rng.normal(np.log(1e6), 0.5, n))
    return dict(notional=notional, strike=strike, first_ex=first, n_ex=n_ex)


@pytest.fixture(scope="module")
def shared(ctx):
    bk = _book()
    n_tr, n_in = 3000, 2500
    zt = np.random.default_rng(11).standard_normal((SPY * HORIZON, n_tr))
    zi = np.random.default_rng(12).standard_normal((SPY * HORIZON, n_in))
    B = BermudanBook(ctx, MODEL, bk["notional"], bk["strike"], bk["first_ex"], bk["n_ex"], n_tr, n_in, SPY, HORIZON,
                     z_train=zt, z_infer=zi)
    rt, lbt = oracle.berm_simulate(MODEL, 1.0 / SPY, zt)
    ri, lbi = oracle.berm_simulate(MODEL, 1.0 / SPY, zi)
    yield dict(B=B, bk=bk, rt=rt, lbt=lbt, ri=ri, lbi=lbi)
    B.close()


def test_policy_matches_oracle(shared):
    B, bk = shared["B"], shared["bk"]
    pol = B.fit_policy()
    ref = oracle.berm_policy(MODEL, bk, SPY, shared["rt"], shared["lbt"])
    tau, y0 = B.pathwise(0)
    assert np.array_equal(tau, ref["tau"])
    assert np.allclose(y0, ref["y0"], rtol=1e-12, atol=1e-9)
    assert np.allclose(pol["ex_coef"], ref["ex_coef"], rtol=1e-8, atol=1e-6 * np.abs(ref["ex_coef"]).max())
    assert np.allclose(pol["price_train"], ref["price"], rtol=1e-11)
    # forward application on the inference set with the oracle's own policy
    t_i, y_i = oracle.berm_apply(MODEL, bk, SPY, shared["ri"], shared["lbi"], ref["ex_coef"])
    tau_i, y0_i = B.pathwise(1)
    assert np.array_equal(tau_i, t_i)
    assert np.allclose(y0_i, y_i, rtol=1e-12, atol=1e-9)
    assert np.allclose(pol["price_infer"], y_i.mean(axis=1), rtol=1e-11)


def test_judge_matches_oracle(shared):
    B, bk = shared["B"], shared["bk"]
    ee, coef = B.exposure_lsm(1)
    tau, y0 = B.pathwise(1)
    for j in range(len(bk["notional"])):
        kmax = (bk["first_ex"][j] + bk["n_ex"][j] - 1) * SPY
        for k in range(1, SPY * HORIZON + 1):
            if k > kmax:
                assert ee[j, k - 1] == 0.0
                continue
            c, n_alive, e = oracle.berm_judge(MODEL, SPY, shared["ri"], shared["lbi"], tau[j], y0[j], k)
            assert np.allclose(coef[j, k - 1], c, rtol=1e-8, atol=1e-7 * max(1.0, np.abs(c).max()))
            assert abs(ee[j, k - 1] - e) <= 1e-9 * max(abs(e), 1.0), (j, k, ee[j, k - 1], e)


def _random_params(n, seed):
    rng = np.random.default_rng(seed)
    th = np.zeros((n, omlp.N_PARAMS), dtype=np.float32)
    th[:, omlp.OFF_W1:omlp.OFF_W2] = rng.normal(0, 0.7, (n, 256))
    th[:, omlp.OFF_W2:omlp.OFF_B2] = rng.normal(0, 0.15, (n, 4096))
    th[:, omlp.OFF_B2:omlp.OFF_W3] = rng.normal(0, 0.1, (n, 64))
    th[:, omlp.OFF_W3:omlp.OFF_B3] = rng.normal(0, 0.15, (n, 4096))
    th[:, omlp.OFF_B3:omlp.OFF_W4] = rng.normal(0, 0.1, (n, 64))
    th[:, omlp.OFF_W4:omlp.OFF_B4] = rng.normal(0, 0.2, (n, 64))
    th[:, omlp.OFF_B4] = 0.1
    return th


def _samples(n_nets, n, seed):
    rng = np.random.default_rng(seed)
    u = rng.uniform(-1.5, 1.5, (n_nets, n, 3)).astype(np.float32)
    t = rng.normal(0.4, 0.6, (n_nets, n)).astype(np.float32)
    m = (rng.uniform(size=(n_nets, n)) < 0.85).astype(np.float32)
    return u, t, m


def test_network_forward_matches_oracle(shared):
    B = shared["B"]
    th = _random_params(B.n_trades, 5)
    B.set_params(MlpConfig(), th)
    u, _, _ = _samples(B.n_trades, 200, 6)
    y = B.mlp_forward(u)
    for j in range(B.n_trades):
        ref, _ = omlp.forward(th[j].astype(np.float64), u[j].astype(np.float64), emulate_bf16=True)
        # FP32 accumulation and the hardware tanh against FP64: a BF16 rounding of an activation can
        # land on the other side of a tie, so allow a few BF16 ulps of the output scale
        assert np.max(np.abs(y[j] - ref)) < 5e-3 * max(1.0, np.abs(ref).max()), j
        assert np.mean(np.abs(y[j] - ref)) < 2e-4 * max(1.0, np.abs(ref).max()), j


def test_network_gradient_matches_oracle(shared):
    B = shared["B"]
    th = _random_params(B.n_trades, 7)
    B.set_params(MlpConfig(), th)
    n = 256
    u, t, m = _samples(B.n_trades, n, 8)
    loss, grad = B.mlp_grad(u, t, m)
    for j in range(B.n_trades):
        rl, rg = omlp.loss_and_grad(th[j].astype(np.float64), u[j].astype(np.float64), t[j].astype(np.float64),
                                    m[j].astype(np.float64), n, emulate_bf16=True)
        assert abs(loss[j] - rl) <= 2e-3 * rl, (j, loss[j], rl)
        for lo, hi in ((0, 256), (256, 4352), (4352, 4416), (4416, 8512), (8512, 8576), (8576, 8641)):
            err = np.linalg.norm(grad[j, lo:hi] - rg[lo:hi]) / np.linalg.norm(rg[lo:hi])
            assert err < 1e-2, (j, lo, hi, err)


def test_adam_step_matches_oracle(shared):
    B = shared["B"]
    th = _random_params(B.n_trades, 9)
    cfg = MlpConfig()
    B.set_params(cfg, th)
    u, t, m = _samples(B.n_trades, 128, 10)
    _, g = B.mlp_grad(u, t, m)
    B.adam_step(u, t, m, 1, 1e-3)
    after = B.get_params()
    for j in range(B.n_trades):
        ref, _, _ = omlp.adam(th[j], np.zeros(omlp.N_PARAMS), np.zeros(omlp.N_PARAMS), g[j], 1, 1e-3,
                              cfg.beta1, cfg.beta2, cfg.eps)
        assert np.allclose(after[j], ref, rtol=1e-6, atol=1e-9)


def test_training_reduces_loss_and_tracks_the_judge(ctx):
    bk = bermudan_book(24, seed=5)
    B = BermudanBook(ctx, MODEL, bk["notional"], bk["strike"], bk["first_ex"], bk["n_ex"], 1 << 15, 1 << 15, 4, 30)
    B.fit_policy()
    ee_poly, _ = B.exposure_lsm(1)
    cfg = MlpConfig.for_model(MODEL, 30, steps=600, batch=512)
    curve = B.train(cfg)
    assert np.all(np.isfinite(curve))
    assert curve[-50:].mean() < 0.7 * curve[:20].mean()
    ee_net = B.exposure_mlp(1)
    err = np.abs(ee_net - ee_poly).sum() / ee_poly.sum()
    assert err < 0.05, err
    B.close()


def test_surrogate_meets_the_per_trade_gate(ctx):
    """100 trades of the benchmark book's generator, 2^18 training paths, 2^18 fresh inference paths:
    the network's EE against the judge's, summed over trades and dates, weighted by notional (every
    EE carries its trade's notional), within 0.5%, both through the node lattice and path by path."""
    bk = bermudan_book(100, seed=7)
    B = BermudanBook(ctx, MODEL, bk["notional"], bk["strike"], bk["first_ex"], bk["n_ex"], 1 << 18, 1 << 18, 4, 30,
                     seed_train=707, seed_infer=1477)
    B.fit_policy()
    ee_lsm, _ = B.exposure_lsm(1)
    cfg = MlpConfig.for_model(MODEL, 30, steps=20000, batch=4096, lr=3e-3, lr_final=1e-5)
    B.train(cfg)
    ee_dir = B.exposure_mlp(1)
    ee_lat = B.exposure_mlp(1, lattice=(128, 8))
    B.close()
    err_dir = np.abs(ee_dir - ee_lsm).sum() / ee_lsm.sum()
    err_lat = np.abs(ee_lat - ee_lsm).sum() / ee_lsm.sum()
    print(f"notional-weighted trade EE error: per path {err_dir:.4%}, lattice {err_lat:.4%}")
    assert err_dir <= 0.005 and err_lat <= 0.005, (err_dir, err_lat)
    assert np.abs(ee_lat - ee_dir).sum() / ee_dir.sum() <= 1e-3


def test_lattice_exposure_against_the_per_path_network(ctx):
    """The lattice pass on a small book with host draws: it converges to the per-path pass as the
    lattice refines, rows after a trade's last date stay zero, and bad lattices are refused."""
    bk = _book(n=5, seed=8)
    n_tr, n_in = 4096, 3000  # a partial last chunk
    zt = np.random.default_rng(31).standard_normal((SPY * HORIZON, n_tr))
    zi = np.random.default_rng(32).standard_normal((SPY * HORIZON, n_in))
    B = BermudanBook(ctx, MODEL, bk["notional"], bk["strike"], bk["first_ex"], bk["n_ex"], n_tr, n_in, SPY, HORIZON,
                     z_train=zt, z_infer=zi)
    B.fit_policy()
    B.train(MlpConfig.for_model(MODEL, HORIZON, steps=300, batch=512))
    ee_dir = B.exposure_mlp(1)
    errs = []
    for lat in ((16, 2), (32, 8), (128, 8)):
        ee = B.exposure_mlp(1, lattice=lat)
        kmax = (bk["first_ex"] + bk["n_ex"] - 1) * SPY
        for j in range(B.n_trades):
            assert np.all(ee[j, kmax[j]:] == 0.0)
        errs.append(np.abs(ee - ee_dir).sum() / ee_dir.sum())
    assert errs[2] < errs[1] < errs[0] and errs[2] < 5e-3, errs
    for bad in ((3, 32), (1, 64), (64, 32)):
        with pytest.raises(Exception, match="lattice"):
            B.exposure_mlp(1, lattice=bad)
    B.close()


def test_network_beats_cubic_on_closed_form_contracts(ctx):
    """One-period payer swaptions (single exercise): the continuation value at t < E is the bond
    put of oracle.swaption. Both estimators are out of sample: the cubic per date fitted on the
    training set, the network trained on it; errors are measured path by path on the inference set."""
    first = np.array([2, 4, 6, 8], dtype=np.int32)
    n_ex = np.ones(4, dtype=np.int32)
    strike = np.array([0.024, 0.027, 0.03, 0.033])
    notional = np.full(4, 1e6)
    n_tr, n_in, spy, hz = 1 << 16, 1 << 13, 4, 10
    zt = oracle.philox_normals(21, n_tr, spy * hz)
    zi = oracle.philox_normals(22, n_in, spy * hz)
    B = BermudanBook(ctx, MODEL, notional, strike, first, n_ex, n_tr, n_in, spy, hz, z_train=zt, z_infer=zi)
    B.fit_policy()
    _, coef_tr = B.exposure_lsm(0)
    cfg = MlpConfig.for_model(MODEL, hz, steps=1500, batch=1024)
    B.train(cfg)
    r, lb = oracle.berm_simulate(MODEL, 1.0 / spy, zi)
    err_net, err_poly = [], []
    for j in range(4):
        ks = np.arange(1, first[j] * spy)          # dates strictly before the exercise date
        t = ks / spy
        u = np.stack([np.stack(cfg.inputs(r[k], lb[k], k / spy), axis=1) for k in ks])  # [dates][paths][3]
        u_all = np.zeros((B.n_trades, u.shape[0] * u.shape[1], 3), dtype=np.float32)
        u_all[j] = u.reshape(-1, 3)
        y = B.mlp_forward(u_all)[j].reshape(len(ks), -1) * notional[j] / cfg.y_scale
        for i, k in enumerate(ks):
            truth = swaption.payer_one_period(MODEL, r[k], t[i], float(first[j]), notional[j], strike[j])
            a, b, sigma, r0 = MODEL
            x = (r[k] - (b + (r0 - b) * np.exp(-a * t[i]))) / (sigma * np.sqrt((1.0 - np.exp(-2.0 * a * t[i])) / (2.0 * a)))
            c = coef_tr[j, k - 1]
            poly = c[0] + x * (c[1] + x * (c[2] + x * c[3]))
            err_net.append((y[i] - truth) / notional[j])
            err_poly.append((poly - truth) / notional[j])
    rmse_net = np.sqrt(np.mean(np.concatenate(err_net) ** 2))
    rmse_poly = np.sqrt(np.mean(np.concatenate(err_poly) ** 2))
    print(f"closed-form contracts: RMSE per unit notional, network {rmse_net:.3e}, cubic {rmse_poly:.3e}")
    assert rmse_net < rmse_poly
    B.close()
