"""Golden fixtures (tests/golden/, written by tests/golden/make_fixtures.py): the oracle reproduces them
exactly (a regression pin of the checker itself), and the HIP path matches them on the fixtures' explicit
host draws within the parity tolerances (1e-12 of the legs for per-path MtM, 1e-10 for EE/PFE/CVA)."""
import json
import os

import numpy as np
import pytest

import oracle

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    with open(os.path.join(HERE, name)) as f:
        return json.load(f)


def _arrays(port):
    return {k: np.asarray(v, dtype=np.int32 if k in ("set_id", "flow_off") else np.float64) for k, v in port.items()}


def test_oracle_reproduces_netting_fixture():
    fx = _load("netting_tiny.json")
    z = np.asarray(fx["z"])
    for scheme, name in ((0, "euler"), (1, "exact")):
        ref = oracle.portfolio_exposure(tuple(fx["model"]), _arrays(fx["portfolio"]), fx["dates"], fx["n_steps"], z,
                                        scheme=scheme, hazard_rate=fx["hazard_rate"],
                                        recovery_rate=fx["recovery_rate"], want_trade=True, want_path=True)
        for k, v in fx["schemes"][name].items():
            assert np.array_equal(np.asarray(ref[k]), np.asarray(v)), (name, k)


def test_oracle_reproduces_reference_profile_fixture():
    fx = _load("reference_profile_tiny.json")
    t, ee, pfe = oracle.exposure_profile(*fx["model"], fx["notional"], fx["fixed_rate"], False, 0.0,
                                         fx["payment_times"], fx["accruals"], fx["monitoring_times"], fx["n_steps"],
                                         np.asarray(fx["z"]))
    assert list(t) == fx["times"] and list(ee) == fx["ee"] and list(pfe) == fx["pfe95"]


def test_philox_fixture():
    fx = _load("philox_kat.json")
    for e in fx["kat"]:
        assert list(oracle.philox4x32_10(e["counter"], e["key"])) == e["out"]
    assert np.array_equal(oracle.philox_normals(7, 4, 4), np.asarray(fx["normals_seed7_4paths_4steps"]))


def test_reference_values_against_oracle():
    rv = _load("reference_values.json")
    pay, acc = [1.0, 2.0, 3.0, 4.0, 5.0], [1.0] * 5
    assert oracle.npv_from_market([1.0, 2.0], [0.02, 0.02], 1e6, 0.02, 0.0, pay, acc) == rv["config1_pv"]["value"]
    assert oracle.dv01([1.0, 2.0], [0.02, 0.02], 1e6, 0.02, 0.0, pay, acc) == rv["config1_dv01"]["value"]
    h = rv["handmade_cva_profiles"]
    a = oracle.unilateral_cva(0.1, 0.03, 0.01, 0.02, h["times"], h["ee_a"], h["hazard_rate"], h["recovery_rate"])
    b = oracle.unilateral_cva(0.1, 0.03, 0.01, 0.02, h["times"], h["ee_b"], h["hazard_rate"], h["recovery_rate"])
    assert abs(b - 2.0 * a) <= 1e-12 * abs(b)  # linear in EE (exposure.rs:729-759)


@pytest.mark.gpu
def test_gpu_matches_netting_fixture(ctx):
    from engine_quant_amd import eqhip
    from engine_quant_amd.eqhip import Portfolio
    from helpers import assert_rel, assert_value_rel, leg_scale
    fx = _load("netting_tiny.json")
    p = _arrays(fx["portfolio"])
    port = Portfolio(set_id=p["set_id"], notional=p["notional"], fixed_rate=p["fixed_rate"], start=p["start"],
                     flow_off=p["flow_off"], pay=p["pay"], acc=p["acc"])
    z = np.asarray(fx["z"])
    model = tuple(fx["model"])
    for disc, name in ((eqhip.DISC_EULER, "euler"), (eqhip.DISC_EXACT, "exact")):
        ref = {k: np.asarray(v) for k, v in fx["schemes"][name].items()}
        got = ctx.exposure_run(model, port, fx["dates"], z.shape[1], fx["n_steps"], rng=eqhip.RNG_HOST_DRAWS,
                               disc=disc, host_draws=z, hazard_rate=fx["hazard_rate"],
                               recovery_rate=fx["recovery_rate"], want_trade_ee=True, want_path=True)
        scale = leg_scale(port, model, fx["dates"])
        set_scale = np.array([scale[port.set_id == s].sum() for s in range(port.n_sets)])
        assert_rel(got["path_mtm"], ref["path_mtm"], 1e-12, scale=scale[None, :, None], what=f"{name} path MtM")
        assert_value_rel(got["set_ee"], ref["set_ee"], leg=set_scale, what=f"{name} set EE")
        assert_value_rel(got["set_pfe"], ref["set_pfe"], leg=set_scale, what=f"{name} set PFE")
        assert_value_rel(got["set_cva"], ref["set_cva"], leg=set_scale, what=f"{name} set CVA")
        assert_value_rel(got["trade_ee"], ref["trade_ee"], leg=scale, what=f"{name} trade EE")


@pytest.mark.gpu
def test_gpu_matches_reference_profile_fixture(ctx):
    from engine_quant_amd import eqhip
    fx = _load("reference_profile_tiny.json")
    z = np.asarray(fx["z"])
    a, b, s, r0 = fx["model"]
    t, ee, pfe = ctx.exposure_profile(a, b, s, r0, fx["notional"], fx["fixed_rate"], False, 0.0, fx["payment_times"],
                                      fx["accruals"], fx["monitoring_times"], fx["n_steps"], z.shape[1], 0,
                                      rng=eqhip.RNG_HOST_DRAWS, host_draws=z)
    assert list(t) == fx["times"]
    assert np.allclose(ee, fx["ee"], rtol=1e-10, atol=1e-6) and np.allclose(pfe, fx["pfe95"], rtol=1e-10, atol=1e-6)
