"""Exposure-cube wire format (quant-domain/src/xva.rs:176-220) and the oracle of the XVA reducer
(quant-engine/src/xva.rs:14-246), on the CPU.

The oracle is pinned by the reference's own tests of the reducer (xva.rs:306-386: zero hazard and full
recovery give zero CVA; symmetric Q and P cubes give CVA = DVA; a 2 x 2 cube is 32 bytes; a P-tagged
Q cube and a foreign source hash are refused) and by closed forms a slip would break: the quantile rule
sorted[round((n - 1) p)] against a brute-force sort, the CSA adjustment case by case, survival
probabilities overriding the hazard, and FVA / MVA / KVA on constant rows.
"""
import numpy as np
import pytest

import oracle
from engine_quant_amd import xva


def cube(values, measure="Q", times=(1.0, 2.0), df=(1.0, 1.0)):
    return xva.ExposureCube(measure, list(times), values, list(df))


def test_reference_reducer_tests():
    # hazard_zero_and_full_recovery_remove_default_adjustments
    _, _, m = oracle.xva_calculate([1.0, 2.0], [[10.0], [10.0]], df=[1, 1], counterparty=(0.0, 1.0, ()),
                                   own=(0.02, 0.4, ()))
    assert m["cva"] == 0.0
    # q_p_symmetry_gives_equal_cva_and_dva
    _, _, m = oracle.xva_calculate([1.0, 2.0], [[10.0, -10.0], [10.0, -10.0]], df=[1, 1],
                                   p=[[-10.0, 10.0], [-10.0, 10.0]], counterparty=(0.02, 0.4, ()),
                                   own=(0.02, 0.4, ()))
    assert abs(m["cva"] - m["dva"]) < 1e-12 and m["cva"] > 0
    # repeated_reduction_is_deterministic_and_exposes_artifact_bytes
    c = cube([[1.0, 3.0], [2.0, 4.0]])
    assert c.estimate_memory_bytes() == 32
    a = oracle.xva_calculate(c.times, c.paths, df=c.discount_factors)
    b = oracle.xva_calculate(c.times, c.paths, df=c.discount_factors)
    assert all(np.array_equal(x, y) for x, y in zip(a[:2], b[:2])) and a[2] == b[2]
    # invalid_measure_is_rejected_instead_of_silently_mixing_q_and_p / exposure_context_mismatch_is_rejected
    req = xva.XvaRequest(context_hash="sha256:ctx")
    with pytest.raises(ValueError, match="tagged Q"):
        xva._check_inputs(req, cube([[1.0], [1.0]], measure="P"), None)
    q = cube([[1.0], [1.0]])
    q.source_hash = "sha256:other"
    with pytest.raises(ValueError, match="source hash"):
        xva._check_inputs(req, q, None)
    with pytest.raises(ValueError, match="grids must match"):
        xva._check_inputs(req, cube([[1.0], [1.0]]), cube([[1.0], [1.0]], "P", times=(1.0, 3.0)))


@pytest.mark.parametrize("n", [1, 2, 19, 20, 21, 1001])
@pytest.mark.parametrize("conf", [0.0, 0.5, 0.95, 1.0])
def test_quantile_rule(n, conf):
    rng = np.random.default_rng(n)
    row = rng.normal(0, 1, n)
    ee, pfe, _ = oracle.xva_calculate([1.0], row[None, :], confidence=conf)
    pos = np.sort(np.maximum(row, 0.0))
    assert pfe[0] == pos[int(np.floor((n - 1) * conf + 0.5))]
    assert abs(ee[0] - pos.mean()) <= 1e-15 * max(1.0, pos.sum())


def test_collateral_adjustment_cases():
    csa = {"threshold": 10.0, "minimum_transfer_amount": 5.0, "independent_amount": 2.0, "variation_margin": True}
    assert oracle.collateral_adjust(100.0, csa) == 100.0 - 90.0 - 2.0
    assert oracle.collateral_adjust(-100.0, csa) == -100.0 + 90.0 + 2.0
    assert oracle.collateral_adjust(13.0, csa) == 13.0            # 3 < minimum transfer: unchanged
    assert oracle.collateral_adjust(15.0, csa) == 15.0 - 5.0 - 2.0
    assert oracle.collateral_adjust(100.0, dict(csa, variation_margin=False)) == 100.0
    assert oracle.collateral_adjust(100.0, None) == 100.0
    assert oracle.collateral_adjust(100.0, dict(csa, threshold=-5.0)) == 100.0 - 100.0 - 2.0  # max(th, 0)


def test_default_increment_and_discount_defaults():
    t = [0.5, 1.0, 2.0]
    q = np.full((3, 4), 100.0)
    # survival probabilities cover the first two dates, the hazard the third
    sp = [0.99, 0.97]
    _, _, m = oracle.xva_calculate(t, q, counterparty=(0.05, 0.4, sp))
    inc = [1 - 0.99, 0.99 - 0.97, np.exp(-0.05 * 1.0) - np.exp(-0.05 * 2.0)]
    assert abs(m["cva"] - sum(100.0 * i * 0.6 for i in inc)) < 1e-12
    # no discount factors: exp(-0 t) = 1; FVA / MVA / KVA on constant rows
    _, _, m = oracle.xva_calculate(t, q, funding_spread=0.01, borrowing_spread=0.02, capital_factor=0.08,
                                   initial_margin_total=50.0)
    yf = np.array([0.5, 0.5, 1.0])
    assert abs(m["fva"] - np.sum(100.0 * 0.01 * yf)) < 1e-12
    assert abs(m["mva"] - np.sum(50.0 * 0.02 * yf)) < 1e-12
    assert abs(m["kva"] - np.sum(100.0 * 0.08 * yf)) < 1e-12
    # DVA from the Q cube's negative parts when no P cube is given
    _, _, m = oracle.xva_calculate(t, -q, own=(0.02, 0.5, ()))
    dinc = np.exp(-0.02 * np.array([0.0, 0.5, 1.0])) - np.exp(-0.02 * np.array(t))
    assert abs(m["dva"] - np.sum(100.0 * dinc * 0.5)) < 1e-12 and m["cva"] == 0.0


def test_cube_validation_messages():
    with pytest.raises(ValueError, match="non-empty"):
        xva.ExposureCube("Q", [], np.zeros((0, 3))).validate()
    with pytest.raises(ValueError, match="equal length"):
        xva.ExposureCube("Q", [1.0, 2.0], np.zeros((3, 2))).validate()
    with pytest.raises(ValueError, match="strictly increasing"):
        xva.ExposureCube("Q", [1.0, 1.0], np.zeros((2, 2))).validate()
    with pytest.raises(ValueError, match="strictly increasing"):
        xva.ExposureCube("Q", [-1.0, 1.0], np.zeros((2, 2))).validate()
    with pytest.raises(ValueError, match="finite rows"):
        xva.ExposureCube("Q", [1.0, 2.0], [[1.0, np.nan], [0.0, 0.0]]).validate()
    with pytest.raises(ValueError, match="discount_factors"):
        xva.ExposureCube("Q", [1.0, 2.0], np.zeros((2, 2)), [1.0]).validate()
    with pytest.raises(ValueError, match="discount_factors"):
        xva.ExposureCube("Q", [1.0, 2.0], np.zeros((2, 2)), [1.0, -0.5]).validate()


def test_wire_forms_round_trip():
    rng = np.random.default_rng(3)
    c = xva.ExposureCube("P", [0.0, 0.25, 1.0], rng.normal(size=(3, 7)), [1.0, 0.99, 0.97], "sha256:abc")
    j = c.to_json()
    import json
    d = json.loads(j)
    assert set(d) == {"measure", "times", "paths", "discount_factors", "source_hash"}  # serde field names
    assert d["measure"] == "P" and len(d["paths"]) == 3 and len(d["paths"][0]) == 7
    for back in (xva.ExposureCube.from_json(j), xva.ExposureCube.from_bytes(c.to_bytes())):
        assert back.measure == "P" and back.source_hash == "sha256:abc"
        assert np.array_equal(back.times, c.times) and np.array_equal(back.paths, c.paths)
        assert np.array_equal(back.discount_factors, c.discount_factors)
    b = c.to_bytes()
    assert len(b) == 32 + 16 + 8 * (3 + 3 + 21)
    with pytest.raises(ValueError, match="truncated"):
        xva.ExposureCube.from_bytes(b[:-8])
    with pytest.raises(ValueError, match="magic"):
        xva.ExposureCube.from_bytes(b"X" * 64)
    nodf = xva.ExposureCube("Q", [1.0], [[2.0]])
    assert xva.ExposureCube.from_bytes(nodf.to_bytes()).discount_factors.size == 0


# serde_json's f64 layout (the shortest round-trip digits; decimal from 1e-5 up to 1e16, else an
# exponent without "+" or leading zeros), worked by hand from its rules for each range
JSON_FLOATS = [(0.0, "0.0"), (-0.0, "-0.0"), (1.0, "1.0"), (100.0, "100.0"), (0.1, "0.1"), (1e-4, "0.0001"),
               (1.5e-5, "0.000015"), (1e-5, "0.00001"), (1e-6, "1e-6"), (-1.5e-6, "-1.5e-6"),
               (1e15, "1000000000000000.0"), (1e16, "1e16"), (1.5e16, "1.5e16"), (123456.789, "123456.789"),
               (1.2345678901234568e17, "1.2345678901234568e17"), (5e-324, "5e-324"),
               (1.7976931348623157e308, "1.7976931348623157e308"), (2.5e-10, "2.5e-10")]


@pytest.mark.parametrize("v, text", JSON_FLOATS)
def test_json_float_layout(v, text):
    assert xva.json_float(v) == text
    assert float(text) == v


def test_artifact_id_hashes_the_reference_json():
    """artifact (quant-engine/src/xva.rs:77-91) hashes serde_json::to_value(cube) written compactly: keys in
    sorted order (serde_json's map without preserve_order), every field present (serde(default) only
    affects reading), source_hash null when absent."""
    import hashlib
    import json
    c = xva.ExposureCube("Q", [0.5, 1.0], [[1.0, -2.5], [0.0, 3e-7]], [0.99, 0.98])
    text = ('{"discount_factors":[0.99,0.98],"measure":"Q","paths":[[1.0,-2.5],[0.0,3e-7]],'
            '"source_hash":null,"times":[0.5,1.0]}')
    assert xva.canonical_json(c).decode() == text
    assert json.loads(text) == json.loads(c.to_json())
    digest = hashlib.sha256(text.encode()).hexdigest()
    assert xva.cube_hash(c) == "sha256:" + digest
    c2 = xva.ExposureCube("P", [1.0], [[2.0]], source_hash="sha256:abc")
    assert xva.canonical_json(c2) == b'{"discount_factors":[],"measure":"P","paths":[[2.0]],"source_hash":"sha256:abc","times":[1.0]}'
    req = xva.XvaRequest(context_hash="sha256:ctx")
    res = xva._result(req, c, None, np.zeros(2), np.zeros(2), np.zeros(5))
    assert res.exposure_cube["artifact_id"] == "exposure-" + digest
    assert res.exposure_cube["hash"] == "sha256:" + digest
