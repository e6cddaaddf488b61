"""Writes the golden fixtures of this directory from the oracle (run from the repository root:
``python tests/golden/make_fixtures.py``). Each fixture is data only: inputs (explicit host draws,
model, trades, dates) and the oracle's outputs on them.

* ``netting_tiny.json``: 3 trades in 2 netting sets (one forward-starting, one receiver), 8 paths x
  12 steps of explicit normals, monthly monitoring dates between resets, exact and Euler schemes:
  per-path MtM, set/trade EE, PFE95, CVA.
* ``reference_profile_tiny.json``: the reference's single-swap profile (exposure.rs:73-138) at its
  reset dates, 8 paths x 16 steps.
* ``philox_kat.json``: Philox4x32-10 known-answer vectors (Salmon et al., Random123 kat_vectors) and
  the first Box-Muller normals of the stream for seed 7.
* ``reference_values.json``: values the reference's own files hold, with their citations.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
import oracle  # noqa: E402

MODEL = (0.10, 0.03, 0.01, 0.02)


def _tiny_portfolio():
    return {"set_id": [0, 0, 1], "notional": [1e6, -5e5, 2e6], "fixed_rate": [0.021, 0.019, 0.024],
            "start": [0.0, 0.0, 0.5], "flow_off": [0, 4, 6, 10],
            "pay": [0.25, 0.5, 0.75, 1.0, 0.5, 1.0, 0.75, 1.0, 1.25, 1.5],
            "acc": [0.25, 0.25, 0.25, 0.25, 0.5, 0.5, 0.25, 0.25, 0.25, 0.25]}


def netting_tiny():
    port = _tiny_portfolio()
    arr = {k: np.asarray(v, dtype=np.int32 if k in ("set_id", "flow_off") else np.float64) for k, v in port.items()}
    dates = [k / 12.0 for k in range(0, 13)]
    z = np.random.default_rng(2024).standard_normal((12, 8))
    out = {"model": MODEL, "portfolio": port, "dates": dates, "n_steps": 12, "z": z.tolist(),
           "hazard_rate": 0.02, "recovery_rate": 0.4, "schemes": {}}
    for scheme, name in ((0, "euler"), (1, "exact")):
        ref = oracle.portfolio_exposure(MODEL, arr, dates, 12, z, scheme=scheme, hazard_rate=0.02, recovery_rate=0.4,
                                        want_trade=True, want_path=True)
        out["schemes"][name] = {k: np.asarray(v).tolist() for k, v in ref.items()}
    return out


def reference_profile_tiny():
    z = np.random.default_rng(7).standard_normal((16, 8))
    pay, acc, times = [1.0, 2.0, 3.0, 4.0, 5.0], [1.0] * 5, [0.0, 1.0, 2.0, 3.0, 4.0]
    t, ee, pfe = oracle.exposure_profile(*MODEL, 1e6, 0.02, False, 0.0, pay, acc, times, 16, z)
    return {"model": MODEL, "notional": 1e6, "fixed_rate": 0.02, "payment_times": pay, "accruals": acc,
            "monitoring_times": times, "n_steps": 16, "z": z.tolist(), "times": list(t), "ee": list(ee),
            "pfe95": list(pfe)}


def philox_kat():
    kat = [  # (counter, key, expected), Random123 kat_vectors for philox4x32-10
        ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
        ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
        ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
         [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
    ]
    for c, k, e in kat:
        assert list(oracle.philox4x32_10(c, k)) == e
    z = oracle.philox_normals(7, 4, 4)
    return {"kat": [{"counter": c, "key": k, "out": e} for c, k, e in kat],
            "normals_seed7_4paths_4steps": z.tolist()}


def reference_values():
    return {
        "config1_pv": {"value": 948.4537547220389, "source": "PLAN_API_REFACTOR.md:330"},
        "config1_dv01": {"value": 480.18800212936185, "source": "PLAN_API_REFACTOR.md:331"},
        "par_swap_dv01": {"value": 480.4686940754473, "source": "clients/excel/README.md:324"},
        "par_swap_seed7_ee": {"value": [0.0, 12862.617942080262, 13673.529752568928, 11957.816098484913,
                                        7124.106240024746], "source": "clients/excel/README.md:321-335",
                              "note": "burn-ndarray draws: a statistical anchor only"},
        "par_swap_seed7_pfe95": {"value": [0.0, 51009.920875152675, 53607.17081592724, 46152.44561612198,
                                           27535.557267142714], "source": "clients/excel/README.md:321-335",
                                 "note": "burn-ndarray draws: a statistical anchor only"},
        "par_swap_seed7_cva": {"value": 503.64194077997536, "source": "clients/excel/README.md:321-335"},
        "config1_cva": {"value": 626.7254432766481, "source": "PLAN_API_REFACTOR.md:331",
                        "note": "burn-ndarray draws: a statistical anchor only"},
        "handmade_cva_profiles": {"times": [0.0, 1.0, 2.0], "ee_a": [0.0, 12000.0, 9500.0],
                                  "ee_b": [0.0, 24000.0, 19000.0], "hazard_rate": 0.02, "recovery_rate": 0.4,
                                  "source": "rust/crates/engine-core/src/exposure.rs:729-759"},
    }


def main():
    files = {"netting_tiny.json": netting_tiny(), "reference_profile_tiny.json": reference_profile_tiny(),
             "philox_kat.json": philox_kat(), "reference_values.json": reference_values()}
    for name, data in files.items():
        with open(os.path.join(HERE, name), "w") as f:
            json.dump(data, f, indent=1)
        print("wrote", name)


if __name__ == "__main__":
    main()
