"""Exposure-cube wire format and the XVA reducer on the GPU.

Mirrors the REST side of engine_quant: the exposure cube contract ``ExposureCubeInput``
(rust/crates/quant-domain/src/xva.rs:176-220), the request's numeric parts (``DefaultSpec``,
``FundingSpec``, ``CollateralAgreement``, xva.rs:55-141) and ``xva::calculate``
(rust/crates/quant-engine/src/xva.rs:97-246), whose per-row reductions run in libeqhip
(``eqhip_xva_calculate``). A Monte Carlo plan created with ``keep_cube=True`` hands its netting-set
cube to the reducer without leaving the device (``plan_xva``), in place of the flat one-path PV cube
of quant-engine/src/lib.rs:751-779.

Wire forms of a cube:
- JSON with the serde field names of ``ExposureCubeInput`` (``measure`` "Q"/"P", ``times``, ``paths``
  as rows [time][path], ``discount_factors``, ``source_hash``), for the REST body;
- a binary form for cubes of production size (100k paths x 360 dates is 288 MB of f64; the JSON
  text of the same cube is about three times that): little-endian, header
  ``b"EQXCUBE1"``, u8 measure (0 Q, 1 P), u8 has discount factors, u16 0, u32 hash length, u64
  n_times, u64 n_paths, the hash text padded to 8 bytes, then f64 times, f64 discount factors (if
  any) and the f64 rows [time][path].
"""
from __future__ import annotations

import ctypes
import hashlib
import json
import struct
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

from . import eqhip

_MAGIC = b"EQXCUBE1"
_HEADER = struct.Struct("<8sBBHIQQ")


@dataclass
class DefaultSpec:
    """quant-domain/src/xva.rs:97-121."""
    hazard_rate: float = 0.0
    recovery_rate: float = 0.4
    survival_probabilities: Sequence[float] = ()
    measure: str = "Q"


@dataclass
class FundingSpec:
    """quant-domain/src/xva.rs:124-141."""
    funding_spread: float = 0.0
    borrowing_spread: float = 0.0
    lending_spread: float = 0.0
    capital_factor: float = 0.08
    margin_period_years: float = 1.0 / 252.0


@dataclass
class CollateralAgreement:
    """quant-domain/src/xva.rs:55-94."""
    id: str
    currency: str = "USD"
    threshold: float = 0.0
    minimum_transfer_amount: float = 0.0
    independent_amount: float = 0.0
    initial_margin: float = 0.0
    variation_margin: bool = True
    collateral_rate: float = 0.0


@dataclass
class XvaRequest:
    """The parts of quant-domain's XvaRequest (xva.rs:234-262) that the reduction reads."""
    operation_id: str = ""
    context_hash: Optional[str] = None
    netting_sets: Sequence[str] = ()
    collateral_agreements: Sequence[CollateralAgreement] = ()
    counterparty_default: DefaultSpec = field(default_factory=DefaultSpec)
    own_default: DefaultSpec = field(default_factory=DefaultSpec)
    funding: FundingSpec = field(default_factory=FundingSpec)
    confidence_level: float = 0.95
    materialize_exposure: bool = False


@dataclass
class XvaResult:
    """quant-domain/src/xva.rs:282-299."""
    operation_id: str
    context_hash: Optional[str]
    times: np.ndarray
    ee: np.ndarray
    pfe: np.ndarray
    cva: float
    dva: float
    fva: float
    mva: float
    kva: float
    measures: dict
    exposure_cube: dict
    artifacts: list
    provenance: dict


class ExposureCube:
    """ExposureCubeInput (quant-domain/src/xva.rs:176-220): rows [time][path] under one measure."""

    def __init__(self, measure, times, paths, discount_factors=(), source_hash=None):
        self.measure = measure
        self.times = np.asarray(times, dtype=np.float64)
        self.paths = np.asarray(paths, dtype=np.float64)
        self.discount_factors = np.asarray(discount_factors, dtype=np.float64)
        self.source_hash = source_hash

    def validate(self):
        """Raises ValueError with the reference's messages (xva.rs:188-220)."""
        t, p = self.times, self.paths
        if len(t) == 0 or p.ndim != 2 or p.shape[0] != len(t):
            raise ValueError("exposure times and rows must be non-empty and have equal length")
        if not np.all(np.isfinite(t[1:])) or np.any(np.diff(t) <= 0) or t[0] < 0.0:
            raise ValueError("exposure times must be finite and strictly increasing")
        if p.shape[1] == 0 or not np.all(np.isfinite(p)):
            raise ValueError("exposure paths must be rectangular finite rows")
        d = self.discount_factors
        if len(d) and (len(d) != len(t) or not np.all(np.isfinite(d)) or np.any(d < 0)):
            raise ValueError("discount_factors must match times and be non-negative")

    def estimate_memory_bytes(self) -> int:
        """quant-engine/src/xva.rs:248-253."""
        return int(self.paths.size * 8)

    # JSON, with the serde field names of ExposureCubeInput
    def to_json(self) -> str:
        """The cube as the reference serialises it for its hash (quant-engine/src/xva.rs:77-91:
        serde_json::to_value, whose map keeps keys sorted, then compact serde_json::to_vec)."""
        return canonical_json(self).decode()

    @classmethod
    def from_json(cls, text: str) -> "ExposureCube":
        d = json.loads(text)
        if d.get("measure", "Q") not in ("Q", "P"):
            raise ValueError(f"unknown probability measure {d.get('measure')!r}")
        return cls(d.get("measure", "Q"), d["times"], d["paths"], d.get("discount_factors", []),
                   d.get("source_hash"))

    # binary form
    def to_bytes(self) -> bytes:
        h = (self.source_hash or "").encode()
        pad = (-len(h)) % 8
        nt, npath = self.paths.shape
        head = _HEADER.pack(_MAGIC, 0 if self.measure == "Q" else 1, 1 if len(self.discount_factors) else 0, 0,
                            len(h), nt, npath)
        parts = [head, h, b"\0" * pad, self.times.astype("<f8").tobytes()]
        if len(self.discount_factors):
            parts.append(self.discount_factors.astype("<f8").tobytes())
        parts.append(np.ascontiguousarray(self.paths, dtype="<f8").tobytes())
        return b"".join(parts)

    @classmethod
    def from_bytes(cls, buf: bytes) -> "ExposureCube":
        magic, meas, has_df, _, hlen, nt, npath = _HEADER.unpack_from(buf, 0)
        if magic != _MAGIC:
            raise ValueError("not an exposure cube (bad magic)")
        off = _HEADER.size
        h = buf[off:off + hlen].decode() or None
        off += hlen + (-hlen) % 8
        times = np.frombuffer(buf, "<f8", nt, off)
        off += 8 * nt
        df = np.zeros(0)
        if has_df:
            df = np.frombuffer(buf, "<f8", nt, off)
            off += 8 * nt
        if len(buf) != off + 8 * nt * npath:
            raise ValueError("exposure cube: truncated or oversized payload")
        paths = np.frombuffer(buf, "<f8", nt * npath, off).reshape(nt, npath)
        return cls("Q" if meas == 0 else "P", times.copy(), paths.copy(), df.copy(), h)


def _params(request: XvaRequest) -> eqhip.XvaParamsC:
    keep = []

    def spec(d: DefaultSpec) -> eqhip.DefaultSpecC:
        sp = np.ascontiguousarray(np.asarray(d.survival_probabilities, dtype=np.float64))
        keep.append(sp)
        ptr = sp.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) if len(sp) else None
        return eqhip.DefaultSpecC(d.hazard_rate, d.recovery_rate, ptr, len(sp))

    csas = list(request.collateral_agreements)
    csa = None
    if csas:
        c = csas[0]
        csa = eqhip.CsaC(c.threshold, c.minimum_transfer_amount, c.independent_amount, int(bool(c.variation_margin)))
        keep.append(csa)
    im = 0.0
    for c in csas:  # xva.rs:160-166 sums max(initial_margin, 0) in CSA order
        im += max(c.initial_margin, 0.0)
    f = request.funding
    p = eqhip.XvaParamsC(spec(request.counterparty_default), spec(request.own_default), f.funding_spread,
                         f.borrowing_spread, f.capital_factor, im, ctypes.pointer(csa) if csa is not None else None,
                         request.confidence_level)
    p._keep = keep
    return p


def _check_inputs(request: XvaRequest, q: ExposureCube, p: Optional[ExposureCube]):
    """The request-level checks of xva.rs:102-130."""
    q.validate()
    if q.measure != "Q":
        raise ValueError("q_exposure must be tagged Q")
    if q.source_hash is not None and request.context_hash is not None and q.source_hash != request.context_hash:
        raise ValueError("Q exposure source hash does not match context hash")
    if p is not None:
        p.validate()
        if p.measure != "P":
            raise ValueError("p_exposure must be tagged P")
        if not np.array_equal(p.times, q.times):
            raise ValueError("Q and P exposure grids must match")
        if p.source_hash is not None and request.context_hash is not None and p.source_hash != request.context_hash:
            raise ValueError("P exposure source hash does not match context hash")
    if not (0.0 <= request.confidence_level <= 1.0):
        raise ValueError("confidence_level must be in [0,1]")


def json_float(v: float) -> str:
    """A finite f64 as serde_json writes it: the shortest digits that read back to the same double, laid
    out as 1234e7 -> "12340000000.0", 1234e-2 -> "12.34", 1234e-6 -> "0.001234" while 1e-5 <= |v| < 1e16,
    otherwise "1e30" / "1.234e33" / "1e-7" (no "+", no leading zeros in the exponent). Python's repr has
    the same shortest digits and the same decimal layout from 1e-4 up to 1e16; below and above, it writes
    an exponent, which is re-laid out here."""
    r = repr(float(v))
    if "n" in r:
        raise ValueError("exposure values must be finite")
    if "e" not in r:
        return r
    mant, exp = r.split("e")
    sign = "-" if mant.startswith("-") else ""
    m = mant.lstrip("-")
    ip, fp = m.split(".") if "." in m else (m, "")
    digits, e10 = ip + fp, int(exp) - len(fp)  # |v| = digits x 10^e10
    stripped = digits.rstrip("0")
    e10 += len(digits) - len(stripped)
    digits = stripped.lstrip("0")
    length = len(digits)
    kk = length + e10  # 10^(kk-1) <= |v| < 10^kk
    if 0 <= e10 and kk <= 16:
        return sign + digits + "0" * e10 + ".0"
    if 0 < kk <= 16:
        return sign + digits[:kk] + "." + digits[kk:]
    if -5 < kk <= 0:
        return sign + "0." + "0" * (-kk) + digits
    tail = "" if length == 1 else "." + digits[1:]
    return sign + digits[0] + tail + "e" + str(kk - 1)


def _json_list(a) -> str:
    return "[" + ",".join(json_float(x) for x in np.asarray(a, dtype=np.float64).ravel().tolist()) + "]"


def canonical_json(cube: "ExposureCube") -> bytes:
    """The exact bytes the reference hashes for an exposure cube: keys in sorted order, no spaces, f64 as
    serde_json formats them, source_hash null or its string."""
    sh = "null" if cube.source_hash is None else json.dumps(cube.source_hash)
    rows = ",".join(_json_list(row) for row in cube.paths)
    return (f'{{"discount_factors":{_json_list(cube.discount_factors)},"measure":{json.dumps(cube.measure)},'
            f'"paths":[{rows}],"source_hash":{sh},"times":{_json_list(cube.times)}}}').encode()


def cube_hash(cube: "ExposureCube") -> str:
    """hash_value of the cube (quant-domain/src/lib.rs:308-313): "sha256:" + hex digest of its JSON."""
    return "sha256:" + hashlib.sha256(canonical_json(cube)).hexdigest()


def _result(request: XvaRequest, q: ExposureCube, p: Optional[ExposureCube], ee, pfe, m) -> XvaResult:
    """The result record of xva.rs:184-245; the artifact id is the reference's: the SHA-256 of the Q cube's
    JSON (artifact, xva.rs:77-91)."""
    cva, dva, fva, mva, kva = (float(v) for v in m)
    approximations = ["piecewise-constant hazard and simple discounted exposure integration",
                      "all requested measures reuse one Q exposure reduction graph"]
    if p is None:
        approximations.append("DVA falls back to Q exposure because no P cube was supplied")
    if request.collateral_agreements:
        approximations.append("collateral is represented as thresholded variation margin from the first CSA")
    lineage = [request.context_hash or "context:unhashed", f"operation:{request.operation_id}"]
    lineage += [f"netting-set:{s}" for s in request.netting_sets]
    lineage += [f"csa:{c.id}" for c in request.collateral_agreements]
    digest = cube_hash(q)[len("sha256:"):]
    cube_ref = {"artifact_id": f"exposure-{digest}", "hash": f"sha256:{digest}", "measure": q.measure,
                "shape": [len(q.times), q.paths.shape[1]], "materialized": bool(request.materialize_exposure),
                "bytes": q.estimate_memory_bytes(), "lineage": list(lineage)}
    provenance = {"context_hash": request.context_hash, "operation_id": request.operation_id, "q_measure": True,
                  "p_measure": p is not None, "shared_exposure_graph": True, "approximations": approximations,
                  "artifact_lineage": lineage + [cube_ref["artifact_id"], cube_ref["hash"]]}
    return XvaResult(request.operation_id, request.context_hash, q.times.copy(), ee, pfe, cva, dva, fva, mva, kva,
                     {"cva": cva, "dva": dva, "fva": fva, "mva": mva, "kva": kva}, cube_ref, [cube_ref], provenance)


def calculate(ctx: "eqhip.Context", request: XvaRequest, q: ExposureCube,
              p: Optional[ExposureCube] = None) -> XvaResult:
    """xva::calculate (quant-engine/src/xva.rs:97-246) with the row reductions on the GPU."""
    _check_inputs(request, q, p)
    df = q.discount_factors if len(q.discount_factors) else None
    ee, pfe, m = ctx.xva_calculate(_params(request), q.times, q.paths, df, None if p is None else p.paths)
    return _result(request, q, p, ee, pfe, m)


def plan_cube(plan: "eqhip.Plan", set_index: int = 0, source_hash: Optional[str] = None) -> ExposureCube:
    """The Q cube of one netting set of a plan's last execute: signed values and model discount factors."""
    values, df = plan.exposure_cube(set_index)
    return ExposureCube("Q", plan.dates.copy(), values, df, source_hash)


def plan_xva(plan: "eqhip.Plan", request: XvaRequest, set_index: int = 0) -> XvaResult:
    """xva::calculate on a plan's resident cube (no device-to-host copy of the cube); DVA from the
    same Q cube, discount factors P(0, t_k) of the plan's model."""
    if not (0.0 <= request.confidence_level <= 1.0):
        raise ValueError("confidence_level must be in [0,1]")
    ee, pfe, m = plan.xva(_params(request), set_index)
    shape_cube = ExposureCube("Q", plan.dates.copy(), np.zeros((len(plan.dates), 1)))
    res = _result(request, shape_cube, None, ee, pfe, m)
    res.exposure_cube.update(shape=[len(plan.dates), plan.n_paths], bytes=8 * len(plan.dates) * plan.n_paths,
                             artifact_id=None, hash=None, materialized=False)
    res.artifacts = [res.exposure_cube]
    res.provenance["artifact_lineage"] = res.provenance["artifact_lineage"][:-2]
    return res
