"""Comparison helpers shared by the parity tests."""

from __future__ import annotations

import numpy as np

from diomedes_amd import records as R


def contacts_to_records(c: dict) -> np.ndarray:
    """Oracle contacts (dict of double arrays) -> GpuAvbdContact records."""
    m = len(c["a"])
    out = np.zeros(m, R.AVBD_CONTACT)
    out["a"], out["b"] = c["a"], c["b"]
    out["friction"] = np.asarray(c["friction"], np.float32)
    for k_src, k_dst in (("normal", "normal"), ("tangent1", "tangent1"), ("tangent2", "tangent2"), ("r_a", "r_a"),
                         ("r_b", "r_b"), ("c0", "c0"), ("penalty", "penalty"), ("lambda", "lambda_")):
        out[k_dst][:, :3] = np.asarray(c[k_src], np.float32)
    return out


def records_to_contacts(rec: np.ndarray) -> dict:
    out = {"a": rec["a"].copy(), "b": rec["b"].copy(), "friction": rec["friction"].astype(np.float64)}
    for k_src, k_dst in (("normal", "normal"), ("tangent1", "tangent1"), ("tangent2", "tangent2"), ("r_a", "r_a"),
                         ("r_b", "r_b"), ("c0", "c0"), ("penalty", "penalty"), ("lambda_", "lambda")):
        out[k_dst] = rec[k_src][:, :3].astype(np.float64)
    return out


def bits_equal(gpu, ref) -> bool:
    """Bit-exact float32 equality (NaNs compare equal to NaNs)."""
    g = np.asarray(gpu, np.float32)
    r = np.asarray(ref, np.float64).astype(np.float32)
    if g.shape != r.shape:
        return False
    return bool(np.all((g.view(np.uint32) == r.view(np.uint32)) | (np.isnan(g) & np.isnan(r))))


def assert_bits(gpu, ref, what: str) -> None:
    g = np.asarray(gpu, np.float32)
    r = np.asarray(ref, np.float64).astype(np.float32)
    assert g.shape == r.shape, f"{what}: shape {g.shape} vs {r.shape}"
    if not bits_equal(g, r):
        diff = np.abs(g.astype(np.float64) - r.astype(np.float64))
        bad = np.argwhere(~((g.view(np.uint32) == r.view(np.uint32)) | (np.isnan(g) & np.isnan(r))))
        raise AssertionError(
            f"{what}: {len(bad)} of {g.size} values differ; max |diff| {np.nanmax(diff):.3e}; first at {bad[0].tolist()}"
            f" gpu={g[tuple(bad[0])]!r} ref={r[tuple(bad[0])]!r}"
        )


def rel_err(a, b) -> float:
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    scale = max(1.0, float(np.max(np.abs(b))) if b.size else 1.0)
    return float(np.max(np.abs(a - b)) / scale) if a.size else 0.0


def primal_tuning(lanes) -> dict:
    """set_tuning arguments for a primal shape of the world step: "packed"
    (the packed primal), "span" (the span primal) or that many lanes per body."""
    if lanes == "packed":
        return {"primal_packed": 1}
    if lanes == "span":
        return {"primal_packed": 2}
    return {"primal_lanes": lanes, "primal_packed": 0}
