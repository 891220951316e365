"""Comparison metrics of receiver IRs between precisions."""
from __future__ import annotations

import numpy as np


def rel_l2(a, b) -> float:
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def decay_curve_db(ir) -> np.ndarray:
    """Schroeder energy-decay curve of each IR row in dB re its total energy:
    EDC(t) = 10 log10(sum_{tau >= t} p(tau)^2 / sum p^2)."""
    e = np.asarray(ir, float) ** 2
    tail = np.cumsum(e[..., ::-1], axis=-1)[..., ::-1]
    tot = tail[..., :1]
    return 10.0 * np.log10(np.maximum(tail, 1e-300) / np.maximum(tot, 1e-300))


def decay_error_db(ir, ref, floor_db: float = -40.0) -> float:
    """Largest |EDC(ir) - EDC(ref)| in dB over the samples where the reference curve is
    above `floor_db` (the late tail down to that level)."""
    a, b = decay_curve_db(ir), decay_curve_db(ref)
    m = b > floor_db
    return float(np.abs(a - b)[m].max()) if m.any() else 0.0
