"""Shared checks for the parity tests."""
import numpy as np


def pixel_radiance(sums, counts):
    c = np.maximum(counts, 1).astype(np.float64)[:, None]
    return sums / c


def mean_relative_difference(a_sums, a_counts, b_sums, b_counts):
    """Mean relative per-pixel radiance difference over rgb: sum |a - b| / sum |b| (the stated tolerance 1e-3)."""
    a = pixel_radiance(a_sums, a_counts)[:, :3]
    b = pixel_radiance(b_sums, b_counts)[:, :3]
    den = np.abs(b).sum()
    return float(np.abs(a - b).sum() / den) if den > 0 else float(np.abs(a - b).sum())


def drain(lane):
    """Polls until every launch in flight has finished; returns the completions."""
    out = []
    while lane.in_flight():
        c = lane.poll()
        if c is not None:
            out.append(c)
    return out
