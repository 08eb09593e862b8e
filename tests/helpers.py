"""Shared checks for the parity tests."""
import numpy as np


def pixel_radiance(sums, counts):
    c = np.maximum(counts, 1).astype(np.float64)[:, None]
    return sums / c


def aggregate_relative_difference(a_sums, a_counts, b_sums, b_counts):
    """Aggregate L1 ratio over rgb: sum |a - b| / sum |b| over the whole image."""
    a = pixel_radiance(a_sums, a_counts)[:, :3]
    b = pixel_radiance(b_sums, b_counts)[:, :3]
    den = np.abs(b).sum()
    return float(np.abs(a - b).sum() / den) if den > 0 else float(np.abs(a - b).sum())


def mean_relative_difference(a_sums, a_counts, b_sums, b_counts):
    """Per-pixel mean relative radiance difference (the stated tolerance is 1e-3): the mean over pixels of
    sum_c |a_c - b_c| / (sum_c |b_c| + eps), eps = 1e-3 of the image's mean sum_c |b_c| (so that a pixel black in b
    counts its absolute difference against a thousandth of the image's level, not a division by zero). Returns the
    larger of that and the aggregate ratio, so either reading meets a tolerance asserted on it."""
    a = pixel_radiance(a_sums, a_counts)[:, :3]
    b = pixel_radiance(b_sums, b_counts)[:, :3]
    level = np.abs(b).sum(axis=1)
    eps = 1e-3 * float(level.mean()) if level.size and level.mean() > 0 else 1e-30
    per_pixel = float((np.abs(a - b).sum(axis=1) / (level + eps)).mean()) if level.size else 0.0
    return max(per_pixel, aggregate_relative_difference(a_sums, a_counts, b_sums, b_counts))


def drain(lane):
    """Polls until every launch in flight has finished; returns the completions."""
    out = []
    while lane.in_flight():
        c = lane.poll()
        if c is not None:
            out.append(c)
    return out
