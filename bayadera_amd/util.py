"""Highest-density intervals from a Histogram's bin ranks.

Mirrors ``src/clojure/uncomplicate/bayadera/util.clj:52-110`` (host-side
post-processing of the EstimateEngine's ``bin-ranks``)."""
from __future__ import annotations

import numpy as np


def hdi_rank_count(bin_rank, pdf, mass: float = 0.95) -> int:
    """util.clj:52-64: the smallest number of top-ranked bins whose pdf sum
    reaches mass * sum(|pdf|)."""
    ranks = np.asarray(bin_rank).astype(np.int64)
    p = np.asarray(pdf, np.float64)
    density = mass * float(np.abs(p).sum())
    acc, i, n = 0.0, 0, ranks.size
    while i < n and acc < density:
        acc += float(p[ranks[i]])
        i += 1
    return i


def hdi_bins(bin_rank, hdi_cnt: int) -> list[float]:
    """util.clj:66-83: the first hdi_cnt ranked bins, sorted, grouped into
    runs; returns [start0, end0, start1, end1, ...]."""
    v = np.sort(np.asarray(bin_rank, np.float64)[:hdi_cnt])
    regions = [float(v[0])]
    last = v[0]
    for b in v[1:]:
        if b - last > 1.5:
            regions += [float(last), float(b)]
        last = b
    regions.append(float(v[hdi_cnt - 1]))
    return regions


def hdi_regions(limits, bin_rank, hdi_cnt: int) -> np.ndarray:
    """util.clj:85-101: the regions of hdi_bins mapped into [lower, upper];
    shape (2, number of regions)."""
    lower, upper = float(limits[0]), float(limits[1])
    width = (upper - lower) / len(bin_rank)
    bins = hdi_bins(bin_rank, hdi_cnt)
    cnt = len(bins) // 2
    out = np.zeros((2, cnt))
    for i in range(cnt):
        out[0, i] = lower + width * bins[2 * i]
        out[1, i] = lower + width * (bins[2 * i + 1] + 1.0)
    return out


def hdi(histogram, index: int, mass: float = 0.95) -> np.ndarray:
    """util.clj:103-110: hdi regions holding `mass` for variable `index`."""
    limits = histogram.limits[index]
    ranks = histogram.bin_ranks[index]
    pdf = histogram.pdf[index]
    return hdi_regions(limits, ranks, hdi_rank_count(ranks, pdf, mass))
