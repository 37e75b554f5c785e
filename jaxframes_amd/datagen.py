"""Seeded synthetic inputs for the benchmark workloads, generated on the GPU.

Shapes follow the workload definitions:
- groupby: int32 keys uniform in [0, card), float32 N(0, 1) values;
- sort: uint64 keys uniform over 64 bits, or Zipf(s=1.1) over 10^6 distinct
  uint64 values, with an int32 row-index payload;
- join: right = 5*10^6 distinct int64 keys each present twice (shuffled);
  left = 80% keys drawn by Zipf(s=1.2) rank over a random permutation of the
  right keys + 20% keys from a disjoint set, so M = 2 * 0.8 * L pairs
  (Zipf duplicates on both sides would give ~10^13 pairs); 2 float32 payload
  columns per side;
- partition: int64 keys uniform over 64 bits;
- batched groupby: frames with log-uniform row counts and cardinalities.
"""

from __future__ import annotations

import math

import torch


def _gen(seed: int, device) -> torch.Generator:
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    return g


def _u64(n: int, g: torch.Generator, device) -> torch.Tensor:
    hi = torch.randint(0, 2**32, (n,), dtype=torch.int64, device=device, generator=g)
    lo = torch.randint(0, 2**32, (n,), dtype=torch.int64, device=device, generator=g)
    return ((hi << 32) | lo).view(torch.uint64)


def zipf_ranks(n: int, nvals: int, s: float, g: torch.Generator, device) -> torch.Tensor:
    """n ranks in [0, nvals) with P(k) proportional to (k+1)^-s."""
    w = torch.arange(1, nvals + 1, dtype=torch.float64, device=device).pow(-s)
    cdf = torch.cumsum(w, 0)
    cdf /= cdf[-1].clone()
    u = torch.rand(n, dtype=torch.float64, device=device, generator=g)
    return torch.searchsorted(cdf, u).clamp_(max=nvals - 1)


def groupby_frame(n: int, card: int = 1000, seed: int = 1234, device="cuda"):
    g = _gen(seed, device)
    keys = torch.randint(0, card, (n,), dtype=torch.int32, device=device, generator=g)
    vals = torch.randn(n, dtype=torch.float32, device=device, generator=g)
    return keys, vals


def sort_keys(n: int, kind: str = "uniform", seed: int = 1234, device="cuda"):
    g = _gen(seed, device)
    if kind == "uniform":
        return _u64(n, g, device)
    if kind == "zipf":
        distinct = _u64(10**6, g, device).view(torch.int64)
        return distinct[zipf_ranks(n, 10**6, 1.1, g, device)].view(torch.uint64)
    raise ValueError(kind)


def join_tables(n_left: int = 10**8, n_right: int = 10**7, seed: int = 1234, device="cuda"):
    """Returns (left_keys, {'a','b'}, right_keys, {'c','d'}, expected_pairs)."""
    g = _gen(seed, device)
    nd = n_right // 2
    # distinct positive right keys: a random permutation of a strided range
    base = torch.randint(1, 2**40, (1,), dtype=torch.int64, device=device, generator=g)
    stride = 2**20 + 7
    distinct = base + torch.randperm(nd, device=device, generator=g).to(torch.int64) * stride
    rk = torch.cat([distinct, distinct])
    rk = rk[torch.randperm(rk.numel(), device=device, generator=g)]
    n_match = int(round(0.8 * n_left))
    order = torch.randperm(nd, device=device, generator=g)
    lm = distinct[order[zipf_ranks(n_match, nd, 1.2, g, device)]]
    lx = -torch.randint(1, 2**62, (n_left - n_match,), dtype=torch.int64, device=device, generator=g)
    lk = torch.cat([lm, lx])
    lk = lk[torch.randperm(n_left, device=device, generator=g)]
    lv = {"a": torch.randn(n_left, device=device, generator=g),
          "b": torch.randn(n_left, device=device, generator=g)}
    rv = {"c": torch.randn(n_right, device=device, generator=g),
          "d": torch.randn(n_right, device=device, generator=g)}
    return lk, lv, rk, rv, 2 * n_match


def partition_keys(n: int, seed: int = 1234, device="cuda"):
    g = _gen(seed, device)
    return _u64(n, g, device).view(torch.int64)


def batched_frames(nframes: int = 4096, min_rows: int = 1000, max_rows: int = 4_000_000,
                   min_card: int = 10, max_card: int = 10**6, seed: int = 1234,
                   total_rows: int | None = None, device="cuda"):
    """Concatenated frames: (keys int32, vals float32, offsets int64[nframes+1],
    cards). Row counts and cardinalities are log-uniform; total_rows rescales
    the row counts (floor at min_rows)."""
    g = torch.Generator()
    g.manual_seed(seed)
    u = torch.rand(nframes, generator=g, dtype=torch.float64)
    rows = torch.floor(torch.exp(math.log(min_rows) + u * (math.log(max_rows) - math.log(min_rows))))
    if total_rows is not None:
        rows = torch.clamp(torch.floor(rows * (total_rows / float(rows.sum()))), min=min_rows)
    rows = rows.to(torch.int64)
    uc = torch.rand(nframes, generator=g, dtype=torch.float64)
    cards = torch.floor(torch.exp(math.log(min_card) + uc * (math.log(max_card) - math.log(min_card))))
    cards = cards.to(torch.int64)
    offsets = torch.zeros(nframes + 1, dtype=torch.int64)
    offsets[1:] = torch.cumsum(rows, 0)
    n = int(offsets[-1])
    gd = _gen(seed + 1, device)
    per_row_card = torch.repeat_interleave(cards.to(device), rows.to(device))
    keys = (torch.rand(n, dtype=torch.float64, device=device, generator=gd) * per_row_card).to(torch.int32)
    vals = torch.randn(n, dtype=torch.float32, device=device, generator=gd)
    return keys, vals, offsets, cards


def served_frames(nframes: int = 8, min_rows: int = 10**6, max_rows: int = 5 * 10**7,
                  min_card: int = 10, max_card: int = 10**6, seed: int = 1234, device="cuda"):
    """Resident frames of the served stream: [(name, int64 keys, float32
    values)], rows and key cardinality log-uniform."""
    g = torch.Generator()
    g.manual_seed(seed)
    u = torch.rand(nframes, generator=g, dtype=torch.float64)
    rows = torch.floor(torch.exp(math.log(min_rows) + u * (math.log(max_rows) - math.log(min_rows))))
    uc = torch.rand(nframes, generator=g, dtype=torch.float64)
    cards = torch.floor(torch.exp(math.log(min_card) + uc * (math.log(max_card) - math.log(min_card))))
    gd = _gen(seed + 1, device)
    out = []
    for i in range(nframes):
        n, c = int(rows[i]), int(cards[i])
        k = torch.randint(0, c, (n,), dtype=torch.int64, device=device, generator=gd)
        v = torch.randn(n, dtype=torch.float32, device=device, generator=gd)
        out.append((f"f{i}", k, v, c))
    return out


def served_queries(frames, nq: int, join_frac: float = 0.3, min_dim: int = 10**5,
                   max_dim: int = 10**6, seed: int = 1234, device="cuda"):
    """The served query mix over frames from served_frames: filter(v > t,
    t ~ U(-1, 1)) -> groupby(k).agg(sum, count, mean), and (join_frac of
    them) inner joins with a dimension table of log-uniform size whose
    distinct int64 keys cover part of the frame's key range."""
    from .serving import FilterGroupBy, Join
    g = torch.Generator()
    g.manual_seed(seed)
    gd = _gen(seed + 2, device)
    qs = []
    for _ in range(nq):
        f = int(torch.randint(0, len(frames), (1,), generator=g))
        name, _, _, card = frames[f]
        if float(torch.rand(1, generator=g, dtype=torch.float64)) >= join_frac:
            t = float(torch.rand(1, generator=g, dtype=torch.float64)) * 2 - 1
            qs.append(FilterGroupBy(name, ">", t))
        else:
            u = float(torch.rand(1, generator=g, dtype=torch.float64))
            nd = int(math.floor(math.exp(math.log(min_dim) + u * (math.log(max_dim) - math.log(min_dim)))))
            dk = torch.randperm(max(card, nd), device=device, generator=gd)[:nd].to(torch.int64)
            qs.append(Join(name, dk, torch.randn(nd, dtype=torch.float32, device=device, generator=gd)))
    return qs
