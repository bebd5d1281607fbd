"""Seeded synthetic rating sets with the shape of the Netflix data.

The real archive cannot be fetched offline, so training runs on data drawn
from a planted model:

  * users' rating counts are lognormal (given median, clipped to a range) and
    rescaled to an exact total;
  * item popularity is lognormal; a user's items are distinct and drawn in
    proportion to popularity;
  * days lie in 0..2242 (uint16 range of the reference's i16 dates), in
    rating sessions inside a per-user active window, sorted per user;
  * a score = mu + b_u + b_i + <p_u, q_i> (rank 16) + a_u * (t - mean_t_u)
    + N(0, noise) is quantised to 1..5 at the quantiles that give the
    marginal (5, 10, 29, 33, 23) %;
  * the last `holdout` ratings of each user by date (at most cnt - 1) form the
    held-out set.

Output follows the reference's on-disk layout (src/bin/ingest.rs:439-477,
read by Dataset::load, src/lib.rs:252-290): user_idxs <i4, user_cnts <i4,
item_idxs <i2, item_cnts <i4, item_years <i4, ratings <i1, dates <i2,
is_test <i1; training rows sorted by (user, date, item).
"""
from __future__ import annotations

import os
from dataclasses import dataclass

import numpy as np

from .tx import Dataset

N_DAYS = 2243
MARGINAL = (0.05, 0.10, 0.29, 0.33, 0.23)


@dataclass(frozen=True)
class SynthSpec:
    n_users: int
    n_items: int
    n_ratings: int
    user_median: float
    user_max: int
    item_sigma: float = 2.0
    item_min: int = 3
    item_max: int = 230_000
    rank: int = 16
    noise: float = 0.9
    holdout: int = 3


# The two dataset shapes of the benchmark configurations.
SMALL = SynthSpec(n_users=10_000, n_items=2_000, n_ratings=1_000_000, user_median=60, user_max=2_000,
                  item_sigma=1.2, item_min=3, item_max=60_000)
FULL = SynthSpec(n_users=480_189, n_items=17_770, n_ratings=100_480_507, user_median=96, user_max=17_000)


def _exact_total(counts: np.ndarray, total: int, lo: int, hi: int) -> np.ndarray:
    """Scale integer counts in [lo, hi] so that they sum to exactly `total`."""
    c = counts.astype(np.float64)
    for _ in range(60):
        s = np.clip(np.round(c), lo, hi).sum()
        if s == total:
            break
        c *= total / s
    out = np.clip(np.round(c), lo, hi).astype(np.int64)
    diff = int(total - out.sum())
    order = np.argsort(-out, kind="stable")
    k = 0
    while diff != 0:
        j = order[k % len(order)]
        step = 1 if diff > 0 else -1
        if lo <= out[j] + step <= hi:
            out[j] += step
            diff -= step
        k += 1
    return out


def _item_weights(spec: SynthSpec, rng: np.random.Generator) -> np.ndarray:
    w = rng.lognormal(0.0, spec.item_sigma, spec.n_items)
    target = _exact_total(w / w.sum() * spec.n_ratings, spec.n_ratings, spec.item_min, spec.item_max)
    return target.astype(np.float64) / target.sum()


def _user_ranges(users_sel: np.ndarray, starts: np.ndarray, counts: np.ndarray) -> np.ndarray:
    """Flat indices of the rating ranges of the selected users."""
    lens = counts[users_sel]
    base = np.repeat(starts[users_sel] - np.concatenate([[0], np.cumsum(lens)[:-1]]), lens)
    return base + np.arange(int(lens.sum()), dtype=np.int64)


def _draw_items(counts: np.ndarray, p: np.ndarray, rng: np.random.Generator) -> np.ndarray:
    """Per user `counts[u]` distinct items drawn ~ p, concatenated by user."""
    n_items = len(p)
    users = np.repeat(np.arange(len(counts), dtype=np.int64), counts)
    items = np.empty(len(users), np.int64)
    starts = np.concatenate([[0], np.cumsum(counts)])
    heavy = counts > n_items // 8
    cdf = np.cumsum(p)
    cdf /= cdf[-1]
    idx = np.nonzero(~heavy[users])[0]
    items[idx] = np.minimum(np.searchsorted(cdf, rng.random(len(idx))), n_items - 1)
    # heavy users: weighted sampling without replacement (Gumbel top-k)
    logp = np.log(p)
    for u in np.nonzero(heavy)[0]:
        keys = logp - np.log(-np.log(rng.random(n_items)))
        k = int(counts[u])
        items[starts[u]:starts[u + 1]] = np.argpartition(-keys, k - 1)[:k]
    # redraw duplicates, re-checking only the users that still have some
    sel = np.arange(len(counts), dtype=np.int64)[~heavy]
    for it in range(200):
        rows = _user_ranges(sel, starts, counts)
        key = users[rows] * n_items + items[rows]
        order = np.argsort(key)
        ks = key[order]
        dup = np.zeros(len(ks), bool)
        dup[1:] = ks[1:] == ks[:-1]
        pos = rows[order[dup]]
        if len(pos) == 0:
            break
        if it < 10:
            items[pos] = np.minimum(np.searchsorted(cdf, rng.random(len(pos))), n_items - 1)
        else:
            items[pos] = rng.integers(0, n_items, len(pos))
        sel = np.unique(users[pos])
    else:
        raise RuntimeError("could not draw distinct items")
    return items


def generate(spec: SynthSpec, seed: int = 2024):
    """Returns (train, heldout) Datasets drawn from the planted model."""
    rng = np.random.default_rng(seed)
    U, I = spec.n_users, spec.n_items
    # user activity
    mean = spec.n_ratings / U
    sigma = float(np.sqrt(max(2.0 * np.log(mean / spec.user_median), 0.05)))
    raw = rng.lognormal(np.log(spec.user_median), sigma, U)
    counts = _exact_total(raw, spec.n_ratings, 1, min(spec.user_max, I))
    p = _item_weights(spec, rng)
    items = _draw_items(counts, p, rng)
    users = np.repeat(np.arange(U, dtype=np.int64), counts)
    n = len(users)
    starts = np.concatenate([[0], np.cumsum(counts)])

    # days: sessions inside an active window per user
    span = np.minimum(rng.lognormal(np.log(400.0), 0.8, U), N_DAYS - 1).astype(np.int64)
    first = (rng.random(U) * (N_DAYS - span)).astype(np.int64)
    n_sess = np.maximum(1, np.round(counts / rng.uniform(2.0, 8.0, U))).astype(np.int64)
    sess_start = np.concatenate([[0], np.cumsum(n_sess)])
    sess_day = first[np.repeat(np.arange(U), n_sess)] + (rng.random(sess_start[-1]) *
                                                        (span[np.repeat(np.arange(U), n_sess)] + 1)).astype(np.int64)
    pick = (rng.random(n) * n_sess[users]).astype(np.int64)
    days = np.minimum(sess_day[sess_start[users] + pick], N_DAYS - 1)

    # planted scores, computed in chunks
    scale = (0.45 / spec.rank) ** 0.5
    P = rng.normal(0.0, scale, (U, spec.rank)).astype(np.float32)
    Q = rng.normal(0.0, 1.0, (I, spec.rank)).astype(np.float32)
    bu = rng.normal(0.0, 0.35, U)
    bi = rng.normal(0.0, 0.45, I)
    slope = rng.normal(0.0, 0.25 / 1000.0, U)
    tmean = np.add.reduceat(days.astype(np.float64), starts[:-1]) / counts
    score = np.empty(n, np.float32)
    CH = 1 << 23
    for lo in range(0, n, CH):
        hi = min(n, lo + CH)
        uu, ii = users[lo:hi], items[lo:hi]
        s = np.einsum("nk,nk->n", P[uu], Q[ii], dtype=np.float32).astype(np.float64)
        s += bu[uu] + bi[ii] + slope[uu] * (days[lo:hi] - tmean[uu]) + rng.normal(0.0, spec.noise, hi - lo)
        score[lo:hi] = s
    cuts = np.quantile(score[:: max(1, n // 4_000_000)], np.cumsum(MARGINAL)[:-1])
    ratings = (np.searchsorted(cuts, score) + 1).astype(np.int8)

    # sort each user's ratings by (day, item); hold out the last ones
    order = np.argsort((users << 32) | (days << 16) | items)
    users, items, days, ratings = users[order], items[order], days[order], ratings[order]
    pos = np.arange(n) - starts[users]
    n_hold = np.minimum(spec.holdout, counts - 1)
    held = pos >= (counts - n_hold)[users]
    tr = Dataset.from_arrays(users[~held], items[~held], ratings[~held], days[~held], U, I, name="train")
    pr = Dataset.from_arrays(users[held], items[held], ratings[held], days[held], U, I, name="heldout")
    tr.item_years[:] = 2000
    pr.item_years[:] = 2000
    return tr, pr


def write_npy_dir(ds: Dataset, path: str) -> None:
    """Write one set in the reference's data/<set>/*.npy layout."""
    os.makedirs(path, exist_ok=True)
    np.save(os.path.join(path, "user_idxs.npy"), ds.user_idxs.astype("<i4"))
    np.save(os.path.join(path, "user_cnts.npy"), ds.user_cnts.astype("<i4"))
    np.save(os.path.join(path, "item_idxs.npy"), ds.item_idxs.astype("<i2"))
    np.save(os.path.join(path, "item_cnts.npy"), ds.item_cnts.astype("<i4"))
    np.save(os.path.join(path, "item_years.npy"), ds.item_years.astype("<i4"))
    np.save(os.path.join(path, "ratings.npy"), ds.raw_ratings.astype("<i1"))
    np.save(os.path.join(path, "dates.npy"), ds.dates.astype("<i2"))
    np.save(os.path.join(path, "is_test.npy"), ds.is_test.astype("<i1"))
