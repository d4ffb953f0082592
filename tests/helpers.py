"""Shared test helpers: seeded initial configurations and comparison metrics."""
from __future__ import annotations

import numpy as np
from scipy.spatial import cKDTree


def rsa_positions(n: int, box: float, seed: int, dmin: float = 1.0,
                  batch: int = 4096, max_rounds: int = 100000) -> np.ndarray:
    """Seeded random sequential addition of n points with min-image pair
    distance >= dmin in a periodic cube [0, box)^3.  Points are accepted in
    proposal order, so the result is a function of (n, box, seed, dmin)."""
    rng = np.random.default_rng(seed)
    acc = np.empty((0, 3))
    for _ in range(max_rounds):
        if len(acc) >= n:
            break
        cand = rng.random((batch, 3)) * box
        cand = np.minimum(cand, np.nextafter(box, 0))
        if len(acc):
            t = cKDTree(acc, boxsize=box)
            d, _ = t.query(cand, k=1)
            cand = cand[d >= dmin]
        if not len(cand):
            continue
        # greedy in proposal order among the surviving candidates
        t2 = cKDTree(cand, boxsize=box)
        pairs = t2.query_pairs(dmin, output_type="ndarray")
        keep = np.ones(len(cand), bool)
        if len(pairs):
            pairs = pairs[np.lexsort((pairs[:, 1], pairs[:, 0]))]
            for a, b in pairs:
                if keep[a]:
                    keep[b] = False
        cand = cand[keep]
        acc = np.vstack([acc, cand[: n - len(acc)]])
    if len(acc) < n:
        raise RuntimeError("rsa_positions: could not place all points")
    return acc


def maxwell_velocities(n: int, kT: float, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    v = rng.standard_normal((n, 3)) * np.sqrt(kT)
    return v - v.mean(axis=0)


def force_rel_err(f, fref, pair_abs_sum):
    """Per-particle |dF| / max(|F|, sum_j |f_ij|) -- net forces cancel, so the
    scale is the sum of pair-force magnitudes."""
    d = np.linalg.norm(np.asarray(f) - np.asarray(fref), axis=1)
    scale = np.maximum(np.linalg.norm(fref, axis=1), pair_abs_sum)
    scale = np.where(scale > 0, scale, 1.0)
    return d / scale


def pair_abs_sum(pos, box, rc, fmag):
    """sum_j |f(r_ij)| per particle over min-image pairs with r < rc."""
    pos = np.asarray(pos)
    t = cKDTree(np.mod(pos, box), boxsize=box)
    pairs = t.query_pairs(rc, output_type="ndarray")
    out = np.zeros(len(pos))
    if len(pairs):
        d = pos[pairs[:, 0]] - pos[pairs[:, 1]]
        d -= box * np.round(d / box)
        r = np.linalg.norm(d, axis=1)
        m = np.abs(fmag(r))
        np.add.at(out, pairs[:, 0], m)
        np.add.at(out, pairs[:, 1], m)
    return out


def pair_abs_sum_all(pos, box, rc, fmag, chunk=8_000_000):
    """sum_j |f(r_ij)| for every particle of a large system: all pairs i < j
    within rc at once (cKDTree.query_pairs), their magnitudes added to both
    ends with bincount, in chunks."""
    pos = np.asarray(pos)
    n = len(pos)
    pairs = cKDTree(np.mod(pos, box), boxsize=box).query_pairs(rc, output_type="ndarray")
    out = np.zeros(n)
    for a in range(0, len(pairs), chunk):
        p = pairs[a:a + chunk]
        d = pos[p[:, 1]] - pos[p[:, 0]]
        d -= box * np.round(d / box)
        w = np.abs(fmag(np.sqrt((d * d).sum(axis=1))))
        out += np.bincount(p[:, 0], weights=w, minlength=n)
        out += np.bincount(p[:, 1], weights=w, minlength=n)
    return out


def pair_abs_sum_subset(pos, box, rc, fmag, idx):
    """sum_j |f(r_ij)| for the particles idx only (large systems)."""
    pos = np.asarray(pos)
    t = cKDTree(np.mod(pos, box), boxsize=box)
    nbrs = t.query_ball_point(np.mod(pos[idx], box), rc)
    out = np.zeros(len(idx))
    for k, (i, js) in enumerate(zip(idx, nbrs)):
        js = np.array([j for j in js if j != i], dtype=np.int64)
        if len(js):
            d = pos[js] - pos[i]
            d -= box * np.round(d / box)
            out[k] = np.abs(fmag(np.linalg.norm(d, axis=1))).sum()
    return out
