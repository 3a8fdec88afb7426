"""Seeded synthetic inputs for the benchmark configurations and the tests.

The reference draws its vectors from cuRAND streams seeded by the clock
(generateProblems.cu:25-428, compareMultiselect.hpp:106-126), which cannot be
reproduced; these generators keep its distribution shapes with a counter-based
SplitMix64 stream, so every input is a pure function of (seed, index) and the
same on every machine.  -0.0 is canonicalised to +0.0 and nothing is NaN.
"""
from __future__ import annotations

import numpy as np

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
M1 = np.uint64(0xBF58476D1CE4E5B9)
M2 = np.uint64(0x94D049BB133111EB)
CHUNK = 1 << 24


def splitmix64(seed: int, start: int, count: int) -> np.ndarray:
    """Outputs start..start+count-1 of the SplitMix64 sequence of `seed`."""
    with np.errstate(over="ignore"):
        i = np.arange(start + 1, start + count + 1, dtype=np.uint64)
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + i * GOLDEN
        z = (z ^ (z >> np.uint64(30))) * M1
        z = (z ^ (z >> np.uint64(27))) * M2
        return z ^ (z >> np.uint64(31))


def _chunked(n: int, fn, dtype) -> np.ndarray:
    out = np.empty(n, dtype)
    for s in range(0, n, CHUNK):
        e = min(n, s + CHUNK)
        out[s:e] = fn(s, e - s)
    return out


def uniform_f32(seed: int, n: int) -> np.ndarray:
    """U[0,1) on the 24-bit grid: (x >> 40) * 2^-24."""
    return _chunked(n, lambda s, c: (splitmix64(seed, s, c) >> np.uint64(40)).astype(np.float32)
                    * np.float32(2.0 ** -24), np.float32)


def uniform_f64(seed: int, n: int) -> np.ndarray:
    """U[0,1) on the 53-bit grid."""
    return _chunked(n, lambda s, c: (splitmix64(seed, s, c) >> np.uint64(11)).astype(np.float64)
                    * (2.0 ** -53), np.float64)


def _normal64(seed: int, s: int, c: int) -> np.ndarray:
    # Box-Muller in double from two independent streams of the same seed
    u1 = ((splitmix64(seed, s, c) >> np.uint64(11)).astype(np.float64) + 1.0) * (2.0 ** -53)
    u2 = (splitmix64(seed ^ 0x5DEECE66D, s, c) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)


def _canon(x: np.ndarray) -> np.ndarray:
    x[x == 0] = 0  # -0.0 -> +0.0
    return x


def normal_f32(seed: int, n: int) -> np.ndarray:
    return _canon(_chunked(n, lambda s, c: _normal64(seed, s, c).astype(np.float32), np.float32))


def normal_f64(seed: int, n: int) -> np.ndarray:
    return _canon(_chunked(n, lambda s, c: _normal64(seed, s, c), np.float64))


def cauchy_f32(seed: int, n: int) -> np.ndarray:
    """Cauchy(0,1) = tan(pi (u - 1/2)), clipped to finite float32
    (generateProblems.cu:212-247 shape)."""
    fmax = np.finfo(np.float32).max

    def gen(s, c):
        u = (splitmix64(seed, s, c) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)
        u = np.clip(u, 2.0 ** -40, 1.0 - 2.0 ** -40)
        return np.clip(np.tan(np.pi * (u - 0.5)), -fmax, fmax).astype(np.float32)
    return _canon(_chunked(n, gen, np.float32))


def uint_range(seed: int, n: int, lo: int, hi: int) -> np.ndarray:
    """uint32 uniform in [lo, hi] (inclusive)."""
    span = np.uint64(hi - lo + 1)
    return _chunked(n, lambda s, c: (np.uint64(lo) + (splitmix64(seed, s, c) >> np.uint64(16)) % span)
                    .astype(np.uint32), np.uint32)


def int_range(seed: int, n: int, lo: int, hi: int) -> np.ndarray:
    span = hi - lo + 1
    return _chunked(n, lambda s, c: (np.int64(lo) + ((splitmix64(seed, s, c) >> np.uint64(16))
                                                     % np.uint64(span)).astype(np.int64))
                    .astype(np.int32), np.int32)


def plateau_f32(seed: int, n: int, value: float = 0.5, frac: float = 0.5) -> np.ndarray:
    """Exactly round(frac*n) entries equal to `value`, the rest U[0,1), in a
    seeded random order."""
    x = uniform_f32(seed, n)
    m = int(round(frac * n))
    # positions: a seeded permutation's first m slots
    keys = splitmix64(seed ^ 0xA5A5A5A5, 0, n)
    idx = np.argpartition(keys, m - 1)[:m] if m > 0 else np.empty(0, np.int64)
    x[idx] = np.float32(value)
    return x


def permutation_u32(seed: int, n: int) -> np.ndarray:
    """A seeded permutation of 1..n (the k-th largest is n - k + 1)."""
    keys = splitmix64(seed, 0, n)
    return (np.argsort(keys, kind="stable") + 1).astype(np.uint32)


def sorted_u32(seed: int, n: int, start: int = 0, total: int | None = None) -> np.ndarray:
    """Positions start..start+n-1 of the non-decreasing sequence of length
    `total` (default start + n): floor((j 2^32 + r_j) / total), r_j = the top
    32 bits of SplitMix64 output j (the device family "sorted"; stands in for
    generateProblems.cu:34-45, which sorts random uints)."""
    total = start + n if total is None else total
    assert 1 <= total < 2 ** 32 and start + n <= total

    def gen(s, c):
        j = np.arange(start + s, start + s + c, dtype=np.uint64)
        r = splitmix64(seed, start + s, c) >> np.uint64(32)
        return ((j << np.uint64(32)) | r) // np.uint64(total)
    return _chunked(n, lambda s, c: gen(s, c).astype(np.uint32), np.uint32)


def sorted_as(dtype, u: np.ndarray) -> np.ndarray:
    """The "sorted" family in another dtype, from its uint32 values."""
    d = np.dtype(dtype)
    if d == np.uint32:
        return u
    if d == np.int32:
        return (u ^ np.uint32(0x80000000)).view(np.int32)
    v = u.astype(np.float64) * 2.0 ** -32
    return v.astype(np.float32) if d == np.float32 else v


def killer_f32(seed: int, n: int) -> np.ndarray:
    """Bucket killer (generateProblems.cu:344-367): uniform floats with bits
    masked by 0x80EFFFFF (tiny magnitudes) and positions 0..252 holding the
    powers of two 2^-126..2^126 (unsorted; no result depends on order)."""
    u = uniform_f32(seed, n).view(np.uint32) & np.uint32(0x80EFFFFF)
    m = min(n, 253)
    u[:m] = (np.arange(m, dtype=np.uint32) + np.uint32(1)) << np.uint32(23)
    return u.view(np.float32)


def killer_f64(seed: int, n: int) -> np.ndarray:
    """Bucket killer doubles (generateProblems.cu:385-423): bits masked by
    0x800FFFFFFFFFFFFF, positions 0..2044 holding 2^-1022..2^1022."""
    u = uniform_f64(seed, n).view(np.uint64) & np.uint64(0x800FFFFFFFFFFFFF)
    m = min(n, 2045)
    u[:m] = (np.arange(m, dtype=np.uint64) + np.uint64(1)) << np.uint64(52)
    return u.view(np.float64)


def huge_f32(seed: int, n: int) -> np.ndarray:
    """U[0,1) * 10^6 in float32 arithmetic (generateProblems.cu:287-298)."""
    return uniform_f32(seed, n) * np.float32(1.0e6)


def ones_twos(seed: int, n: int, dtype=np.float32) -> np.ndarray:
    """1 or 2 with probability 1/2 each (generateProblems.cu:154-177 shape)."""
    return _chunked(n, lambda s, c: (np.uint64(1) + (splitmix64(seed, s, c) >> np.uint64(63)))
                    .astype(dtype), dtype)


# ---------------------------------------------------------------- k sets
def k_percentiles(n: int) -> np.ndarray:
    """101 ranks k = round(p (n-1)) + 1, p = 0.00 .. 1.00 (p = 0 is the max)."""
    p = np.arange(101, dtype=np.float64) / 100.0
    return (np.rint(p * (n - 1)) + 1).astype(np.uint32)


def k_uniform(seed: int, n: int, count: int, dedup: bool = True) -> np.ndarray:
    """count ranks uniform in [1, n]; sorted and deduplicated when dedup."""
    k = (splitmix64(seed, 0, count) % np.uint64(n) + np.uint64(1)).astype(np.uint32)
    return np.unique(k) if dedup else k


def k_cluster(seed: int, n: int, count: int = 256, width: int = 1000) -> np.ndarray:
    """count distinct ranks in [n/2 - width/2, n/2 + width/2 - 1] plus 1 and n."""
    lo = n // 2 - width // 2
    keys = splitmix64(seed, 0, width)
    pick = np.sort(np.argsort(keys, kind="stable")[:count]) + lo
    return np.concatenate([pick.astype(np.uint32), np.array([1, n], np.uint32)])


def _unit_f32(seed: int, count: int) -> np.ndarray:
    """float32 uniforms in [0,1) on the 24-bit grid (stand-in for
    curandGenerateUniform in the reference's k generators)."""
    return (splitmix64(seed, 0, count) >> np.uint64(40)).astype(np.float32) * np.float32(2.0 ** -24)


def k_even(count: int, n: int) -> np.ndarray:
    """"Uniform Ks" (generateProblems.cu:506-513): k_0 = 1, k_i =
    (uint)((i / (float)count) * n) in float32 arithmetic, k_{count-1} = n."""
    i = np.arange(count, dtype=np.float32)
    k = ((i / np.float32(count)) * np.float32(n)).astype(np.uint32)
    k[0] = 1
    k[-1] = n
    return np.clip(k, 1, n).astype(np.uint32)


def k_normal(seed: int, count: int, n: int) -> np.ndarray:
    """"Normal Random Ks" (generateProblems.cu:515-531): (|z| + 7) * (n / 14)
    with z ~ N(0,1); the reference can exceed n for |z| > 7, clamped here."""
    z = _normal64(seed, 0, count)
    k = ((np.abs(z) + 7.0) * (n / 14.0)).astype(np.uint64)
    return np.clip(k, 1, n).astype(np.uint32)


def k_cluster9(seed: int, count: int, n: int) -> np.ndarray:
    """"Cluster Ks" (generateProblems.cu:533-557): clusters of 9 consecutive
    ranks c-4..c+4 around centres c = (uint)(u (n - 10) + 5).  The
    reference leaves the last count mod 9 entries unwritten when count is
    not a multiple of 9; here they continue the last cluster upwards
    (clamped to n)."""
    assert n >= 10
    nc = max(1, (count + 8) // 9)
    u = _unit_f32(seed, nc)
    centres = (u * np.float32(n - 10) + np.float32(5)).astype(np.int64)
    k = centres[:, None] + np.arange(-4, 5, dtype=np.int64)[None, :]
    return np.clip(k.reshape(-1)[:count], 1, n).astype(np.uint32)


def k_sectioned(seed: int, count: int, n: int) -> np.ndarray:
    """"Sectioned Ks" (generateProblems.cu:559-576): count consecutive ranks
    from k_0 = (uint)(u (n - count)), kept >= 1 here (the reference can
    start at 0, an invalid rank)."""
    assert count <= n
    k0 = int((_unit_f32(seed, 1)[0] * np.float32(n - count)).astype(np.uint32))
    k0 = max(1, k0)
    k = np.arange(k0, k0 + count, dtype=np.uint64)
    return np.clip(k, 1, n).astype(np.uint32)


# ---------------------------------------------------------------- batches
def batch_spec(master_seed: int, nseg: int, log2_lo: float = 10.0, log2_hi: float = 22.0,
               max_k: int = 64, max_topk: int = 4096, topk_frac: float = 0.5) -> dict:
    """An uneven batch: segment lengths floor(2^U), U ~ U[log2_lo, log2_hi];
    values U[0,1) or N(0,1) by a seeded coin (per-segment seed from the
    master stream); each segment asks either 1..max_k uniform k values in
    [1, len] or a top-K with K = floor(2^V), V ~ U[0, log2(min(max_topk, len))].
    Returns lengths, kinds (ggks kind codes), num_k, the concatenated k list,
    per-segment value seeds and families."""
    z = splitmix64(master_seed, 0, 6 * nseg).reshape(nseg, 6)
    unit = lambda w: (w >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)  # noqa: E731
    lengths = np.floor(2.0 ** (log2_lo + unit(z[:, 0]) * (log2_hi - log2_lo))).astype(np.uint64)
    lengths = np.maximum(lengths, 1)
    normal = (z[:, 1] & np.uint64(1)).astype(bool)
    topk = unit(z[:, 2]) < topk_frac
    nk = (1 + z[:, 3] % np.uint64(max_k)).astype(np.uint32)
    kmax = np.minimum(lengths, max_topk).astype(np.float64)
    K = np.floor(2.0 ** (unit(z[:, 4]) * np.log2(kmax))).astype(np.uint32)
    K = np.clip(K, 1, kmax.astype(np.uint32))
    seeds = z[:, 5]
    num_k = np.where(topk, K, nk).astype(np.uint32)
    klists = []
    for i in np.nonzero(~topk)[0]:
        klists.append((splitmix64(int(seeds[i]), 0, int(nk[i])) % lengths[i] + np.uint64(1))
                      .astype(np.uint32))
    k = np.concatenate(klists) if klists else np.zeros(0, np.uint32)
    return {"lengths": lengths, "kinds": np.where(topk, 2, 0).astype(np.uint32),
            "num_k": num_k, "k": k, "seeds": seeds, "normal": normal}


def batch_values_f32(spec: dict) -> np.ndarray:
    """The batch vector of a batch_spec on the host (segments back to back)."""
    parts = [(normal_f32 if nm else uniform_f32)(int(sd), int(n))
             for n, sd, nm in zip(spec["lengths"], spec["seeds"], spec["normal"])]
    return np.concatenate(parts) if parts else np.zeros(0, np.float32)
