"""The GPU's max_features selection stream, on the host (test helper).

Sample::sample with at_least_one (utils.rs:193-216) + choose_multiple (model.rs:264-269)
as the GPU draws them: node e of an example (0 = v, 1..cnt = neighbours, then the
negatives) uses Philox4x32-10 words of purpose 8 | (e << 4) at counter
(position in the pass, pass, purpose, i // 4), word i % 4, key = the seed.
Fixed(k), n > k: Floyd's k-subset over draws n-k .. n-1; Probability(p): position i kept
when its draw (24 bits) is below p, else one position from draw 64.
"""
import numpy as np

from cloverleaf_amd import _native as N

FIXED, PROBABILITY = 1, 2


class Draws:
    def __init__(self, seed, pos, pass_, e):
        self.key = [seed & 0xFFFFFFFF, seed >> 32]
        self.ctr = [pos, pass_, 8 | (e << 4)]
        self.cache = {}

    def word(self, i):
        blk = i >> 2
        if blk not in self.cache:
            self.cache[blk] = N.philox4x32(self.ctr + [blk], self.key)
        return int(self.cache[blk][i & 3])


def uniform_below(r, n):
    return (r * n) >> 32


def uniform01(r):
    return np.float32(r >> 8) * np.float32(1.0 / 16777216.0)


def feature_select(kind, k, p, n, seed, pos, pass_, e):
    """(kept positions, scale) of a node with n features."""
    d = Draws(seed, pos, pass_, e)
    if kind == FIXED:
        if n <= k:
            return list(range(n)), 1.0
        chosen = set()
        for i in range(n - k, n):
            t = uniform_below(d.word(i), i + 1)
            chosen.add(i if t in chosen else t)
        return sorted(chosen), float(np.float32(n) / np.float32(k))
    p32 = np.float32(p)
    kept = [i for i in range(n) if uniform01(d.word(i)) < p32]
    if kept:
        return kept, float(np.float32(1.0) / p32)
    t = uniform_below(d.word(64), n)
    return [t], float((np.float32(1.0) - p32 ** np.float32(n)) / p32)


def example_selections(kind, k, p, f_indptr, seed, pos, pass_, v, nbrs, negs):
    nodes = [v] + list(nbrs) + list(negs)
    return [feature_select(kind, k, p, int(f_indptr[u + 1] - f_indptr[u]), seed, pos, pass_, e)
            for e, u in enumerate(nodes)]
