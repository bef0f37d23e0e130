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


class StreamDraws(Draws):
    """Philox words of an explicit purpose word at (position, pass)."""

    def __init__(self, seed, pos, pass_, purpose):
        self.key = [seed & 0xFFFFFFFF, seed >> 32]
        self.ctr = [pos, pass_, purpose]
        self.cache = {}


def random_attention_blocks(kind, k, p, heads, seed, pos, pass_, block_sizes):
    """The GPU's Random attention stream (compute_random_attention_matrix,
    attention.rs:221-250) for the blocks of one example (block e of n items; e = 0 the
    anchor, then the neighbours, then the negatives): per head (mask, scale).  Purpose
    10 | (e << 8) | (h << 4): the Probability size (keep draw j per column); purpose
    11 | ...: row i's Floyd draws at i * 64 + t."""
    out = []
    for e, n in enumerate(block_sizes):
        heads_out = []
        for h in range(heads):
            tag = (e << 8) | (h << 4)
            if n == 1:
                heads_out.append((np.ones((1, 1), dtype=bool), 1.0))  # unused: V of head 0
                continue
            if kind == FIXED:
                kk = min(k, n)
                sc = float(np.float32(n) / np.float32(kk))
            else:
                d = StreamDraws(seed, pos, pass_, 10 | tag)
                p32 = np.float32(p)
                kk = sum(1 for j in range(n) if uniform01(d.word(j)) < p32)
                sc = float(np.float32(1.0) / p32)
                if kk == 0:
                    kk = 1
                    sc = float((np.float32(1.0) - p32 ** np.float32(n)) / p32)
            d = StreamDraws(seed, pos, pass_, 11 | tag)
            mask = np.zeros((n, n), dtype=bool)
            for i in range(n):
                if kk >= n:
                    mask[i, :] = True
                    continue
                chosen = set()
                for t in range(n - kk, n):
                    r = uniform_below(d.word(i * 64 + t), t + 1)
                    chosen.add(t if r in chosen else r)
                mask[i, sorted(chosen)] = True
            heads_out.append((mask, sc))
        out.append(heads_out)
    return out
