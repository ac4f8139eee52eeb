"""Independent restatement, for tests, of the C++ random draws the reference's k-means
initialisations make (kmeans.cpp:1404-1811 use std::mt19937_64 with
std::uniform_int_distribution<int>, std::discrete_distribution<int> and
std::uniform_real_distribution<double> as libstdc++ implements them), and of the two
initialisations themselves in exact arithmetic. On integer-valued data the squared
distances are exact whichever way they are computed, so these sequences are the
reference's own. Nothing here is imported by the library.

Sources restated:
  * mt19937_64: the C++ standard's parameters ([rand.predef]); pinned by the standard's
    requirement that the 10000th output of a default-constructed engine is
    9981545732273789042.
  * uniform_int_distribution: Lemire's nearly divisionless method with 128-bit products
    (D. Lemire, "Fast Random Integer Generation in an Interval", ACM TOMACS 29(1), 2019),
    which libstdc++ uses for a 64-bit engine.
  * generate_canonical<double, 53>: one engine call, x / 2^64 with x rounded to double,
    values >= 1 replaced by the largest double below 1.
  * discrete_distribution: weights normalised by their sum, cumulative sums with the last
    set to 1, then the first cumulative value >= a canonical draw; fewer than two weights
    give 0 without a draw.
"""
import math

import numpy as np

M64 = (1 << 64) - 1


class MT19937_64:
    N, M = 312, 156
    A = 0xB5026F5AA96619E9
    UPPER, LOWER = M64 ^ ((1 << 31) - 1), (1 << 31) - 1

    def __init__(self, seed=5489):
        self.mt = [0] * self.N
        self.mt[0] = seed & M64
        for i in range(1, self.N):
            prev = self.mt[i - 1]
            self.mt[i] = (6364136223846793005 * (prev ^ (prev >> 62)) + i) & M64
        self.idx = self.N

    def _twist(self):
        mt = self.mt
        for i in range(self.N):
            x = (mt[i] & self.UPPER) | (mt[(i + 1) % self.N] & self.LOWER)
            xa = x >> 1
            if x & 1:
                xa ^= self.A
            mt[i] = mt[(i + self.M) % self.N] ^ xa
        self.idx = 0

    def __call__(self):
        if self.idx >= self.N:
            self._twist()
        y = self.mt[self.idx]
        self.idx += 1
        y ^= (y >> 29) & 0x5555555555555555
        y ^= (y << 17) & 0x71D67FFFEDA60000
        y ^= (y << 37) & 0xFFF7EEE000000000
        y ^= y >> 43
        return y & M64


def uniform_int(g, a, b):
    rng = b - a + 1
    product = g() * rng
    low = product & M64
    if low < rng:
        threshold = (-rng) % rng
        while low < threshold:
            product = g() * rng
            low = product & M64
    return a + (product >> 64)


def canonical(g):
    r = float(g()) / 18446744073709551616.0
    return r if r < 1.0 else math.nextafter(1.0, 0.0)


class Discrete:
    def __init__(self, w):
        w = [float(v) for v in w]
        if len(w) < 2:
            self.cp = []
            return
        s = 0.0
        for v in w:
            s += v
        cp, acc = [], 0.0
        for v in w:
            acc += v / s
            cp.append(acc)
        cp[-1] = 1.0
        self.cp = cp

    def __call__(self, g):
        if not self.cp:
            return 0
        p = canonical(g)
        lo, hi = 0, len(self.cp)
        while lo < hi:                      # first position with cp >= p
            mid = (lo + hi) // 2
            if self.cp[mid] < p:
                lo = mid + 1
            else:
                hi = mid
        return lo


def _sq(X, i, j):
    return float(sum((X[i, l] - X[j, l]) ** 2 for l in range(X.shape[1])))


def kmeanspp(X, k, seed, dist=_sq):
    """Indices of the rows k-means++ picks (kmeans.cpp:1524-1664), exact weights dist(X, i, j)
    (squared Euclidean by default)."""
    g = MT19937_64(seed)
    n = X.shape[0]
    first = uniform_int(g, 0, n - 1)
    w = [dist(X, i, first) for i in range(n)]
    w[first] = 0.0
    if not any(v > 0 for v in w):
        return list(range(k))
    chosen = [first]
    trials = 2 + int(math.log(k))
    for _ in range(1, k):
        dd = Discrete(w)
        cand = [dd(g) for _ in range(trials)]
        best, best_cost, best_w = 0, math.inf, None
        for t, c in enumerate(cand):
            if c in cand[:t]:
                continue
            w5 = [max(0.0, min(w[i], dist(X, i, c))) for i in range(n)]
            cost = 0.0
            for v in w5:
                cost += v
            if cost < best_cost:
                best, best_cost, best_w = c, cost, w5
        chosen.append(best)
        w = best_w
        w[best] = 0.0
    return chosen


def afkmc2(X, k, seed, m, dist=_sq):
    """Indices of the rows AFK-MC2 picks (kmeans.cpp:1667-1811), exact distances dist(X, i, j)
    (squared Euclidean by default)."""
    g = MT19937_64(seed)
    n = X.shape[0]
    c1 = uniform_int(g, 0, n - 1)
    w = [max(0.0, dist(X, i, c1)) for i in range(n)]
    s = 0.0
    for v in w:
        s += v
    q = [0.5 * v / s + 1.0 / (2 * n) for v in w] if s > 0 else [1.0 / n] * n
    chosen = [c1]
    qd = Discrete(q)

    def near(i):
        return min(max(0.0, dist(X, i, c)) for c in chosen)

    for _ in range(1, k):
        x = qd(g)
        dx = near(x)
        for _ in range(m):
            y = qd(g)
            dy = near(y)
            dxqy = dx * q[y]
            if dxqy == 0.0 or (dy * q[x]) / dxqy > canonical(g):
                x, dx = y, dy
        chosen.append(x)
    return chosen


def selection_sample(n, k, seed):
    """Rows of the random initialisation (kmeans.cpp:1404-1425 through da_std::sample,
    da_std.hpp:102-118): row i of the `left` not yet visited is taken when a uniform draw
    from [0, left - 1] is below the number still needed."""
    g = MT19937_64(seed)
    out, need, left = [], k, n
    i = 0
    while need > 0:
        if uniform_int(g, 0, left - 1) < need:
            out.append(i)
            need -= 1
        i += 1
        left -= 1
    return out


def rows(X, idx):
    return np.asarray(X, dtype=np.float64)[np.asarray(idx)]
