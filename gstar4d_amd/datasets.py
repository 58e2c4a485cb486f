"""Seeded synthetic point clouds for the benchmark configurations.

Every FP64 uniform is (splitmix64(seed, i) >> 11) * 2^-53, counter based, so any slice of a
stream is reproducible without generating the prefix.  Gaussians use Box-Muller on two
uniform streams.  All generators return (N, 3) float64 arrays without exact duplicates.
"""
import numpy as np

_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)
_GOLD = np.uint64(0x9E3779B97F4A7C15)


def splitmix64(seed, start, count):
    """count consecutive outputs of the SplitMix64 stream `seed`, starting at index start."""
    with np.errstate(over="ignore"):
        i = np.arange(start, start + count, dtype=np.uint64) + np.uint64(1)
        z = np.uint64(seed) * np.uint64(0xD1B54A32D192ED03) + i * _GOLD
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def uniform01(seed, count, start=0):
    return (splitmix64(seed, start, count) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def _stream(seed, tag):
    return (int(seed) * 1000003 + tag) & 0xFFFFFFFFFFFFFFFF


def gaussians(seed, count):
    u1 = uniform01(_stream(seed, 11), count)
    u2 = uniform01(_stream(seed, 12), count)
    u1 = np.where(u1 <= 0.0, 2.0 ** -53, u1)
    r = np.sqrt(-2.0 * np.log(u1))
    return r * np.cos(2.0 * np.pi * u2)


def _dedup(p):
    """Drops exact duplicate rows, keeping first occurrences in order.  Equal rows share x, so
    the full row comparison runs only when some x value repeats."""
    xs = np.sort(p[:, 0])
    if not (xs[1:] == xs[:-1]).any():
        return p
    _, keep = np.unique(p, axis=0, return_index=True)
    return p[np.sort(keep)]


def uniform_cube(n, seed=0):
    """C1/C2: i.i.d. uniform in [0,1)^3 (distinct with overwhelming probability; checked)."""
    p = uniform01(seed, 3 * n).reshape(n, 3)
    return _dedup(p)


def _clusters(n, centres, sigmas, seed):
    """Gaussian clusters, resampled until inside [0,1)^3."""
    k = centres.shape[0]
    pick = (splitmix64(_stream(seed, 21), 0, n) % np.uint64(k)).astype(np.int64)
    out = np.empty((n, 3))
    todo = np.arange(n)
    rnd = 0
    while todo.size:
        g = gaussians(_stream(seed, 100 + rnd), 3 * todo.size).reshape(-1, 3)
        cand = centres[pick[todo]] + sigmas[pick[todo], None] * g
        ok = np.all((cand >= 0.0) & (cand < 1.0), axis=1)
        out[todo[ok]] = cand[ok]
        todo = todo[~ok]
        rnd += 1
    return out


def hard_mix(n=2_000_000, seed=0):
    """C3: 50% Gaussian clusters (64 centres, sigma 0.01), 25% sphere r=0.4 about the cube
    centre, 25% of the 80^3 lattice scaled by 1/80 (subsampled without replacement); exact
    duplicates removed."""
    n_cl, n_sp = n // 2, n // 4
    n_la = n - n_cl - n_sp
    centres = uniform01(_stream(seed, 1), 64 * 3).reshape(64, 3)
    cl = _clusters(n_cl, centres, np.full(64, 0.01), _stream(seed, 2))
    g = gaussians(_stream(seed, 3), 3 * n_sp).reshape(n_sp, 3)
    g /= np.linalg.norm(g, axis=1)[:, None]
    sp = 0.5 + 0.4 * g
    lat_n = 80 ** 3
    keys = splitmix64(_stream(seed, 4), 0, lat_n)
    chosen = np.argsort(keys, kind="stable")[:min(n_la, lat_n)]
    ijk = np.stack(np.unravel_index(chosen, (80, 80, 80)), axis=1).astype(np.float64)
    la = ijk / 80.0
    return _dedup(np.concatenate([cl, sp, la]))


def clustered_mix(n=8_000_000, seed=0):
    """C4: 70% from 256 Gaussian clusters (sigma log-uniform in [0.002, 0.05]) + 30% uniform."""
    n_cl = int(round(0.7 * n))
    centres = uniform01(_stream(seed, 31), 256 * 3).reshape(256, 3)
    u = uniform01(_stream(seed, 32), 256)
    sig = np.exp(np.log(0.002) + u * (np.log(0.05) - np.log(0.002)))
    cl = _clusters(n_cl, centres, sig, _stream(seed, 33))
    un = uniform01(_stream(seed, 34), 3 * (n - n_cl)).reshape(-1, 3)
    return _dedup(np.concatenate([cl, un]))


def batch_sets(nsets=4096, seed=0, lo=1000, hi=200000):
    """C5: set sizes log-uniform in [lo, hi]; each set uniform or one Gaussian (sigma 0.1)."""
    u = uniform01(_stream(seed, 41), nsets)
    sizes = np.floor(np.exp(np.log(lo) + u * (np.log(hi) - np.log(lo)))).astype(np.int64)
    coin = uniform01(_stream(seed, 42), nsets) < 0.5
    sets = []
    for k in range(nsets):
        s = _stream(seed, 1000 + k)
        if coin[k]:
            p = uniform01(s, 3 * int(sizes[k])).reshape(-1, 3)
        else:
            c = np.full((1, 3), 0.5)
            p = _clusters(int(sizes[k]), c, np.array([0.1]), s)
        sets.append(_dedup(p))
    return sets


CONFIGS = {
    "c1_uniform_100k_g128": dict(n=100_000, grid=128, gen="uniform"),
    "c2_uniform_4m_g512": dict(n=4_194_304, grid=512, gen="uniform"),
    "c3_hard_2m_g512": dict(n=2_000_000, grid=512, gen="hard"),
    "c4_clustered_8m_g512": dict(n=8_000_000, grid=512, gen="clustered"),
}


def make(name, seed=0, n=None):
    c = CONFIGS[name]
    n = n or c["n"]
    if c["gen"] == "uniform":
        return uniform_cube(n, seed)
    if c["gen"] == "hard":
        return hard_mix(n, seed)
    return clustered_mix(n, seed)
