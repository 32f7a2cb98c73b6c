"""Seeded synthetic inputs of the benchmark configurations.

SplitMix64 streams, one per quantity (coordinates, strengths, shuffles),
from a per-configuration base seed; uniform coordinates are
``f32(-pi + 2 pi u)`` with ``u = (r >> 40) 2^-24`` (clamped below pi),
normal variates come from Box-Muller in f64, Gaussian clusters are wrapped
into ``[-pi, pi)``.  All generation is vectorized numpy on the host.
"""
import numpy as np

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
BASE_SEED = 0x4E55465400000000  # "NUFT" << 32


def config_seed(config, stream):
    """Base seed of a configuration (1..5) and an independent stream id."""
    return (BASE_SEED | int(config)) ^ (int(stream) * 0x632BE59BD9B4E019 & 0xFFFFFFFFFFFFFFFF)


def splitmix64(seed, n):
    """n outputs of the SplitMix64 generator started at `seed`."""
    with np.errstate(over="ignore"):
        i = np.arange(1, n + 1, dtype=np.uint64)
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + i * GOLDEN
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed, n):
    """Uniforms in [0, 1) with 24-bit resolution."""
    r = splitmix64(seed, n)
    return (r >> np.uint64(40)).astype(np.float64) * (2.0 ** -24)


def uniform_coords(seed, n, dim, lo=-np.pi, hi=np.pi, dtype=np.float32):
    u = uniform01(seed, n * dim).reshape(n, dim)
    x = (lo + (hi - lo) * u).astype(dtype)
    top = np.nextafter(dtype(hi), dtype(lo))
    return np.minimum(x, top)


def normal(seed, n):
    """n standard normals (Box-Muller on two uniform streams)."""
    m = (n + 1) // 2
    u1 = uniform01(seed, m)
    u2 = uniform01(seed ^ 0x5DEECE66D, m)
    u1 = np.maximum(u1, 2.0 ** -25)
    r = np.sqrt(-2.0 * np.log(u1))
    z = np.concatenate([r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)])
    return z[:n]


def complex_normal(seed, n, dtype=np.complex64):
    re = normal(seed, n)
    im = normal(seed ^ 0xA5A5A5A5A5A5A5A5, n)
    return (re + 1j * im).astype(dtype)


def wrap(x):
    return np.mod(x + np.pi, 2 * np.pi) - np.pi


def clustered_coords(seed, n, center, sigma, dtype=np.float32):
    dim = len(center)
    z = normal(seed, n * dim).reshape(n, dim)
    x = wrap(np.asarray(center, dtype=np.float64)[None, :] + sigma * z).astype(dtype)
    top = np.nextafter(dtype(np.pi), dtype(0))
    x = np.where(x >= np.pi, top, x)
    return np.where(x < -np.pi, dtype(-np.pi), x).astype(dtype)


def shuffle_perm(seed, n):
    """A Fisher-Yates-equivalent random permutation (argsort of random keys)."""
    return np.argsort(splitmix64(seed, n), kind="stable")


def config2_points(M=10_000_000, dtype=np.float32):
    """Half uniform in [-pi, pi)^3, half N((0.3, -1.1, 2.0), 0.05^2) wrapped."""
    half = M // 2
    a = uniform_coords(config_seed(2, 1), half, 3, dtype=dtype)
    b = clustered_coords(config_seed(2, 2), M - half, (0.3, -1.1, 2.0), 0.05, dtype=dtype)
    return np.concatenate([a, b])


def radial_spokes(spokes=8192, samples=512, dtype=np.float32):
    """Fibonacci-sphere spokes through the origin (radial 3D trajectory)."""
    s = np.arange(spokes, dtype=np.float64)
    z = 1.0 - (2 * s + 1) / spokes
    r = np.sqrt(1.0 - z * z)
    phi = s * np.pi * (3.0 - np.sqrt(5.0))
    d = np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=1)
    t = -np.pi + 2 * np.pi * np.arange(samples, dtype=np.float64) / samples
    pts = (t[None, :, None] * d[:, None, :]).reshape(-1, 3)
    return pts.astype(dtype)


def batch_problem_specs(count=64):
    """The 64 independent 2D type-2 problems of the batching workload:
    (n0, n1, M, coordinate seed, mode seed) per problem, modes per axis
    log-uniform in [64, 512], M log-uniform in [1e4, 2e6]."""
    r = uniform01(config_seed(4, 1), count * 3).reshape(count, 3)
    out = []
    for i in range(count):
        n0 = int(round(np.exp(np.log(64) + r[i, 0] * (np.log(512) - np.log(64)))))
        n1 = int(round(np.exp(np.log(64) + r[i, 1] * (np.log(512) - np.log(64)))))
        M = int(round(np.exp(np.log(1e4) + r[i, 2] * (np.log(2e6) - np.log(1e4)))))
        out.append((n0, n1, M, config_seed(4, 100 + i), config_seed(4, 200 + i)))
    return out
