"""Synthetic workloads with fixed seeds (numpy PCG64 default_rng).

  c1  6 isotropic Gaussians in 16-d, n=10,000 (means U[-10,10]^16, sigma 1),
      self-kNN k=10, pairwise on the first 2,000 points.
  c2  DB 1,000,000 x 128 uniform [0,1) f32, 100,000 queries, k=10.
  c3  70,000 x 784 integer-valued f32: 80% zeros, nonzeros rint(Beta(2,2)*255);
      self-kNN k=20 feeding LOF.
  c4  512 self-kNN instances, n log-uniform in [1e3, 2e5], d in {8,32,64,256},
      3-20 Gaussian components (means U[-5,5]^d, sigma 1), k=15.
  c5  10,000,000 x 96 from a 50-component Gaussian mixture (means U[-10,10]^96,
      sigma 1), k=10, served under Poisson arrivals.
Every generator takes a `scale` to draw the same distribution at a smaller size
(tests use small scales; benchmarks use 1.0).
"""
import numpy as np


def gaussian_mixture(rng, n, d, comps, mean_lo, mean_hi, sigma=1.0, dtype=np.float32):
    means = rng.uniform(mean_lo, mean_hi, size=(comps, d))
    labels = rng.integers(0, comps, size=n)
    x = means[labels] + rng.normal(0.0, sigma, size=(n, d))
    return x.astype(dtype)


def c1(n=10_000, d=16, seed=1):
    rng = np.random.default_rng(seed)
    return gaussian_mixture(rng, n, d, 6, -10.0, 10.0)


def c2(n=1_000_000, m=100_000, d=128, seed_db=2, seed_q=3):
    db = np.random.default_rng(seed_db).random((n, d), dtype=np.float32)
    q = np.random.default_rng(seed_q).random((m, d), dtype=np.float32)
    return db, q


def c3(n=70_000, d=784, seed=4):
    rng = np.random.default_rng(seed)
    mask = rng.random((n, d)) < 0.2
    vals = np.rint(rng.beta(2.0, 2.0, size=(n, d)) * 255.0)
    return (mask * vals).astype(np.float32)


def c4(instances=512, n_lo=1_000, n_hi=200_000, seed=5, scale=1.0):
    """List of (X, k) self-kNN problems."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(instances):
        n = int(round(np.exp(rng.uniform(np.log(n_lo), np.log(n_hi))) * scale))
        n = max(n, 16)
        d = int(rng.choice([8, 32, 64, 256]))
        comps = int(rng.integers(3, 21))
        out.append(gaussian_mixture(rng, n, d, comps, -5.0, 5.0))
    return out


def c5(n=10_000_000, d=96, comps=50, seed=6, chunk=1_000_000):
    rng = np.random.default_rng(seed)
    means = rng.uniform(-10.0, 10.0, size=(comps, d))
    x = np.empty((n, d), np.float32)
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        lab = rng.integers(0, comps, size=e - s)
        x[s:e] = (means[lab] + rng.normal(0.0, 1.0, size=(e - s, d))).astype(np.float32)
    return x, means
