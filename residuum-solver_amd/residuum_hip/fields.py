"""Seeded synthetic inputs: white noise, Gaussian random fields on element
centroids of a structured block mesh, and sampling without replacement.
All randomness comes from the library's Philox4x32-10 streams, so every
host and language reproduces the same fields."""
import numpy as np

from . import philox_normal, philox_uniform


def gaussian_random_field(seed, n, ell, std, mean=0.0):
    """A Gaussian random field on the n = (nx, ny[, nz]) elements of a
    structured mesh of the unit square / cube, element order x fastest:
    Philox white noise (stream 0) smoothed by a separable Gaussian of
    standard deviation ell (in domain units; ell * n_k cells along axis k,
    reflecting boundaries, truncated at 4 standard deviations), then
    shifted and scaled to the given mean and standard deviation."""
    from scipy.ndimage import gaussian_filter
    n = tuple(int(v) for v in n)
    white = philox_normal(seed, 0, int(np.prod(n))).reshape(n[::-1])
    sigma = [ell * k for k in n[::-1]]
    f = gaussian_filter(white, sigma=sigma, mode="reflect", truncate=4.0)
    f = (f - f.mean()) / f.std()
    return (mean + std * f).reshape(-1)


def sample_without_replacement(seed, n, k):
    """k distinct indices of range(n): the k smallest Philox uniforms (stream
    0) decide, ties broken by index; returned in ascending order."""
    u = philox_uniform(seed, 0, n)
    idx = np.argsort(u, kind="stable")[:k]
    return np.sort(idx)
