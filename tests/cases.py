"""Shared synthetic inputs of the tests (no compute of the product or the oracle)."""
import numpy as np


def two_balls(shape=(20, 18, 16), seed=5):
    """a [0, 255] image with two bright balls of different intensity and noise, and an
    initial contour of two slightly larger balls: the mask stays two 6-connected
    components for the first iterations, so a seeded extraction has something to remove.
    Returns (u0, phi0, centre of the larger ball)."""
    i, j, k = np.meshgrid(*[np.arange(s) for s in shape], indexing="ij")
    c1 = (shape[0] * 3 // 10, shape[1] // 3, shape[2] // 2)
    c2 = (shape[0] * 7 // 10, shape[1] * 2 // 3, shape[2] // 2)
    r = min(shape) / 4

    def ball(c, rad):
        return ((i - c[0]) ** 2 + (j - c[1]) ** 2 + (k - c[2]) ** 2) <= rad * rad

    rng = np.random.default_rng(seed)
    u0 = 0.2 + 0.6 * ball(c1, r) + 0.5 * ball(c2, 0.75 * r) + 0.05 * rng.standard_normal(shape)
    u0 = (u0 - u0.min()) / ((u0.max() - u0.min()) / 255.0)
    phi0 = (ball(c1, 1.25 * r) | ball(c2, r)).astype(np.float64)
    return u0, phi0, c1
