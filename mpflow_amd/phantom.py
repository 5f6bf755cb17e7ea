"""Seeded synthetic conductivity phantoms and measurement noise.

Random numbers come from SplitMix64 (Steele, Lea & Flood, "Fast splittable
pseudorandom number generators", OOPSLA 2014), so every implementation that is given
the same seed draws the same phantom.  Seeds follow ``1000003 * config + frame``.

Conductivity is assigned per element by its centroid.
"""
from __future__ import annotations

import math

import numpy as np

__all__ = ["SplitMix64", "frame_seed", "inclusion_phantom", "ellipse_phantom",
           "element_centroids", "add_noise"]

_M64 = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed: int):
        self.state = seed & _M64

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & _M64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
        return z ^ (z >> 31)

    def uniform(self) -> float:
        """Uniform double in [0, 1) from the top 53 bits."""
        return (self.next_u64() >> 11) * (1.0 / (1 << 53))

    def normal(self) -> float:
        """Standard normal by Box-Muller (one value per call)."""
        u1 = self.uniform()
        u2 = self.uniform()
        return math.sqrt(-2.0 * math.log(1.0 - u1)) * math.cos(2.0 * math.pi * u2)


def frame_seed(config: int, frame: int, offset: int = 0) -> int:
    return 1000003 * config + frame + offset


def element_centroids(nodes: np.ndarray, elements: np.ndarray) -> np.ndarray:
    return nodes[elements[:, :3]].mean(axis=1)


def inclusion_phantom(nodes, elements, seed, radius=0.2, sigma=2.0, background=1.0,
                      max_centre=0.6):
    """One circular inclusion; centre radius ``max_centre*sqrt(u1)``, angle ``2 pi u2``."""
    g = SplitMix64(seed)
    rc = max_centre * math.sqrt(g.uniform())
    th = 2.0 * math.pi * g.uniform()
    c = np.array([rc * math.cos(th), rc * math.sin(th)])
    cen = element_centroids(nodes, elements)
    gamma = np.full(elements.shape[0], background, dtype=np.float64)
    gamma[np.sum((cen - c) ** 2, axis=1) <= radius * radius] = sigma
    return gamma


def ellipse_phantom(nodes, elements, seed, n_min=1, n_max=3, background=1.0,
                    axis_range=(0.1, 0.3), sigma_range=(0.2, 5.0), max_centre=0.7):
    """1..n ellipses: semi-axes U(0.1,0.3), rotation U(0,pi), centre uniform in the disc
    of radius 0.7, sigma log-uniform in [0.2, 5]; later ellipses overwrite earlier ones."""
    g = SplitMix64(seed)
    count = n_min + int(g.uniform() * (n_max - n_min + 1))
    cen = element_centroids(nodes, elements)
    gamma = np.full(elements.shape[0], background, dtype=np.float64)
    for _ in range(count):
        a = axis_range[0] + (axis_range[1] - axis_range[0]) * g.uniform()
        b = axis_range[0] + (axis_range[1] - axis_range[0]) * g.uniform()
        rot = math.pi * g.uniform()
        rc = max_centre * math.sqrt(g.uniform())
        th = 2.0 * math.pi * g.uniform()
        s = math.exp(math.log(sigma_range[0]) +
                     (math.log(sigma_range[1]) - math.log(sigma_range[0])) * g.uniform())
        d = cen - np.array([rc * math.cos(th), rc * math.sin(th)])
        u = d[:, 0] * math.cos(rot) + d[:, 1] * math.sin(rot)
        v = -d[:, 0] * math.sin(rot) + d[:, 1] * math.cos(rot)
        gamma[(u / a) ** 2 + (v / b) ** 2 <= 1.0] = s
    return gamma


def add_noise(U: np.ndarray, snr_db: float, seed: int) -> np.ndarray:
    """U + N(0, s^2) with s = rms(U) * 10^(-SNR/20)."""
    g = SplitMix64(seed ^ 0x5DEECE66D)
    s = math.sqrt(float(np.mean(np.asarray(U, dtype=np.float64) ** 2))) * 10.0 ** (-snr_db / 20.0)
    noise = np.array([g.normal() for _ in range(U.size)]).reshape(U.shape)
    return U + s * noise
