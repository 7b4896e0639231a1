"""Seeded, portable particle-set generators for the five workloads.

The reference ships no generator, seed or fixture (its example hard-codes a
2x2x2 lattice, magnetic_solver/main.cu:5-41).  These generators define the
workloads exactly so every run, test and fixture sees the same inputs:

* RNG: SplitMix64; the k-th draw of a stream seeded with s is
  mix(s + (k+1) * 0x9E3779B97F4A7C15); U[0,1) = (z >> 11) * 2**-53.
* Lattice order: x index outermost, z innermost, as main.cu:25-34.
* Jitter: three uniforms per site, drawn in site order (x, y, z).
* Volume per particle V = spacing**3 = 1; no external dipoles (uniform H).

Workloads (names used by bench.py and the tests):
  lattice8k      N = 8192 jittered 16x16x32 lattice, h=2, chi=1, H=(0,0,1)
  cylinder262k   cylinder radius 32, layers filled in z order until exactly
                 262144 sites, h=2, chi=1, H=(0,0,1)
  pool1m         256x256x16 lattice + 104857 uniform points in
                 [0,256)x[0,256)x[16,32)  (N = 1153433), h=2, chi=1
  pool8m         512x512x32 lattice (N = 8388608), h=2, chi=2
  droplets96     96 jittered-lattice spheres, N_i log-uniform in
                 [2048, 262144] (seed 7), chi_i ~ U[0.5, 3], H_i random unit
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, count: int, offset: int = 0) -> np.ndarray:
    """Draws offset .. offset+count-1 of the SplitMix64 stream seeded with seed."""
    k = np.arange(offset + 1, offset + count + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + k * GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed: int, count: int, offset: int = 0) -> np.ndarray:
    z = splitmix64(seed, count, offset)
    return (z >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


@dataclass
class ParticleSystem:
    """One magnetizable particle set with its solve parameters."""
    name: str
    pos: np.ndarray            # (n, 3) float32
    h: float
    chi: float
    volume: float = 1.0
    hconst: np.ndarray = field(default_factory=lambda: np.array([0.0, 0.0, 1.0]))

    @property
    def n(self) -> int:
        return int(self.pos.shape[0])

    @property
    def kappa(self) -> float:
        """m = kappa * x with x = B/mu0 (magnetic_force.cu:1016)."""
        return self.volume * self.chi / (1.0 + self.chi)


def jittered_lattice(nx: int, ny: int, nz: int, spacing: float = 1.0, jitter: float = 0.1,
                     seed: int = 0, origin=(0.0, 0.0, 0.0)) -> np.ndarray:
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    sites = np.stack([i.ravel(), j.ravel(), k.ravel()], axis=1).astype(np.float64) * spacing
    u = uniform(seed, 3 * sites.shape[0]).reshape(-1, 3)
    sites += (2.0 * u - 1.0) * jitter
    sites += np.asarray(origin, dtype=np.float64)
    return sites.astype(np.float32)


def cylinder_sites(radius: int, count: int) -> np.ndarray:
    """Lattice points (i, j, k), i^2 + j^2 <= radius^2, filled layer by layer
    (k = 0, 1, ...), i outer / j inner within a layer, until `count` sites."""
    r = np.arange(-radius, radius + 1)
    I, J = np.meshgrid(r, r, indexing="ij")
    inside = (I * I + J * J) <= radius * radius
    layer = np.stack([I[inside], J[inside]], axis=1)
    nl = -(-count // layer.shape[0])
    out = np.concatenate([np.column_stack([layer, np.full(layer.shape[0], k)]) for k in range(nl)])
    return out[:count].astype(np.float64)


def lattice8k() -> ParticleSystem:
    return ParticleSystem("lattice8k", jittered_lattice(16, 16, 32, seed=0), h=2.0, chi=1.0)


def cylinder262k(count: int = 262144) -> ParticleSystem:
    sites = cylinder_sites(32, count)
    u = uniform(0, 3 * sites.shape[0]).reshape(-1, 3)
    pos = (sites + (2.0 * u - 1.0) * 0.1).astype(np.float32)
    return ParticleSystem("cylinder262k", pos, h=2.0, chi=1.0)


def pool1m() -> ParticleSystem:
    lat = jittered_lattice(256, 256, 16, seed=0).astype(np.float64)
    nr = 104857
    u = uniform(1, 3 * nr).reshape(-1, 3)
    rnd = np.column_stack([u[:, 0] * 256.0, u[:, 1] * 256.0, 16.0 + u[:, 2] * 16.0])
    return ParticleSystem("pool1m", np.concatenate([lat, rnd]).astype(np.float32), h=2.0, chi=1.0)


def pool8m() -> ParticleSystem:
    return ParticleSystem("pool8m", jittered_lattice(512, 512, 32, seed=0), h=2.0, chi=2.0)


def droplets96(count: int = 96, nmin: int = 2048, nmax: int = 262144, seed: int = 7):
    """96 independent spheres; returns a list of ParticleSystem."""
    u = uniform(seed, 6 * count).reshape(count, 6)
    systems = []
    for s in range(count):
        n_i = int(round(np.exp(np.log(nmin) + u[s, 0] * (np.log(nmax) - np.log(nmin)))))
        chi = 0.5 + 2.5 * u[s, 1]
        # direction: Box-Muller normals from the next draws, normalised
        a = np.sqrt(-2.0 * np.log(max(u[s, 2], 1e-300)))
        g = np.array([a * np.cos(2 * np.pi * u[s, 3]), a * np.sin(2 * np.pi * u[s, 3]),
                      np.sqrt(-2.0 * np.log(max(u[s, 4], 1e-300))) * np.cos(2 * np.pi * u[s, 5])])
        hdir = g / np.linalg.norm(g)
        radius = (3.0 * n_i / (4.0 * np.pi)) ** (1.0 / 3.0)
        R = int(np.ceil(radius)) + 1
        r = np.arange(-R, R + 1)
        I, J, K = np.meshgrid(r, r, r, indexing="ij")
        sites = np.stack([I.ravel(), J.ravel(), K.ravel()], axis=1).astype(np.float64)
        sites = sites[(sites ** 2).sum(1) <= radius * radius]
        ju = uniform(1000 + s, 3 * sites.shape[0]).reshape(-1, 3)
        pos = sites + (2.0 * ju - 1.0) * 0.1
        systems.append(ParticleSystem(f"droplet{s}", pos.astype(np.float32), h=2.0, chi=chi,
                                      hconst=hdir))
    return systems


WORKLOADS = {
    "lattice8k": lattice8k,
    "cylinder262k": cylinder262k,
    "pool1m": pool1m,
    "pool8m": pool8m,
}
