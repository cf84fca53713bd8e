"""Seeded synthetic workloads: the moving-bodies batch (many caches on uneven grids, rigid motion)
and the smooth-union CSG body (hundreds of spheres)."""
import math

import numpy as np

from .geometry import SDFBox, SDFSphere, SDFSphereUnion, SDFTorus, random_rotation
from .grid import CartesianGrid


def _rodrigues(axis, angle):
    axis = np.asarray(axis, dtype=np.float64)
    axis = axis / np.linalg.norm(axis)
    K = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
    return np.eye(3) + math.sin(angle) * K + (1 - math.cos(angle)) * (K @ K)


class MovingBodies:
    """``nbodies`` independent bodies, each on its own isotropic grid with per-axis cell counts drawn
    log-uniformly in [nmin, nmax] (cell size 2/max(n), the grid centred on the origin), a sphere,
    torus or box of size 0.2-0.6 of the domain, a method (marching cubes or PLIC) drawn per body, and
    a rigid motion: a rotation of the body frame about a random axis and a translation, both per
    step (the translation applied as a grid shift, as the reference moves bodies)."""

    def __init__(self, seed, nbodies=64, nmin=64, nmax=256, plic_fraction=0.5):
        rng = np.random.default_rng(seed)
        self.bodies = []
        for b in range(nbodies):
            n = tuple(int(round(math.exp(rng.uniform(math.log(nmin), math.log(nmax))))) for _ in range(3))
            d = 2.0 / max(n)
            x0 = tuple(-0.5 * na * d for na in n)
            size = rng.uniform(0.2, 0.6) * 2.0
            shape = ("sphere", "torus", "box")[rng.integers(3)]
            method = "plic" if rng.uniform() < plic_fraction else "mc"
            R0 = random_rotation(rng)
            axis = rng.standard_normal(3)
            omega = rng.uniform(0.01, 0.05)
            vel = rng.uniform(-0.5, 0.5, 3) * d
            centre = tuple(rng.uniform(-0.1, 0.1, 3))
            half = rng.uniform(0.5, 1.0, 3) * size / 2
            self.bodies.append(dict(n=n, d=d, x0=x0, size=size, shape=shape, method=method, R0=R0,
                                    axis=axis, omega=omega, vel=vel, centre=centre, half=half))

    def grid(self, b, step):
        B = self.bodies[b]
        base = CartesianGrid(B["x0"], B["n"], (B["d"],) * 3)
        return base.moved(tuple(-step * v for v in B["vel"]))

    def geometry(self, b, step):
        B = self.bodies[b]
        R = _rodrigues(B["axis"], B["omega"] * step) @ B["R0"]
        s = B["size"]
        if B["shape"] == "sphere":
            return SDFSphere(B["centre"], s / 2)
        if B["shape"] == "torus":
            return SDFTorus(B["centre"], 0.35 * s, 0.15 * s, R)
        return SDFBox(B["centre"], tuple(B["half"]), R)

    def ncells(self):
        return sum(int(np.prod(B["n"])) for B in self.bodies)


def csg_spheres(seed, count=512, lo=-0.7, hi=0.7, rmin=0.05, rmax=0.15, smoothing=0.02):
    """The smooth union (polynomial smooth minimum, k = smoothing) of ``count`` spheres with
    centres uniform in [lo, hi]^3 and radii uniform in [rmin, rmax]."""
    rng = np.random.default_rng(seed)
    centres = rng.uniform(lo, hi, (count, 3))
    radii = rng.uniform(rmin, rmax, count)
    return SDFSphereUnion(centres, radii, smoothing)
