"""Plain-data body descriptors, the POD form of a geometry that crosses the C ABI.

The reference takes any SDFLibrary geometry or callable ``x -> phi`` (``sdf_value``,
``src/marching_squares/cache.jl:69``) and, for PLIC, ``get_sdf`` at a cell centroid
(``src/plic/plic.jl:74``).  A closure cannot cross a C ABI, so a body here is one of a fixed set of
descriptors (``kind`` + up to 24 parameters + an optional primitive table).  An arbitrary field is
still supported by sampling it into ``phi`` yourself and calling ``update_cache_from_phi``.

SDFLibrary is not vendored with the reference; these are the build's own signed-distance functions
(negative inside, ``||.||`` accumulated as ``((dx^2 + dy^2) + dz^2)``):

* ``SDFSphere`` / ``SDFCircle``: ``||x - c|| - r``; normal ``(x - c)/||x - c||``.
* ``SDFTorus``: in the frame ``p = R^T (x - c)``, ``q = (sqrt(px^2 + py^2) - R, pz)``,
  ``phi = ||q|| - r``; normal is the analytic gradient rotated back.
* ``SDFBox``: the exact box distance in its rotated frame; normal from ``sign(p)`` (zero on a
  medial plane, which PLIC reports as a degenerate fit).
* ``SDFPlane``: ``n . (x - p)``.
* ``SDFSphereUnion``: spheres folded left with the polynomial smooth minimum
  ``h = max(k - |a-b|, 0)/k; min(a, b) - h*h*k/4``.
* ``SDFUnion(children...; smoothing)``: spheres, tori, boxes and planes folded left the same way,
  ``smoothing = 0`` being the plain minimum (``test/test_plic.jl:96-112`` unions a circle and a
  rectangle so).  Its normal is the fold's gradient: the winning child's own normal wherever the fold
  took a plain minimum, ``(1 - h/2) n_min + (h/2) n_max`` where it blended, scaled to unit length.
* ``SDFRectangle(center, dims)``: a 2D box of side lengths ``dims`` (SDFLibrary's own reading of
  ``dims`` is not available; here they are side lengths, not half-widths).
* ``EmptyGeo``: the constant field ``value`` (the reference seeds every cache with ``EmptyGeo(1)``,
  ``src/marching_squares/cache.jl:46``).
* ``SDFBounded(child, pad_distance, fill_distance)``: the child within ``pad_distance`` (Euclidean)
  of a box holding it, ``fill_distance`` (positive) with no normal beyond -- SDFLibrary's
  ``SDFBounded`` as the reference uses it (``src/marching_cubes/surface.jl:70-71``,
  ``src/volume.jl:26-27``, ``test/test_bounded.jl``).
"""
from dataclasses import dataclass, field

import numpy as np

GEO_EMPTY, GEO_SPHERE, GEO_TORUS, GEO_BOX, GEO_PLANE, GEO_SPHERE_UNION, GEO_UNION = range(7)
UNION_RECORD = 16  # doubles per union child: its kind, then its own parameters
GEO_BOUNDED = 0x100  # a flag on a bounded body's kind

_EYE = (1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0)


def _vec3(v):
    v = [float(x) for x in v]
    return v + [0.0] * (3 - len(v))


def _rot(r):
    if r is None:
        return list(_EYE)
    r = np.asarray(r, dtype=np.float64).reshape(-1)
    if r.size == 4:  # a 2x2 rotation, embedded
        return [r[0], r[1], 0.0, r[2], r[3], 0.0, 0.0, 0.0, 1.0]
    return [float(x) for x in r]


@dataclass
class EmptyGeo:
    value: float = 1.0
    kind = GEO_EMPTY
    prims = None

    @property
    def params(self):
        return [float(self.value)]


@dataclass
class SDFSphere:
    center: tuple
    radius: float
    kind = GEO_SPHERE
    prims = None

    @property
    def params(self):
        return _vec3(self.center) + [float(self.radius)]


SDFCircle = SDFSphere  # the same descriptor in 2D (the third centre coordinate is ignored)


@dataclass
class SDFTorus:
    center: tuple
    major_radius: float
    minor_radius: float
    rotation: object = None  # 3x3, row-major; the torus axis is R @ e_z
    kind = GEO_TORUS
    prims = None

    @property
    def params(self):
        return (_vec3(self.center) + [float(self.major_radius), float(self.minor_radius)] +
                _rot(self.rotation))


@dataclass
class SDFBox:
    center: tuple
    half_extents: tuple
    rotation: object = None
    kind = GEO_BOX
    prims = None

    @property
    def params(self):
        return _vec3(self.center) + _vec3(self.half_extents) + _rot(self.rotation)


@dataclass
class SDFPlane:
    normal: tuple
    point: tuple
    kind = GEO_PLANE
    prims = None

    @property
    def params(self):
        return _vec3(self.normal) + _vec3(self.point)


@dataclass
class SDFSphereUnion:
    centers: np.ndarray
    radii: np.ndarray
    smoothing: float = 0.02
    kind = GEO_SPHERE_UNION
    _prims: np.ndarray = field(default=None, repr=False)

    def __post_init__(self):
        c = np.asarray(self.centers, dtype=np.float64).reshape(-1, 3)
        r = np.asarray(self.radii, dtype=np.float64).reshape(-1, 1)
        self._prims = np.ascontiguousarray(np.hstack([c, r]))

    @property
    def prims(self):
        return self._prims

    @property
    def params(self):
        return [float(self.smoothing)]


def SDFRectangle(center, dims, rotation=None):
    """A 2D box of side lengths ``dims`` centred at ``center`` (``test/test_plic.jl:101``)."""
    return SDFBox(tuple(center), tuple(0.5 * float(v) for v in dims), rotation)


class SDFUnion:
    """``SDFUnion(children...; smoothing)``: the children (spheres, tori, boxes, planes) folded left
    with the polynomial smooth minimum, ``smoothing = 0`` the plain minimum."""
    kind = GEO_UNION

    def __init__(self, *children, smoothing=0.0):
        if not children:
            raise ValueError("a union needs at least one child")
        for c in children:
            if int(c.kind) not in (GEO_SPHERE, GEO_TORUS, GEO_BOX, GEO_PLANE):
                raise ValueError("a union's children are spheres, tori, boxes or planes")
        if not float(smoothing) >= 0.0:
            raise ValueError("smoothing must be >= 0")
        self.children = tuple(children)
        self.smoothing = float(smoothing)
        rec = np.zeros((len(children), UNION_RECORD))
        for i, c in enumerate(children):
            p = [float(v) for v in c.params]
            rec[i, 0] = int(c.kind)
            rec[i, 1:1 + len(p)] = p
        self._prims = rec

    def __repr__(self):
        return f"SDFUnion({', '.join(map(repr, self.children))}; smoothing={self.smoothing})"

    @property
    def prims(self):
        return self._prims

    @property
    def params(self):
        return [self.smoothing]


@dataclass
class SDFBounded:
    """SDFLibrary's ``SDFBounded(child; pad_distance, fill_distance)``: within ``pad_distance`` of a
    box holding the child, the child's own value; beyond it ``fill_distance``, which must be positive so
    those nodes and centroids read as outside (``src/marching_cubes/surface.jl:70-71``,
    ``src/volume.jl:26-27``), with no normal.  A pad wider than a cell diagonal leaves every cut cell,
    and so the measured volume and the contour, bit for bit alone (``test/test_bounded.jl``).

    SDFLibrary is not vendored: the box is the one ``bounds()`` gives (a sphere's ``c +- r``, a
    rotated box's world-axis extents, a torus's ``c +- (R + r)``, a sphere union's spheres widened by
    its smoothing) unless ``box = (lo, hi)`` is given, the band is the Euclidean distance to the box,
    and ``fill_distance`` defaults to ``pad_distance`` (a lower bound of the true distance beyond the
    band)."""
    child: object
    pad_distance: float
    fill_distance: float = None
    box: tuple = None

    def __post_init__(self):
        if int(self.child.kind) & GEO_BOUNDED or int(self.child.kind) in (GEO_EMPTY, GEO_PLANE):
            raise ValueError("SDFBounded needs a bounded child: a sphere, torus, box or union")
        if len(list(self.child.params)) > 16:
            raise ValueError("the child's parameters do not fit a bounded descriptor")
        if not float(self.pad_distance) >= 0.0:
            raise ValueError("pad_distance must be >= 0")
        if not self.fill() > 0.0:
            raise ValueError("fill_distance must be positive")

    @property
    def kind(self):
        return int(self.child.kind) | GEO_BOUNDED

    @property
    def prims(self):
        return self.child.prims

    def fill(self):
        if self.fill_distance is not None:
            return float(self.fill_distance)
        return float(self.pad_distance) if float(self.pad_distance) > 0.0 else 1.0

    def bounds(self):
        """(lo, hi), each three floats, of the box the band is measured from."""
        if self.box is not None:
            return _vec3(self.box[0]), _vec3(self.box[1])
        g = self.child
        if int(g.kind) == GEO_SPHERE:
            c, r = np.array(_vec3(g.center)), float(g.radius)
            return list(c - r), list(c + r)
        if int(g.kind) == GEO_TORUS:
            c, e = np.array(_vec3(g.center)), float(g.major_radius) + float(g.minor_radius)
            return list(c - e), list(c + e)
        if int(g.kind) == GEO_BOX:
            c = np.array(_vec3(g.center))
            R = np.array(_rot(g.rotation)).reshape(3, 3)
            e = np.abs(R) @ np.array(_vec3(g.half_extents))  # world = c + R p
            return list(c - e), list(c + e)
        if int(g.kind) == GEO_UNION:
            if any(int(c.kind) == GEO_PLANE for c in g.children):
                raise ValueError("a union with a plane has no bounding box: give box = (lo, hi)")
            b = [SDFBounded(c, 0.0).bounds() for c in g.children]
            lo = np.min([v[0] for v in b], axis=0) - g.smoothing
            hi = np.max([v[1] for v in b], axis=0) + g.smoothing
            return list(lo), list(hi)
        p = np.asarray(g.prims)
        k = float(g.smoothing)
        return list((p[:, :3] - p[:, 3:]).min(axis=0) - k), list((p[:, :3] + p[:, 3:]).max(axis=0) + k)

    @property
    def params(self):
        p = [float(v) for v in self.child.params]
        lo, hi = self.bounds()
        return p + [0.0] * (16 - len(p)) + lo + hi + [float(self.pad_distance), self.fill()]


def random_rotation(rng):
    """A uniformly random rotation matrix from a unit quaternion drawn by ``rng`` (numpy)."""
    q = rng.standard_normal(4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
