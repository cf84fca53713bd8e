"""Host readers over a cache's columns -- ``src/cut_cell.jl:155-347``.

Each reader works on whole numpy columns (vectorised over cells) with the reference's per-cell
expression order, so a value read back is bit-identical to one formed cell by cell:

* ``full_face_area`` -- ``src/cut_cell.jl:243-253``: in 2D the other axis' size; in 3D the product of
  the other two axes in increasing order, starting from one.
* ``face_area_of`` -- ``:276-277``: ``face_fraction * full_face_area``.
* ``interface_normal_area`` -- ``:224-226``: per axis ``A(2a) - A(2a-1)``.
* ``closure_residual`` -- ``:340-347``: face vector areas summed in direction order, minus the
  interface vector area.  Exactly zero by construction.
* ``face_centroid`` -- ``:155-165``: stored in-plane offsets added to the cell's lower lattice corner.
* ``cut_cell_report`` -- ``:302-319``.
"""
import numpy as np

CELL_INSIDE, CELL_OUTSIDE, CELL_CUT = 0, 1, 2


def full_face_area(cellsize, direction, dtype=np.float64):
    """0-based ``direction`` (0 = -x, 1 = +x, 2 = -y, ...), in the element type ``dtype``."""
    T = np.dtype(dtype).type
    cs = [T(v) for v in cellsize]
    ax = direction >> 1
    if len(cs) == 2:
        return cs[1 - ax]
    p = T(1)
    for a in range(len(cs)):
        if a != ax:
            p = T(p * cs[a])
    return p


def face_area_of(face_fraction, direction, cellsize):
    ff = np.asarray(face_fraction)
    return ff[..., direction] * full_face_area(cellsize, direction, ff.dtype)


def interface_normal_area(face_fraction, cellsize):
    D = len(cellsize)
    return np.stack([face_area_of(face_fraction, 2 * a + 1, cellsize) -
                     face_area_of(face_fraction, 2 * a, cellsize) for a in range(D)], axis=-1)


def face_vector_area(face_fraction, direction, cellsize):
    D = len(cellsize)
    T = np.asarray(face_fraction).dtype.type
    s = T(1) if direction & 1 else T(-1)
    out = np.zeros(face_fraction.shape[:-1] + (D,), dtype=T)
    out[..., direction >> 1] = s * face_area_of(face_fraction, direction, cellsize)
    return out


def closure_residual(face_fraction, cellsize):
    D = len(cellsize)
    acc = np.zeros(face_fraction.shape[:-1] + (D,), dtype=np.asarray(face_fraction).dtype)
    for k in range(2 * D):
        acc = acc + face_vector_area(face_fraction, k, cellsize)
    return acc - interface_normal_area(face_fraction, cellsize)


def interface_area(face_fraction, cellsize):
    v = interface_normal_area(face_fraction, cellsize)
    s = v[..., 0] * v[..., 0]
    for a in range(1, v.shape[-1]):
        s = s + v[..., a] * v[..., a]
    return np.sqrt(s)


def interface_normal(face_fraction, cellsize):
    v = interface_normal_area(face_fraction, cellsize)
    a = interface_area(face_fraction, cellsize)
    with np.errstate(invalid="ignore", divide="ignore"):
        out = np.where(a[..., None] > 0, v / np.where(a > 0, a, 1.0)[..., None], 0.0)
    return out


def face_centroid(fcl, direction, lower, upper):
    """Global centroid of face ``direction`` from the stored offsets ``fcl`` (``[..., 2D, D-1]``),
    the cell's lower lattice corner ``lower`` and its upper lattice corner ``upper``."""
    D = lower.shape[-1]
    ax = direction >> 1
    out = np.empty(lower.shape)
    for a in range(D):
        if a == ax:
            out[..., a] = upper[..., a] if direction & 1 else lower[..., a]
        else:
            out[..., a] = lower[..., a] + fcl[..., direction, a if a < ax else a - 1]
    return out


def cut_cell_report(kind, ambiguous, volume_fraction):
    kind = np.asarray(kind)
    vf = np.asarray(volume_fraction)
    cutmask = (kind == CELL_CUT) & (vf > 0)
    alpha_min = min(1.0, float(vf[cutmask].min())) if cutmask.any() else 1.0
    return dict(n=int(kind.size), cut=int((kind == CELL_CUT).sum()),
                inside=int((kind == CELL_INSIDE).sum()), outside=int((kind == CELL_OUTSIDE).sum()),
                ambiguous=int(np.count_nonzero(ambiguous)), min_volume_fraction=alpha_min)
