"""Shared test helpers: nodal fields from callables, single-cell fixtures, shared-face checks."""
import numpy as np

MC_BITS = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                    [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]])
MS_BITS = np.array([[0, 0], [1, 0], [1, 1], [0, 1]])
MC_EDGE = [(0, 1), (1, 2), (3, 2), (0, 3), (4, 5), (5, 6), (7, 6), (4, 7),
           (0, 4), (1, 5), (2, 6), (3, 7)]


def node_coords(x0, d, n):
    """Node coordinates of a grid, shape (nnodes, D), column-major (i fastest)."""
    D = len(n)
    axes = [x0[a] + np.arange(n[a] + 1, dtype=np.float64) * d[a] for a in range(D)]
    mesh = np.meshgrid(*axes, indexing="ij")
    return np.stack([m.reshape(-1, order="F") for m in mesh], axis=1)


def sample(f, x0, d, n):
    """phi at every node of the grid from a Python callable ``f(x) -> float``."""
    X = node_coords(x0, d, n)
    return np.array([float(f(x)) for x in X])


def sample_vec(f, x0, d, n):
    """phi at every node from a vectorised callable ``f(X[:, 0], X[:, 1], ...)``."""
    X = node_coords(x0, d, n)
    return np.asarray(f(*[X[:, a] for a in range(X.shape[1])]), dtype=np.float64)


def cube_nodes(origin=(0.0, 0.0, 0.0), h=(1.0, 1.0, 1.0)):
    return np.array([[origin[a] + b[a] * h[a] for a in range(3)] for b in MC_BITS], dtype=np.float64)


def shared_face_mismatches(col, n, D):
    """How many interior faces of a per-cell face column (ncell, 2D, ...) disagree bitwise between
    the two cells sharing them: cell ci's slot 2c+1 vs cell ci+e_c's slot 2c (0-based)."""
    A = fortran(col, n)
    bad = 0
    for c in range(D):
        lo = [slice(None)] * D
        hi = [slice(None)] * D
        lo[c] = slice(0, n[c] - 1)
        hi[c] = slice(1, n[c])
        a = A[tuple(lo) + (2 * c + 1,)]
        b = A[tuple(hi) + (2 * c,)]
        same = (a.view(np.uint64) == b.view(np.uint64)) if a.dtype == np.float64 else (a == b)
        if same.ndim > D:
            same = same.all(axis=tuple(range(D, same.ndim)))
        bad += int((~same).sum())
    return bad


def fortran(col, n):
    """A flat per-cell column (linear index i + n0*(j + n1*k)) viewed as an array [i, j, (k), ...]."""
    D = len(n)
    A = col.reshape(tuple(reversed(n)) + col.shape[1:])
    return np.transpose(A, tuple(reversed(range(D))) + tuple(range(D, A.ndim)))


# ---- marched surfaces --------------------------------------------------------------------------
def cell_indices_ref(nodes, tris, x0, d, n):
    """cell_indices (src/marching_cubes/surface.jl:128-141): the 1-based linear index of the cell
    each triangle's centroid floors into, clamped; the centroid in the nodes' own precision, the
    rest in Float64 as the reference converts the grid."""
    a, b, c = (nodes[tris[:, k] - 1] for k in range(3))
    cen = ((a + b) + c) / nodes.dtype.type(3)
    idx = []
    for ax in range(3):
        f = np.floor((cen[:, ax].astype(np.float64) - x0[ax]) / d[ax]) + 1.0
        idx.append(np.clip(f, 1.0, float(n[ax])).astype(np.int64))
    return idx[0] + n[0] * ((idx[1] - 1) + n[1] * (idx[2] - 1))


def mesh_topology(tris):
    """(histogram of undirected edge uses, largest directed edge multiplicity, V - E + F)."""
    e = np.concatenate([tris[:, [0, 1]], tris[:, [1, 2]], tris[:, [2, 0]]])
    _, und = np.unique(np.sort(e, axis=1), axis=0, return_counts=True)
    _, dire = np.unique(e, axis=0, return_counts=True)
    V = len(np.unique(tris))
    return np.bincount(und), int(dire.max()) if len(dire) else 0, V - len(und) + len(tris)


def tri_vector_areas(nodes, tris):
    a, b, c = (nodes[tris[:, k] - 1].astype(np.float64) for k in range(3))
    return 0.5 * np.cross(b - a, c - a), a, b, c


def surface_moments(nodes, tris, face_fraction, origin, h, face_area_of):
    """What the triangles marched through one cell imply for its moments, closed against the cell's
    own open faces (test/test_mc_moments.jl:107-143): vector area, interface centroid, outside
    volume fraction and centroid, in the cell's coordinates."""
    o = np.asarray(origin, dtype=np.float64)
    ff = np.asarray(face_fraction)[None, :]
    vol3 = 0.0
    cm2 = np.zeros(3)
    for ax in range(3):
        a_open = float(face_area_of(ff, 2 * ax + 1, h)[0])
        vol3 += h[ax] * a_open
        cm2[ax] += h[ax] ** 2 * a_open
    S = np.zeros(3)
    A = 0.0
    Ac = np.zeros(3)
    for t in tris:
        a, b, c = (nodes[t[k] - 1] - o for k in range(3))
        w = 0.5 * np.cross(b - a, c - a)
        xc = (a + b + c) / 3
        S += w
        ln = np.linalg.norm(w)
        A += ln
        Ac += ln * xc
        vol3 -= w @ xc
        cm2 -= w * ((a * a + b * b + c * c + a * b + b * c + c * a) / 6)
    V = vol3 / 3
    return dict(vector_area=S, interface_centroid=o + Ac / A if A > 0 else None,
                volume_fraction=V / np.prod(h), centroid=o + cm2 / (2 * V) if V != 0 else None)


# ---- quadtrees ---------------------------------------------------------------------------------
GEO_CENTRE, GEO_RADIUS = (0.13, -0.07), 0.9        # test/test_moments.jl:30-32
CACHE_CENTRE, CACHE_RADIUS = (0.013, -0.021), 0.83  # test/test_cache.jl:22


def circle_sdf(x, c, r):
    dx, dy = x[..., 0] - c[0], x[..., 1] - c[1]
    return np.sqrt(dx * dx + dy * dy) - r


def banded_tree(level=4, width=2):
    """test/test_moments.jl:72-95: base [-2, 2]^2 of 8 x 8, initial level 1, a uniformly refined
    band round the circle so every cut leaf sits at the finest level, with genuine level jumps."""
    from cutcell import AdaptiveMesh, CartesianGrid
    base = CartesianGrid((-2.0, -2.0), (8, 8), (0.5, 0.5))
    mesh = AdaptiveMesh(base, max_level=5, initial_level=1)
    for _ in range(level):
        h = mesh.elem_size()
        diag = np.sqrt((h * h).sum(axis=1))
        d = np.abs(circle_sdf(mesh.elem_centroid(), GEO_CENTRE, GEO_RADIUS))
        flags = (mesh.level < level) & (d <= (width + 0.5) * diag)
        if not flags.any():
            break
        mesh.refine(flags)
    return mesh


def cache_tree():
    """test/test_cache.jl:116-124: base [-1.5, 1.5]^2 of 8 x 8, initial level 1, refined twice near
    the circle, so cut leaves come in several sizes."""
    from cutcell import AdaptiveMesh, CartesianGrid
    base = CartesianGrid((-1.5, -1.5), (8, 8), (3.0 / 8, 3.0 / 8))
    tree = AdaptiveMesh(base, max_level=3, initial_level=1)
    for _ in range(2):
        near = np.abs(circle_sdf(tree.elem_centroid(), CACHE_CENTRE, CACHE_RADIUS)) < \
            1.5 * tree.elem_size().max(axis=1)
        tree.refine(near)
    return tree


def same_level_neighbours(mesh):
    """Pairs (a, b, axis) of leaves at one level sharing a face, b on a's + side."""
    at = {(int(l), int(x), int(y)): i for i, (l, (x, y)) in enumerate(zip(mesh.level, mesh.coord))}
    out = []
    for i, (l, (x, y)) in enumerate(zip(mesh.level, mesh.coord)):
        for ax, key in ((0, (int(l), int(x) + 1, int(y))), (1, (int(l), int(x), int(y) + 1))):
            j = at.get(key)
            if j is not None:
                out.append((i, j, ax))
    return out


class DeviceArray:
    """A device copy of a host array, made through the HIP runtime the library itself links (so its
    pointer is valid in the library's kernels).  For tests of the entry points that take device
    pointers."""

    def __init__(self, arr):
        import ctypes
        import cutcell._lib as L
        L.lib()  # the library, and with it its HIP runtime, loaded first
        self._hip = hip = ctypes.CDLL("libamdhip64.so.7")
        hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        hip.hipFree.argtypes = [ctypes.c_void_p]
        a = np.ascontiguousarray(arr)
        self.ptr = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.ptr), max(a.nbytes, 8)) == 0
        assert hip.hipMemcpy(self.ptr, ctypes.c_void_p(a.ctypes.data), a.nbytes, 1) == 0  # host to device
        assert hip.hipDeviceSynchronize() == 0

    def free(self):
        if self.ptr is not None and self.ptr.value:
            self._hip.hipFree(self.ptr)
        self.ptr = None

    def __del__(self):
        self.free()
