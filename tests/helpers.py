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
