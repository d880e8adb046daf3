"""Flow diagnostics on downloaded particle arrays (host side, numpy/scipy).

`front_position` is the dam-break surge front: the largest x over free particles that have at
least half the bulk neighbour count within the kernel support, which excludes isolated spray.
"""
import numpy as np
from scipy.spatial import cKDTree


def bulk_neighbours(dim, h_factor=1.3, cellfac=2.0):
    """Lattice neighbour count within rij <= Cellfac h for a cubic lattice with h = h_factor dx."""
    R = cellfac*h_factor
    n = int(np.ceil(R))
    r = np.arange(-n, n + 1)
    if dim == 2:
        X, Y = np.meshgrid(r, r)
        d2 = X*X + Y*Y
    else:
        X, Y, Z = np.meshgrid(r, r, r)
        d2 = X*X + Y*Y + Z*Z
    return int(((d2 <= R*R) & (d2 > 0)).sum())


def front_position(p, dim, h=None, cellfac=2.0, fraction=0.5, h_factor=1.3):
    """x of the surge front: max x over free particles with >= fraction * bulk neighbours."""
    free = p["IsFree"] == 1
    x = p["x"][free][:, :dim]
    if h is None:
        h = float(p["h"][free][0])
    tree = cKDTree(x)
    counts = tree.query_ball_point(x, r=cellfac*h, return_length=True) - 1
    need = fraction*bulk_neighbours(dim, h_factor, cellfac)
    ok = counts >= need
    if not ok.any():
        return float("nan")
    return float(x[ok, 0].max())


def relative_error(a, b):
    """||a - b|| / ||b|| over all entries."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b)/nb) if nb > 0 else float(np.linalg.norm(a - b))
