"""Seeded synthetic inputs for the solver: periodic Voronoi grains and
uniformly random crystal orientations as Kocks Euler angles (degrees).

The reference reads a per-element angle file (`orientation_input file`,
inmat.f) with Kocks angles in degrees and turns them into a rotation with
mm10_rotation_matrix (mm10_a.f:1287-1345).  Voxel e = ix*N*N + iy*N + iz
(x slowest, as the reference numbers its elements).
"""
import numpy as np


def random_rotations(rng, n):
    """n rotation matrices uniform on SO(3) (Shoemake's quaternion method)."""
    u = rng.random((n, 3))
    a, b = np.sqrt(1 - u[:, 0]), np.sqrt(u[:, 0])
    x = a * np.sin(2 * np.pi * u[:, 1])
    y = a * np.cos(2 * np.pi * u[:, 1])
    z = b * np.sin(2 * np.pi * u[:, 2])
    w = b * np.cos(2 * np.pi * u[:, 2])
    R = np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w),
                  2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w),
                  2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], 1)
    return R.reshape(n, 3, 3)


def kocks_matrix(angles_deg):
    """Rotation of Kocks angles (psi, theta, phi) in degrees, the matrix of
    mm10_rotation_matrix."""
    psi, th, phi = np.radians(np.asarray(angles_deg, float)).T
    s, c = np.sin, np.cos
    g = np.empty(np.shape(psi) + (3, 3))
    g[..., 0, 0] = -s(psi) * s(phi) - c(psi) * c(phi) * c(th)
    g[..., 0, 1] = c(psi) * s(phi) - s(psi) * c(phi) * c(th)
    g[..., 0, 2] = c(phi) * s(th)
    g[..., 1, 0] = s(psi) * c(phi) - c(psi) * s(phi) * c(th)
    g[..., 1, 1] = -c(psi) * c(phi) - s(psi) * s(phi) * c(th)
    g[..., 1, 2] = s(phi) * s(th)
    g[..., 2, 0] = c(psi) * s(th)
    g[..., 2, 1] = s(psi) * s(th)
    g[..., 2, 2] = c(th)
    return g


def kocks_angles(g):
    """Inverse of kocks_matrix (the extraction of mm10_update_euler_angles,
    mm10_a.f:1171-1220), degrees."""
    th = np.arccos(np.clip(g[..., 2, 2], -1.0, 1.0))
    psi = np.arctan2(g[..., 2, 1], g[..., 2, 0])
    phi = np.arctan2(g[..., 1, 2], g[..., 0, 2])
    return np.degrees(np.stack([psi, th, phi], -1))


def voronoi(rng, N, ngrains, method="auto"):
    """Grain id per voxel: nearest of ngrains seeds uniform in [0,1)^3 under
    the periodic (minimum image) metric, voxel centres ((i+1/2)/N, ...).
    method: "brute" (every seed), "kdtree" (periodic k-d tree) or "auto"
    (the tree for large grids)."""
    seeds = rng.random((ngrains, 3))
    c = (np.arange(N) + 0.5) / N
    if method == "kdtree" or (method == "auto" and N ** 3 * ngrains > 5e9):
        # large grids: periodic k-d tree (same nearest seed up to exact ties)
        from scipy.spatial import cKDTree
        tree = cKDTree(seeds, boxsize=1.0)
        gid = np.empty(N ** 3, np.int32)
        for ix in range(N):
            X = np.stack(np.meshgrid([c[ix]], c, c, indexing="ij"), -1).reshape(-1, 3)
            gid[ix * N * N:(ix + 1) * N * N] = tree.query(X, k=1)[1]
        return gid
    gid = np.empty(N ** 3, np.int32)
    best = np.full(N ** 3, np.inf)
    X = np.stack(np.meshgrid(c, c, c, indexing="ij"), -1).reshape(-1, 3)
    for k in range(ngrains):
        d = X - seeds[k]
        d -= np.round(d)
        d2 = np.einsum("ij,ij->i", d, d)
        m = d2 < best
        best[m] = d2[m]
        gid[m] = k
    return gid


def polycrystal(N, ngrains, seed):
    """(grain id [V], angles [3][V] degrees) for a seeded periodic Voronoi
    polycrystal with random orientations."""
    rng = np.random.default_rng(seed)
    gid = voronoi(rng, N, ngrains)
    ang = kocks_angles(random_rotations(rng, ngrains))
    return gid, np.ascontiguousarray(ang[gid].T)
