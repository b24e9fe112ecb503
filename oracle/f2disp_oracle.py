"""Oracle (test infrastructure, not product code): CPU restatement of the
reference's displacement reconstruction f2disp (src/f2disp.f), numpy and
scipy.sparse.  Only tests/ may import it.

The reference builds an (N+1)^3-node mesh of N^3 eight-node bricks, one per
voxel (src/oumodel.f:11-130, blkgen/vblkn/vblke), assembles the Laplace-type
matrix A = sum_e int B^T B (form_BTB, f2disp.f:202-430, 2x2x2 Gauss) and the
right-hand side b = sum_e int B^T F^T (f2disp.f:62-147, one-point Gauss),
removes node 1 (its position is held at 0) and solves A x = b for the three
position components with PARDISO (f2disp.f:163-165, 445-557); the nodal
displacement is u = x - X (f2disp.f:169-174).  Here the solve is a sparse
direct solve (scipy spsolve) of the same reduced system.

Layouts: F is [9][N^3] with voxel index (x N + y) N + z (x slowest, the
reference's element order, oumodel.f vblke loops i, j, k with k innermost)
and component 3 i + j holding F_ij (the reference's Fn1(:, 3(i-1)+j));
nodes are numbered ix + P iy + P^2 iz with P = N + 1 (vblkn, x fastest);
u is returned as [P^3][3] (the reference's u(3(n-1)+k)).
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spl

# intpntb, 8-point rule (f2disp.f:606-621): abscissae and unit weights
G8 = 0.577350269189626
# local node order of a brick (vblke nd(1..8), oumodel.f:640-647): corners
# (i,j,k) (i+1,j,k) (i+1,j+1,k) (i,j+1,k), then the same at k+1
SIGNS = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                  [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], dtype=float)


def spacing(N, lengths):
    return np.asarray(lengths, dtype=float) / N


def node_coords(N, lengths):
    """X of every node, [P^3][3] (vblkn: uniform grid on the box)."""
    h = spacing(N, lengths)
    P = N + 1
    iz, iy, ix = np.meshgrid(np.arange(P), np.arange(P), np.arange(P), indexing="ij")
    return np.stack([ix.ravel() * h[0], iy.ravel() * h[1], iz.ravel() * h[2]], axis=1)


def shape_grad(xi, h):
    """QXY of shgb for an axis-aligned brick: dN_a/dX_k at natural point xi,
    [3][8], and det J."""
    dN = np.empty((3, 8))
    for k in range(3):
        others = [m for m in range(3) if m != k]
        dN[k] = SIGNS[:, k] / 8.0 * (1 + SIGNS[:, others[0]] * xi[others[0]]) * \
            (1 + SIGNS[:, others[1]] * xi[others[1]]) * 2.0 / h[k]
    return dN, h[0] * h[1] * h[2] / 8.0


def element_matrix(h):
    """A_local = sum over the 8 Gauss points of w det J B^T B (form_BTB,
    f2disp.f:398-413)."""
    A = np.zeros((8, 8))
    for l in range(8):  # intpntb point order: k fastest, then j, then i
        i, rem = divmod(l, 4)
        j, k = divmod(rem, 2)
        g = (-G8, G8)
        xi = np.array([g[k], g[j], g[i]])
        B, jdet = shape_grad(xi, h)
        A += 1.0 * jdet * (B.T @ B)
    return A


def element_nodes(N):
    """Node ids of every element in the reference's element order, [N^3][8]."""
    P = N + 1
    ex, ey, ez = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    ex, ey, ez = ex.ravel(), ey.ravel(), ez.ravel()  # voxel (x N + y) N + z
    s = ((SIGNS + 1) // 2).astype(int)
    return (ex[:, None] + s[None, :, 0]) + P * (ey[:, None] + s[None, :, 1]) + \
        P * P * (ez[:, None] + s[None, :, 2])


def assemble_A(N, lengths):
    """Global matrix over all P^3 nodes (node 1 not yet removed)."""
    Ae = element_matrix(spacing(N, lengths))
    en = element_nodes(N)
    rows = np.repeat(en, 8, axis=1).ravel()
    cols = np.tile(en, (1, 8)).ravel()
    vals = np.tile(Ae.ravel(), en.shape[0])
    P3 = (N + 1) ** 3
    return sp.csr_matrix((vals, (rows, cols)), shape=(P3, P3))


def assemble_b(F, N, lengths):
    """b = sum_e w det J B(0)^T F_e^T with the one-point rule (w = 8),
    f2disp.f:97-147; returns [P^3][3]."""
    h = spacing(N, lengths)
    B, jdet = shape_grad(np.zeros(3), h)
    Fm = np.asarray(F, dtype=float).reshape(3, 3, -1)  # [i][j][voxel]
    # b_local[a, i] = sum_k B[k, a] F_ik
    bl = 8.0 * jdet * np.einsum("ka,ikv->vai", B, Fm)
    en = element_nodes(N)
    b = np.zeros(((N + 1) ** 3, 3))
    np.add.at(b, en.ravel(), bl.reshape(-1, 3))
    return b


def f2disp(F, N, lengths=(1.0, 1.0, 1.0)):
    """Nodal displacements [P^3][3], node 1 (id 0) held at zero."""
    A = assemble_A(N, lengths)
    b = assemble_b(F, N, lengths)
    Ar = A[1:, 1:].tocsc()
    x = np.zeros_like(b)
    for k in range(3):
        x[1:, k] = spl.spsolve(Ar, b[1:, k])
    u = x - node_coords(N, lengths)
    u[0] = 0.0
    return u
