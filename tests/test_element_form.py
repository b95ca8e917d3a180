"""CPU: the reference-curl form of the element Jacobian used by the element
operator of the float32 CG (sn_element_newton mode 2).

The kernel applies J_e = S C^T Mr C S with C the integer 3x6 table below
(rows of the header's documentation: c_j for the local edges (0,1), (0,2),
(0,3), (1,2), (1,3), (2,3)) and S the orientation signs.  Pinned here:
the table is half the Whitney reference curls; every element matrix of the
oracle's restatement of the reference kernel (fem_c/ctools.c:960-1388) has
exactly this rank-3 form; and the table annihilates gradients exactly.
"""
import numpy as np

from oracle import oracle as orc

C_INT = np.array([[0, 1, -1, 0, 0, 1],      # x components of c_0..c_5
                  [-1, 0, 1, 0, -1, 0],     # y
                  [1, -1, 0, 1, 0, 0]],     # z
                 dtype=np.float64)          # column j = c_j


def test_integer_table_is_half_the_reference_curls():
    from snoopy_amd.finite_element import reference_curls
    np.testing.assert_array_equal(2.0 * C_INT.T, reference_curls())


def test_oracle_element_matrices_have_the_reference_curl_form():
    from snoopy_amd.magnets import Dipole
    from snoopy_amd import finite_element as fe
    dp = Dipole(h=0.2)
    s = dp.solver(quad_order=2, verbose=False)
    asm = s.curl_curl_factory
    inp = orc.domain_inputs(asm, 0, 2)          # the iron domain
    ne = len(inp['cells'])
    model = orc._model_of(asm.material_list[0])
    rng = np.random.default_rng(3)
    x = rng.standard_normal(asm.num_dofs) * 2e-2
    # zero source field: B = curl A alone, so nu and nu' vary and dK != 0
    Bs = np.zeros((ne * inp['Q'], 3))
    r, c, vK, vdK, _ = orc.element_triplets(inp['nodes'], inp['cells'], inp['gids'],
                                            inp['orient'], inp['curls'], inp['d_phi'], inp['w'],
                                            model, x, Bs)
    g = np.asarray(inp['gids']).reshape(ne, 6)
    # local (j, k) of each of the element's 36 triplets from its global ids
    jj = np.argmax(r.reshape(ne, 36, 1) == g[:, None, :], axis=2)
    kk = np.argmax(c.reshape(ne, 36, 1) == g[:, None, :], axis=2)
    assert (np.take_along_axis(g, jj, 1) == r.reshape(ne, 36)).all()
    assert (np.take_along_axis(g, kk, 1) == c.reshape(ne, 36)).all()
    Je = np.zeros((ne, 6, 6))
    np.add.at(Je, (np.repeat(np.arange(ne), 36), jj.ravel(), kk.ravel()), vK + vdK)
    assert np.abs(vdK).max() > 1e-3 * np.abs(vK).max()
    sgn = np.where((np.asarray(inp['orient'])[:, None] >> np.arange(6)) & 1, -1.0, 1.0)
    A = C_INT[None, :, :] * sgn[:, None, :]                  # [ne, 3, 6] = C S
    AAt_inv = np.linalg.inv(A @ A.transpose(0, 2, 1))
    Mr = AAt_inv @ A @ Je @ A.transpose(0, 2, 1) @ AAt_inv
    rec = A.transpose(0, 2, 1) @ Mr @ A
    assert np.abs(rec - Je).max() <= 1e-12 * np.abs(Je).max()
    assert np.abs(Mr - Mr.transpose(0, 2, 1)).max() <= 1e-12 * np.abs(Mr).max()
    # the kernel's Mr is 4 Jm^T M Jm with M positive definite: so is Mr
    assert (np.linalg.eigvalsh(0.5 * (Mr + Mr.transpose(0, 2, 1)))[:, 0] > 0).all()
    assert fe.LOCAL_EDGES.tolist() == [[0, 1], [0, 2], [0, 3], [1, 2], [1, 3], [2, 3]]


def test_integer_table_annihilates_gradients_exactly():
    from snoopy_amd.finite_element import LOCAL_EDGES
    rng = np.random.default_rng(8)
    for _ in range(200):
        phi = rng.integers(-1000, 1000, size=4).astype(np.float64)
        # edge values of grad phi in local orientation, then any element signs
        xh = np.array([phi[b] - phi[a] for a, b in LOCAL_EDGES])
        s = rng.choice([-1.0, 1.0], size=6)
        x = s * xh                        # stored (global-orientation) values
        u = C_INT @ (s * x)               # what the kernel forms: c (S x)
        assert (u == 0.0).all()
