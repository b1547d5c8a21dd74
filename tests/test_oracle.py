"""The FP64 oracle against what the mathematics fixes: closed forms, brute
force and invariances.  Each test is chosen so that a plausible slip in the
restatement (a dropped term, a wrong sign, a transposed operand, a missing
factor of 2) fails it.  CPU only.
"""
import numpy as np
import pytest

from mbptcode_amd import grids, synth
from oracle import isdf_oracle as O

HA_MEV = 27211.386245988


def small_system(seed=5, M=37, nocc=4, nvir=5, naux=None):
    """The recipe of the reference's row-split test (tests/
    test_isdf_block_action_rows.py:54-74), with a grid-gauge twin."""
    rng = np.random.default_rng(seed)
    norb = nocc + nvir
    eps = np.sort(rng.normal(size=norb))
    eps[nocc:] += 2.0
    X = 0.3 * rng.normal(size=(M, norb))
    naux = naux or 7
    D = 0.3 * rng.normal(size=(M, naux))
    W = rng.normal(size=(naux, naux))
    W = 0.5 * (W + W.T) + naux * np.eye(naux)
    z = rng.normal(size=(3, nocc, nvir))
    return eps, X, D, W, z


@pytest.fixture(scope='module')
def cfg0():
    inst = synth.config(0)
    D = O.grid_factors(inst.V)
    g = grids.static_screening_grid(inst.eps, inst.nocc, ntau=inst.ntau)
    W_aux = O.static_screening_matrix(inst.X, D, inst.eps, inst.nocc,
                                      g.tau_points, g.cosft_wt[0])
    return inst, D, g, W_aux


@pytest.mark.parametrize('spin_kappa', [2.0, 0.0])
def test_block_action_equals_dense_tda_matrix(spin_kappa):
    eps, X, D, W, z = small_system()
    apply_A, d = O.isdf_block_action(eps, 4, X, D, W, kappa=spin_kappa)
    A = O.dense_tda_matrix(eps, 4, X, D, W, kappa=spin_kappa)
    ref = (A @ z.reshape(3, -1).T).T.reshape(z.shape)
    got = apply_A(z)
    assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max()


def test_dense_matrix_has_the_known_pieces():
    """A - diag(d) splits into kappa*(ia|jb) - (ij|W|ab): the Hartree part is
    rank <= naux and PSD, the triplet drops it, and the whole is symmetric."""
    eps, X, D, W, z = small_system()
    A2 = O.dense_tda_matrix(eps, 4, X, D, W, kappa=2.0)
    A0 = O.dense_tda_matrix(eps, 4, X, D, W, kappa=0.0)
    H = 0.5 * (A2 - A0)
    assert np.allclose(A2, A2.T, atol=1e-13)
    ev = np.linalg.eigvalsh(H)
    assert ev.min() > -1e-12
    assert np.sum(ev > 1e-10 * ev.max()) <= D.shape[1]
    # brute force one element of each term
    i, a, j, b = 1, 2, 3, 0
    no = 4
    C = lambda p, q: X[:, p] * X[:, no + q]
    hart = C(i, a) @ (D @ D.T) @ C(j, b)
    Zt = D @ W @ D.T
    direct = (X[:, i] * X[:, j]) @ Zt @ (X[:, no + a] * X[:, no + b])
    nv = 5
    assert np.isclose(A2[i * nv + a, j * nv + b], 2 * hart - direct, rtol=1e-12)


@pytest.mark.parametrize('spin_kappa', [2.0, 0.0])
def test_full_action_b_block_equals_dense_b(spin_kappa):
    eps, X, D, W, z = small_system()
    apply_AB, d = O.isdf_block_action(eps, 4, X, D, W, kappa=spin_kappa, tda=False)
    Az, Bz = apply_AB(z)
    B = O.dense_casida_b_matrix(4, X, D, W, kappa=spin_kappa)
    ref = (B @ z.reshape(3, -1).T).T.reshape(z.shape)
    assert np.abs(Bz - ref).max() <= 1e-13 * np.abs(ref).max()
    apply_A, _ = O.isdf_block_action(eps, 4, X, D, W, kappa=spin_kappa)
    assert np.array_equal(Az, apply_A(z))
    tiled = O.isdf_block_action(eps, 4, X, D, W, kappa=spin_kappa, tda=False,
                                tile_memory_gb=6 * 37 * 8 / 1e9)[0](z)[1]
    assert np.abs(tiled - Bz).max() <= 1e-13 * np.abs(Bz).max()


def test_dense_b_has_the_swapped_exchange():
    """B = kappa (ia|jb) - (ib|W|ja): symmetric, the Hartree part shared with
    A, and one element by brute force against the swapped pair (ib)(ja) --
    the direct pair (ij)(ab) of A would fail it."""
    eps, X, D, W, z = small_system()
    no, nv = 4, 5
    B2 = O.dense_casida_b_matrix(no, X, D, W, kappa=2.0)
    B0 = O.dense_casida_b_matrix(no, X, D, W, kappa=0.0)
    A2 = O.dense_tda_matrix(eps, no, X, D, W, kappa=2.0)
    A0 = O.dense_tda_matrix(eps, no, X, D, W, kappa=0.0)
    assert np.allclose(B2, B2.T, atol=1e-13)
    assert np.allclose(B2 - B0, A2 - A0, atol=1e-13)
    i, a, j, b = 1, 2, 3, 0
    C = lambda p, q: X[:, p] * X[:, no + q]
    Zt = D @ W @ D.T
    hart = C(i, a) @ (D @ D.T) @ C(j, b)
    swap = C(i, b) @ Zt @ C(j, a)
    direct = (X[:, i] * X[:, j]) @ Zt @ (X[:, no + a] * X[:, no + b])
    assert np.isclose(B2[i * nv + a, j * nv + b], 2 * hart - swap, rtol=1e-12)
    assert not np.isclose(B2[i * nv + a, j * nv + b], 2 * hart - direct, rtol=1e-6)


def test_block_action_is_tiling_invariant():
    """tile_memory_gb sets the grid-row tile; the result must not move
    (tests/test_davidson_isdf_bse.py:165-177)."""
    eps, X, D, W, z = small_system()
    whole = O.isdf_block_action(eps, 4, X, D, W)[0](z)
    tiled = O.isdf_block_action(eps, 4, X, D, W, tile_memory_gb=6 * 37 * 8 / 1e9)[0](z)
    assert np.abs(whole - tiled).max() <= 1e-13 * np.abs(whole).max()


def test_screened_diagonal_is_the_dense_direct_diagonal():
    eps, X, D, W, z = small_system()
    apply_A, d = O.isdf_block_action(eps, 4, X, D, W)
    A0 = O.dense_tda_matrix(eps, 4, X, D, W, kappa=0.0)
    assert np.allclose(apply_A.screened_diagonal().ravel(), np.diag(A0), rtol=1e-12)


def test_proj_tau_is_the_pair_sum():
    eps, X, D, W, z = small_system()
    X_o, X_v, e_o, e_v, mu = O.split_branches(X, eps, 4)
    for tau in (0.0, 0.3, 2.5):
        got = O.polarizability_projected_tau(X_o, X_v, e_o, e_v, D, tau)
        ref = O.proj_tau_bruteforce(X, D, eps, 4, tau)
        assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max()
    # the tiled kernel (5 rows per tile) equals the whole one
    tiled = O.polarizability_projected_tau(X_o, X_v, e_o, e_v, D, 0.3,
                                           tile_memory_gb=5 * 3 * 37 * 8 / 1e9)
    whole = O.polarizability_projected_tau(X_o, X_v, e_o, e_v, D, 0.3)
    assert np.abs(tiled - whole).max() <= 1e-13 * np.abs(whole).max()
    # mu anywhere in the gap gives the same result
    other = O.polarizability_projected_tau(*O.split_branches(X, eps, 4, mu=mu + 0.1)[:4],
                                           D, 0.3)
    assert np.abs(other - whole).max() <= 1e-12 * np.abs(whole).max()


def test_chi0_frequency_matches_closed_form(cfg0):
    inst, D, g, W_aux = cfg0
    nu, wt = grids.gauss_legendre_grid(5, 0.5 * (inst.eps[5] - inst.eps[4]))
    gr = grids.TimeFrequencyGrid.minimax_split(16, *grids.particle_hole_window(inst.eps, 5)[:2],
                                              nu, wt)
    chi = O.chi0_imaginary_frequency(inst.X, D, inst.eps, 5, gr.tau_points, gr.cosft_wt)
    for k, w in enumerate(nu):
        ref = O.chi0_omega_closed_form(inst.X, D, inst.eps, 5, w)
        assert np.abs(chi[k] - ref).max() <= 2e-5 * np.abs(ref).max()


def test_static_screening_satisfies_dyson(cfg0):
    inst, D, g, W_aux = cfg0
    chi = O.chi0_imaginary_frequency(inst.X, D, inst.eps, 5, g.tau_points,
                                     g.cosft_wt[:1])[0]
    resid = W_aux @ (np.eye(len(chi)) - chi) - np.eye(len(chi))
    assert np.abs(resid).max() < 1e-8
    # the grid W = D W_aux D^T equals (I - V chi0)^-1 V
    Wg = O.grid_screened_interaction(W_aux, D)
    Chi = np.linalg.solve(D.T, np.linalg.solve(D.T, chi.T).T)   # chi0 on the grid
    ref = np.linalg.solve(np.eye(len(Chi)) - inst.V @ Chi, inst.V)
    assert np.abs(Wg - ref).max() <= 1e-8 * np.abs(ref).max()


def test_rpa_energy_matches_casida_trace():
    """Space-time E_c against Furche's trace formula on the same factors, at
    the quadrature's error (GL-40 frequencies, 20 minimax time points)."""
    inst = synth.make_instance(4, 12, 60, seed=2)
    D = O.grid_factors(inst.V)
    g = grids.rpa_energy_grid(inst.eps, inst.nocc, ntau=20, nfreq=40)
    e_st = O.rpa_correlation_energy_space_time(inst.X, D, inst.eps, 4, g.tau_points,
                                               g.cosft_wt, g.omega_weights)
    e_cas = O.rpa_energy_casida(inst.X, D, inst.eps, 4)
    assert e_cas < 0
    assert abs(e_st - e_cas) < 1e-6 + 1e-4 * abs(e_cas)


def test_rpa_energy_planted_sign_error_is_caught():
    """A sign slip in chi0 (chi0 -> -chi0) moves E_c far off the trace."""
    inst = synth.make_instance(4, 12, 60, seed=2)
    D = O.grid_factors(inst.V)
    g = grids.rpa_energy_grid(inst.eps, inst.nocc, ntau=20, nfreq=40)
    e_cas = O.rpa_energy_casida(inst.X, D, inst.eps, 4)
    e_bad = O.rpa_correlation_energy_space_time(inst.X, D, inst.eps, 4,
                                                g.tau_points, -g.cosft_wt,
                                                g.omega_weights)
    assert abs(e_bad - e_cas) > 1e-3 * abs(e_cas)


def test_davidson_matches_eigh(cfg0):
    inst, D, g, W_aux = cfg0
    apply_A, d = O.isdf_block_action(inst.eps, 5, inst.X, D, W_aux)
    A = O.dense_tda_matrix(inst.eps, 5, inst.X, D, W_aux)
    ref = np.linalg.eigvalsh(A)[:inst.nroots]
    mm = lambda V: apply_A(V.T.reshape(-1, 5, 19)).reshape(V.shape[1], -1).T
    e, X, conv, _ = O.davidson_tda(mm, d, inst.nroots, tol=1e-8,
                                   precond_diag=apply_A.screened_diagonal())
    assert conv.all()
    assert np.abs(e - ref).max() * HA_MEV < 1e-6
    r = A @ X - X * e
    assert np.linalg.norm(r, axis=0).max() <= 1e-8


def test_davidson_finds_a_degenerate_pair():
    """tests/test_davidson_symmetric.py:33-60: 400x400 with a degenerate pair."""
    rng = np.random.default_rng(1)
    n = 400
    d = np.sort(rng.uniform(1, 10, n))
    d[1] = d[0]
    Q = np.linalg.qr(rng.normal(size=(n, n)))[0]
    A = np.diag(d) + 1e-3 * (Q[:, :3] @ Q[:, :3].T)
    ref = np.linalg.eigvalsh(A)[:4]
    e, X, conv, _ = O.davidson_tda(lambda V: A @ V, np.diag(A), 4, tol=1e-9)
    assert conv.all()
    assert np.abs(e - ref).max() < 1e-9


def test_guess_indices_extend_over_degeneracy():
    d = np.array([1.0, 1.0, 1.0, 2.0, 2.0005, 2.0008, 3.0, 4.0, 5.0, 6.0, 7.0])
    idx = O.guess_indices(d, 2, guess_factor=2)
    # n = max(4, 6) = 6 reaches index of 2.0008; nothing degenerate past it
    assert len(idx) == 6
    d2 = np.array([1.0] * 30)
    assert len(O.guess_indices(d2, 2)) == 20   # capped at max(4k, 20)


def test_dense_casida_roots_solve_the_casida_equations(cfg0):
    """The closed-form Casida roots: A X + B Y = w X, B X + A Y = -w Y to
    round-off, X^T X - Y^T Y = 1, B = 0 gives eig(A), and the RPA trace
    formula E_c = (sum w - Tr A)/2 agrees with rpa_energy_casida."""
    inst, D, g, W_aux = cfg0
    A = O.dense_tda_matrix(inst.eps, inst.nocc, inst.X, D, W_aux)
    B = O.dense_casida_b_matrix(inst.nocc, inst.X, D, W_aux)
    w, Xc, Yc = O.dense_casida_roots(A, B)
    assert np.all(np.diff(w) >= 0) and w.min() > 0
    assert np.abs(A @ Xc + B @ Yc - Xc * w).max() < 1e-11 * w.max()
    assert np.abs(B @ Xc + A @ Yc + Yc * w).max() < 1e-11 * w.max()
    assert np.allclose(np.sum(Xc * Xc, 0) - np.sum(Yc * Yc, 0), 1.0, atol=1e-12)
    w0, X0, Y0 = O.dense_casida_roots(A, np.zeros_like(A))
    assert np.allclose(w0, np.linalg.eigvalsh(A), rtol=1e-12)
    assert np.abs(Y0).max() < 1e-12
    # RPA: A = d + 2K, B = 2K (no W) -> Furche's trace
    eps, X, D, W, z = small_system()
    Arpa = O.dense_tda_matrix(eps, 4, X, D, np.zeros_like(W))
    Brpa = O.dense_casida_b_matrix(4, X, D, np.zeros_like(W))
    wr = O.dense_casida_roots(Arpa, Brpa)[0]
    assert np.isclose(0.5 * (wr.sum() - np.trace(Arpa)),
                      O.rpa_energy_casida(X, D, eps, 4), rtol=1e-10)
