"""Pins of the oracle's fine Wilson-clover operator and of CoarsenOperator (no GPU).

The fine operator's convention is this build's own (Grid's is not available);
it is pinned by the Clifford algebra, the free-field plane-wave spectrum
(closed form), gamma5-hermiticity, a numpy restatement of the clover leaves,
and M = Mdiag + sum_p Mdir_p.  CoarsenOperator is pinned by the Galerkin
identity D_c = R D P (MultigridPreconditioner.h:398-421), independent of the
inner/outer mask bookkeeping.
"""
import numpy as np
import pytest

import oracle
from helpers import gauss, rel

MASS, CSW = -0.1, 1.0


def test_gamma_matrices():
    G = [oracle.gamma(mu) for mu in range(4)]
    for a in range(4):
        assert np.allclose(G[a], G[a].conj().T)
        for b in range(4):
            assert np.allclose(G[a] @ G[b] + G[b] @ G[a], 2 * np.eye(4) * (a == b))
    assert np.allclose(G[0] @ G[1] @ G[2] @ G[3], np.diag([1, 1, -1, -1]))


def test_gauge_is_su3():
    U = oracle.su3_gauge((2, 2, 2, 2), 0.4, 7, 1).reshape(-1, 3, 3)
    assert np.abs(U @ U.conj().transpose(0, 2, 1) - np.eye(3)).max() < 1e-13
    assert np.abs(np.linalg.det(U) - 1).max() < 1e-13


def test_free_field_plane_wave():
    """U = 1: D(p) = (4+m) - sum cos p_mu + i sum gamma_mu sin p_mu on a plane wave."""
    dims = (4, 2, 6, 4)
    V = int(np.prod(dims))
    U = np.zeros((4, V, 3, 3), dtype=complex)
    U[:] = np.eye(3)
    diag = oracle.wilson_diag(dims, U, MASS, CSW)
    n = np.array([1, 1, 2, 3])
    p = 2 * np.pi * n / np.array(dims)
    c = np.indices(dims[::-1]).reshape(4, -1)[::-1]  # x_d per site
    phase = np.exp(1j * (p[:, None] * c).sum(0))
    chi = gauss(12, 3).reshape(4, 3)
    psi = phase[:, None, None] * chi[None]
    out = oracle.wilson_M(dims, U, diag, psi.reshape(V, 12)).reshape(V, 4, 3)
    G = [oracle.gamma(mu) for mu in range(4)]
    Dp = (4 + MASS - np.cos(p).sum()) * np.eye(4) + 1j * sum(G[m] * np.sin(p[m]) for m in range(4))
    ref = phase[:, None, None] * (Dp @ chi)[None]
    assert rel(ref, out) < 1e-13


def leaves_numpy(U, dims, x, mu, nu):
    """Clover leaves at site x by numpy (independent restatement)."""
    def sh(y, d, s):
        c = list(np.unravel_index(y, dims[::-1]))[::-1]
        c[d] = (c[d] + s) % dims[d]
        return int(np.ravel_multi_index(c[::-1], dims[::-1]))
    L = lambda m, y: U[m, y]
    D = lambda A: A.conj().T
    xm, xn = sh(x, mu, 1), sh(x, nu, 1)
    xmm, xmn = sh(x, mu, -1), sh(x, nu, -1)
    Q = (L(mu, x) @ L(nu, xm) @ D(L(mu, xn)) @ D(L(nu, x))
         + L(nu, x) @ D(L(mu, sh(xmm, nu, 1))) @ D(L(nu, xmm)) @ L(mu, xmm)
         + D(L(mu, xmm)) @ D(L(nu, sh(xmm, nu, -1))) @ L(mu, sh(xmm, nu, -1)) @ L(nu, xmn)
         + D(L(nu, xmn)) @ L(mu, xmn) @ L(nu, sh(xmn, mu, 1)) @ D(L(mu, x)))
    return Q


def test_clover_term_vs_numpy():
    dims = (2, 4, 2, 4)
    V = int(np.prod(dims))
    U = oracle.su3_gauge(dims, 0.4, 11, 2)
    diag = oracle.wilson_diag(dims, U, MASS, CSW)
    G = [oracle.gamma(mu) for mu in range(4)]
    for x in (0, 5, V - 1):
        C = (4 + MASS) * np.eye(12, dtype=complex)
        for mu in range(4):
            for nu in range(mu + 1, 4):
                Q = leaves_numpy(U, dims, x, mu, nu)
                F = (Q - Q.conj().T) / 8j
                S = 0.5j * (G[mu] @ G[nu] - G[nu] @ G[mu])
                C += -0.5 * CSW * np.kron(S, F)
        assert np.abs(C - diag[x]).max() < 1e-13
        assert np.abs(diag[x] - diag[x].conj().T).max() < 1e-13  # Hermitian
        # block diagonal in chirality (commutes with gamma5)
        assert np.abs(diag[x][:6, 6:]).max() < 1e-14 and np.abs(diag[x][6:, :6]).max() < 1e-14


def dense_fine(dims, U, diag):
    V = int(np.prod(dims))
    D = np.zeros((12 * V, 12 * V), dtype=complex)
    for k in range(12 * V):
        e = np.zeros((V, 12), dtype=complex)
        e.ravel()[k] = 1
        D[:, k] = oracle.wilson_M(dims, U, diag, e).ravel()
    return D


def test_fine_operator_split_and_g5_hermiticity():
    dims = (2, 2, 2, 2)
    V = 16
    U = oracle.su3_gauge(dims, 0.4, 13, 3)
    diag = oracle.wilson_diag(dims, U, MASS, CSW)
    psi = gauss(V * 12, 14).reshape(V, 12)
    total = oracle.wilson_diag_apply(diag, psi)
    for p in range(8):
        mu, disp = oracle.point_dir_disp(p)
        total += oracle.wilson_dir(dims, U, psi, mu, disp)
    assert rel(oracle.wilson_M(dims, U, diag, psi), total) < 1e-14
    D = dense_fine(dims, U, diag)
    g5 = np.tile(np.repeat([1.0, 1.0, -1.0, -1.0], 3), V)
    assert np.abs(D.conj().T - g5[:, None] * D * g5[None, :]).max() < 1e-13


@pytest.mark.parametrize("fd,blk,nb", [((4, 4, 4, 4), (2, 2, 2, 2), 3),
                                       ((6, 4, 2, 2), (2, 2, 2, 2), 2)])
def test_coarsen_operator_is_galerkin_product(fd, blk, nb):
    """D_c = R D P on random coarse vectors; odd coarse extent included (6/2 = 3)."""
    Vf = int(np.prod(fd))
    U = oracle.su3_gauge(fd, 0.4, 21, 4)
    diag = oracle.wilson_diag(fd, U, MASS, CSW)
    B = oracle.orthogonalise(fd, blk, nb, gauss(nb * Vf * 12, 22).reshape(nb, Vf, 12), 2)
    links = oracle.coarsen_operator(fd, blk, nb, U, diag, B)
    cdims = oracle.coarse_dims(fd, blk)
    Vc, N = int(np.prod(cdims)), 2 * nb
    for s in (23, 24):
        vc = gauss(Vc * N, s).reshape(Vc, N)
        lhs = oracle.coarse_apply(cdims, N, links, vc)
        rhs = oracle.project(fd, blk, nb, B,
                             oracle.wilson_M(fd, U, diag, oracle.promote(fd, blk, nb, B, vc)))
        assert rel(rhs, lhs) < 1e-13
    # gamma5-hermitian coarse operator: the forward links follow from the backward ones
    rebuilt = oracle.construct_remaining_links(cdims, N, links, -1).reshape(9, Vc, N, N)
    assert rel(links[0:4], rebuilt[0:4]) < 1e-13


@pytest.mark.parametrize("d1,blk,N1,nb2", [((4, 4, 4, 4), (2, 2, 2, 2), 8, 3),
                                           ((6, 2, 2, 2), (2, 2, 2, 2), 6, 2)])
def test_coarse_of_coarse_is_galerkin_product(d1, blk, N1, nb2):
    """Second-level CoarsenOperator: D_2 = R D_1 P on random vectors (odd extent included)."""
    from helpers import rand_links
    V1 = int(np.prod(d1))
    L1 = rand_links(d1, N1, 71)
    B = oracle.orthogonalise_ef(d1, blk, nb2, N1, gauss(nb2 * V1 * N1, 72).reshape(nb2, V1, N1), 2)
    L2 = oracle.coarsen_coarse_operator(d1, blk, nb2, N1, L1, B)
    d2 = oracle.coarse_dims(d1, blk)
    V2, N2 = int(np.prod(d2)), 2 * nb2
    v2 = gauss(V2 * N2, 73).reshape(V2, N2)
    lhs = oracle.coarse_apply(d2, N2, L2, v2)
    rhs = oracle.project_ef(d1, blk, nb2, N1, B,
                            oracle.coarse_apply(d1, N1, L1,
                                                oracle.promote_ef(d1, blk, nb2, N1, B, v2)))
    assert rel(rhs, lhs) < 1e-13
