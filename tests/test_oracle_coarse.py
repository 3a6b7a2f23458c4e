"""Pins of the CPU oracle's coarse-operator functions (no GPU).

Each test checks the oracle against something it does not share code with:
the reference's explicit tables, a numpy dense brute force, numpy's LAPACK
inverse, or an algebraic identity the reference's own tests assert.
"""
import numpy as np
import pytest

import oracle
from helpers import (dense_operator, g5_links, gamma5_coarse, half_sites, neighbour_table,
                     rand_links, rand_vec, rel)


# ----------------------------------------------------------------- geometry
def test_point_table_matches_reference():
    # include/GeometryRework.h:71-74 (directions-faster, 4d): explicit table
    dirs = [0, 1, 2, 3, 0, 1, 2, 3, 0]
    disps = [+1, +1, +1, +1, -1, -1, -1, -1, 0]
    for p in range(9):
        assert oracle.point_dir_disp(p) == (dirs[p], disps[p])
        # round trip of src/Test_geom.cc:44-72
        assert oracle.point_from_dir_disp(dirs[p], disps[p]) == p


@pytest.mark.parametrize("dims", [(2, 2, 2, 2), (4, 2, 6, 2), (3, 3, 3, 3)])
def test_site_shift_matches_roll(dims):
    nbr = neighbour_table(dims)
    V = int(np.prod(dims))
    for p in range(8):
        d, s = oracle.point_dir_disp(p)
        assert all(oracle.site_shift(dims, x, d, s) == nbr[p, x] for x in range(V))


@pytest.mark.parametrize("dims", [(2, 2, 2, 2), (4, 2, 6, 2), (6, 3, 1, 5)])
def test_half_index_is_x_over_2(dims):
    # the library stores half-lattice site h = x >> 1; the oracle counts
    for cb in (0, 1):
        for h, x in enumerate(half_sites(dims, cb)):
            assert oracle.half_index(dims, int(x)) == h == int(x) >> 1


# ----------------------------------------------------------------- stencil
@pytest.mark.parametrize("dims,N", [((2, 2, 2, 2), 4), ((3, 3, 3, 3), 2), ((4, 2, 2, 2), 6)])
def test_coarse_M_and_points_vs_dense(dims, N):
    links = rand_links(dims, N, 11)
    v = rand_vec(dims, N, 12)
    V = int(np.prod(dims))
    for mask in (0x1FF, 0xFF, 1 << 8, 1 << 2, 1 << 5):
        D = dense_operator(dims, N, links, mask)
        ref = (D @ v.ravel()).reshape(V, N)
        got = oracle.coarse_apply(dims, N, links, v, mask)
        assert rel(ref, got) < 1e-14, hex(mask)


@pytest.mark.parametrize("dims,N", [((2, 2, 2, 2), 4), ((4, 2, 2, 2), 6)])
def test_mdag_is_g5_sandwich_and_adjoint(dims, N):
    V = int(np.prod(dims))
    g5 = gamma5_coarse(V, N)
    links = rand_links(dims, N, 21)
    v = rand_vec(dims, N, 22)
    D = dense_operator(dims, N, links)
    ref = (g5[:, None] * D * g5[None, :]) @ v.ravel()
    got = oracle.coarse_apply(dims, N, links, v, 0x1FF, dag=1)
    assert rel(ref, got) < 1e-14
    # with gamma5-hermitian links, Mdag is the true adjoint (src/Test_coarse_red_black.cc:303-347)
    L5 = g5_links(dims, N, 23)
    D5 = dense_operator(dims, N, L5)
    assert np.abs(D5.conj().T - g5[:, None] * D5 * g5[None, :]).max() < 1e-13
    phi = rand_vec(dims, N, 24)
    lhs = np.vdot(phi.ravel(), oracle.coarse_apply(dims, N, L5, v).ravel())
    rhs = np.vdot(oracle.coarse_apply(dims, N, L5, phi, dag=1).ravel(), v.ravel())
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)


@pytest.mark.parametrize("dims,N", [((4, 2, 2, 2), 4), ((2, 4, 2, 4), 6)])
def test_checkerboard_identities(dims, N):
    """Mee+Moo+Meo+Moe = M and Meo+Moe = Dhop (src/Test_coarse_red_black.cc:198-232, 433-480)."""
    V = int(np.prod(dims))
    links = rand_links(dims, N, 31)
    v = rand_vec(dims, N, 32)
    M = oracle.coarse_apply(dims, N, links, v)
    Dh = oracle.coarse_apply(dims, N, links, v, 0xFF)
    Md = oracle.coarse_apply(dims, N, links, v, 1 << 8)
    assert rel(M, Dh + Md) < 1e-14  # Mdiag + Dhop = M (src/Test_coarse_red_black.cc:170-196)
    assembled = np.zeros_like(v)
    for out_cb in (0, 1):
        in_cb = 1 - out_cb
        vin = v[half_sites(dims, in_cb)]
        hop = oracle.coarse_apply_cb(dims, N, links, vin, out_cb)
        assembled[half_sites(dims, out_cb)] = hop
    assert rel(Dh, assembled) < 1e-14
    mats = links[8]
    for cb in (0, 1):
        vin = v[half_sites(dims, cb)]
        got = oracle.site_matrix_apply(dims, N, mats, vin, cb)
        ref = np.einsum("xrc,xc->xr", mats[half_sites(dims, cb)], vin)
        assert rel(ref, got) < 1e-14
    # dagger on half lattices: G5C sandwich
    g = np.concatenate([np.ones(N // 2), -np.ones(N // 2)])
    vin = v[half_sites(dims, 0)]
    ref = oracle.coarse_apply_cb(dims, N, links, vin * g, 1) * g
    got = oracle.coarse_apply_cb(dims, N, links, vin, 1, dag=1)
    assert rel(ref, got) < 1e-14
    assert V % 2 == 0


@pytest.mark.parametrize("N", [2, 8, 24, 64])
def test_invert_vs_lapack(N):
    A = oracle.gaussian(N * N, 7, N).reshape(N, N) + 2.0 * np.eye(N)
    Ai = oracle.invert(A)
    assert np.abs(Ai - np.linalg.inv(A)).max() < 1e-11 * np.abs(Ai).max()
    # needs pivoting: zero leading diagonal
    P = np.roll(np.eye(N), 1, axis=0) if N > 1 else np.eye(N)
    B = P @ A
    B[0, 0] = 0
    assert np.abs(oracle.invert(B) @ B - np.eye(N)).max() < 1e-10
    with pytest.raises(np.linalg.LinAlgError):
        oracle.invert(np.zeros((N, N)))


def test_mooee_inv_identity():
    """MooeeInv * Mooee = 1 (src/Test_coarse_red_black.cc:280-301)."""
    dims, N = (2, 2, 2, 2), 8
    links = rand_links(dims, N, 41)
    inv = oracle.invert_self(dims, N, links)
    v = rand_vec(dims, N, 42)
    w = oracle.site_matrix_apply(dims, N, links[8], v)
    back = oracle.site_matrix_apply(dims, N, inv, w)
    assert rel(v, back) < 1e-12
    # adjoint application = conj transpose
    wd = oracle.site_matrix_apply(dims, N, links[8], v, adjoint=1)
    ref = np.einsum("xcr,xc->xr", links[8].conj(), v)
    assert rel(ref, wd) < 1e-14


def test_layout_round_trip_and_element():
    V, N = 3, 6
    flat = oracle.gaussian(9 * V * N * N, 5, 5)
    split = oracle.links_flat_to_split(V, N, flat)
    back = oracle.links_split_to_flat(V, N, split)
    assert np.array_equal(flat, back)
    # element check against include/Layout_converters.h:49-55: n = s*Nc_c + c
    nb = N // 2
    f = flat.reshape(9, V, N, N)
    s = split.reshape(9, V, 2, 2, nb, nb)
    assert f[3, 2, 1 * nb + 2, 0 * nb + 1] == s[3, 2, 1, 0, 2, 1]


def test_mrhs_equals_rhs_loop():
    """src/Benchmark_mrhs_coarsening.cc:227-242: the 5d apply equals a per-rhs loop."""
    dims, N, nrhs = (2, 2, 2, 2), 6, 5
    links = rand_links(dims, N, 51)
    v = rand_vec(dims, N, 52, nrhs=nrhs)
    got = oracle.coarse_mrhs_M(dims, N, nrhs, links, v)
    for k in range(nrhs):
        ref = oracle.coarse_apply(dims, N, links, np.ascontiguousarray(v[:, k, :]))
        assert rel(ref, got[:, k, :]) < 1e-14


@pytest.mark.parametrize("disp_have", [-1, +1])
def test_remaining_links_give_g5_hermitian_operator(disp_have):
    dims, N = (3, 2, 2, 2), 4
    V = int(np.prod(dims))
    L = rand_links(dims, N, 61)
    keep = [0, 1, 2, 3] if disp_have == +1 else [4, 5, 6, 7]
    drop = [p for p in range(8) if p not in keep]
    L[drop] = 0
    # self link must itself be gamma5-hermitian: A8 = G5 A8^dag G5
    g = np.concatenate([np.ones(N // 2), -np.ones(N // 2)])
    L[8] = 0.5 * (L[8] + g[None, :, None] * L[8].conj().transpose(0, 2, 1) * g[None, None, :])
    full = oracle.construct_remaining_links(dims, N, L, disp_have).reshape(9, V, N, N)
    assert np.array_equal(full[keep], L[keep])
    D = dense_operator(dims, N, full)
    g5 = gamma5_coarse(V, N)
    assert np.abs(D.conj().T - g5[:, None] * D * g5[None, :]).max() < 1e-13


def test_gaussian_generator_statistics():
    z = oracle.gaussian(200000, 123, 9)
    assert abs(z.real.mean()) < 0.01 and abs(z.imag.mean()) < 0.01
    assert abs(z.real.var() - 0.5) < 0.01 and abs(z.imag.var() - 0.5) < 0.01
    assert abs(np.mean(z.real * z.imag)) < 0.01
    # pure function of (seed, stream, index)
    assert np.array_equal(oracle.gaussian(1000, 123, 9), z[:1000])
    assert not np.array_equal(oracle.gaussian(1000, 123, 10), z[:1000])
