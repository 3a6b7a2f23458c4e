"""Shared test helpers: seeded inputs and an independent dense brute force.

The dense routines assemble the coarse operator as one (V*N) x (V*N) matrix
with numpy index arithmetic of their own (np.roll over the 4d site grid), so a
wrong index, sign or transpose in either the oracle or the library shows up
as a mismatch against them.
"""
from __future__ import annotations

import numpy as np

import oracle

SEED = 1 | (2 << 16) | (3 << 32) | (4 << 48)  # the reference's seeds {1,2,3,4}


def gauss(n: int, stream: int, seed: int = SEED) -> np.ndarray:
    return oracle.gaussian(n, seed, stream)


def rand_links(dims, N, stream=100, seed=SEED) -> np.ndarray:
    V = int(np.prod(dims))
    return gauss(9 * V * N * N, stream, seed).reshape(9, V, N, N)


def rand_vec(dims, N, stream=200, nrhs=None, seed=SEED) -> np.ndarray:
    V = int(np.prod(dims))
    if nrhs is None:
        return gauss(V * N, stream, seed).reshape(V, N)
    return gauss(V * nrhs * N, stream, seed).reshape(V, nrhs, N)


def g5_links(dims, N, stream=300, seed=SEED, herm_self=True) -> np.ndarray:
    """Synthetic links after src/Benchmark_create_coarse_coarse_operator.cc:218-221:
    Gaussian backward and self links, forward links rebuilt from the backward
    ones.  With herm_self the self link is also made gamma5-hermitian
    (A8 = G5 A8^dag G5), so that the whole operator is."""
    L = rand_links(dims, N, stream, seed)
    L[0:4] = 0
    if herm_self:
        g = np.concatenate([np.ones(N // 2), -np.ones(N // 2)])
        L[8] = 0.5 * (L[8] + g[None, :, None] * L[8].conj().transpose(0, 2, 1) * g[None, None, :])
    return oracle.construct_remaining_links(dims, N, L, -1).reshape(9, -1, N, N)


def site_grid(dims):
    """[x3][x2][x1][x0] array of lexicographic site indices."""
    V = int(np.prod(dims))
    return np.arange(V).reshape(dims[3], dims[2], dims[1], dims[0])


def neighbour_table(dims):
    """nbr[p][x] for the 9 points via np.roll (independent of the oracle)."""
    g = site_grid(dims)
    axis = {0: 3, 1: 2, 2: 1, 3: 0}  # lattice dim -> numpy axis of site_grid
    V = int(np.prod(dims))
    nbr = np.empty((9, V), dtype=np.int64)
    for p in range(9):
        if p == 8:
            nbr[p] = np.arange(V)
            continue
        mu = p % 4
        disp = 1 if p < 4 else -1
        # value at x = index of x + disp*mu  <=>  roll by -disp
        nbr[p] = np.roll(g, -disp, axis=axis[mu]).ravel()
    return nbr


def dense_operator(dims, N, links, mask=0x1FF) -> np.ndarray:
    V = int(np.prod(dims))
    L = np.asarray(links).reshape(9, V, N, N)
    nbr = neighbour_table(dims)
    D = np.zeros((V * N, V * N), dtype=np.complex128)
    for p in range(9):
        if not (mask >> p) & 1:
            continue
        for x in range(V):
            y = nbr[p, x]
            D[x * N:(x + 1) * N, y * N:(y + 1) * N] += L[p, x]
    return D


def gamma5_coarse(V, N) -> np.ndarray:
    g = np.concatenate([np.ones(N // 2), -np.ones(N // 2)])
    return np.tile(g, V)


def parity_of(dims) -> np.ndarray:
    V = int(np.prod(dims))
    c = np.indices((dims[3], dims[2], dims[1], dims[0])).reshape(4, V)
    return c.sum(axis=0) % 2


def half_sites(dims, cb) -> np.ndarray:
    """Full indices of the parity-cb sites in lexicographic order."""
    par = parity_of(dims)
    return np.nonzero(par == cb)[0]


def rel(ref, res) -> float:
    return oracle.relative_error(ref, res)
