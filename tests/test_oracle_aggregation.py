"""Pins of the oracle's block-transfer functions (no GPU).

Checks against numpy brute force (einsum per block), numpy QR with the
phase convention of Gram-Schmidt (R with a real positive diagonal), the
reference's transfer identities (MultigridPreconditioner.h:352-396) and the
split == doubled equivalences the reference asserts
(src/Benchmark_aggregation.cc:229-232, 263-266, 334-341).
"""
import numpy as np
import pytest

import oracle
from helpers import gauss, rel

FD, BLK = (4, 4, 4, 8), (2, 2, 2, 4)


def block_index(fine_dims, block):
    """coarse site of every fine site, by numpy index arithmetic."""
    c = np.indices(fine_dims[::-1]).reshape(4, -1)[::-1]  # c[d] = x_d, dim 0 fastest
    cd = oracle.coarse_dims(fine_dims, block)
    cc = [c[d] // block[d] for d in range(4)]
    return cc[0] + cd[0] * (cc[1] + cd[1] * (cc[2] + cd[2] * cc[3]))


def rand_basis(fd, nb, stream):
    Vf = int(np.prod(fd))
    return gauss(nb * Vf * 12, stream).reshape(nb, Vf, 12)


def test_lut_round_trip():
    """src/Test_coarsening_lookuptable.cc:37-87: forward lists and reverse table agree."""
    lut, rev = oracle.lut_populate(FD, BLK)
    Vf = int(np.prod(FD))
    assert sorted(np.concatenate(lut).tolist()) == list(range(Vf))
    for sc, sites in enumerate(lut):
        assert (rev[sites] == sc).all()
    assert (rev == block_index(FD, BLK)).all()


@pytest.mark.parametrize("p", range(9))
def test_inner_outer_luts_partition_blocks(p):
    inner, _ = oracle.lut_populate(FD, BLK, oracle.MASK_INNER, p)
    outer, _ = oracle.lut_populate(FD, BLK, oracle.MASK_OUTER, p)
    full, _ = oracle.lut_populate(FD, BLK)
    d, s = oracle.point_dir_disp(p)
    for sc in range(len(full)):
        assert sorted(inner[sc].tolist() + outer[sc].tolist()) == full[sc].tolist()
        if p == 8:
            assert len(outer[sc]) == 0
        else:  # outer face of a 2 or 4 extent block: 1/b of the block
            assert len(outer[sc]) == len(full[sc]) // BLK[d]


def test_project_promote_vs_einsum():
    nb = 3
    Vf = int(np.prod(FD))
    B = rand_basis(FD, nb, 401)
    f = gauss(Vf * 12, 402).reshape(Vf, 12)
    bi = block_index(FD, BLK)
    Vc = bi.max() + 1
    ref = np.zeros((Vc, 2, nb), dtype=complex)
    for k in range(2):
        sl = slice(6 * k, 6 * k + 6)
        per_site = np.einsum("ixe,xe->xi", B[:, :, sl].conj(), f[:, sl])
        np.add.at(ref[:, k, :], bi, per_site)
    got = oracle.project(FD, BLK, nb, B, f)
    assert rel(ref.reshape(Vc, 2 * nb), got) < 1e-14
    c = gauss(Vc * 2 * nb, 403).reshape(Vc, 2, nb)
    ref_f = np.zeros((Vf, 12), dtype=complex)
    for k in range(2):
        sl = slice(6 * k, 6 * k + 6)
        ref_f[:, sl] = np.einsum("xi,ixe->xe", c[bi, k, :], B[:, :, sl])
    got_f = oracle.promote(FD, BLK, nb, B, c.reshape(Vc, 2 * nb))
    assert rel(ref_f, got_f) < 1e-14


def test_masked_project_sums_to_full():
    nb = 2
    Vf = int(np.prod(FD))
    B = rand_basis(FD, nb, 411)
    f = gauss(Vf * 12, 412).reshape(Vf, 12)
    full = oracle.project(FD, BLK, nb, B, f)
    for p in (0, 5):
        i = oracle.project(FD, BLK, nb, B, f, oracle.MASK_INNER, p)
        o = oracle.project(FD, BLK, nb, B, f, oracle.MASK_OUTER, p)
        assert rel(full, i + o) < 1e-14


def test_orthogonalise_is_phase_fixed_qr():
    """MGS per (block, chirality) equals the thin QR with a positive real R diagonal."""
    nb = 4
    Vf = int(np.prod(FD))
    B = rand_basis(FD, nb, 421)
    Q = oracle.orthogonalise(FD, BLK, nb, B, passes=1)
    bi = block_index(FD, BLK)
    for sc in (0, 7, bi.max()):
        sites = np.nonzero(bi == sc)[0]
        for k in range(2):
            sl = slice(6 * k, 6 * k + 6)
            A = B[:, sites, sl].reshape(nb, -1).T  # rows (site, e), columns i
            q, r = np.linalg.qr(A)
            ph = np.diag(r) / np.abs(np.diag(r))
            q = q * ph[None, :]
            got = Q[:, sites, sl].reshape(nb, -1).T
            assert np.abs(q - got).max() < 1e-12


def _cgs_dgks_sweeps(A, passes):
    """The device kernel's column sweep on one panel A (rows x columns): all
    <u, v> with the finished columns at once, v -= U <U, v>, repeated when
    |v| fell below 1/sqrt(2) of its value before the projection, v /= |v|."""
    Q = A.copy()
    for _ in range(passes):
        for v in range(Q.shape[1]):
            x = Q[:, v].copy()
            n2 = np.vdot(x, x).real
            for _rep in range(2):
                if v == 0:
                    break
                x = x - Q[:, :v] @ (Q[:, :v].conj().T @ x)
                m2 = np.vdot(x, x).real
                if m2 >= 0.5 * n2:
                    break
                n2 = m2
            Q[:, v] = x / np.sqrt(np.vdot(x, x).real)
    return Q


@pytest.mark.parametrize("passes", [1, 2])
def test_classical_sweep_with_reorthogonalisation_equals_mgs(passes):
    """The GPU orthonormalisation sweeps columns with classical projections
    (all inner products with the finished columns reduced together) and a
    second projection when the norm drops below 1/sqrt(2); on full-rank
    panels this gives the modified Gram-Schmidt result of the oracle
    (include/AggregationKernelsRework.h:266-341) to rounding, also for a
    panel with nearly dependent columns, where the second projection acts."""
    nb = 6
    B = rand_basis(FD, nb, 441)
    bi = block_index(FD, BLK)
    sites = np.nonzero(bi == 3)[0]
    # columns 4 and 5 nearly along columns 0 and 1 (cancellation by 1e-6)
    B[4] = B[0] + 1e-6 * B[4]
    B[5] = B[1] - 1e-6 * B[5]
    Q = oracle.orthogonalise(FD, BLK, nb, B, passes=passes)
    for k in range(2):
        sl = slice(6 * k, 6 * k + 6)
        A = B[:, sites, sl].reshape(nb, -1).T
        got = _cgs_dgks_sweeps(A, passes)
        ref = Q[:, sites, sl].reshape(nb, -1).T
        assert np.abs(got.conj().T @ got - np.eye(nb)).max() < 1e-12
        # the nearly dependent columns are determined only to ~1e-16 / 1e-6
        assert np.abs(got[:, :4] - ref[:, :4]).max() < 1e-12
        assert np.abs(got - ref).max() < 1e-8


def test_transfer_identities():
    """(1 - P R) v_i = 0 and (1 - R P) v_c = 0 (MultigridPreconditioner.h:352-396)."""
    nb = 3
    Vf = int(np.prod(FD))
    Q = oracle.orthogonalise(FD, BLK, nb, rand_basis(FD, nb, 431), passes=2)
    Vc = int(np.prod(oracle.coarse_dims(FD, BLK)))
    vc = gauss(Vc * 2 * nb, 432).reshape(Vc, 2 * nb)
    assert rel(vc, oracle.project(FD, BLK, nb, Q, oracle.promote(FD, BLK, nb, Q, vc))) < 1e-13
    for i in range(nb):
        vi = np.ascontiguousarray(Q[i])
        back = oracle.promote(FD, BLK, nb, Q, oracle.project(FD, BLK, nb, Q, vi))
        assert rel(vi, back) < 1e-13
        # CheckOrthogonal: R v_i = e_i in both chiralities (AggregationRework.h:91-108)
        e = np.zeros((Vc, 2, nb), dtype=complex)
        e[:, :, i] = 1.0
        assert rel(e.reshape(Vc, 2 * nb), oracle.project(FD, BLK, nb, Q, vi)) < 1e-13


def test_split_equals_doubled_forms():
    """src/Benchmark_aggregation.cc:229-232, 263-266, 334-341."""
    nb = 3
    Vf = int(np.prod(FD))
    B = rand_basis(FD, nb, 441)
    Bd = oracle.chiral_double(Vf, nb, B)
    # doubled vectors have definite chirality and sum back to the original (undo, :203-213)
    assert np.array_equal(Bd[:nb] + Bd[nb:], B)
    assert np.all(Bd[:nb, :, 6:] == 0) and np.all(Bd[nb:, :, :6] == 0)
    f = gauss(Vf * 12, 442).reshape(Vf, 12)
    assert rel(oracle.project(FD, BLK, nb, B, f),
               oracle.project_doubled(FD, BLK, 2 * nb, Bd, f)) < 1e-14
    Q = oracle.orthogonalise(FD, BLK, nb, B, passes=1)
    Qd = oracle.orthogonalise_doubled(FD, BLK, 2 * nb, Bd, passes=1)
    assert np.abs(Qd[:nb] + Qd[nb:] - Q).max() < 1e-13
