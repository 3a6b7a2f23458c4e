"""GPU parity of the block transfer operators against the FP64 oracle.

Tolerances (relative L2): FP64 1e-13 (project/promote) and 1e-12 (Gram-Schmidt,
which compounds nb^2 rounding steps); FP32 vectors 2e-6.
"""
import numpy as np
import pytest

import oracle
from helpers import gauss, rel

pytestmark = pytest.mark.gpu

CASES = [((8, 8, 8, 8), (4, 4, 4, 4), 12),   # fine 8^4 -> 2^4, N = 24
         ((4, 4, 4, 8), (2, 2, 2, 4), 3),
         ((4, 8, 4, 4), (4, 4, 2, 2), 16),
         ((8, 4, 4, 4), (2, 2, 2, 2), 32)]


def make(gpu, fd, blk, nb, stream=501, ortho=True):
    Vf = int(np.prod(fd))
    B = gauss(nb * Vf * 12, stream).reshape(nb, Vf, 12)
    agg = gpu.Aggregation(fd, blk, nb)
    agg.set_basis(gpu.DeviceBuffer.from_numpy(B))
    return agg, B


@pytest.fixture(params=[0, 1], ids=["cholqr", "sweep"])
def ortho_method(gpu, request):
    """Both orthonormalisation methods of the library: Cholesky QR in two
    rounds (the default) and the column sweep."""
    gpu.call("mgc_aggregation_set_ortho_method", request.param)
    yield request.param
    gpu.call("mgc_aggregation_set_ortho_method", 0)


@pytest.mark.parametrize("fd,blk,nb", CASES + [((4, 4, 4, 4), (2, 2, 2, 2), 20),
                                               ((4, 4, 4, 8), (4, 4, 2, 2), 17)])
def test_orthogonalise_matches_oracle(gpu, fd, blk, nb, ortho_method):
    print(f"[case] ortho fd={fd} blk={blk} nb={nb} method={ortho_method}", flush=True)
    agg, B = make(gpu, fd, blk, nb)
    for passes in (1, 2):
        agg.set_basis(gpu.DeviceBuffer.from_numpy(B))
        agg.orthogonalise(passes)
        d = gpu.DeviceBuffer(B.nbytes)
        agg.get_basis(d)
        gpu.synchronize()
        got = d.to_numpy(np.complex128, B.shape)
        ref = oracle.orthogonalise(fd, blk, nb, B, passes)
        assert rel(ref, got) < 1e-12, passes
        errs = agg.check_orthogonal()  # squared deviations summed over coarse sites
        assert errs.max() < 1e-20, errs.max()


@pytest.mark.parametrize("passes", [1, 2])
def test_orthogonalise_nearly_dependent_columns(gpu, passes, ortho_method):
    """A basis whose vectors 4 and 5 are nearly along 0 and 1 (cancellation
    1e-6), where the sweep's second projection acts.  The well-determined
    columns match the oracle's modified Gram-Schmidt to 1e-12, the nearly
    dependent ones to 1e-8 (they carry ~1e-16 / 1e-6 rounding), and every
    panel is orthonormal."""
    fd, blk, nb = (4, 4, 4, 8), (2, 2, 2, 4), 6
    print(f"[case] ortho nearly dependent passes={passes}", flush=True)
    Vf = int(np.prod(fd))
    B = gauss(nb * Vf * 12, 561).reshape(nb, Vf, 12)
    B[4] = B[0] + 1e-6 * B[4]
    B[5] = B[1] - 1e-6 * B[5]
    agg = gpu.Aggregation(fd, blk, nb)
    agg.set_basis(gpu.DeviceBuffer.from_numpy(B))
    agg.orthogonalise(passes)
    d = gpu.DeviceBuffer(B.nbytes)
    agg.get_basis(d)
    gpu.synchronize()
    got = d.to_numpy(np.complex128, B.shape)
    ref = oracle.orthogonalise(fd, blk, nb, B, passes)
    assert rel(ref[:4], got[:4]) < 1e-12
    assert rel(ref, got) < 1e-8
    assert agg.check_orthogonal().max() < 1e-20


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("fd,blk,nb", CASES)
@pytest.mark.parametrize("nrhs", [1, 5])
def test_project_promote_match_oracle(gpu, fd, blk, nb, prec, nrhs):
    print(f"[case] transfer fd={fd} blk={blk} nb={nb} prec={prec} nrhs={nrhs}", flush=True)
    agg, B = make(gpu, fd, blk, nb, 511)
    Vf = int(np.prod(fd))
    Vc = int(np.prod(oracle.coarse_dims(fd, blk)))
    dt = gpu.complex_dtype(prec)
    tol = 1e-13 if prec == 0 else 2e-6
    f = gauss(Vf * nrhs * 12, 512).reshape(Vf, nrhs, 12)
    fq = f.astype(dt).astype(np.complex128)
    dc = gpu.DeviceBuffer(Vc * nrhs * 2 * nb * np.dtype(dt).itemsize)
    agg.project(gpu.DeviceBuffer.from_numpy(f.astype(dt)), dc, prec, nrhs)
    gpu.synchronize()
    got = dc.to_numpy(dt, (Vc, nrhs, 2 * nb)).astype(np.complex128)
    for r in range(nrhs):
        ref = oracle.project(fd, blk, nb, B, np.ascontiguousarray(fq[:, r]))
        assert rel(ref, got[:, r]) < tol, f"project rhs={r}"
    c = gauss(Vc * nrhs * 2 * nb, 513).reshape(Vc, nrhs, 2 * nb)
    cq = c.astype(dt).astype(np.complex128)
    df = gpu.DeviceBuffer(Vf * nrhs * 12 * np.dtype(dt).itemsize)
    agg.promote(gpu.DeviceBuffer.from_numpy(c.astype(dt)), df, prec, nrhs)
    gpu.synchronize()
    gotf = df.to_numpy(dt, (Vf, nrhs, 12)).astype(np.complex128)
    for r in range(nrhs):
        ref = oracle.promote(fd, blk, nb, B, np.ascontiguousarray(cq[:, r]))
        assert rel(ref, gotf[:, r]) < tol, f"promote rhs={r}"


def test_full_size_transfer_identities(gpu):
    """Fine 32^4, 4^4 blocks, nb = 16 (coarse 8^4, N = 32): R P v_c = v_c after
    orthonormalisation, and R of a basis vector is e_i (CheckOrthogonal), with
    the basis generated on the device."""
    fd, blk, nb = (32, 32, 32, 32), (4, 4, 4, 4), 16
    print(f"[case] full-size transfer fd={fd} nb={nb}", flush=True)
    Vf, Vc, N = 32 ** 4, 8 ** 4, 32
    agg = gpu.Aggregation(fd, blk, nb)
    basis = gpu.DeviceBuffer(nb * Vf * 12 * 16)
    gpu.random_gaussian(basis, nb * Vf * 12, gpu.F64, 99, 1)
    agg.set_basis(basis)
    agg.orthogonalise(1)
    errs = agg.check_orthogonal()
    assert errs.max() < 1e-20 * Vc * N
    vc = gpu.DeviceBuffer(Vc * N * 16)
    gpu.random_gaussian(vc, Vc * N, gpu.F64, 99, 2)
    vf = gpu.DeviceBuffer(Vf * 12 * 16)
    back = gpu.DeviceBuffer(Vc * N * 16)
    agg.promote(vc, vf, gpu.F64)
    agg.project(vf, back, gpu.F64)
    gpu.synchronize()
    a = vc.to_numpy(np.complex128, (Vc, N))
    b = back.to_numpy(np.complex128, (Vc, N))
    assert rel(a, b) < 1e-13


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("nrhs", [1, 3])
def test_project_masked_matches_oracle(gpu, prec, nrhs):
    """ProjectToSubspace over the inner (kind 1) and outer (kind 2) sites of
    every hopping point (include/CoarsenedMatrixRework.h:437-453) on Dirac
    fields, against the oracle's masked projection; inner + outer = all."""
    fd, blk, nb = (8, 4, 4, 8), (4, 2, 2, 4), 6
    print(f"[case] masked project fd={fd} blk={blk} nb={nb} prec={prec} nrhs={nrhs}", flush=True)
    agg, B = make(gpu, fd, blk, nb, 521)
    Vf = int(np.prod(fd))
    Vc = int(np.prod(oracle.coarse_dims(fd, blk)))
    dt = gpu.complex_dtype(prec)
    tol = 1e-13 if prec == 0 else 2e-6
    f = gauss(Vf * nrhs * 12, 522).reshape(Vf, nrhs, 12)
    fq = f.astype(dt).astype(np.complex128)
    din = gpu.DeviceBuffer.from_numpy(f.astype(dt))
    dc = gpu.DeviceBuffer(Vc * nrhs * 2 * nb * np.dtype(dt).itemsize)
    for point in range(8):
        parts = []
        for kind in (1, 2):
            agg.project_masked(din, dc, kind, point, prec, nrhs)
            gpu.synchronize()
            got = dc.to_numpy(dt, (Vc, nrhs, 2 * nb)).astype(np.complex128)
            for r in range(nrhs):
                ref = oracle.project(fd, blk, nb, B, np.ascontiguousarray(fq[:, r]), kind=kind,
                                     point=point)
                assert rel(ref, got[:, r]) < tol, (point, kind, r)
            parts.append(got)
        full = [oracle.project(fd, blk, nb, B, np.ascontiguousarray(fq[:, r]))
                for r in range(nrhs)]
        for r in range(nrhs):
            assert rel(full[r], parts[0][:, r] + parts[1][:, r]) < tol, (point, r)


def test_full_size_transfers_fp32(gpu):
    """FP32 project / promote at fine 32^4, 4^4 blocks, nb = 16 against the
    FP64 transfers of the same orthonormal basis (relative 2e-6 on FP32
    inputs), and R P v_c = v_c in FP32."""
    fd, blk, nb = (32, 32, 32, 32), (4, 4, 4, 4), 16
    print(f"[case] full-size transfer fp32 fd={fd} nb={nb}", flush=True)
    Vf, Vc, N = 32 ** 4, 8 ** 4, 32
    agg = gpu.Aggregation(fd, blk, nb)
    basis = gpu.DeviceBuffer(nb * Vf * 12 * 16)
    gpu.random_gaussian(basis, nb * Vf * 12, gpu.F64, 98, 1)
    agg.set_basis(basis)
    agg.orthogonalise(2)
    basis.close()
    f32 = gpu.DeviceBuffer(Vf * 12 * 8)
    gpu.random_gaussian(f32, Vf * 12, gpu.F32, 98, 2)
    gpu.synchronize()
    fh = f32.to_numpy(np.complex64, (Vf, 12))
    f64 = gpu.DeviceBuffer.from_numpy(fh.astype(np.complex128))
    c32, c64 = gpu.DeviceBuffer(Vc * N * 8), gpu.DeviceBuffer(Vc * N * 16)
    agg.project(f32, c32, gpu.F32)
    agg.project(f64, c64, gpu.F64)
    gpu.synchronize()
    a = c64.to_numpy(np.complex128, (Vc, N))
    b = c32.to_numpy(np.complex64, (Vc, N)).astype(np.complex128)
    assert rel(a, b) < 2e-6
    p32 = gpu.DeviceBuffer(Vf * 12 * 8)
    p64 = gpu.DeviceBuffer(Vf * 12 * 16)
    cin = gpu.DeviceBuffer.from_numpy(b.astype(np.complex64))
    agg.promote(cin, p32, gpu.F32)
    agg.promote(gpu.DeviceBuffer.from_numpy(b), p64, gpu.F64)
    back = gpu.DeviceBuffer(Vc * N * 8)
    agg.project(p32, back, gpu.F32)
    gpu.synchronize()
    x = p64.to_numpy(np.complex128, (Vf, 12))
    y = p32.to_numpy(np.complex64, (Vf, 12)).astype(np.complex128)
    assert rel(x, y) < 2e-6
    assert rel(b, back.to_numpy(np.complex64, (Vc, N)).astype(np.complex128)) < 2e-6


def test_orthogonalise_breakdown_falls_back_to_sweep(gpu):
    """Columns dependent to 1e-9 in every panel: the Cholesky pivot falls
    below its floor, every panel is flagged and orthonormalised by the column
    sweep instead, so the result equals the sweep's bit for bit and is
    orthonormal."""
    fd, blk, nb = (4, 4, 4, 8), (2, 2, 2, 4), 6
    print("[case] ortho breakdown fallback", flush=True)
    Vf = int(np.prod(fd))
    B = gauss(nb * Vf * 12, 571).reshape(nb, Vf, 12)
    B[3] = B[1] + 1e-9 * B[3]
    out = {}
    for method in (0, 1):
        gpu.call("mgc_aggregation_set_ortho_method", method)
        try:
            agg = gpu.Aggregation(fd, blk, nb)
            agg.set_basis(gpu.DeviceBuffer.from_numpy(B))
            agg.orthogonalise(2)
            d = gpu.DeviceBuffer(B.nbytes)
            agg.get_basis(d)
            gpu.synchronize()
            out[method] = d.to_numpy(np.complex128, B.shape)
            assert agg.check_orthogonal().max() < 1e-20
        finally:
            gpu.call("mgc_aggregation_set_ortho_method", 0)
    assert np.array_equal(out[0].view(np.float64), out[1].view(np.float64))


def test_orthogonalise_full_size_cholqr_matches_sweep(gpu):
    """Fine 32^4, 4^4 blocks, nb = 32 (the nbasis-64 set-up): the Cholesky
    form against the column sweep on the same Gaussian basis (1e-12), and
    orthonormal."""
    fd, blk, nb = (32, 32, 32, 32), (4, 4, 4, 4), 32
    print(f"[case] full-size ortho fd={fd} nb={nb}", flush=True)
    Vf = 32 ** 4
    basis = gpu.DeviceBuffer(nb * Vf * 12 * 16)
    gpu.random_gaussian(basis, nb * Vf * 12, gpu.F64, 97, 1)
    res = {}
    for method in (0, 1):
        gpu.call("mgc_aggregation_set_ortho_method", method)
        try:
            agg = gpu.Aggregation(fd, blk, nb)
            agg.set_basis(basis)
            agg.orthogonalise(2)
            d = gpu.DeviceBuffer(nb * Vf * 12 * 16)
            agg.get_basis(d)
            gpu.synchronize()
            res[method] = d.to_numpy(np.complex128, (nb, Vf, 12))
            d.close()
            if method == 0:
                assert agg.check_orthogonal().max() < 1e-20 * 8 ** 4
            agg.close()
        finally:
            gpu.call("mgc_aggregation_set_ortho_method", 0)
    assert rel(res[1], res[0]) < 1e-12
