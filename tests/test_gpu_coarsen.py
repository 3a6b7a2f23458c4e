"""GPU parity of the fine Wilson-clover operator and of CoarsenOperator.

Bounds (relative L2 against the FP64 oracle): fine operator FP64 1e-13,
FP32 2e-6; coarse links FP64 1e-12; the Galerkin identity D_c = R D P through
the library's own kernels 1e-12.
"""
import numpy as np
import pytest

import oracle
from helpers import gauss, rel

pytestmark = pytest.mark.gpu
MASS, CSW = -0.1, 1.0


def fine_setup(gpu, dims, seed=31):
    U = oracle.su3_gauge(dims, 0.4, seed, 1)
    fop = gpu.FineOp(dims, MASS, CSW)
    fop.set_gauge(gpu.DeviceBuffer.from_numpy(U))
    return fop, U


def test_random_gauge_matches_oracle(gpu):
    dims = (4, 2, 2, 4)
    V = int(np.prod(dims))
    d = gpu.DeviceBuffer(4 * V * 9 * 16)
    gpu.random_gauge(d, dims, 0.4, 77, 3)
    got = d.to_numpy(np.complex128, (4, V, 3, 3))
    ref = oracle.su3_gauge(dims, 0.4, 77, 3)
    assert np.abs(got - ref).max() < 1e-13


@pytest.mark.parametrize("dims", [(4, 4, 2, 2), (2, 2, 4, 6)])
def test_clover_term_matches_oracle(gpu, dims):
    fop, U = fine_setup(gpu, dims)
    V = int(np.prod(dims))
    d = gpu.DeviceBuffer(V * 72 * 16)
    fop.get_diag(d)
    gpu.synchronize()
    blocks = d.to_numpy(np.complex128, (V, 2, 6, 6))
    ref = oracle.wilson_diag(dims, U, MASS, CSW)
    for l in range(2):
        assert np.abs(blocks[:, l] - ref[:, 6 * l:6 * l + 6, 6 * l:6 * l + 6]).max() < 1e-13
    assert np.abs(ref[:, :6, 6:]).max() < 1e-14  # off-chirality blocks vanish


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("nrhs", [1, 3, 12])
@pytest.mark.parametrize("dims", [(4, 4, 2, 6), (4, 4, 2, 8)])
def test_fine_operator_matches_oracle(gpu, prec, nrhs, dims):
    """(4, 4, 2, 8) has L3 % 8 == 0, so the apply's tiles are dealt to the
    eight XCDs by time slices (the order used at 32^4 and 48^4); (4, 4, 2, 6)
    takes the plain tile order.  FP32 with 12 rhs runs the prefetching
    multi-rhs kernel."""
    print(f"[case] fine op dims={dims} prec={prec} nrhs={nrhs}", flush=True)
    V = int(np.prod(dims))
    fop, U = fine_setup(gpu, dims, 41)
    diag = oracle.wilson_diag(dims, U, MASS, CSW)
    dt = gpu.complex_dtype(prec)
    psi = gauss(V * nrhs * 12, 42).reshape(V, nrhs, 12)
    pq = psi.astype(dt).astype(np.complex128)
    tol = 1e-13 if prec == 0 else 2e-6
    din = gpu.DeviceBuffer.from_numpy(psi.astype(dt))
    dout = gpu.DeviceBuffer(psi.size * np.dtype(dt).itemsize)

    def got():
        gpu.synchronize()
        return dout.to_numpy(dt, (V, nrhs, 12)).astype(np.complex128)

    g5 = np.repeat([1.0, 1.0, -1.0, -1.0], 3)
    for r in range(nrhs):
        pr = np.ascontiguousarray(pq[:, r])
        fop.M(din, dout, prec, nrhs)
        assert rel(oracle.wilson_M(dims, U, diag, pr), got()[:, r]) < tol
        fop.Mdag(din, dout, prec, nrhs)
        ref = g5 * oracle.wilson_M(dims, U, diag, pr * g5)
        assert rel(ref, got()[:, r]) < tol
        fop.Mdiag(din, dout, prec, nrhs)
        assert rel(oracle.wilson_diag_apply(diag, pr), got()[:, r]) < tol
        for mu in range(4):
            for disp in (1, -1):
                fop.Mdir(din, dout, mu, disp, prec, nrhs)
                assert rel(oracle.wilson_dir(dims, U, pr, mu, disp), got()[:, r]) < tol


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("dims,nrhs", [((4, 4, 2, 6), 1), ((4, 4, 2, 6), 3), ((8, 8, 8, 8), 1),
                                       ((8, 8, 8, 8), 5), ((8, 8, 8, 8), 12),
                                       ((4, 4, 2, 6), 7), ((4, 4, 2, 6), 10), ((4, 4, 2, 6), 100),
                                       ((4, 4, 4, 8), 256)])
def test_fine_operator_kernel_variants_bit_equal(gpu, prec, dims, nrhs):
    """The LDS-gathered apply, its prefetching multi-rhs form (FP32, nrhs
    >= 5) and its single-rhs form with one colour row per thread keep the
    per-thread arithmetic of the apply in which each thread reads its own
    records: every kind (M, Mdag, Mdiag, the eight Mdir) gives the same bits.
    (4, 4, 2, 6) has partly filled tiles (7 and 10 rhs: 36 and 25 sites a
    tile, the last one partial; 100 rhs: two sites of 100 threads each),
    8^4 and 4x4x4x8 the
    XCD-ordered tiles (256 rhs, the largest count: one site per
    workgroup)."""
    print(f"[case] fine variants dims={dims} nrhs={nrhs} prec={prec}", flush=True)
    V = int(np.prod(dims))
    fop, _ = fine_setup(gpu, dims, 43)
    dt = gpu.complex_dtype(prec)
    psi = gauss(V * nrhs * 12, 44).reshape(V, nrhs, 12).astype(dt)
    din = gpu.DeviceBuffer.from_numpy(psi)
    dout = gpu.DeviceBuffer(psi.size * np.dtype(dt).itemsize)
    kinds = [lambda: fop.M(din, dout, prec, nrhs), lambda: fop.Mdag(din, dout, prec, nrhs),
             lambda: fop.Mdiag(din, dout, prec, nrhs)]
    kinds += [lambda mu=mu, d=d: fop.Mdir(din, dout, mu, d, prec, nrhs)
              for mu in range(4) for d in (1, -1)]
    try:
        for run in kinds:
            out = []
            for variant in (0, 1, 2, 3):
                gpu.call("mgc_fine_op_set_variant", variant)
                dout.zero()
                run()
                gpu.synchronize()
                out.append(dout.to_numpy(dt, (V, nrhs, 12)))
            for o in out[1:]:
                assert np.array_equal(out[0], o), float(np.abs(out[0] - o).max())
    finally:
        gpu.call("mgc_fine_op_set_variant", -1)


CO_CASES = [((8, 8, 8, 8), (4, 4, 4, 4), 12),   # the reference's CPU case: 8^4 -> 2^4, N = 24
            ((6, 4, 2, 2), (2, 2, 2, 2), 4),    # odd coarse extent 3
            ((4, 4, 2, 2), (4, 2, 2, 2), 8),    # coarse extents 1 and 2
            ((4, 4, 4, 4), (2, 2, 2, 2), 16),
            ((8, 8, 8, 8), (4, 4, 4, 4), 20),   # the reference's default nbasis 40
            ((4, 4, 4, 4), (2, 2, 2, 2), 30)]


@pytest.fixture(params=[1, 0], ids=["mfma", "vector"])
def coarsen_variant(gpu, request):
    """Both FP64 kernels of the CoarsenOperator build: MFMA (default) and vector ALU."""
    gpu.call("mgc_coarsen_set_variant", request.param)
    yield request.param
    gpu.call("mgc_coarsen_set_variant", -1)


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("fd,blk,nb", CO_CASES + [((4, 4, 4, 4), (2, 2, 2, 2), 7),
                                                  ((4, 4, 4, 4), (2, 2, 2, 2), 17)])
def test_coarsen_operator_matches_oracle(gpu, fd, blk, nb, mode, coarsen_variant):
    if coarsen_variant == 0 and nb in (7, 17):
        pytest.skip("the vector form takes nb = 2 and even nb from 4 to 32")
    print(f"[case] coarsen fd={fd} blk={blk} nb={nb} mode={mode} variant={coarsen_variant}",
          flush=True)
    Vf = int(np.prod(fd))
    fop, U = fine_setup(gpu, fd, 51)
    diag = oracle.wilson_diag(fd, U, MASS, CSW)
    B = oracle.orthogonalise(fd, blk, nb, gauss(nb * Vf * 12, 52).reshape(nb, Vf, 12), 2)
    agg = gpu.Aggregation(fd, blk, nb)
    agg.set_basis(gpu.DeviceBuffer.from_numpy(B))
    cdims = oracle.coarse_dims(fd, blk)
    Vc, N = int(np.prod(cdims)), 2 * nb
    cop = gpu.CoarseOp(cdims, N, gpu.F64)
    gpu.coarsen_operator(cop, fop, agg, mode)
    d = gpu.DeviceBuffer(9 * Vc * N * N * 16)
    cop.get_links(d, gpu.LAYOUT_FLAT)
    gpu.synchronize()
    got = d.to_numpy(np.complex128, (9, Vc, N, N))
    ref = oracle.coarsen_operator(fd, blk, nb, U, diag, B)
    for p in range(9):
        assert rel(ref[p], got[p]) < 1e-12, f"point {p}"


def test_construct_remaining_links_matches_oracle(gpu):
    dims, N = (4, 2, 2, 2), 8
    V = int(np.prod(dims))
    L = gauss(9 * V * N * N, 61).reshape(9, V, N, N)
    for disp_have in (-1, 1):
        cop = gpu.CoarseOp(dims, N, gpu.F64)
        cop.set_links(gpu.DeviceBuffer.from_numpy(L), gpu.LAYOUT_FLAT, invert=False)
        gpu.construct_remaining_links(cop, disp_have)
        d = gpu.DeviceBuffer(9 * V * N * N * 16)
        cop.get_links(d, gpu.LAYOUT_FLAT)
        gpu.synchronize()
        got = d.to_numpy(np.complex128, (9, V, N, N))
        ref = oracle.construct_remaining_links(dims, N, L, disp_have).reshape(9, V, N, N)
        assert np.abs(got - ref).max() < 1e-15


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("d1,blk,N1,nb2", [((4, 4, 4, 4), (2, 2, 2, 2), 8, 3),
                                           ((6, 4, 2, 2), (2, 2, 2, 2), 16, 4),
                                           ((12, 12, 12, 12), (4, 4, 4, 4), 32, 16)])
def test_coarse_of_coarse_matches_oracle(gpu, d1, blk, N1, nb2, mode):
    """Second-level CoarsenOperator (the 12^4 -> 3^4 level of the solve included)."""
    print(f"[case] coarse-of-coarse d1={d1} N1={N1} nb2={nb2} mode={mode}", flush=True)
    from helpers import g5_links
    V1 = int(np.prod(d1))
    L1 = g5_links(d1, N1, 81)  # gamma5-hermitian, so the rebuilt forward links agree
    B = gauss(nb2 * V1 * N1, 82).reshape(nb2, V1, N1)
    cop1 = gpu.CoarseOp(d1, N1, gpu.F64)
    cop1.set_links(gpu.DeviceBuffer.from_numpy(L1), gpu.LAYOUT_FLAT, invert=False)
    agg = gpu.Aggregation(d1, blk, nb2, site_elems=N1)
    agg.set_basis(gpu.DeviceBuffer.from_numpy(B))
    agg.orthogonalise(2)
    db = gpu.DeviceBuffer(B.nbytes)
    agg.get_basis(db)
    gpu.synchronize()
    Bo = db.to_numpy(np.complex128, B.shape)
    assert rel(oracle.orthogonalise_ef(d1, blk, nb2, N1, B, 2), Bo) < 1e-12
    d2 = oracle.coarse_dims(d1, blk)
    V2, N2 = int(np.prod(d2)), 2 * nb2
    cop2 = gpu.CoarseOp(d2, N2, gpu.F64)
    gpu.coarsen_coarse_operator(cop2, cop1, agg, mode)
    d = gpu.DeviceBuffer(9 * V2 * N2 * N2 * 16)
    cop2.get_links(d, gpu.LAYOUT_FLAT)
    gpu.synchronize()
    got = d.to_numpy(np.complex128, (9, V2, N2, N2))
    ref = oracle.coarsen_coarse_operator(d1, blk, nb2, N1, L1, Bo)
    for p in range(9):
        assert rel(ref[p], got[p]) < 1e-12, f"point {p}"


CFG4_SETUP = dict(nlevels=2, block=[(4, 4, 4, 4)], nb=[32], precision=0, setup_iters=8,
                  setup_solve_iters=4)


def test_full_size_galerkin_identity_cfg4(gpu, coarsen_variant):
    """Fine 32^4, 4^4 blocks, nbasis 64 (coarse 8^4, nb = 32): the basis is
    generated by 8 seeded inverse iterations (setup rounds of BiCGSTAB,
    4 iterations each, on the Wilson-clover operator, pre- and
    post-orthonormalised, then block-orthonormalised), and D_c v = R D P v on
    random coarse vectors for the coarse operator CoarsenOperator builds from
    it; every operator the library's own, all fields generated on the device."""
    fd, blk, nb = (32, 32, 32, 32), (4, 4, 4, 4), 32
    Vf, Vc, N = 32 ** 4, 8 ** 4, 64
    print(f"[case] full-size coarsen fd={fd} nb={nb} variant={coarsen_variant}", flush=True)
    s = gpu.Stream()
    g = gpu.DeviceBuffer(4 * Vf * 9 * 16)
    gpu.random_gauge(g, fd, 0.4, 5, 1, s)
    fop = gpu.FineOp(fd, MASS, CSW)
    fop.set_gauge(g, s)
    s.sync()
    mg = gpu.Multigrid(fop, gpu.mg_params(**CFG4_SETUP))
    mg.setup()
    agg = gpu.Aggregation.borrow(mg.aggregation_handle(0), fd, blk, nb)
    assert agg.check_orthogonal().max() < 1e-20 * Vc
    cop = gpu.CoarseOp((8, 8, 8, 8), N, gpu.F64)
    e0, e1 = gpu.Event(), gpu.Event()
    e0.record(s)
    gpu.coarsen_operator(cop, fop, agg, gpu.COARSEN_DIRSAVE, s)
    e1.record(s)
    s.sync()
    print(f"  CoarsenOperator (dirsave, FP64) {e0.elapsed_ms(e1):.1f} ms", flush=True)
    vc = gpu.DeviceBuffer(Vc * N * 16)
    gpu.random_gaussian(vc, Vc * N, gpu.F64, 5, 3, s)
    lhs = gpu.DeviceBuffer(Vc * N * 16)
    cop.M(vc, lhs, s)
    f1 = gpu.DeviceBuffer(Vf * 12 * 16)
    f2 = gpu.DeviceBuffer(Vf * 12 * 16)
    rhs = gpu.DeviceBuffer(Vc * N * 16)
    agg.promote(vc, f1, gpu.F64, 1, s)
    fop.M(f1, f2, gpu.F64, 1, s)
    agg.project(f2, rhs, gpu.F64, 1, s)
    s.sync()
    a = lhs.to_numpy(np.complex128, (Vc, N))
    b = rhs.to_numpy(np.complex128, (Vc, N))
    assert rel(b, a) < 1e-12
    # the hierarchy's own coarse operator is the same build
    own = gpu.CoarseOp.borrow(mg.coarse_op_handle(1), (8, 8, 8, 8), N, gpu.F64)
    own.M(vc, lhs, s)
    s.sync()
    assert rel(b, lhs.to_numpy(np.complex128, (Vc, N))) < 1e-12
    mg.close()
