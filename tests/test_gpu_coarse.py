"""GPU parity of the single-rhs coarse operator against the FP64 oracle.

Tolerances (relative L2, ||ref-res||/||ref||):
  F64 operator: 1e-13 (the FP64 parity bound of this path is 1e-12; the
                reference's own bound 1e-15 on the squared ratio equals
                3.2e-8 here)
  F32 operator: 2e-6 * sqrt(N/8) (one FP32 rounding per multiply-add over
                9*N terms of random sign)
"""
import math

import numpy as np
import pytest

import oracle
from helpers import g5_links, half_sites, rand_links, rand_vec, rel

pytestmark = pytest.mark.gpu

CASES = [((2, 2, 2, 2), 24), ((4, 2, 2, 2), 8), ((2, 4, 2, 2), 32), ((4, 4, 2, 2), 64),
         ((2, 2, 2, 4), 4), ((4, 2, 4, 2), 12), ((2, 2, 2, 2), 48), ((4, 2, 2, 2), 16),
         ((2, 2, 2, 2), 40), ((4, 2, 2, 2), 40), ((2, 2, 2, 2), 2), ((2, 2, 2, 2), 6),
         ((2, 4, 2, 2), 18), ((2, 2, 2, 2), 62)]


def tol(prec, N):
    import mgc
    return 1e-13 if prec == mgc.F64 else 2e-6 * math.sqrt(max(N, 8) / 8)


def setup(gpu, dims, N, prec, links=None):
    if links is None:
        links = rand_links(dims, N, 101)
    op = gpu.CoarseOp(dims, N, prec)
    dl = gpu.DeviceBuffer.from_numpy(links)
    op.set_links(dl, gpu.LAYOUT_FLAT, invert=True)
    return op, links


def run(gpu, fn, vin, prec, nout_shape):
    dt = gpu.complex_dtype(prec)
    din = gpu.DeviceBuffer.from_numpy(vin.astype(dt))
    dout = gpu.DeviceBuffer(int(np.prod(nout_shape)) * np.dtype(dt).itemsize)
    fn(din, dout)
    gpu.synchronize()
    return dout.to_numpy(dt, nout_shape).astype(np.complex128)


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("dims,N", CASES)
def test_full_lattice_kinds(gpu, dims, N, prec):
    print(f"[case] full dims={dims} N={N} prec={prec}", flush=True)
    op, links = setup(gpu, dims, N, prec)
    V = int(np.prod(dims))
    v = rand_vec(dims, N, 102)
    vq = v.astype(gpu.complex_dtype(prec)).astype(np.complex128)
    t = tol(prec, N)
    cases = {
        "M": (lambda i, o: op.M(i, o), 0x1FF, 0),
        "Mdag": (lambda i, o: op.Mdag(i, o), 0x1FF, 1),
        "Mdiag": (lambda i, o: op.Mdiag(i, o), 1 << 8, 0),
        "Dhop": (lambda i, o: op.Dhop(i, o, 0), 0xFF, 0),
        "Dhopdag": (lambda i, o: op.Dhop(i, o, 1), 0xFF, 1),
    }
    for name, (fn, mask, dag) in cases.items():
        ref = oracle.coarse_apply(dims, N, links, vq, mask, dag)
        got = run(gpu, fn, v, prec, (V, N))
        assert rel(ref, got) < t, name
    for p in range(9):
        d, s = oracle.point_dir_disp(p)
        ref = oracle.coarse_apply(dims, N, links, vq, 1 << p)
        got = run(gpu, lambda i, o: op.Mdir(i, o, d, s), v, prec, (V, N))
        assert rel(ref, got) < t, f"Mdir p={p}"
    got8 = run(gpu, lambda i, o: op.MdirAll(i, o), v, prec, (8, V, N))
    for p in range(8):
        ref = oracle.coarse_apply(dims, N, links, vq, 1 << p)
        assert rel(ref, got8[p]) < t, f"MdirAll p={p}"


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("dims,N", [((2, 4, 2, 2), 40), ((4, 2, 2, 2), 64)])
def test_hermitian_operator_dagger_is_M(gpu, dims, N, prec):
    """Petrov-Galerkin operators (hermitian = 1, include/CoarsenedMatrixUpstream.h:609-621):
    Mdag and the multi-rhs dag apply return M unchanged, while Dhop(dag) keeps
    the G5C sandwich for either kind, as the reference's DhopInternal does
    (include/CoarsenedMatrixRework.h:293-300)."""
    print(f"[case] hermitian dims={dims} N={N} prec={prec}", flush=True)
    links = rand_links(dims, N, 131)
    op = gpu.CoarseOp(dims, N, prec, hermitian=1)
    op.set_links(gpu.DeviceBuffer.from_numpy(links), gpu.LAYOUT_FLAT, invert=False)
    V = int(np.prod(dims))
    v = rand_vec(dims, N, 132)
    vq = v.astype(gpu.complex_dtype(prec)).astype(np.complex128)
    t = tol(prec, N)
    ref = oracle.coarse_apply(dims, N, links, vq)
    assert rel(ref, run(gpu, lambda i, o: op.Mdag(i, o), v, prec, (V, N))) < t
    refh = oracle.coarse_apply(dims, N, links, vq, 0xFF, dag=1)
    assert rel(refh, run(gpu, lambda i, o: op.Dhop(i, o, 1), v, prec, (V, N))) < t
    if prec == 1:
        nrhs = 5
        vm = rand_vec(dims, N, 133, nrhs=nrhs).astype(np.complex64)
        din = gpu.DeviceBuffer.from_numpy(vm)
        dout = gpu.DeviceBuffer(V * nrhs * N * 8)
        op.M_mrhs(din, dout, nrhs, mrhs_prec=gpu.F32, dag=1)
        gpu.synchronize()
        got = dout.to_numpy(np.complex64, (V, nrhs, N)).astype(np.complex128)
        for k in range(nrhs):
            r = oracle.coarse_apply(dims, N, links, vm[:, k].astype(np.complex128))
            assert rel(r, got[:, k]) < 2e-6 * math.sqrt(N / 8), k
    op.close()


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("dims,N", [((4, 2, 2, 2), 8), ((2, 2, 4, 2), 32), ((4, 2, 2, 2), 64),
                                    ((2, 2, 2, 4), 40)])
def test_checkerboarded_kinds(gpu, dims, N, prec):
    print(f"[case] cb dims={dims} N={N} prec={prec}", flush=True)
    links = rand_links(dims, N, 111)
    links[8] += 4.0 * np.eye(N)  # well-conditioned self links
    op, _ = setup(gpu, dims, N, prec, links)
    H = int(np.prod(dims)) // 2
    v = rand_vec(dims, N, 112)
    vq = v.astype(gpu.complex_dtype(prec)).astype(np.complex128)
    t = tol(prec, N)
    inv = oracle.invert_self(dims, N, links)
    for in_cb in (0, 1):
        out_cb = 1 - in_cb
        vin = np.ascontiguousarray(v[half_sites(dims, in_cb)])
        vinq = np.ascontiguousarray(vq[half_sites(dims, in_cb)])
        for dag in (0, 1):
            ref = oracle.coarse_apply_cb(dims, N, links, vinq, out_cb, 0xFF, dag)
            fn = op.DhopOE if in_cb == 0 else op.DhopEO
            got = run(gpu, lambda i, o: fn(i, o, dag), vin, prec, (H, N))
            assert rel(ref, got) < t, f"Dhop cb in={in_cb} dag={dag}"
            mfn = op.MeooeDag if dag else op.Meooe
            got = run(gpu, lambda i, o: mfn(i, o, in_cb), vin, prec, (H, N))
            assert rel(ref, got) < t, f"Meooe in={in_cb} dag={dag}"
    for cb in (gpu.FULL, gpu.EVEN, gpu.ODD):
        if cb == gpu.FULL:
            vin, vinq, n = v, vq, 2 * H
        else:
            vin = np.ascontiguousarray(v[half_sites(dims, cb)])
            vinq = np.ascontiguousarray(vq[half_sites(dims, cb)])
            n = H
        for name, mats, adj in (("Mooee", links[8], 0), ("MooeeDag", links[8], 1),
                                ("MooeeInv", inv, 0), ("MooeeInvDag", inv, 1)):
            ref = oracle.site_matrix_apply(dims, N, mats, vinq, cb, adj)
            got = run(gpu, lambda i, o: getattr(op, name)(i, o, cb), vin, prec, (n, N))
            assert rel(ref, got) < (t if "Inv" not in name else 10 * t), f"{name} cb={cb}"


@pytest.mark.parametrize("N", [8, 24, 64])
def test_self_inverse_matches_oracle(gpu, N):
    dims = (2, 2, 2, 2)
    links = rand_links(dims, N, 121)
    op, _ = setup(gpu, dims, N, gpu.F64, links)
    V = 16
    d = gpu.DeviceBuffer(V * N * N * 16)
    op.get_self_inverse(d)
    got = d.to_numpy(np.complex128, (V, N, N))
    ref = oracle.invert_self(dims, N, links)
    assert rel(ref, got) < 1e-11
    # and MooeeInv * Mooee = 1 through the library
    v = rand_vec(dims, N, 122)
    w = run(gpu, lambda i, o: op.Mooee(i, o), v, gpu.F64, (V, N))
    back = run(gpu, lambda i, o: op.MooeeInv(i, o), w, gpu.F64, (V, N))
    assert rel(v, back) < 1e-11


def test_singular_self_link_is_an_error(gpu):
    dims, N = (2, 2, 2, 2), 8
    links = rand_links(dims, N, 131)
    links[8, 3] = 0.0  # one site's self link is zero
    op = gpu.CoarseOp(dims, N, gpu.F64)
    with pytest.raises(gpu.MgcError) as e:
        op.set_links(gpu.DeviceBuffer.from_numpy(links), gpu.LAYOUT_FLAT, invert=True)
    assert e.value.code == -6


def test_split_layout_round_trip(gpu):
    dims, N = (2, 2, 2, 2), 8
    V = 16
    links = rand_links(dims, N, 141)
    split = oracle.links_flat_to_split(V, N, links)
    op = gpu.CoarseOp(dims, N, gpu.F64)
    op.set_links(gpu.DeviceBuffer.from_numpy(split), gpu.LAYOUT_SPLIT, invert=False)
    d = gpu.DeviceBuffer(9 * V * N * N * 16)
    op.get_links(d, gpu.LAYOUT_FLAT)
    assert np.array_equal(d.to_numpy(np.complex128, (9, V, N, N)), links)
    op.get_links(d, gpu.LAYOUT_SPLIT)
    assert np.array_equal(d.to_numpy(np.complex128, (9 * V * N * N,)), split)


def test_gaussian_generator_matches_oracle(gpu):
    n = 4099
    d = gpu.DeviceBuffer(n * 16)
    gpu.random_gaussian(d, n, gpu.F64, 77, 5)
    got = d.to_numpy(np.complex128, (n,))
    ref = oracle.gaussian(n, 77, 5)
    assert np.abs(got - ref).max() < 1e-13


@pytest.mark.parametrize("prec", [0, 1])
def test_full_size_single_rhs_cfg2(gpu, prec):
    """8^4 coarse lattice, nbasis 32 (the single-rhs drop-in workload), M and Mdag."""
    dims, N = (8, 8, 8, 8), 32
    print(f"[case] full-size dims={dims} N={N} prec={prec}", flush=True)
    links = g5_links(dims, N, 151, herm_self=False)
    op, _ = setup(gpu, dims, N, prec, links)
    V = 4096
    v = rand_vec(dims, N, 152)
    vq = v.astype(gpu.complex_dtype(prec)).astype(np.complex128)
    for name, dag in (("M", 0), ("Mdag", 1)):
        ref = oracle.coarse_apply(dims, N, links, vq, 0x1FF, dag)
        got = run(gpu, lambda i, o: getattr(op, name)(i, o), v, prec, (V, N))
        assert rel(ref, got) < tol(prec, N), name


def test_full_size_single_rhs_n64_f64(gpu):
    """8^4 coarse lattice, nbasis 64, FP64 (the single-rhs reference point of the multi-rhs case)."""
    dims, N = (8, 8, 8, 8), 64
    print(f"[case] full-size dims={dims} N={N} prec=0", flush=True)
    links = g5_links(dims, N, 161, herm_self=False)
    op, _ = setup(gpu, dims, N, gpu.F64, links)
    v = rand_vec(dims, N, 162)
    ref = oracle.coarse_apply(dims, N, links, v)
    got = run(gpu, lambda i, o: op.M(i, o), v, gpu.F64, (4096, N))
    assert rel(ref, got) < 1e-13
