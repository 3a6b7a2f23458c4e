"""GPU parity of the multi-rhs MFMA coarse stencil against the FP64 oracle.

Per-rhs relative L2 bounds (stated against the FP64 oracle on the same
inputs, with the inputs rounded to FP32 for the FP32 path):
  FP32 MFMA                : 2e-6 * sqrt(N/8)
  FP16 MFMA, FP32 accumulate: 4e-3   (u = 2^-11, block scales set each
                                      link/vector block's max to 1)
  BF16 MFMA, FP32 accumulate: 3e-2   (u = 2^-8)
"""
import math

import numpy as np
import pytest

import oracle
from helpers import g5_links, rand_links, rand_vec, rel

pytestmark = pytest.mark.gpu

F32, F16, BF16 = 1, 2, 3
TOL = {F32: lambda N: 2e-6 * math.sqrt(N / 8), F16: lambda N: 4e-3, BF16: lambda N: 3e-2}


def run_mrhs(gpu, op, v, nrhs, mp, dag=0, half_input=False):
    V, N = op.V, op.N
    din = gpu.DeviceBuffer.from_numpy(v.astype(np.complex64))
    dout = gpu.DeviceBuffer(V * nrhs * N * 8)
    if half_input:
        q = gpu.DeviceBuffer(V * nrhs * N * 4)
        sc = gpu.DeviceBuffer(V * nrhs * 4)
        gpu.quantize(din, V * nrhs, N, mp, q, sc)
        op.M_mrhs(q, dout, nrhs, mrhs_prec=mp, in_prec=mp, in_scales=sc, dag=dag)
    else:
        op.M_mrhs(din, dout, nrhs, mrhs_prec=mp, in_prec=F32, dag=dag)
    gpu.synchronize()
    return dout.to_numpy(np.complex64, (V, nrhs, N)).astype(np.complex128)


@pytest.mark.parametrize("mp", [F32, F16, BF16])
@pytest.mark.parametrize("dims,N,nrhs", [((2, 2, 2, 2), 32, 12), ((4, 2, 2, 2), 64, 5),
                                         ((2, 4, 2, 2), 64, 24), ((2, 2, 2, 2), 32, 70)])
def test_mrhs_matches_oracle(gpu, dims, N, nrhs, mp):
    print(f"[case] mrhs dims={dims} N={N} nrhs={nrhs} prec={mp}", flush=True)
    links = rand_links(dims, N, 601)
    op = gpu.CoarseOp(dims, N, gpu.F64)
    op.set_links(gpu.DeviceBuffer.from_numpy(links), gpu.LAYOUT_FLAT, invert=False)
    v = rand_vec(dims, N, 602, nrhs=nrhs)
    vq = v.astype(np.complex64).astype(np.complex128)
    ref = oracle.coarse_mrhs_M(dims, N, nrhs, links, vq)
    got = run_mrhs(gpu, op, v, nrhs, mp)
    errs = [rel(ref[:, k], got[:, k]) for k in range(nrhs)]
    print(f"  max per-rhs rel err {max(errs):.3e}", flush=True)
    assert max(errs) < TOL[mp](N)
    if mp != F32:
        got_h = run_mrhs(gpu, op, v, nrhs, mp, half_input=True)
        assert max(rel(ref[:, k], got_h[:, k]) for k in range(nrhs)) < TOL[mp](N)


@pytest.mark.parametrize("mp", [F32, F16, BF16])
@pytest.mark.parametrize("N", [8, 16, 24, 40, 48, 56])
def test_mrhs_other_nbasis_match_oracle(gpu, N, mp):
    """nbasis values other than 32 and 64 (40 is the reference's default,
    src/Benchmark_coarse_operator.cc:46-47): the operand is zero-padded to
    16-row tiles; M and Mdag against the oracle, in one and two column passes
    (nrhs 70 > 64), with FP32 and pre-quantised half-precision inputs."""
    dims = (2, 4, 2, 2)
    for nrhs in (12, 70):
        print(f"[case] mrhs N={N} nrhs={nrhs} prec={mp}", flush=True)
        links = rand_links(dims, N, 661 + N)
        op = gpu.CoarseOp(dims, N, gpu.F64)
        op.set_links(gpu.DeviceBuffer.from_numpy(links), gpu.LAYOUT_FLAT, invert=False)
        v = rand_vec(dims, N, 662 + N, nrhs=nrhs)
        vq = v.astype(np.complex64).astype(np.complex128)
        for dag in (0, 1):
            got = run_mrhs(gpu, op, v, nrhs, mp, dag=dag)
            goth = run_mrhs(gpu, op, v, nrhs, mp, dag=dag, half_input=True) if mp != F32 else got
            for k in range(nrhs):
                r = oracle.coarse_apply(dims, N, links, np.ascontiguousarray(vq[:, k]), dag=dag)
                assert rel(r, got[:, k]) < TOL[mp](N), (dag, k)
                assert rel(r, goth[:, k]) < TOL[mp](N), (dag, k)
        op.close()


@pytest.mark.parametrize("variant", [0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10])
@pytest.mark.parametrize("mp", [F32, F16, BF16])
@pytest.mark.parametrize("N", [40])
def test_mrhs_every_kernel_variant_n40(gpu, variant, mp, N):
    """Every variant request at nbasis 40 (2, 7 and 8 run as 0, 3 and 4 there)."""
    dims, nrhs = (4, 2, 2, 2), 20
    print(f"[case] mrhs N={N} variant={variant} prec={mp}", flush=True)
    links = rand_links(dims, N, 671)
    op = gpu.CoarseOp(dims, N, gpu.F64)
    op.set_links(gpu.DeviceBuffer.from_numpy(links), gpu.LAYOUT_FLAT, invert=False)
    v = rand_vec(dims, N, 672, nrhs=nrhs)
    vq = v.astype(np.complex64).astype(np.complex128)
    gpu.call("mgc_mrhs_set_variant", variant)
    try:
        for dag in (0, 1):
            got = run_mrhs(gpu, op, v, nrhs, mp, dag=dag)
            for k in range(nrhs):
                r = oracle.coarse_apply(dims, N, links, np.ascontiguousarray(vq[:, k]), dag=dag)
                assert rel(r, got[:, k]) < TOL[mp](N), (dag, k)
    finally:
        gpu.call("mgc_mrhs_set_variant", -1)
    op.close()


@pytest.mark.parametrize("variant", [0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10])
@pytest.mark.parametrize("mp", [F32, F16, BF16])
def test_mrhs_every_kernel_variant(gpu, variant, mp):
    """Each kernel variant the autotuner may pick is checked on its own."""
    dims, N, nrhs = (4, 2, 2, 2), 64, 20
    print(f"[case] mrhs variant={variant} prec={mp}", flush=True)
    links = rand_links(dims, N, 641)
    op = gpu.CoarseOp(dims, N, gpu.F64)
    op.set_links(gpu.DeviceBuffer.from_numpy(links), gpu.LAYOUT_FLAT, invert=False)
    v = rand_vec(dims, N, 642, nrhs=nrhs)
    vq = v.astype(np.complex64).astype(np.complex128)
    ref = oracle.coarse_mrhs_M(dims, N, nrhs, links, vq)
    gpu.call("mgc_mrhs_set_variant", variant)
    try:
        for dag in (0, 1):
            got = run_mrhs(gpu, op, v, nrhs, mp, dag=dag)
            for k in range(nrhs):
                r = ref[:, k] if not dag else oracle.coarse_apply(
                    dims, N, links, np.ascontiguousarray(vq[:, k]), dag=1)
                assert rel(r, got[:, k]) < TOL[mp](N), (dag, k)
    finally:
        gpu.call("mgc_mrhs_set_variant", -1)


@pytest.mark.parametrize("variant", [7, 8, 9, 10])
@pytest.mark.parametrize("mp", [F32, F16, BF16])
@pytest.mark.parametrize("dims,N,nrhs", [((2, 2, 2, 2), 32, 70), ((2, 4, 2, 2), 64, 48),
                                         ((4, 2, 2, 2), 64, 5), ((3, 2, 2, 2), 32, 12)])
def test_mrhs_direct_lds_variants_bit_equal(gpu, dims, N, nrhs, mp, variant):
    """The direct-to-LDS forms and the XCD-block site order keep the
    arithmetic order of the LDS-staged form: outputs equal variant 0's bit
    for bit (FP16/BF16 inputs quantised, column passes with rhs0 > 0, partly
    filled tiles, and an odd extent, where the XCD-block order falls back to
    contiguous site ranges, included)."""
    print(f"[case] mrhs direct-lds dims={dims} N={N} nrhs={nrhs} prec={mp} v={variant}",
          flush=True)
    links = rand_links(dims, N, 651)
    op = gpu.CoarseOp(dims, N, gpu.F64)
    op.set_links(gpu.DeviceBuffer.from_numpy(links), gpu.LAYOUT_FLAT, invert=False)
    v = rand_vec(dims, N, 652, nrhs=nrhs)
    half = mp != F32
    try:
        for dag in (0, 1):
            gpu.call("mgc_mrhs_set_variant", 0)
            base = run_mrhs(gpu, op, v, nrhs, mp, dag=dag, half_input=half)
            gpu.call("mgc_mrhs_set_variant", variant)
            got = run_mrhs(gpu, op, v, nrhs, mp, dag=dag, half_input=half)
            assert np.array_equal(base, got), (dag, float(np.abs(base - got).max()))
    finally:
        gpu.call("mgc_mrhs_set_variant", -1)


@pytest.mark.parametrize("mp", [F32, F16])
def test_mrhs_dagger(gpu, mp):
    dims, N, nrhs = (2, 2, 2, 2), 32, 6
    links = rand_links(dims, N, 611)
    op = gpu.CoarseOp(dims, N, gpu.F64)
    op.set_links(gpu.DeviceBuffer.from_numpy(links), gpu.LAYOUT_FLAT, invert=False)
    v = rand_vec(dims, N, 612, nrhs=nrhs).astype(np.complex64).astype(np.complex128)
    got = run_mrhs(gpu, op, v, nrhs, mp, dag=1)
    for k in range(nrhs):
        ref = oracle.coarse_apply(dims, N, links, np.ascontiguousarray(v[:, k]), dag=1)
        assert rel(ref, got[:, k]) < TOL[mp](N)


def test_mrhs_equals_single_rhs_loop_fp32(gpu):
    """The batched apply equals nrhs single-rhs FP32 applies (src/Benchmark_mrhs_coarsening.cc:227-242)."""
    dims, N, nrhs = (4, 2, 2, 2), 32, 8
    links = g5_links(dims, N, 621)
    op = gpu.CoarseOp(dims, N, gpu.F32)
    op.set_links(gpu.DeviceBuffer.from_numpy(links), gpu.LAYOUT_FLAT, invert=False)
    v = rand_vec(dims, N, 622, nrhs=nrhs)
    got = run_mrhs(gpu, op, v, nrhs, F32)
    V = int(np.prod(dims))
    for k in range(nrhs):
        vk = np.ascontiguousarray(v[:, k]).astype(np.complex64)
        di = gpu.DeviceBuffer.from_numpy(vk)
        do = gpu.DeviceBuffer(V * N * 8)
        op.M(di, do)
        gpu.synchronize()
        single = do.to_numpy(np.complex64, (V, N)).astype(np.complex128)
        assert rel(single, got[:, k]) < 2e-6


@pytest.mark.parametrize("nrhs", [12, 24, 48])
def test_mrhs_full_size_cfg3(gpu, nrhs):
    """8^4 coarse lattice, nbasis 64 (the multi-rhs precision workload): every
    precision against the FP64 oracle on gamma5-hermitian synthetic links."""
    dims, N = (8, 8, 8, 8), 64
    print(f"[case] mrhs full size nrhs={nrhs}", flush=True)
    links = g5_links(dims, N, 631, herm_self=False)
    op = gpu.CoarseOp(dims, N, gpu.F64)
    op.set_links(gpu.DeviceBuffer.from_numpy(links), gpu.LAYOUT_FLAT, invert=False)
    v = rand_vec(dims, N, 632, nrhs=nrhs)
    vq = v.astype(np.complex64).astype(np.complex128)
    ref = oracle.coarse_mrhs_M(dims, N, nrhs, links, vq)
    for mp in (F32, F16, BF16):
        got = run_mrhs(gpu, op, v, nrhs, mp)
        e = max(rel(ref[:, k], got[:, k]) for k in range(nrhs))
        print(f"  prec={mp} max per-rhs rel err {e:.3e}", flush=True)
        assert e < TOL[mp](N)


def _batched_case(gpu, dims_list, N, nrhs, mp, seed, herm=None):
    ops, vs, links_l = [], [], []
    for i, dims in enumerate(dims_list):
        links = rand_links(dims, N, seed + 10 * i)
        op = gpu.CoarseOp(dims, N, gpu.F64, hermitian=(herm[i] if herm else 0))
        op.set_links(gpu.DeviceBuffer.from_numpy(links), gpu.LAYOUT_FLAT, invert=False)
        ops.append(op)
        links_l.append(links)
        vs.append(rand_vec(dims, N, seed + 10 * i + 1, nrhs=nrhs))
    ins, scs = [], []
    for op, v in zip(ops, vs):
        din = gpu.DeviceBuffer.from_numpy(v.astype(np.complex64))
        if mp == F32:
            ins.append(din)
        else:
            q = gpu.DeviceBuffer(op.V * nrhs * N * 4)
            sc = gpu.DeviceBuffer(op.V * nrhs * 4)
            gpu.quantize(din, op.V * nrhs, N, mp, q, sc)
            ins.append(q)
            scs.append(sc)
    return ops, vs, links_l, ins, (scs if mp != F32 else None)


@pytest.mark.parametrize("mp", [F32, F16, BF16])
@pytest.mark.parametrize("N,nrhs", [(32, 12), (64, 24), (32, 70), (40, 12), (24, 20)])
def test_mrhs_batched_lattices(gpu, N, nrhs, mp):
    """Lattices of uneven volume in one launch: each matches the oracle, and,
    with an LDS-staged variant forced, matches its own single-lattice launch
    bit for bit (same kernel body, same accumulation order)."""
    dims_list = [(2, 2, 2, 2), (4, 2, 2, 2), (2, 4, 2, 4)]
    print(f"[case] mrhs batched N={N} nrhs={nrhs} prec={mp}", flush=True)
    ops, vs, links_l, ins, scs = _batched_case(gpu, dims_list, N, nrhs, mp, 700)
    outs = [gpu.DeviceBuffer(op.V * nrhs * N * 8) for op in ops]
    for dag in (0, 1):
        gpu.M_mrhs_batched(ops, ins, outs, nrhs, mrhs_prec=mp, in_scales=scs, dag=dag)
        gpu.synchronize()
        for op, v, links, o, dims in zip(ops, vs, links_l, outs, dims_list):
            got = o.to_numpy(np.complex64, (op.V, nrhs, N)).astype(np.complex128)
            vq = v.astype(np.complex64).astype(np.complex128)
            for k in range(nrhs):
                r = oracle.coarse_apply(dims, N, links, np.ascontiguousarray(vq[:, k]), dag=dag)
                assert rel(r, got[:, k]) < TOL[mp](N), (dims, dag, k)
    for variant in (0, 1, 3, 4, 5, 6):  # LDS-staged variants, shared with the batched launch
        gpu.call("mgc_mrhs_set_variant", variant)
        try:
            gpu.M_mrhs_batched(ops, ins, outs, nrhs, mrhs_prec=mp, in_scales=scs)
            singles = []
            for i, op in enumerate(ops):
                so = gpu.DeviceBuffer(op.V * nrhs * N * 8)
                op.M_mrhs(ins[i], so, nrhs, mrhs_prec=mp, in_prec=mp,
                          in_scales=(scs[i] if scs else None))
                singles.append(so)
            gpu.synchronize()
            for op, o, so in zip(ops, outs, singles):
                a = o.to_numpy(np.float32, (op.V * nrhs * N * 2,))
                b = so.to_numpy(np.float32, (op.V * nrhs * N * 2,))
                assert np.array_equal(a, b), variant
        finally:
            gpu.call("mgc_mrhs_set_variant", -1)


def test_mrhs_batched_rejects_bad_batches(gpu):
    ops, vs, links_l, ins, scs = _batched_case(gpu, [(2, 2, 2, 2)], 32, 4, F32, 720)
    other = gpu.CoarseOp((2, 2, 2, 2), 64, gpu.F64)
    other.set_links(gpu.DeviceBuffer.from_numpy(rand_links((2, 2, 2, 2), 64, 721)),
                    gpu.LAYOUT_FLAT, invert=False)
    oin = gpu.DeviceBuffer(other.V * 4 * 64 * 8)
    outs = [gpu.DeviceBuffer(16 * 4 * 64 * 8) for _ in range(2)]
    with pytest.raises(gpu.MgcError):
        gpu.M_mrhs_batched([ops[0], other], [ins[0], oin], outs, 4, mrhs_prec=F32)
    with pytest.raises(gpu.MgcError):
        gpu.M_mrhs_batched(ops, ins, outs[:1], 4, mrhs_prec=F16)  # half precision needs scales
    with pytest.raises(gpu.MgcError):
        gpu.M_mrhs_batched(ops * 9, ins * 9, outs[:1] * 9, 4, mrhs_prec=F32)
