"""The no-allocation multi-rhs apply (mgc_coarse_op_M_mrhs_noalloc) against
the allocating one (mgc_coarse_op_M_mrhs) on the same operator and FP32
inputs: bit-equal outputs (same quantiser, same tuned kernel variant), on two
calls in a row where the second has a smaller input, plus its error paths.
"""
import ctypes

import numpy as np
import pytest

import oracle
from helpers import rand_links, rand_vec, rel

pytestmark = pytest.mark.gpu

F32, F16, BF16 = 1, 2, 3


def _apply_pair(gpu, op, v, nrhs, mp, dag):
    V, N = op.V, op.N
    din = gpu.DeviceBuffer.from_numpy(v.astype(np.complex64))
    ref = gpu.DeviceBuffer(V * nrhs * N * 8)
    got = gpu.DeviceBuffer(V * nrhs * N * 8)
    op.M_mrhs(din, ref, nrhs, mrhs_prec=mp, in_prec=F32, dag=dag)  # tunes the shape
    op.M_mrhs_noalloc(din, got, nrhs, mrhs_prec=mp, dag=dag)
    gpu.synchronize()
    a = ref.to_numpy(np.complex64, (V, nrhs, N))
    b = got.to_numpy(np.complex64, (V, nrhs, N))
    return a, b


@pytest.mark.parametrize("mp", [F16, BF16, F32])
def test_noalloc_equals_allocating_apply_twice(gpu, mp):
    dims, N = (4, 4, 2, 2), 64
    links = rand_links(dims, N, 901)
    op = gpu.CoarseOp(dims, N, gpu.F64)
    op.set_links(gpu.DeviceBuffer.from_numpy(links), gpu.LAYOUT_FLAT, invert=False)
    op.prepare_mrhs(mp)
    for call_no, (nrhs, dag) in enumerate([(24, 0), (7, 1)]):
        print(f"[case] noalloc mrhs dims={dims} N={N} prec={mp} call={call_no} nrhs={nrhs} "
              f"dag={dag}", flush=True)
        v = rand_vec(dims, N, 902 + call_no, nrhs=nrhs)
        a, b = _apply_pair(gpu, op, v, nrhs, mp, dag)
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), "not bit-equal"
        # and the result is the stencil's (FP16/BF16/FP32 bounds of the multi-rhs tests)
        vq = v.astype(np.complex64).astype(np.complex128)
        tol = {F32: 4e-6, F16: 4e-3, BF16: 3e-2}[mp]
        for k in range(nrhs):
            r = oracle.coarse_apply(dims, N, links, np.ascontiguousarray(vq[:, k]), dag=dag)
            assert rel(r, b[:, k].astype(np.complex128)) < tol, (call_no, k)
    op.close()


def test_noalloc_rejects_unprepared_operand_and_oversize_input(gpu):
    dims, N = (2, 2, 2, 2), 32
    links = rand_links(dims, N, 911)
    op = gpu.CoarseOp(dims, N, gpu.F64)
    op.set_links(gpu.DeviceBuffer.from_numpy(links), gpu.LAYOUT_FLAT, invert=False)
    buf = gpu.DeviceBuffer(16 * 8 * N * 8)
    L = gpu.lib()
    # operand not prepared for F16: a state error before any launch
    st = L.mgc_coarse_op_M_mrhs_noalloc(ctypes.c_void_p(op.handle), ctypes.c_void_p(buf.ptr),
                                        ctypes.c_void_p(buf.ptr + 8), 8, F16, 0, None)
    assert st == -5
    op.prepare_mrhs(F16)
    gpu.synchronize()
    # 16 sites x 1e6 rhs x (4*32 + 4) bytes is far beyond the 64 MiB scratch:
    # rejected on the host before anything is launched
    st = L.mgc_coarse_op_M_mrhs_noalloc(ctypes.c_void_p(op.handle), ctypes.c_void_p(buf.ptr),
                                        ctypes.c_void_p(buf.ptr + 8), 1000000, F16, 0, None)
    assert st == -2
    assert b"64 MiB" in L.mgc_last_error()
    op.close()
