"""A caller's fine operator behind the callback interface (the reference's
LinearOperatorBase, include/CoarsenedMatrixRework.h:401-403,
include/MultigridPreconditioner.h:123-124).

* The library's own operator wrapped as a callback builds the same coarse
  links through the point-wise masked projections as the fused build does
  (relative 1e-12 per point).
* A Python callback applying a Wilson-clover operator of another mass (FP64
  only, at most 4 rhs per call) gives the oracle's coarse links for that
  operator (1e-12), in the FULL and the DIRSAVE forms.
* A multigrid hierarchy over that Python operator solves to 1e-10 (true
  residual recomputed with the operator), its FP32 hierarchy applying the
  operator in FP64 through conversions.
"""
import ctypes

import numpy as np
import pytest

import oracle
from helpers import gauss, rel

pytestmark = pytest.mark.gpu
CSW = 1.0


def fine_setup(gpu, dims, mass, seed=31):
    U = oracle.su3_gauge(dims, 0.4, seed, 1)
    fop = gpu.FineOp(dims, mass, CSW)
    fop.set_gauge(gpu.DeviceBuffer.from_numpy(U))
    return fop, U


def python_linop(gpu, fop, dims, max_nrhs=4):
    """A callback operator implemented in Python on top of the library's
    apply calls, accepting F64 only (calls record the points they saw)."""
    L = gpu.lib()
    seen = set()

    def fn(point, prec, i, o, nrhs, stream):
        seen.add(point)
        if prec != gpu.F64 or nrhs > max_nrhs:
            return -3
        h = ctypes.c_void_p(fop.handle)
        if point < 8:
            d, s = oracle.point_dir_disp(point)
            return L.mgc_fine_op_Mdir(h, ctypes.c_void_p(i), ctypes.c_void_p(o), d, s, prec, nrhs,
                                      ctypes.c_void_p(stream))
        if point == 8:
            return L.mgc_fine_op_Mdiag(h, ctypes.c_void_p(i), ctypes.c_void_p(o), prec, nrhs,
                                       ctypes.c_void_p(stream))
        return L.mgc_fine_op_M(h, ctypes.c_void_p(i), ctypes.c_void_p(o), prec, nrhs,
                               ctypes.c_void_p(stream))

    lop = gpu.make_linop(fn, dims, precisions=(gpu.F64,), max_nrhs=max_nrhs)
    lop._seen = seen
    return lop


def _links(gpu, cop, Vc, N):
    d = gpu.DeviceBuffer(9 * Vc * N * N * 16)
    cop.get_links(d, gpu.LAYOUT_FLAT)
    gpu.synchronize()
    return d.to_numpy(np.complex128, (9, Vc, N, N))


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("fd,blk,nb", [((8, 8, 8, 8), (4, 4, 4, 4), 12),
                                       ((8, 8, 8, 8), (4, 4, 4, 4), 20),
                                       ((6, 4, 2, 2), (2, 2, 2, 2), 4)])
def test_wrapped_builtin_equals_fused_build(gpu, fd, blk, nb, mode):
    print(f"[case] linop wrapped builtin fd={fd} nb={nb} mode={mode}", flush=True)
    Vf = int(np.prod(fd))
    fop, U = fine_setup(gpu, fd, -0.1)
    B = oracle.orthogonalise(fd, blk, nb, gauss(nb * Vf * 12, 801).reshape(nb, Vf, 12), 2)
    agg = gpu.Aggregation(fd, blk, nb)
    agg.set_basis(gpu.DeviceBuffer.from_numpy(B))
    cdims = oracle.coarse_dims(fd, blk)
    Vc, N = int(np.prod(cdims)), 2 * nb
    a = gpu.CoarseOp(cdims, N, gpu.F64)
    b = gpu.CoarseOp(cdims, N, gpu.F64)
    gpu.coarsen_operator(a, fop, agg, mode)
    gpu.coarsen_operator_linop(b, gpu.fine_op_linop(fop), agg, mode)
    la, lb = _links(gpu, a, Vc, N), _links(gpu, b, Vc, N)
    for p in range(9):
        assert rel(la[p], lb[p]) < 1e-12, p


@pytest.mark.parametrize("mode", [0, 1])
def test_python_operator_coarsens_to_oracle(gpu, mode):
    fd, blk, nb, mass = (4, 4, 4, 4), (2, 2, 2, 2), 6, 0.3
    print(f"[case] linop python operator mass={mass} mode={mode}", flush=True)
    Vf = int(np.prod(fd))
    fop, U = fine_setup(gpu, fd, mass, 811)
    lop = python_linop(gpu, fop, fd, max_nrhs=4)
    B = oracle.orthogonalise(fd, blk, nb, gauss(nb * Vf * 12, 812).reshape(nb, Vf, 12), 2)
    agg = gpu.Aggregation(fd, blk, nb)
    agg.set_basis(gpu.DeviceBuffer.from_numpy(B))
    cdims = oracle.coarse_dims(fd, blk)
    Vc, N = int(np.prod(cdims)), 2 * nb
    cop = gpu.CoarseOp(cdims, N, gpu.F64)
    gpu.coarsen_operator_linop(cop, lop, agg, mode)
    got = _links(gpu, cop, Vc, N)
    diag = oracle.wilson_diag(fd, U, mass, CSW)
    ref = oracle.coarsen_operator(fd, blk, nb, U, diag, B)
    for p in range(9):
        assert rel(ref[p], got[p]) < 1e-12, p
    assert lop._seen == set(range(9))


def test_callback_error_is_reported(gpu):
    fd, blk, nb = (4, 4, 4, 4), (2, 2, 2, 2), 4
    lop = gpu.make_linop(lambda *a: 7, fd, max_nrhs=2)
    Vf = int(np.prod(fd))
    agg = gpu.Aggregation(fd, blk, nb)
    agg.set_basis(gpu.DeviceBuffer.from_numpy(gauss(nb * Vf * 12, 821).reshape(nb, Vf, 12)))
    cop = gpu.CoarseOp(oracle.coarse_dims(fd, blk), 2 * nb, gpu.F64)
    with pytest.raises(gpu.MgcError):
        gpu.coarsen_operator_linop(cop, lop, agg, 0)
    assert b"callback" in gpu.lib().mgc_last_error()


@pytest.mark.parametrize("prec", [0, 1])
def test_multigrid_over_python_operator(gpu, prec):
    dims = (16, 16, 16, 16)
    V = int(np.prod(dims))
    print(f"[case] linop multigrid {dims} prec={prec}", flush=True)
    g = gpu.DeviceBuffer(4 * V * 9 * 16)
    gpu.random_gauge(g, dims, 0.4, 91, 1)
    fop = gpu.FineOp(dims, -0.1, CSW)
    fop.set_gauge(g)
    lop = python_linop(gpu, fop, dims, max_nrhs=16)
    cfg = dict(nlevels=3, block=[(4, 4, 4, 4), (2, 2, 2, 2)], nb=[8, 8], precision=prec)
    b = gpu.DeviceBuffer(V * 12 * 16)
    gpu.random_gaussian(b, V * 12, gpu.F64, 92, 1)
    its = {}
    for name, op in (("builtin", fop), ("callback", lop)):
        mg = gpu.Multigrid(op, gpu.mg_params(**cfg))
        mg.setup()
        x = gpu.DeviceBuffer(V * 12 * 16)
        st = mg.solve(b, x, tol=1e-10, maxit=200, restart=20)
        r = gpu.DeviceBuffer(V * 12 * 16)
        fop.M(x, r, gpu.F64)
        gpu.synchronize()
        res = rel(b.to_numpy(np.complex128, (V, 12)), r.to_numpy(np.complex128, (V, 12)))
        print(f"  {name}: {st} recomputed residual {res:.3e}", flush=True)
        assert st["converged"] == 1 and res <= 1.2e-10
        its[name] = st["iterations"]
        mg.close()
    assert abs(its["callback"] - its["builtin"]) <= 1, its
