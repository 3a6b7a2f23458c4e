"""Multigrid solves at the sizes of the batched mixed-precision solve
configuration (fine 48^4 -> 12^4 -> 3^4, nbasis 32/32, Wilson-clover
m = -0.1, csw = 1.0, SU(3) field with eps = 0.4).

Acceptance follows src/Test_wilsonclover_mg.cc:121-151: FGMRES(20)
preconditioned by the K-cycle hierarchy to a relative residual of 1e-10.
The true residual ||b - D x|| / ||b|| is recomputed here with the FP64 fine
operator (accepted up to 1.2e-10 for the final rounding of x), and the FP32
("mixed") hierarchy must not take more outer iterations than the FP64 plain
form of the same hierarchy.

The batched solve of 12 right-hand sides runs on 36^4 -> 9^4 -> 3^4: the
largest lattice of this family whose FP64 outer FGMRES(20) fits one GPU with
12 rhs (43 batched FP64 vectors of 3.9 GB, ~0.2 TB in all; at 40^4 they take
0.25 TB and at 48^4 0.5 TB, beyond the 288 GB of HBM).
"""
import numpy as np
import pytest

from helpers import gauss, rel

pytestmark = pytest.mark.gpu
MASS, CSW, EPS = -0.1, 1.0, 0.4
SEED = 1 | (2 << 16) | (3 << 32) | (4 << 48)


def make_fine(gpu, dims, stream_id):
    V = int(np.prod(dims))
    g = gpu.DeviceBuffer(4 * V * 9 * 16)
    gpu.random_gauge(g, dims, EPS, SEED, stream_id)
    fop = gpu.FineOp(dims, MASS, CSW)
    fop.set_gauge(g)
    g.close()
    return fop


def true_residual(gpu, fop, b, x, V, nrhs=1):
    r = gpu.DeviceBuffer(V * nrhs * 12 * 16)
    fop.M(x, r, gpu.F64, nrhs=nrhs)
    gpu.synchronize()
    Dx = r.to_numpy(np.complex128, (V, nrhs, 12))
    bb = b.to_numpy(np.complex128, (V, nrhs, 12))
    r.close()
    return [rel(bb[:, k], Dx[:, k]) for k in range(nrhs)]


def test_cfg5_fp64_plain_and_mixed_reach_1e10(gpu):
    dims = (48, 48, 48, 48)
    V = int(np.prod(dims))
    print(f"[case] mg solve {dims} -> 12^4 -> 3^4 nbasis 32/32", flush=True)
    fop = make_fine(gpu, dims, 21)
    b = gpu.DeviceBuffer(V * 12 * 16)
    gpu.random_gaussian(b, V * 12, gpu.F64, SEED, 22)
    its = {}
    for prec in (gpu.F64, gpu.F32):
        p = gpu.mg_params(nlevels=3, block=[(4, 4, 4, 4), (4, 4, 4, 4)], nb=[16, 16],
                          precision=prec)
        mg = gpu.Multigrid(fop, p)
        mg.setup()
        x = gpu.DeviceBuffer(V * 12 * 16)
        st = mg.solve(b, x, tol=1e-10, maxit=500, restart=20)
        res = true_residual(gpu, fop, b, x, V)[0]
        print(f"  prec={prec}: {st} recomputed residual {res:.3e}", flush=True)
        assert st["converged"] == 1
        assert res <= 1.2e-10
        its[prec] = st["iterations"]
        x.close()
        mg.close()
    assert its[gpu.F32] <= its[gpu.F64], its
    b.close()
    fop.close()


def test_batched_12rhs_largest_lattice(gpu):
    dims, nrhs = (36, 36, 36, 36), 12
    V = int(np.prod(dims))
    print(f"[case] batched mg solve {dims} -> 9^4 -> 3^4 nbasis 32/32 nrhs={nrhs} f16 coarse",
          flush=True)
    fop = make_fine(gpu, dims, 31)
    p = gpu.mg_params(nlevels=3, block=[(4, 4, 4, 4), (3, 3, 3, 3)], nb=[16, 16],
                      precision=gpu.F32)
    mg = gpu.Multigrid(fop, p)
    mg.setup()
    b = gpu.DeviceBuffer(V * nrhs * 12 * 16)
    gpu.random_gaussian(b, V * nrhs * 12, gpu.F64, SEED, 32)
    x = gpu.DeviceBuffer(V * nrhs * 12 * 16)
    sts = mg.solve_mrhs(b, x, nrhs, coarse_prec=gpu.F16, tol=1e-10, maxit=500, restart=20)
    res = true_residual(gpu, fop, b, x, V, nrhs)
    print(f"  iterations {[s['iterations'] for s in sts]} max residual {max(res):.3e} "
          f"solve {sts[0]['seconds']:.2f} s", flush=True)
    assert all(s["converged"] == 1 for s in sts)
    assert max(res) <= 1.2e-10
    # rhs 0 alone with the same hierarchy: the batch takes about as many iterations
    bh = b.to_numpy(np.complex128, (V, nrhs, 12))
    b.close()
    x.close()
    b0 = gpu.DeviceBuffer.from_numpy(np.ascontiguousarray(bh[:, 0]))
    x0 = gpu.DeviceBuffer(V * 12 * 16)
    st0 = mg.solve(b0, x0, tol=1e-10, maxit=500, restart=20)
    print(f"  rhs 0 alone: {st0}", flush=True)
    assert sts[0]["iterations"] <= st0["iterations"] + 3
    mg.close()
    fop.close()
