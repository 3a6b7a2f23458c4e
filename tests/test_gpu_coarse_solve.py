"""Coarse solves on several lattices at once (mgc_coarse_solve_batched,
mgc_coarse_batch_*): every operator application of all of them is one
multi-lattice launch of the multi-rhs stencil.

Checks, on three lattices of uneven volume with diagonally dominant random
links (the solve converges in a few tens of iterations):
* every (lattice, rhs) column meets tol against the FP64 oracle's operator
  (||b - M x|| / ||b|| with M applied by the oracle to the returned x, bound
  tol + 2e-6 for the FP32 vectors);
* the batch gives every lattice the result of a solve of that lattice alone,
  bit for bit (each lattice runs its own lockstep GMRES, and the batched
  stencil launch is bit-equal to per-lattice launches), iteration counts
  included;
* an FP16 stencil solve to 1e-3 whose true residual, recomputed with the FP32
  stencil, is below 1e-2.
"""
import numpy as np
import pytest

import oracle
from helpers import gauss, rel

pytestmark = pytest.mark.gpu

F32, F16 = 1, 2
LATS = [(4, 4, 4, 4), (2, 2, 2, 4), (4, 2, 2, 2)]
N = 32
NRHS = 6


def dominant_links(dims, stream):
    V = int(np.prod(dims))
    L = 0.1 / np.sqrt(N) * gauss(9 * V * N * N, stream).reshape(9, V, N, N)
    L[8] += 4.0 * np.eye(N)[None, :, :]
    return L


@pytest.fixture(scope="module")
def problem(gpu):
    ops, links, bs, bdev = [], [], [], []
    for k, d in enumerate(LATS):
        L = dominant_links(d, 700 + k)
        op = gpu.CoarseOp(d, N, gpu.F64)
        op.set_links(gpu.DeviceBuffer.from_numpy(L), gpu.LAYOUT_FLAT, invert=False)
        V = int(np.prod(d))
        b = gauss(V * NRHS * N, 720 + k).reshape(V, NRHS, N).astype(np.complex64)
        ops.append(op)
        links.append(L)
        bs.append(b)
        bdev.append(gpu.DeviceBuffer.from_numpy(b))
    return ops, links, bs, bdev


def solve(gpu, ops, bdev, idx, prec, tol):
    xs = [gpu.DeviceBuffer(int(np.prod(LATS[k])) * NRHS * N * 8) for k in idx]
    st = gpu.coarse_solve_batched([ops[k] for k in idx], [bdev[k] for k in idx], xs, NRHS,
                                  mrhs_prec=prec, tol=tol, maxit=300, restart=20)
    gpu.synchronize()
    out = [x.to_numpy(np.complex64, (int(np.prod(LATS[k])), NRHS, N)) for x, k in zip(xs, idx)]
    return st, out


def test_batched_coarse_solve_fp32_meets_tolerance_and_matches_separate(gpu, problem):
    ops, links, bs, bdev = problem
    tol = 1e-6
    print("[case] batched coarse solve F32, 3 lattices x 6 rhs", flush=True)
    st, xs = solve(gpu, ops, bdev, [0, 1, 2], F32, tol)
    for k, d in enumerate(LATS):
        x = xs[k].astype(np.complex128)
        Mx = oracle.coarse_mrhs_M(d, N, NRHS, links[k], x)
        for r in range(NRHS):
            s = st[k * NRHS + r]
            res = rel(bs[k][:, r].astype(np.complex128), Mx[:, r])
            assert s["converged"] == 1, (k, r, s)
            assert res < tol + 2e-6, (k, r, res)
            assert s["true_residual"] < tol + 2e-6, (k, r, s)
    its = [s["iterations"] for s in st]
    print(f"  iterations per column {its}", flush=True)
    for k in range(len(LATS)):
        st1, x1 = solve(gpu, ops, bdev, [k], F32, tol)
        assert [s["iterations"] for s in st1] == its[k * NRHS:(k + 1) * NRHS], k
        assert np.array_equal(x1[0].view(np.float32), xs[k].view(np.float32)), k


def test_batched_coarse_solve_fp16_stencil(gpu, problem):
    ops, links, bs, bdev = problem
    print("[case] batched coarse solve F16, 3 lattices x 6 rhs", flush=True)
    st, xs = solve(gpu, ops, bdev, [0, 1, 2], F16, 1e-3)
    worst = max(s["true_residual"] for s in st)
    print(f"  worst recomputed residual {worst:.3e}", flush=True)
    assert all(s["converged"] == 1 for s in st)
    assert worst < 1e-2


def test_batched_coarse_solve_rejects_bad_batches(gpu, problem):
    ops, links, bs, bdev = problem
    x = gpu.DeviceBuffer(256 * NRHS * N * 8)
    with pytest.raises(gpu.MgcError, match="nrhs must be 1..64"):
        gpu.coarse_solve_batched([ops[0]], [bdev[0]], [x], 65)
    with pytest.raises(gpu.MgcError, match="appears twice"):
        gpu.coarse_solve_batched([ops[0], ops[0]], [bdev[0], bdev[0]], [x, x], NRHS)
    other = gpu.CoarseOp((2, 2, 2, 2), 16, gpu.F64)
    other.set_links(gpu.DeviceBuffer.from_numpy(gauss(9 * 16 * 16 * 16, 9).reshape(9, 16, 16, 16)),
                    gpu.LAYOUT_FLAT, invert=False)
    with pytest.raises(gpu.MgcError, match="same nbasis"):
        gpu.coarse_solve_batched([ops[0], other], [bdev[0], bdev[1]], [x, x], NRHS)


def test_coarse_batch_handle_reuses_state(gpu, problem):
    """A prepared batch solves twice with new right-hand sides, each time to
    the same result as the one-shot call (the kept Krylov basis and
    workspaces carry nothing over between solves)."""
    ops, links, bs, bdev = problem
    print("[case] prepared coarse batch, two solves", flush=True)
    cb = gpu.CoarseBatch(ops, NRHS, mrhs_prec=F32, restart=20)
    try:
        for rep in range(2):
            b2 = [gpu.DeviceBuffer.from_numpy((b * (1.0 + rep)).astype(np.complex64)) for b in bs]
            xs = [gpu.DeviceBuffer(int(np.prod(d)) * NRHS * N * 8) for d in LATS]
            st = cb.solve(b2, xs, tol=1e-6, maxit=300)
            ref = [gpu.DeviceBuffer(int(np.prod(d)) * NRHS * N * 8) for d in LATS]
            st1 = gpu.coarse_solve_batched(ops, b2, ref, NRHS, mrhs_prec=F32, tol=1e-6, maxit=300)
            gpu.synchronize()
            assert [s["iterations"] for s in st] == [s["iterations"] for s in st1], rep
            for k, d in enumerate(LATS):
                V = int(np.prod(d))
                a = xs[k].to_numpy(np.complex64, (V, NRHS, N))
                r = ref[k].to_numpy(np.complex64, (V, NRHS, N))
                assert np.array_equal(a.view(np.float32), r.view(np.float32)), (rep, k)
    finally:
        cb.close()


def test_batched_coarse_solve_lattice_leaving_early(gpu, problem):
    """One lattice far better conditioned than the others (its self link 40
    times the identity) converges in a couple of iterations and leaves the
    batch; the others keep batching without it.  Every lattice's result
    still equals its solve alone, bit for bit."""
    ops, links, bs, bdev = problem
    print("[case] batched coarse solve, one lattice leaving early", flush=True)
    d = LATS[1]
    V = int(np.prod(d))
    L = 0.1 / np.sqrt(N) * gauss(9 * V * N * N, 731).reshape(9, V, N, N)
    L[8] += 40.0 * np.eye(N)[None, :, :]
    easy = gpu.CoarseOp(d, N, gpu.F64)
    easy.set_links(gpu.DeviceBuffer.from_numpy(L), gpu.LAYOUT_FLAT, invert=False)
    batch_ops = [ops[0], easy, ops[2]]
    batch_b = [bdev[0], bdev[1], bdev[2]]
    xs = [gpu.DeviceBuffer(int(np.prod(LATS[k])) * NRHS * N * 8) for k in range(3)]
    st = gpu.coarse_solve_batched(batch_ops, batch_b, xs, NRHS, mrhs_prec=F32, tol=1e-6,
                                  maxit=300)
    gpu.synchronize()
    its = [max(s["iterations"] for s in st[k * NRHS:(k + 1) * NRHS]) for k in range(3)]
    print(f"  iterations per lattice {its}", flush=True)
    assert its[1] < min(its[0], its[2])
    for k in range(3):
        x1 = gpu.DeviceBuffer(int(np.prod(LATS[k])) * NRHS * N * 8)
        st1 = gpu.coarse_solve_batched([batch_ops[k]], [batch_b[k]], [x1], NRHS, mrhs_prec=F32,
                                       tol=1e-6, maxit=300)
        gpu.synchronize()
        shp = (int(np.prod(LATS[k])), NRHS, N)
        assert [s["iterations"] for s in st1] == [s["iterations"] for s in st[k * NRHS:(k + 1) * NRHS]]
        a = xs[k].to_numpy(np.complex64, shp)
        b = x1.to_numpy(np.complex64, shp)
        assert np.array_equal(a.view(np.float32), b.view(np.float32)), k


def fixed_solve(gpu, ops, bdevs, idx, prec, steps, graph, stream):
    cb = gpu.CoarseBatch([ops[k] for k in idx], NRHS, mrhs_prec=prec, stream=stream,
                         fixed_steps=steps, graph=graph)
    try:
        xs = [gpu.DeviceBuffer(int(np.prod(LATS[k])) * NRHS * N * 8) for k in idx]
        st = cb.solve([bdevs[k] for k in idx], xs, tol=1e-4)
        gpu.synchronize()
        out = [x.to_numpy(np.complex64, (int(np.prod(LATS[k])), NRHS, N)) for x, k in zip(xs, idx)]
        # a second solve on the same handle (in graph mode a replay) with other rhs
        b2 = [gpu.DeviceBuffer.from_numpy((2.0 * bdevs[k].to_numpy(np.complex64,
                                                                  (int(np.prod(LATS[k])), NRHS, N)))
                                          .astype(np.complex64)) for k in idx]
        xs2 = [gpu.DeviceBuffer(int(np.prod(LATS[k])) * NRHS * N * 8) for k in idx]
        cb.solve(b2, xs2, tol=1e-4)
        gpu.synchronize()
        out2 = [x.to_numpy(np.complex64, (int(np.prod(LATS[k])), NRHS, N)) for x, k in zip(xs2, idx)]
    finally:
        cb.close()
    return st, out, out2


def test_fixed_step_batch_graph_equals_eager_and_separate(gpu, problem):
    """The fixed-step batch (32 GMRES steps on every lattice in lockstep, one
    multi-lattice launch per step): the captured and replayed HIP graph gives
    the x of the eager form bit for bit, on its first solve and on a replay
    with other right-hand sides; each lattice gets the x of a fixed-step batch
    of that lattice alone; and 32 steps bring every column's residual, against
    the FP64 oracle's operator, below 1e-4."""
    ops, links, bs, bdev = problem
    s = gpu.Stream()
    steps = 32
    print("[case] fixed-step coarse batch F32, graph 1 and 2", flush=True)
    st1, x1, y1 = fixed_solve(gpu, ops, bdev, [0, 1, 2], F32, steps, 1, s)
    st2, x2, y2 = fixed_solve(gpu, ops, bdev, [0, 1, 2], F32, steps, 2, s)
    for k in range(3):
        assert np.array_equal(x1[k].view(np.float32), x2[k].view(np.float32)), k
        assert np.array_equal(y1[k].view(np.float32), y2[k].view(np.float32)), k
    assert all(st["iterations"] == steps for st in st2)
    worst = 0.0
    for k, d in enumerate(LATS):
        Mx = oracle.coarse_mrhs_M(d, N, NRHS, links[k], x2[k].astype(np.complex128))
        for r in range(NRHS):
            res = rel(bs[k][:, r].astype(np.complex128), Mx[:, r])
            worst = max(worst, res)
            assert res < 1e-4, (k, r, res)
            assert abs(st2[k * NRHS + r]["true_residual"] - res) < 1e-5, (k, r)
    print(f"  worst residual after {steps} steps {worst:.3e}", flush=True)
    for k in range(3):
        _, xk, _ = fixed_solve(gpu, ops, bdev, [k], F32, steps, 2, s)
        assert np.array_equal(xk[0].view(np.float32), x2[k].view(np.float32)), k


def test_fixed_step_batch_graph_fp16(gpu, problem):
    ops, links, bs, bdev = problem
    s = gpu.Stream()
    print("[case] fixed-step coarse batch F16, graph 2", flush=True)
    st, xs, _ = fixed_solve(gpu, ops, bdev, [0, 1, 2], F16, 24, 2, s)
    for k, d in enumerate(LATS):
        Mx = oracle.coarse_mrhs_M(d, N, NRHS, links[k], xs[k].astype(np.complex128))
        for r in range(NRHS):
            res = rel(bs[k][:, r].astype(np.complex128), Mx[:, r])
            assert res < 1e-2, (k, r, res)


def test_fixed_step_batch_rejects_bad_arguments(gpu, problem):
    ops, links, bs, bdev = problem
    with pytest.raises(gpu.MgcError, match="graph must be 1 or 2"):
        gpu.CoarseBatch([ops[0]], NRHS, fixed_steps=8, graph=3, stream=gpu.Stream())
    with pytest.raises(gpu.MgcError, match="null stream cannot be captured"):
        gpu.CoarseBatch([ops[0]], NRHS, fixed_steps=8, graph=2, stream=None)
