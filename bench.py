#!/usr/bin/env python3
"""Benchmark of the multigrid coarse-level hot path on one MI355X.

A step is one application of the coarse Dirac stencil to one batch of
synthetic right-hand sides (inputs already resident in HBM).  The headline
(``value``) is rhs*sites/s of the headline leg; every other leg is reported
under ``legs``.  ``roofline`` is computed from the algorithmic bytes/flops of
the headline kernel and its HIP-event-timed launch duration; ``cpu_baseline``
is the FP64 CPU oracle timed on the host cores on the same workload.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (os.path.join(ROOT, "grid-mg-bench_amd", "python"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402

METRIC = ("coarse-stencil rhs·sites/s (MRHS FP16/FP32 vs FP64) "
          "+ MG solve s @ res 1e-10")
HBM_PEAK_GBS = 8000.0
PEAK_TFLOPS = {"f64": 78.6, "f32": 157.3, "f16": 2500.0, "bf16": 2500.0}
SEED = 1 | (2 << 16) | (3 << 32) | (4 << 48)


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


class Timer:
    """Times K launches of `step` on `stream` with HIP events and the host clock."""

    def __init__(self, mgc, stream):
        self.mgc, self.stream = mgc, stream

    def run(self, step, steps: int, warmup: int):
        mgc = self.mgc
        for _ in range(warmup):
            step()
        mgc.synchronize()
        e0, e1 = mgc.Event(), mgc.Event()
        t0 = time.perf_counter()
        e0.record(self.stream)
        for _ in range(steps):
            step()
        e1.record(self.stream)
        mgc.synchronize()
        t1 = time.perf_counter()
        return (t1 - t0) / steps, e0.elapsed_ms(e1) * 1e-3 / steps


def single_rhs_leg(mgc, stream, dims, N, prec, steps, warmup):
    """Coarse M, one rhs: out(x) = sum_p A_p(x) in(x+d_p)."""
    V = int(np.prod(dims))
    wl = 16 if prec == mgc.F64 else 8
    links = mgc.DeviceBuffer(9 * V * N * N * 16)
    mgc.random_gaussian(links, 9 * V * N * N, mgc.F64, SEED, 1, stream)
    op = mgc.CoarseOp(dims, N, prec)
    op.set_links(links, mgc.LAYOUT_FLAT, invert=False, stream=stream)
    vin = mgc.DeviceBuffer(V * N * wl)
    vout = mgc.DeviceBuffer(V * N * wl)
    mgc.random_gaussian(vin, V * N, prec, SEED, 2, stream)
    stream.sync()
    t_host, t_dev = Timer(mgc, stream).run(lambda: op.M(vin, vout, stream), steps, warmup)
    bytes_alg = 9 * V * N * N * wl + V * N * 2 * wl
    flops_alg = 8 * 9 * V * N * N
    res = {
        "workload": f"coarse M single-rhs {'x'.join(map(str, dims))} nbasis={N} "
                    f"{'fp64' if prec == mgc.F64 else 'fp32'}",
        "rhs_sites_per_s": V / t_dev,
        "ms_per_step": t_host * 1e3,
        "kernel_us": t_dev * 1e6,
        "GBps": bytes_alg / t_dev / 1e9,
        "hbm_frac": bytes_alg / t_dev / 1e9 / HBM_PEAK_GBS,
        "bytes_per_launch": bytes_alg,
        "flops_per_launch": flops_alg,
        "dtype": "f64" if prec == mgc.F64 else "f32",
    }
    for b in (links, vin, vout):
        b.close()
    op.close()
    return res


def transfer_leg(mgc, stream, fine_dims, blk, nb, prec, steps, warmup):
    """Aggregation project (R) and promote (P), one rhs: the block panel of the
    basis [block][site][e][nb] is streamed once per call."""
    Vf = int(np.prod(fine_dims))
    Vc = Vf // int(np.prod(blk))
    N = 2 * nb
    wl = 16 if prec == mgc.F64 else 8
    agg = mgc.Aggregation(fine_dims, blk, nb)
    basis = mgc.DeviceBuffer(nb * Vf * 12 * 16)
    mgc.random_gaussian(basis, nb * Vf * 12, mgc.F64, SEED, 11, stream)
    agg.set_basis(basis, stream)
    agg.orthogonalise(1, stream)
    vf = mgc.DeviceBuffer(Vf * 12 * wl)
    vc = mgc.DeviceBuffer(Vc * N * wl)
    mgc.random_gaussian(vf, Vf * 12, prec, SEED, 12, stream)
    mgc.random_gaussian(vc, Vc * N, prec, SEED, 13, stream)
    stream.sync()
    bytes_alg = Vf * 12 * nb * wl + Vf * 12 * wl + Vc * N * wl
    res = {"workload": f"Aggregation {'x'.join(map(str, fine_dims))} blocks "
                       f"{'x'.join(map(str, blk))} nbasis={N} ({nb} vectors x 2 chiralities) "
                       f"{'fp64' if prec == mgc.F64 else 'fp32'}",
           "bytes_per_call": bytes_alg}
    for name, fn in (("project", lambda: agg.project(vf, vc, prec, 1, stream)),
                     ("promote", lambda: agg.promote(vc, vf, prec, 1, stream))):
        _, t_dev = Timer(mgc, stream).run(fn, steps, warmup)
        res[name] = {"us": t_dev * 1e6, "GBps": bytes_alg / t_dev / 1e9,
                     "hbm_frac": bytes_alg / t_dev / 1e9 / HBM_PEAK_GBS}
    for b in (basis, vf, vc):
        b.close()
    agg.close()
    return res


def fine_op_leg(mgc, stream, dims, prec, nrhs, steps, warmup):
    """Wilson-clover apply: gauge 4 links and packed hermitian clover (2 x 18
    complex) per site read once, spinors in and out per rhs."""
    V = int(np.prod(dims))
    w = 16 if prec == mgc.F64 else 8
    g = mgc.DeviceBuffer(4 * V * 9 * 16)
    mgc.random_gauge(g, dims, 0.4, SEED, 21, stream)
    fop = mgc.FineOp(dims, -0.1, 1.0)
    fop.set_gauge(g, stream)
    vi = mgc.DeviceBuffer(V * nrhs * 12 * w)
    vo = mgc.DeviceBuffer(V * nrhs * 12 * w)
    mgc.random_gaussian(vi, V * nrhs * 12, prec, SEED, 22, stream)
    stream.sync()
    _, t_dev = Timer(mgc, stream).run(lambda: fop.M(vi, vo, prec, nrhs, stream), steps, warmup)
    bytes_alg = V * (36 * w + 36 * w + 8 * 4 + nrhs * 2 * 12 * w)
    res = {"workload": f"Wilson-clover M {'x'.join(map(str, dims))} nrhs={nrhs} "
                       f"{'fp64' if prec == mgc.F64 else 'fp32'}",
           "kernel_us": t_dev * 1e6, "GBps": bytes_alg / t_dev / 1e9,
           "hbm_frac": bytes_alg / t_dev / 1e9 / HBM_PEAK_GBS, "bytes_per_launch": bytes_alg}
    for b in (g, vi, vo):
        b.close()
    fop.close()
    return res


def coarsen_leg(mgc, stream, fine_dims, blk, nb, reps=3):
    """The nbasis-2nb construction workload: basis vectors from 8 seeded inverse
    iterations (a 2-level hierarchy's setup: 8 rounds of BiCGSTAB, 4 iterations
    each, on the Wilson-clover operator, FP64, pre- and post-orthonormalised),
    then

    * the block orthonormalisation of that basis: Cholesky QR (two Gram rounds
      and one apply, 3 reads + 1 write of the panel) against the column sweep;
      algorithmic bytes = one read and one write of the panel;
    * CoarsenOperator (FP64, backward points + self, remaining links by
      gamma5-hermiticity), MFMA and vector-ALU builds.  Contraction flops per
      coarse site: every site of each computed point's face (the whole block
      for self) contributes 4 nb^2 link entries x 6 spin-colour terms x 8 flops.
    """
    Vf = int(np.prod(fine_dims))
    cd = tuple(f // b for f, b in zip(fine_dims, blk))
    Vc = int(np.prod(cd))
    Vb = int(np.prod(blk))
    N = 2 * nb
    g = mgc.DeviceBuffer(4 * Vf * 9 * 16)
    mgc.random_gauge(g, fine_dims, 0.4, SEED, 31, stream)
    fop = mgc.FineOp(fine_dims, -0.1, 1.0)
    fop.set_gauge(g, stream)
    stream.sync()
    p = mgc.mg_params(nlevels=2, block=[blk], nb=[nb], precision=mgc.F64, setup_iters=8,
                      setup_solve_iters=4)
    mg = mgc.Multigrid(fop, p)
    t0 = time.perf_counter()
    mg.setup(stream)
    t_setup = time.perf_counter() - t0
    src = mgc.Aggregation.borrow(mg.aggregation_handle(0), fine_dims, blk, nb)
    basis = mgc.DeviceBuffer(nb * Vf * 12 * 16)
    src.get_basis(basis, stream)
    agg = mgc.Aggregation(fine_dims, blk, nb)
    e0, e1 = mgc.Event(), mgc.Event()
    panel_bytes = nb * Vf * 12 * 16
    orth = {}
    for method, name in ((0, "cholqr"), (1, "sweep")):
        mgc.call("mgc_aggregation_set_ortho_method", method)
        agg.set_basis(basis, stream)
        e0.record(stream)
        agg.orthogonalise(2, stream)
        e1.record(stream)
        e1.sync()
        t = e0.elapsed_ms(e1) * 1e-3
        orth[name] = {"ms": t * 1e3, "GBps": 2 * panel_bytes / t / 1e9,
                      "hbm_frac": 2 * panel_bytes / t / 1e9 / HBM_PEAK_GBS}
    mgc.call("mgc_aggregation_set_ortho_method", 0)
    agg.set_basis(basis, stream)
    agg.orthogonalise(2, stream)
    face_sites = sum(Vb // b for b in blk) + Vb
    flops = Vc * face_sites * 4 * nb * nb * 6 * 8
    cop = mgc.CoarseOp(cd, N, mgc.F64)
    build = {}
    for variant, name in ((1, "mfma"), (0, "vector")):
        mgc.call("mgc_coarsen_set_variant", variant)
        mgc.coarsen_operator(cop, fop, agg, mgc.COARSEN_DIRSAVE, stream)  # warm-up
        e0.record(stream)
        for _ in range(reps):
            mgc.coarsen_operator(cop, fop, agg, mgc.COARSEN_DIRSAVE, stream)
        e1.record(stream)
        e1.sync()
        t = e0.elapsed_ms(e1) * 1e-3 / reps
        build[name] = {"ms": t * 1e3, "TFLOPs": flops / t / 1e12,
                       "fp64_frac": flops / t / 1e12 / PEAK_TFLOPS["f64"]}
    mgc.call("mgc_coarsen_set_variant", -1)
    res = {"workload": f"CoarsenOperator {'x'.join(map(str, fine_dims))} -> "
                       f"{'x'.join(map(str, cd))} nbasis={N} fp64 (backward points + self), "
                       f"basis from 8 inverse iterations",
           "basis_setup_s": t_setup,
           "flops": flops,
           "ms": build["mfma"]["ms"], "TFLOPs": build["mfma"]["TFLOPs"],
           "fp64_frac": build["mfma"]["fp64_frac"],
           "build": build,
           "orthogonalise": orth,
           "orthogonalise_panel_bytes": panel_bytes}
    mg.close()
    for b in (g, basis):
        b.close()
    cop.close()
    agg.close()
    fop.close()
    return res


# HBM bytes moved per launch by the headline kernel on the headline workload,
# per autotuned variant (F16, N=64, 6 column tiles): two rocprofv3 --pmc
# runs, FETCH_SIZE then WRITE_SIZE, each scaled by what the counter reads for
# a 1 GiB copy kernel (FETCH x 2.000, WRITE x 1.000).  Each figure is the mean
# of the variant's two autotuner trials in those runs (the tuner times every
# variant in dispatch order, twice, six launches a trial), so variants that
# share a kernel are still told apart.  Variant 9 (the one the tuner picks)
# was read again at commit 0f953b4 with the variant forced (852.9 MB, within
# 0.07% of the first reading), and variant 10 with it forced when it was added.  The command lines and readings are in
# profiles/k_mrhs_f16_8x8x8x8_n64_nrhs48.md.
HEADLINE_TRAFFIC_BYTES = {0: 987.5e6, 1: 898.2e6, 2: 996.4e6, 3: 1023.9e6, 4: 925.4e6,
                          5: 987.6e6, 6: 921.8e6, 7: 1008.1e6, 8: 921.7e6, 9: 852.9e6,
                          10: 849.5e6}

MRHS_KERNELS = {0: "k_mrhs (LDS-staged B')",
                1: "k_mrhs (LDS-staged B', non-temporal link loads)",
                2: "k_mrhs_reg (register B', 2 m-tiles per wavefront)",
                3: "k_mrhs (LDS-staged B', XCD-contiguous sites)",
                4: "k_mrhs (LDS-staged B', non-temporal link loads, XCD-contiguous sites)",
                5: "k_mrhs_persistent (LDS-staged B')",
                6: "k_mrhs_persistent (LDS-staged B', non-temporal link loads, "
                   "XCD-contiguous sites)",
                7: "k_mrhs_dma (input copied to LDS, 2 m-tiles per wavefront)",
                8: "k_mrhs_dma (input copied to LDS, 2 m-tiles per wavefront, non-temporal "
                   "link loads, XCD-contiguous sites)",
                9: "k_mrhs_xblk (LDS-staged B', non-temporal link loads, XCD-block sites)",
                10: "k_mrhs_xblk2 (as 9, input and link loads two points ahead)"}


def mrhs_leg(mgc, stream, dims, N, nrhs, mp, steps, warmup, half_input=True, keep=None):
    """Multi-rhs coarse M on MFMA: one launch applies M to nrhs right-hand sides.
    half_input: the FP16/BF16 inputs are already quantised (+ per (site, rhs)
    FP32 scales); otherwise FP32 inputs are quantised inside the step.
    keep: a dict that receives host copies of the links, the FP32 inputs and
    the output (for the CPU-baseline leg, which checks them)."""
    V = int(np.prod(dims))
    names = {mgc.F32: "f32", mgc.F16: "f16", mgc.BF16: "bf16"}
    links = mgc.DeviceBuffer(9 * V * N * N * 16)
    mgc.random_gaussian(links, 9 * V * N * N, mgc.F64, SEED, 11, stream)
    op = mgc.CoarseOp(dims, N, mgc.F64)
    op.set_links(links, mgc.LAYOUT_FLAT, invert=False, stream=stream)
    op.prepare_mrhs(mp, stream)
    vin = mgc.DeviceBuffer(V * nrhs * N * 8)
    mgc.random_gaussian(vin, V * nrhs * N, mgc.F32, SEED, 12, stream)
    vout = mgc.DeviceBuffer(V * nrhs * N * 8)
    wl = 8 if mp == mgc.F32 else 4
    use_half = half_input and mp != mgc.F32
    if use_half:
        q = mgc.DeviceBuffer(V * nrhs * N * 4)
        qs = mgc.DeviceBuffer(V * nrhs * 4)
        mgc.quantize(vin, V * nrhs, N, mp, q, qs, stream)

        def step():
            op.M_mrhs(q, vout, nrhs, mrhs_prec=mp, in_prec=mp, in_scales=qs, stream=stream)
        bytes_in = V * nrhs * N * 4 + V * nrhs * 4
    else:
        def step():
            op.M_mrhs(vin, vout, nrhs, mrhs_prec=mp, in_prec=mgc.F32, stream=stream)
        bytes_in = V * nrhs * N * 8
    stream.sync()
    t_host, t_dev = Timer(mgc, stream).run(step, steps, warmup)
    variant = ctypes.c_int(-1)
    mgc.call("mgc_mrhs_get_variant", mp, N, nrhs, ctypes.byref(variant))
    bytes_alg = 9 * V * N * N * wl + (0 if mp == mgc.F32 else 9 * V * 4) + bytes_in + \
        V * nrhs * N * 8
    flops_alg = 8 * 9 * V * N * N * nrhs
    peak_tf = PEAK_TFLOPS[names[mp]]
    t_hbm = bytes_alg / (HBM_PEAK_GBS * 1e9)
    t_mfma = flops_alg / (peak_tf * 1e12)
    res = {
        "workload": f"coarse M multi-rhs {'x'.join(map(str, dims))} nbasis={N} nrhs={nrhs} "
                    f"{names[mp]} MFMA (fp32 accumulate), "
                    f"{'pre-quantised half' if use_half else 'fp32'} input, fp32 output",
        "rhs_sites_per_s": V * nrhs / t_dev,
        "ms_per_step": t_host * 1e3,
        "kernel_us": t_dev * 1e6,
        "GBps": bytes_alg / t_dev / 1e9,
        "TFLOPs": flops_alg / t_dev / 1e12,
        "hbm_frac": bytes_alg / t_dev / 1e9 / HBM_PEAK_GBS,
        "mfma_frac": flops_alg / t_dev / 1e12 / peak_tf,
        "bound": "hbm" if t_hbm >= t_mfma else "mfma",
        "roofline_frac": max(t_hbm, t_mfma) / t_dev,
        "bytes_per_launch": bytes_alg,
        "flops_per_launch": flops_alg,
        "dtype": names[mp],
        "kernel_variant": variant.value,
        "kernel_name": MRHS_KERNELS.get(variant.value, "untuned"),
    }
    if keep is not None:
        keep["links"] = links.to_numpy(np.complex128, (9, V, N, N))
        keep["in"] = vin.to_numpy(np.complex64, (V, nrhs, N)).astype(np.complex128)
        keep["out"] = vout.to_numpy(np.complex64, (V, nrhs, N)).astype(np.complex128)
    for b in (links, vin, vout):
        b.close()
    op.close()
    return res


def mg_leg(mgc, stream, dims, blocks, nbs, tol=1e-10,
           forms=("fp64_plain", "mixed_fp32", "mixed_fp32_graph"), setup_nrhs=8):
    """Three-level K-cycle multigrid solve of the Wilson-clover system
    (m = -0.1, csw = 1.0, random SU(3) field eps = 0.4) on one Gaussian rhs:
    FGMRES(20) outer in FP64 with true-residual recomputation, hierarchy in
    FP64 ("plain form") and in FP32 ("mixed"), the reference's default setup
    (BiCGSTAB 1e-3 / 100 test vectors).  "mixed_fp32_graph" runs the K-cycle
    as a captured HIP graph of device-scalar GMRES with fixed step counts
    (smoother 4, K-cycle wrapper 1, coarsest 6).
    Setup, the first preconditioner application (graph capture) and the
    solve timed apart.  The plain form generates its test vectors one at a
    time, as the reference does; the mixed forms solve them setup_nrhs at a
    time in lockstep through the multi-rhs operators, and "mixed_fp32" also
    times its set-up one at a time for comparison."""
    V = int(np.prod(dims))
    g = mgc.DeviceBuffer(4 * V * 9 * 16)
    mgc.random_gauge(g, dims, 0.4, SEED, 21, stream)
    fop = mgc.FineOp(dims, -0.1, 1.0)
    fop.set_gauge(g, stream)
    b = mgc.DeviceBuffer(V * 12 * 16)
    mgc.random_gaussian(b, V * 12, mgc.F64, SEED, 22, stream)
    x = mgc.DeviceBuffer(V * 12 * 16)
    stream.sync()
    lat = [tuple(dims)]
    for blk in blocks:
        lat.append(tuple(a // b for a, b in zip(lat[-1], blk)))
    out = {"workload": "Wilson-clover " + " -> ".join("x".join(map(str, d)) for d in lat)
                       + f", nbasis {'/'.join(str(2 * n) for n in nbs)}, K-cycle, FGMRES(20) "
                         f"to {tol:g}"}
    for name, prec, graph in (("fp64_plain", mgc.F64, 0), ("mixed_fp32", mgc.F32, 0),
                              ("mixed_fp32_graph", mgc.F32, 2)):
        if name not in forms:
            continue
        snrhs = 0 if prec == mgc.F64 else setup_nrhs
        t_setup1 = None
        if name == "mixed_fp32" and snrhs > 1:
            p = mgc.mg_params(nlevels=len(blocks) + 1, block=blocks, nb=nbs, precision=prec,
                              graph=graph)
            mg = mgc.Multigrid(fop, p)
            t0 = time.perf_counter()
            mg.setup(stream)
            t_setup1 = time.perf_counter() - t0
            mg.close()
        # graph form: fixed step counts 4 (smoother), 1 (K-cycle wrapper), 6 (coarsest)
        steps = dict(wrapper_restart=1, coarsest_restart=6) if graph else {}
        p = mgc.mg_params(nlevels=len(blocks) + 1, block=blocks, nb=nbs, precision=prec,
                          graph=graph, setup_nrhs=snrhs, **steps)
        mg = mgc.Multigrid(fop, p)
        t0 = time.perf_counter()
        mg.setup(stream)
        t_setup = time.perf_counter() - t0
        # one preconditioner application before the timed solve: first-use
        # work (workspaces; in graph mode the capture and instantiation of
        # the K-cycle graph) is timed apart
        t0 = time.perf_counter()
        mg.precondition(b, x)
        stream.sync()
        t_first = time.perf_counter() - t0
        st = mg.solve(b, x, tol=tol, maxit=500, restart=20)
        out[name] = {"setup_s": t_setup,
                     "setup_form": (f"test vectors {snrhs} at a time in lockstep" if snrhs > 1
                                    else "test vectors one at a time"),
                     "first_precondition_s": t_first,
                     "solve_s": st["seconds"],
                     "outer_iterations": st["iterations"], "converged": bool(st["converged"]),
                     "true_residual": st["true_residual"],
                     "precond_applies": st["precond_applies"]}
        if t_setup1 is not None:
            out[name]["setup_one_at_a_time_s"] = t_setup1
        if graph:
            out[name]["graph_steps"] = {"smoother": p.smoother_iters, "wrapper": p.wrapper_restart,
                                        "coarsest": p.coarsest_restart}
        mg.close()
    for buf in (g, b, x):
        buf.close()
    fop.close()
    return out


def mg_batched_leg(mgc, stream, dims, blocks, nbs, nrhs, coarse_prec, tol=1e-10):
    """The same K-cycle multigrid solve for nrhs Gaussian right-hand sides in
    lockstep (FP64 outer FGMRES(20), FP32 hierarchy, coarse levels on the
    multi-rhs MFMA stencil in coarse_prec), next to one single-rhs solve of
    the first rhs with the same hierarchy.  The set-up solves its test
    vectors 8 at a time in lockstep."""
    V = int(np.prod(dims))
    g = mgc.DeviceBuffer(4 * V * 9 * 16)
    mgc.random_gauge(g, dims, 0.4, SEED, 21, stream)
    fop = mgc.FineOp(dims, -0.1, 1.0)
    fop.set_gauge(g, stream)
    b = mgc.DeviceBuffer(V * nrhs * 12 * 16)
    mgc.random_gaussian(b, V * nrhs * 12, mgc.F64, SEED, 23, stream)
    x = mgc.DeviceBuffer(V * nrhs * 12 * 16)
    stream.sync()
    names = {mgc.F32: "f32", mgc.F16: "f16", mgc.BF16: "bf16"}
    p = mgc.mg_params(nlevels=len(blocks) + 1, block=blocks, nb=nbs, precision=mgc.F32,
                      setup_nrhs=8)
    mg = mgc.Multigrid(fop, p)
    t0 = time.perf_counter()
    mg.setup(stream)
    t_setup = time.perf_counter() - t0
    sts = mg.solve_mrhs(b, x, nrhs, coarse_prec=coarse_prec, tol=tol, maxit=500, restart=20)
    its = [st["iterations"] for st in sts]
    res = [st["true_residual"] for st in sts]
    # one rhs alone with the same hierarchy (single-rhs kernels throughout)
    b1 = mgc.DeviceBuffer(V * 12 * 16)
    mgc.random_gaussian(b1, V * 12, mgc.F64, SEED, 22, stream)
    x1 = mgc.DeviceBuffer(V * 12 * 16)
    st1 = mg.solve(b1, x1, tol=tol, maxit=500, restart=20)
    mg.close()
    # the batched K-cycle captured as a HIP graph of device-scalar batched
    # GMRES with fixed step counts (smoother 4, wrapper 1, coarsest 6)
    pg = mgc.mg_params(nlevels=len(blocks) + 1, block=blocks, nb=nbs, precision=mgc.F32,
                       setup_nrhs=8, graph=2, wrapper_restart=1, coarsest_restart=6)
    mgg = mgc.Multigrid(fop, pg)
    mgg.setup(stream)
    xg = mgc.DeviceBuffer(V * nrhs * 12 * 16)
    # one outer iteration first: the capture and instantiation of the graph
    # (at the first preconditioner application) are timed apart
    t0 = time.perf_counter()
    mgg.solve_mrhs(b, xg, nrhs, coarse_prec=coarse_prec, tol=tol, maxit=1, restart=20)
    t_first_g = time.perf_counter() - t0
    stg = mgg.solve_mrhs(b, xg, nrhs, coarse_prec=coarse_prec, tol=tol, maxit=500, restart=20)
    mgg.close()
    for buf in (g, b, x, b1, x1, xg):
        buf.close()
    fop.close()
    lat = [tuple(dims)]
    for blk in blocks:
        lat.append(tuple(a // b for a, b in zip(lat[-1], blk)))
    return {
        "workload": "Wilson-clover " + " -> ".join("x".join(map(str, d)) for d in lat)
                    + f", nbasis {'/'.join(str(2 * n) for n in nbs)}, {nrhs} rhs in lockstep, "
                      f"coarse levels {names[coarse_prec]} MFMA, FGMRES(20) to {tol:g}",
        "nrhs": nrhs,
        "setup_s": t_setup,
        "solve_s": sts[0]["seconds"],
        "solve_s_per_rhs": sts[0]["seconds"] / nrhs,
        "outer_iterations_per_rhs": its,
        "true_residual_max": max(res),
        "all_converged": all(st["converged"] for st in sts),
        "single_rhs_solve_s": st1["seconds"],
        "single_rhs_outer_iterations": st1["iterations"],
        "speedup_per_rhs_vs_single": st1["seconds"] / (sts[0]["seconds"] / nrhs),
        "graph": {
            "graph_steps": {"smoother": pg.smoother_iters, "wrapper": pg.wrapper_restart,
                            "coarsest": pg.coarsest_restart},
            "first_outer_iteration_s": t_first_g,
            "solve_s": stg[0]["seconds"],
            "solve_s_per_rhs": stg[0]["seconds"] / nrhs,
            "outer_iterations_per_rhs": [st["iterations"] for st in stg],
            "true_residual_max": max(st["true_residual"] for st in stg),
            "all_converged": all(st["converged"] for st in stg),
            "speedup_per_rhs_vs_single": st1["seconds"] / (stg[0]["seconds"] / nrhs),
        },
    }


def coarse_batch_leg(mgc, stream, fines=((48, 48, 48, 48), (32, 32, 32, 32), (24, 24, 24, 24)),
                     nrhs=12, mp=None, tol=1e-3):
    """Coarse solves of three lattices in one lockstep batch
    (mgc_coarse_solve_batched): the level-1 operators (4^4 blocks, nbasis 32)
    of FP32 Wilson-clover hierarchies on the given fine lattices, nrhs
    Gaussian rhs each, GMRES(20) to tol with the stencil in mp (FP16 MFMA by
    default) and one multi-lattice launch per application; against the same
    solves run one lattice at a time.  Both with prepared handles (the
    solve's state kept between calls), mean of 5 solves after a warm-up."""
    mp = mgc.F16 if mp is None else mp
    mgs, fops, gs, ops, bs, xs, coarse = [], [], [], [], [], [], []
    for k, d in enumerate(fines):
        V = int(np.prod(d))
        g = mgc.DeviceBuffer(4 * V * 9 * 16)
        mgc.random_gauge(g, d, 0.4, SEED, 41 + k, stream)
        fop = mgc.FineOp(d, -0.1, 1.0)
        fop.set_gauge(g, stream)
        p = mgc.mg_params(nlevels=2, block=[(4, 4, 4, 4)], nb=[16], precision=mgc.F32)
        mg = mgc.Multigrid(fop, p)
        mg.setup(stream)
        cd = tuple(a // 4 for a in d)
        Vc = int(np.prod(cd))
        op = mgc.CoarseOp.borrow(mg.coarse_op_handle(1), cd, 32, mgc.F64)
        b = mgc.DeviceBuffer(Vc * nrhs * 32 * 8)
        mgc.random_gaussian(b, Vc * nrhs * 32, mgc.F32, SEED, 51 + k, stream)
        xs.append(mgc.DeviceBuffer(Vc * nrhs * 32 * 8))
        mgs.append(mg); fops.append(fop); gs.append(g); ops.append(op); bs.append(b)
        coarse.append(cd)
    stream.sync()
    # prepared batches (workspaces and Krylov basis kept), timed after a warm-up
    # solve that also runs the stencil's variant tuner
    cb = mgc.CoarseBatch(ops, nrhs, mrhs_prec=mp, restart=20, stream=stream)
    cb.solve(bs, xs, tol=tol, maxit=300)
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        st = cb.solve(bs, xs, tol=tol, maxit=300)
    t_batch = (time.perf_counter() - t0) / reps
    cb.close()
    # the same batch with a fixed step count (the most steps any lattice's
    # tolerance-driven solve took), captured into a HIP graph at its first
    # solve and replayed
    its = [s_["iterations"] for s_ in st]
    steps = min(63, max(its))
    cg = mgc.CoarseBatch(ops, nrhs, mrhs_prec=mp, stream=stream, fixed_steps=steps, graph=2)
    cg.solve(bs, xs, tol=tol)
    t0 = time.perf_counter()
    for _ in range(reps):
        stg = cg.solve(bs, xs, tol=tol)
    t_graph = (time.perf_counter() - t0) / reps
    cg.close()
    t_sep, st_sep = 0.0, []
    for k in range(len(ops)):
        c1 = mgc.CoarseBatch([ops[k]], nrhs, mrhs_prec=mp, restart=20, stream=stream)
        c1.solve([bs[k]], [xs[k]], tol=tol, maxit=300)
        t0 = time.perf_counter()
        for _ in range(reps):
            s1 = c1.solve([bs[k]], [xs[k]], tol=tol, maxit=300)
        t_sep += (time.perf_counter() - t0) / reps
        st_sep += s1
        c1.close()
    for lst in (xs, bs, gs):
        for b in lst:
            b.close()
    for mg in mgs:
        mg.close()
    for fop in fops:
        fop.close()
    names = {mgc.F32: "f32", mgc.F16: "f16", mgc.BF16: "bf16"}
    return {
        "workload": "coarse GMRES(20) to %g on %s (level 1 of %s, nbasis 32), %d rhs each, "
                    "%s MFMA stencil, all lattices in one lockstep batch" % (
                        tol, " + ".join("x".join(map(str, c)) for c in coarse),
                        " / ".join("x".join(map(str, d)) for d in fines), nrhs, names[mp]),
        "batched_s": t_batch,
        "separate_s": t_sep,
        "speedup": t_sep / t_batch,
        "iterations_per_lattice": [max(its[k * nrhs:(k + 1) * nrhs]) for k in range(len(ops))],
        "separate_iterations_per_lattice": [max(s["iterations"] for s in st_sep[k * nrhs:(k + 1) * nrhs])
                                            for k in range(len(ops))],
        "all_converged": all(s["converged"] for s in st),
        "true_residual_max_fp32_stencil": max(s["true_residual"] for s in st),
        "fixed_steps_graph": {
            "steps": steps,
            "batched_s": t_graph,
            "speedup_vs_separate": t_sep / t_graph,
            "all_converged": all(s_["converged"] for s_ in stg),
            "true_residual_max": max(s_["true_residual"] for s_ in stg),
        },
    }


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def time_protocol(fn, ncal=10, min_sec=5.0, max_calls=1000):
    """include/Benchmark_helpers.h:36-59 with the defaults of
    src/Benchmark_coarse_operator.cc:60-61: ncal calibration calls, then
    min(max_calls, min_sec * ncal / t_cal + 1) timed calls; seconds per call."""
    t0 = time.perf_counter()
    for _ in range(ncal):
        fn()
    tcal = time.perf_counter() - t0
    n = int(min(max_calls, min_sec * ncal / tcal + 1))
    t0 = time.perf_counter()
    for _ in range(n):
        fn()
    return (time.perf_counter() - t0) / n, n


def cpu_baseline(dims, N, nrhs, head):
    """FP64 oracle multi-rhs coarse M (one OpenMP-parallel site loop, dense
    N x N matvec per point and rhs) on the host cores, on the headline's own
    links and FP32 inputs, timed with the reference's protocol (10 calibration
    calls, then enough calls for 5 s).  The same per-site work is timed on one
    thread on a 4^4 sample of the lattice (rhs·sites/s is an intensive rate).
    The oracle's result also checks the headline's output (per-rhs relative
    L2 error)."""
    import oracle
    V = int(np.prod(dims))
    links, v = head["links"], head["in"]
    ref = oracle.coarse_mrhs_M(dims, N, nrhs, links, v)
    errs = [oracle.relative_error(ref[:, k], head["out"][:, k]) for k in range(nrhs)]
    # the host threads this process may use (the OpenMP setting where given)
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    oracle.set_num_threads(threads)
    t_all, n_all = time_protocol(lambda: oracle.coarse_mrhs_M(dims, N, nrhs, links, v))
    # one thread: the same operator on a 4^4 sample (its links and inputs)
    sd = (4, 4, 4, 4)
    Vs = int(np.prod(sd))
    sl = np.ascontiguousarray(links.reshape(9, V, N, N)[:, :Vs])
    sv = np.ascontiguousarray(v.reshape(V, nrhs, N)[:Vs])
    oracle.set_num_threads(1)
    t_one, n_one = time_protocol(lambda: oracle.coarse_mrhs_M(sd, N, nrhs, sl, sv))
    oracle.set_num_threads(threads)
    return {
        "value": V * nrhs / t_all,
        "unit": "rhs·sites/s",
        "cores": threads,
        "kind": "port",
        "cpu_model": cpu_model(),
        "sample": f"FP64 oracle multi-rhs coarse M, {'x'.join(map(str, dims))} nbasis={N} "
                  f"nrhs={nrhs}, {threads} OpenMP threads: {n_all} timed calls after 10 "
                  f"calibration calls ({n_all * t_all:.1f} s)",
        "one_core": {"value": Vs * nrhs / t_one, "unit": "rhs·sites/s", "cores": 1,
                     "sample": f"the same operator on a 4^4 part of the lattice, one thread: "
                               f"{n_one} timed calls after 10 calibration calls "
                               f"({n_one * t_one:.1f} s)"},
        "headline_max_rel_err_vs_oracle": max(errs),
    }


def aux_legs(mgc, stream, legs, args):
    """Single-rhs transfers and fine operator on the 32^4 -> 8^4 drop-in
    workload, and the CoarsenOperator build on 32^4 with nbasis 64."""
    f32 = (32, 32, 32, 32)
    b4 = (4, 4, 4, 4)
    for prec, nm in ((mgc.F64, "f64"), (mgc.F32, "f32")):
        log(f"leg: project / promote 32^4 nbasis=32 {nm}")
        legs[f"transfer_{nm}_n32"] = transfer_leg(mgc, stream, f32, b4, 16, prec, args.steps // 4,
                                                  args.warmup)
    for prec, nm, nrhs in ((mgc.F64, "f64", 1), (mgc.F32, "f32", 1), (mgc.F32, "f32", 12)):
        log(f"leg: Wilson-clover 32^4 {nm} nrhs={nrhs}")
        legs[f"wilson_{nm}_nrhs{nrhs}"] = fine_op_leg(mgc, stream, f32, prec, nrhs,
                                                      args.steps // 4, args.warmup)
    log("leg: CoarsenOperator 32^4 nbasis=64")
    legs["coarsen_f64_n64"] = coarsen_leg(mgc, stream, f32, b4, 32)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    # untimed launches before each timed leg: enough for the GPU clocks to
    # reach their steady state (20 launches of the 0.17 ms headline kernel
    # leave it about 3 % slower than 200 do)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU-baseline leg")
    ap.add_argument("--no-mg", action="store_true", help="skip the multigrid-solve leg")
    ap.add_argument("--mg-dims", default="48,48,48,48",
                    help="fine lattice of the multigrid solve (blocks 4^4 then 4^4)")
    ap.add_argument("--only-headline", action="store_true",
                    help="run only the headline leg (for profiling)")
    ap.add_argument("--only-mrhs", type=int, default=0, metavar="NRHS",
                    help="run only the f16 nbasis-64 multi-rhs leg with NRHS right-hand sides "
                         "and print it (for profiling)")
    ap.add_argument("--mg-batch", default="32,32,32,32:12",
                    help="fine lattice and rhs count of the batched multigrid solve "
                         "(DIMS:NRHS; empty to skip)")
    ap.add_argument("--only-mg-batch", action="store_true",
                    help="run only the batched multigrid-solve leg and print it (for profiling)")
    ap.add_argument("--only-aux", action="store_true",
                    help="run only the transfer, fine-operator and CoarsenOperator legs")
    ap.add_argument("--only-mg-small", action="store_true",
                    help="only run the 16^4 eager / HIP-graph multigrid comparison")
    ap.add_argument("--only-coarse-batch", action="store_true",
                    help="only run the multi-lattice batched coarse-solve leg")
    ap.add_argument("--only-wilson", action="store_true",
                    help="only run the Wilson-clover apply legs")
    ap.add_argument("--only-coarsen", action="store_true",
                    help="only run the CoarsenOperator / block-orthonormalisation leg")
    ap.add_argument("--mg-forms", default="fp64_plain,mixed_fp32,mixed_fp32_graph",
                    help="forms of the multigrid leg to run (comma-separated)")
    ap.add_argument("--only-mg", action="store_true",
                    help="run only the multigrid-solve leg and print its result (for profiling)")
    ap.add_argument("--mrhs-variant", type=int, default=-1,
                    help="force one multi-rhs kernel variant instead of autotuning (profiling)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run a device copy kernel of BYTES bytes (calibrates the "
                         "FETCH_SIZE / WRITE_SIZE counters of a profiler run), then exit")
    args = ap.parse_args()
    if args.copy_calibration:
        import mgc
        nbytes = (args.copy_calibration + 15) // 16 * 16
        src, dst = mgc.DeviceBuffer(nbytes), mgc.DeviceBuffer(nbytes)
        src.zero()
        for _ in range(max(1, args.steps)):
            mgc.call("mgc_device_copy", dst.ptr, src.ptr, ctypes.c_size_t(nbytes), None)
        mgc.synchronize()
        print(json.dumps({"copy_calibration_bytes": nbytes, "launches": max(1, args.steps)}))
        return 0
    if args.gpus != 1:
        print("bench.py measures one GPU only: the multi-GPU form (4d domain decomposition "
              "of the coarse lattice with halo exchange) is not built",
              file=sys.stderr)
        return 2

    import mgc
    log(f"device: {mgc.device_name()}")
    if args.mrhs_variant >= 0:
        mgc.call("mgc_mrhs_set_variant", args.mrhs_variant)
    stream = mgc.Stream()
    if args.only_mrhs:
        res = mrhs_leg(mgc, stream, (8, 8, 8, 8), 64, args.only_mrhs, mgc.F16, args.steps,
                       args.warmup)
        print(json.dumps({"mrhs": res}), flush=True)
        return 0
    def batch_leg():
        dstr, nstr = args.mg_batch.split(":")
        d = tuple(int(v) for v in dstr.split(","))
        log(f"leg: batched multigrid solve {d} nrhs={nstr}")
        return mg_batched_leg(mgc, stream, d, [(4, 4, 4, 4), (4, 4, 4, 4)], [16, 16], int(nstr),
                              mgc.F16)
    if args.only_mg_batch:
        print(json.dumps({"mg_batched": batch_leg()}), flush=True)
        return 0
    if args.only_mg_small:
        log("leg: multigrid solve 16^4, eager against HIP-graph K-cycle")
        print(json.dumps({"small_graph": mg_leg(mgc, stream, (16, 16, 16, 16),
                                                [(4, 4, 4, 4), (2, 2, 2, 2)], [16, 16],
                                                forms=("mixed_fp32", "mixed_fp32_graph"))}),
              flush=True)
        return 0
    if args.only_coarse_batch:
        log("leg: batched coarse solves 12^4 + 8^4 + 6^4, 12 rhs each")
        print(json.dumps({"coarse_batched": coarse_batch_leg(mgc, stream),
                          "coarse_batched_f32": coarse_batch_leg(mgc, stream, mp=mgc.F32,
                                                                 tol=1e-6)}), flush=True)
        return 0
    if args.only_mg:
        d = tuple(int(v) for v in args.mg_dims.split(","))
        log(f"leg: multigrid solve {d}")
        print(json.dumps({"mg_solve": mg_leg(mgc, stream, d, [(4, 4, 4, 4), (4, 4, 4, 4)],
                                             [16, 16], forms=tuple(args.mg_forms.split(",")))}),
              flush=True)
        return 0
    legs = {}
    if args.only_wilson:
        for prec, nm, nrhs in ((mgc.F64, "f64", 1), (mgc.F32, "f32", 1), (mgc.F32, "f32", 12)):
            log(f"leg: Wilson-clover 32^4 {nm} nrhs={nrhs}")
            legs[f"wilson_{nm}_nrhs{nrhs}"] = fine_op_leg(mgc, stream, (32, 32, 32, 32), prec, nrhs,
                                                          args.steps // 4, args.warmup)
        print(json.dumps({"legs": legs}), flush=True)
        return 0
    if args.only_coarsen:
        log("leg: CoarsenOperator 32^4 nbasis=64")
        legs["coarsen_f64_n64"] = coarsen_leg(mgc, stream, (32, 32, 32, 32), (4, 4, 4, 4), 32)
        print(json.dumps({"legs": legs}), flush=True)
        return 0
    if args.only_aux:
        aux_legs(mgc, stream, legs, args)
        print(json.dumps({"legs": legs}), flush=True)
        return 0
    dims8 = (8, 8, 8, 8)
    # headline: multi-rhs coarse stencil, 8^4 coarse lattice, nbasis 64, 48 rhs,
    # FP16 MFMA with FP32 accumulation (pre-quantised FP16 inputs, FP32 output)
    head_key = "M_mrhs48_f16_n64"
    log("leg (headline): multi-rhs f16 nbasis=64 nrhs=48")
    head_host = {} if not args.no_cpu else None
    legs[head_key] = mrhs_leg(mgc, stream, dims8, 64, 48, mgc.F16, args.steps, args.warmup,
                              keep=head_host)
    if not args.only_headline:
        log("leg: single-rhs fp64 nbasis=64 (reference point of the multi-rhs ratio)")
        legs["M_1rhs_f64_n64"] = single_rhs_leg(mgc, stream, dims8, 64, mgc.F64, args.steps,
                                                args.warmup)
        log("leg: single-rhs fp64 / fp32 nbasis=32")
        legs["M_1rhs_f64_n32"] = single_rhs_leg(mgc, stream, dims8, 32, mgc.F64, args.steps,
                                                args.warmup)
        legs["M_1rhs_f32_n32"] = single_rhs_leg(mgc, stream, dims8, 32, mgc.F32, args.steps,
                                                args.warmup)
        for nrhs in (12, 24, 48):
            for mp, nm in ((mgc.F32, "f32"), (mgc.F16, "f16"), (mgc.BF16, "bf16")):
                key = f"M_mrhs{nrhs}_{nm}_n64"
                if key in legs:
                    continue
                log(f"leg: multi-rhs {nm} nbasis=64 nrhs={nrhs}")
                legs[key] = mrhs_leg(mgc, stream, dims8, 64, nrhs, mp, args.steps, args.warmup)
        aux_legs(mgc, stream, legs, args)
        log("leg: multi-rhs f16 nbasis=64 nrhs=48, fp32 input quantised in the step")
        legs["M_mrhs48_f16_n64_f32in"] = mrhs_leg(mgc, stream, dims8, 64, 48, mgc.F16,
                                                  args.steps, args.warmup, half_input=False)
    head = legs[head_key]
    mg = None
    if not args.no_mg and not args.only_headline:
        d = tuple(int(v) for v in args.mg_dims.split(","))
        log(f"leg: multigrid solve {d}")
        mg = mg_leg(mgc, stream, d, [(4, 4, 4, 4), (4, 4, 4, 4)], [16, 16])
        if args.mg_batch:
            mg["batched"] = batch_leg()
        log("leg: multigrid solve 16^4, eager against HIP-graph K-cycle")
        mg["small_graph"] = mg_leg(mgc, stream, (16, 16, 16, 16), [(4, 4, 4, 4), (2, 2, 2, 2)],
                                   [16, 16], forms=("mixed_fp32", "mixed_fp32_graph"))
        log("leg: batched coarse solves 12^4 + 8^4 + 6^4, 12 rhs each")
        mg["coarse_batched"] = coarse_batch_leg(mgc, stream)
        mg["coarse_batched_f32"] = coarse_batch_leg(mgc, stream, mp=mgc.F32, tol=1e-6)
    ref64 = legs.get("M_1rhs_f64_n64")
    line = {
        "metric": METRIC,
        "value": head["rhs_sites_per_s"],
        "unit": "rhs·sites/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f16",
        "data": "synthetic (counter-based Gaussian links and vectors, seed {1,2,3,4})",
        "config": {"workload": head["workload"], "headline_leg": head_key,
                   "coarse_lattice": "8^4", "nbasis": 64, "nrhs": 48,
                   "accumulate": "fp32", "output": "fp32"},
        "roofline": {
            "bound": head["bound"],
            "achieved": head["GBps"] if head["bound"] == "hbm" else head["TFLOPs"],
            "peak": HBM_PEAK_GBS if head["bound"] == "hbm" else PEAK_TFLOPS["f16"],
            "unit": "GB/s" if head["bound"] == "hbm" else "TFLOP/s",
            "frac": head["hbm_frac"] if head["bound"] == "hbm" else head["mfma_frac"],
            "traffic": HEADLINE_TRAFFIC_BYTES.get(head["kernel_variant"]),
            "traffic_source": "rocprofv3 FETCH_SIZE / WRITE_SIZE of this kernel variant, "
                              "calibrated on a 1 GiB copy, read at commit 0f953b4 "
                              "(profiles/k_mrhs_f16_8x8x8x8_n64_nrhs48.md)",
            "kernel": head["kernel_name"],
            "bytes_per_launch": head["bytes_per_launch"],
            "flops_per_launch": head["flops_per_launch"],
        },
        "speedup_vs_fp64_single_rhs": (head["rhs_sites_per_s"] / ref64["rhs_sites_per_s"]
                                       if ref64 else None),
        "mg_solve": mg,
        "legs": legs,
    }
    if not args.no_cpu:
        log("cpu baseline (FP64 oracle, same multi-rhs workload)")
        line["cpu_baseline"] = cpu_baseline(dims8, 64, 48, head_host)
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
