"""numpy face of the CPU FP64 oracle (liboracle.so, built from oracle/*.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by
the CPU-baseline leg of bench.py, never by the library it checks.  Each
wrapper names the reference file:line its C function restates.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} not built (make -C oracle)")
        _lib = ctypes.CDLL(LIB_PATH)
    return _lib


def _d(a: np.ndarray):
    assert a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def _dims(dims):
    return (ctypes.c_int * 4)(*[int(v) for v in dims])


def num_threads() -> int:
    return int(lib().orc_num_threads())


def set_num_threads(n: int) -> None:
    lib().orc_set_num_threads(int(n))


def volume(dims) -> int:
    return int(np.prod(dims))


def gaussian(n: int, seed: int, stream: int) -> np.ndarray:
    """Counter-based Gaussian complex numbers, re/im ~ N(0, 1/2)."""
    out = np.empty(n, dtype=np.complex128)
    lib().orc_gaussian(_d(out.view(np.float64)), ctypes.c_size_t(n), ctypes.c_uint64(seed),
                       ctypes.c_uint64(stream))
    return out


def point_from_dir_disp(dir_: int, disp: int) -> int:
    return lib().orc_point_from_dir_disp(dir_, disp)


def point_dir_disp(p: int):
    d, s = ctypes.c_int(), ctypes.c_int()
    lib().orc_point_dir_disp(p, ctypes.byref(d), ctypes.byref(s))
    return d.value, s.value


def site_shift(dims, x: int, dir_: int, disp: int) -> int:
    return lib().orc_site_shift(_dims(dims), x, dir_, disp)


def parity(dims, x: int) -> int:
    return lib().orc_parity(_dims(dims), x)


def half_index(dims, x: int) -> int:
    return lib().orc_half_index(_dims(dims), x)


def coarse_apply(dims, N, links, vin, mask=0x1FF, dag=0) -> np.ndarray:
    """out = sum_{p in mask} A_p(x) in(x+d_p), G5C-sandwiched when dag.
    include/CoarsenedMatrixBaseline.h:314-344, include/CoarsenedMatrixUpstream.h:609-742."""
    links = np.ascontiguousarray(links, dtype=np.complex128)
    vin = np.ascontiguousarray(vin, dtype=np.complex128)
    out = np.zeros_like(vin)
    lib().orc_coarse_apply(_dims(dims), N, _d(links.view(np.float64)), _d(vin.view(np.float64)),
                           _d(out.view(np.float64)), ctypes.c_uint(mask), dag)
    return out


def coarse_apply_cb(dims, N, links, vin_half, out_cb, mask=0xFF, dag=0) -> np.ndarray:
    """Half-lattice hopping, include/CoarsenedMatrixBaseline.h:448-554."""
    links = np.ascontiguousarray(links, dtype=np.complex128)
    vin_half = np.ascontiguousarray(vin_half, dtype=np.complex128)
    out = np.zeros_like(vin_half)
    lib().orc_coarse_apply_cb(_dims(dims), N, _d(links.view(np.float64)),
                              _d(vin_half.view(np.float64)), _d(out.view(np.float64)), out_cb,
                              ctypes.c_uint(mask), dag)
    return out


def site_matrix_apply(dims, N, mats, vin, cb=-1, adjoint=0) -> np.ndarray:
    """Mooee family, include/CoarsenedMatrixBaseline.h:468-489."""
    mats = np.ascontiguousarray(mats, dtype=np.complex128)
    vin = np.ascontiguousarray(vin, dtype=np.complex128)
    out = np.zeros_like(vin)
    lib().orc_site_matrix_apply(_dims(dims), N, _d(mats.view(np.float64)),
                                _d(vin.view(np.float64)), _d(out.view(np.float64)), cb, adjoint)
    return out


def invert(A: np.ndarray) -> np.ndarray:
    A = np.ascontiguousarray(A, dtype=np.complex128)
    N = A.shape[0]
    out = np.empty_like(A)
    if lib().orc_invert(N, _d(A.view(np.float64)), _d(out.view(np.float64))) != 0:
        raise np.linalg.LinAlgError("singular")
    return out


def invert_self(dims, N, links) -> np.ndarray:
    """InvertSelfStencilPoint, include/CoarsenedMatrixBaseline.h:735-768."""
    links = np.ascontiguousarray(links, dtype=np.complex128)
    out = np.empty((volume(dims), N, N), dtype=np.complex128)
    if lib().orc_invert_self(_dims(dims), N, _d(links.view(np.float64)),
                             _d(out.view(np.float64))) != 0:
        raise np.linalg.LinAlgError("singular self link")
    return out


def links_flat_to_split(V, N, flat) -> np.ndarray:
    """include/Layout_converters.h:62-89."""
    flat = np.ascontiguousarray(flat, dtype=np.complex128)
    out = np.empty(flat.size, dtype=np.complex128)
    lib().orc_links_flat_to_split(V, N, _d(flat.view(np.float64)), _d(out.view(np.float64)))
    return out


def links_split_to_flat(V, N, split) -> np.ndarray:
    """include/Layout_converters.h:33-60."""
    split = np.ascontiguousarray(split, dtype=np.complex128)
    out = np.empty(split.size, dtype=np.complex128)
    lib().orc_links_split_to_flat(V, N, _d(split.view(np.float64)), _d(out.view(np.float64)))
    return out


def coarse_mrhs_M(dims, N, nrhs, links, vin) -> np.ndarray:
    """include/CoarsenedMatrixUpstreamImprovedDirsaveLutMRHS.h:736-804."""
    links = np.ascontiguousarray(links, dtype=np.complex128)
    vin = np.ascontiguousarray(vin, dtype=np.complex128)
    out = np.zeros_like(vin)
    lib().orc_coarse_mrhs_M(_dims(dims), N, nrhs, _d(links.view(np.float64)),
                            _d(vin.view(np.float64)), _d(out.view(np.float64)))
    return out


def construct_remaining_links(dims, N, links, disp_have=-1) -> np.ndarray:
    """include/CoarsenedMatrixUpstreamImprovedDirsave.h:999-1032 (disp_have=-1),
    include/CoarsenedMatrixKernelsRework.h:141-167 (disp_have=+1)."""
    out = np.array(links, dtype=np.complex128, copy=True, order="C")
    lib().orc_construct_remaining_links(_dims(dims), N, _d(out.view(np.float64)), disp_have)
    return out


def relative_error(ref: np.ndarray, res: np.ndarray) -> float:
    """||ref - res||_2 / ||ref||_2 (the square root of the reference's
    norm2(diff)/norm2(ref), include/Benchmark_helpers.h:215-236)."""
    ref = np.asarray(ref).ravel()
    res = np.asarray(res).ravel()
    n = np.linalg.norm(ref)
    return float(np.linalg.norm(ref - res) / (n if n > 0 else 1.0))


# ----------------------------------------------------------------------------
# block transfer (aggregation.c)
# ----------------------------------------------------------------------------
MASK_ALL, MASK_INNER, MASK_OUTER = 0, 1, 2


def coarse_dims(fine_dims, block):
    return tuple(int(f) // int(b) for f, b in zip(fine_dims, block))


def lut_populate(fine_dims, block, kind=MASK_ALL, point=8):
    """include/CoarseningLookupTableRework.h:121-179: (lut lists, reverse table)."""
    Vf = volume(fine_dims)
    Vc = volume(coarse_dims(fine_dims, block))
    bv = volume(block)
    lut = np.full(Vc * bv, -1, dtype=np.int64)
    sizes = np.zeros(Vc, dtype=np.int64)
    rev = np.zeros(Vf, dtype=np.int64)
    ip = ctypes.POINTER(ctypes.c_int64)
    rc = lib().orc_lut_populate(_dims(fine_dims), _dims(block), kind, point,
                                lut.ctypes.data_as(ip), sizes.ctypes.data_as(ip),
                                rev.ctypes.data_as(ip))
    if rc != 0:
        raise ValueError("block does not divide the lattice")
    lut = lut.reshape(Vc, bv)
    return [lut[s, :sizes[s]].copy() for s in range(Vc)], rev


def project(fine_dims, block, nb, basis, fine, kind=MASK_ALL, point=8):
    """aggregateProject, include/AggregationKernelsRework.h:49-94 (split chirality)."""
    basis = np.ascontiguousarray(basis, dtype=np.complex128)
    fine = np.ascontiguousarray(fine, dtype=np.complex128)
    Vc = volume(coarse_dims(fine_dims, block))
    out = np.zeros((Vc, 2 * nb), dtype=np.complex128)
    lib().orc_project(_dims(fine_dims), _dims(block), nb, _d(basis.view(np.float64)),
                      _d(fine.view(np.float64)), _d(out.view(np.float64)), kind, point)
    return out


def promote(fine_dims, block, nb, basis, coarse):
    """aggregatePromote, include/AggregationKernelsRework.h:159-205."""
    basis = np.ascontiguousarray(basis, dtype=np.complex128)
    coarse = np.ascontiguousarray(coarse, dtype=np.complex128)
    out = np.zeros((volume(fine_dims), 12), dtype=np.complex128)
    lib().orc_promote(_dims(fine_dims), _dims(block), nb, _d(basis.view(np.float64)),
                      _d(coarse.view(np.float64)), _d(out.view(np.float64)))
    return out


def orthogonalise(fine_dims, block, nb, basis, passes=2):
    """aggregateOrthogonalise, include/AggregationKernelsRework.h:266-341."""
    out = np.array(basis, dtype=np.complex128, copy=True, order="C")
    lib().orc_orthogonalise(_dims(fine_dims), _dims(block), nb, _d(out.view(np.float64)), passes)
    return out


def project_ef(fine_dims, block, nb, Ef, basis, fine, kind=MASK_ALL, point=8):
    """aggregateProject on fields of Ef elements per site."""
    basis = np.ascontiguousarray(basis, dtype=np.complex128)
    fine = np.ascontiguousarray(fine, dtype=np.complex128)
    Vc = volume(coarse_dims(fine_dims, block))
    out = np.zeros((Vc, 2 * nb), dtype=np.complex128)
    lib().orc_project_ef(_dims(fine_dims), _dims(block), nb, Ef, _d(basis.view(np.float64)),
                         _d(fine.view(np.float64)), _d(out.view(np.float64)), kind, point)
    return out


def promote_ef(fine_dims, block, nb, Ef, basis, coarse):
    basis = np.ascontiguousarray(basis, dtype=np.complex128)
    coarse = np.ascontiguousarray(coarse, dtype=np.complex128)
    out = np.zeros((volume(fine_dims), Ef), dtype=np.complex128)
    lib().orc_promote_ef(_dims(fine_dims), _dims(block), nb, Ef, _d(basis.view(np.float64)),
                         _d(coarse.view(np.float64)), _d(out.view(np.float64)))
    return out


def orthogonalise_ef(fine_dims, block, nb, Ef, basis, passes=2):
    out = np.array(basis, dtype=np.complex128, copy=True, order="C")
    lib().orc_orthogonalise_ef(_dims(fine_dims), _dims(block), nb, Ef, _d(out.view(np.float64)),
                               passes)
    return out


def coarsen_coarse_operator(dims1, block, nb2, N1, links1, basis):
    """CoarsenOperator (mask form, include/CoarsenedMatrixBaseline.h:609-675) of a
    coarse operator onto the next level.  Returns FLAT links [9][V2][N2][N2]."""
    links1 = np.ascontiguousarray(links1, dtype=np.complex128)
    basis = np.ascontiguousarray(basis, dtype=np.complex128)
    V2 = volume(coarse_dims(dims1, block))
    N2 = 2 * nb2
    out = np.zeros((9, V2, N2, N2), dtype=np.complex128)
    lib().orc_coarsen_coarse_operator(_dims(dims1), _dims(block), nb2, N1,
                                      _d(links1.view(np.float64)), _d(basis.view(np.float64)),
                                      _d(out.view(np.float64)))
    return out


def chiral_double(Vf, nb, basis):
    """performChiralDoublingG5C, include/Benchmark_helpers.h:156-181."""
    basis = np.ascontiguousarray(basis, dtype=np.complex128)
    out = np.zeros((2 * nb,) + basis.shape[1:], dtype=np.complex128)
    lib().orc_chiral_double(Vf, nb, _d(basis.view(np.float64)), _d(out.view(np.float64)))
    return out


def project_doubled(fine_dims, block, N, basis, fine):
    """Grid blockProject over N chirally doubled vectors (include/CoarsenedMatrixUpstream.h:172)."""
    basis = np.ascontiguousarray(basis, dtype=np.complex128)
    fine = np.ascontiguousarray(fine, dtype=np.complex128)
    out = np.zeros((volume(coarse_dims(fine_dims, block)), N), dtype=np.complex128)
    lib().orc_project_doubled(_dims(fine_dims), _dims(block), N, _d(basis.view(np.float64)),
                              _d(fine.view(np.float64)), _d(out.view(np.float64)))
    return out


def orthogonalise_doubled(fine_dims, block, N, basis, passes=1):
    """blockLutedOrthonormalise, include/CoarsenedMatrixUpstreamImprovedDirsaveLut.h:176-239."""
    out = np.array(basis, dtype=np.complex128, copy=True, order="C")
    lib().orc_orthogonalise_doubled(_dims(fine_dims), _dims(block), N, _d(out.view(np.float64)),
                                    passes)
    return out


# ----------------------------------------------------------------------------
# fine Wilson-clover operator, gauge field, coarse operator construction
# ----------------------------------------------------------------------------
def gamma(mu: int) -> np.ndarray:
    out = np.zeros((4, 4), dtype=np.complex128)
    lib().orc_gamma(mu, _d(out.view(np.float64)))
    return out


def su3_gauge(dims, eps=0.4, seed=1, stream=0) -> np.ndarray:
    """U_mu(x) = exp(i eps H), H = traceless Hermitian part of a Gaussian 3x3
    (a_ij from the counter-based generator, index (mu*V + x)*9 + 3i + j);
    exponential through numpy's Hermitian eigensolver.  [4][V][3][3]."""
    V = volume(dims)
    a = gaussian(4 * V * 9, seed, stream).reshape(4 * V, 3, 3)
    H = 0.5 * (a + a.conj().transpose(0, 2, 1))
    tr = np.trace(H, axis1=1, axis2=2).real / 3.0
    H = H - tr[:, None, None] * np.eye(3)[None]
    w, Q = np.linalg.eigh(H)
    U = np.einsum("nij,nj,nkj->nik", Q, np.exp(1j * eps * w), Q.conj())
    return np.ascontiguousarray(U.reshape(4, V, 3, 3))


def wilson_diag(dims, gauge, mass, csw) -> np.ndarray:
    """(4+m) + clover term per site, [V][12][12]."""
    gauge = np.ascontiguousarray(gauge, dtype=np.complex128)
    out = np.zeros((volume(dims), 12, 12), dtype=np.complex128)
    lib().orc_wilson_diag(_dims(dims), _d(gauge.view(np.float64)), ctypes.c_double(mass),
                          ctypes.c_double(csw), _d(out.view(np.float64)))
    return out


def wilson_dir(dims, gauge, psi, mu, disp) -> np.ndarray:
    gauge = np.ascontiguousarray(gauge, dtype=np.complex128)
    psi = np.ascontiguousarray(psi, dtype=np.complex128)
    out = np.zeros_like(psi)
    lib().orc_wilson_dir(_dims(dims), _d(gauge.view(np.float64)), _d(psi.view(np.float64)),
                         _d(out.view(np.float64)), mu, disp)
    return out


def wilson_diag_apply(diag, psi) -> np.ndarray:
    diag = np.ascontiguousarray(diag, dtype=np.complex128)
    psi = np.ascontiguousarray(psi, dtype=np.complex128)
    out = np.zeros_like(psi)
    lib().orc_wilson_diag_apply(diag.shape[0], _d(diag.view(np.float64)),
                                _d(psi.view(np.float64)), _d(out.view(np.float64)))
    return out


def wilson_M(dims, gauge, diag, psi) -> np.ndarray:
    gauge = np.ascontiguousarray(gauge, dtype=np.complex128)
    diag = np.ascontiguousarray(diag, dtype=np.complex128)
    psi = np.ascontiguousarray(psi, dtype=np.complex128)
    out = np.zeros_like(psi)
    lib().orc_wilson_M(_dims(dims), _d(gauge.view(np.float64)), _d(diag.view(np.float64)),
                       _d(psi.view(np.float64)), _d(out.view(np.float64)))
    return out


def coarsen_operator(fine_dims, block, nb, gauge, diag, basis) -> np.ndarray:
    """CoarsenOperator, include/CoarsenedMatrixBaseline.h:575-705 (mask form).
    Returns FLAT links [9][Vc][N][N], N = 2*nb."""
    gauge = np.ascontiguousarray(gauge, dtype=np.complex128)
    diag = np.ascontiguousarray(diag, dtype=np.complex128)
    basis = np.ascontiguousarray(basis, dtype=np.complex128)
    Vc = volume(coarse_dims(fine_dims, block))
    N = 2 * nb
    out = np.zeros((9, Vc, N, N), dtype=np.complex128)
    lib().orc_coarsen_operator(_dims(fine_dims), _dims(block), nb, _d(gauge.view(np.float64)),
                               _d(diag.view(np.float64)), _d(basis.view(np.float64)),
                               _d(out.view(np.float64)))
    return out
