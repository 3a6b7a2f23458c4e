"""The C ABI library loads and exports every symbol include/mgc.h declares (no GPU)."""
import ctypes

import mgc


def test_library_exports_every_declared_symbol():
    names = mgc.declared_symbols()
    assert len(names) > 30
    L = ctypes.CDLL(mgc.LIB_PATH)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_binding_covers_declared_symbols():
    # every declared entry point has ctypes argument types in the binding
    names = mgc.declared_symbols()
    assert not [n for n in names if n not in mgc._SIGS]


def test_mg_params_layout_and_reference_defaults():
    """The ctypes mirror of mgc_mg_params matches the C layout: the C side fills
    the reference defaults (include/MultigridParams.h:95-133) field by field."""
    p = mgc.mg_params()
    assert p.nlevels == 3
    assert [list(p.block[0]), list(p.block[1])] == [[4, 4, 4, 4], [2, 2, 2, 2]]
    assert (p.smoother_tol, p.smoother_iters, p.smoother_restart) == (1e-3, 4, 4)
    assert (p.wrapper_tol, p.wrapper_iters, p.wrapper_restart) == (1e-1, 10, 5)
    assert (p.coarsest_tol, p.coarsest_iters, p.coarsest_restart) == (5e-2, 50, 10)
    assert p.seed == 1 | (2 << 16) | (3 << 32) | (4 << 48)
    # setup: BiCGSTAB 1e-3 / 100, test vectors, pre- and post-orthonormalised,
    # no refinement (include/MultigridParams.h:97-107); GMRES smoother and
    # coarsest solver; nb = 20 (NBASIS 40, src/Test_wilsonclover_mg.cc:39-40)
    assert (p.setup_solver, p.setup_tol, p.setup_solve_iters) == (mgc.SOLVER_BICGSTAB, 1e-3, 100)
    assert (p.setup_iters, p.setup_refine_iters) == (1, 0)
    assert p.setup_basis_type == mgc.BASIS_TEST_VECTORS
    assert (p.setup_pre_orthonormalise, p.setup_post_orthonormalise) == (1, 1)
    assert (p.smoother_solver, p.coarsest_solver) == (mgc.SOLVER_GMRES, mgc.SOLVER_GMRES)
    assert list(p.nb) == [20, 20, 20]
    q = mgc.mg_params(precision=mgc.F64, nb=[12, 8], block=[(2, 2, 2, 2), (3, 3, 3, 3)])
    assert q.precision == mgc.F64 and list(q.nb)[:2] == [12, 8] and list(q.block[1]) == [3] * 4


def test_errors_are_status_codes_without_device():
    L = mgc.lib()
    # argument validation happens before any device call
    assert L.mgc_coarse_op_create(None, None, 8, mgc.F64, 0) == -1
    assert b"null" in L.mgc_last_error()
    assert L.mgc_coarse_op_M(None, None, None, None) == -1
    assert L.mgc_version() >= 10000


def test_kernel_variant_ranges_without_device():
    """The variant controls only set a process-wide choice: every documented
    value is accepted and the first value past the range is refused."""
    L = mgc.lib()
    try:
        for v in (-1, 0, 1, 2, 3):
            assert L.mgc_fine_op_set_variant(v) == 0
        assert L.mgc_fine_op_set_variant(4) == -1
        assert b"variant" in L.mgc_last_error()
        assert L.mgc_fine_op_set_variant(-2) == -1
    finally:
        assert L.mgc_fine_op_set_variant(-1) == 0


def test_multigrid_parameter_checks_without_device():
    """Hierarchy parameters are validated before any device work: a solver
    type the chosen mode cannot run, an unknown graph mode or basis type is a
    status code with a message, not a failure later in the set-up."""
    L = mgc.lib()
    lop = mgc.make_linop(lambda *a: 0, (4, 4, 4, 4))
    cases = [
        (dict(graph=3), b"graph must be 0, 1 or 2"),
        (dict(graph=2, smoother_solver=mgc.SOLVER_MR), b"graph mode runs GMRES"),
        (dict(graph=1, coarsest_solver=mgc.SOLVER_BICGSTAB), b"graph mode runs GMRES"),
        (dict(smoother_solver=mgc.SOLVER_BICGSTAB), b"the smoother is GMRES or MR"),
        (dict(setup_solver=99), b"unknown setup or coarsest solver type"),
        (dict(setup_basis_type=7), b"unknown basis-vector type"),
        (dict(setup_iters=-1), b"negative setup iteration count"),
        (dict(setup_nrhs=65), b"setup_nrhs must be 0..64"),
        (dict(setup_nrhs=-2), b"setup_nrhs must be 0..64"),
    ]
    for kw, msg in cases:
        p = mgc.mg_params(nlevels=2, block=[(2, 2, 2, 2)], nb=[4], **kw)
        h = ctypes.c_void_p()
        rc = L.mgc_mg_create_linop(ctypes.byref(h), ctypes.byref(lop), ctypes.byref(p))
        assert rc == -1, kw
        assert msg in L.mgc_last_error(), (kw, L.mgc_last_error())
        assert not h.value
