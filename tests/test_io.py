"""On-disk formats (host code of the library, no GPU): the reference drivers'
XML parameter file (Grid's serialisation of MGTestParams,
include/MultigridParams.h:73-185) and NERSC / ILDG gauge configurations
(src/Test_wilsonclover_mg.cc:94-101).

The gauge tests pin the readers against writers written here in numpy (the
NERSC text header and big-endian two-row or 3x3 records; LIME record headers
with struct) and the plaquette against a numpy einsum restatement, so a
wrong byte order, site order, row reconstruction or record layout fails.
Grid's own files cannot be produced here: the exact bytes Grid's XmlWriter
and NerscIO emit are parity unpinned beyond these format rules.
"""
import os
import struct

import numpy as np
import pytest

import mgc
import oracle

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def plaquette_np(U, dims):
    """Average Re tr(U_mu(x) U_nu(x+mu) U_mu(x+nu)^H U_nu(x)^H) / 3 with np.roll."""
    L = list(dims)
    G = U.reshape(4, L[3], L[2], L[1], L[0], 3, 3)  # [mu][t][z][y][x]
    axis = {0: 3, 1: 2, 2: 1, 3: 0}
    tot = 0.0
    for mu in range(4):
        for nu in range(mu + 1, 4):
            a = G[mu]
            b = np.roll(G[nu], -1, axis=axis[mu])
            c = np.roll(G[mu], -1, axis=axis[nu])
            d = G[nu]
            p = np.einsum("...ij,...jk,...lk,...ml->...im", a, b, c.conj(), d.conj())
            tot += np.einsum("...ii", p).real.sum()
    return tot / (3 * 6 * np.prod(L))


def test_xml_reference_format_fixture():
    mg, t = mgc.read_params_xml(os.path.join(GOLDEN, "mg_params_reference_format.xml"))
    assert mg.nlevels == 3
    assert [list(mg.block[0]), list(mg.block[1])] == [[4, 4, 4, 2], [3, 3, 3, 3]]
    assert (mg.setup_iters, mg.setup_refine_iters) == (2, 1)
    assert mg.setup_basis_type == mgc.BASIS_NULL_VECTORS
    assert (mg.setup_pre_orthonormalise, mg.setup_post_orthonormalise) == (0, 1)
    assert mg.setup_recursive_updates == 0
    assert (mg.setup_solver, mg.setup_tol, mg.setup_solve_iters, mg.setup_restart) == \
        (mgc.SOLVER_GMRES, 0.005, 30, 12)
    assert (mg.wrapper_tol, mg.wrapper_iters, mg.wrapper_restart) == (0.05, 8, 4)
    assert (mg.smoother_solver, mg.smoother_tol, mg.smoother_iters, mg.smoother_restart) == \
        (mgc.SOLVER_MR, 1e-4, 6, 6)
    assert (mg.coarsest_solver, mg.coarsest_tol, mg.coarsest_iters, mg.coarsest_restart) == \
        (mgc.SOLVER_BICGSTAB, 0.02, 80, 16)
    assert (t.outer_tol, t.outer_maxit, t.outer_restart) == (1e-10, 500, 25)
    assert (t.mass_setup, t.mass_solve, t.csw) == (-0.05, -0.1, 1.25)
    assert t.config == b"/data/ckpoint_lat.1000" and t.source_type == b"gaussian"
    assert t.antiperiodic_t == 1


def test_xml_round_trip(tmp_path):
    mg = mgc.mg_params(nlevels=4, block=[(4, 4, 4, 4), (2, 2, 2, 2), (3, 3, 3, 3)],
                       smoother_solver=mgc.SOLVER_MR, setup_tol=2.5e-4, coarsest_iters=33)
    t = mgc.test_params()
    t.csw = 1.5
    t.config = b"gauge.lime"
    p = str(tmp_path / "params.xml")
    mgc.write_params_xml(p, mg, t)
    mg2, t2 = mgc.read_params_xml(p)
    for f, _ in mgc.MGParams._fields_:
        if f in ("block", "nb", "precision", "coarsen_mode", "seed"):
            continue  # not part of MultiGridParams
        assert getattr(mg2, f) == getattr(mg, f), f
    assert [list(mg2.block[l]) for l in range(3)] == [list(mg.block[l]) for l in range(3)]
    assert (t2.csw, t2.config) == (1.5, b"gauge.lime")


def test_xml_rejects_per_level_differences(tmp_path):
    s = open(os.path.join(GOLDEN, "mg_params_reference_format.xml")).read()
    s = s.replace("<elem>2</elem>\n        <elem>2</elem>", "<elem>2</elem>\n        <elem>3</elem>", 1)
    p = tmp_path / "bad.xml"
    p.write_text(s)
    with pytest.raises(mgc.MgcError, match="differs between levels"):
        mgc.read_params_xml(str(p))
    p.write_text("<grid><Params><mg><nLevels>3</nLevels>")
    with pytest.raises(mgc.MgcError, match="XML"):
        mgc.read_params_xml(str(p))


@pytest.fixture(scope="module")
def field():
    dims = (4, 2, 2, 6)
    U = oracle.su3_gauge(dims, 0.4, 123, 1).reshape(4, -1, 3, 3)
    return dims, U


@pytest.mark.parametrize("fmt", [mgc.GAUGE_NERSC, mgc.GAUGE_ILDG])
@pytest.mark.parametrize("bits", [64, 32])
def test_gauge_round_trip(tmp_path, field, fmt, bits):
    dims, U = field
    p = str(tmp_path / "cfg")
    mgc.write_gauge(p, dims, U, fmt, bits)
    d2, U2, plaq = mgc.read_gauge(p)
    assert d2 == dims
    if bits == 64:
        assert np.array_equal(U2.view(np.float64), U.view(np.float64))
    else:
        assert np.abs(U2 - U).max() < 1e-7
    assert abs(plaq - plaquette_np(U2, dims)) < 1e-13


def nersc_header(dims, plaq, trace, datatype, fp, checksum=None):
    h = ["BEGIN_HEADER", "HDR_VERSION = 1.0", f"DATATYPE = {datatype}"]
    h += [f"DIMENSION_{d + 1} = {dims[d]}" for d in range(4)]
    h += [f"PLAQUETTE = {plaq:.15f}", f"LINK_TRACE = {trace:.15f}", f"FLOATING_POINT = {fp}"]
    if checksum is not None:
        h.append(f"CHECKSUM = {checksum:x}")
    h.append("END_HEADER")
    return ("\n".join(h) + "\n").encode()


def test_nersc_two_row_file_written_in_numpy(tmp_path, field):
    """A 4D_SU3_GAUGE file (two stored rows per link, big endian, sites x
    fastest with the four links of a site together): the reader rebuilds
    row 3 = conj(row 1 x row 2), which equals the SU(3) field's own."""
    dims, U = field
    V = int(np.prod(dims))
    rows = np.ascontiguousarray(U.transpose(1, 0, 2, 3)[:, :, :2, :])  # [site][mu][2][3]
    data = rows.astype(">c16").tobytes()
    trace = np.einsum("mxii", U).real.sum() / (3 * 4 * V)
    p = tmp_path / "two_row.nersc"
    p.write_bytes(nersc_header(dims, plaquette_np(U, dims), trace, "4D_SU3_GAUGE", "IEEE64BIG")
                  + data)
    d2, U2, plaq = mgc.read_gauge(str(p))
    assert d2 == dims
    assert np.abs(U2 - U).max() < 1e-14
    # a header plaquette that does not match the data is an error
    p.write_bytes(nersc_header(dims, 0.123, trace, "4D_SU3_GAUGE", "IEEE64BIG") + data)
    with pytest.raises(mgc.MgcError, match="plaquette"):
        mgc.read_gauge(str(p))


def test_nersc_checksum_is_verified(tmp_path, field):
    dims, U = field
    p = str(tmp_path / "cfg.nersc")
    mgc.write_gauge(p, dims, U, mgc.GAUGE_NERSC, 64)
    raw = open(p, "rb").read()
    # the checksum (sum of the 32-bit words of the little-endian doubles)
    words = np.ascontiguousarray(U.transpose(1, 0, 2, 3)).astype("<c16").view("<u4")
    want = int(words.astype(np.uint64).sum() % (1 << 32))
    assert f"CHECKSUM = {want:x}\n".encode() in raw
    bad = raw.replace(f"CHECKSUM = {want:x}".encode(), f"CHECKSUM = {(want + 1) % (1 << 32):x}".encode())
    open(p, "wb").write(bad)
    with pytest.raises(mgc.MgcError, match="checksum"):
        mgc.read_gauge(p)


def lime_record(rtype, data, mb, me):
    flags = (0x8000 if mb else 0) | (0x4000 if me else 0)
    head = struct.pack(">IHHQ", 0x456789AB, 1, flags, len(data)) + rtype.encode().ljust(128, b"\0")
    return head + data + b"\0" * ((8 - len(data) % 8) % 8)


def test_ildg_file_written_in_numpy(tmp_path, field):
    """LIME records built with struct (record headers, 8-byte padding, an
    extra record the reader skips) around big-endian 3x3 links."""
    dims, U = field
    fmt = ("<?xml version='1.0'?><ildgFormat><version>1.0</version><field>su3gauge</field>"
           f"<precision>32</precision><lx>{dims[0]}</lx><ly>{dims[1]}</ly><lz>{dims[2]}</lz>"
           f"<lt>{dims[3]}</lt></ildgFormat>").encode()
    data = np.ascontiguousarray(U.transpose(1, 0, 2, 3)).astype(">c8").tobytes()
    blob = (lime_record("scidac-private-file-xml", b"<x>1</x>", True, False)
            + lime_record("ildg-format", fmt, False, False)
            + lime_record("ildg-binary-data", data, False, False)
            + lime_record("ildg-data-lfn", b"lfn://test", False, True))
    p = tmp_path / "cfg.lime"
    p.write_bytes(blob)
    d2, U2, _ = mgc.read_gauge(str(p))
    assert d2 == dims
    assert np.abs(U2 - U.astype(np.complex64)).max() == 0.0
