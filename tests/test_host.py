"""Host-side logic: the SIMsalabim reader, the grids, the device arrays, the packing for the
kernels, and that the C-ABI library loads and exports what include/ddhip.h declares."""

from __future__ import annotations

import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

from conftest import GOLDEN, ROOT, fixture_device


# ---- grids (reference tests test_solver.py:67-172) ------------------------------------------
def test_simss_grid_on_a_thin_stack():
    from jaxdd_amd.device import make_grid

    x, h, i0, i1, _ = make_grid([1e-9, 500e-9, 1e-9], 400, 4.0)
    assert len(x) == 402 and np.all(np.diff(x) > 0)
    assert abs(h[:-1].sum() - 1.0) < 1e-3
    assert i0 == [0, 6, 397] and i1 == [5, 396, 401]


@pytest.mark.parametrize(("stack", "simss", "exact"), [
    (("7e-10", "3.5e-07", "7e-10"), 389, 390),
    (("1.4e-09", "7e-07", "1.4e-09"), 389, 390),
    (("1e-07", "5e-07", "2e-07"), 250, 250),
    (("1e-06", "7e-07", "5e-07"), 127, 127),
])
def test_grid_split_on_a_whole_share(stack, simss, exact):
    """SimSS 5.36's absorber points where FLOOR meets a whole number (reference
    test_solver.py:77-94, from SimSS's own Var.dat)."""
    from jaxdd_amd.device import make_grid

    for split, absorber in (("simss", simss), ("exact", exact)):
        _, _, i0, i1, lid = make_grid(stack, 400, 4.0, split)
        assert i1[1] - i0[1] == absorber
        assert i1[-1] == 401 and np.all(np.diff(lid) >= 0)
    assert make_grid(("1e-07", "5e-07", "2e-07"), 400, 4.0, "simss")[3][0] == 49
    assert make_grid(("1e-07", "5e-07", "2e-07"), 400, 4.0, "exact")[3][0] == 50


def test_graded_grid_gives_every_layer_its_ends():
    from jaxdd_amd.grid import GRADED_SMALLEST, graded_layout

    stack = ["1e-09", "5e-07", "3e-08"]
    L = np.array([float(s) for s in stack])
    for NP in (400, 1600):
        lay = graded_layout(stack, NP)
        i0, i1, raw = lay.i0, lay.i1, lay.unit
        x, h = lay.place(L, L.sum())
        assert len(x) == NP + 2 and np.all(np.diff(x) > 0)
        assert i0[0] == 0 and i1[-1] == NP + 1
        assert all(b + 1 == a for a, b in zip(i0[1:], i1[:-1]))
        for k in range(3):
            end = x[i0[k + 1]] if k < 2 else L.sum()
            assert abs((end - x[i0[k]]) - L[k]) < 1e-3 * L[k] + 2 * h[i1[k]] * L.sum()
        absorber = raw[i0[1]:i1[1] + 1]
        assert absorber[1] < absorber[len(absorber) // 2] / 100
        assert raw[i0[1]] < 10 * GRADED_SMALLEST


# ---- the reader -------------------------------------------------------------------------
def test_setup_from_texts_builds_the_same_device():
    import dataclasses

    from jaxdd_amd.device import build_device
    from jaxdd_amd.simss_input import read_setup, setup_from_texts

    case = GOLDEN / "frozen_1e17_x1_thin_light"
    files = {f.name: f.read_text() for f in case.iterdir() if f.suffix == ".txt"}
    a = build_device(read_setup(case / "simulation_setup.txt"))
    b = build_device(setup_from_texts(files))
    for field in dataclasses.fields(a):
        assert np.array_equal(getattr(a, field.name), getattr(b, field.name)), field.name
    profile = setup_from_texts(files).raw["genProfile"]
    without = {k: v for k, v in files.items() if k != profile}
    with pytest.raises(FileNotFoundError, match="genProfile"):
        build_device(setup_from_texts(without))


def test_simss_voltages_of_the_fixtures():
    from jaxdd_amd.simss_input import read_setup, simss_voltages

    for case in ("thin_l1_light", "logv_light"):
        setup = read_setup(GOLDEN / case / "simulation_setup.txt")
        ref = np.genfromtxt(GOLDEN / case / "JV.dat", names=True)
        assert np.allclose(simss_voltages(setup), ref["Vint"], atol=1e-9)


def test_unsupported_input_is_refused_by_name(tmp_path):
    from jaxdd_amd.simss_input import check_supported, read_setup

    for f in (GOLDEN / "thin_l1_light").iterdir():
        (tmp_path / f.name).write_text(f.read_text())
    setup_file = tmp_path / "simulation_setup.txt"
    text = setup_file.read_text()
    setup_file.write_text(text.replace("currDiffInt = 2 ", "currDiffInt = 3 ", 1))
    with pytest.raises(NotImplementedError, match="currDiffInt"):
        check_supported(read_setup(setup_file))
    setup_file.write_text(text.replace("S_n_L = -1 ", "S_n_L = 20 ", 1))
    with pytest.raises(NotImplementedError, match="S_n_L"):
        check_supported(read_setup(setup_file))
    check_supported(read_setup(GOLDEN / "thin_l1_light" / "simulation_setup.txt"))


def test_held_ions_are_fixed_charges():
    dev = fixture_device("held_thin_light")
    assert not dev.anion_mask.any() and not dev.cation_mask.any()
    assert dev.anion_fixed.max() == 8e22 and dev.cation_fixed.max() == 1e23


# ---- packing for the kernels --------------------------------------------------------------
def test_pack_segments_reproduce_every_point():
    from jaxdd_amd.batch import SEG_FIELDS, _point_fields, pack

    devs = [fixture_device(c) for c in ("thin_l1_light", "steady_1e17_thin_light", "resist_light")]
    packed = pack(devs)
    n = len(devs)
    for k, d in enumerate(devs):
        fields = _point_fields(d)
        sid = packed["sid"][:, k]
        assert np.all(np.diff(sid.astype(int)) >= 0)
        got = packed["seg"][sid, :, k]
        assert np.array_equal(got, fields), "segment tables differ from the point values"
        assert np.array_equal(packed["pt"][:, 0, k], d.h)
        assert np.array_equal(packed["pt"][:, 2, k], d.G)
    assert packed["seg"].shape == (packed["nseg"], len(SEG_FIELDS), n)
    # three layers, the two interface intervals and the two contacts (no recombination there)
    assert packed["nseg"] == 7


def test_prepare_refuses_mixed_grid_sizes():
    from jaxdd_amd.batch import pack
    from jaxdd_amd.synthetic import c1_device

    with pytest.raises(ValueError, match="one grid size"):
        pack([fixture_device("thin_l1_light"), c1_device()])


def test_synthetic_table_is_stable():
    """The drawn parameter table's fingerprint (seed 20261107), and prefix stability."""
    from jaxdd_amd import synthetic as S

    t = S.draw(16)
    assert S.table_digest(t) == S.table_digest(S.draw(16))
    big = S.draw(64)
    assert all(np.array_equal(np.asarray(big[k])[:16], np.asarray(t[k])) for k in t)
    assert S.table_digest(S.draw(16384)) == SYNTHETIC_DIGEST


#: SHA-256 of the 16384-device table every workload draws from.
SYNTHETIC_DIGEST = "2c771161f99bc8b656e67ec91e5f116222f5fccb285f2cd7c558e1e53d7d8cf9"


# ---- the C-ABI library ---------------------------------------------------------------------
def _declared_symbols():
    text = (ROOT / "include" / "ddhip.h").read_text()
    return sorted(set(re.findall(r"\b(ddh_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from jaxdd_amd import _build, _lib

    if not _lib.LIB.exists():
        _build.build()
    lib = ctypes.CDLL(str(_lib.LIB))
    declared = _declared_symbols()
    assert set(declared) == set(_lib.SYMBOLS)
    for name in declared:
        assert hasattr(lib, name), name
    lib.ddh_abi_version.restype = ctypes.c_int
    assert lib.ddh_abi_version() == 5
    lib.ddh_workspace_bytes.restype = ctypes.c_size_t
    assert lib.ddh_workspace_bytes(16384, 400, 61, 1) > 0
    assert lib.ddh_workspace_bytes(0, 400, 61, 1) == 0


def test_library_rejects_bad_shapes_before_any_launch():
    """Argument checks happen on the host: a NULL or a bad size returns an error code."""
    from jaxdd_amd import _build, _lib

    if not _lib.LIB.exists():
        _build.build()
    lib = _lib.load()
    rc = lib.ddh_solve_batch(None, None, 0, None, None, None, None, None, None, 0, None)
    assert rc == -4
    assert b"NULL" in lib.ddh_last_error()
    rc = lib.ddh_block_sweep(4, 10, 4, 1, 1, 1, 1, 1, 0, 1, None)
    assert rc == -1
    # more segments than the LDS holds: refused before any device memory is touched
    b = _lib.Batch(64, 400, 15, 0, 1, 1, 1, 1)
    o = _lib.Opts(1e-10, 60, 60, 0, 0, 1, 0)
    rc = lib.ddh_solve_batch(ctypes.byref(b), 1, 61, ctypes.byref(o), 1, 1, 1, 1, 1, 1 << 40, None)
    assert rc == -1 and b"nseg" in lib.ddh_last_error()
    rc = lib.ddh_assemble(ctypes.byref(b), 1, 1, None, 0, 1, 1, 1, 1, 1, 1, None)
    assert rc == -1 and b"nseg" in lib.ddh_last_error()
    rc = lib.ddh_assemble(None, None, None, None, 0, None, None, None, None, None, None, None)
    assert rc == -4
    rc = lib.ddh_assemble_ions(ctypes.byref(b), 1, 1, 1, 0, 1, 1, 1, 1, 1, None)
    assert rc == -1 and b"nseg" in lib.ddh_last_error()
    rc = lib.ddh_assemble_ions(ctypes.byref(b), 1, 1, None, 0, 1, 1, 1, 1, 1, None)
    assert rc == -4 and b"NULL" in lib.ddh_last_error()
    # lanes: 1, 2 or 4; four lanes need NP >= 16
    b = _lib.Batch(64, 400, 7, 0, 1, 1, 1, 1)
    o = _lib.Opts(1e-10, 60, 60, 0, 0, 1, 0, 3)
    rc = lib.ddh_solve_batch(ctypes.byref(b), 1, 61, ctypes.byref(o), 1, 1, 1, 1, 1, 1 << 40, None)
    assert rc == -4 and b"lanes" in lib.ddh_last_error()
    b = _lib.Batch(64, 12, 7, 0, 1, 1, 1, 1)
    o = _lib.Opts(1e-10, 60, 60, 0, 0, 1, 0, 4)
    rc = lib.ddh_solve_batch(ctypes.byref(b), 1, 61, ctypes.byref(o), 1, 1, 1, 1, 1, 1 << 40, None)
    assert rc == -1 and b"NP >= 16" in lib.ddh_last_error()
    # the adjoint stages a device's arrays in LDS: a grid too fine for it is refused
    b = _lib.Batch(64, 5000, 7, 0, 1, 1, 1, 1)
    rc = lib.ddh_adjoint(ctypes.byref(b), 1, 61, 1, None, 1, 16, 1, 1, 1, 0, 1, 1, 1, 1 << 45,
                         None)
    assert rc == -1 and b"LDS" in lib.ddh_last_error()


def test_four_lanes_are_refused_for_mobile_ions_on_the_host():
    """solve_batch(lanes=4) and jv_gradients check the batch before any launch: mobile ions
    with four lanes and NP < 16 (ValueError)."""
    from jaxdd_amd import adjoint as A
    from jaxdd_amd import synthetic as S
    from jaxdd_amd.batch import from_packed, pack, solve_batch

    t = S.draw(3)
    ions = from_packed(pack([S.device_from_row(t, k, 40, True, False) for k in range(3)]), "cpu")
    with pytest.raises(ValueError, match="mobile ions"):
        solve_batch(ions, [0.0, 0.5], lanes=4)
    coarse = from_packed(pack([S.device_from_row(t, k, 12, False, False) for k in range(3)]), "cpu")
    with pytest.raises(ValueError, match="NP >= 16"):
        solve_batch(coarse, [0.0, 0.5], lanes=4)
    with pytest.raises(ValueError, match="lanes"):
        solve_batch(coarse, [0.0, 0.5], lanes=3)
    tan = {"pt": np.zeros((1,) + tuple(ions["_pt"].shape)), "seg": None, "sc": None}
    with pytest.raises(ValueError, match="mobile ions"):
        A.jv_gradients(ions, tan, [0.0, 0.5], lanes=4)
    with pytest.raises(ValueError, match="precision"):
        A.jv_gradients(ions, tan, [0.0, 0.5], precision="fp32")
