"""Parity of the HIP path (through the C-ABI, libddhip.so) with the CPU oracle and the SimSS
fixtures, on an MI355X. Sizes the oracle finishes in seconds are compared point by point;
the full 16384-device workload is checked through size-independent properties."""

from __future__ import annotations

import numpy as np
import pytest

from conftest import (CHAOTIC_STEPS, FIXTURES, chaotic_devices, curve_scale, fixture_device,
                      fixture_jv, flags_agree, short_points)

pytestmark = pytest.mark.gpu

V_FULL = np.round(np.arange(-0.2, 1.3 + 1e-9, 0.025), 6)


def _np(t):
    return t.detach().cpu().numpy()


# ---- the assembly operator (reference test pattern test_solver.py:1578-1617) --------------
@pytest.mark.parametrize("case", ["fixtures", "c3"])
def test_assemble_matches_oracle(torch_cuda, oracle, emu, case):
    """ddh_assemble against the oracle's `_assemble` and current at the same states, and
    against its own CPU replay."""
    from jaxdd_amd import synthetic as S
    from jaxdd_amd.batch import assemble, pack, prepare
    from test_kernel_emu import _asm_state

    rng = np.random.default_rng(11)
    if case == "fixtures":
        devs = [fixture_device(c) for c in ("thin_l1_light", "frozen_1e17_x1_thin_light")]
        logs, bern = None, "simss"
    else:
        devs = S.devices("c3", 5)
        logs = np.array([[-30.0, -31.0], [-29.0, -32.5], [-33.0, -28.0], [-31.0, -30.0],
                         [-30.5, -29.5]])
        bern = "exact"
    Vint = np.linspace(-0.2, 1.2, len(devs))
    y, fulls = _asm_state(devs, Vint, rng)
    out = {k: _np(v) for k, v in assemble(prepare(devs, "cuda"), y, Vint, logs, bern).items()}
    rep = emu.assemble_packed(pack(devs), y, Vint, logs, bern)
    shunt = np.array([1 / d.R_shunt if d.R_shunt > 0 else 0.0 for d in devs])
    for k, dev in enumerate(devs):
        la, lc = (0.0, 0.0) if logs is None else logs[k]
        Fo, Ao, Bo, Co, Uo = oracle.assemble(dev, fulls[k], la=la, lc=lc, bernoulli_name=bern)
        rs = np.max(np.abs(np.concatenate([Ao, Bo, Co], axis=2)), axis=2) + 1e-300
        assert np.all(np.abs(out["F"][k] - Fo) <= 1e-11 * (rs + np.abs(Fo)))
        for key, ref in (("A", Ao), ("B", Bo), ("C", Co)):
            assert np.all(np.abs(out[key][k] - ref) <= 1e-12 * rs[:, :, None]), key
        assert np.all(np.abs(out["U"][k] - Uo) <= 1e-12 * (np.abs(Uo) + rs[:, :, None]))
        Jo = oracle.current(dev, fulls[k]) + shunt[k] * Vint[k]
        assert abs(out["J"][k] - Jo) <= 1e-10 * (abs(Jo) + oracle.j_scale(dev))
        # the GPU and its CPU replay: the same arithmetic up to FMA contraction
        assert np.all(np.abs(out["A"][k] - rep[1][k]) <= 1e-13 * rs[:, :, None])


@pytest.mark.parametrize("bern", ["exact", "simss"])
def test_assemble_ions_matches_oracle(torch_cuda, oracle, emu, bern):
    """ddh_assemble_ions (the steady 5x5 system, ions as unknowns) against the oracle's
    `_assemble_ions` restatement and against its own CPU replay, at ions off their Boltzmann
    profiles."""
    from jaxdd_amd import synthetic as S
    from jaxdd_amd.batch import assemble_ions, pack, prepare
    from test_kernel_emu import _asm5_state, check_assemble_ions

    rng = np.random.default_rng(12)
    devs = S.devices("c3", 5) + [fixture_device("steady_1e17_thin_light")]
    devs = devs if bern == "exact" else devs[5:]
    logs = np.array([[-30.0, -31.0], [-29.0, -32.5], [-33.0, -28.0], [-31.0, -30.0],
                     [-30.5, -29.5], [40.0, 41.0]])[-len(devs):]
    Vint = np.linspace(-0.2, 1.2, 6)[-len(devs):]
    y, fulls = _asm5_state(devs, Vint, logs, rng)
    out = assemble_ions(prepare(devs, "cuda"), y, Vint, logs, bern)
    got = tuple(_np(out[k]) for k in ("F", "A", "B", "C", "U"))
    check_assemble_ions(got, devs, fulls, logs, oracle, bern)
    rep = emu.assemble_ions_packed(pack(devs), y, Vint, logs, bern)
    rs = np.max(np.abs(np.concatenate(rep[1:4], axis=3)), axis=3) + 1e-300
    for g, r in zip(got[1:4], rep[1:4]):  # the same arithmetic up to FMA contraction
        assert np.all(np.abs(g - r) <= 1e-13 * rs[..., None])


# ---- the block sweep operator (reference tests test_solver.py:1620-1641, 1753-1773) ------
@pytest.mark.parametrize("b,k", [(3, 1), (3, 3), (5, 4), (3, 8)])
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("mixed", 2e-5)])
def test_block_sweep_is_the_scan(torch_cuda, oracle, b, k, precision, tol):
    from jaxdd_amd.sweep import block_sweep

    torch = torch_cuda
    rng = np.random.default_rng(3)
    n, NP = 70, 12  # not a whole number of wavefronts
    A, C = rng.normal(size=(n, NP, b, b)), rng.normal(size=(n, NP, b, b))
    B = rng.normal(size=(n, NP, b, b)) + 8 * np.eye(b)
    d = rng.normal(size=(n, NP, b, k))
    ref = np.stack([oracle.block_thomas(A[i], B[i], C[i], d[i]) for i in range(n)])
    x = block_sweep(*(torch.as_tensor(t, device="cuda") for t in (A, B, C, d)), precision=precision)
    assert np.max(np.abs(_np(x) - ref)) < tol * np.max(np.abs(ref))


# ---- the fixtures: oracle and SimSS -----------------------------------------------------
@pytest.fixture(scope="module")
def fixtures_on_gpu(torch_cuda, oracle):
    from jaxdd_amd.batch import prepare

    cases = sorted(FIXTURES)
    devs = [fixture_device(c) for c in cases]
    ref = oracle.solve_batch(devs, V_FULL, bernoulli_name="simss", first_max_iter=None)
    return cases, devs, prepare(devs), ref


@pytest.mark.parametrize("lanes", [1, 2])
@pytest.mark.parametrize("precision,jtol", [("fp64", 1e-12), ("mixed", 1e-9)])
def test_solve_batch_matches_oracle_on_fixtures(fixtures_on_gpu, precision, jtol, lanes):
    from jaxdd_amd.batch import solve_batch

    cases, devs, batch, (Jo, oko, itso) = fixtures_on_gpu
    J, ok, its = solve_batch(batch, V_FULL, bernoulli_name="simss", precision=precision,
                             lanes=lanes)
    J, ok, its = _np(J), _np(ok), _np(its)
    assert np.array_equal(ok, oko)
    if precision == "fp64":
        assert np.array_equal(its, itso)
    for k, dev in enumerate(devs):
        assert np.max(np.abs(J[k] - Jo[k])) < jtol * curve_scale(dev, Jo[k]), cases[k]


@pytest.mark.parametrize("case", sorted(FIXTURES))
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_solve_jv_reproduces_simss(torch_cuda, case, precision):
    from jaxdd_amd.solver import solve_jv

    dev = fixture_device(case)
    ref = fixture_jv(case)
    res = solve_jv(dev, ref["Vint"], bernoulli_name="simss", precision=precision)
    assert res.converged.all()
    scale = np.max(np.abs(ref["Jext"]))
    assert np.max(np.abs(res.J - ref["Jext"])) < FIXTURES[case] * scale
    if case == "resist_light":
        assert np.max(np.abs(res.V_ext - ref["Vext"])) < 1e-4
    if case == "thin_l1_voc_light" and precision == "fp64":
        from conftest import open_circuit_voltage

        assert abs(open_circuit_voltage(ref["Vext"], res.J)
                   - open_circuit_voltage(ref["Vext"], ref["Jext"])) < 1e-10


def test_zimt_steady_states_through_the_steady_solver(torch_cuda, oracle):
    """ZimT 5.36's steady states (`zimt_steady_light`: t = 0 on each of seven lines from -0.2
    to 1.1 V, the last at half light) by `solve_jv` on the GPU, one line at a time as the
    reference's own test solves them (tests/test_solver.py:454-470): ZimT's Jext to 1e-9 of
    its scale, and the oracle's steps and currents (1e-12)."""
    import dataclasses

    from conftest import GOLDEN
    from jaxdd_amd.device import build_device
    from jaxdd_amd.simss_input import check_supported, read_setup, read_tvg
    from jaxdd_amd.solver import solve_jv

    setup = read_setup(GOLDEN / "zimt_steady_light" / "simulation_setup.txt")
    check_supported(setup)
    tvg = read_tvg(setup)
    assert np.all(tvg.t == 0) and not tvg.oc.any()
    ref = np.genfromtxt(GOLDEN / "zimt_steady_light" / "tj.dat", names=True)
    dev = build_device(setup, grid="simss")
    scale = np.max(np.abs(ref["Jext"]))
    for v, g, want in zip(tvg.Vext, tvg.G_frac, ref["Jext"]):
        d = dataclasses.replace(dev, G_frac=float(g))
        res = solve_jv(d, np.array([v]), bernoulli_name="simss")
        Jo, oko, itso = oracle.solve_jv(d, np.array([v]), bernoulli_name="simss")
        assert res.converged.all() and np.array_equal(res.iterations, itso)
        assert abs(res.J[0] - Jo[0]) < 1e-12 * scale
        assert abs(res.J[0] - want) < 1e-9 * scale, (v, g)


def test_solve_jv_matches_oracle(torch_cuda, oracle):
    from jaxdd_amd.solver import solve_jv

    for case in ("thin_l1_light", "frozen_1e17_x1_thin_light", "held_thin_light"):
        dev = fixture_device(case)
        V = V_FULL[::-1].copy()
        Jo, oko, itso = oracle.solve_jv(dev, V, bernoulli_name="simss")
        res = solve_jv(dev, V, bernoulli_name="simss")
        assert np.array_equal(res.converged, oko)
        assert np.array_equal(res.iterations, itso)
        assert np.max(np.abs(res.J - Jo)) < 1e-12 * curve_scale(dev, Jo)


def test_solve_jv_noise_floor_stop(torch_cuda, oracle):
    """At tol 1e-14 every solve_jv solve ends at the round-off floor or below its noise floor
    (src/jaxdd/solver.py:1728-1751, 2162-2166): the GPU's flags are the oracle's and its
    curve within 1e-10 of the scale."""
    from jaxdd_amd import synthetic as S
    from jaxdd_amd.solver import solve_jv

    V = S.VOLTAGES[::3].copy()
    for dev in [fixture_device("thin_l1_light")] + S.devices("c2", 2):
        Jo, oko, itso = oracle.solve_jv(dev, V, tol=1e-14, max_iter=200)
        res = solve_jv(dev, V, tol=1e-14)
        assert np.array_equal(res.converged, oko)
        assert np.max(np.abs(res.J - Jo)) < 1e-10 * curve_scale(dev, Jo)
        assert np.max(np.abs(res.iterations - itso)) <= 2


def test_solve_jv_raises_on_a_failed_pre_bias(torch_cuda):
    from jaxdd_amd.solver import solve_jv

    dev = fixture_device("frozen_1e17_x1_thin_light")
    with pytest.raises(RuntimeError, match="pre-bias"):
        solve_jv(dev, V_FULL[:3], max_iter=1)


def test_voltage_order_and_g_frac(torch_cuda, fixtures_on_gpu):
    """Results come back in the caller's voltage order; dict(batch, g_frac=...) is honoured
    (reference tests test_solver.py:1521-1553)."""
    from jaxdd_amd.batch import solve_batch

    _, devs, batch, _ = fixtures_on_gpu
    up = [_np(a) for a in solve_batch(batch, V_FULL)]
    down = [_np(a) for a in solve_batch(batch, V_FULL[::-1])]
    for a, b in zip(up, down):
        assert np.array_equal(a, b[:, ::-1])
    dark = dict(batch, g_frac=batch["g_frac"] * 0.0)
    J, ok, _ = solve_batch(dark, np.array([0.0, 0.4]))
    assert _np(ok).all()
    scale = _np(batch["J_scale"])
    assert np.all(np.abs(_np(J)[:, 0]) < 1e-5 * scale)  # no light, no bias: no current


def test_deterministic(torch_cuda, fixtures_on_gpu):
    from jaxdd_amd.batch import solve_batch

    _, _, batch, _ = fixtures_on_gpu
    a = [_np(t) for t in solve_batch(batch, V_FULL, precision="mixed")]
    b = [_np(t) for t in solve_batch(batch, V_FULL, precision="mixed")]
    for x, y in zip(a, b):
        assert np.array_equal(x, y, equal_nan=True)


# ---- the synthetic workloads against the oracle --------------------------------------------
@pytest.mark.parametrize("lanes", [1, 2])
@pytest.mark.parametrize("config,variant,n", [("c2", "free", 48), ("c3", "free", 48),
                                              ("c3", "frozen", 48)])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_synthetic_matches_oracle(torch_cuda, oracle, config, variant, n, precision, lanes):
    from jaxdd_amd import synthetic as S
    from jaxdd_amd.batch import prepare, solve_batch

    devs = S.devices(config, n, variant=variant)
    Jo, oko, itso = oracle.solve_batch(devs, S.VOLTAGES, first_max_iter=None)
    J, ok, its = (_np(t) for t in solve_batch(prepare(devs), S.VOLTAGES, precision=precision,
                                              lanes=lanes))
    wild = chaotic_devices(oracle, devs, S.VOLTAGES, Jo)
    assert wild.mean() <= 0.05
    agree, chaotic = flags_agree(ok[~wild], its[~wild], oko[~wild], itso[~wild])
    assert agree and chaotic <= 0.02 * ok[~wild].size
    calm = (its < CHAOTIC_STEPS) & (itso < CHAOTIC_STEPS) & ~wild[:, None]
    short = short_points(devs, S.VOLTAGES, its, itso, wild)
    differ = int(np.sum(its[calm] != itso[calm]))
    differ_short = int(np.sum(its[short] != itso[short]))
    print(f"{config}-{variant} {precision} lanes={lanes}: devices excluded as chaotic "
          f"{np.flatnonzero(wild).tolist()}; Newton steps differ from the oracle's on "
          f"{differ_short} of {int(short.sum())} short solves, {differ} of {int(calm.sum())} "
          "calm points", flush=True)
    if precision == "fp64":  # every short solve takes the oracle's steps
        assert differ_short == 0
    else:  # FP32 factors: an inexact Newton, a step more or less here and there
        assert differ <= 0.05 * calm.sum()
    both = ok & oko
    for k, dev in enumerate(devs):
        if wild[k]:
            # the oracle itself moves here when the generation changes by 1e-13: the GPU's
            # curve must still be a solve of the device, finite where converged
            assert np.all(np.isfinite(J[k][ok[k]]))
            continue
        err = np.max(np.abs(J[k] - Jo[k])[both[k]], initial=0.0)
        assert err < 1e-6 * curve_scale(dev, Jo[k])  # the workload's accuracy bar


@pytest.mark.parametrize("lanes", [1, 2])
@pytest.mark.parametrize("ions", [False, True])
def test_odd_grid_matches_oracle(torch_cuda, oracle, ions, lanes):
    """NP = 301 interior points (the two-lane split's halves differ by a row), 37 devices (a
    partial wavefront): the oracle's flags and iterations, and its curves."""
    from jaxdd_amd import synthetic as S
    from jaxdd_amd.batch import prepare, solve_batch

    t = S.draw(37)
    devs = [S.device_from_row(t, k, 301, ions, False) for k in range(37)]
    V = S.VOLTAGES[::4].copy()
    Jo, oko, itso = oracle.solve_batch(devs, V, first_max_iter=None)
    J, ok, its = (_np(x) for x in solve_batch(prepare(devs), V, lanes=lanes))
    wild = chaotic_devices(oracle, devs, V, Jo)
    assert wild.mean() <= 0.05
    agree, chaotic = flags_agree(ok[~wild], its[~wild], oko[~wild], itso[~wild])
    assert agree and chaotic <= 0.02 * ok[~wild].size
    short = short_points(devs, V, its, itso, wild)
    differ = int(np.sum(its[short] != itso[short]))
    print(f"NP=301 ions={ions} lanes={lanes}: excluded {np.flatnonzero(wild).tolist()}; "
          f"steps differ on {differ} of {int(short.sum())} short solves", flush=True)
    assert differ == 0
    for k, dev in enumerate(devs):
        if wild[k]:
            continue
        both = ok[k] & oko[k]
        assert np.max(np.abs(J[k] - Jo[k])[both], initial=0.0) < 1e-9 * curve_scale(dev, Jo[k])


# ---- four lanes per device: the grid in quarters (devices without mobile ions) -------------
@pytest.mark.parametrize("precision,jtol", [("fp64", 1e-12), ("mixed", 1e-9)])
def test_four_lanes_match_oracle_on_fixtures(fixtures_on_gpu, precision, jtol):
    """ddh_solve_batch with opts.lanes = 4 on the whole fixture batch: the devices without
    mobile ions get the oracle's flags, in FP64 its iterations, and its curves; the others are
    refused by the kernel (status 2, no converged voltage, NaN currents)."""
    from jaxdd_amd import _lib
    from jaxdd_amd.batch import SC_FIELDS, _run

    cases, devs, batch, (Jo, oko, itso) = fixtures_on_gpu
    sc = _np(batch["_sc"])
    ions = (sc[SC_FIELDS.index("a_total")] > 0) | (sc[SC_FIELDS.index("c_total")] > 0)
    assert ions.any() and not ions.all()
    opts = _lib.Opts(1e-10, 60, 60, 0, 1, _lib.PRECISIONS[precision], 0, 4)
    J, ok, its, status = (_np(t) for t in _run(batch, V_FULL, opts, _lib.PRECISIONS[precision]))
    shunt = _np(batch["G_shunt"])
    J = J + shunt[:, None] * V_FULL[None, :]
    assert np.array_equal(status, np.where(ions, 2, 0))
    assert not ok[ions].any() and np.isnan(J[ions]).all() and not its[ions].any()
    free = np.flatnonzero(~ions)
    assert np.array_equal(ok[free].astype(bool), oko[free])
    if precision == "fp64":
        assert np.array_equal(its[free], itso[free])
    for k in free:
        assert np.max(np.abs(J[k] - Jo[k])) < jtol * curve_scale(devs[k], Jo[k]), cases[k]


@pytest.mark.parametrize("np_,n", [(400, 48), (301, 37), (16, 21)])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_four_lanes_match_oracle_on_synthetic(torch_cuda, oracle, np_, n, precision):
    """The gradient workload's devices (no ions) on its grid (NP = 400), an odd grid (301:
    quarters of 75, 75, 75 and 76 rows) and the smallest the split takes (16), in partial
    wavefronts (16 devices a wavefront): the oracle's flags and iterations, and its curves."""
    from jaxdd_amd import synthetic as S
    from jaxdd_amd.batch import prepare, solve_batch

    t = S.draw(n)
    devs = [S.device_from_row(t, k, np_, False, False) for k in range(n)]
    V = S.VOLTAGES[::2].copy()
    Jo, oko, itso = oracle.solve_batch(devs, V, first_max_iter=None)
    J, ok, its = (_np(x) for x in solve_batch(prepare(devs), V, precision=precision, lanes=4))
    wild = chaotic_devices(oracle, devs, V, Jo)
    assert wild.mean() <= 0.05
    agree, chaotic = flags_agree(ok[~wild], its[~wild], oko[~wild], itso[~wild])
    assert agree and chaotic <= 0.02 * ok[~wild].size
    bar = 1e-9 if precision == "fp64" else 1e-6  # the workload's accuracy bar in mixed
    for k, dev in enumerate(devs):
        if wild[k]:
            continue
        both = ok[k] & oko[k]
        assert np.max(np.abs(J[k] - Jo[k])[both], initial=0.0) < bar * curve_scale(dev, Jo[k])


def test_one_device_workload_c1(torch_cuda, oracle):
    from jaxdd_amd import synthetic as S
    from jaxdd_amd.solver import solve_jv

    dev = S.c1_device()
    Jo, oko, itso = oracle.solve_jv(dev, S.VOLTAGES)
    res = solve_jv(dev, S.VOLTAGES)
    assert res.converged.all() and oko.all()
    assert np.max(np.abs(res.J - Jo)) < 1e-10 * curve_scale(dev, Jo)


def test_nan_input_ends_unconverged(torch_cuda):
    import dataclasses

    from jaxdd_amd.batch import prepare, solve_batch

    good = fixture_device("thin_l1_light")
    G = good.G.copy()
    G[200] = np.nan
    bad = dataclasses.replace(good, G=G)
    J, ok, its = (_np(t) for t in solve_batch(prepare([bad, good]), V_FULL[::6].copy()))
    assert not ok[0].any() and ok[1].all()
    assert its[0].max() <= 1


# ---- the full-size workload: size-independent properties ---------------------------------
def test_full_batch_properties(torch_cuda):
    """4096 devices of the ion workload: every converged voltage's current finite, the short-circuit
    current never above the generation's, mixed within 1e-6 of the all-FP64 solve per
    converged point, the same converged flags away from chaotic solves, and a second solve
    bit-identical to the first."""
    from jaxdd_amd import synthetic as S
    from jaxdd_amd.batch import from_packed, solve_batch

    # built in this process: no fork after the GPU is initialised
    batch = from_packed(S.packed("c3", 4096, variant="free", workers=1))
    J64, ok64, it64 = (_np(t) for t in solve_batch(batch, S.VOLTAGES))
    Jm, okm, itm = (_np(t) for t in solve_batch(batch, S.VOLTAGES, precision="mixed"))
    Jm2 = _np(solve_batch(batch, S.VOLTAGES, precision="mixed")[0])
    assert np.array_equal(Jm, Jm2, equal_nan=True)
    # one lane per device against two: the same curves to round-off on every calm point
    J1, ok1, it1 = (_np(t) for t in solve_batch(batch, S.VOLTAGES, lanes=1))
    agree1, chaotic1 = flags_agree(ok1, it1, ok64, it64)
    assert agree1 and chaotic1 <= 0.02 * ok1.size
    assert np.isfinite(J64[ok64]).all()
    assert ok64.mean() > 0.99
    agree, chaotic = flags_agree(okm, itm, ok64, it64)
    assert agree and chaotic <= 0.02 * okm.size
    scale = _np(batch["J_scale"])[:, None] + np.max(np.abs(J64), axis=1, keepdims=True)
    both = okm & ok64
    assert np.max((np.abs(Jm - J64) / scale)[both]) < 1e-6
    i0 = int(np.argmin(np.abs(S.VOLTAGES)))
    assert np.all(-J64[:, i0][ok64[:, i0]] <= _np(batch["J_scale"])[ok64[:, i0]] * (1 + 1e-9))


# ---- the adjoint: dJ/dtheta on the GPU ----------------------------------------------------
@pytest.fixture(scope="module")
def c4_gpu(torch_cuda):
    from jaxdd_amd import adjoint as A
    from jaxdd_amd import synthetic as S

    th = A.theta_from_table(S.draw(4), range(4))
    packed, tan = A.tangents(th)
    V = S.VOLTAGES[::6].copy()
    J, ok, its, grad, states, _ = A.jv_gradients(packed, tan, V, tol=1e-12, strict=True,
                                                 max_iter=200, return_states=True)
    return th, packed, tan, V, _np(J), _np(ok), _np(grad), _np(states)


def test_adjoint_kernel_is_its_cpu_replay(c4_gpu, emu):
    """The HIP adjoint through the C-ABI against the same per-lane code on the CPU, both
    linearised at the GPU's converged states. The bar is set by the conditioning at high
    forward bias, where a 1e-15 change of the state moves single gradient components by up to
    1e-4 (the GPU's and the host's exp and log differ in the last bit)."""
    th, packed, tan, V, J, ok, grad, states = c4_gpu
    assert ok.all()
    ge, status = emu.adjoint_packed(packed, V, np.ascontiguousarray(states), ok, tan)
    assert not status.any()
    rng = np.random.default_rng(0)
    w, d = rng.normal(size=ok.shape), rng.normal(size=th.shape)
    scale = th.copy()
    scale[:, 15] = 0.01
    a = np.sum(w[:, :, None] * grad * (scale * d)[:, None, :])
    b = np.sum(w[:, :, None] * ge * (scale * d)[:, None, :])
    assert abs(a - b) < 1e-6 * abs(b)


def test_adjoint_is_the_finite_difference_of_the_gpu_curves(c4_gpu):
    """The reference's gradient test (tests/test_solver.py:1784-1811) on the GPU: a weighted sum
    of the curves, a random direction in theta, the adjoint's directional derivative against
    central differences of the GPU's own FP64 curves."""
    from jaxdd_amd import adjoint as A
    from jaxdd_amd.batch import prepare, solve_batch

    th, packed, tan, V, J, ok, grad, _ = c4_gpu
    rng = np.random.default_rng(1)
    w, d = rng.normal(size=ok.shape), rng.normal(size=th.shape)
    scale = th.copy()
    scale[:, 15] = 0.01
    by_adjoint = np.sum(w[:, :, None] * grad * (scale * d)[:, None, :])

    structs = [A.structure_of(t) for t in th]  # each device's layout, held as theta moves

    def loss(a):
        x = th * (1.0 + a * d)
        x[:, 15] = th[:, 15] + a * d[:, 15] * 0.01
        batch = prepare([A.device_from_theta(t, structure=st) for t, st in zip(x, structs)])
        Jx, okx, _ = solve_batch(batch, V, tol=1e-12, strict=True, max_iter=200)
        assert _np(okx).all()
        return float(np.sum(w * _np(Jx)))

    h = 1e-5
    fd = (loss(h) - loss(-h)) / (2 * h)
    assert abs(by_adjoint - fd) < 1e-5 * abs(fd)


def test_adjoint_directional_derivatives_of_200_devices(torch_cuda):
    """The gradient workload's check at its own size (200 devices): per device, the adjoint's
    derivative along a random direction in the 16 parameters (one tangent direction) against
    central differences of the GPU's FP64 curves, every voltage where all three solves
    converged, measured against the curve's current scale: every device to 2e-5, half of them
    to 1e-6. Near open circuit the adjoint's two terms nearly cancel, so the derivative moves
    with the converged state within the solve's tolerance: for device 0 at 1.225 V the adjoints
    at the states of its one-, two- and four-lane solves differ by up to 7.6e-6 of the scale
    (the kernels' CPU replay), and the differences of the curves lie 1.2e-5 from the GPU's."""
    from jaxdd_amd import adjoint as A
    from jaxdd_amd import synthetic as S
    from jaxdd_amd.batch import j_scale, pack, prepare, solve_batch

    n = 200
    th = A.theta_from_table(S.draw(n), range(n))
    d = np.random.default_rng(5).normal(size=th.shape)

    def theta_at(a):
        x = th * (1.0 + a * d)
        x[:, 15] = th[:, 15] + a * d[:, 15] * 0.01  # the work-function offset: 10 mV units
        return x

    structs = [A.structure_of(t) for t in th]  # each device's layout, held as theta moves

    def devices_at(a):
        return [A.device_from_theta(t, structure=st) for t, st in zip(theta_at(a), structs)]

    devs0 = devices_at(0.0)
    base = pack(devs0)
    e = A.REL_STEP
    hi, lo = (pack(devices_at(s)) for s in (e, -e))
    for side in (hi, lo):
        assert side["nseg"] == base["nseg"] and np.array_equal(side["sid"], base["sid"])
    tan = {k: ((hi[k] - lo[k]) / (2 * e))[None] for k in ("pt", "seg", "sc")}
    V = S.VOLTAGES[::3].copy()
    J, ok, _, grad = A.jv_gradients(base, tan, V, tol=1e-12, strict=True, max_iter=200)
    grad, ok = _np(grad)[:, :, 0], _np(ok)

    h = 1e-5
    curves = []
    for s_ in (h, -h):
        # the kernel the adjoint's states came from (four lanes per device), so that the
        # differences see the same Newton steps and stops
        Jx, okx, _ = solve_batch(prepare(devices_at(s_)), V, tol=1e-12, strict=True,
                                 max_iter=200, lanes=4)
        curves.append((_np(Jx), _np(okx)))
    fd = (curves[0][0] - curves[1][0]) / (2 * h)
    both = ok & curves[0][1] & curves[1][1]
    assert both.mean() > 0.95
    scale = np.array([j_scale(dv) for dv in devs0]) + np.max(np.abs(_np(J)), axis=1)
    err = np.max(np.where(both, np.abs(grad - fd), 0.0), axis=1) / scale
    print(f"per-device error quantiles 50/90/100%: {np.quantile(err, [0.5, 0.9, 1.0])}", flush=True)
    assert np.max(err) < 2e-5, (np.argmax(err), np.max(err))
    assert np.median(err) < 1e-6


def test_gradients_of_the_mixed_precision_solve(c4_gpu):
    """The solve with its Newton factors in FP32 converges to the FP64 solve's states (its
    residual is FP64), so the adjoint linearised there gives the same gradients: the same
    directional metric as the replay test, against the FP64 solve's gradients."""
    from jaxdd_amd import adjoint as A

    th, packed, tan, V, J, ok, grad, states = c4_gpu
    Jm, okm, _, gm, sm, _ = A.jv_gradients(packed, tan, V, tol=1e-12, strict=True, max_iter=200,
                                           precision="mixed", return_states=True)
    assert _np(okm).all()
    assert np.max(np.abs(_np(sm) - states)) < 1e-9
    assert np.max(np.abs(_np(Jm) - J)) <= 1e-9 * np.max(np.abs(J))
    rng = np.random.default_rng(3)
    w, d = rng.normal(size=ok.shape), rng.normal(size=th.shape)
    scale = th.copy()
    scale[:, 15] = 0.01
    a = np.sum(w[:, :, None] * _np(gm) * (scale * d)[:, None, :])
    b = np.sum(w[:, :, None] * grad * (scale * d)[:, None, :])
    assert abs(a - b) < 1e-6 * abs(b)


def test_autograd_function_backward_is_the_adjoint(c4_gpu):
    import torch

    from jaxdd_amd import adjoint as A

    th, packed, tan, V, J, ok, grad, _ = c4_gpu
    theta = torch.tensor(th, dtype=torch.float64, requires_grad=True)
    Jt, okt = A.curves(theta, V, tol=1e-12)
    w = torch.as_tensor(np.random.default_rng(2).normal(size=Jt.shape), dtype=torch.float64)
    (w * Jt).sum().backward()
    # the same curves' adjoint with the tangents built on the GPU, as the Function builds them
    batch, gtan = A.gpu_tangents(th)
    _, _, _, g = A.jv_gradients(batch, gtan, V, tol=1e-12)
    expect = np.einsum("nv,nvk->nk", w.numpy(), _np(g))
    got = theta.grad.numpy()
    assert np.allclose(got, expect, rtol=1e-9, atol=0.0)


def test_charged_traps_gradients_and_the_impedance_refusal(torch_cuda, emu):
    """A batch holding charged bulk traps (SimSS's run `bulk_acceptor_light`): its gradients
    in the traps' numbers (`jv_gradients`, host tangents) are the CPU replay's to 1e-9 of their
    scale (the replay's against the oracle's differences: test_adjoint.py); the impedance
    refuses the batch before any launch."""
    from conftest import bumped, fixture_inputs, number_of
    from jaxdd_amd.adjoint import jv_gradients
    from jaxdd_amd.batch import pack, prepare
    from jaxdd_amd.device import device_arrays
    from jaxdd_amd.impedance import solve_impedance
    from test_adjoint import CHARGED_WRT

    st, p = fixture_inputs("bulk_acceptor_light")
    packed = pack([device_arrays(st, p)])
    tan = {key: np.zeros((len(CHARGED_WRT),) + packed[key].shape) for key in ("pt", "seg", "sc")}
    for k, w in enumerate(CHARGED_WRT):
        h = 1e-6 * abs(number_of(p, w))
        up, dn = (pack([device_arrays(st, bumped(p, w, s * h))]) for s in (1.0, -1.0))
        for key in tan:
            tan[key][k] = (up[key] - dn[key]) / (2 * h)
    V = np.array([-0.2, 0.3, 0.6, 0.8, 0.95, 1.05])
    print("charged traps: jv_gradients on the GPU", flush=True)
    J, ok, _, g = jv_gradients(packed, tan, V, tol=1e-12, strict=True, max_iter=200, lanes=2)
    J, ok, g = _np(J), _np(ok), _np(g)
    states = np.zeros((len(V), st.NP, 3, 1))
    Jr, okr, _, _ = emu.solve_packed(packed, V, tol=1e-12, strict=True, max_iter=200,
                                     states=states)
    gr, status = emu.adjoint_packed(packed, V, states, okr, tan)
    assert ok.all() and okr.all() and not status.any()
    assert np.max(np.abs(g - gr)) <= 1e-9 * np.max(np.abs(gr))
    with pytest.raises(NotImplementedError, match="charged"):
        solve_impedance(prepare([device_arrays(st, p)]), [1e3])
