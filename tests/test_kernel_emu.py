"""The HIP kernels' per-lane logic (jaxdd_amd/csrc/ddhip_core.h), built for the CPU by
tests/kernel_emu and replayed wavefront by wavefront, against the oracle. These run without a
GPU; the same comparisons on the GPU are in test_gpu.py."""

from __future__ import annotations

import numpy as np
import pytest

from conftest import FIXTURES, chaotic_devices, curve_scale, fixture_device, flags_agree

V_FULL = np.round(np.arange(-0.2, 1.3 + 1e-9, 0.025), 6)


@pytest.fixture(scope="module")
def fixture_runs(oracle):
    from jaxdd_amd.batch import pack

    cases = sorted(FIXTURES)
    devs = [fixture_device(c) for c in cases]
    ref = oracle.solve_batch(devs, V_FULL, bernoulli_name="simss", first_max_iter=None)
    shunt = np.array([1 / d.R_shunt if d.R_shunt > 0 else 0.0 for d in devs])
    return cases, devs, pack(devs), (ref[0] - shunt[:, None] * V_FULL, ref[1], ref[2])


@pytest.mark.parametrize("lanes", [1, 2])
@pytest.mark.parametrize("precision,jtol", [("fp64", 1e-12), ("mixed", 1e-9)])
def test_replay_matches_oracle_on_fixtures(emu, fixture_runs, precision, jtol, lanes):
    """One lane per device (the reference's elimination order) and two (the solve split at the
    grid's middle): the oracle's flags, in FP64 its iterations, and its curves."""
    cases, devs, packed, (Jo, oko, itso) = fixture_runs
    J, ok, its, status = emu.solve_packed(packed, V_FULL, bernoulli="simss", precision=precision,
                                          lanes=lanes)
    assert not status.any()
    assert np.array_equal(ok, oko)
    if precision == "fp64":
        assert np.array_equal(its, itso)  # the same Newton steps, voltage by voltage
    for k, dev in enumerate(devs):
        assert np.max(np.abs(J[k] - Jo[k])) < jtol * curve_scale(dev, Jo[k]), cases[k]


@pytest.mark.parametrize("lanes", [1, 2])
@pytest.mark.parametrize("config,variant", [("c2", "free"), ("c3", "free"), ("c3", "frozen")])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_replay_matches_oracle_on_synthetic(emu, oracle, config, variant, precision, lanes):
    from jaxdd_amd import synthetic as S
    from jaxdd_amd.batch import pack

    devs = S.devices(config, 12, variant=variant)
    Jo, oko, itso = oracle.solve_batch(devs, S.VOLTAGES, first_max_iter=None)
    J, ok, its, _ = emu.solve_packed(pack(devs), S.VOLTAGES, precision=precision, lanes=lanes)
    wild = chaotic_devices(oracle, devs, S.VOLTAGES, Jo)
    assert wild.mean() <= 0.1
    agree, chaotic = flags_agree(ok[~wild], its[~wild], oko[~wild], itso[~wild])
    assert agree and chaotic <= 0.02 * ok[~wild].size
    both = ok & oko
    for k, dev in enumerate(devs):
        if wild[k]:
            continue
        err = np.max(np.abs(J[k] - Jo[k])[both[k]], initial=0.0)
        assert err < 1e-9 * curve_scale(dev, Jo[k])


def test_replay_solve_jv_mode_matches_oracle(emu, oracle):
    from jaxdd_amd.batch import pack

    for case in ("thin_l1_light", "frozen_1e17_x1_thin_light", "steady_1e17_thin_light"):
        dev = fixture_device(case)
        V = V_FULL[::-1].copy()  # the given order, not sorted
        Jo, oko, itso = oracle.solve_jv(dev, V, bernoulli_name="simss")
        J, ok, its, status = emu.solve_packed(pack([dev]), V, bernoulli="simss", jv=True,
                                              max_iter=200)
        assert not status.any()
        assert np.array_equal(ok[0], oko)
        assert np.max(np.abs(J[0] - Jo)) < 1e-10 * curve_scale(dev, Jo)


def test_replay_pre_bias_failure_is_reported(emu):
    """jv mode: a pre-bias that cannot converge (one Newton step allowed) sets the status the
    host turns into RuntimeError, as the reference's solve_jv raises."""
    from jaxdd_amd.batch import pack

    dev = fixture_device("frozen_1e17_x1_thin_light")
    _, _, _, status = emu.solve_packed(pack([dev]), V_FULL[:3], jv=True, max_iter=1)
    assert status[0] == 1


def test_replay_nan_input_ends_unconverged(emu):
    """A NaN in a device's generation ends every Newton solve at its first non-number step:
    the lane finishes, its flags False, the other lanes untouched."""
    import dataclasses

    from jaxdd_amd.batch import pack

    good = fixture_device("thin_l1_light")
    G = good.G.copy()
    G[200] = np.nan
    bad = dataclasses.replace(good, G=G)
    J, ok, its, _ = emu.solve_packed(pack([bad, good]), V_FULL[::6].copy())
    assert not ok[0].any() and ok[1].all()
    assert its[0].max() <= 1


def test_replay_noise_floor_stop_is_the_oracles(emu, oracle):
    """solve_jv's noise-floor stop (src/jaxdd/solver.py:1728-1751, 2162-2166): at tol 1e-14
    every solve ends at the round-off floor or below its own noise floor, and the replay takes
    the oracle's steps exactly (without the noise floor it takes more on every device)."""
    from jaxdd_amd import synthetic as S
    from jaxdd_amd.batch import pack

    V = S.VOLTAGES[::3].copy()
    devs = [fixture_device(c) for c in ("thin_l1_light", "steady_1e17_thin_light")]
    devs += S.devices("c2", 3)
    for dev in devs:
        Jo, oko, itso = oracle.solve_jv(dev, V, tol=1e-14, max_iter=200)
        J, ok, its, status = emu.solve_packed(pack([dev]), V, tol=1e-14, max_iter=200,
                                              strict=True, jv=True)
        assert not status.any()
        assert np.array_equal(ok[0], oko) and np.array_equal(its[0], itso)
        shunt = 1.0 / dev.R_shunt if dev.R_shunt > 0 else 0.0
        assert np.max(np.abs(J[0] + shunt * V - Jo)) < 1e-10 * curve_scale(dev, Jo)


def _asm_state(devs, Vint, rng):
    """Interior states (n, NP, 3): V and the log densities linear between the contacts plus
    noise, and each device's full state with its contacts (the oracle's input)."""
    ys, fulls = [], []
    for dev, v in zip(devs, Vint):
        NP = len(dev.x) - 2
        lo = np.array([dev.V0 + 0.5 * dev.left_elec * v, np.log(dev.n_left), np.log(dev.p_left)])
        hi = np.array([dev.VL - 0.5 * dev.left_elec * v, np.log(dev.n_right), np.log(dev.p_right)])
        t = (dev.x / dev.x[-1])[:, None]
        full = lo + (hi - lo) * t
        full[1:-1] += rng.normal(scale=[0.02, 0.5, 0.5], size=(NP, 3))
        ys.append(full[1:-1])
        fulls.append(full)
    return np.array(ys), fulls


@pytest.mark.parametrize("case", ["fixtures", "c3"])
def test_replay_assemble_matches_oracle(emu, oracle, case):
    """The assembly operator (ddhip_assemble.h) against the oracle's `_assemble` restatement
    and its current at the same states: residual, blocks, ion columns and J."""
    from jaxdd_amd import synthetic as S
    from jaxdd_amd.batch import pack

    rng = np.random.default_rng(7)
    if case == "fixtures":
        devs = [fixture_device(c) for c in ("thin_l1_light", "frozen_1e17_x1_thin_light")]
        logs = None
        bern = "simss"
    else:
        devs = S.devices("c3", 3)
        logs = np.array([[-30.0, -31.0], [-29.0, -32.5], [-33.0, -28.0]])
        bern = "exact"
    Vint = np.array([0.3, 0.9, -0.1][: len(devs)])
    y, fulls = _asm_state(devs, Vint, rng)
    F, A, B, C, U, J = emu.assemble_packed(pack(devs), y, Vint, logs, bern)
    for k, dev in enumerate(devs):
        la, lc = (0.0, 0.0) if logs is None else logs[k]
        Fo, Ao, Bo, Co, Uo = oracle.assemble(dev, fulls[k], la=la, lc=lc, bernoulli_name=bern)
        # entries relative to their row's scale (the rows are scaled residuals)
        rs = np.max(np.abs(np.concatenate([Ao, Bo, Co], axis=2)), axis=2) + 1e-300
        assert np.all(np.abs(F[k] - Fo) <= 1e-11 * (rs + np.abs(Fo))), case
        for got, ref in ((A[k], Ao), (B[k], Bo), (C[k], Co)):
            assert np.all(np.abs(got - ref) <= 1e-12 * rs[:, :, None]), case
        assert np.all(np.abs(U[k] - Uo) <= 1e-12 * (np.abs(Uo) + rs[:, :, None]))
        Jo = oracle.current(dev, fulls[k])  # the ions do not enter the current
        assert abs(J[k] - Jo) <= 1e-10 * (abs(Jo) + oracle.j_scale(dev)), (J[k], Jo)


def _asm5_state(devs, Vint, logs, rng):
    """5x5 interior states (n, NP, 5): `_asm_state`'s carriers and each ion's log density off
    its Boltzmann profile of the slot's normalisation in its region (0 outside), and each
    device's full state (NP + 2 x 5, the oracle's input)."""
    y3, fulls3 = _asm_state(devs, Vint, rng)
    ys, fulls = [], []
    for dev, full, (la, lc) in zip(devs, fulls3, logs):
        V = full[:, 0]
        am, cm = np.asarray(dev.anion_mask) > 0, np.asarray(dev.cation_mask) > 0
        lna = np.where(am, la + V / dev.Vt + rng.normal(0, 0.3, V.shape), 0.0)
        lnc = np.where(cm, lc - V / dev.Vt + rng.normal(0, 0.3, V.shape), 0.0)
        f5 = np.column_stack([full, lna, lnc])
        ys.append(f5[1:-1])
        fulls.append(f5)
    return np.array(ys), fulls


def check_assemble_ions(out, devs, fulls, logs, oracle, bern):
    """The 5x5 operator's outputs (F, A, B, C, U as arrays (n, ...)) against the oracle's
    `_assemble_ions` restatement at the same states, relative to each row's scale."""
    F, A, B, C, U = out
    for k, dev in enumerate(devs):
        Fo, Ao, Bo, Co, Uo = oracle.assemble_ions(dev, fulls[k], logs[k], bernoulli_name=bern)
        rs = np.max(np.abs(np.concatenate([Ao, Bo, Co], axis=2)), axis=2) + 1e-300
        assert np.all(np.abs(F[k] - Fo) <= 1e-11 * (rs + np.abs(Fo)))
        for got, ref in ((A[k], Ao), (B[k], Bo), (C[k], Co)):
            assert np.all(np.abs(got - ref) <= 1e-12 * rs[:, :, None])
        assert np.array_equal(U[k], Uo)


@pytest.mark.parametrize("bern", ["exact", "simss"])
def test_replay_assemble_ions_matches_oracle(emu, oracle, bern):
    """The 5x5 assembly operator (ions as unknowns, steady) against the oracle's
    `_assemble_ions` restatement: residual, blocks and the slots' columns."""
    from jaxdd_amd import synthetic as S
    from jaxdd_amd.batch import pack

    rng = np.random.default_rng(9)
    devs = S.devices("c3", 3) + [fixture_device("steady_1e17_thin_light")]
    devs = devs if bern == "exact" else devs[3:]
    logs = np.array([[-30.0, -31.0], [-29.0, -32.5], [-33.0, -28.0], [40.0, 41.0]])[-len(devs):]
    Vint = np.array([0.3, 0.9, -0.1, 0.5])[-len(devs):]
    y, fulls = _asm5_state(devs, Vint, logs, rng)
    out = emu.assemble_ions_packed(pack(devs), y, Vint, logs, bern)
    check_assemble_ions(out, devs, fulls, logs, oracle, bern)


@pytest.mark.parametrize("lanes", [1, 2])
@pytest.mark.parametrize("ions", [False, True])
def test_replay_odd_grid_matches_oracle(emu, oracle, ions, lanes):
    """An odd number of interior points (NP = 301): the two-lane split's halves differ by a
    row; the oracle's flags and iterations, and its curves to the synthetic tests' bar (each
    solve stops at a step below 1e-10, so the last bits of two FP64 orders differ there)."""
    from jaxdd_amd import synthetic as S
    from jaxdd_amd.batch import pack

    t = S.draw(4)
    devs = [S.device_from_row(t, k, 301, ions, False) for k in range(4)]
    V = S.VOLTAGES[::4].copy()
    Jo, oko, itso = oracle.solve_batch(devs, V, first_max_iter=None)
    J, ok, its, status = emu.solve_packed(pack(devs), V, precision="fp64", lanes=lanes)
    assert not status.any()
    agree, chaotic = flags_agree(ok, its, oko, itso)
    assert agree and chaotic == 0
    for k, dev in enumerate(devs):
        both = ok[k] & oko[k]
        assert np.max(np.abs(J[k] - Jo[k])[both], initial=0.0) < 1e-9 * curve_scale(dev, Jo[k])


# ---- four lanes per device (ddhip_quad.h): devices without mobile ions -------------------
def _ion_free(packed):
    from jaxdd_amd.batch import SC_FIELDS

    sc = packed["sc"]
    ions = (sc[SC_FIELDS.index("a_total")] > 0) | (sc[SC_FIELDS.index("c_total")] > 0)
    return np.flatnonzero(~ions)


@pytest.mark.parametrize("precision,jtol", [("fp64", 1e-12), ("mixed", 1e-9)])
def test_replay_four_lanes_match_oracle_on_fixtures(emu, fixture_runs, precision, jtol):
    """The grid in quarters, the unknowns at the middle kept as extra columns: the oracle's
    flags, in FP64 its iterations, and its curves; devices with mobile ions refused (status 2,
    no converged voltage, NaN currents)."""
    cases, devs, packed, (Jo, oko, itso) = fixture_runs
    J, ok, its, status = emu.solve_packed(packed, V_FULL, bernoulli="simss", precision=precision,
                                          lanes=4)
    free = _ion_free(packed)
    assert 0 < len(free) < len(devs)
    refused = np.setdiff1d(np.arange(len(devs)), free)
    assert np.all(status[refused] == 2) and not status[free].any()
    assert not ok[refused].any() and np.isnan(J[refused]).all() and not its[refused].any()
    assert np.array_equal(ok[free], oko[free])
    if precision == "fp64":
        assert np.array_equal(its[free], itso[free])
    for k in free:
        assert np.max(np.abs(J[k] - Jo[k])) < jtol * curve_scale(devs[k], Jo[k]), cases[k]


@pytest.mark.parametrize("np_", [301, 16, 19])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_replay_four_lanes_match_oracle_on_synthetic(emu, oracle, np_, precision):
    """The gradient workload's devices (no ions) on an odd grid (NP = 301: quarters of 75, 75,
    75, 76 rows) and on the smallest grids the split takes (16: four rows a lane; 19): the
    oracle's flags and iterations, and its curves."""
    from jaxdd_amd import synthetic as S
    from jaxdd_amd.batch import pack

    t = S.draw(12)
    devs = [S.device_from_row(t, k, np_, False, False) for k in range(12)]
    V = S.VOLTAGES[::2].copy()
    Jo, oko, itso = oracle.solve_batch(devs, V, first_max_iter=None)
    J, ok, its, status = emu.solve_packed(pack(devs), V, precision=precision, lanes=4)
    assert not status.any()
    wild = chaotic_devices(oracle, devs, V, Jo)
    assert wild.mean() <= 0.1
    agree, chaotic = flags_agree(ok[~wild], its[~wild], oko[~wild], itso[~wild])
    assert agree and chaotic <= 0.02 * ok[~wild].size
    for k, dev in enumerate(devs):
        if wild[k]:
            continue
        both = ok[k] & oko[k]
        assert np.max(np.abs(J[k] - Jo[k])[both], initial=0.0) < 1e-9 * curve_scale(dev, Jo[k])


@pytest.mark.parametrize("lanes", [2, 4])
def test_replay_kept_states_are_the_one_lane_states(emu, lanes):
    """ddh_solve_batch_states with a device's state copy split over its lanes: every voltage's
    kept state (NP = 301: shares of 903 values, not a multiple of the copy's batches of eight)
    is the one-lane solve's state to the step tolerance, and the curves agree."""
    from jaxdd_amd import synthetic as S
    from jaxdd_amd.batch import pack

    t = S.draw(5)
    devs = [S.device_from_row(t, k, 301, False, False) for k in range(5)]
    packed = pack(devs)
    V = S.VOLTAGES[::5].copy()
    shape = (len(V), 301, 3, len(devs))
    ref, got = np.full(shape, np.nan), np.full(shape, np.nan)
    J1, ok1, _, _ = emu.solve_packed(packed, V, tol=1e-12, strict=True, max_iter=200, states=ref)
    J, ok, _, _ = emu.solve_packed(packed, V, tol=1e-12, strict=True, max_iter=200, states=got,
                                   lanes=lanes)
    assert ok1.all() and ok.all()
    assert not np.isnan(got).any()
    # the two eliminations round differently: the states agree to the round-off floor of the
    # log-densities (~1e-11 of |ln p| ~ 50 deep in a doped layer)
    assert np.all(np.abs(got - ref) <= 1e-9 + 5e-11 * np.abs(ref))
    for k, dev in enumerate(devs):
        assert np.max(np.abs(J[k] - J1[k])) < 1e-9 * curve_scale(dev, J1[k])


@pytest.mark.parametrize("n", [1, 3, 7, 8, 9, 37, 64, 1023, 1024, 1025, 4096, 5001])
def test_xcd_map_gives_every_device_to_one_workgroup(emu, n):
    """The impedance tables' launch (as many workgroups as devices, rounded up to whole XCD
    rounds) and the adjoint's (at most 1024 workgroups, each looping over its devices): every
    device to exactly one (workgroup, turn), none outside the batch."""
    for nblocks in sorted({((n + 7) // 8) * 8, min(n, 1024), 1, 5, 8, 13, 1024}):
        hits, outside = emu.xcd_hits(n, nblocks)
        assert outside == 0 and np.all(hits == 1), (n, nblocks)


@pytest.mark.parametrize("lanes", [1, 2])
@pytest.mark.parametrize("config,variant", [("c2", "free"), ("c3", "free"), ("c3", "frozen")])
def test_replay_takes_the_oracles_newton_steps(emu, oracle, config, variant, lanes):
    """The kernels' per-lane code in FP64 on 16 devices of each synthetic workload: the
    oracle's Newton step count on every short solve (conftest.short_points), its flags, and
    its curves to 1e-9 of each curve's scale."""
    from conftest import short_points

    from jaxdd_amd import synthetic as S
    from jaxdd_amd.batch import pack

    devs = S.devices(config, 16, variant=variant)
    V = S.VOLTAGES[::2].copy()
    Jo, oko, itso = oracle.solve_batch(devs, V, first_max_iter=None)
    J, ok, its, _ = emu.solve_packed(pack(devs), V, lanes=lanes)
    wild = chaotic_devices(oracle, devs, V, Jo)
    short = short_points(devs, V, its, itso, wild)
    assert short.sum() > 0.8 * its.size * (1 - wild.mean())
    assert np.array_equal(its[short], itso[short])
    agree, _ = flags_agree(ok[~wild], its[~wild], oko[~wild], itso[~wild])
    assert agree
    for k, dev in enumerate(devs):
        if not wild[k]:
            both = ok[k] & oko[k]
            assert np.max(np.abs(J[k] - Jo[k])[both], initial=0.0) < 1e-9 * curve_scale(dev, Jo[k])
