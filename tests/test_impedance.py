"""The impedance at a steady state (reference: src/jaxdd/impedance.py; its tests
tests/test_impedance.py): the oracle pinned by the limits the reference's own tests hold it to
(the steady J-V curve's slope at low frequency, the geometric capacitance at high frequency),
by the steady current it must reduce to and by the exact solution of the traps' step; the
kernels' CPU replay and the GPU against the oracle."""

import numpy as np
import pytest

from conftest import fixture_device

V0 = 0.6
FREQS = np.logspace(-3, 14, 12)
#: SimSS fixture devices without mobile ions (three layers, two interfaces, bulk traps)
NO_ION_FIXTURES = ("thin_l1_light", "none_light", "held_thin_light", "resist_light")


def _synthetic(n, NP, seed_rows=None):
    from jaxdd_amd import synthetic as S

    t = S.draw(max(n, 4))
    return [S.device_from_row(t, k, NP, False, False) for k in range(n)]


def _steady(oracle, dev, V, tol=1e-12):
    J, ok, its, st = oracle.solve_one(dev, [V], tol=tol, max_iter=200, return_states=True)
    assert ok.all()
    return st[0]


def _close_to_oracle(Y, dev, y, V, laps, dt, bern="exact"):
    """Y (one device's, per lap) against the oracle at the same state; the bar is 20 times the
    spread of two eliminations that are equal in exact arithmetic (the reference's adjugate
    inverses and pivoted solves), relative to each |Y|: what the conditioning of the device's
    systems lets rounding move (1e-6 of |Y| for the worst-conditioned devices at low
    frequency, 1e-10 for most)."""
    from oracle import ddimpedance as I

    full = I.full_state(dev, y, V)
    Yo = I.admittance(dev, full, laps, dt=dt, bernoulli_name=bern)
    Yp = I.admittance(dev, full, laps, dt=dt, bernoulli_name=bern, pivoted=True)
    spread = float(np.max(np.abs(Yo - Yp) / np.abs(Yo)))
    err = np.abs(np.asarray(Y) - Yo) / np.abs(Yo)
    assert np.max(err) <= 20 * spread + 1e-10, (np.max(err), spread)
    return float(np.max(err))


# ---- the oracle ---------------------------------------------------------------------------
def test_transient_current_at_rest_is_the_steady_current(oracle):
    """The restated transient current with no time step is the C oracle's steady current (its
    own restatement of batch.py:321-389, pinned by the SimSS fixtures), with the steady
    widths and with the control volumes (equal at a steady state, where both carriers' rates
    are the same)."""
    from oracle import ddimpedance as I

    for dev in _synthetic(2, 120) + [fixture_device("none_light")]:
        full = I.full_state(dev, _steady(oracle, dev, V0), V0)
        before = (np.exp(full[:, 1]), np.exp(full[:, 2]), full[:, 0])
        NP = len(dev.x) - 2
        Jc = oracle.current(dev, full, dev.G_frac)
        Js = I.transient_current(dev, full, dev.G_frac, 0.0, before, width=np.asarray(dev.h)[: NP + 1])
        Jv = I.transient_current(dev, full, dev.G_frac, 0.0, before)
        scale = abs(Jc) + 1e-3
        assert abs(Js - Jc) < 1e-12 * scale and abs(Jv - Jc) < 1e-12 * scale


def test_low_frequency_limit_is_the_steady_slope(oracle):
    """The reference's tests/test_impedance.py:21-33: at 1e-7 Hz the admittance is the steady
    current's slope in the voltage (the oracle's solves at +-10 uV), with bulk traps, to 1e-5,
    its imaginary part below 1e-4 of it."""
    from oracle import ddimpedance as I

    for dev in _synthetic(2, 150) + [fixture_device("thin_l1_light")]:
        dV = 1e-5
        J, ok, _ = oracle.solve_one(dev, [V0 - dV, V0 + dV], tol=1e-12, max_iter=200)
        assert ok.all()
        slope = (J[1] - J[0]) / (2 * dV)
        Y = I.admittance(dev, I.full_state(dev, _steady(oracle, dev, V0), V0), I.laps_of([1e-7]))[0]
        assert Y.real == pytest.approx(slope, rel=1e-5)
        assert abs(Y.imag) < 1e-4 * abs(slope)


def test_high_frequency_limit_is_the_geometric_capacitance(oracle):
    """The reference's tests/test_impedance.py:36-47: far above every relaxation (1e14 Hz) the
    device is its layers' dielectrics in series, 1/integral dx/eps, to 1e-5."""
    from oracle import ddimpedance as I

    for dev in _synthetic(2, 150) + [fixture_device("none_light")]:
        full = I.full_state(dev, _steady(oracle, dev, 0.0), 0.0)
        Y = I.admittance(dev, full, I.laps_of([1e14]))[0]
        C = Y.imag / (2 * np.pi * 1e14)
        assert C == pytest.approx(I.geometric_capacitance(dev), rel=1e-5)


def test_step_transfer_is_the_exact_step_of_the_occupation():
    """A trap level's occupation relaxes to its steady value f_inf at its rate a. Over implicit
    Euler's steps of dt with f_inf held at each step's end (z^k for a sine), the exact step is
    f_k = f_inf,k + (f_(k-1) - f_inf,k) e^(-a dt), with mean f_inf,k + (f_(k-1) - f_inf,k) m
    over the step: after the start has decayed, (f - f_inf)/f_inf at the step's end and of the
    mean are step_transfer's charge and rate parts; dt -> 0 gives the continuous -iw/(iw + a)."""
    from oracle import ddimpedance as I

    for a, f, dt in ((3e5, 1e4, 2.5e-6), (1e3, 1e5, 1e-7), (2e7, 1e6, 5e-8)):
        w = 2 * np.pi * f
        z = np.exp(1j * w * dt)
        E = np.exp(-a * dt)
        m = -np.expm1(-a * dt) / (a * dt)
        fk, last_end, last_mean = 0.0, None, None
        steps = int(60 / (a * dt)) + 200
        for k in range(1, steps + 1):
            finf = z ** k
            mean = finf + (fk - finf) * m
            fk = finf + (fk - finf) * E
            last_end, last_mean = fk / finf - 1, mean / finf - 1
        lap = I.laps_of([f], dt)[0]
        h_rate, h_charge = I.step_transfer(a, lap, dt)
        assert abs(h_charge - last_end) < 1e-9 * max(1.0, abs(last_end))
        assert abs(h_rate - last_mean) < 1e-9 * max(1.0, abs(last_mean))
    h_cont, _ = I.step_transfer(3e5, 1j * 2 * np.pi * 1e4)
    h_small, _ = I.step_transfer(3e5, I.laps_of([1e4], 1e-13)[0], 1e-13)
    assert abs(h_small - h_cont) < 1e-6 * abs(h_cont)


def test_implicit_euler_steps_tend_to_the_continuous_response(oracle):
    """Implicit Euler's admittance differs from the continuous one by its first-order error:
    ten times smaller steps, ten times closer (the reference quotes 4-8 % at 40 steps a
    period)."""
    from oracle import ddimpedance as I

    dev = _synthetic(1, 150)[0]
    full = I.full_state(dev, _steady(oracle, dev, V0), V0)
    f = 1e5
    Yc = I.admittance(dev, full, I.laps_of([f]))[0]
    gaps = []
    for dt in (1 / (f * 40), 1 / (f * 400), 1 / (f * 4000)):
        gaps.append(abs(I.admittance(dev, full, I.laps_of([f], dt), dt=dt)[0] / Yc - 1))
    assert 1e-3 < gaps[0] < 0.2
    assert 7 < gaps[0] / gaps[1] < 13 and 7 < gaps[1] / gaps[2] < 13


def test_external_impedance_of_the_series_resistance_and_the_shunt():
    """The reference's tests/test_impedance.py:100-109: Z = 1/(Y + G_shunt) + R_series."""
    from jaxdd_amd.impedance import external_impedance

    Y = np.array([[1.0 + 2.0j, 3.0 - 1.0j], [0.5 + 0.1j, 2.0 + 7.0j]])
    Z = external_impedance(Y, [10.0, 0.0], [1e-4, 2e-4])
    assert np.allclose(Z, 1.0 / (Y + np.array([[10.0], [0.0]])) + np.array([[1e-4], [2e-4]]),
                       rtol=1e-15)


# ---- the oracle's branch with mobile ions (impedance.py:242-262) ----------------------------
def _ion_devices(n, NP):
    from jaxdd_amd import synthetic as S

    t = S.draw(max(n, 4) + 1)
    return [S.device_from_row(t, k, NP, True, False) for k in range(1, n + 1)]


def _ion_state(oracle, dev, V):
    """The steady state at V with the ions' logarithms (`ddimpedance.ion_state`), the ions'
    densities on the grid from the oracle's `solve_states`."""
    from oracle import ddimpedance as I

    st = oracle.solve_states(dev, [V], tol=1e-12, max_iter=200)
    assert st["converged"].all()
    full = np.stack([st["V"][0], np.log(st["n"][0]), np.log(st["p"][0])], -1)
    return I.ion_state(dev, full, st["anion"][0], st["cation"][0])


def test_ion_branch_of_a_device_without_ions_is_the_3x3_branch(oracle):
    """A device without mobile ions through the 5 x 5 branch (its ion rows held, its slots
    identities, as in a batch where other devices have ions) gives the 3 x 3 branch's admittance
    to 20 times the spread of the 3 x 3 branch's two eliminations (adjugate and pivoted): what
    the low frequencies' conditioning lets rounding move (up to 1e-6 of Y on a dark device)."""
    from oracle import ddimpedance as I

    for dev in _synthetic(2, 100):
        full = I.full_state(dev, _steady(oracle, dev, V0), V0)
        full5 = np.concatenate([full, np.zeros((full.shape[0], 2))], 1)
        for dt in (None, 1e-6):
            laps = I.laps_of(np.logspace(-3, 12, 6), dt)
            Y3 = I.admittance(dev, full, laps, dt=dt)
            Yp = I.admittance(dev, full, laps, dt=dt, pivoted=True)
            Y5 = I.admittance_ions(dev, full5, laps, dt=dt)
            spread = np.abs(Yp / Y3 - 1)
            assert np.all(np.abs(Y5 / Y3 - 1) <= 20 * spread + 1e-10)


def test_ion_masses_are_the_slots_columns(oracle):
    """Each ion row's time term per unit of dti (`ion_masses`, restated from the row's form) is
    minus its slot's column (the C oracle's U: the row's derivative in the slot's log, the
    density before the step scaled by exp(log)), at the steady state and dti = 1; 0 outside the
    region; the carriers' the 3 x 3 branch's."""
    from oracle import ddimpedance as I

    for dev in _ion_devices(2, 90):
        f5 = _ion_state(oracle, dev, V0)
        _, _, _, _, U, _, _ = oracle.assemble_step(dev, f5, dev.G_frac, 1.0, f5, None)
        M = I.ion_masses(dev, f5)
        for k in (0, 1):
            col = M[:, 3 + k]
            assert np.all(np.abs(col + U[:, 3 + k, k]) <= 1e-14 * np.max(np.abs(col)))
            assert np.count_nonzero(col) == np.count_nonzero(np.asarray(
                (dev.anion_mask, dev.cation_mask)[k])[1:-1])


def test_ion_low_frequency_limit_is_the_steady_slope(oracle):
    """The reference's tests/test_impedance.py:20-33 with ions: at 1e-9 Hz the ions follow the
    voltage to their steady state, and the admittance is the slope of the steady current with
    free ions (the oracle's solves at +-10 uV) to 1e-5, its imaginary part below 1e-4 of it."""
    from oracle import ddimpedance as I

    for dev in _ion_devices(2, 120):
        dV = 1e-5
        J, ok, _ = oracle.solve_one(dev, [V0 - dV, V0 + dV], tol=1e-12, max_iter=200)
        assert ok.all()
        slope = (J[1] - J[0]) / (2 * dV)
        Y = I.admittance_ions(dev, _ion_state(oracle, dev, V0), I.laps_of([1e-9]))[0]
        assert Y.real == pytest.approx(slope, rel=1e-5)
        assert abs(Y.imag) < 1e-4 * abs(slope)


def test_ion_high_frequency_limit_is_the_geometric_capacitance(oracle):
    """With ions too: at 1e14 Hz nothing but the dielectrics responds (1e-5)."""
    from oracle import ddimpedance as I

    for dev in _ion_devices(2, 120):
        Y = I.admittance_ions(dev, _ion_state(oracle, dev, 0.0), I.laps_of([1e14]))[0]
        assert Y.imag / (2 * np.pi * 1e14) == pytest.approx(I.geometric_capacitance(dev), rel=1e-5)


def test_ion_branch_is_one_implicit_euler_step_of_the_transient(oracle):
    """The mid-band with ions: one implicit-Euler step of dt from the steady state with a
    voltage step dV changes the transient oracle's current (its ions as unknowns, each species'
    number kept by its slot; pinned to ZimT's runs with ions) by Y(lap = 1/dt) dV, central in
    +-dV, to 1e-6, from 1 ns to 1 s."""
    from oracle import ddimpedance as I

    for dev in _ion_devices(2, 120):
        f5 = _ion_state(oracle, dev, V0)
        dV = 1e-4
        for dt in (1e-9, 1e-6, 1e-3, 1.0):
            Y = I.admittance_ions(dev, f5, np.array([1.0 / dt + 0j]), dt=dt)[0]
            J = []
            for s in (1, -1):
                r = oracle.transient(dev, [0.0, dt], [V0, V0 + s * dV], [1.0, 1.0], tol=1e-14)
                assert r["converged"].all()
                J.append(r["J"][1])
            assert abs((J[0] - J[1]) / (2 * dV) / Y.real - 1) < 1e-6, dt


def test_ion_branch_is_the_transients_response_to_a_sine(oracle):
    """The reference's tests/test_impedance.py:50-84 with ions: through implicit Euler at steps
    of dt (40 a period, 24 periods: the ions' slow settling leaves 1.4e-6 after 12) the
    transient oracle's response to a 0.1 mV sine at 0.6 V, fitted over the last 4 periods, is
    the admittance with dt to 1e-7; the continuous response differs from it by implicit
    Euler's error (> 1e-2)."""
    from oracle import ddimpedance as I

    dev = _ion_devices(1, 80)[0]
    f, amp, per, periods, fitted = 1e2, 1e-4, 40, 24, 4
    dt = 1.0 / (f * per)
    k = np.arange(1, per * periods + 1)
    t = np.concatenate([[0.0], k * dt])
    V = np.concatenate([[V0], V0 + amp * np.sin(2 * np.pi * f * k * dt)])
    r = oracle.transient(dev, t, V, np.ones(len(t)), tol=1e-13, max_iter=60)
    assert r["converged"].all()
    m = k > per * (periods - fitted)
    ph = 2 * np.pi * f * k[m] * dt
    X = np.stack([np.ones(m.sum()), np.sin(ph), np.cos(ph)], -1)
    c = np.linalg.lstsq(X, r["J"][1:][m], rcond=None)[0]
    Y_t = (c[1] + 1j * c[2]) / amp
    f5 = _ion_state(oracle, dev, V0)
    discrete = I.admittance_ions(dev, f5, I.laps_of([f], dt), dt=dt)[0]
    continuous = I.admittance_ions(dev, f5, I.laps_of([f]))[0]
    assert abs(discrete / Y_t - 1) < 1e-7
    assert abs(continuous / Y_t - 1) > 1e-2


# ---- the kernels' CPU replay ----------------------------------------------------------------
def _cases():
    """(devices, voltage, Bernoulli): synthetic devices lit at three voltages and dark at one,
    the SimSS fixture devices without ions with SimSS's Bernoulli (their interfaces move the
    current's weights)."""
    import dataclasses

    lit = _synthetic(4, 150)
    dark = [dataclasses.replace(d, G_frac=0.0) for d in lit]
    fix = [fixture_device(c) for c in NO_ION_FIXTURES]
    small = _synthetic(3, 20)  # fewer points than lanes: most of the table's segments empty
    return ([(lit, V, "exact") for V in (0.0, V0, 1.0)] + [(dark, V0, "exact")]
            + [(fix, V0, "simss"), (small, V0, "exact")])


def _steady_b(oracle, dev, V, bern):
    J, ok, its, st = oracle.solve_one(dev, [V], tol=1e-12, max_iter=200, bernoulli=bern,
                                      return_states=True)
    assert ok.all()
    return st[0]


@pytest.mark.parametrize("dt", [None, 1e-7])
def test_replay_matches_the_oracle(oracle, emu, dt):
    """The impedance kernels' code (the per-device table, the per-frequency elimination with the
    current contracted on the way) on the CPU against the oracle, which solves by the
    reference's block elimination and back-substitution and takes the current's derivatives by
    the complex step."""
    from jaxdd_amd.batch import pack

    from oracle import ddimpedance as I

    laps = I.laps_of(FREQS, dt)
    for devs, V, bern in _cases():
        ys = np.stack([_steady_b(oracle, d, V, bern) for d in devs])
        Y, status = emu.impedance_packed(pack(devs), ys, np.full(len(devs), V), laps, dt, bern)
        assert not status.any()
        for k, dev in enumerate(devs):
            _close_to_oracle(Y[k], dev, ys[k], V, laps, dt, bern)


def _close_to_oracle_ions(Y, dev, full5, laps, dt, bern="exact"):
    """`_close_to_oracle` for the 5 x 5 path: Y against the oracle's ion branch at the same
    state, to 20 times the spread of its two eliminations (adjugate and pivoted)."""
    from oracle import ddimpedance as I

    Yo = I.admittance_ions(dev, full5, laps, dt=dt, bernoulli_name=bern)
    Yp = I.admittance_ions(dev, full5, laps, dt=dt, bernoulli_name=bern, pivoted=True)
    spread = float(np.max(np.abs(Yo - Yp) / np.abs(Yo)))
    err = np.abs(np.asarray(Y) - Yo) / np.abs(Yo)
    assert np.max(err) <= 20 * spread + 1e-10, (np.max(err), spread)
    return float(np.max(err))


def _ion_cases():
    """(devices, voltage, Bernoulli) for the 5 x 5 path: synthetic devices with ions in their
    absorber at three voltages (one batch also holding a device without ions, which the 5 x 5
    path must give the 3 x 3 result), and ZimT's device with ions (SimSS's Bernoulli)."""
    from test_transient import zimt_case

    ion = _ion_devices(3, 90)
    mixed = ion[:2] + _synthetic(1, 90)
    zimt = [zimt_case("zimt_ions_fast_light")[0]]
    return ([(ion, V, "exact") for V in (0.0, V0, 1.0)] + [(mixed, V0, "exact")]
            + [(zimt, 0.5, "simss")])


def _ion_full_state(oracle, dev, V, bern):
    """The oracle's steady state at V as the 5 x 5 path's full state (`ion_state`; zero ion
    columns for a device without ions)."""
    from oracle import ddimpedance as I

    st = oracle.solve_states(dev, [V], tol=1e-12, max_iter=200, bernoulli_name=bern)
    assert st["converged"].all()
    full = np.stack([st["V"][0], np.log(st["n"][0]), np.log(st["p"][0])], -1)
    return I.ion_state(dev, full, st["anion"][0], st["cation"][0])


@pytest.mark.parametrize("dt", [None, 1e-7])
def test_replay_with_ions_matches_the_oracle(oracle, emu, dt):
    """The 5 x 5 path's code (the ions' rows, constraint rows and current coefficients in the
    table; the bordered elimination with three right-hand sides and three functionals) on the
    CPU against the oracle's ion branch at the same states."""
    from jaxdd_amd.batch import pack

    from oracle import ddimpedance as I

    laps = I.laps_of(FREQS, dt)
    for devs, V, bern in _ion_cases():
        fulls = [_ion_full_state(oracle, d, V, bern) for d in devs]
        ys = np.stack([f[1:-1, :3] for f in fulls])
        Y, status = emu.impedance_packed(pack(devs), ys, np.full(len(devs), V), laps, dt, bern,
                                         ions=True)
        assert not status.any()
        for k, dev in enumerate(devs):
            _close_to_oracle_ions(Y[k], dev, fulls[k], laps, dt, bern)


def test_replay_ion_current_coefficients_are_the_complex_step_derivatives(oracle, emu):
    """The table's coefficients of the ions' logarithms in the linearised current against the
    complex-step derivatives of the oracle's transient current with ions, point by point: at
    dti = 0 (only the region's two ends move K) and per unit of lap (the densities' change
    through the double integral and the direct sum)."""
    from jaxdd_amd.batch import pack
    from oracle import ddimpedance as I

    dev = _ion_devices(1, 70)[0]
    f5 = _ion_full_state(oracle, dev, V0, "exact")
    _, _, tab, _ = emu.impedance_packed(pack([dev]), f5[None, 1:-1, :3], np.full(1, V0),
                                        I.laps_of([1e3]), table=True, ions=True)
    tab = tab[0]
    regions = I.ion_regions(dev)
    before = (np.exp(f5[:, 1]), np.exp(f5[:, 2]), f5[:, 0].copy())
    ions = (regions, tuple(np.exp(f5[:, 3 + k]) * r["mask"] for k, r in enumerate(regions)))
    NP = len(dev.x) - 2

    def dJ(i, c, dti):
        v = np.zeros_like(f5)
        v[i, c] = 1.0
        return I._cstep(lambda s: I.transient_current(dev, f5 + s * v, dev.G_frac, dti, before,
                                                      ions=ions), 1.0)

    F = emu.IMP_FIELDS
    for q, (c0, cl) in enumerate(((F["CA0"], F["CAL"]), (F["CC0"], F["CCL"]))):
        d0 = np.array([dJ(i + 1, 3 + q, 0.0) for i in range(NP)])
        d1 = np.array([dJ(i + 1, 3 + q, 1.0) for i in range(NP)])
        scale = np.max(np.abs(d1 - d0))
        assert scale > 0
        assert np.max(np.abs(tab[:, cl] - (d1 - d0))) < 1e-9 * scale
        assert np.max(np.abs(tab[:, c0] - d0)) < 1e-9 * scale


def test_replay_refuses_an_ion_region_at_a_contact(oracle, emu):
    """A device whose anions reach a contact is outside the 5 x 5 path: status 2, Y NaN; the
    other devices of the batch are solved."""
    import dataclasses

    from jaxdd_amd.batch import pack
    from oracle import ddimpedance as I

    devs = _ion_devices(2, 60)
    m = np.asarray(devs[1].anion_mask, float).copy()
    m[: int(np.argmax(m > 0)) + 1] = 1.0  # the region widened to the left contact
    devs[1] = dataclasses.replace(devs[1], anion_mask=m)
    with pytest.raises(ValueError, match="contact"):
        I.ion_regions(devs[1])
    f0 = _ion_full_state(oracle, devs[0], V0, "exact")
    ys = np.stack([f0[1:-1, :3], f0[1:-1, :3]])
    Y, status = emu.impedance_packed(pack(devs), ys, np.full(2, V0), I.laps_of([1e3]), ions=True)
    assert list(status) == [0, 2]
    assert np.isfinite(Y[0]).all() and np.isnan(Y[1]).all()


def test_replay_current_coefficients_are_the_complex_step_derivatives(oracle, emu):
    """The per-device table's coefficients of the linearised current against the complex-step
    derivatives of the oracle's transient current, point by point, on a fixture device with
    two interfaces: each carrier's density-change coefficient (the lap part of dln n, dln p:
    where the interfaces move the weights, by 7 % and 2e-4 at this device's two), the steady
    rates' (dti = 0) parts of dln n, dln p, the displacement current's dV and the contacts'
    direct terms. dJ/dV through the integrals alone is left to the admittance's own parity: it
    is a difference of nearly equal terms, 1e-13 of the others."""
    from jaxdd_amd.batch import pack
    from oracle import ddimpedance as I

    dev = fixture_device("none_light")
    y = _steady(oracle, dev, V0)
    _, _, tab, dsc = emu.impedance_packed(pack([dev]), y[None], np.full(1, V0),
                                          I.laps_of([1e3]), table=True)
    tab, dsc = tab[0], dsc[0]
    full = I.full_state(dev, y, V0)
    before = (np.exp(full[:, 1]), np.exp(full[:, 2]), full[:, 0].copy())
    NP = len(dev.x) - 2

    def dJ(v, dti):
        return I._cstep(lambda s: I.transient_current(dev, full + s * v, dev.G_frac, dti, before),
                        float(np.max(np.abs(v))))

    def unit(i, c):
        v = np.zeros_like(full)
        v[i, c] = 1.0
        return v

    d0 = np.array([[dJ(unit(i + 1, c), 0.0) for c in range(3)] for i in range(NP)])
    d1 = np.array([[dJ(unit(i + 1, c), 1.0) for c in range(3)] for i in range(NP)])
    lap_part = d1 - d0
    F = emu.IMP_FIELDS
    # GLN (dln n), GLP (dln p) per unit of lap; G0N, G0P at dti = 0
    for col, ref in ((F["GLN"], lap_part[:, 1]), (F["GLP"], lap_part[:, 2]),
                     (F["G0N"], d0[:, 1]), (F["G0P"], d0[:, 2])):
        err = np.abs(tab[:, col] - ref)
        assert np.all(err <= 1e-6 * np.abs(ref) + 1e-12 * np.max(np.abs(ref))), col
    # KD (dV per unit of lap): eps's jumps over L H, 0 inside a layer (the complex step's
    # rounding there: 1e-15 of eps/(L H))
    scale = np.max(dev.eps) / (dev.L * np.sum(np.asarray(dev.h)[: NP + 1]))
    assert np.all(np.abs(tab[:, F["KD"]] - lap_part[:, 0]) <= 1e-12 * scale)
    contacts = np.zeros_like(full)
    contacts[0, 0], contacts[-1, 0] = 0.5 * dev.left_elec, -0.5 * dev.left_elec
    jv0, jv1 = dJ(contacts, 0.0), dJ(contacts, 1.0)
    assert abs(dsc[6] - jv0) <= 1e-9 * abs(jv0)
    assert abs(dsc[7] - (jv1 - jv0)) <= 1e-9 * abs(jv1 - jv0)


# ---- the host -------------------------------------------------------------------------------
def test_what_the_host_refuses():
    """An unknown ion flux, a step that is not positive; with mobile ions, ions frozen at a
    pre-bias (as the reference) and the modified drift: before any launch."""
    from jaxdd_amd import synthetic as S
    from jaxdd_amd.batch import from_packed, pack
    from jaxdd_amd.impedance import solve_impedance

    t = S.draw(3)
    plain = from_packed(pack([S.device_from_row(t, k, 40, False, False) for k in range(3)]), "cpu")
    with pytest.raises(ValueError, match="ion_flux"):
        solve_impedance(plain, [1e3], ion_flux="drift")
    with pytest.raises(ValueError, match="dt"):
        solve_impedance(plain, [1e3], dt=0.0)
    frozen = from_packed(pack([S.device_from_row(t, k, 40, True, True) for k in range(3)]), "cpu")
    with pytest.raises(NotImplementedError, match="frozen"):
        solve_impedance(frozen, [1e3])
    ions = from_packed(pack([S.device_from_row(t, k, 40, True, False) for k in range(3)]), "cpu")
    with pytest.raises(NotImplementedError, match="modified drift"):
        solve_impedance(ions, [1e3], ion_flux="modified-drift")
    with pytest.raises(RuntimeError, match="GPU"):
        solve_impedance(ions, [1e3])


def test_the_c_abi_rejects_bad_arguments():
    """ddh_impedance checks its arguments on the host and launches nothing on an error."""
    import ctypes

    from jaxdd_amd import _lib

    lib = _lib.load()
    b = _lib.Batch(4, 100, 3, 0, 1, 1, 1, 1)
    ws = lib.ddh_impedance_workspace_bytes(4, 100)
    assert ws == (4 * 100 * _lib.IMP_TABLE_FIELDS + 4 * _lib.IMP_SCALARS) * 8
    args = dict(y=1, Vint=1, iface=1, laps=1, nf=3, dt=0.0, bern=0, Y=1, status=1, ws=1, nb=ws)

    def call(**kw):
        a = dict(args, **kw)
        return lib.ddh_impedance(ctypes.byref(b), a["y"], a["Vint"], a["iface"], a["laps"],
                                 a["nf"], a["dt"], a["bern"], a["Y"], a["status"], a["ws"],
                                 a["nb"], None)

    assert call(iface=None) == -4  # DDH_E_ARG
    assert call(nf=0) == -1 and b"nf" in lib.ddh_last_error()  # DDH_E_SHAPE
    assert call(dt=-1.0) == -4 and b"dt" in lib.ddh_last_error()
    assert call(dt=float("nan")) == -4
    assert call(bern=2) == -4
    assert call(nb=ws - 8) == -3 and b"workspace" in lib.ddh_last_error()  # DDH_E_WORKSPACE


# ---- the GPU --------------------------------------------------------------------------------
@pytest.fixture(scope="module")
def gpu_batches(torch_cuda):
    """(devices, batch, Bernoulli): synthetic devices lit and dark, the SimSS fixture devices
    without ions with SimSS's Bernoulli, devices of 20 points (fewer than a wavefront's
    lanes) and a batch of 37 devices (the tables' XCD map giving each XCD a run of five)."""
    import dataclasses

    from jaxdd_amd.batch import prepare

    syn = _synthetic(6, 150)
    dark = [dataclasses.replace(d, G_frac=0.0) for d in syn]
    fix = [fixture_device(c) for c in NO_ION_FIXTURES]
    small = _synthetic(3, 20)
    # more devices than XCDs: each XCD's workgroups take a run of several devices
    many = _synthetic(37, 60)
    return [(syn, prepare(syn), "exact"), (dark, prepare(dark), "exact"),
            (fix, prepare(fix), "simss"), (small, prepare(small), "exact"),
            (many, prepare(many), "exact")]


@pytest.mark.gpu
@pytest.mark.parametrize("dt", [None, 1e-7])
def test_gpu_impedance_matches_the_oracle(gpu_batches, dt):
    """ddh_impedance through the C-ABI against the oracle at the GPU's own steady states."""
    from jaxdd_amd.impedance import laps_of, solve_impedance

    for devs, batch, bern in gpu_batches:
        r, states = solve_impedance(batch, FREQS, voltage=V0, tol=1e-12, max_iter=200, dt=dt,
                                    bernoulli_name=bern, return_states=True)
        assert r.converged.all()
        ys = states.cpu().numpy()
        for k, dev in enumerate(devs):
            _close_to_oracle(r.Y[k], dev, ys[k], V0, laps_of(FREQS, dt), dt, bern)
        Z = 1.0 / (r.Y + batch["G_shunt"].cpu().numpy()[:, None]) + batch["R_series"].cpu().numpy()[:, None]
        assert np.allclose(r.Z, Z, rtol=1e-14)


@pytest.mark.gpu
def test_gpu_limits(gpu_batches):
    """On the GPU alone: at 1e-7 Hz the slope of the GPU's own steady J-V curve (+-10 uV, to
    1e-5), at 1e14 Hz the geometric capacitance (to 1e-5 on the grids of 150 and 400 points;
    the discrete displacement current and Poisson stencil reach the continuum's 1/int dx/eps
    only as the grid is refined)."""
    from jaxdd_amd.batch import solve_batch
    from jaxdd_amd.impedance import solve_impedance
    from oracle import ddimpedance as I

    for devs, batch, bern in gpu_batches:
        r = solve_impedance(batch, [1e-7, 1e14], voltage=V0, tol=1e-12, max_iter=200,
                            bernoulli_name=bern)
        dV = 1e-5
        J, ok, _ = solve_batch(batch, [V0 - dV, V0 + dV], tol=1e-12, max_iter=200,
                               bernoulli_name=bern)
        J = J.cpu().numpy()
        G_shunt = batch["G_shunt"].cpu().numpy()
        slope = (J[:, 1] - J[:, 0]) / (2 * dV) - G_shunt  # Y is the internal admittance
        assert np.allclose(r.Y[:, 0].real, slope, rtol=1e-5, atol=0)
        if batch["NP"] < 100:  # the continuum's limit: 1e-3 apart on 20 points, 1e-5 by 150
            continue
        geo = np.array([I.geometric_capacitance(d) for d in devs])
        assert np.allclose(r.capacitance[:, 1], geo, rtol=1e-5, atol=0)


@pytest.mark.gpu
def test_gpu_series_resistance_and_shunt_leave_the_admittance(gpu_batches):
    """The reference's tests/test_impedance.py:100-109 on the GPU: a series resistance and a
    shunt change Z = 1/(Y + G_shunt) + R_series, not Y."""
    from jaxdd_amd.impedance import solve_impedance

    devs, batch, _ = gpu_batches[0]
    plain = solve_impedance(batch, [1e3, 1e6], voltage=V0)
    b2 = dict(batch, R_series=batch["R_series"] + 1e-4, G_shunt=batch["G_shunt"] + 10.0)
    r = solve_impedance(b2, [1e3, 1e6], voltage=V0)
    np.testing.assert_allclose(r.Y, plain.Y, rtol=1e-12)
    np.testing.assert_allclose(
        r.Z, 1.0 / (plain.Y + b2["G_shunt"].cpu().numpy()[:, None])
        + b2["R_series"].cpu().numpy()[:, None], rtol=1e-12)


@pytest.fixture(scope="module")
def gpu_ion_batches(torch_cuda):
    """(devices, batch, Bernoulli, voltage) for the 5 x 5 path: synthetic devices with ions
    (one batch of 37, the tables' XCD map giving each XCD a run of several, with two devices
    without ions among them), ZimT's device with ions (SimSS's Bernoulli)."""
    from jaxdd_amd.batch import prepare
    from test_transient import zimt_case

    ion = _ion_devices(6, 120)
    many = _ion_devices(35, 60) + _synthetic(2, 60)
    zimt = [zimt_case("zimt_ions_fast_light")[0]]
    return [(ion, prepare(ion), "exact", V0), (many, prepare(many), "exact", V0),
            (zimt, prepare(zimt), "simss", 0.5)]


def _gpu_ion_state(oracle, dev, y, V):
    """The 5 x 5 path's full state at the GPU's own steady state y (NP x 3): the ions' closed-form
    profiles of its potential, as `solve_states` gives them."""
    from oracle import ddimpedance as I
    from oracle.ddoracle import _ions_on_grid

    full = I.full_state(dev, y, V)
    anion, cation = _ions_on_grid(dev, full[None, :, 0])
    return I.ion_state(dev, full, anion[0], cation[0])


@pytest.mark.gpu
@pytest.mark.parametrize("dt", [None, 1e-7])
def test_gpu_impedance_with_ions_matches_the_oracle(gpu_ion_batches, oracle, dt):
    """ddh_impedance_ions through the C-ABI against the oracle's ion branch at the GPU's own
    steady states (devices without ions in the batch: the 3 x 3 result)."""
    from jaxdd_amd.impedance import laps_of, solve_impedance

    for devs, batch, bern, V in gpu_ion_batches:
        print(f"impedance with ions: {len(devs)} devices, NP = {batch['NP']}, dt = {dt}",
              flush=True)
        r, states = solve_impedance(batch, FREQS, voltage=V, tol=1e-12, max_iter=200, dt=dt,
                                    bernoulli_name=bern, return_states=True)
        assert r.converged.all()
        ys = states.cpu().numpy()
        for k, dev in enumerate(devs):
            _close_to_oracle_ions(r.Y[k], dev, _gpu_ion_state(oracle, dev, ys[k], V),
                                  laps_of(FREQS, dt), dt, bern)


@pytest.mark.gpu
def test_gpu_impedance_with_ions_limits(gpu_ion_batches):
    """With ions on the GPU alone: at 1e-9 Hz the slope of the GPU's own steady curve with free
    ions (+-10 uV, to 1e-5), at 1e14 Hz the geometric capacitance (1e-5, 120 points)."""
    from jaxdd_amd.batch import solve_batch
    from jaxdd_amd.impedance import solve_impedance
    from oracle import ddimpedance as I

    devs, batch, bern, V = gpu_ion_batches[0]
    r = solve_impedance(batch, [1e-9, 1e14], voltage=V, tol=1e-12, max_iter=200)
    dV = 1e-5
    J, ok, _ = solve_batch(batch, [V - dV, V + dV], tol=1e-12, max_iter=200)
    assert ok.cpu().numpy().all()
    J = J.cpu().numpy()
    slope = (J[:, 1] - J[:, 0]) / (2 * dV) - batch["G_shunt"].cpu().numpy()
    assert np.allclose(r.Y[:, 0].real, slope, rtol=1e-5, atol=0)
    geo = np.array([I.geometric_capacitance(d) for d in devs])
    assert np.allclose(r.capacitance[:, 1], geo, rtol=1e-5, atol=0)
