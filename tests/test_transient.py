"""Transients (reference: src/jaxdd/transient.py; its tests tests/test_solver.py:504-580): the
oracle's restatement pinned on ZimT 5.36's own runs at the reference's tolerances, its
time-step Jacobian against central differences of its residual, a held voltage's transient
against the steady solve, and the impedance oracle's time terms against one implicit-Euler
step of the transient (the mid-band of the small-signal response)."""

from __future__ import annotations


import numpy as np
import pytest

from conftest import GOLDEN

#: ZimT 5.36 runs copied from the reference's tests/data, with the tolerance (of the largest
#: |Jext|) the reference's own tests hold them to (tests/test_solver.py:454-470, 504-616):
#: electrons and holes to 1e-9 (`zimt_sweep_spv_light` is the sweep with ZimT's layer voltages
#: written, `zimt_steady_light` seven steady states, t = 0 on every line, the last at half
#: light), bulk traps in time to 1e-9, mobile ions (5 x 5) to 1e-7.
ZIMT = {
    "zimt_photo_light": 1e-9,
    "zimt_sweep_light": 1e-9,
    "zimt_sweep_spv_light": 1e-9,
    "zimt_steady_light": 1e-9,
    "zimt_traps_photo_light": 1e-9,
    "zimt_ions_slow_light": 1e-7,
    "zimt_ions_fast_light": 1e-7,
}


def zimt_case(case):
    from jaxdd_amd.device import build_device
    from jaxdd_amd.simss_input import read_setup, read_tvg

    setup = read_setup(GOLDEN / case / "simulation_setup.txt")
    tvg = read_tvg(setup)
    ref = np.genfromtxt(GOLDEN / case / "tj.dat", names=True)
    return build_device(setup, grid="simss"), tvg, ref


@pytest.mark.parametrize("case", sorted(ZIMT))
def test_transient_oracle_reproduces_zimt(oracle, case):
    """ZimT's Jext through 240 steps (as ZimT computes: SimSS's Bernoulli function, ZimT's
    sums in the integral current; light switched at short circuit over six decades of time, a
    sweep to 1.2 V and back, traps following in time, ions in the absorber swept slowly and
    fast)."""
    dev, tvg, ref = zimt_case(case)
    r = oracle.transient(dev, tvg.t, np.where(tvg.oc, 0.0, tvg.Vext), tvg.G_frac,
                         bernoulli_name="simss", zimt_current=True,
                         failure_mode=tvg.failure_mode)
    assert r["converged"].all() and r["kept"].all()
    scale = np.max(np.abs(ref["Jext"]))
    assert np.max(np.abs(r["J_ext"] - ref["Jext"])) < ZIMT[case] * scale


def _state(oracle, dev, V, ions):
    """A converged steady state at V as a full state ((NP + 2) x 3 or 5)."""
    from oracle.ddimpedance import full_state

    _, ok, _, st = oracle.solve_one(dev, [V], return_states=True)
    assert ok.all()
    full = full_state(dev, st[0], V)
    if not ions:
        return full
    Vt = dev.Vt
    am, cm = np.asarray(dev.anion_mask, float), np.asarray(dev.cation_mask, float)
    Vf = full[:, 0]
    hL = dev.h[:-1] * dev.L
    cols = []
    for sign, m, tot in ((1.0, am, dev.anion_total), (-1.0, cm, dev.cation_total)):
        prof = m * np.exp(sign * Vf / Vt)
        num = np.sum(0.5 * (prof[1:] + prof[:-1]) * m[1:] * m[:-1] * hL)
        cols.append(np.where(m > 0, np.log(tot / num) + sign * Vf / Vt, 0.0))
    return np.concatenate([full, np.stack(cols, -1)], axis=1)


@pytest.mark.parametrize("ions", [False, True])
@pytest.mark.parametrize("dti", [0.0, 1e6])
def test_time_step_jacobian_is_the_residuals_derivative(oracle, ions, dti):
    """The time step's blocks (A, B, C) and the slots' columns (U) against central differences of
    the oracle's own residual, with the traps' occupation lagging and (ions) the ions as
    unknowns of their own."""
    from jaxdd_amd import synthetic as S

    t = S.draw(6)
    dev = S.device_from_row(t, 3, 24, ions, False)
    full = _state(oracle, dev, 0.4, ions)
    rng = np.random.default_rng(5)
    prev = full + 1e-2 * rng.standard_normal(full.shape) * (full != 0)
    now = full + 1e-3 * rng.standard_normal(full.shape) * (full != 0)
    NP = len(dev.x) - 2
    fb = rng.uniform(0.1, 0.9, NP)
    logs = np.array([0.03, -0.02]) if ions else np.zeros(2)
    F, A, B, C, U, _, _ = oracle.assemble_step(dev, now, 1.0, dti, prev, fb, logs)
    nc = now.shape[1]
    rs = np.max(np.abs(np.concatenate([A, B, C], axis=2)), axis=2) + 1e-300
    for j in range(NP + 2):
        for c in range(nc):
            h = 1e-6 * max(1.0, abs(now[j, c]))
            up, dn = now.copy(), now.copy()
            up[j, c] += h
            dn[j, c] -= h
            dF = (oracle.assemble_step(dev, up, 1.0, dti, prev, fb, logs)[0]
                  - oracle.assemble_step(dev, dn, 1.0, dti, prev, fb, logs)[0]) / (2 * h)
            for r in range(NP):
                off = j - (r + 1)
                if abs(off) > 1:
                    continue
                got = {-1: A, 0: B, 1: C}[off][r, :, c]
                assert np.all(np.abs(got - dF[r]) <= 1e-6 * rs[r]), (j, c, r)
    for k in range(2):
        h = 1e-6
        lu, ld = logs.copy(), logs.copy()
        lu[k] += h
        ld[k] -= h
        dF = (oracle.assemble_step(dev, now, 1.0, dti, prev, fb, lu)[0]
              - oracle.assemble_step(dev, now, 1.0, dti, prev, fb, ld)[0]) / (2 * h)
        assert np.all(np.abs(U[:, :, k] - dF) <= 1e-6 * (rs + np.abs(dF)))


def test_a_held_voltage_stays_at_the_steady_current(oracle):
    """With the voltage and the light held, every step of a transient is the steady state: its
    current the steady solve's at every time (traps and ions included)."""
    from jaxdd_amd import synthetic as S

    t = S.draw(4)
    for ions in (False, True):
        dev = S.device_from_row(t, 1, 80, ions, False)
        V = 0.7
        times = np.concatenate([[0.0], np.logspace(-9, 1, 30)])
        r = oracle.transient(dev, times, np.full(len(times), V), np.ones(len(times)))
        Js, oks, _ = oracle.solve_batch([dev], [V], first_max_iter=None)
        assert r["converged"].all() and oks.all()
        scale = oracle.j_scale(dev) + abs(Js[0, 0])
        assert np.max(np.abs(r["J_ext"] - Js[0, 0])) < 1e-8 * scale


def test_one_implicit_euler_step_is_the_impedance_at_its_rate(oracle):
    """The impedance oracle's time terms (the densities' -fac n dti, the traps' lag, the
    current's densities' change and displacement current) against the transient: one
    implicit-Euler step of length dt from a steady state with a voltage step dV changes the
    current by Y(lap = 1/dt) dV, to the step's linearisation (central in +-dV). The impedance
    oracle's two limits leave this mid-band free."""
    from jaxdd_amd import synthetic as S
    from oracle import ddimpedance as I

    t = S.draw(6)
    for k in (2, 4):
        dev = S.device_from_row(t, k, 120, False, False)
        V0, dV = 0.6, 1e-4
        full = _state(oracle, dev, V0, False)
        Y_dc = I.admittance(dev, full, np.array([1e-9 + 0j]))[0].real
        for dt in (1e-9, 1e-7, 1e-5, 1e-3):
            Y = I.admittance(dev, full, np.array([1.0 / dt + 0j]), dt=dt)[0]
            J = []
            for s in (1, -1):
                r = oracle.transient(dev, [0.0, dt], [V0, V0 + s * dV], [1.0, 1.0], tol=1e-14)
                assert r["converged"].all()
                J.append(r["J"][1])
            slope = (J[0] - J[1]) / (2 * dV)
            assert abs(Y.imag) < 1e-12 * abs(Y.real)
            assert abs(slope / Y.real - 1) < 1e-6, (k, dt, slope, Y)
            if dt <= 1e-5:  # the mid-band: far from the steady slope
                assert Y.real > 4 * Y_dc


# ---- the kernel's per-lane code on the CPU, and the GPU -------------------------------------
def c5_steps(nsteps, V_step=0.9):
    """The transient workload's steps: the steady state at 0 V, then a step to ``V_step`` held
    through ``nsteps`` log-spaced implicit-Euler steps from 1 ns to 10 s, lit."""
    t = np.concatenate([[0.0], np.logspace(-9, 1, nsteps)])
    V = np.concatenate([[0.0], np.full(nsteps, V_step)])
    return t, V, np.ones(nsteps + 1)


def _replay(emu, devs, t, V, G, bern="exact", zimt=False, failure_mode=1):
    from jaxdd_amd.batch import pack

    packed = pack(devs)
    start = dict(packed, sc=packed["sc"].copy())
    start["sc"][2] = G[0]
    NP = len(devs[0].x) - 2
    st = np.zeros((1, NP, 3, len(devs)))
    emu.solve_packed(start, np.array([V[0]]), max_iter=200, bernoulli=bern, states=st)
    return emu.transient_packed(packed, t, V, G, st[0], bern=bern, zimt=zimt,
                                failure_mode=failure_mode)


@pytest.mark.parametrize("case", sorted(ZIMT))
def test_replay_transient_is_the_oracles_on_zimt_runs(oracle, emu, case):
    """The transient kernel's code on the CPU against the oracle on ZimT's runs: the same
    Newton steps, the same currents to 1e-10 of the scale, and ZimT's to its tolerance."""
    dev, tvg, ref = zimt_case(case)
    V = np.where(tvg.oc, 0.0, tvg.Vext)
    J, Jx, ok, its, kept, status = _replay(emu, [dev], tvg.t, V, tvg.G_frac, "simss", True,
                                           tvg.failure_mode)
    r = oracle.transient(dev, tvg.t, V, tvg.G_frac, bernoulli_name="simss", zimt_current=True,
                         failure_mode=tvg.failure_mode)
    scale = np.max(np.abs(ref["Jext"]))
    assert not status.any() and ok.all() and kept.all()
    assert np.array_equal(its[0], r["iterations"])
    assert np.max(np.abs(Jx[0] - r["J_ext"])) < 1e-10 * scale
    assert np.max(np.abs(Jx[0] - ref["Jext"])) < ZIMT[case] * scale


@pytest.mark.parametrize("ions", [True, False])
def test_replay_transient_is_the_oracles_on_the_workload(oracle, emu, ions):
    """The transient workload's devices (ions and traps in the perovskite, NP = 300) through a
    voltage step: the kernel's code against the oracle, curve by curve (1e-9 of each curve's
    scale), flags equal."""
    from jaxdd_amd import synthetic as S

    tab = S.draw(4)
    devs = [S.device_from_row(tab, k, 300, ions, False) for k in range(4)]
    t, V, G = c5_steps(80)
    J, Jx, ok, its, kept, status = _replay(emu, devs, t, V, G)
    assert not status.any()
    for k, dev in enumerate(devs):
        r = oracle.transient(dev, t, V, G)
        assert np.array_equal(ok[k], r["converged"]) and np.array_equal(kept[k], r["kept"])
        scale = oracle.j_scale(dev) + np.max(np.abs(r["J_ext"]))
        assert np.max(np.abs(Jx[k] - r["J_ext"])) < 1e-9 * scale


def test_what_the_transient_refuses():
    from jaxdd_amd.simss_input import TVG
    from jaxdd_amd.transient import _steps

    base = dict(t=np.array([0.0, 1e-6]), Vext=np.array([0.0, 0.5]), oc=np.array([False, False]),
                G_frac=np.ones(2), track=np.zeros(2, int), dti=np.array([0.0, 1e6]))
    _steps(TVG(**base), 2)
    with pytest.raises(NotImplementedError, match="open circuit"):
        _steps(TVG(**dict(base, oc=np.array([False, True]))), 2)
    with pytest.raises(NotImplementedError, match="Track"):
        _steps(TVG(**dict(base, track=np.array([0, 1]))), 2)
    with pytest.raises(ValueError, match="one per device"):
        _steps([TVG(**base)] * 3, 2)


GPU_OFF = ("the GPU transient kernel is switched off: its first run faulted and the cause is not "
           "yet shown (the CPU replay of its code passes bounds, ASan and UBSan checks)")


@pytest.mark.gpu
@pytest.mark.skip(reason=GPU_OFF)
@pytest.mark.parametrize("case", sorted(ZIMT))
def test_gpu_transient_reproduces_zimt(torch_cuda, oracle, case):
    """solve_transient on the GPU (ddh_transient through the C-ABI): ZimT's Jext to its
    tolerance and the oracle's to 1e-9 of the scale, every step converged and kept."""
    from jaxdd_amd.batch import prepare
    from jaxdd_amd.transient import solve_transient

    dev, tvg, ref = zimt_case(case)
    print(f"transient {case} on the GPU", flush=True)
    r = solve_transient(prepare([dev]), tvg, bernoulli_name="simss", zimt_current=True)
    o = oracle.transient(dev, tvg.t, np.where(tvg.oc, 0.0, tvg.Vext), tvg.G_frac,
                         bernoulli_name="simss", zimt_current=True, failure_mode=tvg.failure_mode)
    scale = np.max(np.abs(ref["Jext"]))
    assert r.converged.all() and r.accepted.all()
    assert np.max(np.abs(r.J_ext[0] - o["J_ext"])) < 1e-9 * scale
    assert np.max(np.abs(r.J_ext[0] - ref["Jext"])) < ZIMT[case] * scale


@pytest.mark.gpu
@pytest.mark.skip(reason=GPU_OFF)
@pytest.mark.parametrize("ions", [True, False])
def test_gpu_transient_matches_the_oracle_on_the_workload(torch_cuda, oracle, ions):
    """The transient workload's devices (NP = 300, 24 devices: a partial wavefront) through a
    voltage step on the GPU against the oracle: flags equal, each curve to 1e-9 of its scale."""
    from jaxdd_amd import synthetic as S
    from jaxdd_amd.batch import prepare
    from jaxdd_amd.simss_input import TVG
    from jaxdd_amd.transient import solve_transient

    tab = S.draw(24)
    devs = [S.device_from_row(tab, k, 300, ions, False) for k in range(24)]
    t, V, G = c5_steps(120)
    tvg = TVG(t=t, Vext=V, oc=np.zeros(len(t), bool), G_frac=G, track=np.zeros(len(t), int),
              dti=np.zeros(len(t)))
    print(f"transient workload ions={ions}: 24 devices x {len(t)} steps", flush=True)
    r = solve_transient(prepare(devs), tvg)
    for k, dev in enumerate(devs):
        o = oracle.transient(dev, t, V, G)
        assert np.array_equal(r.converged[k], o["converged"])
        assert np.array_equal(r.accepted[k], o["kept"])
        scale = oracle.j_scale(dev) + np.max(np.abs(o["J_ext"]))
        assert np.max(np.abs(r.J_ext[k] - o["J_ext"])) < 1e-9 * scale


def test_solve_transient_refuses_while_the_gpu_kernel_is_off():
    from jaxdd_amd.transient import solve_transient

    with pytest.raises(NotImplementedError, match="switched off"):
        solve_transient({}, None)
