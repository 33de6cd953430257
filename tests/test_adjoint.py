"""The adjoint (dJ/dtheta by the implicit function theorem) on the CPU: the adjoint kernel's
per-lane code replayed (tests/kernel_emu) against central finite differences of the oracle's
curves, in the reference's own form (tests/test_solver.py:1784-1811: a random weighting of
the curves, a random direction in the parameters, the directional derivative against central
differences of the weighted sum)."""

from __future__ import annotations

import numpy as np
import pytest
from conftest import ION_WRT, named_fd_errors, number_of

N_DEV = 4
V_SUB = slice(None, None, 6)  # 11 of the 61 voltages


@pytest.fixture(scope="module")
def c4_case(emu):
    from jaxdd_amd import adjoint as A
    from jaxdd_amd import synthetic as S

    th = A.theta_from_table(S.draw(N_DEV), range(N_DEV))
    packed, tan = A.tangents(th)
    V = S.VOLTAGES[V_SUB].copy()
    states = np.zeros((len(V), 400, 3, N_DEV))
    J, ok, its, _ = emu.solve_packed(packed, V, tol=1e-12, strict=True, max_iter=200,
                                     states=states)
    grad, status = emu.adjoint_packed(packed, V, states, ok, tan)
    return th, packed, tan, V, J, ok, grad, status


def _theta_at(th, d, a):
    out = th * (1.0 + a * d)
    out[:, 15] = th[:, 15] + a * d[:, 15] * 0.01  # the work-function offset: 10 mV units
    return out


def test_tangents_are_the_builders_derivatives():
    """The packed arrays' tangents by the builder's central differences agree with one-sided
    differences of a ten times larger step (a smooth builder at a fixed grid layout)."""
    from jaxdd_amd import adjoint as A
    from jaxdd_amd import synthetic as S
    from jaxdd_amd.batch import pack

    th = A.theta_from_table(S.draw(1), [0])
    st = A.structure_of(th[0])
    base, tan = A.tangents(th)
    for k in (1, 3, 10, 14, 15):
        h = 10 * A._steps(th)[0, k]
        e = np.zeros(16)
        e[k] = h
        hi = pack([A.device_from_theta(th[0] + e, structure=st)])
        for key in ("pt", "seg", "sc"):
            one = (hi[key] - base[key]) / h
            ref = np.max(np.abs(tan[key][k])) + 1e-300
            assert np.max(np.abs(one - tan[key][k])) < 1e-3 * ref, (A.THETA[k], key)


def test_adjoint_is_the_finite_difference_of_the_curves(c4_case, oracle):
    from jaxdd_amd import adjoint as A

    th, _, _, V, J, ok, grad, status = c4_case
    assert not status.any() and ok.all()
    rng = np.random.default_rng(0)
    w = rng.normal(size=ok.shape)
    d = rng.normal(size=th.shape)
    scale = th.copy()
    scale[:, 15] = 0.01
    by_adjoint = np.sum(w[:, :, None] * grad * (scale * d)[:, None, :])

    structs = [A.structure_of(t) for t in th]  # each device's layout, held as theta moves

    def loss(a):
        devs = [A.device_from_theta(x, structure=st) for x, st in zip(_theta_at(th, d, a), structs)]
        Jx, okx, _ = oracle.solve_batch(devs, V, tol=1e-12, strict=True, first_max_iter=None,
                                        max_iter=200)
        assert okx.all()
        return np.sum(w * Jx)

    h = 1e-5
    fd = (loss(h) - loss(-h)) / (2 * h)
    assert abs(by_adjoint - fd) < 1e-5 * abs(fd)


def test_each_parameter_against_finite_differences(c4_case, oracle):
    """Per device and parameter, dJ/dtheta_k of the whole curve against central differences
    (relative step 1e-5), measured as the reference measures J: the error of dJ/dln(theta_k)
    relative to the curve's current scale."""
    from jaxdd_amd import adjoint as A
    from jaxdd_amd.batch import j_scale

    th, _, _, V, J, ok, grad, _ = c4_case
    dev_idx = 1
    st = A.structure_of(th[dev_idx])  # the device's layout, held as theta moves
    dev = A.device_from_theta(th[dev_idx], structure=st)
    cscale = j_scale(dev) + np.max(np.abs(J[dev_idx]))
    for k in range(16):
        step = 1e-5 * abs(th[dev_idx, k]) if k != 15 else 1e-5
        e = np.zeros(16)
        e[k] = step
        Jp = oracle.solve_batch([A.device_from_theta(th[dev_idx] + e, structure=st)], V,
                                tol=1e-12, strict=True, first_max_iter=None, max_iter=200)[0][0]
        Jm = oracle.solve_batch([A.device_from_theta(th[dev_idx] - e, structure=st)], V,
                                tol=1e-12, strict=True, first_max_iter=None, max_iter=200)[0][0]
        fd = (Jp - Jm) / (2 * step)
        unit = abs(th[dev_idx, k]) if k != 15 else 1.0
        err = np.max(np.abs(grad[dev_idx, :, k] - fd)) * unit / cscale
        assert err < 1e-5, (A.THETA[k], err)


def test_thickness_moves_the_points_by_scaling_only():
    """The reference's semantics of a thickness's derivative (src/jaxdd/batch.py:142-148
    `device_tree(structure, params)`, src/jaxdd/device.py:225-233 `_spacing`): the layout is the
    structure's, so the points of the layer whose thickness grows move by scaling its spacing
    (with the interface means); here the tangents of x against the closed form of `place`."""
    from jaxdd_amd import adjoint as A
    from jaxdd_amd import synthetic as S

    th = A.theta_from_table(S.draw(2), [1])
    _, tan = A.tangents(th)
    lay = A.structure_of(th[0]).layout
    for k in range(3):
        h = 1e-6 * th[0, k]
        Lp, Lm = th[0, 0:3].copy(), th[0, 0:3].copy()
        Lp[k] += h
        Lm[k] -= h
        ref = (lay.place(Lp)[0] - lay.place(Lm)[0]) / (2 * h)
        got = tan["pt"][k, :, 1, 0]  # d x / d L_k
        assert np.max(np.abs(got - ref)) < 1e-7 * np.max(np.abs(ref))


def test_unconverged_voltages_have_zero_gradient(emu):
    """As the reference's (tests/test_solver.py:1814-1829): a voltage that misses max_iter
    contributes a zero, finite gradient."""
    from jaxdd_amd import adjoint as A
    from jaxdd_amd import synthetic as S

    th = A.theta_from_table(S.draw(2), [0, 1])
    packed, tan = A.tangents(th)
    V = S.VOLTAGES[::10].copy()
    states = np.zeros((len(V), 400, 3, 2))
    J, ok, its, _ = emu.solve_packed(packed, V, max_iter=5, start_iter=60, states=states)
    assert ok.any() and not ok.all()
    grad, status = emu.adjoint_packed(packed, V, states, ok, tan)
    assert np.isfinite(grad).all()
    assert np.all(grad[~ok] == 0.0) and np.any(grad[ok] != 0.0)


def test_devices_with_frozen_ions_need_the_pre_bias_solution(emu):
    """Ions frozen at a pre-bias: the scan's fixed charge is the pre-bias solution's ions, so
    the adjoint needs that solution; without it such a device is refused (status 1, NaN)."""
    from jaxdd_amd import synthetic as S
    from jaxdd_amd.batch import pack

    packed = pack(S.devices("c3", 1, variant="frozen"))
    V = S.VOLTAGES[::20].copy()
    states = np.zeros((len(V), 400, 3, 1))
    J, ok, its, _ = emu.solve_packed(packed, V, states=states)
    tan = {"pt": np.zeros((1,) + packed["pt"].shape), "seg": np.zeros((1,) + packed["seg"].shape),
           "sc": np.zeros((1,) + packed["sc"].shape)}
    grad, status = emu.adjoint_packed(packed, V, states, ok, tan)
    assert status[0] == 1 and np.isnan(grad).all()


#: The frozen stack's numbers: the free-ion ones and the pre-bias voltage the ions freeze at.
FROZEN_WRT = ION_WRT + ["V_pre"]


@pytest.fixture(scope="module")
def frozen_case(emu):
    from jaxdd_amd import adjoint as A
    from jaxdd_amd import synthetic as S

    t = S.draw(2)
    structs, params = zip(*[S.row_inputs(t, k, 100, ions=True, frozen=True) for k in range(2)])
    dirs = A.directions_for(FROZEN_WRT, 2, len(structs[0].i0))
    packed, tan = emu.build_packed(list(structs), list(params), dirs)
    V = S.VOLTAGES[::10].copy()
    states = np.zeros((len(V), 100, 3, 2))
    pre = np.zeros((100, 3, 2))
    J, ok, its, _ = emu.solve_packed(packed, V, tol=1e-12, strict=True, max_iter=200,
                                     states=states, pre=pre)
    grad, status = emu.adjoint_packed(packed, V, states, ok, tan, pre=pre)
    return structs, params, packed, tan, V, J, ok, states, pre, grad, status


def test_frozen_ions_gradients_are_finite_differences(frozen_case, oracle):
    """Ions frozen at a pre-bias of 1.2 V (reference: src/jaxdd/batch.py:937-960, differentiated
    through the pre-bias solution as `_freeze` is): the scan's adjoint and a second transposed
    bordered solve at the pre-bias solution, against central differences of the oracle's curves
    in the named numbers and the pre-bias voltage. 1e-5 of the curve's scale below 1.3 V; at
    1.3 V, where a 1e-15 relative change of the scan's state moves a thickness's or the work
    function's component by 5e-5, 1e-4."""
    structs, params, _, _, V, J, ok, _, pre, grad, status = frozen_case
    assert not status.any() and ok.all() and np.abs(pre).max() > 0
    err = named_fd_errors(oracle, structs, params, FROZEN_WRT, V, J, grad, per_voltage=True)
    low = V < 1.25
    assert np.max(err[:, :, low]) < 1e-5, np.max(err[:, :, low], axis=2)
    assert np.max(err) < 1e-4, np.max(err, axis=2)
    assert np.median(err) < 3e-8, np.median(err)


def test_frozen_ions_reverse_pass_is_the_directions_passes(frozen_case, emu):
    """The two passes of a frozen device (the scan's, then the pre-bias solution's) by the
    reverse pass and by one dual pass per direction."""
    from jaxdd_amd.batch import j_scale
    from jaxdd_amd.device import device_arrays

    structs, params, packed, tan, V, J, ok, states, pre, grad, _ = frozen_case
    by_pass, st2 = emu.adjoint_packed(packed, V, states, ok, tan, passes=True, pre=pre)
    assert not st2.any()
    unit = np.array([[abs(number_of(p, w)) for w in FROZEN_WRT] for p in params])
    scale = np.array([j_scale(device_arrays(st, p)) for st, p in zip(structs, params)])
    scale = scale + np.max(np.abs(J), axis=1)
    err = np.abs(grad - by_pass) * unit[:, None, :] / scale[:, None, None]
    assert np.max(err) < 5e-11, np.unravel_index(np.argmax(err), err.shape)


@pytest.mark.parametrize("mode", ["layer", "simss"])
def test_langevin_gradients_are_finite_differences(emu, oracle, mode):
    """Langevin's rate (reference: src/jaxdd/device.py:900-911, its gradient in `preLangevin`
    tests/test_solver.py:1104-1122): the adjoint replay's gradients in the rate's numbers
    (prefactors, mobilities, permittivity, an interface velocity, a thickness) against central
    differences of the oracle's curves, each device's structure held."""
    from conftest import LANGEVIN_WRT, langevin_rows
    from jaxdd_amd import adjoint as A
    from jaxdd_amd import synthetic as S

    inputs = langevin_rows(2, 100, mode)
    structs, params = zip(*inputs)
    dirs = A.directions_for(LANGEVIN_WRT, 2, len(structs[0].i0))
    packed, tan = emu.build_packed(list(structs), list(params), dirs)
    V = S.VOLTAGES[::10].copy()
    states = np.zeros((len(V), 100, 3, 2))
    J, ok, its, _ = emu.solve_packed(packed, V, tol=1e-12, strict=True, max_iter=200,
                                     states=states)
    grad, status = emu.adjoint_packed(packed, V, states, ok, tan)
    assert not status.any() and ok.all()
    err = named_fd_errors(oracle, structs, params, LANGEVIN_WRT, V, J, grad)
    assert np.max(err) < 1e-5, err
    assert np.median(err) < 3e-7, err


@pytest.fixture(scope="module")
def ion_case(emu):
    from jaxdd_amd import adjoint as A
    from jaxdd_amd import synthetic as S

    t = S.draw(2)
    structs, params = zip(*[S.row_inputs(t, k, 100, ions=True) for k in range(2)])
    dirs = A.directions_for(ION_WRT, 2, len(structs[0].i0))
    packed, tan = emu.build_packed(list(structs), list(params), dirs)
    V = S.VOLTAGES[::10].copy()
    states = np.zeros((len(V), 100, 3, 2))
    J, ok, its, _ = emu.solve_packed(packed, V, tol=1e-12, strict=True, max_iter=200,
                                     states=states)
    grad, status = emu.adjoint_packed(packed, V, states, ok, tan)
    return structs, params, packed, tan, V, J, ok, states, grad, status


def test_mobile_ions_gradients_are_finite_differences(ion_case, oracle):
    """Free mobile ions (the border of each species' normalisation and number, transposed:
    reference src/jaxdd/batch.py:595-618 `_transposed`): the gradients in the named numbers,
    the ions' own densities among them, against central differences of the oracle's curves."""
    structs, params, _, _, V, J, ok, _, grad, status = ion_case
    assert not status.any() and ok.all()
    err = named_fd_errors(oracle, structs, params, ION_WRT, V, J, grad)
    assert np.max(err) < 1e-5, err
    assert np.median(err) < 3e-7, err


def test_mobile_ions_reverse_pass_is_the_directions_passes(ion_case, emu):
    """The reverse pass (the constraints' partials reversed with the rows') against one dual
    pass per direction (the numbers' tangents forward), per device, voltage and number."""
    from jaxdd_amd.batch import j_scale
    from jaxdd_amd.device import device_arrays

    structs, params, packed, tan, V, J, ok, states, grad, _ = ion_case
    by_pass, st2 = emu.adjoint_packed(packed, V, states, ok, tan, passes=True)
    assert not st2.any()
    unit = np.array([[abs(number_of(p, w)) for w in ION_WRT] for p in params])
    scale = np.array([j_scale(device_arrays(st, p)) for st, p in zip(structs, params)])
    scale = scale + np.max(np.abs(J), axis=1)
    err = np.abs(grad - by_pass) * unit[:, None, :] / scale[:, None, None]
    # (the temperature's terms, through every exponential, cancel to 1e-11 of the scale)
    assert np.max(err) < 5e-11, np.unravel_index(np.argmax(err), err.shape)


def test_reverse_pass_is_the_directions_passes(c4_case, emu):
    """The one reverse pass over the parameters (every direction contracted on the way) and one
    forward dual-number pass per direction: the same derivative by two independent
    differentiations of the same rows and current, per device, voltage and parameter."""
    th, packed, tan, V, J, ok, grad, status = c4_case
    states = np.zeros((len(V), 400, 3, N_DEV))
    emu.solve_packed(packed, V, tol=1e-12, strict=True, max_iter=200, states=states)
    from jaxdd_amd import adjoint as A
    from jaxdd_amd.batch import j_scale

    by_pass, st2 = emu.adjoint_packed(packed, V, states, ok, tan, passes=True)
    assert not st2.any()
    # dJ/dln(theta_k) against each curve's current scale, as the finite-difference tests
    # measure (a parameter J hardly depends on, such as a trap density far from where carriers
    # recombine, has its two terms cancel to round-off of the larger; a thickness's terms,
    # through every point's position, cancel to 1.2e-11 of the scale at worst)
    unit = np.abs(th)
    unit[:, 15] = 1.0
    cscale = np.array([j_scale(A.device_from_theta(x)) for x in th]) + np.max(np.abs(J), axis=1)
    err = np.abs(grad - by_pass) * unit[:, None, :] / cscale[:, None, None]
    assert np.max(err) < 5e-11, np.unravel_index(np.argmax(err), err.shape)
    assert np.max(np.abs(by_pass) * unit[:, None, :] / cscale[:, None, None]) > 0.1


@pytest.mark.parametrize("ions", [False, True])
def test_gradient_in_the_voltage_is_the_curves_slope(ions, emu, oracle):
    """dJ/dV, the applied voltage as a direction (`adjoint.voltage_tangents`: its tangent on the
    contacts' potentials), against central differences of the oracle's curves in the voltage
    (h = 1e-5 V), the shunt's G_shunt added, to 1e-5 of the curve's largest slope as the other
    per-component tests: devices without and with free mobile ions. (At 1.3 V the transposed
    system's condition number reaches 1e33; the block elimination's dJ/dV is then 5.8e-6 from
    the differences where a pivoted dense solve at the same state is 8e-7 from them.)"""
    import torch

    from jaxdd_amd import adjoint as A
    from jaxdd_amd import synthetic as S
    from jaxdd_amd.batch import SC_FIELDS
    from jaxdd_amd.device import device_arrays

    t = S.draw(2)
    structs, params = zip(*[S.row_inputs(t, k, 100, ions=ions) for k in range(2)])
    packed, tan = emu.build_packed(list(structs), list(params),
                                   A.directions_for(["V", ("mu_n", 1)], 2, len(structs[0].i0)))
    tan = A.voltage_tangents({"_sc": torch.as_tensor(packed["sc"])}, tan, [0])
    V = S.VOLTAGES[::10].copy()
    states = np.zeros((len(V), 100, 3, 2))
    J, ok, _, _ = emu.solve_packed(packed, V, tol=1e-12, strict=True, max_iter=200,
                                   states=states)
    grad, status = emu.adjoint_packed(packed, V, states, ok, tan)
    assert ok.all() and not status.any()
    dJdV = grad[:, :, 0] + packed["sc"][SC_FIELDS.index("G_shunt")][:, None]
    h = 1e-5
    for d in range(2):
        dev = device_arrays(structs[d], params[d])
        Jp, Jm = (oracle.solve_batch([dev], V + s * h, tol=1e-12, strict=True,
                                     first_max_iter=None, max_iter=200)[0][0] for s in (1, -1))
        fd = (Jp - Jm) / (2 * h)
        assert np.max(np.abs(dJdV[d] - fd)) < 1e-5 * np.max(np.abs(fd)), (d, dJdV[d], fd)
        assert np.median(np.abs(dJdV[d] - fd)) < 1e-6 * np.max(np.abs(fd))



#: Numbers the charged traps' charge depends on: their density, capture coefficients and
#: level, and a mobility beside them.
CHARGED_WRT = (("N_t_bulk", 1), ("C_n_bulk", 1), ("C_p_bulk", 1), ("E_t_bulk", 1), ("mu_n", 1))


def test_charged_trap_gradients_are_finite_differences(emu, oracle):
    """Acceptor bulk traps charged in the absorber (SimSS's run `bulk_acceptor_light`): the
    adjoint replay's gradients in the traps' numbers, their charge's partials included
    (reference: solver.py:762-773, 1136-1140), with the host builder's tangents (central
    differences, relative step 1e-6), against central differences of the oracle's curves."""
    from conftest import bumped, fixture_inputs, named_fd_errors, number_of
    from jaxdd_amd.batch import pack
    from jaxdd_amd.device import device_arrays

    st, p = fixture_inputs("bulk_acceptor_light")
    packed = pack([device_arrays(st, p)])
    assert np.any(packed["seg"][:, -1, 0] != 0)
    tan = {key: np.zeros((len(CHARGED_WRT),) + packed[key].shape) for key in ("pt", "seg", "sc")}
    for k, w in enumerate(CHARGED_WRT):
        h = 1e-6 * abs(number_of(p, w))
        up, dn = (pack([device_arrays(st, bumped(p, w, s * h))]) for s in (1.0, -1.0))
        for key in tan:
            tan[key][k] = (up[key] - dn[key]) / (2 * h)
    V = np.array([-0.2, 0.3, 0.6, 0.8, 0.95, 1.05])
    states = np.zeros((len(V), st.NP, 3, 1))
    J, ok, _, _ = emu.solve_packed(packed, V, tol=1e-12, strict=True, max_iter=200,
                                   states=states)
    grad, status = emu.adjoint_packed(packed, V, states, ok, tan)
    assert not status.any() and ok.all()
    err = named_fd_errors(oracle, [st], [p], CHARGED_WRT, V, J, grad)
    print(f"charged traps, replay vs differences: {err}")
    assert np.max(err) < 1e-8, err  # measured: 5e-10 at most
    assert np.median(err) < 1e-9, err
