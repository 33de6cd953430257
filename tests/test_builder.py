"""The batch built on the GPU (jaxdd_amd.builder, ddh_build; reference: src/jaxdd/device.py:877-1063
`device_arrays`, src/jaxdd/batch.py:142-223): its lanes' code on the CPU against the host
builder's packed arrays, value by value, and its forward-mode tangents against central
differences of the host builder on each device's structure; the GPU against both."""

import numpy as np
import pytest

V_FULL = np.round(np.arange(-0.2, 1.3 + 1e-9, 0.025), 6)


def _np(t):
    return t.detach().cpu().numpy()


def _rows(n, NP, ions=False, frozen=False):
    from jaxdd_amd import synthetic as S

    t = S.draw(n)
    return [S.row_inputs(t, k, NP, ions, frozen) for k in range(n)]


def _same_packed(got, ref, rtol=1e-14):
    """The builder's arrays against `batch.pack`'s: the segments and every field, each to rtol
    of its largest magnitude over the batch (as packed fields span 40 decades)."""
    assert got["nseg"] == ref["nseg"]
    assert np.array_equal(np.asarray(got["sid"]), ref["sid"])
    for key in ("pt", "seg", "sc", "imu"):
        g, r = np.asarray(got[key]), ref[key]
        axis = tuple(i for i in range(r.ndim) if i != r.ndim - 2)  # per field
        scale = np.max(np.abs(r), axis=axis, keepdims=True)
        assert np.all(np.abs(g - r) <= rtol * np.where(scale > 0, scale, 1.0)), key


@pytest.mark.parametrize("config", ["c2", "c3", "c3-frozen", "theta"])
def test_replay_builds_the_host_batch(emu, config):
    """The builder's lanes on the CPU give `batch.pack`'s arrays of the host builder's devices,
    to 1e-14 of each field's scale: the workloads' stacks without ions, with mobile ions, with
    ions held after a pre-bias, and the gradient workload's parameter vectors."""
    from jaxdd_amd import adjoint as A
    from jaxdd_amd import synthetic as S
    from jaxdd_amd.batch import pack
    from jaxdd_amd.device import device_arrays

    if config == "theta":
        th = A.theta_from_table(S.draw(6), range(6))
        inputs = [(A.structure_of(t, 120), A.params_of(t)) for t in th]
    else:
        inputs = _rows(6, 120, ions=config != "c2", frozen=config == "c3-frozen")
    structs, params = zip(*inputs)
    ref = pack([device_arrays(st, p) for st, p in inputs])
    _same_packed(emu.build_packed(list(structs), list(params)), ref)


@pytest.mark.parametrize("mode", ["layer", "simss"])
def test_replay_builds_langevin_devices(emu, mode):
    """Langevin's rate in every layer (reference: src/jaxdd/device.py:900-911), at each layer's
    mobilities or SIMsalabim's (the interface's at a layer's last point): the builder's lanes
    give the host builder's packed arrays (1e-14 of each field's scale), and their tangents in
    the rate's numbers the host builder's central differences (relative step 1e-6; 1e-8 of
    each field's scale per unit of ln x)."""
    from conftest import LANGEVIN_WRT, bumped, langevin_rows, number_of
    from jaxdd_amd import adjoint as A
    from jaxdd_amd.batch import pack
    from jaxdd_amd.device import device_arrays

    inputs = langevin_rows(4, 80, mode)
    structs, params = zip(*inputs)
    ref = pack([device_arrays(st, p) for st, p in inputs])
    dirs = A.directions_for(LANGEVIN_WRT, len(inputs), len(structs[0].i0))
    got, tan = emu.build_packed(list(structs), list(params), dirs)
    _same_packed(got, ref)
    # the rate (SEG_KD) at every interior point
    kd = np.take_along_axis(ref["seg"][:, 5, :], ref["sid"].astype(np.int64), axis=0)
    assert np.all(kd[1:-1] > 0)
    for key in ("pt", "seg", "sc"):
        axis = tuple(i for i in range(ref[key].ndim) if i != ref[key].ndim - 2)
        scale = np.max(np.abs(ref[key]), axis=axis, keepdims=True)[..., 0]
        for k, w in enumerate(LANGEVIN_WRT):
            for d, (st, p) in enumerate(inputs):
                x = number_of(p, w)
                h = 1e-6 * abs(x)
                up, dn = (pack([device_arrays(st, bumped(p, w, s * h))])[key][..., 0]
                          for s in (1.0, -1.0))
                fd = (up - dn) / (2 * h)
                g = tan[key][k][..., d][:fd.shape[0]]
                err = np.abs(g - fd) * abs(x)
                assert np.all(err <= 1e-8 * scale[:fd.shape[0]] + 1e-300), (key, w, d)


def test_replay_builds_the_simss_runs_devices(emu):
    """The ten SimSS runs' devices in one batch (one with charged bulk traps, their type a
    segment field), nine generating by the profile file
    (reference: src/jaxdd/device.py:804-814 `_generation`, its axis stretched to the stack),
    one by each layer's G_ehp: the builder's lanes give the host builder's packed arrays (1e-14
    of each field's scale), the profile's G at every point to 1e-14 of itself, and the
    tangents in the thicknesses and G_frac the host builder's central differences (relative
    step 1e-6; 1e-8 of each field's scale per unit of ln x). G has a kink at each of the
    profile's nodes, where grid points often lie: there the reference's derivative is the
    slope of the interval the point belongs to (the one after the node), a one-sided
    difference (second order: steps h and 2h), which G's tangent, and that of the current scale
    (a sum over G), match instead. G_frac 0 (a dark run) takes the absolute step 1e-6."""
    from conftest import PROFILE_WRT, SIMSS_RUNS, bumped, fixture_inputs, number_of
    from jaxdd_amd import adjoint as A
    from jaxdd_amd.batch import SC_FIELDS, pack
    from jaxdd_amd.device import device_arrays

    inputs = [fixture_inputs(c) for c in SIMSS_RUNS]
    structs, params = zip(*inputs)
    assert sum(st.gen_mode == "file" for st in structs) == 9
    ref = pack([device_arrays(st, p) for st, p in inputs])
    dirs = A.directions_for(PROFILE_WRT, len(inputs), 3)
    got, tan = emu.build_packed(list(structs), list(params), dirs)
    _same_packed(got, ref)
    G, Gr = got["pt"][:, 2, :], ref["pt"][:, 2, :]
    assert np.all(np.abs(G - Gr) <= 1e-14 * np.abs(Gr))
    assert np.all(G[1:-1, SIMSS_RUNS.index("logv_light")] >= 0)
    for key in ("pt", "seg", "sc"):
        axis = tuple(i for i in range(ref[key].ndim) if i != ref[key].ndim - 2)
        scale = np.max(np.abs(ref[key]), axis=axis, keepdims=True)[..., 0]
        for k, w in enumerate(PROFILE_WRT):
            for d, (st, p) in enumerate(inputs):
                x = number_of(p, w)
                unit = abs(x) if x else 1.0
                h = 1e-6 * unit
                f = {s: pack([device_arrays(st, bumped(p, w, s * h))])[key][..., 0]
                     for s in (2.0, 1.0, -1.0, -2.0)}
                fd = (f[1.0] - f[-1.0]) / (2 * h)
                g = tan[key][k][..., d][:fd.shape[0]]
                err = np.abs(g - fd) * unit
                if key != "seg":
                    f0 = ref[key][..., d][:fd.shape[0]]
                    right = (4 * f[1.0] - f[2.0] - 3 * f0) / (2 * h)
                    left = (3 * f0 - 4 * f[-1.0] + f[-2.0]) / (2 * h)
                    one_sided = np.minimum(np.abs(g - right), np.abs(g - left)) * unit
                    kinked = (np.s_[:, 2] if key == "pt" else SC_FIELDS.index("J_scale"))
                    err[kinked] = np.minimum(err[kinked], one_sided[kinked])
                assert np.all(err <= 1e-8 * scale[:fd.shape[0]] + 1e-300), (key, w, SIMSS_RUNS[d])


def test_replay_tangents_are_the_host_builders_differences(emu):
    """The forward-mode tangents of the 16 parameters against `adjoint.tangents`' central
    differences (relative step 1e-6) of the host builder on each device's structure, in the
    units the differences resolve: d(field)/d ln(theta_k) (d/d W_L for the work function's
    offset) to 1e-8 of the field's own scale over the batch. A field that hardly depends on a
    parameter (the current scale on a contact layer's thickness: d ln J/d ln L ~ 1e-7) is known
    to the differences only to their round-off, 1e-10 of the field."""
    from jaxdd_amd import adjoint as A
    from jaxdd_amd import synthetic as S

    th = A.theta_from_table(S.draw(5), range(5))
    structs = [A.structure_of(t, 100) for t in th]
    out, tan = emu.build_packed(structs, [A.params_of(t) for t in th], A.directions_of(th))
    base, ref = A.tangents(th, NP=100)
    unit = np.abs(th)
    unit[:, 15] = 1.0
    for key in ("pt", "seg", "sc"):
        g, r, v = tan[key], ref[key], base[key]
        axis = tuple(i for i in range(v.ndim) if i != v.ndim - 2)
        scale = np.max(np.abs(v), axis=axis, keepdims=True)  # each field's own scale
        for k in range(len(A.THETA)):
            err = np.abs(g[k] - r[k]) * unit[:, k]
            assert np.all(err <= 1e-8 * scale + 1e-300), (key, A.THETA[k])
            # and the tangents are not all zero where the differences are not
            assert np.all((np.abs(r[k]) * unit[:, k] > 1e-6 * scale) <= (g[k] != 0)), (key, A.THETA[k])


def test_replay_layout_is_the_host_graded_layout(emu):
    """The GPU layout's lanes on the CPU against the host's graded rule (`grid.graded_layout`)
    for 400 drawn stacks: every layer's point count and first point, the segments and the
    interfaces exactly; each layer's spacing sum to 1e-15; the spacings to 1e-10 of themselves
    (each point stops its Newton steps on its own, the host a layer's points together: the
    positions agree to their last bits, which the smallest spacings magnify)."""
    from jaxdd_amd import adjoint as A
    from jaxdd_amd import builder as B
    from jaxdd_amd import synthetic as S
    from jaxdd_amd.grid import GRADED_POINTS_PER_EFOLD, GRADED_SMALLEST

    n = 400
    th = A.theta_from_table(S.draw(n), range(n))
    out = emu.graded_layout_packed(th[:, 0:3], 400, GRADED_POINTS_PER_EFOLD, GRADED_SMALLEST)
    for k in range(n):
        st = A.structure_of(th[k])
        assert tuple(out["i0"][:, k]) == tuple(st.i0)
        assert np.array_equal(out["lid"][:, k], st.lid)
        assert np.array_equal(out["sid"][:, k], B.segment_ids(st))
        assert out["nseg"][k] == B.segment_ids(st)[-1] + 1
        iface = np.zeros(401, np.uint8)
        iface[list(st.i1[:-1])] = 1
        assert np.array_equal(out["iface"][:, k], iface)
        assert np.max(np.abs(out["unit_sum"][:, k] / st.norm - 1)) < 1e-15
        assert np.max(np.abs(out["unit"][:, k] / st.raw - 1)) < 1e-10


def test_vectorised_numbers_are_the_devices_numbers():
    """`adjoint.numbers_of` (all devices at once) is `builder.number_arrays` of `params_of`
    (one device at a time), exactly."""
    from jaxdd_amd import adjoint as A
    from jaxdd_amd import builder as B
    from jaxdd_amd import synthetic as S

    th = A.theta_from_table(S.draw(7), range(7))
    th[:, 15] = np.linspace(-0.02, 0.03, 7)
    structs = [A.structure_of(t, 60) for t in th]
    got = A.numbers_of(th)
    ref = B.number_arrays(structs, [A.params_of(t) for t in th])
    assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])
    flags = B.structure_arrays(structs)["flags"]
    assert np.array_equal(flags, np.broadcast_to(np.array(A.STACK_FLAGS, np.uint8)[:, None], flags.shape))


def test_the_builder_refuses_what_it_does_not_build():
    """Computed generation, flat-band contacts, optional modules: before any launch."""
    import dataclasses

    from jaxdd_amd import builder as B

    st, _ = _rows(1, 40)[0]
    B.check_buildable(st)
    for change, match in ((dict(gen_mode="calc"), "generation"), (dict(sfb=(True, False)), "flat-band"),
                          (dict(extra={"langevin": True}), "optional")):
        with pytest.raises(NotImplementedError, match=match):
            B.check_buildable(dataclasses.replace(st, **change))


@pytest.mark.gpu
def test_gpu_build_is_the_host_batch_and_its_tangents(torch_cuda, emu):
    """ddh_build through the C-ABI: the packed arrays of 300 gradient-workload devices (more
    than a wavefront of lanes per direction) against the host builder's to 1e-14, and the
    tangents against the CPU replay's to 1e-13."""
    from jaxdd_amd import adjoint as A
    from jaxdd_amd import synthetic as S
    from jaxdd_amd.batch import pack

    n = 300
    th = A.theta_from_table(S.draw(n), range(n))
    print(f"ddh_build: {n} devices x 16 directions", flush=True)
    structs = [A.structure_of(t, 120) for t in th]
    params = [A.params_of(t) for t in th]
    batch, tan = A.gpu_tangents(th, NP=120, structs=structs)
    from jaxdd_amd.device import device_arrays

    ref = pack([device_arrays(st, p) for st, p in zip(structs, params)])
    got = {"pt": _np(batch["_pt"]), "seg": _np(batch["_seg"]), "sc": _np(batch["_sc"]),
           "imu": _np(batch["_imu"]), "sid": _np(batch["_sid"]), "nseg": batch["nseg"]}
    _same_packed(got, ref)
    _, rtan = emu.build_packed(structs, params, A.directions_of(th))
    for key in ("pt", "seg", "sc"):
        g, r = _np(tan[key]), rtan[key]
        scale = np.max(np.abs(r)) + 1e-300
        assert np.max(np.abs(g - r)) <= 1e-13 * scale, key


@pytest.mark.gpu
def test_gpu_gradients_from_built_tangents(torch_cuda):
    """The gradients of 64 gradient-workload devices with the tangents built on the GPU against
    those with the host builder's central differences: the same curves and flags, and
    dJ/dln(theta_k) to 1e-5 of each curve's current scale, the gradient tests' bar: the
    differences' tangents are 1e-10 off in places, and near open circuit the adjoint's two
    terms nearly cancel and lift that to 3e-6 on one device (1e-7 on most)."""
    from jaxdd_amd import adjoint as A
    from jaxdd_amd import synthetic as S
    from jaxdd_amd.batch import j_scale

    n, NP = 64, 120
    th = A.theta_from_table(S.draw(n), range(n))
    V = S.VOLTAGES[::4].copy()
    print(f"gradients from built tangents: {n} devices, NP = {NP}", flush=True)
    batch, tan = A.gpu_tangents(th, NP=NP)
    J1, ok1, _, g1 = A.jv_gradients(batch, tan, V, tol=1e-12, max_iter=200)
    packed, htan = A.tangents(th, NP=NP)
    J2, ok2, _, g2 = A.jv_gradients(packed, htan, V, tol=1e-12, max_iter=200)
    J1, J2, g1, g2 = _np(J1), _np(J2), _np(g1), _np(g2)
    assert np.array_equal(_np(ok1), _np(ok2)) and _np(ok1).mean() > 0.95
    devs = [A.device_from_theta(t, NP) for t in th]
    cscale = np.array([j_scale(d) for d in devs]) + np.max(np.abs(J2), axis=1)
    assert np.max(np.abs(J1 - J2) / cscale[:, None]) < 1e-10
    unit = np.abs(th)
    unit[:, 15] = 1.0
    err = np.abs(g1 - g2) * unit[:, None, :] / cscale[:, None, None]
    assert np.max(err) < 1e-5, np.unravel_index(np.argmax(err), err.shape)
    assert np.median(err) < 1e-9


@pytest.mark.gpu
def test_gpu_graded_build_is_the_host_batch(torch_cuda, emu):
    """The whole build on the GPU (ddh_graded_layout, then ddh_build) for 300 gradient-workload
    devices against the host builder's packed arrays (to 1e-13 of each field's scale: the
    layouts' last bits) and its tangents against the host-structure build's (1e-12)."""
    from jaxdd_amd import adjoint as A
    from jaxdd_amd import synthetic as S
    from jaxdd_amd.batch import pack
    from jaxdd_amd.device import device_arrays

    n = 300
    th = A.theta_from_table(S.draw(n), range(n))
    print(f"graded build on the GPU: {n} devices x 16 directions", flush=True)
    batch, tan = A.gpu_tangents(th, NP=120)
    structs = [A.structure_of(t, 120) for t in th]
    ref = pack([device_arrays(st, A.params_of(t)) for st, t in zip(structs, th)])
    got = {"pt": _np(batch["_pt"]), "seg": _np(batch["_seg"]), "sc": _np(batch["_sc"]),
           "imu": _np(batch["_imu"]), "sid": _np(batch["_sid"]), "nseg": batch["nseg"]}
    _same_packed(got, ref, rtol=1e-13)
    assert np.array_equal(_np(batch["_iface"]), ref["iface"])
    _, htan = A.gpu_tangents(th, NP=120, structs=structs)
    for key in ("pt", "seg", "sc"):
        g, r = _np(tan[key]), _np(htan[key])
        assert np.max(np.abs(g - r)) <= 1e-12 * np.max(np.abs(r)), key


def test_directions_in_named_numbers_are_the_parameters_directions():
    """`adjoint.directions_for` names the same directions as the 16 parameters where they are
    single numbers: the perovskite's electron mobility, the ETL's thickness, the left work
    function, the HTL's acceptors."""
    from jaxdd_amd import adjoint as A
    from jaxdd_amd import synthetic as S

    th = A.theta_from_table(S.draw(3), range(3))
    dl, ds = A.directions_for([("mu_n", 1), ("L", 0), "W_L", ("N_A", 2)], 3, 3)
    tl, ts = A.directions_of(th)
    for k, j in enumerate((5, 0, 15, 11)):
        assert np.array_equal(dl[k], tl[j]) and np.array_equal(ds[k], ts[j])
    with pytest.raises(ValueError, match="per-layer"):
        A.directions_for([("mu_x", 0)], 3, 3)
    with pytest.raises(ValueError, match="per-device"):
        A.directions_for(["nope"], 3, 3)


@pytest.mark.gpu
def test_gpu_named_gradients_and_the_vjp(torch_cuda):
    """`param_gradients` in named numbers equals the parameters' gradient in the same
    directions (the same build, solve and adjoint), and ddh_adjoint_vjp's contraction on the
    device equals the dense gradient contracted on the host (1e-12)."""
    from jaxdd_amd import adjoint as A
    from jaxdd_amd import synthetic as S

    n, NP = 40, 100
    th = A.theta_from_table(S.draw(n), range(n))
    V = S.VOLTAGES[::5].copy()
    structs = [A.structure_of(t, NP) for t in th]
    print(f"named gradients and the VJP: {n} devices", flush=True)
    J, ok, _, g = A.param_gradients(structs, [A.params_of(t) for t in th],
                                    [("mu_n", 1), "W_L"], V, tol=1e-12, max_iter=200)
    batch, tan = A.gpu_tangents(th, NP=NP, structs=structs)
    J2, ok2, _, g2 = A.jv_gradients(batch, tan, V, tol=1e-12, max_iter=200)
    g, g2 = _np(g), _np(g2)
    assert np.array_equal(_np(ok), _np(ok2))
    np.testing.assert_allclose(g[..., 0], g2[..., 5], rtol=1e-12, atol=0)
    np.testing.assert_allclose(g[..., 1], g2[..., 15], rtol=1e-12, atol=0)
    w = np.random.default_rng(3).normal(size=(n, len(V)))
    tbar, J3, ok3 = A.jv_vjp(batch, tan, V, w, tol=1e-12, max_iter=200)
    expect = np.einsum("nv,nvk->nk", w, g2)
    np.testing.assert_allclose(_np(tbar), expect, rtol=1e-12, atol=1e-12 * np.abs(expect).max())


@pytest.mark.gpu
def test_gpu_gradients_with_mobile_ions(torch_cuda, oracle, emu):
    """Free mobile ions on the GPU (the bordered transposed solve): gradients in named numbers,
    the ions' own densities, the temperature and the generation's factor among them, against
    the adjoint's CPU replay at the GPU's own states and against central differences of the
    oracle's curves. As the ion-free gradient's tests: the replay on a random weighting of the
    curves and directions (1e-6), since at high forward bias the transposed system's
    conditioning turns the GPU's and the host's last-bit differences of exp and log into
    differences of single components; the differences per component (1e-5 of the curve's
    scale, as the CPU replay's test)."""
    from conftest import ION_WRT, named_fd_errors, number_of
    from jaxdd_amd import adjoint as A
    from jaxdd_amd import synthetic as S
    from jaxdd_amd.builder import build

    n, NP = 8, 100
    t = S.draw(n)
    structs, params = zip(*[S.row_inputs(t, k, NP, ions=True) for k in range(n)])
    V = S.VOLTAGES[::10].copy()
    print(f"gradients with mobile ions: {n} devices, {len(ION_WRT)} numbers", flush=True)
    dirs = A.directions_for(ION_WRT, n, len(structs[0].i0))
    batch, tan = build(list(structs), list(params), dirs)
    J, ok, _, g, states, _ = A.jv_gradients(batch, tan, V, tol=1e-12, max_iter=200, strict=True,
                                            return_states=True)
    J, ok, g = _np(J), _np(ok), _np(g)
    assert ok.all() and np.isfinite(g).all()
    packed = {"pt": _np(batch["_pt"]), "sid": _np(batch["_sid"]), "seg": _np(batch["_seg"]),
              "sc": _np(batch["_sc"]), "nseg": batch["nseg"]}
    rtan = {k: _np(tan[k]) for k in ("pt", "seg", "sc")}
    rg, rst = emu.adjoint_packed(packed, V, np.ascontiguousarray(_np(states)), ok, rtan)
    assert not rst.any()
    unit = np.array([[abs(number_of(p, w)) for w in ION_WRT] for p in params])
    rng = np.random.default_rng(0)
    w, d = rng.normal(size=ok.shape), rng.normal(size=unit.shape)
    a = np.sum(w[:, :, None] * g * (unit * d)[:, None, :])
    b = np.sum(w[:, :, None] * rg * (unit * d)[:, None, :])
    scale = np.max(np.abs(J), axis=1)[:, None, None]
    comp = np.abs(g - rg) * unit[:, None, :] / scale
    print(f"GPU vs replay: weighted sum {abs(a - b) / abs(b):.2e}, components max "
          f"{comp.max():.2e} median {np.median(comp):.2e}", flush=True)
    assert abs(a - b) < 1e-6 * abs(b)
    err = named_fd_errors(oracle, structs[:3], params[:3], ION_WRT, V, J, g)
    print(f"GPU vs differences: max {err.max():.2e} median {np.median(err):.2e}", flush=True)
    assert np.max(err) < 1e-5, err
    assert np.median(err) < 3e-7, err


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["layer", "simss"])
def test_gpu_langevin(torch_cuda, oracle, emu, mode):
    """Langevin's rate on the GPU: the SimSS run `langevin_light` through `solve_jv` (within
    1e-9 of its scale at SIMsalabim's mobilities, the reference's bar; the oracle's steps and
    flags, 1e-12) and through `solve_batch` (the oracle's flags and steps, 1e-12); drawn stacks
    built by ddh_build with the rate in every layer against the host builder, and their
    gradients in the rate's numbers (`param_gradients`) against central differences of the
    oracle's curves (1e-5 of each curve's scale)."""
    from conftest import GOLDEN, LANGEVIN_WRT, fixture_jv, langevin_rows, named_fd_errors
    from jaxdd_amd import adjoint as A
    from jaxdd_amd.batch import pack, prepare, solve_batch
    from jaxdd_amd.builder import build
    from jaxdd_amd.device import build_device, device_arrays
    from jaxdd_amd.simss_input import read_setup
    from jaxdd_amd.solver import solve_jv

    print(f"Langevin ({mode}): the SimSS run on the GPU", flush=True)
    dev = build_device(read_setup(GOLDEN / "langevin_light" / "simulation_setup.txt"),
                       grid="simss", langevin=mode)
    ref = fixture_jv("langevin_light")
    scale = np.max(np.abs(ref["Jext"]))
    res = solve_jv(dev, ref["Vint"], bernoulli_name="simss")
    Jo, oko, itso = oracle.solve_jv(dev, ref["Vint"], bernoulli_name="simss")
    assert res.converged.all() and np.array_equal(res.iterations, itso)
    assert np.max(np.abs(res.J - Jo)) < 1e-12 * scale
    if mode == "simss":
        assert np.max(np.abs(res.J - ref["Jext"])) < 1e-9 * scale
    J, ok, its = (_np(x) for x in solve_batch(prepare([dev]), V_FULL, bernoulli_name="simss"))
    Jb, okb, itsb = oracle.solve_batch([dev], V_FULL, bernoulli_name="simss", first_max_iter=None)
    assert np.array_equal(ok, okb) and np.array_equal(its, itsb)
    assert np.max(np.abs(J - Jb)) < 1e-12 * scale

    inputs = langevin_rows(8, 100, mode)
    structs, params = zip(*inputs)
    print(f"Langevin ({mode}): 8 built devices, {len(LANGEVIN_WRT)} numbers", flush=True)
    batch = build(list(structs), list(params))
    want = pack([device_arrays(st, p) for st, p in inputs])
    got = {"pt": _np(batch["_pt"]), "seg": _np(batch["_seg"]), "sc": _np(batch["_sc"]),
           "imu": _np(batch["_imu"]), "sid": _np(batch["_sid"]), "nseg": batch["nseg"]}
    _same_packed(got, want, rtol=1e-13)
    V = V_FULL[::10].copy()
    J, ok, _, g = A.param_gradients(structs, params, LANGEVIN_WRT, V, tol=1e-12, max_iter=200)
    J, ok, g = _np(J), _np(ok), _np(g)
    assert ok.all() and np.isfinite(g).all()
    err = named_fd_errors(oracle, structs[:3], params[:3], LANGEVIN_WRT, V, J, g)
    print(f"GPU vs differences: max {err.max():.2e} median {np.median(err):.2e}", flush=True)
    assert np.max(err) < 1e-5, err
    assert np.median(err) < 3e-7, err


@pytest.mark.gpu
def test_gpu_simss_runs_built_with_their_profiles(torch_cuda, oracle, emu):
    """ddh_build of the ten SimSS runs' devices in one batch (nine generating by the profile
    file, one by each layer's G_ehp, one with charged bulk traps): the packed arrays against the host builder's to 1e-14 and
    the tangents in the thicknesses and G_frac against the CPU replay's to 1e-13; then
    `param_gradients` of three of them against central differences of the oracle's curves
    (relative step 1e-5), to 1e-7 of each curve's scale (measured: 3e-9 at most, 3e-10 at the
    median). The curves' derivatives in the thicknesses have kinks where a grid point crosses
    one of the profile's nodes (several points lie on nodes), too small in the curves to show
    at these bounds."""
    from conftest import PROFILE_WRT, SIMSS_RUNS, fixture_inputs, named_fd_errors
    from jaxdd_amd import adjoint as A
    from jaxdd_amd.batch import pack
    from jaxdd_amd.builder import build
    from jaxdd_amd.device import device_arrays

    inputs = [fixture_inputs(c) for c in SIMSS_RUNS]
    structs, params = zip(*inputs)
    print(f"ddh_build: the {len(inputs)} SimSS runs' devices, {len(PROFILE_WRT)} directions",
          flush=True)
    dirs = A.directions_for(PROFILE_WRT, len(inputs), 3)
    batch, tan = build(list(structs), list(params), dirs)
    got = {"pt": _np(batch["_pt"]), "seg": _np(batch["_seg"]), "sc": _np(batch["_sc"]),
           "imu": _np(batch["_imu"]), "sid": _np(batch["_sid"]), "nseg": batch["nseg"]}
    _same_packed(got, pack([device_arrays(st, p) for st, p in inputs]))
    _, rtan = emu.build_packed(list(structs), list(params), dirs)
    for key in ("pt", "seg", "sc"):
        g, r = _np(tan[key]), rtan[key]
        assert np.max(np.abs(g - r)) <= 1e-13 * (np.max(np.abs(r)) + 1e-300), key

    runs = ("held_thin_light", "steady_1e17_thin_light", "bulk_acceptor_light")
    sub = [inputs[SIMSS_RUNS.index(c)] for c in runs]
    structs, params = zip(*sub)
    V = V_FULL[::10].copy()
    print(f"param_gradients: {runs}", flush=True)
    J, ok, _, g = A.param_gradients(structs, params, PROFILE_WRT, V, tol=1e-12, max_iter=200)
    J, ok, g = _np(J), _np(ok), _np(g)
    assert ok.all() and np.isfinite(g).all()
    err = named_fd_errors(oracle, structs, params, PROFILE_WRT, V, J, g)
    print(f"GPU vs differences: max {err.max():.2e} median {np.median(err):.2e}\n{err}",
          flush=True)
    assert np.max(err) < 1e-7, err
    assert np.median(err) < 1e-8, err


@pytest.mark.gpu
def test_gpu_gradient_in_the_voltage(torch_cuda):
    """`param_gradients` with "V": each voltage's dJ/dV of devices with free mobile ions (the
    shunt's G_shunt included) against central differences of the GPU's own curves in the
    voltage (h = 1e-5 V), to 1e-5 of each curve's largest slope."""
    from jaxdd_amd import adjoint as A
    from jaxdd_amd import synthetic as S
    from jaxdd_amd.batch import solve_batch
    from jaxdd_amd.builder import build

    n, NP = 16, 100
    t = S.draw(n)
    structs, params = zip(*[S.row_inputs(t, k, NP, ions=True) for k in range(n)])
    V = S.VOLTAGES[::10].copy()
    print(f"the gradient in the voltage: {n} devices", flush=True)
    J, ok, _, g = A.param_gradients(structs, params, ["V", ("mu_n", 1)], V, tol=1e-12,
                                    max_iter=200)
    batch = build(list(structs), list(params))
    h = 1e-5
    Jp, okp, _ = solve_batch(batch, V + h, tol=1e-12, max_iter=200, strict=True)
    Jm, okm, _ = solve_batch(batch, V - h, tol=1e-12, max_iter=200, strict=True)
    assert _np(ok).all() and _np(okp).all() and _np(okm).all()
    fd = (_np(Jp) - _np(Jm)) / (2 * h)
    err = np.abs(_np(g)[:, :, 0] - fd) / np.max(np.abs(fd), axis=1, keepdims=True)
    print(f"dJ/dV against differences: max {err.max():.2e} median {np.median(err):.2e}", flush=True)
    assert err.max() < 1e-5 and np.median(err) < 1e-6


@pytest.mark.gpu
def test_gpu_gradients_with_frozen_ions(torch_cuda, oracle, emu):
    """Ions frozen at a 1.2 V pre-bias on the GPU: the solve keeps the pre-bias solutions
    (ddh_solve_batch_states' pre_states) and the adjoint adds its second pass at them. Against
    the CPU replay at the GPU's own states (a random weighting of the components, 1e-6) and
    against central differences of the oracle's curves (the CPU replay's bars: 1e-5 of the
    curve's scale below 1.3 V, 1e-4 at 1.3 V)."""
    from conftest import ION_WRT, named_fd_errors, number_of
    from jaxdd_amd import adjoint as A
    from jaxdd_amd import synthetic as S
    from jaxdd_amd.builder import build

    wrt = ION_WRT + ["V_pre"]
    n, NP = 8, 100
    t = S.draw(n)
    structs, params = zip(*[S.row_inputs(t, k, NP, ions=True, frozen=True) for k in range(n)])
    V = S.VOLTAGES[::10].copy()
    print(f"gradients with frozen ions: {n} devices, {len(wrt)} numbers", flush=True)
    batch, tan = build(list(structs), list(params), A.directions_for(wrt, n, len(structs[0].i0)))
    J, ok, _, g, states, pre = A.jv_gradients(batch, tan, V, tol=1e-12, max_iter=200,
                                              strict=True, return_states=True)
    J, ok, g = _np(J), _np(ok), _np(g)
    assert ok.all() and np.isfinite(g).all() and pre is not None
    packed = {"pt": _np(batch["_pt"]), "sid": _np(batch["_sid"]), "seg": _np(batch["_seg"]),
              "sc": _np(batch["_sc"]), "nseg": batch["nseg"]}
    rtan = {k: _np(tan[k]) for k in ("pt", "seg", "sc")}
    rg, rst = emu.adjoint_packed(packed, V, np.ascontiguousarray(_np(states)), ok, rtan,
                                 pre=np.ascontiguousarray(_np(pre)))
    assert not rst.any()
    unit = np.array([[abs(number_of(p, w)) for w in wrt] for p in params])
    rng = np.random.default_rng(1)
    w, d = rng.normal(size=ok.shape), rng.normal(size=unit.shape)
    a = np.sum(w[:, :, None] * g * (unit * d)[:, None, :])
    b = np.sum(w[:, :, None] * rg * (unit * d)[:, None, :])
    print(f"GPU vs replay: weighted sum {abs(a - b) / abs(b):.2e}", flush=True)
    assert abs(a - b) < 1e-6 * abs(b)
    err = named_fd_errors(oracle, structs[:2], params[:2], wrt, V, J, g, per_voltage=True)
    low = V < 1.25
    print(f"GPU vs differences: below 1.3 V max {err[:, :, low].max():.2e}, at 1.3 V max "
          f"{err[:, :, ~low].max():.2e}, median {np.median(err):.2e}", flush=True)
    assert np.max(err[:, :, low]) < 1e-5 and np.max(err) < 1e-4 and np.median(err) < 1e-7

