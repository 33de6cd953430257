"""Gradients of J-V curves: dJ(V)/dtheta by the implicit function theorem (the reference's
`_converged_jvp`, src/jaxdd/batch.py:650-737, taken through `jax.grad` over
`device_tree(structure, params)`, src/jaxdd/batch.py:142-148).

The parameters are those of the gradient workload, 16 per device (`THETA`): the three layer
thicknesses, the electron and hole mobility of each layer, the ETL donor and the perovskite's
and HTL's acceptor densities, the perovskite's bulk trap density, its SRH lifetime
tau = 1 / (N_t C) (through the capture coefficient C_n = C_p = C), its generation rate and an
offset of the left work function. They reach the adjoint as the tangents of the packed batch
arrays, d(array)/d(theta_k), built on the GPU with the batch itself (`gpu_tangents`,
ddh_build's forward mode; `tangents` keeps the host builder's central differences as their
check): as the reference's `device_tree(structure, params)`, the grid's layout (points per
layer, their unscaled spacing) is the structure's, fixed at the device's parameters, and the
thicknesses move the points only by scaling each layer's spacing (src/jaxdd/device.py:225-233
`_spacing`).
The HIP adjoint kernel (`ddh_adjoint`) then linearises each converged voltage's residual and
current in those directions and solves one transposed block-tridiagonal system per voltage;
with free mobile ions that system is bordered by each species' normalisation and number
constraint (src/jaxdd/batch.py:595-618 `_transposed`), and `param_gradients` takes any of the
builder's named numbers, the ions' densities and the temperature among them, and the applied
voltage. Ions frozen at a pre-bias are differentiated through the pre-bias solution (a second
transposed bordered solve at it; the solve keeps that solution).

`curves(theta, V)` is the differentiable entry point: a `torch.autograd.Function` whose
forward solves the curves and whose backward contracts the incoming dL/dJ with dJ/dtheta.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from . import synthetic as S
from .batch import SC_FIELDS, _sc_now, check_four_lanes, from_packed, pack

THETA = ("L_etl", "L_pvk", "L_htl",
         "mu_n_etl", "mu_p_etl", "mu_n_pvk", "mu_p_pvk", "mu_n_htl", "mu_p_htl",
         "N_D_etl", "N_A_pvk", "N_A_htl", "N_t", "tau", "G", "dW_L")
#: The perovskite traps' capture coefficient at the drawn devices (BASE C_n_bulk, C_p_bulk).
C_TRAP = 1e-14
#: Relative step of the builder's central differences (an absolute step in eV for dW_L).
REL_STEP = 1e-6


def theta_from_table(t: dict, rows) -> np.ndarray:
    """theta (len(rows), 16) of drawn devices (synthetic.draw): tau = 1 / (N_t C_TRAP), dW_L 0."""
    rows = np.asarray(rows)
    th = np.zeros((len(rows), len(THETA)))
    th[:, 0], th[:, 1], th[:, 2] = t["L_etl"][rows], t["L_pvk"][rows], t["L_htl"][rows]
    th[:, 3:9] = t["mu"][rows]
    th[:, 9:12] = t["dop"][rows]
    th[:, 12] = t["N_t"][rows]
    th[:, 13] = 1.0 / (t["N_t"][rows] * C_TRAP)
    th[:, 14] = t["G"][rows]
    return th


def structure_of(theta, NP: int = 400):
    """The structure of the device of ``theta``: the drawn workload's stack on the graded
    layout of its own thicknesses (reference: src/jaxdd/device.py:385-438 `Structure`)."""
    return S._structure(np.asarray(theta, dtype=float)[0:3].copy(), NP)


def params_of(theta) -> dict:
    """The numbers (`device_parameters`' dict) of one parameter vector: the drawn workload's
    stack, no ions."""
    th = np.asarray(theta, dtype=float)
    L = th[0:3].copy()
    mu = th[3:9]
    C = 1.0 / (th[12] * th[13])
    p = dict(S.BASE, **S.SCALARS)
    p.update(
        L=L, mu_n=mu[0::2].copy(), mu_p=mu[1::2].copy(),
        N_D=np.array([th[9], 0.0, 0.0]), N_A=np.array([0.0, th[10], th[11]]),
        N_t_bulk=np.array([0.0, th[12], 0.0]), G_ehp=np.array([0.0, th[14], 0.0]),
        C_n_bulk=np.array([C_TRAP, C, C_TRAP]), C_p_bulk=np.array([C_TRAP, C, C_TRAP]),
        W_L=S.SCALARS["W_L"] + th[15],
    )
    return p


def device_from_theta(theta, NP: int = 400, structure=None):
    """The device of one parameter vector (the drawn workload's stack, no ions) on
    ``structure`` (default: `structure_of(theta)`, the layout of its own thicknesses)."""
    st = structure_of(theta, NP) if structure is None else structure
    return S.device_arrays(st, params_of(theta))


def directions_of(thetas) -> tuple[np.ndarray, np.ndarray]:
    """The 16 parameters as directions in the GPU builder's numbers (`builder.LAYER_FIELDS`,
    `SCALAR_FIELDS`): dlayers (16, 3, 23, n), dscalars (16, 10, n). Each is its number's unit
    direction, but for the trap density and the lifetime, which also move the perovskite's
    capture coefficients C = 1/(N_t tau)."""
    from .builder import LAYER_FIELDS, SCALAR_FIELDS

    th = np.atleast_2d(np.asarray(thetas, dtype=float))
    n = len(th)
    F = {f: i for i, f in enumerate(LAYER_FIELDS)}
    dl = np.zeros((len(THETA), 3, len(LAYER_FIELDS), n))
    ds = np.zeros((len(THETA), len(SCALAR_FIELDS), n))
    for k in range(3):
        dl[k, k, F["L"]] = 1.0
    for q in range(6):
        dl[3 + q, q // 2, F["mu_n" if q % 2 == 0 else "mu_p"]] = 1.0
    dl[9, 0, F["N_D"]] = 1.0
    dl[10, 1, F["N_A"]] = 1.0
    dl[11, 2, F["N_A"]] = 1.0
    C = 1.0 / (th[:, 12] * th[:, 13])
    dl[12, 1, F["N_t_bulk"]] = 1.0
    for f in ("C_n_bulk", "C_p_bulk"):
        dl[12, 1, F[f]] = -C / th[:, 12]
        dl[13, 1, F[f]] = -C / th[:, 13]
    dl[14, 1, F["G_ehp"]] = 1.0
    ds[15, SCALAR_FIELDS.index("W_L")] = 1.0
    return dl, ds


def _structure_rows(args):
    thetas, NP = args
    return [structure_of(t, NP) for t in thetas]


def structures(thetas, NP: int = 400, workers: int = 1, chunk: int = 64):
    """The structures of the devices of ``thetas`` (n, 16): each on the graded layout of its
    own thicknesses. ``workers`` > 1 forks processes (call it before the process touches the
    GPU)."""
    th = np.atleast_2d(np.asarray(thetas, dtype=float))
    jobs = [(th[s:s + chunk], NP) for s in range(0, len(th), chunk)]
    if workers <= 1:
        return [st for j in jobs for st in _structure_rows(j)]
    import multiprocessing as mp

    pool = mp.get_context("fork").Pool(workers)
    try:
        parts = pool.map(_structure_rows, jobs)
    finally:
        pool.close()  # let the workers exit on their own (no SIGTERM from terminate())
        pool.join()
    return [st for part in parts for st in part]


#: The stack's layers' flags for the GPU builder: bulk traps in the perovskite, no ions.
STACK_FLAGS = (0, 1, 0)


def numbers_of(thetas) -> tuple[np.ndarray, np.ndarray]:
    """The GPU builder's numbers (`builder.number_arrays`' layout: layers [3][23][n], scalars
    [10][n]) of parameter vectors (n, 16), vectorised: `params_of` for every device at once."""
    from .builder import LAYER_FIELDS, SCALAR_FIELDS

    th = np.atleast_2d(np.asarray(thetas, dtype=float))
    n = len(th)
    base = dict(S.BASE, **S.SCALARS)
    layers = np.zeros((3, len(LAYER_FIELDS), n))
    for i, f in enumerate(LAYER_FIELDS):
        if f in base:
            layers[:, i, :] = np.asarray(base[f], dtype=float)[:, None]
    F = {f: i for i, f in enumerate(LAYER_FIELDS)}
    layers[:, F["preLangevin"], :] = 1.0
    layers[:, F["L"], :] = th[:, 0:3].T
    layers[:, F["mu_n"], :] = th[:, 3:9:2].T
    layers[:, F["mu_p"], :] = th[:, 4:9:2].T
    layers[:, F["N_D"], :] = 0.0
    layers[0, F["N_D"], :] = th[:, 9]
    layers[:, F["N_A"], :] = 0.0
    layers[1, F["N_A"], :] = th[:, 10]
    layers[2, F["N_A"], :] = th[:, 11]
    layers[:, F["N_t_bulk"], :] = 0.0
    layers[1, F["N_t_bulk"], :] = th[:, 12]
    C = 1.0 / (th[:, 12] * th[:, 13])
    for f in ("C_n_bulk", "C_p_bulk"):
        layers[:, F[f], :] = C_TRAP
        layers[1, F[f], :] = C
    layers[:, F["G_ehp"], :] = 0.0
    layers[1, F["G_ehp"], :] = th[:, 14]
    scalars = np.zeros((len(SCALAR_FIELDS), n))
    for i, f in enumerate(SCALAR_FIELDS):
        if f in base:
            scalars[i] = float(base[f])
    scalars[SCALAR_FIELDS.index("W_L")] = S.SCALARS["W_L"] + th[:, 15]
    scalars[SCALAR_FIELDS.index("G_shunt")] = 0.0
    scalars[SCALAR_FIELDS.index("pre_cond")] = 0.0
    scalars[SCALAR_FIELDS.index("fix_ions")] = 0.0
    scalars[SCALAR_FIELDS.index("left_elec")] = -1.0
    return layers, scalars


def gpu_tangents(thetas, NP: int = 400, device="cuda", structs=None):
    """(batch, tangents) on the GPU: the devices of ``thetas`` (n, 16) built on the GPU with
    their derivatives in each parameter (forward mode, each device's structure held), in place
    of `tangents`' host central differences. By default everything is built there: each
    device's graded layout (ddh_graded_layout) and its arrays (ddh_build); ``structs``: the
    host's structures (`structures`) instead of the GPU's layouts."""
    from .builder import build, build_graded

    th = np.atleast_2d(np.asarray(thetas, dtype=float))
    if structs is not None:
        return build(structs, [params_of(t) for t in th], directions_of(th), device=device)
    layers, scalars = numbers_of(th)
    return build_graded(th[:, 0:3], STACK_FLAGS, layers, scalars, directions_of(th), NP=NP,
                        device=device)


def _steps(th: np.ndarray) -> np.ndarray:
    h = REL_STEP * np.abs(th)
    h[..., 15] = REL_STEP
    return h


def _tangent_rows(args):
    thetas, NP = args
    structs = [structure_of(t, NP) for t in thetas]
    base = pack([device_from_theta(t, NP, st) for t, st in zip(thetas, structs)])
    n, K = len(thetas), thetas.shape[1]
    out = {"pt": np.zeros((K,) + base["pt"].shape), "seg": np.zeros((K,) + base["seg"].shape),
           "sc": np.zeros((K,) + base["sc"].shape)}
    h = _steps(thetas)
    for k in range(K):
        e = np.zeros(K)
        e[k] = 1.0
        hi = pack([device_from_theta(t + hk * e, NP, st)
                   for t, hk, st in zip(thetas, h[:, k], structs)])
        lo = pack([device_from_theta(t - hk * e, NP, st)
                   for t, hk, st in zip(thetas, h[:, k], structs)])
        for side in (hi, lo):
            if side["nseg"] != base["nseg"] or not np.array_equal(side["sid"], base["sid"]):
                raise ValueError(f"the segments change within the step of {THETA[k]}")
        for key in ("pt", "seg", "sc"):
            out[key][k] = (hi[key] - lo[key]) / (2.0 * h[:, k])
    return base, out


def tangents(thetas, NP: int = 400, workers: int = 1, chunk: int = 64):
    """(packed batch, tangents): the packed arrays of the devices of ``thetas`` (n, 16) and
    their derivatives in each parameter, {"pt": (16, NP+2, 4, n), "seg": (16, nseg, 22, n),
    "sc": (16, 20, n)}, by central differences of the builder on each device's structure (its
    layout fixed). ``workers`` > 1 forks processes (call it before the process touches the
    GPU)."""
    from .batch import merge

    thetas = np.atleast_2d(np.asarray(thetas, dtype=float))
    jobs = [(thetas[s:s + chunk], NP) for s in range(0, len(thetas), chunk)]
    if workers <= 1:
        parts = [_tangent_rows(j) for j in jobs]
    else:
        import multiprocessing as mp

        pool = mp.get_context("fork").Pool(workers)
        try:
            parts = pool.map(_tangent_rows, jobs)
        finally:
            pool.close()  # let the workers exit on their own (no SIGTERM from terminate())
            pool.join()
    nseg = max(p[0]["nseg"] for p in parts)
    if any(p[0]["nseg"] != nseg for p in parts):
        raise ValueError("devices with different segment counts in one gradient batch")
    packed = merge([p[0] for p in parts])
    tan = {k: np.concatenate([p[1][k] for p in parts], axis=-1) for k in ("pt", "seg", "sc")}
    return packed, tan


def resident(packed, tan, device="cuda"):
    """The packed batch (`from_packed`) and the tangents moved to ``device`` once, for calls of
    jv_gradients that should not copy them again."""
    batch = from_packed(packed, device)
    dtan = {k: torch.as_tensor(np.ascontiguousarray(tan[k]), device=batch["_pt"].device)
            for k in ("pt", "seg", "sc")}
    return batch, dtan


def _lanes(batch, lanes):
    """The solve's lanes per device: ``lanes``, or for None 4 where the four-lane kernel takes
    the batch (no mobile ions, NP >= 16), else 2."""
    if lanes is not None:
        return lanes
    try:
        check_four_lanes(batch)
    except ValueError:
        return 2
    return 4


FROZEN_REFUSED = "the adjoint refused a device whose ions are frozen at a pre-bias"


def _pre_states(batch):
    """Room for the pre-bias solutions (NP, 3, n) where some device of the batch has its
    mobile ions frozen at its pre-bias (the adjoint's second linearisation point), else None."""
    sc = batch["_sc"]
    ions = (sc[SC_FIELDS.index("a_total")] > 0) | (sc[SC_FIELDS.index("c_total")] > 0)
    frozen = (sc[SC_FIELDS.index("pre_cond")] > 0) & (sc[SC_FIELDS.index("fix_ions")] > 0)
    if not bool((ions & frozen).any()):
        return None
    return torch.zeros((batch["NP"], 3, batch["n"]), dtype=torch.float64,
                       device=batch["_pt"].device)


def _shunt_tangents(dsc):
    """(K, n): the tangents of G_shunt = 1/R_shunt. The shunt's current G_shunt V is added
    outside the kernels (as the reference adds it to every voltage's J), and so its part of
    dJ/dtheta, V dG_shunt/dtheta, here."""
    return dsc[:, SC_FIELDS.index("G_shunt"), :]


def jv_gradients(packed, tan, voltages, tol=1e-10, max_iter=60, strict=False,
                 bernoulli_name="exact", device="cuda", return_states=False, events=None,
                 lanes=None, precision="fp64"):
    """J (n, nV), converged, iterations and dJ/dtheta (n, nV, 16) on the GPU for ascending
    voltages: the solve keeping each voltage's final state, then the adjoint kernel.
    Unconverged voltages get zero gradients (but for the shunt's current G_shunt V, added to
    every voltage's J). ``packed``, ``tan``: host arrays (`tangents`), or
    the device-resident pair `resident` returns. ``return_states``: also the linearisation
    points (nV, NP, 3, n) and the pre-bias solutions (NP, 3, n) of a batch with ions frozen
    at a pre-bias (None for another). ``events``: four torch.cuda.Event recorded on the launch stream
    around the solve kernel and around the adjoint kernel. ``lanes``: GPU lanes per device of
    the solve (4: the grid in quarters, the fastest for the batches gradients are taken of;
    2; 1; None: 4 where the four-lane kernel takes the batch, else 2), as
    `jaxdd_amd.batch.solve_batch`'s. ``precision``: the solve's, as solve_batch's
    ("mixed": the Newton correction's factors in FP32; the residual, the states and the
    adjoint stay FP64, and the states converge to the same ``tol``)."""
    lib = _lib.load()
    V = np.asarray(voltages, dtype=float)
    if np.any(np.diff(V) <= 0):
        raise ValueError("jv_gradients takes ascending voltages")
    batch = packed if "_pt" in packed else from_packed(packed, device)
    n, NP, nV = batch["n"], batch["NP"], len(V)
    lanes = _lanes(batch, lanes)
    if lanes not in (1, 2, 4):
        raise ValueError(f"lanes {lanes!r}: 1, 2 or 4")
    if precision not in _lib.PRECISIONS:
        raise ValueError(f"precision {precision!r}: one of {tuple(_lib.PRECISIONS)}")
    prec = _lib.PRECISIONS[precision]
    if lanes == 4:
        check_four_lanes(batch)
    opts = _lib.Opts(float(tol), int(max_iter), int(max_iter), int(bool(strict)),
                     0 if bernoulli_name == "exact" else 1, prec, 0, int(lanes))
    dev = batch["_pt"].device
    Vd = torch.as_tensor(V, device=dev)
    J = torch.empty((n, nV), dtype=torch.float64, device=dev)
    ok = torch.zeros((n, nV), dtype=torch.uint8, device=dev)
    its = torch.zeros((n, nV), dtype=torch.int32, device=dev)
    status = torch.zeros(n, dtype=torch.int32, device=dev)
    states = torch.empty((nV, NP, 3, n), dtype=torch.float64, device=dev)
    pre = _pre_states(batch)
    nbytes = lib.ddh_workspace_bytes(n, NP, nV, prec)
    ws = torch.empty(nbytes, dtype=torch.uint8, device=dev)
    # the per-device keys the caller may have replaced (g_frac, J_scale, V_pre), as solve_batch
    sc = _sc_now(batch)
    b = _lib.Batch(n, NP, batch["nseg"], 0, batch["_pt"].data_ptr(), batch["_sid"].data_ptr(),
                   batch["_seg"].data_ptr(), sc.data_ptr())
    stream = torch.cuda.current_stream(dev).cuda_stream
    K = tan["pt"].shape[0]
    dpt, dseg, dsc = (tan[k] if isinstance(tan[k], torch.Tensor)
                      else torch.as_tensor(np.ascontiguousarray(tan[k]), device=dev)
                      for k in ("pt", "seg", "sc"))
    grad = torch.empty((n, nV, K), dtype=torch.float64, device=dev)
    astatus = torch.zeros(n, dtype=torch.int32, device=dev)
    abytes = lib.ddh_adjoint_workspace_bytes(n, NP, nV)
    aws = torch.empty(abytes, dtype=torch.uint8, device=dev)
    if events is not None:
        events[0].record()
    pp = None if pre is None else pre.data_ptr()
    _lib.check(lib.ddh_solve_batch_states(ctypes.byref(b), Vd.data_ptr(), nV, ctypes.byref(opts),
                                          J.data_ptr(), ok.data_ptr(), its.data_ptr(),
                                          status.data_ptr(), states.data_ptr(), pp, ws.data_ptr(),
                                          nbytes, stream), "ddh_solve_batch_states")
    if events is not None:
        events[1].record()
        events[2].record()
    _lib.check(lib.ddh_adjoint(ctypes.byref(b), Vd.data_ptr(), nV, states.data_ptr(), pp,
                               ok.data_ptr(), K, dpt.data_ptr(), dseg.data_ptr(), dsc.data_ptr(),
                               opts.bernoulli, grad.data_ptr(), astatus.data_ptr(), aws.data_ptr(),
                               abytes, stream), "ddh_adjoint")
    if events is not None:
        events[3].record()
    torch.cuda.current_stream(dev).synchronize()
    if int(astatus.max()) != 0:
        raise NotImplementedError(FROZEN_REFUSED)
    del sc, ws, aws  # the launches that read them are done
    J = J + batch["G_shunt"][:, None] * Vd[None, :]
    grad = grad + Vd[None, :, None] * _shunt_tangents(dsc).T[:, None, :]
    if return_states:
        return J, ok.bool(), its.long(), grad, states, pre
    return J, ok.bool(), its.long(), grad


#: The name of the applied voltage among the numbers a gradient is taken in: the reference
#: differentiates in the voltages too (`jax.grad` with respect to `solve_batch`'s V).
VOLTAGE = "V"


def directions_for(wrt, n: int, nl: int) -> tuple[np.ndarray, np.ndarray]:
    """Unit directions in the builder's numbers: ``wrt`` names per-device numbers
    (`builder.SCALAR_FIELDS`, e.g. "W_L") or (name, layer) pairs of per-layer numbers
    (`builder.LAYER_FIELDS`, e.g. ("mu_n", 1)), or "V" (`VOLTAGE`, a zero row here: it is no
    number of the builder's, and `voltage_tangents` sets its tangents); returns (dlayers
    (K, nl, 23, n), dscalars (K, 10, n))."""
    from .builder import LAYER_FIELDS, SCALAR_FIELDS

    K = len(wrt)
    dl = np.zeros((K, nl, len(LAYER_FIELDS), n))
    ds = np.zeros((K, len(SCALAR_FIELDS), n))
    for k, w in enumerate(wrt):
        if isinstance(w, str) and w == VOLTAGE:
            continue  # no number of the builder's: `voltage_tangents` sets its row
        if isinstance(w, str):
            if w not in SCALAR_FIELDS:
                raise ValueError(f"{w!r}: not one of the per-device numbers {SCALAR_FIELDS}")
            ds[k, SCALAR_FIELDS.index(w)] = 1.0
            continue
        name, layer = w
        if name not in LAYER_FIELDS or not 0 <= int(layer) < nl:
            raise ValueError(f"{w!r}: not a per-layer number {LAYER_FIELDS} of a layer < {nl}")
        dl[k, int(layer), LAYER_FIELDS.index(name)] = 1.0
    return dl, ds


def param_gradients(structures, params, wrt, voltages, tol=1e-10, max_iter=60, lanes=None,
                    precision="fp64", device="cuda"):
    """dJ(V)/d(number) of each device's curve in the numbers ``wrt`` (`directions_for`): the
    counterpart of `jax.grad` over `device_tree(structure, params)` (reference:
    src/jaxdd/batch.py:142-148, 650-737) for any of the numbers the GPU builder takes, each
    device's structure held. ``structures``, ``params``: the devices' (`device_parameters`);
    returns J (n, nV), converged, iterations and the gradient (n, nV, len(wrt)) as
    `jv_gradients`; "V" in ``wrt`` gives each voltage's dJ/dV (the shunt's G_shunt included)."""
    from .builder import build

    dl, ds = directions_for(wrt, len(structures), len(structures[0].i0))
    batch, tan = build(list(structures), list(params), (dl, ds), device=device)
    vk = [k for k, w in enumerate(wrt) if isinstance(w, str) and w == VOLTAGE]
    if vk:
        tan = voltage_tangents(batch, tan, vk)
    J, ok, its, grad = jv_gradients(batch, tan, voltages, tol=tol, max_iter=max_iter,
                                    lanes=lanes, precision=precision)
    if vk:  # the shunt's current G_shunt V
        grad[:, :, vk] += batch["G_shunt"][:, None, None].to(grad.dtype)
    return J, ok, its, grad


def voltage_tangents(batch, tan, rows):
    """The tangents ``tan`` with directions ``rows`` set to the applied voltage's: V enters a
    voltage's solve only through the contacts' potentials, V0 + elec V / 2 and VL - elec V / 2
    (src/jaxdd/batch.py:255-259 `_bounds`), so its tangent is elec / 2 on the left contact's
    and -elec / 2 on the right's, and each voltage's dJ/dV is that direction's derivative
    (J(V) depends on its own voltage only)."""
    sc = batch["_sc"]
    elec = sc[SC_FIELDS.index("left_elec")]
    out = {k: (v.clone() if isinstance(v, torch.Tensor) else np.array(v)) for k, v in tan.items()}
    for k in rows:
        for key in ("pt", "seg", "sc"):
            out[key][k] = 0
        e = elec if isinstance(out["sc"], torch.Tensor) else elec.cpu().numpy()
        out["sc"][k][SC_FIELDS.index("V0")] = 0.5 * e
        out["sc"][k][SC_FIELDS.index("VL")] = -0.5 * e
    return out


def jv_vjp(batch, tan, voltages, dJbar, tol=1e-10, max_iter=60, bernoulli_name="exact",
           lanes=None):
    """The loss's gradient in the parameters from its derivative in the curves: theta_bar
    (n, K) = sum over voltages of dJbar (n, nV) times dJ/dtheta, contracted on the device
    (ddh_adjoint_vjp); also J (n, nV) and converged. ``batch``, ``tan``: device-resident
    (`gpu_tangents`, `builder.build`)."""
    lib = _lib.load()
    V = np.asarray(voltages, dtype=float)
    if np.any(np.diff(V) <= 0):
        raise ValueError("jv_vjp takes ascending voltages")
    n, NP, nV = batch["n"], batch["NP"], len(V)
    K = tan["pt"].shape[0]
    dev = batch["_pt"].device
    Jb = torch.as_tensor(np.asarray(dJbar, dtype=float), device=dev).contiguous()
    if tuple(Jb.shape) != (n, nV):
        raise ValueError(f"dJbar: ({n}, {nV})")
    lanes = _lanes(batch, lanes)
    if lanes == 4:
        check_four_lanes(batch)
    opts = _lib.Opts(float(tol), int(max_iter), int(max_iter), 0,
                     0 if bernoulli_name == "exact" else 1, _lib.DDH_FP64, 0, int(lanes))
    Vd = torch.as_tensor(V, device=dev)
    J = torch.empty((n, nV), dtype=torch.float64, device=dev)
    ok = torch.zeros((n, nV), dtype=torch.uint8, device=dev)
    its = torch.zeros((n, nV), dtype=torch.int32, device=dev)
    status = torch.zeros(n, dtype=torch.int32, device=dev)
    states = torch.empty((nV, NP, 3, n), dtype=torch.float64, device=dev)
    pre = _pre_states(batch)
    pp = None if pre is None else pre.data_ptr()
    nbytes = lib.ddh_workspace_bytes(n, NP, nV, _lib.DDH_FP64)
    ws = torch.empty(nbytes, dtype=torch.uint8, device=dev)
    sc = _sc_now(batch)
    b = _lib.Batch(n, NP, batch["nseg"], 0, batch["_pt"].data_ptr(), batch["_sid"].data_ptr(),
                   batch["_seg"].data_ptr(), sc.data_ptr())
    stream = torch.cuda.current_stream(dev).cuda_stream
    _lib.check(lib.ddh_solve_batch_states(ctypes.byref(b), Vd.data_ptr(), nV, ctypes.byref(opts),
                                          J.data_ptr(), ok.data_ptr(), its.data_ptr(),
                                          status.data_ptr(), states.data_ptr(), pp, ws.data_ptr(),
                                          nbytes, stream), "ddh_solve_batch_states")
    tbar = torch.empty((n, K), dtype=torch.float64, device=dev)
    astatus = torch.zeros(n, dtype=torch.int32, device=dev)
    vbytes = lib.ddh_adjoint_vjp_workspace_bytes(n, NP, nV, K)
    vws = torch.empty(vbytes, dtype=torch.uint8, device=dev)
    _lib.check(lib.ddh_adjoint_vjp(ctypes.byref(b), Vd.data_ptr(), nV, states.data_ptr(), pp,
                                   ok.data_ptr(), K, tan["pt"].data_ptr(), tan["seg"].data_ptr(),
                                   tan["sc"].data_ptr(), opts.bernoulli, Jb.data_ptr(),
                                   tbar.data_ptr(), astatus.data_ptr(), vws.data_ptr(), vbytes,
                                   stream), "ddh_adjoint_vjp")
    torch.cuda.current_stream(dev).synchronize()
    if int(astatus.max()) != 0:
        raise NotImplementedError(FROZEN_REFUSED)
    tbar = tbar + (Jb @ Vd)[:, None] * _shunt_tangents(tan["sc"]).T
    return tbar, J + batch["G_shunt"][:, None] * Vd[None, :], ok.bool()


class _Curves(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta, voltages, tol, max_iter):
        th = np.atleast_2d(theta.detach().cpu().numpy())
        dev = theta.device if theta.is_cuda else "cuda"
        if not ctx.needs_input_grad[0]:  # no backward will follow: the curves only
            from .batch import solve_batch
            from .builder import build_graded

            layers, scalars = numbers_of(th)
            batch = build_graded(th[:, 0:3], STACK_FLAGS, layers, scalars, device=dev)
            J, ok, _ = solve_batch(batch, voltages, tol=tol, max_iter=max_iter)
            ctx.mark_non_differentiable(ok)
            return J.to(theta.device), ok.to(theta.device)
        batch, tan = gpu_tangents(th, device=dev)
        J, ok, its, grad = jv_gradients(batch, tan, voltages, tol=tol, max_iter=max_iter)
        ctx.save_for_backward(grad.to(theta.device))
        ctx.mark_non_differentiable(ok)
        return J.to(theta.device), ok.to(theta.device)

    @staticmethod
    def backward(ctx, J_bar, ok_bar):
        (grad,) = ctx.saved_tensors
        return torch.einsum("nv,nvk->nk", J_bar.to(grad.dtype), grad), None, None, None


def curves(theta: torch.Tensor, voltages, tol=1e-10, max_iter=60):
    """J-V curves (n, nV) and converged flags of the devices of ``theta`` (n, 16) float64, at
    ascending ``voltages``, differentiable in theta (backward: dJ/dtheta from the adjoint
    kernel; zero at unconverged voltages, as the reference's). The devices, their graded
    layouts and, for a backward, their tangents are built on the GPU."""
    return _Curves.apply(theta, voltages, tol, max_iter)
