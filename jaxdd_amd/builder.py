"""The batch built on the GPU (`ddh_build`): each device's packed arrays from its structure and
per-layer numbers, and their derivatives along given directions in those numbers (reference:
src/jaxdd/device.py:877-1063 `device_arrays`, src/jaxdd/batch.py:142-223 `device_tree`,
`prepare`).

The host keeps what is discrete, the structures (each point's layer, the grid's unscaled
spacing, which layers hold traps or ions), and hands the GPU each layer's numbers; the GPU
places the grid, fills every point's and segment's fields and the device's scalars, and, per
direction, their derivatives (forward mode, the structure held: the thicknesses scale each
layer's spacing, as the reference differentiates `device_tree(structure, params)`). The result
is the batch dict `batch.prepare` returns, on the GPU, and the tangents `adjoint.jv_gradients`
takes.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from .batch import MAX_SEGMENTS, SC_FIELDS, SEG_FIELDS
from .device import Structure
from .grid import GRADED_POINTS_PER_EFOLD, GRADED_SMALLEST

#: The per-layer numbers of ddh_build, in its order.
LAYER_FIELDS = ("L", "eps_r", "E_c", "E_v", "N_c", "N_D", "N_A", "mu_n", "mu_p", "nu_int_n",
                "nu_int_p", "k_direct", "N_t_bulk", "C_n_bulk", "C_p_bulk", "E_t_bulk", "G_ehp",
                "N_anion", "N_cation", "mu_anion", "mu_cation", "preLangevin", "bulkTrapType")
#: The per-device numbers of ddh_build, in its order (G_shunt = 1/R_shunt, 0 without a shunt).
#: gen_profile: 1 where the device's generation is its profile (`gen_x`, `gen_G`), else 0.
SCALAR_FIELDS = ("T", "W_L", "W_R", "G_frac", "V_pre", "G_shunt", "pre_cond", "fix_ions",
                 "left_elec", "gen_profile")
MAX_LAYERS = 8
_TRAPS, _A_MOBILE, _A_FIXED, _C_MOBILE, _C_FIXED = 1, 2, 4, 8, 16
_LANGEVIN, _LANGEVIN_INTERVAL, _LAYER_GEN = 32, 64, 128


def check_buildable(st: Structure) -> None:
    """`NotImplementedError` for a structure outside ddh_build's physics: generation other than
    each layer's or a profile file's, flat-band contacts, more than eight layers, optional
    modules."""
    if st.gen_mode not in ("none", "file"):
        raise NotImplementedError(f"generation mode {st.gen_mode!r}: the GPU builder takes "
                                  "each layer's G_ehp or a profile")
    if any(st.sfb):
        raise NotImplementedError("flat-band work functions are outside the GPU builder")
    if st.extra:
        raise NotImplementedError(f"optional modules outside the GPU builder: {sorted(st.extra)}")
    if len(st.i0) > MAX_LAYERS:
        raise NotImplementedError(f"at most {MAX_LAYERS} layers")


def layer_flags(st: Structure, k: int) -> int:
    """Layer ``k``'s bits for ddh_build: bulk traps, mobile and fixed anions and cations,
    Langevin's rate (at the interval's mobilities with ``langevin_mobility`` "simss"), and
    whether it takes the generation profile (its `layerGen`)."""
    langevin = st.langevin is not None and bool(st.langevin[k])
    gen = st.gen_mode == "file" and bool(st.layer_gen[k])
    return ((_TRAPS if st.traps[k] else 0) | (_A_MOBILE if st.anion_mobile[k] else 0)
            | (_A_FIXED if st.anion_fixed[k] else 0) | (_C_MOBILE if st.cation_mobile[k] else 0)
            | (_C_FIXED if st.cation_fixed[k] else 0) | (_LANGEVIN if langevin else 0)
            | (_LANGEVIN_INTERVAL if langevin and st.langevin_mobility == "simss" else 0)
            | (_LAYER_GEN if gen else 0))


def segment_ids(st: Structure) -> np.ndarray:
    """Each point's segment: a new one at every layer's first point, at each contact and its
    neighbour, and at the last point of a layer before an interface (its interval takes the
    interface's mobility), the points where a packed field may change."""
    P = st.NP + 2
    lid = np.asarray(st.lid)
    key = np.stack([lid, np.arange(P) == 0, np.arange(P) == P - 1,
                    np.isin(np.arange(P), list(st.i1[:-1]))], -1)
    change = np.any(key[1:] != key[:-1], axis=1)
    return np.concatenate([[0], np.cumsum(change)]).astype(np.int64)


def structure_arrays(structures) -> dict:
    """The lane-innermost structure arrays of ddh_build and the batch's iface (host NumPy)."""
    nl = {len(st.i0) for st in structures}
    NPs = {st.NP for st in structures}
    if len(nl) != 1 or len(NPs) != 1:
        raise ValueError("a built batch needs one grid size and one number of layers")
    nl, NP = nl.pop(), NPs.pop()
    for st in structures:
        check_buildable(st)
    sids = [segment_ids(st) for st in structures]
    nseg = int(max(s[-1] for s in sids)) + 1
    if nseg > MAX_SEGMENTS:
        raise NotImplementedError(f"{nseg} segments: at most {MAX_SEGMENTS}")
    flags = np.array([[layer_flags(st, k) for k in range(nl)] for st in structures],
                     dtype=np.uint8).T
    iface = np.zeros((NP + 1, len(structures)), dtype=np.uint8)
    for k, st in enumerate(structures):
        iface[list(st.i1[:-1]), k] = 1
    c = np.ascontiguousarray
    return {
        "NP": NP, "nl": nl, "nseg": nseg,
        "lid": c(np.stack([np.asarray(st.lid) for st in structures], -1).astype(np.uint8)),
        "sid": c(np.stack(sids, -1).astype(np.uint8)),
        "flags": c(flags),
        "i0": c(np.array([st.i0 for st in structures], dtype=np.int32).T),
        "unit": c(np.stack([np.asarray(st.raw, float) for st in structures], -1)),
        "unit_sum": c(np.stack([np.asarray(st.norm, float) for st in structures], -1)),
        "iface": iface,
    }


def number_arrays(structures, params) -> tuple[np.ndarray, np.ndarray]:
    """(layers [nl][23][n], scalars [10][n]) of the devices' numbers (`device_parameters`'
    dicts: per-layer arrays and scalars; a missing `preLangevin` is 1) and the bulk traps'
    charge type of each structure (`bulkTrapType`, where the layer holds traps)."""
    def number(st, p, f):
        if f == "preLangevin" and f not in p:
            return np.ones(len(np.atleast_1d(p["L"])))
        if f == "bulkTrapType":
            kind = np.zeros(len(st.i0)) if st.trap_type is None else st.trap_type
            return np.where(st.traps, kind, 0).astype(float)
        return np.asarray(p[f], dtype=float)

    layers = np.stack([np.stack([number(st, p, f) for f in LAYER_FIELDS], -1)
                       for st, p in zip(structures, params)], -1)  # (nl, 23, n)
    scal = []
    for st, p in zip(structures, params):
        R_shunt = float(p.get("R_shunt", -1.0))
        scal.append([float(p["T"]), float(p["W_L"]), float(p["W_R"]), float(p["G_frac"]),
                     float(p["V_pre"]), 1.0 / R_shunt if R_shunt > 0 else 0.0,
                     float(st.pre_cond), float(st.fix_ions), float(st.left_elec),
                     float(st.gen_mode == "file")])
    return np.ascontiguousarray(layers), np.ascontiguousarray(np.array(scal).T)


def profile_arrays(structures, params) -> np.ndarray | None:
    """ddh_build's generation profiles ``gen`` [2][m][n] (positions, then G), m the longest
    table, a shorter one padded by repeating its last node (the same G at every point); a
    device without a profile holds the positions 0..1 and G 0 (unread). None when no device
    takes a profile."""
    tables = [(np.asarray(p["gen_x"], dtype=float), np.asarray(p["gen_G"], dtype=float))
              if st.gen_mode == "file" else None for st, p in zip(structures, params)]
    m = max((len(t[0]) for t in tables if t is not None), default=0)
    if m == 0:
        return None
    if m < 2:
        raise ValueError("a generation profile needs two nodes at least")
    gen = np.zeros((2, m, len(tables)))
    gen[0] = np.linspace(0.0, 1.0, m)[:, None]
    for d, t in enumerate(tables):
        if t is not None:
            k = len(t[0])
            gen[0, :k, d], gen[1, :k, d] = t
            gen[0, k:, d], gen[1, k:, d] = t[0][-1], t[1][-1]
    return np.ascontiguousarray(gen)


def build(structures, params, directions=None, device="cuda"):
    """The batch (`batch.prepare`'s dict, on ``device``) of devices given by their structures
    and numbers, built by ddh_build. ``directions``: None, or (dlayers [K][nl][23][n], dscalars
    [K][10][n]) host arrays, directions in the numbers; then also the tangents {"pt", "seg",
    "sc"} (torch, on the device) along each, as `adjoint.jv_gradients` takes them."""
    dev = _cuda(device)
    sa = structure_arrays(structures)
    layers, scalars = number_arrays(structures, params)

    def t(a, dtype=torch.float64):
        return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype).to(dev)

    keep = {k: t(sa[k], torch.uint8) for k in ("lid", "sid", "flags", "iface")}
    keep["i0"] = t(sa["i0"], torch.int32)
    keep["unit"], keep["unit_sum"] = t(sa["unit"]), t(sa["unit_sum"])
    R_series = [float(p.get("R_series", 0.0)) for p in params]
    return _build(keep, sa["NP"], sa["nl"], sa["nseg"], layers, scalars, directions, R_series, dev,
                  profile_arrays(structures, params))


def _cuda(device):
    dev = torch.device(device)
    if dev.type != "cuda":
        raise RuntimeError("the batch is built on the GPU: device='cuda'")
    return dev


def build_graded(L, flags, layers, scalars, directions=None, NP: int = 400,
                 beta: float = GRADED_POINTS_PER_EFOLD, d0: float = GRADED_SMALLEST,
                 R_series=None, profile=None, device="cuda"):
    """The batch of devices on graded grids built entirely on the GPU: each device's layout
    from its thicknesses (ddh_graded_layout; ``L`` (n, nl)), then its arrays (ddh_build).
    ``flags`` (nl,) or (nl, n): each layer's bits (1 bulk traps, 2 mobile anions, 4 fixed
    anions, 8 mobile cations, 16 fixed cations, 32 Langevin recombination, 64 Langevin at the
    interval's mobilities, 128 the layer takes the profile); ``layers`` [nl][23][n] and
    ``scalars`` [10][n] the numbers (`number_arrays`' layout); ``profile`` None or the profiles
    [2][m][n] (`profile_arrays`' layout) of the devices whose `gen_profile` is 1; ``directions``
    as `build`'s. Each point stops its
    Newton steps when they no longer move it; the host restatement (`grid.graded_layout`) stops
    a layer's points together: positions agree to the last bits, the smallest spacings to
    1e-11 of themselves."""
    lib = _lib.load()
    dev = _cuda(device)
    L = np.asarray(L, dtype=float)
    n, nl = L.shape
    P = NP + 2

    def t(a, dtype=torch.float64):
        return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype).to(dev)

    keep = {"L": t(L.T)}
    keep["lid"] = torch.empty((P, n), dtype=torch.uint8, device=dev)
    keep["sid"] = torch.empty((P, n), dtype=torch.uint8, device=dev)
    keep["i0"] = torch.empty((nl, n), dtype=torch.int32, device=dev)
    keep["unit"] = torch.empty((P, n), dtype=torch.float64, device=dev)
    keep["unit_sum"] = torch.empty((nl, n), dtype=torch.float64, device=dev)
    keep["iface"] = torch.empty((P - 1, n), dtype=torch.uint8, device=dev)
    nseg = torch.empty(n, dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream(dev).cuda_stream
    _lib.check(lib.ddh_graded_layout(n, NP, nl, keep["L"].data_ptr(), float(beta), float(d0),
                                     keep["lid"].data_ptr(), keep["sid"].data_ptr(),
                                     keep["i0"].data_ptr(), keep["unit"].data_ptr(),
                                     keep["unit_sum"].data_ptr(), keep["iface"].data_ptr(),
                                     nseg.data_ptr(), stream), "ddh_graded_layout")
    counts = nseg.cpu().numpy()
    if np.any(counts < 1):
        raise ValueError("the graded layout's counts do not fill the grid: too few points for "
                         "the layers")
    if int(counts.max()) > MAX_SEGMENTS:
        raise NotImplementedError(f"{int(counts.max())} segments: at most {MAX_SEGMENTS}")
    fl = np.asarray(flags, dtype=np.uint8)
    keep["flags"] = t(np.broadcast_to(fl.reshape(nl, -1), (nl, n)), torch.uint8)
    R = np.zeros(n) if R_series is None else np.asarray(R_series, dtype=float)
    return _build(keep, NP, nl, int(counts.max()), np.asarray(layers, dtype=float),
                  np.asarray(scalars, dtype=float), directions, R, dev, profile)


def _build(keep, NP, nl, nseg, layers, scalars, directions, R_series, dev, profile=None):
    lib = _lib.load()
    n = keep["lid"].shape[1]
    P = NP + 2

    def t(a, dtype=torch.float64):
        return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype).to(dev)

    if layers.shape != (nl, len(LAYER_FIELDS), n) or scalars.shape != (len(SCALAR_FIELDS), n):
        raise ValueError(f"layers ({nl}, {len(LAYER_FIELDS)}, {n}) and scalars "
                         f"({len(SCALAR_FIELDS)}, {n})")
    gen_m = 0
    if profile is not None:
        profile = np.asarray(profile, dtype=float)
        gen_m = profile.shape[1] if profile.ndim == 3 else 0
        if profile.shape != (2, gen_m, n) or gen_m < 2:
            raise ValueError(f"profile: (2, m >= 2, {n})")
        keep["gen"] = t(profile)
    elif np.any(scalars[SCALAR_FIELDS.index("gen_profile")] > 0):
        raise ValueError("gen_profile 1 without a profile")
    keep["layers"], keep["scalars"] = t(layers), t(scalars)
    K = 0
    if directions is not None:
        dl, ds = (np.asarray(x, dtype=float) for x in directions)
        K = dl.shape[0]
        if dl.shape != (K, nl, len(LAYER_FIELDS), n) or ds.shape != (K, len(SCALAR_FIELDS), n):
            raise ValueError(f"directions: (K, {nl}, {len(LAYER_FIELDS)}, {n}) and "
                             f"(K, {len(SCALAR_FIELDS)}, {n})")
        keep["dlayers"], keep["dscalars"] = t(dl), t(ds)
    pt = torch.empty((P, 4, n), dtype=torch.float64, device=dev)
    seg = torch.empty((nseg, len(SEG_FIELDS), n), dtype=torch.float64, device=dev)
    sc = torch.empty((len(SC_FIELDS), n), dtype=torch.float64, device=dev)
    imu = torch.empty((nseg, 2, n), dtype=torch.float64, device=dev)
    tan = None
    if K:
        tan = {"pt": torch.empty((K, P, 4, n), dtype=torch.float64, device=dev),
               "seg": torch.empty((K, nseg, len(SEG_FIELDS), n), dtype=torch.float64,
                                  device=dev),
               "sc": torch.empty((K, len(SC_FIELDS), n), dtype=torch.float64, device=dev)}
    s = _lib.Structure(n, NP, nl, nseg, keep["lid"].data_ptr(), keep["sid"].data_ptr(),
                       keep["flags"].data_ptr(), keep["i0"].data_ptr(), keep["unit"].data_ptr(),
                       keep["unit_sum"].data_ptr())
    ptr = (lambda x: x.data_ptr() if x is not None else None)
    stream = torch.cuda.current_stream(dev).cuda_stream
    _lib.check(lib.ddh_build(ctypes.byref(s), keep["layers"].data_ptr(),
                             keep["scalars"].data_ptr(), ptr(keep.get("gen")), gen_m, K,
                             ptr(keep.get("dlayers")),
                             ptr(keep.get("dscalars")), pt.data_ptr(), seg.data_ptr(),
                             sc.data_ptr(), imu.data_ptr(),
                             *(ptr(tan[k]) if tan else None for k in ("pt", "seg", "sc")),
                             stream), "ddh_build")
    torch.cuda.current_stream(dev).synchronize()
    col = {name: sc[k] for k, name in enumerate(SC_FIELDS)}
    R_series = t(R_series)
    batch = {
        "n": n, "NP": NP, "nseg": nseg,
        "_pt": pt, "_sid": keep["sid"], "_seg": seg, "_sc": sc, "_iface": keep["iface"],
        "_imu": imu,
        "J_scale": col["J_scale"].clone(), "g_frac": col["g_frac"].clone(), "Vt": col["Vt"].clone(),
        "L": col["L"].clone(), "V_pre": col["V_pre"].clone(), "pre_cond": col["pre_cond"].clone(),
        "fix_ions": col["fix_ions"].clone(), "left_elec": col["left_elec"].clone(),
        "G_shunt": col["G_shunt"].clone(), "R_series": R_series,
        "x": pt[:, 1, :].T.contiguous(), "h": pt[:, 0, :].T.contiguous(),
        "_fixed": None,
    }
    return (batch, tan) if K else batch


__all__ = ["LAYER_FIELDS", "SCALAR_FIELDS", "build", "build_graded", "check_buildable",
           "number_arrays", "profile_arrays", "segment_ids", "structure_arrays"]
