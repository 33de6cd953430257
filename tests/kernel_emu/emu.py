"""ctypes binding of the test-only CPU build of the kernels' per-lane logic (emu.cpp).

It runs the same per-lane code as the HIP kernels on packed host arrays (jaxdd_amd.batch.pack),
so tests can compare the kernels' algorithm with the oracle without a GPU.
"""

from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "libddemu.so"
_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


def lib():
    global _lib
    if _lib is None:
        build()
        _lib = ctypes.CDLL(str(LIB))
        vp, i = ctypes.c_void_p, ctypes.c_int
        _lib.emu_solve_batch.argtypes = [i, i, i, vp, vp, vp, vp, vp, i, ctypes.c_double, i, i,
                                         i, i, i, i, vp, vp, vp, vp]
        _lib.emu_keep_states.argtypes = [vp, vp]
        _lib.emu_set_lanes.argtypes = [i]
        _lib.emu_set_adjoint_passes.argtypes = [i]
        _lib.emu_adjoint.argtypes = [i, i, i, vp, vp, vp, vp, vp, i, vp, vp, vp, i, vp, vp, vp,
                                     i, vp, vp]
        _lib.emu_assemble.argtypes = [i, i, i, vp, vp, vp, vp, vp, vp, vp, i, vp, vp, vp, vp,
                                      vp, vp]
        _lib.emu_impedance.argtypes = [i, i, i, vp, vp, vp, vp, vp, vp, vp, vp, i,
                                       ctypes.c_double, i, vp, vp, vp, vp, i, vp]
        _lib.emu_assemble_ions.argtypes = [i, i, i, vp, vp, vp, vp, vp, vp, vp, i, vp, vp, vp,
                                           vp, vp]
        _lib.emu_xcd_hits.argtypes = [i, i, vp]
        _lib.emu_build.argtypes = [i, i, i, i, i] + [vp] * 8 + [vp, i] + [vp] * 9
        _lib.emu_graded_layout.argtypes = [i, i, i, vp, ctypes.c_double, ctypes.c_double] + [vp] * 7
        _lib.emu_transient.argtypes = [i, i, i, i, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp,
                                       ctypes.c_double, i, i, i, i, i, vp, vp, vp, vp, vp, vp]
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def solve_packed(packed, V, tol=1e-10, max_iter=60, start_iter=None, strict=False,
                 bernoulli="exact", precision="fp64", jv=False, states=None, lanes=1, pre=None):
    """(J without the shunt, ok, its, status) for packed devices at voltages V (ascending in
    batch mode, the given order in jv mode). ``states``: an array (nV, NP, 3, n) that receives
    each voltage's final state; ``pre``: one (NP, 3, n) that receives the pre-bias solutions."""
    n = packed["sc"].shape[1]
    NP = packed["pt"].shape[0] - 2
    V = np.ascontiguousarray(V, dtype=np.float64)
    nV = len(V)
    J = np.zeros((n, nV))
    ok = np.zeros((n, nV), dtype=np.uint8)
    its = np.zeros((n, nV), dtype=np.int32)
    status = np.zeros(n, dtype=np.int32)
    arrs = [np.ascontiguousarray(packed[k]) for k in ("pt", "sid", "seg", "sc")]
    if states is not None:
        assert states.shape == (nV, NP, 3, n) and states.dtype == np.float64
        assert states.flags.c_contiguous
    if pre is not None:
        assert pre.shape == (NP, 3, n) and pre.dtype == np.float64 and pre.flags.c_contiguous
    lib().emu_keep_states(None if states is None else _p(states), None if pre is None else _p(pre))
    lib().emu_set_lanes(int(lanes))
    lib().emu_solve_batch(n, NP, packed["nseg"], *(_p(a) for a in arrs), _p(V), nV, tol,
                          max_iter, start_iter or max_iter, int(strict),
                          0 if bernoulli == "exact" else 1, 0 if precision == "fp64" else 1,
                          int(jv), _p(J), _p(ok), _p(its), _p(status))
    return J, ok.astype(bool), its, status


def adjoint_packed(packed, V, states, ok, tangents, bernoulli="exact", passes=False, pre=None):
    """dJ/dtheta (n, nV, ntheta) and status (n,) of the adjoint kernel's lanes, replayed;
    ``passes``: one pass per direction instead of the reverse pass over the parameters;
    ``pre``: the pre-bias solutions (NP, 3, n) of devices whose ions are frozen at them."""
    n = packed["sc"].shape[1]
    NP = packed["pt"].shape[0] - 2
    V = np.ascontiguousarray(V, dtype=np.float64)
    nV = len(V)
    dpt, dseg, dsc = (np.ascontiguousarray(tangents[k]) for k in ("pt", "seg", "sc"))
    K = dpt.shape[0]
    grad = np.zeros((n, nV, K))
    status = np.zeros(n, dtype=np.int32)
    arrs = [np.ascontiguousarray(packed[k]) for k in ("pt", "sid", "seg", "sc")]
    okb = np.ascontiguousarray(ok, dtype=np.uint8)
    lib().emu_set_adjoint_passes(int(bool(passes)))
    lib().emu_adjoint(n, NP, packed["nseg"], *(_p(a) for a in arrs), _p(V), nV,
                      _p(np.ascontiguousarray(states)),
                      None if pre is None else _p(np.ascontiguousarray(pre)), _p(okb), K, _p(dpt),
                      _p(dseg), _p(dsc),
                      0 if bernoulli == "exact" else 1, _p(grad), _p(status))
    return grad, status


def assemble_packed(packed, y, Vint, logs=None, bernoulli="exact"):
    """The assembly operator's CPU replay: y (n, NP, 3), Vint (n,), logs (n, 2) or None ->
    F (n, NP, 3), A, B, C (n, NP, 3, 3), U (n, NP, 3, 2), J (n,) without the shunt."""
    n = packed["sc"].shape[1]
    NP = packed["pt"].shape[0] - 2
    arrs = [np.ascontiguousarray(packed[k]) for k in ("pt", "sid", "seg", "sc")]
    yl = np.ascontiguousarray(np.asarray(y, dtype=float).transpose(1, 2, 0))
    Vd = np.ascontiguousarray(Vint, dtype=float)
    lg = None if logs is None else np.ascontiguousarray(np.asarray(logs, dtype=float).T)
    F = np.zeros((NP, 3, n)); A = np.zeros((NP, 9, n)); B = np.zeros((NP, 9, n))
    C = np.zeros((NP, 9, n)); U = np.zeros((NP, 6, n)); J = np.zeros(n)
    lib().emu_assemble(n, NP, packed["nseg"], *(_p(a) for a in arrs), _p(yl), _p(Vd),
                       _p(lg) if lg is not None else None,
                       0 if bernoulli == "exact" else 1, _p(F), _p(A), _p(B), _p(C), _p(U),
                       _p(J))
    return (F.transpose(2, 0, 1), A.reshape(NP, 3, 3, n).transpose(3, 0, 1, 2),
            B.reshape(NP, 3, 3, n).transpose(3, 0, 1, 2), C.reshape(NP, 3, 3, n).transpose(3, 0, 1, 2),
            U.reshape(NP, 3, 2, n).transpose(3, 0, 1, 2), J)


#: The impedance table's fields (jaxdd_amd/csrc/ddhip_impedance.h `ImpField`).
IMP_FIELDS = {name: k for k, name in enumerate(
    ["A0", "A3", "A4", "A6", "A8"] + [f"B{q}" for q in range(9)] + ["C0", "C3", "C4", "C6", "C8"]
    + ["MN", "MP", "RATE", "E", "M", "UN", "UP", "TN", "TP", "KV", "KD", "G0N", "G0P", "GLN",
       "GLP", "GT", "DRN", "DRP", "NN", "PP", "TMP"]
    # the 5 x 5 path's fields after the 40 above (the table's 64-double rows)
    + ["B03", "B04"] + [f"IA{q}" for q in range(6)] + [f"IC{q}" for q in range(6)]
    + ["MA", "MC", "WA", "WC", "CA0", "CAL", "CC0", "CCL", "LA", "LC"])}


def impedance_packed(packed, y, Vint, laps, dt=None, bernoulli="exact", table=False,
                     ions=False):
    """The impedance operator's CPU replay: y (n, NP, 3) steady states, Vint (n,), laps
    (complex, nf) -> Y (n, nf) complex, status (n,); ``table``: also the per-device tables
    (n, NP, fields) and scalars (n, 8); ``ions``: the 5 x 5 path (the ions as unknowns)."""
    n = packed["sc"].shape[1]
    NP = packed["pt"].shape[0] - 2
    arrs = [np.ascontiguousarray(packed[k]) for k in ("pt", "sid", "seg", "sc")]
    yl = np.ascontiguousarray(np.asarray(y, dtype=float).transpose(1, 2, 0))
    Vd = np.ascontiguousarray(Vint, dtype=float)
    iface = np.ascontiguousarray(packed["iface"], dtype=np.uint8)
    lp = np.ascontiguousarray(np.stack([np.real(laps), np.imag(laps)], -1), dtype=float)
    nf = lp.shape[0]
    imu = np.ascontiguousarray(packed["imu"], dtype=float)
    from jaxdd_amd._lib import IMP_TABLE_FIELDS, IMP_ION_TABLE_FIELDS, IMP_SCALARS

    fields = IMP_ION_TABLE_FIELDS if ions else IMP_TABLE_FIELDS
    tab = np.zeros((n, NP, fields)); dsc = np.zeros((n, IMP_SCALARS))
    Y = np.zeros((n, nf, 2)); status = np.zeros(n, dtype=np.int32)
    lib().emu_impedance(n, NP, packed["nseg"], *(_p(a) for a in arrs), _p(yl), _p(Vd),
                        _p(iface), _p(lp), nf, 0.0 if dt is None else float(dt),
                        0 if bernoulli == "exact" else 1, _p(tab), _p(dsc), _p(Y), _p(status),
                        int(bool(ions)), _p(imu))
    if table:  # the table per point: (n, NP, fields)
        return Y[..., 0] + 1j * Y[..., 1], status, tab, dsc
    return Y[..., 0] + 1j * Y[..., 1], status


def assemble_ions_packed(packed, y, Vint, logs, bernoulli="exact"):
    """The 5x5 assembly operator's CPU replay: y (n, NP, 5), Vint (n,), logs (n, 2) ->
    F (n, NP, 5), A, B, C (n, NP, 5, 5), U (n, NP, 5, 2)."""
    n = packed["sc"].shape[1]
    NP = packed["pt"].shape[0] - 2
    arrs = [np.ascontiguousarray(packed[k]) for k in ("pt", "sid", "seg", "sc")]
    yl = np.ascontiguousarray(np.asarray(y, dtype=float).transpose(1, 2, 0))
    Vd = np.ascontiguousarray(Vint, dtype=float)
    lg = np.ascontiguousarray(np.asarray(logs, dtype=float).T)
    F = np.zeros((NP, 5, n)); A = np.zeros((NP, 25, n)); B = np.zeros((NP, 25, n))
    C = np.zeros((NP, 25, n)); U = np.zeros((NP, 10, n))
    lib().emu_assemble_ions(n, NP, packed["nseg"], *(_p(a) for a in arrs), _p(yl), _p(Vd),
                            _p(lg), 0 if bernoulli == "exact" else 1, _p(F), _p(A), _p(B),
                            _p(C), _p(U))
    return (F.transpose(2, 0, 1), A.reshape(NP, 5, 5, n).transpose(3, 0, 1, 2),
            B.reshape(NP, 5, 5, n).transpose(3, 0, 1, 2), C.reshape(NP, 5, 5, n).transpose(3, 0, 1, 2),
            U.reshape(NP, 5, 2, n).transpose(3, 0, 1, 2))


def xcd_hits(n: int, nblocks: int):
    """(hits per device, pairs outside [0, n)) of the kernels' workgroup-to-device XCD map."""
    hits = np.zeros(n, dtype=np.int32)
    outside = lib().emu_xcd_hits(n, nblocks, _p(hits))
    return hits, outside


def transient_packed(packed, t, V, G, y0, tol=1e-10, max_iter=200, bern="exact", zimt=False,
                     failure_mode=1, ions=None):
    """The transient kernel's per-lane code on the CPU: ``t``, ``V``, ``G`` (n, nt) per device,
    ``y0`` the steady start (NP, 3, n) lane-innermost. Returns J, J_ext, ok, its, kept (n, nt)
    and status (n)."""
    n = packed["sc"].shape[1]
    NP = packed["pt"].shape[0] - 2
    nt = np.asarray(t).shape[-1]
    t, V, G = (np.ascontiguousarray(np.broadcast_to(np.asarray(a, float), (n, nt))) for a in (t, V, G))
    if ions is None:
        ions = bool(np.any(packed["sc"][12] > 0) or np.any(packed["sc"][13] > 0))
    J, Jext = np.zeros((n, nt)), np.zeros((n, nt))
    ok, kept = np.zeros((n, nt), np.uint8), np.zeros((n, nt), np.uint8)
    its, status = np.zeros((n, nt), np.int32), np.zeros(n, np.int32)
    arrs = [np.ascontiguousarray(packed[k]) for k in ("pt", "sid", "iface", "seg", "sc", "imu")]
    y0 = np.ascontiguousarray(y0, dtype=np.float64)
    lib().emu_transient(n, NP, packed["nseg"], nt, *(_p(a) for a in arrs), _p(t), _p(V), _p(G),
                        _p(y0), tol, max_iter, {"exact": 0, "simss": 1}[bern], int(zimt),
                        failure_mode, int(ions), _p(J), _p(Jext), _p(ok), _p(its), _p(kept),
                        _p(status))
    return J, Jext, ok.astype(bool), its, kept.astype(bool), status


def build_packed(structures, params, directions=None):
    """The GPU builder's lanes on the CPU (jaxdd_amd.builder's arrays): the packed arrays
    {"pt", "seg", "sc", "imu", "sid", "nseg"} and, with ``directions`` (dlayers, dscalars), the
    tangents {"pt", "seg", "sc"}."""
    from jaxdd_amd import builder as B
    from jaxdd_amd.batch import SC_FIELDS

    sa = B.structure_arrays(structures)
    layers, scalars = B.number_arrays(structures, params)
    gen = B.profile_arrays(structures, params)
    gen_m = 0 if gen is None else gen.shape[1]
    n, NP, nl, nseg = len(structures), sa["NP"], sa["nl"], sa["nseg"]
    P = NP + 2
    K = 0 if directions is None else np.asarray(directions[0]).shape[0]
    pt, seg = np.zeros((P, 4, n)), np.zeros((nseg, 23, n))
    sc, imu = np.zeros((len(SC_FIELDS), n)), np.zeros((nseg, 2, n))
    dpt, dseg = np.zeros((max(K, 1), P, 4, n)), np.zeros((max(K, 1), nseg, 23, n))
    dsc = np.zeros((max(K, 1), len(SC_FIELDS), n))
    dl = np.ascontiguousarray(directions[0], dtype=float) if K else np.zeros(1)
    ds = np.ascontiguousarray(directions[1], dtype=float) if K else np.zeros(1)
    lib().emu_build(n, NP, nl, nseg, K, _p(sa["lid"]), _p(sa["sid"]), _p(sa["flags"]),
                    _p(sa["i0"]), _p(sa["unit"]), _p(sa["unit_sum"]), _p(layers), _p(scalars),
                    _p(gen) if gen_m else None, gen_m, _p(dl), _p(ds), _p(pt), _p(seg), _p(sc), _p(imu), _p(dpt), _p(dseg), _p(dsc))
    out = {"pt": pt, "seg": seg, "sc": sc, "imu": imu, "sid": sa["sid"], "nseg": nseg}
    if K:
        return out, {"pt": dpt, "seg": dseg, "sc": dsc}
    return out


def graded_layout_packed(L, NP, beta, d0):
    """The GPU layout's lanes on the CPU: L (n, nl) thicknesses -> dict of lid, sid (NP + 2, n),
    i0 (nl, n), unit (NP + 2, n), unit_sum (nl, n), iface (NP + 1, n), nseg (n)."""
    L = np.ascontiguousarray(np.asarray(L, dtype=float).T)
    nl, n = L.shape
    P = NP + 2
    out = {"lid": np.zeros((P, n), np.uint8), "sid": np.zeros((P, n), np.uint8),
           "i0": np.zeros((nl, n), np.int32), "unit": np.zeros((P, n)),
           "unit_sum": np.zeros((nl, n)), "iface": np.zeros((P - 1, n), np.uint8),
           "nseg": np.zeros(n, np.int32)}
    lib().emu_graded_layout(n, NP, nl, _p(L), float(beta), float(d0),
                            *(_p(out[k]) for k in ("lid", "sid", "i0", "unit", "unit_sum",
                                                    "iface", "nseg")))
    return out
