"""Grids of a layer stack: how many points each layer gets and where they sit.

Two rules decide the points (reference interface: src/jaxdd/device.py `layer_points` :73-100,
`make_grid` :236-248, the graded layout :160-214):

- SIMsalabim's `Make_Grid` ("simss"): each layer gets ``max(5, floor(free * L / L_rest))``
  points beyond its first, where a layer whose ``NP * L / L_tot`` falls below 5 is "thin" and
  takes exactly 5, ``free`` is what the thin layers leave of NP and ``L_rest`` the other layers'
  thickness; inside a layer the spacing grows geometrically from both ends
  (``exp(grad * (1/m_half - 1/(i + 1)))`` at the i-th interval from the nearer end).
- The graded rule ("graded"): the point density ``1/H + beta/(d + d0)`` in the distance d to
  either end of the layer (H the stack's thickness over NP) is integrated in closed form; each
  layer gets its share of the integral (at least 5 points beyond its first) and its points
  equidistribute the integral, found here by a safeguarded Newton iteration on the integral.

A layout (`Layout`) is what depends on the thicknesses only through these discrete choices and
the unscaled spacing; `Layout.place` scales it to the thicknesses, and across each interface the
three spacings around it are replaced by their mean (SIMsalabim's rule), so that the grid is a
smooth function of the thicknesses at a fixed layout.
"""

from __future__ import annotations

import math
from collections.abc import Sequence
from dataclasses import dataclass
from fractions import Fraction

import numpy as np

__all__ = ["GRIDS", "GRID_SPLITS", "GRADED_POINTS_PER_EFOLD", "GRADED_SMALLEST", "Layout",
           "graded_layout", "layer_points", "make_grid", "simss_layout", "stack_sum"]

#: SIMsalabim's `MinGridPointsPerLayer`.
MIN_POINTS = 5
#: How the SIMsalabim rule evaluates a layer's share of the points: rational arithmetic on the
#: input's decimals ("exact"), or each operation rounded to the 64-bit significand of the x87
#: extended reals SimSS 5.36 computes in ("simss"). They differ only where the share is a
#: whole number and the floor lands on either side of it.
GRID_SPLITS = ("exact", "simss")
#: The grids a device can be built on.
GRIDS = ("simss", "graded")
#: The graded rule's points per e-fold of the distance to a layer's end, and the distance (m)
#: below which it stops refining (about V_T / E at the fields ions set up at an interface).
GRADED_POINTS_PER_EFOLD = 6.0
GRADED_SMALLEST = 3e-12

_NEWTON_STEPS = 80


def stack_sum(values) -> float:
    """The stack's total thickness: the correctly rounded sum (Shewchuk's algorithm, as
    `math.fsum`), so that it does not depend on the order of the layers."""
    return math.fsum(float(v) for v in values)


def _as_fraction(t) -> Fraction:
    """A thickness as the decimal it was written as (a float by its shortest repr)."""
    return Fraction(t if isinstance(t, str) else repr(float(t)))


def _round_significand(q: Fraction, bits: int = 64) -> Fraction:
    """``q`` rounded to ``bits`` significant bits, halfway cases to even (x87's default)."""
    if q == 0:
        return q
    mag = abs(q)
    e = mag.numerator.bit_length() - mag.denominator.bit_length()  # 2^e <= mag < 2^(e+2)
    if mag < Fraction(2) ** e:
        e -= 1
    scale = Fraction(2) ** (bits - 1 - e)  # mag * scale in [2^(bits-1), 2^bits)
    return Fraction(round(q * scale)) / scale  # Python's round: halves to even


def layer_points(thicknesses: Sequence[str | float], NP: int, split: str = "exact") -> list[int]:
    """Points the SIMsalabim rule gives each layer beyond its first (see the module's
    docstring). ``split``: one of `GRID_SPLITS`."""
    if split not in GRID_SPLITS:
        raise ValueError(f"split {split!r}: one of {GRID_SPLITS}")
    rnd = _round_significand if split == "simss" else (lambda q: q)

    def running_sum(vals):
        acc = Fraction(0)
        for v in vals:
            acc = rnd(acc + v)
        return acc

    L = [rnd(_as_fraction(t)) for t in thicknesses]
    total = running_sum(L)
    is_thin = [rnd(rnd(NP * v) / total) < MIN_POINTS for v in L]
    rest = rnd(total - running_sum(v for v, thin in zip(L, is_thin) if thin))
    free = NP - MIN_POINTS * sum(is_thin)
    return [max(MIN_POINTS, math.floor(rnd(free * rnd(v / rest)))) for v in L]


def _geometric_unit(m: int, grad: float) -> np.ndarray:
    """SIMsalabim's `Make_Sub_Grid` (unscaled) of a layer with ``m + 1`` spacings: spacing j is
    f(i) = exp(grad (1/c - 1/(i + 1))), c = round(m / 2) (halves to even), for the larger of j
    and m - j that does not exceed c (the two halves mirror; where both exceed c, 1)."""
    j = np.arange(m + 1)
    c = round(0.5 * m)
    lo = np.where(j <= c, j, -1)
    hi = np.where(m - j <= c, m - j, -1)
    i = np.maximum(lo, hi)
    f = np.exp(grad * (1.0 / c - 1.0 / (np.maximum(i, 0) + 1.0)))
    return np.where(i >= 0, f, 1.0)


@dataclass(frozen=True)
class Layout:
    """The discrete part of a grid on NP + 2 points: each layer's first and last point (``i0``,
    ``i1``; consecutive layers share no point, the last ends at NP + 1), each point's layer
    (``lid``), and each layer's spacing before scaling (``unit``, one per point: the interval
    after it) with its sum per layer (``unit_sum``)."""

    NP: int
    i0: tuple[int, ...]
    i1: tuple[int, ...]
    lid: np.ndarray
    unit: np.ndarray
    unit_sum: np.ndarray

    @staticmethod
    def from_counts(counts, NP: int, unit_of) -> "Layout":
        """A layout from the points each layer has beyond its first; ``unit_of(k, m)`` gives
        layer k's ``m + 1`` unscaled spacings (the last layer runs to point NP + 1)."""
        nl = len(counts)
        starts = np.concatenate([[0], np.cumsum(np.asarray(counts[:-1]) + 1)]).astype(int)
        ends = starts + np.asarray(counts, dtype=int)
        ends[-1] = NP + 1
        unit = np.zeros(NP + 2)
        unit_sum = np.zeros(nl)
        lid = np.repeat(np.arange(nl), ends - starts + 1)
        if len(lid) != NP + 2:
            raise ValueError(f"layer counts {list(counts)} do not fill NP = {NP}")
        for k in range(nl):
            u = unit_of(k, int(ends[k] - starts[k]))
            unit[starts[k]:ends[k] + 1] = u
            unit_sum[k] = u.sum()
        i1 = [int(starts[k] + counts[k]) for k in range(nl - 1)] + [NP + 1]
        return Layout(NP=NP, i0=tuple(int(s) for s in starts), i1=tuple(i1), lid=lid,
                      unit=unit, unit_sum=unit_sum)

    def place(self, L, total: float | None = None):
        """``(x, h)`` for layer thicknesses ``L``: ``h`` each point's spacing over the stack's
        thickness ``total`` (default `stack_sum`), ``x`` the positions (m). Each layer's unit
        spacing is scaled to its thickness; across every interface the three spacings that
        meet there (the last two of the layer before and the first of the next) take their
        mean."""
        L = np.asarray(L, dtype=float)
        total = stack_sum(L) if total is None else total
        h = self.unit * (L[self.lid] / (total * self.unit_sum[self.lid]))
        for j in self.i0[1:]:
            h[j - 2:j + 1] = (h[j] + h[j - 1] + h[j - 2]) / 3.0
        x = np.zeros(self.NP + 2)
        np.cumsum(total * h[:-1], out=x[1:])
        return x, h


def simss_layout(thicknesses, NP: int, grad: float, split: str = "exact") -> Layout:
    """SIMsalabim's `Make_Grid` layout."""
    counts = layer_points(thicknesses, NP, split)
    return Layout.from_counts(counts, NP, lambda k, m: _geometric_unit(m, grad))


def _graded_integral(u, Lk, H, beta, d0):
    """The graded density's integral over [0, u] of a layer of thickness Lk, and its
    derivative (the density at u)."""
    F = u / H + beta * (np.log1p(u / d0) - np.log1p((Lk - u) / d0) + np.log1p(Lk / d0))
    dF = 1.0 / H + beta / (d0 + u) + beta / (d0 + Lk - u)
    return F, dF


def _invert(targets, Lk, H, beta, d0):
    """Positions u in [0, Lk] with F(u) = targets: Newton steps kept inside a shrinking
    bracket (a bisection step where Newton would leave it), to full precision."""
    lo, hi = np.zeros_like(targets), np.full_like(targets, Lk)
    u = targets / targets[-1] * Lk if targets[-1] > 0 else lo.copy()
    for _ in range(_NEWTON_STEPS):
        F, dF = _graded_integral(u, Lk, H, beta, d0)
        below = F < targets
        lo, hi = np.where(below, u, lo), np.where(below, hi, u)
        step = u - (F - targets) / dF
        inside = (step > lo) & (step < hi)
        new = np.where(inside, step, 0.5 * (lo + hi))
        if np.array_equal(new, u):
            break
        u = new
    return u


def graded_layout(thicknesses, NP: int, beta: float = GRADED_POINTS_PER_EFOLD,
                  d0: float = GRADED_SMALLEST) -> Layout:
    """The graded rule's layout (see the module's docstring)."""
    L = np.array([float(t) for t in thicknesses])
    H = L.sum() / NP
    weight = np.array([_graded_integral(Lk, Lk, H, beta, d0)[0] for Lk in L])
    nl = len(L)
    counts = np.maximum(MIN_POINTS, np.floor((NP + 2 - nl) * weight / weight.sum()).astype(int))
    over = int(counts.sum()) + nl - (NP + 2)  # points the minimum added beyond NP + 2
    if over > 0:
        counts[np.argmax(counts)] -= over

    def unit_of(k, m):
        targets = weight[k] * np.arange(m + 2) / (m + 1)
        u = _invert(targets, L[k], H, beta, d0)
        u[0], u[-1] = 0.0, L[k]
        return np.diff(u)

    return Layout.from_counts(counts, NP, unit_of)


def make_grid(thicknesses, NP: int, grad: float, split: str = "exact"):
    """``(x, h, i0, i1, lid)`` of SIMsalabim's `Make_Grid` for the thicknesses (m, numbers or
    the input's decimal strings): NP + 2 points, ``h`` the spacing over the stack's thickness,
    ``i0``/``i1`` each layer's first and last point (lists), ``lid`` each point's layer."""
    lay = simss_layout(thicknesses, NP, grad, split)
    x, h = lay.place([float(t) for t in thicknesses])
    return x, h, list(lay.i0), list(lay.i1), lay.lid
