"""The batched block-tridiagonal sweep as an operator (reference: src/jaxdd/sweep.py:313-330
`block_sweep`, the custom_vmap over src/jaxdd/solver.py:1534-1560 `_block_thomas`).

``block_sweep(A, B, C, d)`` solves ``A[i] x[i-1] + B[i] x[i] + C[i] x[i+1] = d[i]`` for a batch
of systems, one lane of the HIP kernel per system, blocks b x b (3 or 5) with k <= 8
right-hand sides. Inputs are torch tensors on the GPU shaped as the reference's batched
operands, ``(n, NP, b, b)`` and ``(n, NP, b, k)``; they are laid out lane-innermost for the
kernel (the reference's Pallas kernel transposes the same way, src/jaxdd/sweep.py:258-284).
"""

from __future__ import annotations

import torch

from . import _lib

__all__ = ["block_sweep"]


def _lane_inner(t: torch.Tensor) -> torch.Tensor:
    n, NP = t.shape[:2]
    return t.reshape(n, NP, -1).permute(1, 2, 0).contiguous()


def block_sweep(A, B, C, d, precision: str = "fp64") -> torch.Tensor:
    """x with the shape of ``d``. ``precision`` "fp64" (float64 throughout) or "mixed"
    (float32 blocks and right-hand sides)."""
    if A.dim() == 3:  # one system: the reference's unbatched call
        return block_sweep(A[None], B[None], C[None], d[None], precision)[0]
    n, NP, b, _ = A.shape
    k = d.shape[-1]
    if A.device.type != "cuda":
        raise RuntimeError("block_sweep runs on the GPU")
    if tuple(B.shape) != (n, NP, b, b) or tuple(C.shape) != (n, NP, b, b) or tuple(d.shape[:3]) != (n, NP, b):
        raise ValueError("block_sweep: A, B, C (n, NP, b, b) and d (n, NP, b, k)")
    dtype = torch.float64 if precision == "fp64" else torch.float32
    lib = _lib.load()
    Al, Bl, Cl, dl = (_lane_inner(t.to(dtype)) for t in (A, B, C, d))
    ws = torch.empty(NP * b * b * n, dtype=dtype, device=A.device)
    stream = torch.cuda.current_stream(A.device).cuda_stream
    rc = lib.ddh_block_sweep(n, NP, b, k, Al.data_ptr(), Bl.data_ptr(), Cl.data_ptr(),
                             dl.data_ptr(), _lib.PRECISIONS[precision], ws.data_ptr(), stream)
    _lib.check(rc, "ddh_block_sweep")
    x = dl.permute(2, 0, 1).reshape(n, NP, b, k)
    return x.to(d.dtype)
