"""PCG32 random number generators on the device: the `sampler` argument of sample_boundary.

gQuery's BVH3D.sample_boundary (gquery/shapes/bvh3d.py:246-266) takes a DrJit PCG32
(drjit.auto.ad.PCG32, imported at gquery/core/fwd.py:5) with one generator per query lane.  This
class holds the same two 64-bit fields per lane -- `state` and the odd stream increment `inc` --
as int64 tensors (the uint64 bits) on the scene's device; the sampling kernel advances `state` in
place, as drawing from the DrJit object does.

The generator is O'Neill's pcg32 (XSH RR, 64-bit state, 32-bit output):
    state' = state * 6364136223846793005 + inc  (mod 2^64)
    out    = rotr32(uint32(((old >> 18) ^ old) >> 27), old >> 59)
    next_float32 = float with bits (out >> 9) | 0x3f800000, minus 1      -> [0, 1)
seed(initstate, initseq): inc = (initseq << 1) | 1, state = 0, one step, state += initstate, one
step.  A generator of `size` lanes seeds lane i with stream initseq + i (DrJit's PCG32(size)).
"""
from __future__ import annotations

import numpy as np
import torch

DEFAULT_STATE = 0x853C49E6748FEA9B
DEFAULT_STREAM = 0xDA3E39CB94B95BDB
MULT = 6364136223846793005
_M64 = (1 << 64) - 1


def _step(state: np.ndarray, inc: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        return state * np.uint64(MULT) + inc


def _srl(x: torch.Tensor, s: int) -> torch.Tensor:
    """Logical right shift of the uint64 bits held in an int64 tensor."""
    return (x >> s) & ((1 << (64 - s)) - 1)


def seed_arrays(size: int, initstate: int = DEFAULT_STATE, initseq: int = DEFAULT_STREAM):
    """(state, inc) uint64 numpy arrays of `size` lanes, lane i on stream initseq + i."""
    with np.errstate(over="ignore"):
        seq = np.uint64(initseq & _M64) + np.arange(size, dtype=np.uint64)
        inc = (seq << np.uint64(1)) | np.uint64(1)
        state = _step(np.zeros(size, np.uint64), inc)
        state = _step(state + np.uint64(initstate & _M64), inc)
    return state, inc


class PCG32:
    """`size` PCG32 generators on `device` (DrJit's PCG32(size, initstate, initseq))."""

    def __init__(self, size: int = 1, initstate: int = DEFAULT_STATE, initseq: int = DEFAULT_STREAM,
                 device=None):
        self.device = torch.device(device) if device is not None else torch.device(
            "cuda", torch.cuda.current_device())
        self.seed(size, initstate, initseq)

    def seed(self, size: int, initstate: int = DEFAULT_STATE, initseq: int = DEFAULT_STREAM) -> None:
        state, inc = seed_arrays(int(size), initstate, initseq)
        self.state = torch.from_numpy(state.view(np.int64)).to(self.device)
        self.inc = torch.from_numpy(inc.view(np.int64)).to(self.device)

    def __len__(self) -> int:
        return int(self.state.numel())

    def next_uint32(self) -> torch.Tensor:
        """One output per lane (int64 tensor holding the uint32 value), computed where the state
        lives with wrapping int64 arithmetic; advances every lane."""
        old = self.state
        self.state = old * MULT + self.inc                      # mod 2^64: int64 arithmetic wraps
        xs = _srl(_srl(old, 18) ^ old, 27) & 0xFFFFFFFF
        rot = _srl(old, 59)
        return ((xs >> rot) | (xs << ((32 - rot) & 31))) & 0xFFFFFFFF

    def next_float32(self) -> torch.Tensor:
        """One uniform float32 in [0, 1) per lane; advances every lane."""
        bits = (self.next_uint32() >> 9) | 0x3F800000
        return bits.to(torch.int32).view(torch.float32) - 1.0
