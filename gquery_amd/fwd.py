"""Array types with the names of gquery/core/fwd.py (Array2, Array3, Array2i, Array3i, Float), for
callers written against the reference.

The reference's arrays are DrJit structure-of-arrays: Array3(np_array.T) holds 3 components of N
entries, shape (3, N) (examples/example_3d.py:23,27).  These classes keep that meaning explicit, so
(D, N) input is read as components first even when N == D (a one-triangle mesh's vertices are
(3, 3)).  Untagged numpy / torch input to the library is (N, D) rows; it is read as (D, N) only
when that is unambiguous (first dimension D, second not D).

    Array3(np_array.T)           # (3, N) components, as in the reference examples
    Array3(xs, ys, zs)           # three component arrays of N entries
    Array3i(faces.T)             # int32 indices, components first
    Float(float("inf"))          # a scalar r_max (the examples' Float(dr.inf))
"""
from __future__ import annotations

import numpy as np

__all__ = ["Array2", "Array3", "Array2i", "Array3i", "Float", "Int", "SoA"]


class SoA:
    """D components of N entries (components first), float32 or int32; numpy or torch storage."""
    dim = 0
    dtype = np.float32

    def __init__(self, *args):
        if len(args) == self.dim and self.dim > 1:
            comps = [np.asarray(_host(a), dtype=self.dtype).reshape(-1) for a in args]
            if len({c.size for c in comps}) != 1:
                raise ValueError(f"{type(self).__name__}: components of different lengths")
            self.value = np.stack(comps)
            return
        if len(args) != 1:
            raise TypeError(f"{type(self).__name__} takes one (D, N) array or {self.dim} component arrays")
        a = args[0]
        if isinstance(a, SoA):
            a = a.value
        try:
            import torch
            if isinstance(a, torch.Tensor):
                t = a.to(_torch_dtype(self.dtype))
                if t.ndim == 1 and t.shape[0] == self.dim:
                    t = t.reshape(self.dim, 1)
                if t.ndim != 2 or t.shape[0] != self.dim:
                    raise ValueError(f"{type(self).__name__}: expected ({self.dim}, N), got {tuple(t.shape)}")
                self.value = t
                return
        except ImportError:  # pragma: no cover
            pass
        v = np.asarray(a, dtype=self.dtype)
        if v.ndim == 1 and v.shape[0] == self.dim:
            v = v.reshape(self.dim, 1)
        if v.ndim != 2 or v.shape[0] != self.dim:
            raise ValueError(f"{type(self).__name__}: expected ({self.dim}, N), got {v.shape}")
        self.value = v

    def rows(self):
        """The (N, D) row layout the library works in (a transposed copy)."""
        v = self.value
        if isinstance(v, np.ndarray):
            return np.ascontiguousarray(v.T)
        return v.T.contiguous()

    def __len__(self):
        return int(self.value.shape[1])

    @property
    def x(self):
        return self.value[0]

    @property
    def y(self):
        return self.value[1]

    @property
    def z(self):
        return self.value[2]

    def __repr__(self):
        return f"{type(self).__name__}({self.value!r})"


def _host(a):
    try:
        import torch
        if isinstance(a, torch.Tensor):
            return a.detach().cpu().numpy()
    except ImportError:  # pragma: no cover
        pass
    return a


def _torch_dtype(dt):
    import torch
    return torch.float32 if dt == np.float32 else torch.int32


class Array2(SoA):
    dim, dtype = 2, np.float32


class Array3(SoA):
    dim, dtype = 3, np.float32


class Array2i(SoA):
    dim, dtype = 2, np.int32


class Array3i(SoA):
    dim, dtype = 3, np.int32


def Float(v=0.0):
    """A float32 scalar (returned as a Python float) or a float32 array of per-query values."""
    a = np.asarray(_host(v), dtype=np.float32)
    return float(a) if a.ndim == 0 else a


def Int(v=0):
    """An int32 scalar (Python int) or int32 array."""
    a = np.asarray(_host(v), dtype=np.int32)
    return int(a) if a.ndim == 0 else a
