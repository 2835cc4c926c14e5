"""Query records (gquery/shapes/primitive.py:9-84) as dataclasses of torch tensors.

Field names and meaning match the reference; vectors are (N, dim) tensors instead of DrJit
Array2/Array3.  valid / on_boundary are bool, prim_id int32, everything else float32.
A miss is all zero, as dr.zeros(record) is in the reference traversals.
"""
from __future__ import annotations

from dataclasses import dataclass

import torch


@dataclass
class Intersection:
    valid: torch.Tensor
    p: torch.Tensor
    n: torch.Tensor
    t: torch.Tensor
    d: torch.Tensor
    prim_id: torch.Tensor
    on_boundary: torch.Tensor


@dataclass
class Intersection3D:
    valid: torch.Tensor
    p: torch.Tensor
    n: torch.Tensor
    uv: torch.Tensor
    d: torch.Tensor
    prim_id: torch.Tensor
    on_boundary: torch.Tensor


@dataclass
class ClosestPointRecord:
    valid: torch.Tensor
    p: torch.Tensor
    n: torch.Tensor
    t: torch.Tensor
    d: torch.Tensor
    prim_id: torch.Tensor


@dataclass
class ClosestPointRecord3D:
    valid: torch.Tensor
    p: torch.Tensor
    n: torch.Tensor
    uv: torch.Tensor
    d: torch.Tensor
    prim_id: torch.Tensor


@dataclass
class BoundarySamplingRecord:
    p: torch.Tensor
    n: torch.Tensor
    t: torch.Tensor
    pdf: torch.Tensor
    prim_id: torch.Tensor
    pmf: torch.Tensor


@dataclass
class BoundarySamplingRecord3D:
    p: torch.Tensor
    n: torch.Tensor
    uv: torch.Tensor
    pdf: torch.Tensor
    prim_id: torch.Tensor
    pmf: torch.Tensor


@dataclass
class ClosestSilhouettePointRecord:
    valid: torch.Tensor
    p: torch.Tensor
    d: torch.Tensor
    prim_id: torch.Tensor


@dataclass
class ClosestSilhouettePointRecord3D:
    valid: torch.Tensor
    p: torch.Tensor
    d: torch.Tensor
    prim_id: torch.Tensor
