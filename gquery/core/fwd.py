"""gquery.core.fwd (gquery/core/fwd.py): the array types the reference's examples build their
inputs with -- Array2 / Array3 (float32) and Array2i / Array3i (int32), components first, and Float.
DrJit itself is not part of this package; records come back as torch tensors."""
from gquery_amd.fwd import Array2, Array2i, Array3, Array3i, Float, Int  # noqa: F401
