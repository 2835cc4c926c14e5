"""GPU: gq_closest_point_bvh_noalloc (scratch in a fixed device array, no allocation) against
gq_closest_point_bvh on the same inputs, through the C ABI: bit-identical records on two calls in a
row, the second with a smaller batch (the array keeps the first call's contents), for both engine
node formats in 3D and 2D; a batch beyond the array's capacity is refused with a clear error."""
import ctypes as C

import numpy as np
import pytest
import torch

from gquery_amd import _lib, synth

pytestmark = pytest.mark.gpu

FIELDS = ["valid", "p", "n", "uv", "d", "prim_id"]


def _call(fn, shape, method, x):
    sc = shape.scene
    out, r = sc._alloc(x.shape[0], FIELDS)
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = fn(sc._h, _lib.METHODS[method], C.c_void_p(x.data_ptr()), x.shape[0], C.byref(r), stream)
    return rc, out


@pytest.mark.parametrize("dim", [3, 2])
@pytest.mark.parametrize("method", ["engine", "engine_fp32"])
def test_noalloc_equals_closest_point_bvh(gpu, dim, method):
    from gquery_amd.shapes.mesh import Mesh
    from gquery_amd.shapes.polyline import Polyline
    if dim == 3:
        V, F = synth.displaced_icosphere(6)
        shape = Mesh(V, F, device="cuda:0")
    else:
        V, F = synth.star_polyline(65536)
        shape = Polyline(V, F, device="cuda:0")
    L = _lib.lib()
    for k, n in enumerate((1 << 20, 20000)):   # the second, smaller call reuses the array's old contents
        x = torch.from_numpy(synth.uniform_points(n, dim, seed=70 + k)).cuda()
        print(f"[noalloc] dim={dim} method={method} n={n}", flush=True)
        rc, got = _call(L.gq_closest_point_bvh_noalloc, shape, method, x)
        assert rc == 0, _lib.lib().gq_last_error()
        rc, ref = _call(L.gq_closest_point_bvh, shape, method, x)
        assert rc == 0
        torch.cuda.synchronize()
        for f in FIELDS:
            assert torch.equal(got[f].view(torch.uint8), ref[f].view(torch.uint8)), (n, f)
        assert got["valid"].all().item()


def test_noalloc_refuses_oversized_batch(gpu):
    from gquery_amd.shapes.mesh import Mesh
    V, F = synth.displaced_icosphere(3)
    shape = Mesh(V, F, device="cuda:0")
    L = _lib.lib()
    x = torch.zeros((3 << 20, 3), device="cuda:0")
    rc, _ = _call(L.gq_closest_point_bvh_noalloc, shape, "engine", x)
    assert rc == _lib.GQ_ERR_INVALID
    assert b"scratch" in L.gq_last_error()
    # the plain method is not offered without allocation
    rc, _ = _call(L.gq_closest_point_bvh_noalloc, shape, "plain", x[:10])
    assert rc == _lib.GQ_ERR_UNSUPPORTED
