"""Scene files: a structure written with save() loads back identical, without the builder, and a
damaged file is rejected."""
import numpy as np
import pytest

from gquery_amd import ext, synth

ARRAYS = ("node_data", "bvh_node_data", "primitive_data")


@pytest.mark.parametrize("cls,shape", [(ext.SNCH3D, "ico"), (ext.SNCH, "star"), (ext.BVH3D, "ico"), (ext.BVH, "star")])
def test_round_trip(tmp_path, cls, shape):
    V, F = synth.displaced_icosphere(3, seed=2) if shape == "ico" else synth.star_polyline(2000, seed=2)
    s = cls(V, F)
    path = tmp_path / "scene.gqs"
    s.save(path)
    t = ext.load(path)
    assert type(t) is cls
    assert t.info() == s.info()
    for a in ARRAYS + (("silhouette_data", "raw_silhouette_data") if cls in (ext.SNCH, ext.SNCH3D) else ()):
        np.testing.assert_array_equal(getattr(t, a)(), getattr(s, a)())
    np.testing.assert_array_equal(t._vertices, V)
    np.testing.assert_array_equal(t._indices, F)
    for w in range(11):   # the engine layout built from the loaded structure is the same
        if w in (2, 3, 6, 7, 8) and cls in (ext.BVH, ext.BVH3D):
            continue
        np.testing.assert_array_equal(t.engine_layout(w), s.engine_layout(w))


def test_damaged_files_are_rejected(tmp_path):
    V, F = synth.displaced_icosphere(2, seed=3)
    path = tmp_path / "scene.gqs"
    ext.SNCH3D(V, F).save(path)
    data = bytearray(path.read_bytes())
    flipped = bytearray(data)
    flipped[len(data) // 2] ^= 0x40                 # one bit in the middle: checksum mismatch
    (tmp_path / "flip.gqs").write_bytes(bytes(flipped))
    (tmp_path / "short.gqs").write_bytes(bytes(data[: len(data) - 100]))
    (tmp_path / "magic.gqs").write_bytes(b"XXXXXXXX" + bytes(data[8:]))
    for name in ("flip.gqs", "short.gqs", "magic.gqs", "missing.gqs"):
        with pytest.raises(ValueError):
            ext.load(tmp_path / name)


def _fnv1a64(b: bytes) -> int:
    h = 0xCBF29CE484222325
    for x in b:
        h = ((h ^ x) * 0x100000001B3) & 0xFFFFFFFFFFFFFFFF
    return h


def _resealed(data: bytearray) -> bytes:
    """The file with its closing checksum recomputed over the edited bytes."""
    body = bytes(data[:-8])
    return body + _fnv1a64(body).to_bytes(8, "little")


@pytest.mark.parametrize("dim", [2, 3])
def test_consistent_checksum_but_bad_offsets_are_rejected(tmp_path, dim):
    """A file whose checksum is valid but whose tree indexes outside its arrays (or whose stored
    depth is stale) is rejected before any kernel could read it."""
    V, F = synth.displaced_icosphere(2, seed=3) if dim == 3 else synth.star_polyline(300, seed=3)
    s = (ext.SNCH3D if dim == 3 else ext.SNCH)(V, F)
    path = tmp_path / "scene.gqs"
    s.save(path)
    data = bytearray(path.read_bytes())
    stride = 2 * dim + 3
    nodes_at = 8 + 64 + 8            # magic, header, element count of bvh_node_data
    node = lambda i, f: nodes_at + 4 * (i * stride + f)
    assert (tmp_path / "ok.gqs").write_bytes(_resealed(bytearray(data))) and ext.load(tmp_path / "ok.gqs")
    cases = {}
    d = bytearray(data)
    d[node(0, 2 * dim + 1): node(0, 2 * dim + 1) + 4] = np.float32(1e6).tobytes()      # root's second child
    cases["child"] = d
    leaf = next(i for i in range(len(s.bvh_node_data()) // stride) if s.bvh_node_data()[i * stride + 2 * dim + 2] > 0)
    d = bytearray(data)
    d[node(leaf, 2 * dim): node(leaf, 2 * dim) + 4] = np.float32(F.shape[0]).tobytes()    # leaf past the primitives
    cases["leaf"] = d
    d = bytearray(data)
    d[node(1, 2 * dim + 1): node(1, 2 * dim + 1) + 4] = np.float32(0.5).tobytes()        # non-integer offset
    cases["fraction"] = d
    d = bytearray(data)
    depth = int.from_bytes(d[8 + 48: 8 + 56], "little", signed=True)
    d[8 + 48: 8 + 56] = (depth - 1).to_bytes(8, "little", signed=True)                   # stale max_depth
    cases["depth"] = d
    for name, d in cases.items():
        p = tmp_path / f"{name}.gqs"
        p.write_bytes(_resealed(d))
        with pytest.raises(ValueError, match="inconsistent"):
            ext.load(p)
