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
