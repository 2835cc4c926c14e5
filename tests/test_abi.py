"""CPU: the C ABI library loads, exports every function include/gquery_amd.h declares, and the
ctypes mirrors of its structs have the C layout (checked against a C compiler)."""
import ctypes as C
import re
import subprocess
from pathlib import Path

import pytest

from gquery_amd import _lib

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "gquery_amd.h"


def declared_functions():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"\b(gq_\w+)\s*\(", text)) - {"gq_structure", "gq_scene"})


def test_every_declared_function_is_exported():
    lib = _lib.lib()
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_bindings_cover_the_header():
    assert set(declared_functions()) <= set(_lib.SIGNATURES), set(declared_functions()) - set(_lib.SIGNATURES)


def test_struct_layouts_match_c(tmp_path):
    src = tmp_path / "sizes.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "gquery_amd.h"\nint main(void) {\n'
                   '  printf("%zu %zu %zu %zu %zu %zu %zu %zu\\n", sizeof(gq_records), sizeof(gq_scene_options), sizeof(gq_job),\n'
                   '         offsetof(gq_job, n), offsetof(gq_job, out), offsetof(gq_job, traversal),\n'
                   '         sizeof(gq_sample_records), offsetof(gq_sample_records, prim_id));\n  return 0;\n}\n')
    exe = tmp_path / "sizes"
    r = subprocess.run(["gcc", "-I", str(ROOT / "include"), str(src), "-o", str(exe)], capture_output=True, text=True)
    if r.returncode != 0:
        pytest.skip("no C compiler: " + r.stderr[:200])
    got = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    want = [C.sizeof(_lib.Records), C.sizeof(_lib.SceneOptions), C.sizeof(_lib.Job),
            _lib.Job.n.offset, _lib.Job.out.offset, _lib.Job.traversal.offset,
            C.sizeof(_lib.SampleRecords), _lib.SampleRecords.prim_id.offset]
    assert got == want


def test_version_and_error_calls_need_no_device():
    assert _lib.lib().gq_version() > 0
    assert isinstance(_lib.lib().gq_last_error(), bytes)
