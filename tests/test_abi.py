"""The C-ABI library loads and exports every function include/*.h declares.
No compute call is made here (this suite also runs without a GPU); only
argument validation that fails before any HIP call."""
import ctypes as C
import re
from pathlib import Path

import pytest

import pbm

ROOT = Path(__file__).resolve().parents[1]
DECL = re.compile(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(pbm_[a-z0-9_]+)\s*\(", re.M)


def declared_functions() -> set[str]:
    names = set()
    for h in sorted((ROOT / "include").glob("*.h")):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names.update(m.group(1) for m in DECL.finditer(text)
                     if not m.group(0).lstrip().startswith(("typedef", "#")))
    return names


def test_header_declares_entry_points():
    names = declared_functions()
    assert {"pbm_create", "pbm_destroy", "pbm_get_pages", "pbm_free_pages"} <= names


def test_library_exports_every_declared_function():
    lib = pbm.lib()
    missing = [n for n in sorted(declared_functions()) if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_every_declared_function():
    assert declared_functions() == set(pbm.exported_names())


def test_version_string():
    assert pbm.lib().pbm_version().decode().startswith("pbm ")


@pytest.mark.parametrize("field,value,code", [
    ("n_pages", 0, -5), ("n_pages", (1 << 30) + 1, -5), ("page_size", 4, -1),
    ("strategy", 99, -1), ("order", 7, -1),
])
def test_create_rejects_bad_config_before_touching_the_gpu(field, value, code):
    cfg = pbm.Config(1024, 4096, pbm.RW_BM, 0, 1, 0, 0, 0, 0)
    setattr(cfg, field, value)
    h = C.c_void_p()
    rc = pbm.lib().pbm_create(C.byref(cfg), C.byref(h))
    assert rc == code
    assert h.value is None
    assert pbm.lib().pbm_last_error().decode()


def test_null_arguments_are_rejected():
    L = pbm.lib()
    assert L.pbm_create(None, None) == -1
    assert L.pbm_destroy(None) == 0
    assert L.pbm_get_ctx(None, None) == -1
    assert L.pbm_reset(None, None) == -1
    assert L.pbm_get_pages(None, 1, 0, None, None, None, None) == -1
