"""The C ABI library: it loads on a machine without a GPU, exports every
entry point include/serenx.h declares, reports a missing device as an error
instead of crashing, and the product package never reaches the oracle."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "serenx.h")
LIB = os.path.join(ROOT, "seren_amd", "libserenx.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(serenx_[a-z0-9_]+)\s*\(", text)))


@pytest.mark.skipif(not os.path.exists(LIB), reason="libserenx.so not built")
def test_library_exports_every_declared_symbol():
    lib = C.CDLL(LIB)
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_lists_the_header():
    from seren_amd import engine
    assert sorted(engine.EXPORTED) == declared_functions()


@pytest.mark.skipif(not os.path.exists(LIB), reason="libserenx.so not built")
def test_create_without_gpu_fails_cleanly():
    import torch
    if torch.cuda.device_count() > 0:
        pytest.skip("a GPU is present")
    from seren_amd import engine
    with pytest.raises(engine.SerenxError):
        engine.ForceEngine()


def test_default_options_and_errors_without_device():
    if not os.path.exists(LIB):
        pytest.skip("libserenx.so not built")
    from seren_amd import engine
    lib = engine.load_library()
    opt = engine.Options()
    assert lib.serenx_default_options(C.byref(opt)) == 0
    assert opt.multipole == engine.QUADRUPOLE and opt.mac == engine.MAC_GEOMETRIC
    assert opt.h_fac == 1.2 and opt.alpha == 1.0
    bad = engine.Options(opt.multipole, 7, 0, 0, 1.2, 0.5, 0.1, 1.0, 1.0)
    ctx = C.c_void_p()
    assert lib.serenx_create(C.byref(bad), 0, C.byref(ctx)) == 1   # SERENX_EINVAL
    assert b"mac" in lib.serenx_last_error()


def test_missing_library_fails_loudly(tmp_path):
    from seren_amd import engine
    saved = engine._lib
    engine._lib = None
    try:
        with pytest.raises(engine.SerenxError):
            engine.load_library(str(tmp_path / "absent.so"))
    finally:
        engine._lib = saved


def test_product_package_does_not_use_the_oracle():
    pkg = os.path.join(ROOT, "seren_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".h", ".cpp", ".F90")):
                text = open(os.path.join(dirpath, f)).read()
                assert "oracle" not in text.lower(), f
