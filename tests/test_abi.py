"""The C ABI library loads and exports every symbol include/*.h declares; the
pybind11 module imports; without a device the engine reports an error
instead of computing anything (CPU only)."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "falconnpp_amd", "libfalconnpp_amd.so")


def _ensure_built():
    if not os.path.exists(LIB):
        subprocess.check_call(["make", "-j", "8", "-C", os.path.join(ROOT, "falconnpp_amd", "csrc")])


def _declared():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", fn)).read()
            names |= set(re.findall(r"\b(fpp_[a-z0-9_]+)\s*\(", txt))
    return names


def test_library_exports_every_declared_symbol():
    _ensure_built()
    from falconnpp_amd import engine
    lib = ctypes.CDLL(LIB)
    declared = _declared()
    assert declared, "no declarations found"
    for name in declared:
        assert hasattr(lib, name), name
    assert declared == set(engine.EXPORTS)


def _no_gpu():
    try:
        import torch
        return not torch.cuda.is_available()
    except Exception:
        return True


@pytest.mark.skipif(not _no_gpu(), reason="a GPU is visible")
def test_no_device_is_an_error_not_a_fallback():
    _ensure_built()
    from falconnpp_amd import engine
    with pytest.raises(engine.FppError) as e:
        engine.Index(100, 8)
    assert e.value.code == engine.FPP_EHIP


def test_pybind_module_imports_with_reference_names():
    _ensure_built()
    import falconnpp_amd
    m = falconnpp_amd.FalconnPP
    cls = m.FalconnPP
    for meth in ("setIndexParam", "set_qProbes", "set_threads", "clear", "build", "query", "build2D", "query2D"):
        assert hasattr(cls, meth), meth
    doc = cls.setIndexParam.__doc__
    for kw in ("n_tables", "n_proj", "bucket_minSize", "bucket_scale", "iProbes", "n_threads", "random_seed"):
        assert kw in doc, kw
    if _no_gpu():
        with pytest.raises(RuntimeError):
            cls(100, 8)


@pytest.mark.parametrize("entry", ["from falconnpp_amd import engine; engine.load()",
                                   "import falconnpp_amd; falconnpp_amd.FalconnPP"])
def test_one_hip_runtime_per_process(entry):
    """Loading the library before torch leaves one HIP runtime mapped (torch's,
    imported first by the loaders); a second copy would see no GPU."""
    _ensure_built()
    pytest.importorskip("torch")
    code = (entry + "\nimport torch\n"
            "maps = {l.split()[-1] for l in open('/proc/self/maps') if 'libamdhip64' in l}\n"
            "print(len(maps))\n")
    out = subprocess.run(["python3", "-c", code], cwd=ROOT, capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    assert out.stdout.split()[-1] == "1", out.stdout
