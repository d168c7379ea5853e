"""The C ABI: every function declared in include/*.h is exported by
libprr.so, and the ctypes structures match the C layout (checked by compiling
a probe against the header with gcc).  No GPU needed.
"""
import ctypes as C
import os
import re
import subprocess

import pytest

from parrelray_amd import _abi as A
from parrelray_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", h) for h in ("prr.h", "prr_debug.h")]


def declared_functions():
    names = []
    for h in HEADERS:
        text = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        names += re.findall(r"\b(prr_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_every_declared_function_is_exported():
    lib = C.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    assert len(declared_functions()) >= 40


STRUCTS = {"prr_tracer": A.Tracer, "prr_sphere": A.Sphere, "prr_frame": A.Frame, "prr_image": A.Image,
           "prr_ray_out": A.RayOut, "prr_launch": A.Launch, "prr_stats": A.Stats, "prr_orbit": A.Orbit}


def test_ctypes_layout_matches_the_header(tmp_path):
    src = ['#include <stdio.h>', '#include <stddef.h>', '#include "prr.h"', "int main(void) {"]
    for cname, cls in STRUCTS.items():
        src.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for field, _ in cls._fields_:
            src.append(f'printf("{cname}.{field} %zu\\n", offsetof({cname}, {field}));')
    src.append("return 0; }")
    c = tmp_path / "probe.c"
    c.write_text("\n".join(src))
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(c), "-o", str(exe)], check=True)
    got = dict(line.split() for line in subprocess.run([str(exe)], capture_output=True, text=True,
                                                       check=True).stdout.splitlines())
    for cname, cls in STRUCTS.items():
        assert int(got[cname]) == C.sizeof(cls), cname
        for field, _ in cls._fields_:
            assert int(got[f"{cname}.{field}"]) == getattr(cls, field).offset, (cname, field)


def test_enum_values_match_the_header():
    text = open(HEADERS[0]).read()
    for name, value in [("PRR_KERR_BL", A.PRR_KERR_BL), ("PRR_RKF45", A.PRR_RKF45), ("PRR_FP32_BULK", A.PRR_FP32_BULK),
                        ("PRR_SPHERE", A.PRR_SPHERE), ("PRR_NAN", A.PRR_NAN),
                        ("PRR_FILTER_BILINEAR", A.PRR_FILTER_BILINEAR), ("PRR_SSAA_RANDOM", A.PRR_SSAA_RANDOM),
                        ("PRR_SSAA_JITTERED", A.PRR_SSAA_JITTERED), ("PRR_RK4", A.PRR_RK4),
                        ("PRR_DORMAND", A.PRR_DORMAND)]:
        m = re.search(rf"\b{name}\s*=\s*(\d+)", text)
        if m:
            assert int(m.group(1)) == value, name
        else:  # implicit value: position in its enum
            body = re.search(rf"\{{([^}}]*\b{name}\b[^}}]*)\}}", text).group(1)
            items = [s.split("=")[0].strip() for s in body.split(",") if s.strip()]
            items = [re.sub(r"/\*.*?\*/", "", s).strip() for s in items]
            assert items.index(name) == value, name


def test_no_context_calls_fail_cleanly():
    lib = _lib.lib()
    assert lib.prr_abi_version() == A.PRR_ABI_VERSION
    with pytest.raises(_lib.PrrError):
        _lib.check(lib.prr_set_tracer(None, None))


def test_product_never_touches_the_oracle():
    """The package and libprr.so neither import, link nor name the oracle
    (test infrastructure), and the library links no CPU fallback: importing
    the package and every module in it leaves oracle unimported, and the
    shared library's dynamic dependencies do not include liboracle."""
    import importlib
    import pkgutil
    import sys

    import parrelray_amd
    before = "oracle" in sys.modules
    for m in pkgutil.iter_modules(parrelray_amd.__path__):
        if m.name == "autograd":
            pytest.importorskip("torch")
        importlib.import_module("parrelray_amd." + m.name)
    if not before:
        assert "oracle" not in sys.modules
    pkg = os.path.join(ROOT, "parrelray_amd")
    for d, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".cpp", ".hpp", ".hip", ".h")):
                text = open(os.path.join(d, f), errors="replace").read()
                assert not re.search(r"\boracle\b|liboracle|orc_", text), os.path.join(d, f)
    deps = subprocess.run(["ldd", _lib.LIB_PATH], capture_output=True, text=True).stdout
    assert "oracle" not in deps
