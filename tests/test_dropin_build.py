"""The C++ boundary as a drop-in (CPU): the engine's header compiles beside the reference's
headers, and the reference's own Python module builds over the engine."""
import os
import subprocess
import sys
import sysconfig

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.environ.get("VS_REFERENCE", "/root/reference")
DROPIN_DIR = os.path.join(ROOT, "oracle", "_ref", "dropin")
needs_ref = pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "include", "vecsearch")),
                               reason="the reference sources are not present")


@needs_ref
def test_header_compiles_beside_reference_bindings():
    """python/bindings.cpp of the reference (with its GPU class switched on, bindings.cpp:12-14,
    151-197) compiles against include/vecsearch/gpu_brute_force.h: Metric comes from the
    reference's distance.h, not a second definition."""
    import pybind11

    cmd = ["g++", "-std=gnu++17", "-fsyntax-only", "-DVECSEARCH_HAS_CUDA=1", "-I", os.path.join(ROOT, "include"),
           "-I", os.path.join(REF, "include"), "-I", pybind11.get_include(), "-I", sysconfig.get_paths()["include"],
           os.path.join(REF, "python", "bindings.cpp")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]


def test_header_compiles_standalone():
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-I", os.path.join(ROOT, "include"),
                        os.path.join(ROOT, "vecsearch_amd", "csrc", "gpu_brute_force.cpp")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]


def test_reference_module_loads_over_engine():
    mods = [f for f in os.listdir(DROPIN_DIR) if f.startswith("_vecsearch")] if os.path.isdir(DROPIN_DIR) else []
    if not mods:
        pytest.skip("oracle/_ref/dropin was not built (needs the reference sources)")
    code = ("import importlib.util,sys;"
            "s=importlib.util.spec_from_file_location('_vecsearch',sys.argv[1]);"
            "m=importlib.util.module_from_spec(s);s.loader.exec_module(m);"
            "print(sorted(n for n in dir(m) if not n.startswith('_')));"
            "print(m.GpuBruteForceIndex.search.__doc__.splitlines()[0]);"
            "print(type(m.cuda_available()).__name__)")
    r = subprocess.run([sys.executable, "-c", code, os.path.join(DROPIN_DIR, mods[0])], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    names, sig, avail = r.stdout.splitlines()[:3]
    assert "GpuBruteForceIndex" in names and "BruteForceIndex" in names
    assert "k: " in sig and "method: str = 'auto'" in sig
    assert avail == "bool"
