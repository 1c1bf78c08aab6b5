"""The C ABI library loads and exports every function include/ntlp_gpu.h
declares (no GPU needed: nothing is called)."""
import ctypes
import os
import re

from ntlp_amd import LIB_PATH, load_library

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for fn in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names += re.findall(r"\b(ntlp_\w+)\s*\(", text)
    return sorted(set(names))


def test_header_declares_the_entry_points():
    names = declared_functions()
    for must in ("ntlp_gpu_init", "ntlp_gpu_particle_update_rk3", "ntlp_gpu_particle_update_be",
                 "ntlp_gpu_particle_coupling_update", "ntlp_gpu_bin", "ntlp_gpu_step"):
        assert must in names


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB_PATH), "build with python -m ntlp_amd.build"
    lib = ctypes.CDLL(LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_loads():
    lib = load_library()
    assert lib.ntlp_gpu_last_error() is not None
