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


def test_fortran_module_binds_every_entry_point_a_time_loop_calls():
    """Every C entry point has a bind(c) interface in the Fortran module,
    except the serving loop and its pinned host buffers, which serve callers
    other than the Fortran time loop."""
    text = open(os.path.join(ROOT, "fortran", "ntlp_gpu_mod.f90")).read()
    bound = set(re.findall(r'bind\(c,\s*name="(ntlp_\w+)"\)', text))
    unbound = [n for n in declared_functions()
               if n not in bound and not n.startswith("ntlp_serve_") and n not in ("ntlp_host_alloc", "ntlp_host_free")]
    assert not unbound, unbound
    assert not bound - set(declared_functions())


def test_python_binding_loads():
    lib = load_library()
    assert lib.ntlp_gpu_last_error() is not None


# ------------------------------------------------- argument checks, no GPU
# ntlp_gpu_init checks every argument before it touches the device, and the
# other entry points check the context's state first, so these errors come
# back on a machine without a GPU.

def _init(lib, grid, consts=None, flags=None, null=None, tweak=None):
    from ntlp_amd import Consts, Flags, _CGrid
    cg = grid.to_c(_CGrid)
    if tweak:
        tweak(cg)
    cc = (consts or Consts()).to_c()
    cf = (flags or Flags()).to_c()
    args = [ctypes.byref(cg), ctypes.byref(cc), ctypes.byref(cf)]
    if null is not None:
        args[null] = None
    return lib.ntlp_gpu_init(*args)


def _uniform(nnx=8, nny=8, nnz=8):
    from ntlp_amd.grid import Grid
    return Grid.uniform(nnx, nny, nnz, 0.05, 0.05, 0.05)


def test_init_rejects_bad_arguments_before_touching_the_device():
    import numpy as np
    from ntlp_amd import Flags, NTLP_ERRORS
    code = {v: k for k, v in NTLP_ERRORS.items()}
    lib = load_library()
    assert lib.ntlp_gpu_finalize() == 0  # no context: nothing to release
    cases = []
    cases.append(("null consts", _init(lib, _uniform(), null=1), "NTLP_E_INVALID"))
    cases.append(("nnx = 1", _init(lib, _uniform(nnx=1)), "NTLP_E_INVALID"))
    cases.append(("nnz = 2", _init(lib, _uniform(nnz=2)), "NTLP_E_INVALID"))
    cases.append(("nnz + 2 > 1024", _init(lib, _uniform(nnz=1023)), "NTLP_E_RANGE"))
    g = _uniform()
    g.xl = 0.0  # dx = 0
    cases.append(("dx = 0", _init(lib, g), "NTLP_E_INVALID"))
    cases.append(("no z levels", _init(lib, _uniform(), tweak=lambda c: setattr(c, "z", None)), "NTLP_E_INVALID"))
    cases.append(("icase = 1", _init(lib, _uniform(), flags=Flags(icase=1)), "NTLP_E_UNSUPPORTED"))
    cases.append(("2^31 cell keys", _init(lib, _uniform(2048, 2048, 600)), "NTLP_E_RANGE"))
    cases.append(("ext arrays above 4 GB", _init(lib, _uniform(1024, 1024, 600)), "NTLP_E_RANGE"))
    g = _uniform()
    g.z = np.array(g.z)
    g.z[3] = g.z[2]  # not strictly increasing
    cases.append(("z not increasing", _init(lib, g), "NTLP_E_INVALID"))
    for what, rc, want in cases:
        assert rc == code[want], (what, rc, lib.ntlp_gpu_last_error().decode())
        assert lib.ntlp_gpu_last_error().decode(), what


def test_calls_without_a_context_report_not_init():
    from ntlp_amd import NTLP_ERRORS
    code = {v: k for k, v in NTLP_ERRORS.items()}
    lib = load_library()
    assert lib.ntlp_gpu_finalize() == 0
    assert lib.ntlp_gpu_step(2, 1, 1, 2) == code["NTLP_E_NOT_INIT"]
    assert lib.ntlp_gpu_particle_update_be(2) == code["NTLP_E_NOT_INIT"]
    assert lib.ntlp_gpu_particle_coupling_update(None, 0) == code["NTLP_E_NOT_INIT"]
    assert lib.ntlp_gpu_particles_put(None, None, None, None, 0) == code["NTLP_E_NOT_INIT"]
    assert lib.ntlp_gpu_particles_count() == 0
    assert "ntlp_gpu_init" in lib.ntlp_gpu_last_error().decode()
