"""The C-ABI library loads on a CPU-only machine and exports every entry
point include/born_hip.h declares (no compute: there is no GPU here)."""
import ctypes
import os
import re

from conftest import PKG, ROOT


def declared_functions():
    text = open(os.path.join(ROOT, "include", "born_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(born_\w+)\s*\(", text)))


def test_header_declares_the_reference_interface():
    names = declared_functions()
    # gpu_funcs_3d.h:4-37 counterparts
    for n in ["born_create", "born_create_space", "born_set_ntblock",
              "born_transfer_sinc_table_s", "born_transfer_sinc_table_d",
              "born_transfer_source_func", "born_transfer_vel_func1",
              "born_transfer_vel_func2", "born_transfer_receiver_func",
              "born_rtm_forward", "born_source_prop", "born_rtm_adjoint"]:
        assert n in names


def test_library_exports_every_declared_symbol(pyborn):
    lib = ctypes.CDLL(os.path.join(PKG, "libborn_hip.so"))
    for name in declared_functions():
        assert hasattr(lib, name), name
        assert name in pyborn.SIGNATURES, name
    assert set(pyborn.SIGNATURES) == set(declared_functions())


def test_binding_loads_without_a_gpu(pyborn):
    lib = pyborn.load_library()
    h = ctypes.c_void_p()
    # without a device the context cannot be created: a clean error code,
    # never a crash or a silent fallback
    rc = lib.born_create(ctypes.byref(h), 0)
    if rc != 0:
        assert rc in (-1, -3)
    else:  # a GPU is present
        assert lib.born_destroy(h) == 0
