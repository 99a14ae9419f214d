"""CPU checks of the C ABI surface: the engine library loads, exports every entry point that
include/mpm.h declares, and its host-only paths (defaults, argument validation, status strings) behave.
No call here touches a GPU."""
import ctypes as C
import os
import re

import pytest

import mpm_abi as A

HEADER = os.path.join(A.REPO, "include", "mpm.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mpm_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_the_drop_in_surface():
    names = declared_functions()
    for must in ("mpm_create", "mpm_destroy", "mpm_add_plate", "mpm_add_object", "mpm_set_particles", "mpm_step",
                 "mpm_get_particles", "mpm_get_grid", "mpm_get_plates", "mpm_config_defaults"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(A.PRODUCT_LIB)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, f"libmpm_hip.so lacks {missing}"


def test_every_entry_point_cites_the_reference():
    text = open(HEADER).read()
    blocks = re.findall(r"/\*((?:(?!\*/).)*)\*/\s*\n[^\n]*\b(mpm_(?:create|destroy|add_plate|add_object|set_particles|step|"
                        r"get_particles|get_grid|get_plates|config_defaults|num_particles))\s*\(", text, flags=re.S)
    cited = {name for doc, name in blocks if re.search(r"\.(?:cpp|hpp):\d+", doc)}
    for name in ("mpm_create", "mpm_add_plate", "mpm_add_object", "mpm_set_particles", "mpm_step", "mpm_get_particles",
                 "mpm_get_grid", "mpm_get_plates", "mpm_config_defaults"):
        assert name in cited, f"{name} has no reference file:line citation"


def test_config_defaults_match_reference_members():
    """mpm_config_defaults fills simulation.hpp:28-106's defaults; also checks the ctypes layout."""
    lib = A.product_library()
    c = A.Config()
    lib.lib.mpm_config_defaults(C.byref(c))
    ref = A.default_config()
    for name, _ in A.Config._fields_:
        a, b = getattr(c, name), getattr(ref, name)
        if hasattr(a, "__len__"):
            assert list(a) == list(b), name
        else:
            assert a == b, name


def test_invalid_arguments_are_rejected_without_a_device():
    lib = A.product_library().lib
    h = C.c_void_p()
    assert lib.mpm_create(None, C.byref(h)) == A.ERR["INVALID_ARGUMENT"]
    cfg = A.default_config(dx=0.01, particle_volume=1e-6, particle_mass=1e-3)
    cfg.dim = 4
    assert lib.mpm_create(C.byref(cfg), C.byref(h)) == A.ERR["INVALID_ARGUMENT"]
    cfg.dim = 3
    cfg.pbc, cfg.Lx = 1, 0.0  # a periodic run needs its interval [0, Lx)
    assert lib.mpm_create(C.byref(cfg), C.byref(h)) == A.ERR["INVALID_ARGUMENT"]
    cfg.pbc, cfg.Lx = 0, 1.0
    cfg.plastic_model = A.PLASTIC["Snow"]
    assert lib.mpm_create(C.byref(cfg), C.byref(h)) == A.ERR["UNSUPPORTED"]
    cfg.plastic_model = A.PLASTIC["DP"]
    cfg.elastic_model = A.ELASTIC["NeoHookean"]  # simulation.cpp:80-84 refuses plasticity without Hencky
    assert lib.mpm_create(C.byref(cfg), C.byref(h)) == A.ERR["UNSUPPORTED"]
    cfg.elastic_model = 0
    cfg.spline_degree = 1
    assert lib.mpm_create(C.byref(cfg), C.byref(h)) == A.ERR["UNSUPPORTED"]
    assert lib.mpm_step(None, 1, 0, None) == A.ERR["INVALID_ARGUMENT"]
    assert lib.mpm_num_particles(None) == 0


def test_status_strings_and_version():
    L = A.product_library().lib
    assert L.mpm_status_string(0) == b"ok"
    assert b"Newton" in L.mpm_status_string(A.ERR["NEWTON_MAXITER"])
    assert L.mpm_abi_version() == 2


def test_ctypes_struct_sizes_match_c_layout():
    # mpm_config: 16 int32 + 3+3+6+3+3+4+3+2+2+2+5 doubles
    assert C.sizeof(A.Config) == 16 * 4 + (3 + 3 + 6 + 3 + 3 + 4 + 3 + 2 + 2 + 2 + 5) * 8
    assert C.sizeof(A.Plate) == 8 + 4 + 4 + 8 * 3 + 8 * 3 + 8 * 2
    assert C.sizeof(A.Object) == 4 + 4 + 8 + 8 + 8 * 15
    assert C.sizeof(A.StepInfo) == 8 + 8 + 8 + 4 * 4 + 4 * 2
