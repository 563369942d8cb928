"""The C ABI library: loads without a GPU, exports every symbol declared in
include/*.h, struct layouts match the header, host-side helpers and error
paths behave (CPU only; no compute call needs a device)."""
import ctypes as C
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]


def declared_functions():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names |= set(re.findall(r"\b(fluid_\w+)\s*\(", text))
    return sorted(names)


def test_header_declares_the_surface():
    names = declared_functions()
    for must in ("fluid_create", "fluid_step", "fluid_destroy", "fluid_update_sph_params",
                 "fluid_spawn_particles", "fluid_read_particles", "fluid_read_cell_starts"):
        assert must in names


def test_library_exports_every_declared_symbol(pkg):
    lib = pkg.abi.load_library()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", str(pkg.abi.LIB_PATH)],
                         capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\bT (fluid_\w+)", out))
    assert set(declared_functions()) <= exported
    # the ctypes prototype table covers the whole header
    assert set(declared_functions()) == set(pkg.abi.PROTOTYPES)


def c_layout(expr_list):
    """sizeof/offsetof values computed by the C compiler from the header."""
    src = '#include <stdio.h>\n#include <stddef.h>\n#include "fluid_abi.h"\nint main(void){\n'
    for e in expr_list:
        src += f'  printf("%zu\\n", (size_t)({e}));\n'
    src += "  return 0;\n}\n"
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        cfile = Path(d) / "l.c"
        cfile.write_text(src)
        exe = Path(d) / "l"
        subprocess.run(["gcc", "-I", str(ROOT / "include"), str(cfile), "-o", str(exe)],
                       check=True)
        out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    return [int(x) for x in out.split()]


@pytest.mark.parametrize("name,size", [
    ("SphParticle3D", 64), ("GpuSphParams3D", 64), ("GpuGridParams", 48),
    ("GpuContainerGeometry", 128), ("GpuGravity", 16), ("GpuMouseForce", 48),
    ("GpuRigidBody", 96), ("GpuRigidBodyAccum", 32), ("PredictedState", 32),
    ("GpuSprayParams", 80), ("GpuSprayParticle", 48), ("GpuMcGridParams", 48),
    ("GpuAnisoParams", 32), ("GpuParticleAniso", 48)])
def test_struct_sizes_match_reference_layouts(pkg, name, size):
    cls = getattr(pkg.abi, name)
    assert C.sizeof(cls) == size
    assert c_layout([f"sizeof({name})"]) == [size]
    fields = [f[0] for f in cls._fields_]
    # a Python keyword field carries a trailing underscore (lambda_)
    offs = c_layout([f"offsetof({name}, {f.rstrip('_') or f})" for f in fields])
    assert offs == [getattr(cls, f).offset for f in fields]


def test_create_info_layout(pkg):
    fields = [f[0] for f in pkg.abi.FluidCreateInfo._fields_]
    offs = c_layout([f"offsetof(fluid_create_info, {f})" for f in fields] +
                    ["sizeof(fluid_create_info)"])
    assert offs[:-1] == [getattr(pkg.abi.FluidCreateInfo, f).offset for f in fields]
    assert offs[-1] == C.sizeof(pkg.abi.FluidCreateInfo)


def test_spray_state_layout(pkg):
    fields = [f[0] for f in pkg.abi.FluidSprayState._fields_]
    offs = c_layout([f"offsetof(fluid_spray_state, {f})" for f in fields] +
                    ["sizeof(fluid_spray_state)"])
    assert offs[:-1] == [getattr(pkg.abi.FluidSprayState, f).offset for f in fields]
    assert offs[-1] == C.sizeof(pkg.abi.FluidSprayState)


def test_spray_error_paths_without_device(pkg):
    lib = pkg.abi.load_library()
    h = C.c_void_p()
    prm = pkg.abi.default_spray_params(dt=1 / 60)
    assert lib.fluid_spray_create(None, 16, C.byref(prm), C.byref(h)) == \
        pkg.abi.FLUID_ERR_INVALID_ARG
    assert lib.fluid_spray_step(None, 0) == pkg.abi.FLUID_ERR_INVALID_ARG
    assert lib.fluid_spray_read_stats(None, None) == pkg.abi.FLUID_ERR_INVALID_ARG
    assert lib.fluid_spray_destroy(None) == pkg.abi.FLUID_OK


def test_error_paths_without_device(pkg):
    lib = pkg.abi.load_library()
    assert lib.fluid_abi_version() == 1
    h = C.c_void_p()
    assert lib.fluid_create(None, C.byref(h)) == pkg.abi.FLUID_ERR_INVALID_ARG
    assert b"null" in lib.fluid_last_error()
    info = pkg.abi.FluidCreateInfo()
    info.sph_params.kernel_radius = 0.0
    assert lib.fluid_create(C.byref(info), C.byref(h)) == pkg.abi.FLUID_ERR_INVALID_ARG
    assert lib.fluid_step(None) == pkg.abi.FLUID_ERR_INVALID_ARG
    assert lib.fluid_destroy(None) == pkg.abi.FLUID_OK
    assert lib.fluid_num_particles(None) == 0
    out = pkg.abi.GpuSphParams3D()
    assert lib.fluid_derive_sph_params(-1.0, 0, 0, 0, 1, 0, 0, 0.008, 0, 0.0,
                                       C.byref(out)) == pkg.abi.FLUID_ERR_INVALID_ARG


def test_derived_params_fields(pkg):
    p = pkg.derive_sph_params(0.08, 35.0, 0.45, 15.0, 1.0, 6.5, 0.305, 0.008, 1234,
                              rest_density=8000.0)
    f = np.float32
    assert p.kernel_radius == f(0.08)
    assert p.kernel_radius_sq == f(0.08) * f(0.08)
    assert p.rest_density == 8000.0 and p.num_particles == 1234
    assert p.stiffness == 35.0 and p.near_stiffness == f(0.45)
    assert p.xsph_epsilon == f(0.305) and p.surface_tension == 6.5
    assert abs(p.pcisph_delta - 10.2266) < 1e-3


def test_scenes_are_deterministic_and_sized(pkg):
    a = pkg.scenes.dam_break_8k().particles
    b = pkg.scenes.dam_break_8k().particles
    np.testing.assert_array_equal(a, b)
    assert a.shape == (8192, 16)
    j = a[:, 0:3] - (pkg.scenes.lattice(16, 32, 16, 0.025) + [-0.8, 0.0, -0.4])
    assert np.abs(j).max() <= 0.002 + 1e-6
    assert pkg.scenes.dam_break_50k().n == 50_000
    s = pkg.scenes.block_1m()
    assert s.n == 1 << 20 and s.bounds_extent == 2.0
    assert abs(s.n * 0.025 ** 3 / (3.2 * 4.0 * 3.2) - 0.4) < 1e-9
    u = pkg.scenes.uniform01(9, 100000)
    assert 0.0 <= u.min() and u.max() < 1.0 and abs(u.mean() - 0.5) < 0.01
