"""Sensitivity of the oracle's closed-form pins (CPU only).

Each case plants one plausible mistake in a copy of oracle/fluid_oracle.cpp
(a flipped sign, a swapped index or operand, a wrong constant, a transposed
rotation, a wrong texture layout), builds that copy, and requires the named
pins of tests/test_oracle_terms.py to fail against it.  A pin that still
passes with the mistake in place would not catch it.  The unmodified oracle
passes every pin (test_oracle_terms.py itself).
"""
import shutil
import subprocess
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

import pytest

import test_oracle_terms as terms

ROOT = Path(__file__).resolve().parents[1]
SRC = ROOT / "oracle" / "fluid_oracle.cpp"

# (name, text in fluid_oracle.cpp, replacement, pins that must fail)
MUTANTS = [
    ("pressure_sign", "T sc = -mass * (old_p / (sd * sd) + np / (nd * nd)) * spiky_grad(r);",
     "T sc = mass * (old_p / (sd * sd) + np / (nd * nd)) * spiky_grad(r);",
     ["pcisph_pair_repels", "pcisph_triple_asymmetric"]),
    ("pressure_neighbour_index",
     "T sc = -mass * (old_p / (sd * sd) + np / (nd * nd)) * spiky_grad(r);",
     "T sc = -mass * (old_p / (sd * sd) + old_p / (nd * nd)) * spiky_grad(r);",
     ["pcisph_triple_asymmetric", "pcisph_cluster_momentum"]),
    ("pressure_no_clamp", "if (amag > max_a2) ap = muls(ap, max_a / std::sqrt(amag));",
     "if (amag > max_a2 * T(4)) ap = muls(ap, max_a / std::sqrt(amag));",
     ["pcisph_cfl_clamp"]),
    ("pressure_error_sign", "T err = std::max(T(0), pd - rho0);",
     "T err = std::max(T(0), rho0 - pd);", ["pcisph_pair_repels"]),
    ("separation_scale", "T sep = T(200.0) * (T(1.0) - r / min_dist);",
     "T sep = T(200.0) * (T(1.0) - r / h);", ["separation_adjacent_slots"]),
    ("separation_sign", "ap = add(ap, smul(-sep, dir));", "ap = add(ap, smul(sep, dir));",
     ["separation_adjacent_slots", "separation_far_slots"]),
    ("hash_dir_component", "std::cos(idx * T(2.0))", "std::sin(idx * T(2.0))",
     ["separation_adjacent_slots", "separation_far_slots"]),
    ("hash_dir_order", "T idx = T((float)s) - T((float)j);",
     "T idx = T((float)j) - T((float)s);", ["separation_adjacent_slots"]),
    ("wall_damping_sign", "if (lv.x < T(0)) wa.x -= lv.x * wd * t;",
     "if (lv.x < T(0)) wa.x += lv.x * wd * t;", ["walls_upright", "wall_corner"]),
    ("wall_layer_width", "T bl = h * T(0.4);", "T bl = h * T(0.5);",
     ["walls_upright", "walls_tilted"]),
    ("wall_frame", "accel = add(accel, local_dir_to_world(wa));",
     "accel = add(accel, world_dir_to_local(wa));", ["walls_tilted"]),
    ("wall_penalty_power", "wa.y += k * t * t;", "wa.y += k * t;",
     ["walls_upright", "wall_corner"]),
    ("backstop_restitution", "lvn.x = std::fabs(lvn.x) * rest;", "lvn.x = std::fabs(lvn.x);",
     ["backstop_upright", "backstop_tilted"]),
    ("backstop_frame", "S3(particles[i].pos, local_to_world(lpn));",
     "S3(particles[i].pos, mk(lpn.x, lpn.y + c.cy, lpn.z));", ["backstop_tilted"]),
    ("vortex_cross_order", "V3<T> tg = cross(fd, axis);", "V3<T> tg = cross(axis, fd);",
     ["mouse_modes"]),
    ("drain_direction", "V3<T> dd = normalize(add(fd, mk(T(0), T(-1.0), T(0))));",
     "V3<T> dd = normalize(add(fd, mk(T(0), T(1.0), T(0))));", ["mouse_modes"]),
    ("explode_factor", "accel = sub(accel, muls(muls(fd, s), T(3.0)));",
     "accel = sub(accel, muls(muls(fd, s), T(2.0)));", ["mouse_modes"]),
    ("mouse_falloff", "T fall = T(1.0) - (dist / rad);", "T fall = T(1.0) - dist;",
     ["mouse_modes"]),
    ("pull_sign", "case 1:\n              accel = add(accel, muls(fd, s));",
     "case 1:\n              accel = sub(accel, muls(fd, s));", ["mouse_modes"]),
    ("body_normal_transpose", "V3<T> nw = mk(r0.x * ln.x + r1.x * ln.y + r2.x * ln.z,",
     "V3<T> nw = mk(r0.x * ln.x + r0.y * ln.y + r0.z * ln.z,", ["rigid_shapes"]),
    ("torus_minor", "T minor = he * T(0.3);", "T minor = he * T(0.25);", ["rigid_shapes"]),
    ("torus_normal", "ln = mk(rdx * rl.x / xz, rdy, rdx * rl.z / xz);",
     "ln = mk(rdx * rl.x / xz, rdy, rdx * rl.x / xz);", ["rigid_shapes"]),
    ("cylinder_cap_abs", "T d_cap = std::fabs(rl.y) - he;", "T d_cap = rl.y - he;",
     ["rigid_shapes"]),
    ("cube_normal_axis", "ln = mk(wsign(rl.x), T(0), T(0));", "ln = mk(T(0), wsign(rl.x), T(0));",
     ["rigid_shapes"]),
    ("sdf_texel_layout", "return sdf[(size_t)x + (size_t)y * r + (size_t)z * r * r];",
     "return sdf[(size_t)z + (size_t)y * r + (size_t)x * r * r];", ["rigid_shapes"]),
    ("body_range", "T ir = h * T(0.7);", "T ir = h * T(0.6);", ["rigid_shapes"]),
    ("body_penalty_power", "V3<T> pen_a = muls(muls(muls(nw, T(rb.stiffness)), t), t);",
     "V3<T> pen_a = muls(muls(nw, T(rb.stiffness)), t);", ["rigid_shapes"]),
    ("body_damping", "damp = muls(muls(mk(-nw.x, -nw.y, -nw.z), vn), T(5.0));",
     "damp = muls(muls(mk(-nw.x, -nw.y, -nw.z), vn), T(4.0));", ["rigid_shapes"]),
    ("accum_scale", "acc_add(0, to_i32(re.x * T(1000.0)));",
     "acc_add(0, to_i32(re.x * T(100.0)));", ["rigid_shapes"]),
    ("accum_torque_order", "V3<T> tq = cross(wr, re);", "V3<T> tq = cross(re, wr);",
     ["rigid_shapes"]),
    ("accum_count", "accum[3] = (int32_t)((uint32_t)accum[3] + 1u);",
     "accum[3] = (int32_t)((uint32_t)accum[3] + 2u);", ["rigid_shapes"]),
]


def _build(tmp: Path, name: str, old: str, new: str) -> Path:
    text = SRC.read_text()
    assert text.count(old) == 1, f"{name}: anchor must occur once"
    text = text.replace(old, new).replace('#include "../include/fluid_abi.h"',
                                          f'#include "{ROOT / "include" / "fluid_abi.h"}"')
    src = tmp / f"{name}.cpp"
    src.write_text(text)
    lib = tmp / f"lib{name}.so"
    subprocess.run(["g++", "-O1", "-std=c++17", "-fPIC", "-shared", "-fopenmp",
                    "-ffp-contract=off", "-fno-fast-math", "-o", str(lib), str(src)],
                   check=True, capture_output=True)
    return lib


@pytest.fixture(scope="module")
def mutant_libs(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    tmp = tmp_path_factory.mktemp("oracle_mutants")
    with ThreadPoolExecutor(max_workers=8) as ex:
        futs = {m[0]: ex.submit(_build, tmp, m[0], m[1], m[2]) for m in MUTANTS}
        return {k: f.result() for k, f in futs.items()}


@pytest.mark.parametrize("name,old,new,pins", MUTANTS, ids=[m[0] for m in MUTANTS])
def test_planted_mistake_fails_its_pins(pkg, orc, mutant_libs, name, old, new, pins):
    with orc.library(mutant_libs[name]):
        for pin in pins:
            with pytest.raises(AssertionError):
                terms.PINS[pin](pkg, orc)

