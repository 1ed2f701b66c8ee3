"""CPU checks of the C-ABI library: it loads, exports every symbol include/sfsim_atmo.h declares,
and its host-only entry points behave (no GPU compute is called here)."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from sfsim_amd import atmosphere as atm

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    text = open(os.path.join(ROOT, "include", "sfsim_atmo.h")).read()
    return sorted(set(re.findall(r"\b(sfsim_atmo_[a-z0-9_]+)\s*\(", text)))


def test_header_and_python_export_list_agree():
    assert sorted(atm.EXPORTS) == _declared_symbols()


def test_library_exports_every_declared_symbol():
    lib = atm.lib()
    for name in _declared_symbols():
        assert hasattr(lib, name), name


def test_null_context_is_an_argument_error():
    lib = atm.lib()
    P = atm.workload("earth100-64")
    out = np.zeros(64 * 64 * 3, np.float32)
    assert lib.sfsim_atmo_transmittance_table(None, ctypes.byref(P), atm._fp(out)) == atm.ERR_ARGUMENT
    assert lib.sfsim_atmo_destroy(None) == atm.ERR_ARGUMENT
    assert lib.sfsim_atmo_last_error(None) == b"context is NULL"
    assert lib.sfsim_atmo_create(0, None) == atm.ERR_ARGUMENT


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="checks the no-device path")
def test_create_without_device_fails_loudly():
    with pytest.raises(atm.AtmoError) as e:
        atm.Atmosphere(0)
    assert e.value.code in (atm.ERR_DEVICE, atm.ERR_ARGUMENT)


def test_scatter_file_round_trip(tmp_path):
    data = np.arange(3 * 7, dtype=np.float32) * 0.25 - 1.0
    path = tmp_path / "t.scatter"
    atm.write_scatter(str(path), data)
    raw = path.read_bytes()
    assert raw == data.astype("<f4").tobytes()  # util.clj:227-240: little-endian float32, no header
    assert np.array_equal(atm.read_scatter(str(path), data.size), data)
    with pytest.raises(atm.AtmoError):
        atm.read_scatter(str(path), data.size + 1)
    with pytest.raises(atm.AtmoError):
        atm.read_scatter(str(path), data.size - 1)


def test_sphere_direction_counts_match_reference_quadrature():
    assert atm.sphere_directions(False, 14) == 66
    assert atm.sphere_directions(False, 15) == 71
    assert atm.sphere_directions(False, 28) == 256
    assert atm.sphere_directions(True, 64) == 660
    assert atm.sphere_directions(True, 100) == 1605
    assert atm.sphere_directions(True, 3) == 0


def test_stage_names():
    lib = atm.lib()
    names = [lib.sfsim_atmo_stage_name(k).decode() for k in range(9)]
    assert names == ["transmittance", "surface_radiance_base", "single_scatter", "direction_geometry", "point_scatter",
                     "surface_radiance", "ray_scatter", "accumulate", "final_resample"]
    assert lib.sfsim_atmo_stage_name(99) == b""


def test_entries_count_of_workloads():
    assert atm.entries(atm.workload("full-5-order")) == 64 * 256 + 16 * 64 + 2 * 32 * 64 * 32 * 8
    assert atm.entries(atm.workload("production")) == 2098096


PARAM_FIELDS = [  # (C member designator, bytes) in declaration order
    ("planet.centre", 24), ("planet.radius", 8), ("planet.height", 8), ("planet.brightness", 24),
    ("mie.scatter_base", 24), ("mie.scatter_scale", 8), ("mie.scatter_quotient", 8), ("mie.scatter_g", 8),
    ("rayleigh.scatter_base", 24), ("rayleigh.scatter_scale", 8), ("rayleigh.scatter_quotient", 8),
    ("rayleigh.scatter_g", 8), ("intensity", 24), ("transmittance_shape", 8), ("surface_radiance_shape", 8),
    ("scatter_shape", 16), ("ray_steps", 4), ("sphere_steps", 4), ("iterations", 4), ("reserved", 4)]


def test_params_struct_has_no_padding(tmp_path):
    """sfsim_atmo_params as the C compiler lays it out equals the packed field list that coffi's
    [::mem/struct ...] and ctypes describe: every offset is the sum of the sizes before it and the size is
    their total (a multiple of 8, so arrays of parameter sets stride alike)."""
    src = tmp_path / "probe.c"
    lines = ['#include <stddef.h>', '#include <stdio.h>', f'#include "{ROOT}/include/sfsim_atmo.h"', "int main(void) {"]
    for name, _ in PARAM_FIELDS:
        lines.append(f'  printf("%zu\\n", offsetof(sfsim_atmo_params, {name}));')
    lines += ['  printf("%zu\\n", sizeof(sfsim_atmo_params));', "  return 0;", "}"]
    src.write_text("\n".join(lines))
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-std=c11", "-Wall", "-o", str(exe), str(src)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
    packed = np.cumsum([0] + [n for _, n in PARAM_FIELDS]).tolist()
    assert got == packed
    assert got[-1] == 232 and got[-1] % 8 == 0
    assert ctypes.sizeof(atm.Params) == 232
    for (name, _), off in zip(PARAM_FIELDS, packed):
        outer, _, inner = name.partition(".")
        base = getattr(atm.Params, outer).offset
        sub = getattr(type(getattr(atm.Params(), outer)), inner).offset if inner else 0
        assert base + sub == off, name
