"""CPU checks of the C-ABI library: it loads, exports every symbol include/sfsim_atmo.h declares,
and its host-only entry points behave (no GPU compute is called here)."""
import ctypes
import os
import re

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
