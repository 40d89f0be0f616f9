"""C ABI checks that need no GPU: the library loads, exports every entry point the
header declares, and the ctypes mirror has the header's exact struct layouts."""
import ctypes as C
import re
import subprocess
from pathlib import Path

import pytest

from conftest import REPO
from numericalearth_hip import _ctypes_abi as abi
from numericalearth_hip import _lib
from numericalearth_hip.build import library_path

HEADER = REPO / "include" / "numericalearth_hip.h"
HEADERS = sorted((REPO / "include").glob("*.h"))   # the product header and the test-only one

STRUCTS = {
    "ne_view": abi.View, "ne_grid": abi.Grid, "ne_source_grid": abi.SourceGrid, "ne_time_interp": abi.TimeInterp,
    "ne_prescribed_atmosphere": abi.PrescribedAtmosphereDesc, "ne_prescribed_radiation": abi.PrescribedRadiationDesc,
    "ne_exchange_state": abi.ExchangeState, "ne_ocean_state": abi.OceanState,
    "ne_interface_fluxes": abi.InterfaceFluxes, "ne_net_ocean_fluxes": abi.NetOceanFluxes,
    "ne_radiative_fluxes": abi.RadiativeFluxes, "ne_params": abi.Params, "ne_counters": abi.Counters,
    "ne_model_desc": abi.ModelDesc, "ne_batch_tuning": abi.BatchTuning,
}


def declared_functions(headers=HEADERS):
    names = set()
    for h in headers:
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names |= set(re.findall(r"^\s*int\s+(ne_\w+)\s*\(", text, flags=re.M))
    return sorted(names)


def test_test_hooks_stay_out_of_the_product_header():
    assert "ne_log_f64" not in declared_functions([HEADER])
    assert "ne_log_f64" in declared_functions()


def test_library_exports_every_declared_entry_point():
    lib = C.CDLL(str(library_path()))
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"
    # and the Python binding declares a signature for each of them
    assert sorted(_lib.SIGNATURES) == names


def test_integration_notes_index_every_declared_entry_point():
    """The maintainer's notes list each exported symbol (a `name_{f64,f32}` entry names both)."""
    text = (REPO / "INTEGRATION.md").read_text()
    documented = set()
    for name in re.findall(r"`(ne_[\w{},]+)`", text):
        m = re.fullmatch(r"(ne_\w*?)\{([\w,]+)\}", name)
        documented |= {m.group(1) + s for s in m.group(2).split(",")} if m else {name}
    undocumented = [n for n in declared_functions() if n not in documented]
    assert not undocumented, f"not in INTEGRATION.md: {undocumented}"


def test_version_and_defaults_match_python_mirror():
    L = _lib.load()
    assert L.ne_version() == 1
    p = abi.Params()
    assert L.ne_params_default(C.byref(p)) == 0
    from numericalearth_hip.params import default_params, params_dict
    assert params_dict(p) == params_dict(default_params())


def test_null_arguments_are_rejected_without_a_device():
    L = _lib.load()
    assert L.ne_params_default(None) == abi.NE_ERR_NULL
    assert L.ne_ctx_destroy(None) == abi.NE_ERR_NULL
    assert L.ne_model_destroy(None) == abi.NE_ERR_NULL
    assert L.ne_model_time_step(None, 1.0, None) == abi.NE_ERR_NULL
    h = C.c_void_p()
    assert L.ne_batch_create(None, abi.NE_FP64, 1, None, C.byref(h), None) == abi.NE_ERR_NULL
    assert L.ne_batch_destroy(None) == abi.NE_ERR_NULL
    assert L.ne_batch_add(None, None) == abi.NE_ERR_NULL
    assert L.ne_batch_remove(None, None) == abi.NE_ERR_NULL
    assert L.ne_model_workspace_bytes(None, None) == abi.NE_ERR_NULL
    assert L.ne_batch_time_step(None, 1.0, None) == abi.NE_ERR_NULL
    assert L.ne_batch_set_graph(None, 1) == abi.NE_ERR_NULL
    assert L.ne_batch_autotune(None, None, None) == abi.NE_ERR_NULL
    assert L.ne_batch_set_tuning(None, None) == abi.NE_ERR_NULL
    assert L.ne_batch_info(None, None, None, None, None) == abi.NE_ERR_NULL
    assert L.ne_model_rounding_residual(None, None, None) == abi.NE_ERR_NULL
    assert L.ne_stream_probe_f64(None, None, 5, None, 2, 10, None) == abi.NE_ERR_NULL
    assert L.ne_fill_halo_f64(None, None, abi.View(), 1, 1, 1, None) == abi.NE_ERR_NULL
    v = abi.View()
    assert L.ne_tracer_ab2_update_f64(None, None, v, v, v, v, v, v, 1.0, abi.NE_AB2_CHI, 0, None) == abi.NE_ERR_NULL


def test_quasi_adams_bashforth_options_are_checked_on_the_host():
    """The timestepper names of ocean_simulation (SplitRungeKutta3, QuasiAdamsBashforth2); the
    Quasi-Adams-Bashforth 2 step is FP64 only, refused for the mixed form before any device work."""
    from numericalearth_hip.model import OceanOnlyModel
    with pytest.raises(ValueError):
        OceanOnlyModel(None, precision="mixed", timestepper="QuasiAdamsBashforth2")
    with pytest.raises(ValueError):
        OceanOnlyModel(None, timestepper="RungeKutta4")
    assert (abi.NE_TIMESTEPPER_SPLIT_RK3, abi.NE_TIMESTEPPER_QUASI_ADAMS_BASHFORTH2) == (0, 1)


def test_julia_mirror_has_the_header_params_size():
    text = (REPO / "numericalearth.jl_amd" / "julia" / "NumericalEarthHIPExt.jl").read_text()
    m = re.search(r"const NE_PARAMS_BYTES = (\d+)", text)
    assert m and int(m.group(1)) == C.sizeof(abi.Params)


def test_struct_layouts_match_header(tmp_path):
    lines = ["#include <stdio.h>", "#include <stddef.h>", f'#include "{HEADER}"', "int main(void) {"]
    for cname, cls in STRUCTS.items():
        lines.append(f'printf("{cname} sizeof %zu\\n", sizeof({cname}));')
        for fname, _ in cls._fields_:
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {}
    for line in out:
        if line:
            s, f, v = line.split()
            got[(s, f)] = int(v)
    for cname, cls in STRUCTS.items():
        assert got[(cname, "sizeof")] == C.sizeof(cls), cname
        for fname, _ in cls._fields_:
            assert got[(cname, fname)] == getattr(cls, fname).offset, (cname, fname)
