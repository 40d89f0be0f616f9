"""The Julia extension (numericalearth.jl_amd/julia/NumericalEarthHIPExt.jl) cannot run here (no
Julia), so its binding is checked as text against the C ABI: every ABI struct it mirrors has the
header's fields in the header's order with the matching types (through the ctypes mirror, whose
layout test_abi compiles against the header), every entry point it calls is declared, each
ccall passes as many arguments as the header declares parameters, and its enum constants have the
header's values."""
import ctypes as C
import re

import pytest

from conftest import REPO
from numericalearth_hip import _ctypes_abi as abi

JL = REPO / "numericalearth.jl_amd" / "julia" / "NumericalEarthHIPExt.jl"
HEADERS = sorted((REPO / "include").glob("*.h"))

MIRRORS = {
    "NEView": abi.View, "NEGrid": abi.Grid, "NESourceGrid": abi.SourceGrid, "NETimeInterp": abi.TimeInterp,
    "NEPrescribedAtmosphere": abi.PrescribedAtmosphereDesc, "NEPrescribedRadiation": abi.PrescribedRadiationDesc,
    "NEExchangeState": abi.ExchangeState, "NEOceanState": abi.OceanState, "NEInterfaceFluxes": abi.InterfaceFluxes,
    "NENetOceanFluxes": abi.NetOceanFluxes, "NERadiativeFluxes": abi.RadiativeFluxes, "NEParams": abi.Params,
}
SCALARS = {"Float64": C.c_double, "Int32": C.c_int32, "Int64": C.c_int64, "UInt32": C.c_uint32,
           "UInt64": C.c_uint64}


def _text():
    return JL.read_text()


def _structs(text):
    out = {}
    for m in re.finditer(r"^struct (\w+)\n(.*?)^end", text, flags=re.M | re.S):
        fields = []
        for line in m.group(2).splitlines():
            f = re.match(r"\s+(\w+) :: (.+?)\s*(#.*)?$", line)
            if f:
                fields.append((f.group(1), f.group(2)))
        out[m.group(1)] = fields
    return out


def _same_type(jl: str, ct) -> bool:
    if jl.startswith("Ptr{"):
        return ct is C.c_void_p
    m = re.match(r"NTuple\{(\d+), (\w+)\}", jl)
    if m:
        return hasattr(ct, "_length_") and ct._length_ == int(m.group(1)) and \
            C.sizeof(ct._type_) == C.sizeof(SCALARS[m.group(2)])
    if jl in SCALARS:
        return C.sizeof(ct) == C.sizeof(SCALARS[jl]) and not hasattr(ct, "_fields_")
    return jl in MIRRORS and ct is MIRRORS[jl]


@pytest.mark.parametrize("name", sorted(MIRRORS))
def test_struct_mirrors_the_header(name):
    fields = _structs(_text())[name]
    want = MIRRORS[name]._fields_
    assert [f for f, _ in fields] == [f for f, _ in want], name
    for (f, jt), (_, ct) in zip(fields, want):
        assert _same_type(jt, ct), (name, f, jt, ct)


def _header_params():
    params = {}
    for h in HEADERS:
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        for m in re.finditer(r"^\s*int\s+(ne_\w+)\s*\(([^;]*?)\)\s*;", text, flags=re.M | re.S):
            args = [a for a in m.group(2).split(",") if a.strip() and a.strip() != "void"]
            params[m.group(1)] = len(args)
    return params


def _split_top(s):
    parts, depth, cur = [], 0, ""
    for ch in s:
        if ch in "([{":
            depth += 1
        elif ch in ")]}":
            depth -= 1
        if ch == "," and depth == 0:
            parts.append(cur)
            cur = ""
        else:
            cur += ch
    if cur.strip():
        parts.append(cur)
    return parts


def _ccalls(text):
    """(function names, number of argument types) of every ccall, the names of a call through a
    symbol built as Symbol(:ne_x_, suffix(…)) expanded to its f64 and f32 forms."""
    calls = []
    for m in re.finditer(r"ccall\(", text):
        depth, i = 1, m.end()
        while depth:
            depth += {"(": 1, ")": -1}.get(text[i], 0)
            i += 1
        parts = _split_top(text[m.end():i - 1])
        target, types = parts[0].strip(), parts[2].strip()
        ntypes = len(_split_top(types[1:-1]))
        lit = re.match(r"\(:(ne_\w+), libne\)", target)
        if lit:
            names = [lit.group(1)]
        else:   # the nearest preceding Symbol(:ne_…_, suffix(…)) names the call
            pre = re.findall(r"Symbol\(:(ne_\w+_), suffix", text[:m.start()])
            names = [pre[-1] + "f64", pre[-1] + "f32"]
        calls.append((names, ntypes))
    return calls


def test_every_call_is_declared_with_its_arity():
    declared = _header_params()
    calls = _ccalls(_text())
    assert len(calls) >= 12
    called = set()
    for names, ntypes in calls:
        for n in names:
            assert n in declared, n
            assert declared[n] == ntypes, (n, declared[n], ntypes)
            called.add(n)
    # the phases of the coupled step and the ocean step are all bound
    for n in ("ne_interp_atmos_f64", "ne_interp_radiation_f64", "ne_ao_flux_solve_f64", "ne_freeze_clamp_f64",
              "ne_assemble_net_ocean_fluxes_f64", "ne_apply_air_sea_radiation_f64", "ne_tracer_rk_stage_f64",
              "ne_tracer_rk_stage_f32", "ne_interp_atmos_f32", "ne_grid_immersed_levels",
              "ne_fractional_indices_nodes_f64", "ne_snapshot_time_indices", "ne_params_default"):
        assert n in called, n


def test_enum_constants_have_the_header_values():
    text = _text()
    checked = 0
    for m in re.finditer(r"^const ((?:NE_\w+, )*NE_\w+) = ((?:Int32\(\d+\), )*Int32\(\d+\))$", text, flags=re.M):
        names = m.group(1).split(", ")
        values = [int(v) for v in re.findall(r"Int32\((\d+)\)", m.group(2))]
        for n, v in zip(names, values):
            assert getattr(abi, n) == v, n
            checked += 1
    assert checked >= 15


def test_methods_dispatch_only_on_handed_over_models():
    """No method of the extension applies to a model whose ocean is not a HIPOcean."""
    text = _text()
    sigs = re.findall(r"^function (?:EarthSystemModels|InterfaceComputations|Oceananigans\.TimeSteppers)\.\w+!?\((.*?)\)$",
                      text, flags=re.M)
    assert len(sigs) >= 7
    for s in sigs:
        assert re.search(r"::HIP(Model|Ocean|FreezingLimitedModel|RadiatedModel)\b", s), s
    assert "const HIPModel = EarthSystemModel{<:Any, <:Any, <:Any, <:Any, <:HIPOcean}" in text
