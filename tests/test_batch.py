"""Ensembles (ne_batch_*): a batch step must give, member by member, exactly what stepping each
member alone with ne_model_time_step gives (the kernels share the per-tile and per-column code and
every member keeps its own parameters), with or without the HIP graph, before and after the
autotuner, and across atmosphere snapshots (device-side clock)."""
import numpy as np
import pytest
import torch

from numericalearth_hip import _ctypes_abi as abi
from numericalearth_hip._lib import NEError
from numericalearth_hip.ensemble import (Ensemble, MemberSpec, build_members, draw_ensemble, member_case,
                                         member_params)
from numericalearth_hip.model import OceanOnlyModel

pytestmark = pytest.mark.gpu

DT = 1200.0


def _specs(n=8, seed=5):
    # a small ensemble with the generator's spread of shapes (single columns to regional grids)
    specs = draw_ensemble(n, seed, largest=(150, 60, 20))
    print("[batch] shapes", [(s.Nx, s.Ny, s.Nz) for s in specs], flush=True)
    return specs


def _build(specs, precision, general=(), n_times=1, two_color=(), large_yeager=(), skin=(), time_indexing=None):
    """Members; those whose index is in `general` get a non-default Edson constant (general flux path),
    those in `two_color` the TwoColorRadiation forcing, those in `large_yeager` the Large-Yeager
    coefficient-based fluxes, those in `skin` the skin temperature."""
    from numericalearth_hip import _ctypes_abi as abi_
    out = []
    for k, s in enumerate(specs):
        p = member_params(s)
        if k in general:
            p.edson_momentum[0] = 45.0
        if k in two_color:
            p.radiative_forcing = abi_.NE_RADIATIVE_FORCING_TWO_COLOR
        if k in large_yeager:
            p.turbulent_fluxes = abi_.NE_FLUXES_COEFFICIENT_BASED
            p.transfer_coefficients = abi_.NE_TRANSFER_LARGE_YEAGER
            p.maxiter = 20
        if k in skin:
            p.interface_temperature = abi_.NE_TEMPERATURE_SKIN
            p.skin_diffusivity, p.skin_depth = 1e-4, 1.0
        case = member_case(s) if n_times == 1 else _case_times(s, n_times)
        out.append(OceanOnlyModel(case, params=p, precision=precision, time_indexing=time_indexing))
    return out


def _case_times(s, n_times):
    from numericalearth_hip.generators import NE_BOUNDED, make_case
    return make_case(s.Nx, s.Ny, s.Nz, seed=s.seed, longitude=s.longitude, latitude=s.latitude,
                     topology_x=NE_BOUNDED, n_times=n_times, dt_times=3600.0)


def _state(m):
    d = {"T": m.T, "S": m.S, "J0": m.shortwave_surface_flux}
    d.update({f"flux.{k}": v for k, v in m.fluxes.items()})
    d.update({f"net.{k}": v for k, v in m.net.items()})
    d["frazil"] = m.frazil_heat
    return {k: v.detach().cpu().numpy() for k, v in d.items()}


def _assert_same(a_models, b_models, what):
    torch.cuda.synchronize()
    for n, (a, b) in enumerate(zip(a_models, b_models)):
        sa, sb = _state(a), _state(b)
        for k in sa:
            if not np.array_equal(sa[k], sb[k], equal_nan=True):
                d = np.nanmax(np.abs(sa[k] - sb[k]))
                raise AssertionError(f"{what}: member {n} {a.case.shape} field {k} differs (max |Δ| {d:.3e})")
        assert a.clock == b.clock, (what, n, a.clock, b.clock)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_equals_members_stepped_alone(require_gpu, precision):
    specs = _specs()
    mix = dict(general=(2,), two_color=(4, 5), large_yeager=(1, 6), skin=(6, 7))
    alone = _build(specs, precision, **mix)
    batched = _build(specs, precision, **mix)
    ens = Ensemble(batched)
    print("[batch]", ens.info(), ens.tuning, flush=True)
    for _ in range(3):
        ens.time_step(DT)
        for m in alone:
            m.time_step(DT)
    _assert_same(alone, batched, f"{precision} eager")
    ens.close()
    # released members step on their own again
    for m in batched:
        m.time_step(DT)
    for m in alone:
        m.time_step(DT)
    _assert_same(alone, batched, f"{precision} after release")


def _edge_specs():
    """Members of edge shapes: a single column, one level, one row, two levels, ragged sizes."""
    shapes = [(1, 1, 12), (7, 5, 1), (70, 1, 6), (9, 6, 2), (33, 17, 3), (64, 4, 5)]
    out = []
    for k, (nx, ny, nz) in enumerate(shapes):
        lon0, lat0 = 20.0 + 40.0 * k, -30.0 + 5.0 * k
        out.append(MemberSpec(nx, ny, nz, 100 + k, (lon0, lon0 + 0.25 * nx), (lat0, lat0 + 0.25 * ny),
                              0.011 + 0.001 * k, 1.0))
    print("[batch] edge shapes", shapes, flush=True)
    return out


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_with_edge_members_equals_members_alone(require_gpu, precision):
    specs = _edge_specs()
    alone = _build(specs, precision, two_color=(2,), skin=(4,))
    batched = _build(specs, precision, two_color=(2,), skin=(4,))
    ens = Ensemble(batched)
    for _ in range(3):
        ens.time_step(DT)
        for m in alone:
            m.time_step(DT)
    _assert_same(alone, batched, f"{precision} edge members")
    ens.close()


def test_graph_replays_equal_eager_steps(require_gpu):
    specs = _specs(6, seed=8)
    eager = _build(specs, "fp64")
    graphed = _build(specs, "fp64")
    e1, e2 = Ensemble(eager), Ensemble(graphed)
    e2.use_graph(True)
    for n in range(4):
        e1.time_step(DT)
        e2.time_step(DT)
        # the program's own eager work and a synchronization between replays
        x = torch.randn(1 << 16, device="cuda", dtype=torch.float64)
        float((x * x).sum())
        torch.cuda.synchronize()
        _assert_same(eager, graphed, f"graph replay {n}")


def test_autotune_keeps_state_and_results(require_gpu, tmp_path):
    specs = _specs(6, seed=9)
    alone = _build(specs, "fp64")
    batched = _build(specs, "fp64")
    ens = Ensemble(batched)
    torch.cuda.synchronize()
    before = [_state(m) for m in batched]
    chosen = ens.autotune(cache=tmp_path / "tune.json")
    print("[batch] autotune", chosen, flush=True)
    assert chosen["tile_rows"] in (4, 8, 16) and chosen["surface_block"] in (64, 128, 256)
    assert chosen["stage_ms"] > 0 and chosen["surface_ms"] > 0
    torch.cuda.synchronize()
    after = [_state(m) for m in batched]
    for b, a in zip(before, after):
        for k in b:
            assert np.array_equal(b[k], a[k], equal_nan=True), k
    again = ens.autotune(cache=tmp_path / "tune.json")
    assert again["cached"] and again["tile_rows"] == chosen["tile_rows"]
    for _ in range(2):
        ens.time_step(DT)
        for m in alone:
            m.time_step(DT)
    _assert_same(alone, batched, "autotuned")


@pytest.mark.parametrize("tile_rows", [4, 16])
def test_every_tile_shape_gives_the_same_step(require_gpu, tile_rows):
    specs = _specs(5, seed=10)
    alone = _build(specs, "fp64")
    batched = _build(specs, "fp64")
    ens = Ensemble(batched)
    t = ens.tuning
    t.update(tile_rows=tile_rows, stage_blocks=16, surface_block=64, surface_blocks=3)
    ens.set_tuning(t)
    for _ in range(2):
        ens.time_step(DT)
        for m in alone:
            m.time_step(DT)
    _assert_same(alone, batched, f"tile_rows={tile_rows}")


def test_device_clock_crosses_snapshots(require_gpu):
    specs = _specs(4, seed=11)
    alone = _build(specs, "fp64", n_times=3)
    batched = _build(specs, "fp64", n_times=3)
    ens = Ensemble(batched)
    ens.use_graph(True)
    for _ in range(8):   # 8 × 20 min crosses the 1 h and 2 h snapshots and clamps after the last
        ens.time_step(DT)
        for m in alone:
            m.time_step(DT)
    _assert_same(alone, batched, "time interpolation")


def test_device_clock_cycles_through_the_wrap(require_gpu):
    from numericalearth_hip.time_series import Cyclical
    specs = _specs(4, seed=13)
    # members with different periods: 3 h inferred, and 4.5 h (a 1.5 h wrap interval)
    alone = _build(specs[:2], "fp64", n_times=3, time_indexing=Cyclical())
    alone += _build(specs[2:], "fp64", n_times=3, time_indexing=Cyclical(4.5 * 3600.0))
    batched = _build(specs[:2], "fp64", n_times=3, time_indexing=Cyclical())
    batched += _build(specs[2:], "fp64", n_times=3, time_indexing=Cyclical(4.5 * 3600.0))
    ens = Ensemble(batched)
    ens.use_graph(True)
    for _ in range(24):  # 8 h: through the wrap of both periods, twice for the 3 h one
        ens.time_step(DT)
        for m in alone:
            m.time_step(DT)
    _assert_same(alone, batched, "cyclical time interpolation")


def test_members_are_locked_while_batched(require_gpu):
    specs = _specs(2, seed=12)
    models = _build(specs, "fp64")
    ens = Ensemble(models)
    with pytest.raises(NEError) as e:
        models[0].time_step(DT)
    assert e.value.status == abi.NE_ERR_STATE
    with pytest.raises(NEError):
        Ensemble(models)                      # a model belongs to one batch at a time
    ens.close()
    models[0].time_step(DT)


def test_quasi_adams_bashforth_members_are_refused(require_gpu):
    """The batched step is the split RK3 one: a model with the Quasi-Adams-Bashforth 2 timestepper
    is refused (NE_ERR_UNSUPPORTED) and still steps on its own."""
    from numericalearth_hip.generators import make_case
    case = make_case(24, 12, 4, seed=3)
    m = OceanOnlyModel(case, timestepper="QuasiAdamsBashforth2")
    ens = Ensemble([], precision="fp64")
    with pytest.raises(NEError) as e:
        ens.add(m)
    assert e.value.status == abi.NE_ERR_UNSUPPORTED
    ens.close()
    m.time_step(1200.0)
    torch.cuda.synchronize()
    assert m.clock == (1200.0, 1)
    m.close()


def test_mixed_precisions_are_refused(require_gpu):
    specs = _specs(2, seed=13)
    a = _build(specs[:1], "fp64")[0]
    b = _build(specs[1:], "mixed")[0]
    with pytest.raises(ValueError):
        Ensemble([a, b])
    import ctypes as C
    arr = (C.c_void_p * 2)(a.handle.value, b.handle.value)
    h = C.c_void_p()
    assert a.lib.ne_batch_create(a.ctx.handle, abi.NE_FP64, 2, arr, C.byref(h), None) == abi.NE_ERR_UNSUPPORTED
    ens = Ensemble([], precision="fp64")
    with pytest.raises(NEError) as e:
        ens.add(b)
    assert e.value.status == abi.NE_ERR_UNSUPPORTED
    ens.close()


def test_mixed_deferred_columns_equal_the_inline_finish(require_gpu):
    """The mixed form's single model runs the FP32 flux iteration in a kernel of its own and finishes
    the columns it leaves (drift still above the FP32 threshold) in a second kernel; the batch
    finishes them inline in one kernel. With maxiter = 6 the FP32 iteration stops after 5 and most
    columns are left to the FP64 iteration: both paths must give the same bits."""
    specs = _specs(n=4, seed=11)
    alone, batched = [], []
    for s in specs:
        p = member_params(s)
        p.maxiter = 6
        alone.append(OceanOnlyModel(member_case(s), params=p, precision="mixed"))
        batched.append(OceanOnlyModel(member_case(s), params=p, precision="mixed"))
    ens = Ensemble(batched)
    for _ in range(2):
        ens.time_step(DT)
        for m in alone:
            m.time_step(DT)
    _assert_same(alone, batched, "mixed, maxiter 6")
    its = np.concatenate([m.iterations.cpu().numpy().ravel() for m in alone])
    deferred = int(np.sum(its == 6))
    print(f"[batch] columns finished by the FP64 iteration: {deferred} of {its.size}", flush=True)
    assert deferred > 0
    ens.close()
