"""Vertical velocity from continuity (compute_w_from_continuity!, Oceananigans, called by the
HydrostaticFreeSurfaceModel that ocean_simulation.jl:487-499 builds): oracle pins on the CPU and
the HIP kernel against the oracle on the GPU."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from numericalearth_hip import _ctypes_abi as abi
from numericalearth_hip.generators import NE_BOUNDED, NE_PERIODIC, continuity_w, make_case


def _oracle_w(case, u=None, v=None):
    om = O.OracleModel(case)
    u = np.ascontiguousarray(case.u if u is None else u)
    v = np.ascontiguousarray(case.v if v is None else v)
    w = np.full_like(case.w, np.nan)
    O.lib().neo_compute_w_from_continuity(C.byref(om.grid), O.np_view(u), O.np_view(v), O.np_view(w))
    return w


def _divergence(case, u, v, w):
    """Per cell: the net outflow δx(Ax u) + δy(Ay v) + Az δz w, and the magnitude of its terms."""
    g = case.grid
    Nx, Ny, Nz = case.shape
    per_x = g.spec.topology_x == NE_PERIODIC
    up = np.concatenate([u[:, :, 1:Nx], u[:, :, :1]], axis=2) if per_x else u[:, :, 1:]
    out = np.zeros((Nz, Ny, Nx))
    mag = np.zeros((Nz, Ny, Nx))
    for k in range(Nz):
        Ax = g.dy_xface[:, None] * g.dz[k]
        Ays = g.dx_yface[:Ny, None] * g.dz[k]
        Ayn = g.dx_yface[1:, None] * g.dz[k]
        terms = (Ax * up[k], -Ax * u[k, :, :Nx], Ayn * v[k, 1:], -Ays * v[k, :Ny],
                 g.Az[:, None] * w[k + 1], -g.Az[:, None] * w[k])
        out[k] = sum(terms)
        mag[k] = sum(np.abs(t) for t in terms)
    return out, mag


@pytest.mark.parametrize("topology_x", [NE_PERIODIC, NE_BOUNDED])
def test_continuity_is_divergence_free_and_zero_below(topology_x):
    case = make_case(48, 30, 12, seed=31, topology_x=topology_x)
    rng = np.random.default_rng(5)
    g = case.grid
    active = np.arange(g.spec.Nz)[:, None, None] >= g.kmin[None]
    # new horizontal velocities, zero on faces touching an inactive cell
    u = rng.normal(size=case.u.shape) * 0.2
    v = rng.normal(size=case.v.shape) * 0.2
    Nx = case.shape[0]
    act_u = np.zeros(case.u.shape, bool)
    left = np.concatenate([active[:, :, -1:], active[:, :, :]], axis=2) if topology_x == NE_PERIODIC else \
        np.concatenate([np.zeros_like(active[:, :, :1]), active], axis=2)
    right = np.concatenate([active, active[:, :, :1]], axis=2) if topology_x == NE_PERIODIC else \
        np.concatenate([active, np.zeros_like(active[:, :, :1])], axis=2)
    act_u = left & right
    u = np.where(act_u, u, 0.0)
    if topology_x == NE_PERIODIC:
        u[:, :, Nx] = u[:, :, 0]
    south = np.concatenate([np.zeros_like(active[:, :1]), active], axis=1)
    north = np.concatenate([active, np.zeros_like(active[:, :1])], axis=1)
    v = np.where(south & north, v, 0.0)
    w = _oracle_w(case, u, v)
    assert np.all(w[0] == 0.0)
    # faces above inactive cells carry nothing
    assert np.all(w[1:][~active] == 0.0)
    div, mag = _divergence(case, u, v, w)
    assert np.all(np.abs(div[active]) <= 1e-13 * mag[active])
    # the surface value is the column's net horizontal inflow over Az
    hor = div - g.Az[None, :, None] * (w[1:] - w[:-1])
    assert np.allclose(w[-1] * g.Az[:, None], -hor.sum(axis=0), rtol=1e-12, atol=1e-12 * np.abs(hor).sum(axis=0).max())
    # zero velocities give zero w
    assert np.all(_oracle_w(case, 0 * u, 0 * v) == 0.0)


def test_continuity_reproduces_the_generated_w():
    """The generator builds w from continuity with the same face arithmetic: the oracle gives it back
    bit for bit (and, its velocities having no barotropic part, w vanishes at the top)."""
    case = make_case(40, 22, 9, seed=32)
    w = _oracle_w(case)
    assert np.array_equal(w, continuity_w(case.grid, case.u, case.v,
                                          np.arange(9)[:, None, None] >= case.grid.kmin[None]))
    assert np.array_equal(w, case.w)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_continuity_kernel_matches_oracle(require_gpu, precision):
    import torch
    from numericalearth_hip.model import OceanOnlyModel, Phases
    case = make_case(90, 40, 10, seed=33)
    gm = OceanOnlyModel(case, precision=precision)
    # new horizontal velocities (zero where the case's are), as a caller evolving them would set
    gm.u.copy_(torch.as_tensor(1.7 * case.u, dtype=gm.u.dtype, device=gm.u.device))
    gm.v.copy_(torch.as_tensor(-0.6 * case.v, dtype=gm.v.dtype, device=gm.v.device))
    gm.update_vertical_velocity()
    torch.cuda.synchronize()
    u_used = gm.u.detach().cpu().numpy().astype(np.float64)
    v_used = gm.v.detach().cpu().numpy().astype(np.float64)
    ref = _oracle_w(case, u_used, v_used)
    got = gm.w.detach().cpu().numpy().astype(np.float64)
    if precision == "fp64":
        assert np.array_equal(got, ref)
        ph = Phases(gm)
        w2 = torch.full_like(gm.w, float("nan"))
        ph.compute_w(gm.u, gm.v, w2)
        torch.cuda.synchronize()
        assert np.array_equal(w2.cpu().numpy(), ref)
    else:   # accumulated in FP64, each face rounded once to FP32
        assert np.array_equal(got, ref.astype(np.float32).astype(np.float64))


@pytest.mark.gpu
def test_model_steps_with_evolving_velocities_match_oracle(require_gpu):
    """A caller that evolves u and v sets them before each step and recomputes w from continuity
    (the model's update_vertical_velocity); four such steps equal the oracle doing the same, to
    1e-12 in T and S, and w is the oracle's bit for bit after every update."""
    import torch
    from helpers import assert_rel
    from numericalearth_hip.model import OceanOnlyModel
    case = make_case(90, 40, 10, seed=34)
    print("[case] 90x40x10 seed=34 evolving velocities", flush=True)
    gm = OceanOnlyModel(case)
    om = O.OracleModel(case)
    om.reconcile()
    for n in range(4):
        f = 1.0 + 0.5 * np.sin(0.9 * n + 0.3)      # same factor for u and v: no barotropic transport
        u, v = f * case.u, f * case.v
        gm.u.copy_(torch.as_tensor(u, dtype=gm.u.dtype, device=gm.u.device))
        gm.v.copy_(torch.as_tensor(v, dtype=gm.v.dtype, device=gm.v.device))
        gm.update_vertical_velocity()
        om.u[...] = u
        om.v[...] = v
        om.w[...] = _oracle_w(case, u, v)
        torch.cuda.synchronize()
        assert np.array_equal(gm.w.cpu().numpy(), om.w), f"w after update {n + 1}"
        gm.time_step(1200.0)
        om.time_step(1200.0)
        torch.cuda.synchronize()
        assert_rel(gm.T.cpu().numpy(), om.T, 1e-12, f"T after step {n + 1}")
        assert_rel(gm.S.cpu().numpy(), om.S, 1e-12, f"S after step {n + 1}")
    gm.close()
