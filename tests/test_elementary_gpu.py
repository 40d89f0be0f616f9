"""The flux solve's FP64 logarithm (ne_log_f64) against numpy's, on the GPU: within 2 ulp over
the ranges the solve passes it (arguments between ½ and 2 after reduction, ratios of heights to
roughness lengths, Reynolds numbers) and over the whole normal range, and exact where the value
is: log 1 = 0, and log 2^k = k·ln 2 rounded once. Its arctangent for arguments ≥ 1
(ne_atan_ge1_f64) against numpy's: within 2 ulp on both sides of the switch at 1 + √2 and up to
the largest doubles, π/4 at 1 and π/2 at +∞. Its cube root for arguments ≥ 1 (ne_cbrt_ge1_f64)
against numpy's: within 1 ulp up to 1e38, exact at small cubes and powers of eight."""
from decimal import Decimal, getcontext

import numpy as np
import pytest
import torch

from numericalearth_hip._lib import check
from numericalearth_hip.model import Context, stream_handle

pytestmark = pytest.mark.gpu
getcontext().prec = 50


def _device_log(x):
    ctx = Context.get(0)
    xd = torch.from_numpy(np.ascontiguousarray(x, dtype=np.float64)).cuda()
    yd = torch.empty_like(xd)
    check(ctx.lib.ne_log_f64(ctx.handle, xd.data_ptr(), yd.data_ptr(), xd.numel(), stream_handle()), "ne_log_f64")
    torch.cuda.synchronize()
    return yd.cpu().numpy()


def _device_atan_ge1(x):
    ctx = Context.get(0)
    xd = torch.from_numpy(np.ascontiguousarray(x, dtype=np.float64)).cuda()
    yd = torch.empty_like(xd)
    check(ctx.lib.ne_atan_ge1_f64(ctx.handle, xd.data_ptr(), yd.data_ptr(), xd.numel(), stream_handle()),
          "ne_atan_ge1_f64")
    torch.cuda.synchronize()
    return yd.cpu().numpy()


def _device_cbrt_ge1(x):
    ctx = Context.get(0)
    xd = torch.from_numpy(np.ascontiguousarray(x, dtype=np.float64)).cuda()
    yd = torch.empty_like(xd)
    check(ctx.lib.ne_cbrt_ge1_f64(ctx.handle, xd.data_ptr(), yd.data_ptr(), xd.numel(), stream_handle()),
          "ne_cbrt_ge1_f64")
    torch.cuda.synchronize()
    return yd.cpu().numpy()


def _device_sqrt(x):
    ctx = Context.get(0)
    xd = torch.from_numpy(np.ascontiguousarray(x, dtype=np.float64)).cuda()
    yd = torch.empty_like(xd)
    check(ctx.lib.ne_sqrt_f64(ctx.handle, xd.data_ptr(), yd.data_ptr(), xd.numel(), stream_handle()), "ne_sqrt_f64")
    torch.cuda.synchronize()
    return yd.cpu().numpy()


def _device_exp(x):
    ctx = Context.get(0)
    xd = torch.from_numpy(np.ascontiguousarray(x, dtype=np.float64)).cuda()
    yd = torch.empty_like(xd)
    check(ctx.lib.ne_exp_f64(ctx.handle, xd.data_ptr(), yd.data_ptr(), xd.numel(), stream_handle()), "ne_exp_f64")
    torch.cuda.synchronize()
    return yd.cpu().numpy()


def _ulps(y, ref):
    return np.abs(y - ref) / np.spacing(np.abs(ref))


def test_log_within_two_ulp(require_gpu):
    rng = np.random.default_rng(41)
    samples = {
        "reduced range [0.5, 2]": rng.uniform(0.5, 2.0, 1 << 20),
        "near 1": 1.0 + rng.uniform(-1e-3, 1e-3, 1 << 18) * 10.0 ** rng.uniform(-12, 0, 1 << 18),
        "heights / roughness": 10.0 ** rng.uniform(0.3, 7.0, 1 << 18),
        "Reynolds numbers": 10.0 ** rng.uniform(-4.0, 4.0, 1 << 18),
        "normal range": 10.0 ** rng.uniform(-307.0, 308.0, 1 << 20),
    }
    for name, x in samples.items():
        x = x[x != 1.0]
        y = _device_log(x)
        ref = np.log(x)
        worst = float(np.max(_ulps(y, ref)))
        print(f"[log] {name}: worst {worst:.2f} ulp over {x.size} arguments", flush=True)
        assert worst <= 2.0, name


LN2 = Decimal("0.6931471805599453094172321214581765680755001343602552541206800094933936")


def test_log_exact_values(require_gpu):
    k = np.arange(-1020, 1021)
    y = _device_log(np.ldexp(1.0, k))
    ref = np.array([float(int(kk) * LN2) for kk in k])          # k·ln 2 rounded once
    assert np.max(_ulps(y[k != 0], ref[k != 0])) <= 1.0
    assert y[k == 0][0] == 0.0 and _device_log(np.array([1.0]))[0] == 0.0
    assert np.isnan(_device_log(np.array([np.nan]))[0])


def test_atan_ge1_within_two_ulp(require_gpu):
    rng = np.random.default_rng(43)
    s = 1.0 + np.sqrt(2.0)
    samples = {
        "[1, 1 + √2)": rng.uniform(1.0, s, 1 << 20),
        "around 1 + √2": s + rng.uniform(-1e-6, 1e-6, 1 << 16),
        "stability arguments (1 − 15ζ)^¼, (1 + 2(1 − cζ)^⅓)/√3": np.concatenate(
            [(1.0 + 15.0 * 10.0 ** rng.uniform(-4, 3, 1 << 18)) ** 0.25,
             (1.0 + 2.0 * np.cbrt(1.0 + 34.15 * 10.0 ** rng.uniform(-4, 3, 1 << 18))) / np.sqrt(3.0)]),
        "[1 + √2, 1e308]": 10.0 ** rng.uniform(np.log10(s), 308.0, 1 << 20),
    }
    for name, x in samples.items():
        y = _device_atan_ge1(x)
        worst = float(np.max(_ulps(y, np.arctan(x))))
        print(f"[atan] {name}: worst {worst:.2f} ulp over {x.size} arguments", flush=True)
        assert worst <= 2.0, name
    y = _device_atan_ge1(np.array([1.0, np.inf, np.nan, np.nextafter(1.0, 2.0)]))
    assert y[0] == np.pi / 4 and y[1] == np.pi / 2 and np.isnan(y[2])
    assert _ulps(y[3:], np.arctan(np.array([np.nextafter(1.0, 2.0)])))[0] <= 1.0


def test_cbrt_ge1_within_one_ulp(require_gpu):
    rng = np.random.default_rng(47)
    samples = {
        "[1, 8]": rng.uniform(1.0, 8.0, 1 << 20),
        "stability arguments 1 − cζ": 1.0 + 34.15 * 10.0 ** rng.uniform(-4, 4, 1 << 19),
        "[8, 1e38]": 10.0 ** rng.uniform(np.log10(8.0), 38.0, 1 << 20),
        "cubes": np.arange(1.0, 1 << 17) ** 3,
    }
    for name, x in samples.items():
        y = _device_cbrt_ge1(x)
        worst = float(np.max(_ulps(y, np.cbrt(x))))
        print(f"[cbrt] {name}: worst {worst:.2f} ulp over {x.size} arguments", flush=True)
        assert worst <= 1.0, name
    y = _device_cbrt_ge1(np.array([1.0, 8.0, 27.0, 2.0 ** 120]))
    assert y[0] == 1.0 and y[1] == 2.0 and y[2] == 3.0 and y[3] == 2.0 ** 40


def test_sqrt_within_one_ulp(require_gpu):
    rng = np.random.default_rng(53)
    samples = {
        "[1, 4)": rng.uniform(1.0, 4.0, 1 << 20),
        "wind speeds squared": 10.0 ** rng.uniform(-4.0, 3.5, 1 << 18),
        "stability arguments 1 − 15ζ and their roots": np.concatenate(
            [1.0 + 15.0 * 10.0 ** rng.uniform(-4, 4, 1 << 18), np.sqrt(1.0 + 15.0 * 10.0 ** rng.uniform(-4, 4, 1 << 18))]),
        "normal range": 10.0 ** rng.uniform(-300.0, 300.0, 1 << 20),
    }
    for name, x in samples.items():
        y = _device_sqrt(x)
        worst = float(np.max(_ulps(y, np.sqrt(x))))
        print(f"[sqrt] {name}: worst {worst:.2f} ulp over {x.size} arguments", flush=True)
        assert worst <= 1.0, name
    sq = np.arange(1.0, 1 << 16) ** 2
    assert np.array_equal(_device_sqrt(sq), np.sqrt(sq))          # exact at perfect squares


def test_exp_within_two_ulp(require_gpu):
    rng = np.random.default_rng(59)
    samples = {
        "[-ln 2/2, ln 2/2]": rng.uniform(-0.35, 0.35, 1 << 20),
        "the solve's arguments": rng.uniform(-60.0, 60.0, 1 << 20),
        "[-700, 700]": rng.uniform(-700.0, 700.0, 1 << 20),
    }
    for name, x in samples.items():
        y = _device_exp(x)
        worst = float(np.max(_ulps(y, np.exp(x))))
        print(f"[exp] {name}: worst {worst:.2f} ulp over {x.size} arguments", flush=True)
        assert worst <= 2.0, name
    assert _device_exp(np.array([0.0]))[0] == 1.0


@pytest.mark.parametrize("n_in,n_out,n", [(5, 2, 1000003), (0, 1, 77), (8, 4, 4096)])
def test_stream_probe_writes_the_sums(require_gpu, n_in, n_out, n):
    """ne_stream_probe_f64, the kernel of known traffic the memory counters are read against:
    out_j = Σ_k in_k + j, summed in order from 0 (so equal to torch's same-order sum)."""
    import ctypes as C
    ctx = Context.get(0)
    g = torch.Generator(device="cuda").manual_seed(n)
    ins = [torch.rand(n, dtype=torch.float64, device="cuda", generator=g) for _ in range(n_in)]
    outs = [torch.full((n,), -1.0, dtype=torch.float64, device="cuda") for _ in range(n_out)]
    pin = (C.c_void_p * max(n_in, 1))(*[a.data_ptr() for a in ins])
    pout = (C.c_void_p * n_out)(*[a.data_ptr() for a in outs])
    check(ctx.lib.ne_stream_probe_f64(ctx.handle, pin, n_in, pout, n_out, n, stream_handle()), "ne_stream_probe_f64")
    torch.cuda.synchronize()
    s = torch.zeros(n, dtype=torch.float64, device="cuda")
    for a in ins:
        s = s + a
    for j, o in enumerate(outs):
        assert torch.equal(o, s + float(j)), j
