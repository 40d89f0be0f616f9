"""The flux solve's FP64 logarithm (ne_log_f64) against numpy's, on the GPU: within 2 ulp over
the ranges the solve passes it (arguments between ½ and 2 after reduction, ratios of heights to
roughness lengths, Reynolds numbers) and over the whole normal range, and exact where the value
is: log 1 = 0, and log 2^k = k·ln 2 rounded once."""
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
