"""GPU checks of the learned-guess U-Net kernels (libhfsi hfsi_unet_*) against
a float64 CPU model of the same network built from torch.nn.functional
(conv3d / avg_pool3d / nearest interpolation / autograd) -- a library
routine, not reference code.

Tolerances (bf16 activations and weights on the device, float32
accumulation; float64 on the CPU):
  forward output            : ||y_gpu - y_cpu|| / ||y_cpu|| < 2e-2
  parameter gradients       : per layer ||g_gpu - g_cpu|| / ||g_cpu|| < 5e-2
"""

import ctypes
import sys

import numpy as np
import pytest
import torch
import torch.nn.functional as TF

pytestmark = pytest.mark.gpu

LAYERS = [("E1A", 16, 16), ("E1B", 16, 16), ("E2A", 16, 32), ("E2B", 32, 32), ("B3A", 32, 64), ("B3B", 64, 64),
          ("D2A", 96, 32), ("D2B", 32, 32), ("D1A", 48, 16), ("D1B", 16, 16)]


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _lib():
    from easyfsi_amd import _lib

    return _lib, _lib.lib()


def _ok(L, rc):
    assert rc == 0, L.hfsi_last_error().decode()


def _split(flat):
    """Parameter vector -> {name: (W [cout, cin, 3, 3, 3], b)} + output (w[16], b[1])."""
    out, off = {}, 0
    for name, cin, cout in LAYERS:
        n = cout * 27 * cin
        w = flat[off:off + n].reshape(cout, 27, cin).transpose(0, 2, 1).reshape(cout, cin, 3, 3, 3)
        off += n
        b = flat[off:off + cout]
        off += cout
        out[name] = (w, b)
    out["OUT"] = (flat[off:off + 16], flat[off + 16:off + 17])
    assert off + 17 == flat.size
    return out


def _cpu_net(params, x):
    """x: [1, 16, D, H, W] float64; params: name -> (W, b) float64 tensors."""
    def conv(name, t):
        w, b = params[name]
        return torch.relu(TF.conv3d(t, w, b, padding=1))

    a1 = conv("E1A", x)
    s1 = conv("E1B", a1)
    a2 = conv("E2A", TF.avg_pool3d(s1, 2))
    s2 = conv("E2B", a2)
    a3 = conv("B3A", TF.avg_pool3d(s2, 2))
    b3 = conv("B3B", a3)
    u2 = torch.cat([TF.interpolate(b3, scale_factor=2, mode="nearest"), s2], dim=1)
    d2 = conv("D2B", conv("D2A", u2))
    u1 = torch.cat([TF.interpolate(d2, scale_factor=2, mode="nearest"), s1], dim=1)
    d1 = conv("D1B", conv("D1A", u1))
    w, b = params["OUT"]
    return torch.einsum("c,ncdhw->ndhw", w, d1) + b


def _bf16(a):
    return torch.from_numpy(np.asarray(a, np.float32)).to(torch.bfloat16).to(torch.float64)


def _setup(shape, seed):
    _lib, L = _lib()
    h = ctypes.c_void_p()
    stream = torch.cuda.current_stream().cuda_stream
    _ok(L, L.hfsi_unet_create(ctypes.byref(h), *shape, 1, seed, stream))
    n = L.hfsi_unet_num_params(h)
    params = np.zeros(n, np.float32)
    _ok(L, L.hfsi_unet_get_params(h, params.ctypes.data, stream))
    # non-zero biases exercise the bias paths
    rng = np.random.default_rng(seed)
    split = _split(params)
    off = 0
    for name, cin, cout in LAYERS:
        off += cout * 27 * cin
        params[off:off + cout] = rng.normal(0, 0.05, cout)
        off += cout
    params[-1] = 0.1
    _ok(L, L.hfsi_unet_set_params(h, params.ctypes.data, stream))
    return _lib, L, h, params, stream


def _inputs(shape, seed):
    rng = np.random.default_rng(seed)
    x5 = rng.normal(size=shape + (5,)).astype(np.float32)
    x5[..., 1] = rng.random(shape) < 0.1
    return x5


def _cpu_params(params, grad=False):
    split = _split(params.astype(np.float64))
    out = {}
    for k, (w, b) in split.items():
        # the device computes with bf16 copies of the 3x3x3 weights; biases and the output layer stay float32
        wt = _bf16(w) if k != "OUT" else torch.from_numpy(np.asarray(w, np.float64))
        bt = torch.from_numpy(np.asarray(b, np.float64))
        wt.requires_grad_(grad)
        bt.requires_grad_(grad)
        out[k] = (wt, bt)
    return out


@pytest.mark.parametrize("shape", [(8, 12, 16), (16, 16, 8)])
def test_forward_matches_cpu_model(shape):
    _say(f"unet forward {shape}")
    _lib, L, h, params, stream = _setup(shape, 3)
    x5 = _inputs(shape, 4)
    xin = torch.from_numpy(x5).cuda()
    y = torch.zeros(shape, dtype=torch.float32, device="cuda")
    _ok(L, L.hfsi_unet_forward(h, xin.data_ptr(), y.data_ptr(), stream))
    torch.cuda.synchronize()
    x = torch.zeros((1, 16) + shape, dtype=torch.float64)
    x[0, :5] = _bf16(x5).permute(3, 0, 1, 2)
    ref = _cpu_net(_cpu_params(params), x)[0].detach().numpy()
    got = y.cpu().numpy()
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    _say(f"  rel err {err:.2e}")
    assert err < 2e-2
    L.hfsi_unet_destroy(h)


def test_gradients_match_cpu_autograd():
    _say("unet gradients")
    shape = (8, 8, 16)
    _lib, L, h, params, stream = _setup(shape, 7)
    x5 = _inputs(shape, 8)
    rng = np.random.default_rng(9)
    target = rng.normal(size=shape).astype(np.float32)
    mask = (rng.random(shape) < 0.9).astype(np.uint8)
    xin = torch.from_numpy(x5).cuda()
    tt = torch.from_numpy(target).cuda()
    mt = torch.from_numpy(mask).cuda()
    y = torch.zeros(shape, dtype=torch.float32, device="cuda")
    dy = torch.zeros_like(y)
    loss = ctypes.c_double()
    _ok(L, L.hfsi_unet_zero_grads(h, stream))
    _ok(L, L.hfsi_unet_forward(h, xin.data_ptr(), y.data_ptr(), stream))
    _ok(L, L.hfsi_unet_loss(h, y.data_ptr(), tt.data_ptr(), mt.data_ptr(), dy.data_ptr(), ctypes.byref(loss), stream))
    _ok(L, L.hfsi_unet_backward(h, dy.data_ptr(), stream))
    grads = np.zeros_like(params)
    _ok(L, L.hfsi_unet_get_grads(h, grads.ctypes.data, stream))

    cp = _cpu_params(params, grad=True)
    x = torch.zeros((1, 16) + shape, dtype=torch.float64)
    x[0, :5] = _bf16(x5).permute(3, 0, 1, 2)
    out = _cpu_net(cp, x)[0]
    m = torch.from_numpy(mask.astype(bool))
    t = torch.from_numpy(target.astype(np.float64))
    ref_loss = ((out - t)[m] ** 2).sum() / (t[m] ** 2).sum()
    assert loss.value == pytest.approx(float(ref_loss), rel=3e-2)
    ref_loss.backward()
    gsplit = _split(grads.astype(np.float64))
    for name, cin, cout in LAYERS + [("OUT", 16, 1)]:
        gw_ref = cp[name][0].grad.numpy()
        gb_ref = cp[name][1].grad.numpy()
        gw, gb = gsplit[name]
        if name == "E1A":  # padded input channels carry zero input: compare the five real ones
            gw, gw_ref = gw[:, :5], gw_ref[:, :5]
        ew = np.linalg.norm(gw - gw_ref) / max(np.linalg.norm(gw_ref), 1e-30)
        eb = np.linalg.norm(gb - gb_ref) / max(np.linalg.norm(gb_ref), 1e-30)
        _say(f"  {name}: weight grad rel err {ew:.2e}, bias grad rel err {eb:.2e}")
        assert ew < 5e-2 and eb < 5e-2, (name, ew, eb)
    L.hfsi_unet_destroy(h)


def test_adam_steps_reduce_loss():
    _say("unet adam")
    shape = (16, 16, 16)
    _lib, L, h, params, stream = _setup(shape, 11)
    x5 = _inputs(shape, 12)
    rng = np.random.default_rng(13)
    target = np.cumsum(rng.normal(size=shape), axis=2).astype(np.float32)
    xin = torch.from_numpy(x5).cuda()
    tt = torch.from_numpy(target).cuda()
    y = torch.zeros(shape, dtype=torch.float32, device="cuda")
    dy = torch.zeros_like(y)
    losses = []
    for _ in range(30):
        loss = ctypes.c_double()
        _ok(L, L.hfsi_unet_zero_grads(h, stream))
        _ok(L, L.hfsi_unet_forward(h, xin.data_ptr(), y.data_ptr(), stream))
        _ok(L, L.hfsi_unet_loss(h, y.data_ptr(), tt.data_ptr(), None, dy.data_ptr(), ctypes.byref(loss), stream))
        _ok(L, L.hfsi_unet_backward(h, dy.data_ptr(), stream))
        _ok(L, L.hfsi_unet_adam_step(h, 3e-3, 0.9, 0.999, 1e-8, 0.0, stream))
        losses.append(loss.value)
    _say(f"  loss {losses[0]:.3f} -> {losses[-1]:.3f}")
    assert losses[-1] < 0.5 * losses[0]
    L.hfsi_unet_destroy(h)
