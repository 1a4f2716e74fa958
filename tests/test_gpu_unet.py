"""GPU checks of the learned-guess U-Net kernels (libhfsi hfsi_unet_*) against
a float64 CPU model of the same network built from torch.nn.functional
(conv3d / avg_pool3d / nearest interpolation / autograd) -- a library
routine, not reference code.

The device keeps weights, activations and activation gradients in bf16 with
float32 accumulation.  The CPU model rounds to bf16 at the same places (the
stored activations in the forward pass, the stored activation gradients in
the backward pass), so that what remains is float32 accumulation order and
the odd value that lands on the other side of a bf16 rounding boundary.
Tolerances:
  forward output            : ||y_gpu - y_cpu|| / ||y_cpu|| < 5e-3
  parameter gradients       : per layer ||g_gpu - g_cpu|| / ||g_cpu|| < 3e-2
  Adam step                 : float32 closed form of the first update, 2e-6
(The same comparisons against the model without the rounding are printed:
the bf16 storage alone moves the first layer's weight gradient by ~7 %.)
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


def _load():
    from easyfsi_amd import _lib as lib_mod

    return lib_mod, lib_mod.lib()


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


class _RoundBF16(torch.autograd.Function):
    """Identity up to bf16 storage: rounds the value on the way forward and
    its gradient on the way back."""

    @staticmethod
    def forward(ctx, x):
        return x.to(torch.bfloat16).to(x.dtype)

    @staticmethod
    def backward(ctx, g):
        return g.to(torch.bfloat16).to(g.dtype)


def _cpu_net(params, x, bf16_storage=True):
    """x: [1, 16, D, H, W] float64; params: name -> (W, b) float64 tensors."""
    R = _RoundBF16.apply if bf16_storage else (lambda t: t)

    def conv(name, t):
        w, b = params[name]
        return R(torch.relu(TF.conv3d(t, w, b, padding=1)))

    a1 = conv("E1A", x)
    s1 = conv("E1B", a1)
    a2 = conv("E2A", R(TF.avg_pool3d(s1, 2)))
    s2 = conv("E2B", a2)
    a3 = conv("B3A", R(TF.avg_pool3d(s2, 2)))
    b3 = conv("B3B", a3)
    u2 = R(torch.cat([TF.interpolate(b3, scale_factor=2, mode="nearest"), s2], dim=1))
    d2 = conv("D2B", conv("D2A", u2))
    u1 = R(torch.cat([TF.interpolate(d2, scale_factor=2, mode="nearest"), s1], dim=1))
    d1 = conv("D1B", conv("D1A", u1))
    w, b = params["OUT"]
    return (w[None, :, None, None, None] * d1).sum(dim=1) + b


def _bf16(a):
    return torch.from_numpy(np.asarray(a, np.float32)).to(torch.bfloat16).to(torch.float64)


def _setup(shape, seed):
    _lib, L = _load()
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
    exact = _cpu_net(_cpu_params(params), x, bf16_storage=False)[0].detach().numpy()
    got = y.cpu().numpy()
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    err_exact = np.linalg.norm(got - exact) / np.linalg.norm(exact)
    _say(f"  rel err {err:.2e} (against the model without bf16 storage: {err_exact:.2e})")
    assert err < 5e-3
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

    x = torch.zeros((1, 16) + shape, dtype=torch.float64)
    x[0, :5] = _bf16(x5).permute(3, 0, 1, 2)
    m = torch.from_numpy(mask.astype(bool))
    t = torch.from_numpy(target.astype(np.float64))
    ref = {}
    for storage in (True, False):
        cp = _cpu_params(params, grad=True)
        out = _cpu_net(cp, x, bf16_storage=storage)[0]
        ref_loss = ((out - t)[m] ** 2).sum() / (t[m] ** 2).sum()
        ref_loss.backward()
        ref[storage] = (float(ref_loss), cp)
    assert loss.value == pytest.approx(ref[True][0], rel=5e-3)
    gsplit = _split(grads.astype(np.float64))

    def rel(a, b):
        return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)

    for name, cin, cout in LAYERS + [("OUT", 16, 1)]:
        gw, gb = gsplit[name]
        errs = []
        for storage in (True, False):
            cp = ref[storage][1]
            gw_ref, gb_ref = cp[name][0].grad.numpy(), cp[name][1].grad.numpy()
            if name == "E1A":  # padded input channels carry zero input: compare the five real ones
                errs.append((rel(gw[:, :5], gw_ref[:, :5]), rel(gb, gb_ref)))
            else:
                errs.append((rel(gw, gw_ref), rel(gb, gb_ref)))
        (ew, eb), (ew0, eb0) = errs
        _say(f"  {name}: weight grad rel err {ew:.2e}, bias grad rel err {eb:.2e} "
             f"(without bf16 storage: {ew0:.2e}, {eb0:.2e})")
        assert ew < 3e-2 and eb < 3e-2, (name, ew, eb)
    L.hfsi_unet_destroy(h)


def test_adam_first_step_closed_form_and_training_reduces_loss():
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
    lr, b1, b2, eps, wd = 3e-3, 0.9, 0.999, 1e-8, 1e-4
    losses = []
    for step in range(30):
        loss = ctypes.c_double()
        _ok(L, L.hfsi_unet_zero_grads(h, stream))
        _ok(L, L.hfsi_unet_forward(h, xin.data_ptr(), y.data_ptr(), stream))
        _ok(L, L.hfsi_unet_loss(h, y.data_ptr(), tt.data_ptr(), None, dy.data_ptr(), ctypes.byref(loss), stream))
        _ok(L, L.hfsi_unet_backward(h, dy.data_ptr(), stream))
        if step == 0:
            p0 = np.zeros_like(params)
            g0 = np.zeros_like(params)
            _ok(L, L.hfsi_unet_get_params(h, p0.ctypes.data, stream))
            _ok(L, L.hfsi_unet_get_grads(h, g0.ctypes.data, stream))
        _ok(L, L.hfsi_unet_adam_step(h, lr, b1, b2, eps, wd, stream))
        if step == 0:
            p1 = np.zeros_like(params)
            _ok(L, L.hfsi_unet_get_params(h, p1.ctypes.data, stream))
            # first step from m = v = 0, float32: m = (1-b1) g', v = (1-b2) g'^2, g' = g + wd p,
            # p -= lr (m / (1-b1)) / (sqrt(v / (1-b2)) + eps)
            f = np.float32
            gp = g0 + f(wd) * p0
            m = f(1 - b1) * gp
            v = f(1 - b2) * gp * gp
            want = p0 - f(lr) * (m / f(1 - b1)) / (np.sqrt(v / f(1 - b2)) + f(eps))
            np.testing.assert_allclose(p1, want, rtol=2e-6, atol=2e-6 * lr)
        losses.append(loss.value)
    _say(f"  loss {losses[0]:.3f} -> {losses[-1]:.3f}")
    assert losses[-1] < 0.9 * losses[0]
    L.hfsi_unet_destroy(h)
