"""The warm-start network's MFMA kernels against torch's CPU conv3d in float64.

The GPU rounds inputs, weights, activations and the ReLU-masked upstream gradient to BF16 and
accumulates in FP32; the references round the same tensors to BF16 and compute in float64, so
a layer's outputs and gradients agree to FP32 accumulation error.
"""
import sys

import numpy as np
import pytest
import torch

from natray_amd import net


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()

VOX = 4096


def bf16(a: np.ndarray) -> torch.Tensor:
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to(torch.bfloat16).to(torch.float64)


def to_ncdhw(x: torch.Tensor) -> torch.Tensor:
    """[T][4096][C] (voxel 256 x + 16 y + z) -> [T][C][16][16][16] (x, y, z)."""
    t, _, c = x.shape
    return x.reshape(t, 16, 16, 16, c).permute(0, 4, 1, 2, 3)


def from_ncdhw(x: torch.Tensor) -> torch.Tensor:
    t, c = x.shape[:2]
    return x.permute(0, 2, 3, 4, 1).reshape(t, VOX, c)


def torch_weight(W: torch.Tensor) -> torch.Tensor:
    """W[tap][cin][cout], tap = 9 (dx+1) + 3 (dy+1) + dz+1 -> torch [cout][cin][3][3][3]."""
    cin, cout = W.shape[1:]
    return W.reshape(3, 3, 3, cin, cout).permute(4, 3, 0, 1, 2).contiguous()


def rel_err(a, b) -> float:
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


@pytest.mark.gpu
@pytest.mark.parametrize("cin,cout,relu", [(2, 16, True), (16, 16, True), (16, 1, False)])
def test_layer_matches_torch_conv3d(cin, cout, relu):
    say(f"[net] layer {cin}->{cout} relu={relu}")
    rng = np.random.default_rng(cin * 100 + cout)
    T = 3
    X = rng.standard_normal((T, VOX, cin)).astype(np.float32)
    W = (rng.standard_normal((27, cin, cout)) * np.sqrt(2.0 / (27 * cin))).astype(np.float32)
    b = rng.standard_normal(cout).astype(np.float32) * 0.1
    dY = rng.standard_normal((T, VOX, cout)).astype(np.float32)
    Y, dX, dW, db = net.layer_check(cin, cout, relu, X, W, b, dY)

    x = to_ncdhw(bf16(X)).requires_grad_(True)
    w = torch_weight(bf16(W)).requires_grad_(True)
    bb = torch.from_numpy(b.astype(np.float64)).requires_grad_(True)
    pre = torch.nn.functional.conv3d(x, w, bb, padding=1)
    y = torch.relu(pre) if relu else pre
    y_np = from_ncdhw(y.detach()).numpy()
    assert rel_err(Y, y_np) < 1e-5
    # The upstream gradient through the ReLU (on the BF16 output the GPU keeps), in BF16.
    pos = torch.from_numpy((from_ncdhw(y.detach()).to(torch.bfloat16).to(torch.float64) > 0).numpy()) if relu else None
    dz = torch.from_numpy(dY.astype(np.float64))
    if relu:
        dz = dz * pos
    dz = dz.to(torch.bfloat16).to(torch.float64)
    gx, gw, gb = torch.autograd.grad(pre, (x, w, bb), grad_outputs=to_ncdhw(dz))
    assert rel_err(dX, from_ncdhw(gx).numpy()) < 1e-5
    gw_ours = gw.permute(2, 3, 4, 1, 0).reshape(27, cin, cout).numpy()
    assert rel_err(dW, gw_ours) < 1e-4
    assert rel_err(db, gb.numpy()) < 1e-4


def reference_net(params: np.ndarray, X: np.ndarray):
    """The network in torch float64 with the GPU's BF16 roundings of input, weights and
    activations; returns (output [T][4096], parameter tensors)."""
    tens = []
    x = to_ncdhw(bf16(X))
    for i, (cin, cout, relu, wo, bo) in enumerate(net.LAYERS):
        W = bf16(params[wo:wo + 27 * cin * cout].reshape(27, cin, cout)).requires_grad_(True)
        b = torch.from_numpy(params[bo:bo + cout].astype(np.float64)).requires_grad_(True)
        tens += [W, b]
        x = torch.nn.functional.conv3d(x, torch_weight(W), b, padding=1)
        if relu:
            x = torch.relu(x)
            x = x + (x.to(torch.bfloat16).to(torch.float64) - x).detach()  # the stored BF16 activation
    return from_ncdhw(x)[:, :, 0], tens


@pytest.mark.gpu
def test_network_gradient_matches_torch():
    """A train step with lr = 0 gives the gradient of the masked MSE of the whole network;
    torch float64 on the same BF16-rounded forward gives the same within BF16 rounding of the
    backward pass (cosine above 0.999, relative error below 3%)."""
    say("[net] whole-network gradient")
    rng = np.random.default_rng(5)
    T = 2
    n = net.Net(max_tiles=T, seed=11)
    X = rng.standard_normal((T, VOX, 2)).astype(np.float32)
    X[:, :, 1] = np.sign(X[:, :, 1])
    target = rng.standard_normal((T, VOX)).astype(np.float32)
    mask = (rng.random((T, VOX)) < 0.7).astype(np.float32)
    loss = n.train_step(X, target, mask, lr=0.0)
    g = n.grads()
    p = n.params()
    out, tens = reference_net(p, X)
    m = torch.from_numpy(mask.astype(np.float64))
    ref_loss = (((out - torch.from_numpy(target.astype(np.float64))) * m) ** 2).sum() / m.sum()
    ref_loss.backward()
    gr = np.concatenate([t.grad.numpy().ravel() for t in tens])
    assert abs(loss - float(ref_loss)) <= 1e-3 * float(ref_loss)
    cos = float(np.dot(g, gr) / (np.linalg.norm(g) * np.linalg.norm(gr)))
    assert cos > 0.999, cos
    assert np.linalg.norm(g - gr) / np.linalg.norm(gr) < 3e-2
    fwd = n.forward(X)
    assert rel_err(fwd, out.detach().numpy()) < 1e-2
    n.close()


@pytest.mark.gpu
def test_training_reduces_the_loss():
    """Adam on a fixed batch: the network learns a smooth local filter of its first input."""
    say("[net] training on a fixed batch")
    rng = np.random.default_rng(9)
    T = 4
    n = net.Net(max_tiles=T, seed=3)
    X = np.zeros((T, VOX, 2), np.float32)
    X[:, :, 0] = rng.standard_normal((T, VOX))
    X[:, :, 1] = 1.0
    xt = torch.from_numpy(X[:, :, 0].reshape(T, 1, 16, 16, 16).astype(np.float64))
    target = (torch.nn.functional.avg_pool3d(xt, 3, stride=1, padding=1) * 2.0).reshape(T, VOX).numpy().astype(np.float32)
    mask = np.ones((T, VOX), np.float32)
    first = n.train_step(X, target, mask, lr=3e-3)
    for _ in range(150):
        last = n.train_step(X, target, mask, lr=3e-3)
    assert last < 0.2 * first, (first, last)
    n.close()


def random_operator(rng, T, frac=0.7, kc=7e-4):
    """A tile operator as the solve builds it: water voxels at random, diag = kc + 6 (no solid
    neighbours), bits for water neighbours inside the tile."""
    water = rng.random((T, 16, 16, 16)) < frac
    bits = np.zeros((T, 16, 16, 16), np.uint32)
    for axis in range(3):
        lo = [slice(None)] * 4
        hi = [slice(None)] * 4
        lo[axis + 1] = slice(1, None)
        hi[axis + 1] = slice(None, -1)
        lo, hi = tuple(lo), tuple(hi)
        bits[lo] |= np.where(water[lo] & water[hi], np.uint32(1 << (2 * axis)), np.uint32(0))
        bits[hi] |= np.where(water[hi] & water[lo], np.uint32(1 << (2 * axis + 1)), np.uint32(0))
    diag = np.where(water, np.float32(kc + 6.0), np.float32(0.0)).astype(np.float32)
    return diag.reshape(T, VOX), bits.reshape(T, VOX)


def torch_operator(diag, bits, o):
    T = o.shape[0]
    o3 = o.reshape(T, 16, 16, 16)
    b = torch.from_numpy(bits.reshape(T, 16, 16, 16).astype(np.int64))
    s = torch.zeros_like(o3)
    for axis in range(3):
        n = 16 - 1
        lo = torch.narrow(o3, axis + 1, 1, n)
        hi = torch.narrow(o3, axis + 1, 0, n)
        m_lo = (torch.narrow(b, axis + 1, 1, n) & (1 << (2 * axis))) != 0
        m_hi = (torch.narrow(b, axis + 1, 0, n) & (1 << (2 * axis + 1))) != 0
        pad_lo = [0, 0] * 3
        pad_hi = [0, 0] * 3
        # place contributions back at their voxel positions
        k = 2 * (2 - axis)
        pad_lo[k] = 1
        pad_hi[k + 1] = 1
        s = s + torch.nn.functional.pad(torch.where(m_lo, hi, torch.zeros_like(hi)), pad_lo)
        s = s + torch.nn.functional.pad(torch.where(m_hi, lo, torch.zeros_like(lo)), pad_hi)
    d = torch.from_numpy(diag.astype(np.float64)).reshape(T, 16, 16, 16)
    return torch.where(d > 0, d * o3 - s, torch.zeros_like(o3)).reshape(T, VOX)


@pytest.mark.gpu
def test_residual_loss_gradient_matches_torch():
    """The residual loss mean_water (x - A o)^2 and its gradient through the whole network
    against torch float64 on the same BF16-rounded forward pass."""
    say("[net] residual-loss gradient")
    rng = np.random.default_rng(21)
    T = 2
    n = net.Net(max_tiles=T, seed=4)
    X = rng.standard_normal((T, VOX, 2)).astype(np.float32)
    diag, bits = random_operator(rng, T)
    X[:, :, 1] = np.where(diag > 0, 1.0, 0.0)
    loss = n.train_step_residual(X, diag, bits, lr=0.0)
    g = n.grads()
    out, tens = reference_net(n.params(), X)
    x = torch.from_numpy(X[:, :, 0].astype(np.float64))
    w = torch.from_numpy(diag > 0)
    e = torch.where(w, x - torch_operator(diag, bits, out), torch.zeros_like(x))
    ref_loss = (e ** 2).sum() / w.sum()
    ref_loss.backward()
    gr = np.concatenate([t.grad.numpy().ravel() for t in tens])
    assert abs(loss - float(ref_loss.detach())) <= 1e-3 * float(ref_loss.detach())
    cos = float(np.dot(g, gr) / (np.linalg.norm(g) * np.linalg.norm(gr)))
    assert cos > 0.999, cos
    assert np.linalg.norm(g - gr) / np.linalg.norm(gr) < 3e-2
    n.close()
