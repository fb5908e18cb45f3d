"""Differentiation: adjoint kernels, autograd functions and the FFT convolution.

* adjoint identity <A x, y> = <x, A* y> (real inner products) for r2c, c2r and
  c2c over every algorithm branch, in FP64;
* torch.autograd.gradcheck (central finite differences, FP64) of rfftn,
  irfftn, fftn and fft_conv2d;
* the convolution against its textbook definition (direct sum);
* the full-size convolution (16 x 2048^2 images, 63 x 63 kernel, loss
  sum((y - t)^2)): gradients against FP64 central finite differences at 64
  random coordinates.  The loss is exactly quadratic along any coordinate, so
  a central difference has no truncation error for any h; what limits it is
  the rounding noise of two 67M-term FP64 losses, about n^(1/2) eps |y| |2(y-t)|
  ~ 1e-10, divided by 2h.  With h = 1e-6 that floor is ~1e-4 absolute (bar
  2e-3), with h = 1e-2 it is ~1e-8 (bar 1e-7 relative).
"""
import numpy as np
import pytest
import torch

from impulse_amd import _lib, synth
from impulse_amd import autograd as AG

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _device():
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    torch.cuda.set_device(0)
    yield
    torch.cuda.synchronize()


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _dot_c(a, b):
    return float(torch.sum(a.real * b.real + a.imag * b.imag))


def _dot_r(a, b):
    return float(torch.sum(a * b))


CASES = [([3, 1], [1]), ([3, 2], [1]), ([3, 15], [1]), ([3, 16], [1]), ([3, 17], [1]), ([2, 4095], [1]),
         ([2, 4096], [1]), ([3, 1009], [1]), ([2, 20000], [1]), ([3, 20014], [1]), ([3, 10007], [1]),
         ([6, 10], [0, 1]), ([9, 8], [1, 0]), ([4, 6, 7], [0, 2]), ([5, 4, 9], [2, 1, 0])]


@pytest.mark.parametrize("shape,axes", CASES)
def test_adjoint_identity_real(shape, axes):
    print(f"case adjoint {shape} {axes}", flush=True)
    hshape = list(shape)
    hshape[axes[-1]] = shape[axes[-1]] // 2 + 1
    x = _t(synth.uniform(shape, -1, 1, seed=1))
    G = _t(synth.complex_normal(hshape, seed=2))
    for fwd in (True, False):
        X = torch.empty(hshape, dtype=torch.complex128, device="cuda")
        AG._run(_lib.R2C, shape, x, X, axes, fwd)
        xa = torch.empty(shape, dtype=torch.float64, device="cuda")
        AG._run(_lib.R2C_ADJOINT, shape, G, xa, axes, fwd)
        lhs, rhs = _dot_c(X, G), _dot_r(x, xa)
        assert abs(lhs - rhs) <= 1e-12 * (abs(lhs) + float(torch.linalg.vector_norm(X) * torch.linalg.vector_norm(G))), (fwd, lhs, rhs)
        # c2r and its adjoint
        y = _t(synth.uniform(shape, -1, 1, seed=3))
        z = torch.empty(shape, dtype=torch.float64, device="cuda")
        AG._run(_lib.C2R, shape, G, z, axes, fwd, 0.5)
        Ga = torch.empty(hshape, dtype=torch.complex128, device="cuda")
        AG._run(_lib.C2R_ADJOINT, shape, y, Ga, axes, fwd, 0.5)
        lhs, rhs = _dot_r(z, y), _dot_c(G, Ga)
        assert abs(lhs - rhs) <= 1e-12 * (abs(lhs) + float(torch.linalg.vector_norm(z) * torch.linalg.vector_norm(y))), (fwd, lhs, rhs)


@pytest.mark.parametrize("shape,axes", [([3, 17], [1]), ([2, 4096], [1]), ([3, 1009], [1]), ([2, 20000], [1]),
                                        ([6, 10], [0, 1])])
def test_adjoint_identity_c2c(shape, axes):
    x = _t(synth.complex_normal(shape, seed=4))
    y = _t(synth.complex_normal(shape, seed=5))
    for fwd in (True, False):
        Fx = torch.empty_like(x)
        AG._run(_lib.C2C, shape, x, Fx, axes, fwd)
        Fy = torch.empty_like(y)
        AG._run(_lib.C2C, shape, y, Fy, axes, not fwd)
        lhs, rhs = _dot_c(Fx, y), _dot_c(x, Fy)
        assert abs(lhs - rhs) <= 1e-12 * float(torch.linalg.vector_norm(Fx) * torch.linalg.vector_norm(y))


def test_gradcheck_transforms():
    x = torch.rand(3, 10, dtype=torch.float64, device="cuda", requires_grad=True)
    assert torch.autograd.gradcheck(lambda t: AG.rfftn(t, [1]), (x,), eps=1e-6, atol=1e-9)
    x2 = torch.rand(4, 9, dtype=torch.float64, device="cuda", requires_grad=True)
    assert torch.autograd.gradcheck(lambda t: AG.rfftn(t, [0, 1], False), (x2,), eps=1e-6, atol=1e-9)
    X = torch.randn(3, 6, dtype=torch.complex128, device="cuda", requires_grad=True)
    assert torch.autograd.gradcheck(lambda t: AG.irfftn(t, 10, [1], False, 0.1), (X,), eps=1e-6, atol=1e-9)
    X2 = torch.randn(5, 5, dtype=torch.complex128, device="cuda", requires_grad=True)
    assert torch.autograd.gradcheck(lambda t: AG.irfftn(t, 9, [0, 1]), (X2,), eps=1e-6, atol=1e-9)
    z = torch.randn(4, 7, dtype=torch.complex128, device="cuda", requires_grad=True)
    assert torch.autograd.gradcheck(lambda t: AG.fftn(t, [0, 1], True, 0.5), (z,), eps=1e-6, atol=1e-9)


def _direct_conv_same(x, k):
    B, H, W = x.shape
    KH, KW = k.shape
    cy, cx = (KH - 1) // 2, (KW - 1) // 2
    y = np.zeros_like(x)
    for a in range(KH):
        for b in range(KW):
            # y[i, j] += x[i + cy - a, j + cx - b] * k[a, b]
            di, dj = cy - a, cx - b
            i0, i1 = max(0, -di), min(H, H - di)
            j0, j1 = max(0, -dj), min(W, W - dj)
            y[:, i0:i1, j0:j1] += x[:, i0 + di:i1 + di, j0 + dj:j1 + dj] * k[a, b]
    return y


@pytest.mark.parametrize("shape,ks", [((2, 20, 24), (5, 7)), ((3, 17, 16), (4, 3)), ((1, 33, 31), (9, 9))])
def test_conv_matches_definition(shape, ks):
    x = synth.uniform(shape, 0, 1, seed=7)
    k = synth.normal(ks, seed=8)
    y = AG.fft_conv2d(_t(x), _t(k))
    torch.cuda.synchronize()
    ref = _direct_conv_same(x, k)
    assert np.linalg.norm(y.cpu().numpy() - ref) <= 1e-12 * np.linalg.norm(ref)
    # float32 in/out: FP64 arithmetic, float32 rounding only
    y32 = AG.fft_conv2d(_t(x.astype(np.float32)), _t(k.astype(np.float32)))
    ref32 = _direct_conv_same(x.astype(np.float32).astype(np.float64), k.astype(np.float32).astype(np.float64))
    assert np.linalg.norm(y32.cpu().numpy() - ref32) <= 1e-6 * np.linalg.norm(ref32)


def _direct_adjoints(x, k, g):
    """gx[p, q] = sum_ab g[p - cy + a, q - cx + b] k[a, b] and gk[a, b] = sum g[i, j] x[i + cy - a, j + cx - b]:
    the gradients of sum(y * g) for the 'same' convolution y, as direct sums."""
    B, H, W = x.shape
    KH, KW = k.shape
    cy, cx = (KH - 1) // 2, (KW - 1) // 2
    gx, gk = np.zeros_like(x), np.zeros_like(k)
    for a in range(KH):
        for b in range(KW):
            di, dj = cy - a, cx - b
            i0, i1 = max(0, -di), min(H, H - di)
            j0, j1 = max(0, -dj), min(W, W - dj)
            gx[:, i0 + di:i1 + di, j0 + dj:j1 + dj] += g[:, i0:i1, j0:j1] * k[a, b]
            gk[a, b] = np.sum(g[:, i0:i1, j0:j1] * x[:, i0 + di:i1 + di, j0 + dj:j1 + dj])
    return gx, gk


# the last two shapes' grids (columns or rows beyond 2,560 points) are beyond the fused LDS
# kernels and take the plan-based path
@pytest.mark.parametrize("shape,ks", [((2, 20, 24), (5, 7)), ((3, 17, 16), (4, 3)), ((1, 2050, 8), (63, 3)),
                                      ((2, 4, 2048), (3, 63)), ((1, 4700, 6), (9, 3)), ((1, 3, 5200), (3, 3))])
def test_conv_forward_and_backward_match_direct_sums(shape, ks):
    print(f"case conv direct sums {shape} {ks}", flush=True)
    x = synth.uniform(shape, 0, 1, seed=17)
    k = synth.normal(ks, seed=18)
    g = synth.normal(shape, seed=19)
    xt, kt = _t(x).requires_grad_(True), _t(k).requires_grad_(True)
    y = AG.fft_conv2d(xt, kt)
    torch.sum(y * _t(g)).backward()
    rx, rk = _direct_adjoints(x, k, g)
    assert np.linalg.norm(y.detach().cpu().numpy() - _direct_conv_same(x, k)) <= 1e-12 * np.linalg.norm(
        _direct_conv_same(x, k))
    assert np.linalg.norm(xt.grad.cpu().numpy() - rx) <= 1e-12 * np.linalg.norm(rx)
    assert np.linalg.norm(kt.grad.cpu().numpy() - rk) <= 1e-12 * np.linalg.norm(rk)


def test_gradcheck_conv():
    x = torch.rand(2, 12, 11, dtype=torch.float64, device="cuda", requires_grad=True)
    k = torch.randn(5, 3, dtype=torch.float64, device="cuda", requires_grad=True)
    assert torch.autograd.gradcheck(AG.fft_conv2d, (x, k), eps=1e-6, atol=1e-8)


def test_conv_full_size_finite_differences():
    # 16 images of 2048^2 (pixels U[0,1]), a 63x63 kernel N(0, 1/63^2), target N(0,1)
    B, H, K = 16, 2048, 63
    print("case conv 16x2048^2 k63", flush=True)
    x = _t(synth.uniform((B, H, H), 0.0, 1.0, seed=11).astype(np.float32)).double().requires_grad_(True)
    k = _t(synth.normal((K, K), seed=12, sigma=1.0 / K)).requires_grad_(True)
    t = _t(synth.normal((B, H, H), seed=13))
    y = AG.fft_conv2d(x, k)
    loss = torch.sum((y - t) ** 2)
    loss.backward()
    gx, gk = x.grad.detach(), k.grad.detach()
    rng = np.random.default_rng(0)
    worst = {1e-6: 0.0, 1e-2: 0.0}
    xd, kd = x.detach(), k.detach()
    with torch.no_grad():
        for c in range(64):
            img = c < 32
            if img:
                coord = (int(rng.integers(B)), int(rng.integers(H)), int(rng.integers(H)))
                g = float(gx[coord])
            else:
                coord = (int(rng.integers(K)), int(rng.integers(K)))
                g = float(gk[coord])
            for h in worst:
                if img:
                    xp, xm = xd.clone(), xd.clone()
                    xp[coord] += h
                    xm[coord] -= h
                    yp, ym = AG.fft_conv2d(xp, kd), AG.fft_conv2d(xm, kd)
                else:
                    kp, km = kd.clone(), kd.clone()
                    kp[coord] += h
                    km[coord] -= h
                    yp, ym = AG.fft_conv2d(xd, kp), AG.fft_conv2d(xd, km)
                # L(+) - L(-) = sum (y+ - y-)(y+ + y- - 2t)
                fd = float(torch.sum((yp - ym) * (yp + ym - 2 * t))) / (2 * h)
                worst[h] = max(worst[h], abs(fd - g) / max(1.0, abs(g)))
    print(f"conv full-size FD check over 64 coordinates: worst error h=1e-6 {worst[1e-6]:.3e}, "
          f"h=1e-2 {worst[1e-2]:.3e}")
    assert worst[1e-6] <= 2e-3
    assert worst[1e-2] <= 1e-7


def test_conv_full_size_float32_io():
    """The C5 workload with float32 images, kernel and gradients (the bench's dtype) against the
    float64 path on the same values: FP64 arithmetic, so only the float32 rounding of the outputs
    differs (rel-L2 bar 1e-6)."""
    B, H, K = 16, 2048, 63
    print("case conv 16x2048^2 k63 float32 io", flush=True)
    x32 = synth.uniform_device(B * H * H, 0.0, 1.0, seed=11, dtype=torch.float32).view(B, H, H)
    k32 = _t(synth.normal((K, K), seed=12, sigma=1.0 / K).astype(np.float32))
    g32 = synth.uniform_device(B * H * H, -1.0, 1.0, seed=13, dtype=torch.float32).view(B, H, H)
    c32 = AG._conv(B, H, H, K, K, torch.float32)
    c64 = AG._conv(B, H, H, K, K, torch.float64)
    y32, gx32, gk32 = torch.empty_like(x32), torch.empty_like(x32), torch.empty_like(k32)
    c32.forward(x32, k32, y32)
    c32.backward(x32, k32, g32, gx32, gk32)
    x64, k64, g64 = x32.double(), k32.double(), g32.double()
    y64, gx64, gk64 = torch.empty_like(x64), torch.empty_like(x64), torch.empty_like(k64)
    c64.forward(x64, k64, y64)
    c64.backward(x64, k64, g64, gx64, gk64)
    torch.cuda.synchronize()
    for a, b in ((y32, y64), (gx32, gx64), (gk32, gk64)):
        assert a.dtype == torch.float32
        err = float(torch.linalg.vector_norm(a.double() - b) / torch.linalg.vector_norm(b))
        assert err <= 1e-6, err


def test_gradients_through_conjugate_views():
    """torch passes lazily conjugated gradients (a conj bit, e.g. from a loss built on Y.conj()); the
    backward passes must read their values, not the raw memory behind the view.  Linear functionals:
    sum(x * grad) equals the loss exactly."""
    torch.manual_seed(1)
    for n in (64, 4095, 10007):
        x = torch.randn(3, n, dtype=torch.float64, device="cuda", requires_grad=True)
        Y = AG.rfftn(x, [1])
        G = torch.randn_like(Y)
        loss = (Y.conj() * G).real.sum()
        (gx,) = torch.autograd.grad(loss, x)
        assert abs(float(loss.detach()) - float((x.detach() * gx).sum())) <= 1e-10 * float(x.detach().norm() * gx.norm()), n
        Z = torch.randn(3, n // 2 + 1, dtype=torch.complex128, device="cuda", requires_grad=True)
        z = AG.irfftn(Z.conj(), n, [1])          # a conjugate view as the input, too
        w = torch.randn_like(z)
        loss2 = (z * w).sum()
        (gZ,) = torch.autograd.grad(loss2, Z)
        assert abs(float(loss2.detach()) - float((Z.detach().conj() * gZ).real.sum())) <= \
            1e-10 * float(Z.detach().abs().norm() * gZ.abs().norm()), n
    with pytest.raises(ValueError):
        from impulse_amd import fft as F
        F.DataDesc(torch.zeros(4, dtype=torch.complex128, device="cuda").conj(), [4])
