"""Chunking.gpu_batch and the BLAS-1 helpers on the GPU against NumPy.

The pipeline only moves data, so chunked results match the unchunked
computation to rounding (torch may pick differently vectorised elementwise
kernels for a chunk and for the whole array); everything is checked against
NumPy at 1e-6 relative."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def cdata(rng, shape):
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)


def test_gpu_batch_axis0_host_and_device_args(gpu):
    from holotomocupy_amd.chunking import Chunking
    print("[case] gpu_batch axis 0", flush=True)
    rng = np.random.default_rng(1)
    x = cdata(rng, (37, 8, 16))
    y = cdata(rng, (37, 8, 16))
    w = torch.from_numpy(cdata(rng, (8, 16))).cuda()  # non-chunked device input
    out = np.zeros_like(x)
    ch = Chunking(1 << 24, 5)

    @ch.gpu_batch(axis_out=0, axis_inp=0, nout=1)
    def f(self, out, x, y, w, a):
        out.copy_(a * x + y * w)

    f(ch, out, x, y, w, 2.0)
    # all-device arguments: direct call on the whole arrays
    xd, yd = torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda()
    outd = torch.zeros_like(xd)
    f(ch, outd, xd, yd, w, 2.0)
    assert np.allclose(out, outd.cpu().numpy(), rtol=1e-6, atol=1e-6)
    ref = (2.0 * x.astype(np.complex128) + y * w.cpu().numpy())
    assert np.allclose(out, ref, rtol=1e-6, atol=1e-6)


def test_gpu_batch_axis1_and_padding(gpu):
    from holotomocupy_amd.chunking import Chunking
    print("[case] gpu_batch axis 1 and inp_pad", flush=True)
    rng = np.random.default_rng(2)
    x = rng.standard_normal((3, 23, 12)).astype(np.float32)
    out = np.zeros((4, 23, 12), dtype=np.float32)
    ch = Chunking(1 << 22, 4)

    @ch.gpu_batch(axis_out=1, axis_inp=1, nout=1)
    def g(self, out, x):
        out[:3].copy_(x * 3)
        out[3].copy_(x.sum(0))

    g(ch, out, x)
    assert np.allclose(out[:3], 3 * x) and np.allclose(out[3], x.sum(0), rtol=1e-6)
    # padded first input: each chunk sees `pad` extra rows (a 3-point stencil)
    pad = 2
    z = rng.standard_normal((30 + pad, 5, 4)).astype(np.float32)
    res = np.zeros((30, 5, 4), dtype=np.float32)

    @ch.gpu_batch(axis_out=0, axis_inp=0, nout=1, inp_pad=pad)
    def h(self, out, zp):
        n = out.shape[0]
        out.copy_(zp[:n] + zp[1:n + 1] + zp[2:n + 2])

    h(ch, res, z)
    assert np.allclose(res, z[:-2] + z[1:-1] + z[2:], rtol=1e-6)


@pytest.mark.parametrize("host", [True, False])
def test_blas1_helpers(gpu, host):
    from holotomocupy_amd.chunking import Chunking
    print(f"[case] blas1 helpers host={host}", flush=True)
    rng = np.random.default_rng(3)
    x = cdata(rng, (29, 16, 24))
    y = cdata(rng, (29, 16, 24))
    ch = Chunking(1 << 24, 6)
    a, b = 0.75, -1.25
    X = x if host else torch.from_numpy(x).cuda()
    Y = y if host else torch.from_numpy(y).cuda()
    got = ch.redot_batch(X, Y)
    ref = float(np.sum(x.real.astype(np.float64) * y.real + x.imag.astype(np.float64) * y.imag))
    assert abs(got - ref) < 1e-6 * abs(ref) + 1e-3
    out = np.zeros_like(x) if host else torch.zeros_like(X)
    ch.linear_batch(X, Y, a, b, out=out)
    o = out if host else out.cpu().numpy()
    assert np.allclose(o, a * x + b * y, rtol=1e-6, atol=1e-6)
    out2 = np.zeros_like(x) if host else torch.zeros_like(X)
    ch.mulc_batch(out2, X, a)
    o2 = out2 if host else out2.cpu().numpy()
    assert np.allclose(o2, a * x, rtol=1e-6, atol=1e-7)
    X2 = x.copy() if host else torch.from_numpy(x.copy()).cuda()
    r = ch.linear_redot_batch(X2, Y, a, b)
    xn = a * x.astype(np.complex128) + b * y
    assert np.allclose(X2 if host else X2.cpu().numpy(), xn, rtol=1e-5, atol=1e-6)
    rref = float(np.sum(y.real * xn.real + y.imag * xn.imag))
    assert abs(r - rref) < 1e-5 * abs(rref) + 1e-3


def test_redot_deterministic_and_unaligned(gpu):
    from holotomocupy_amd.chunking import Chunking
    print("[case] redot deterministic / unaligned views", flush=True)
    rng = np.random.default_rng(4)
    x = torch.from_numpy(rng.standard_normal(100003).astype(np.float32)).cuda()
    y = torch.from_numpy(rng.standard_normal(100003).astype(np.float32)).cuda()
    ch = Chunking(1 << 20, 4)
    vals = {ch.redot_batch(x, y) for _ in range(3)}
    assert len(vals) == 1
    xs, ys = x[1:], y[1:]  # 4-byte offset: scalar path
    ref = float(np.dot(xs.cpu().numpy().astype(np.float64), ys.cpu().numpy()))
    assert abs(ch.redot_batch(xs, ys) - ref) < 1e-6 * abs(ref) + 1e-3
    o = torch.empty_like(xs)
    ch.linear_batch(xs, ys, 2.0, 3.0, out=o)
    assert torch.allclose(o, 2 * xs + 3 * ys, rtol=1e-6, atol=1e-6)
