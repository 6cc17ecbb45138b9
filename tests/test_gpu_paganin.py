"""GPU parity of the Paganin initial guess (holotomocupy_amd.paganin) against
the FP64 oracle (oracle/paganin.py): relative L2 <= 1e-5 on the phase (FP32
device FFTs)."""
import numpy as np
import pytest
import torch

from oracle import paganin as opg

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape", [(1, 64, 64), (4, 96, 80)])
def test_paganin_matches_oracle(gpu, shape):
    from holotomocupy_amd.paganin import multi_paganin, paganin_projection
    print(f"[case] paganin shape={shape}", flush=True)
    nd, ny, nx = shape
    rng = np.random.default_rng(nd + ny)
    yy, xx = np.mgrid[0:ny, 0:nx]
    base = 0.9 + 0.05 * np.cos(2 * np.pi * xx / nx) * np.sin(2 * np.pi * yy / ny)
    data = (base[None] * (1 + 0.01 * rng.standard_normal(shape))).astype(np.float32)
    dist = np.array([4.7e-3, 5.0e-3, 6.3e-3, 9.0e-3])[:nd]
    args = (dist, 7.25e-11, 3.0e-8, 120.0, 0.01)
    d = torch.from_numpy(data).cuda()
    ph = multi_paganin(d, *args).cpu().numpy()
    ref = opg.multi_paganin(data.astype(np.float64), *args)
    assert np.linalg.norm(ph - ref) / np.linalg.norm(ref) < 1e-5
    ph2 = paganin_projection(d, *args[:4], alpha=0.01).cpu().numpy()
    ref2 = opg.paganin_projection(data.astype(np.float64), *args[:4], alpha=0.01)
    assert ph2.shape == (ny, nx)
    assert np.linalg.norm(ph2 - ref2) / np.linalg.norm(ref2) < 1e-5
