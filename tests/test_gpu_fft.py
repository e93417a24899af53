"""The FFT wrapper's transforms on the GPU: the in-place C2R (execute_inplace_z2d,
cuda_wrapper.cu:252-254) in the FFTW padded layout against the out-of-place C2R
(bit for bit) and against numpy's inverse real FFT."""
import numpy as np
import pytest
import torch

from test_gpu_parity import host

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("N", [8, 32, 96])
def test_inplace_z2d_padded_layout(cfl, N):
    print(f"case inplace z2d N={N}", flush=True)
    rng = np.random.default_rng(N)
    f = torch.from_numpy(rng.standard_normal(N ** 3)).cuda()
    plan = cfl.SpectrumPlan(N)
    fk = plan.d2z(f)
    out_of_place = host(plan.z2d(fk.clone()))
    padded = host(plan.z2d_inplace(fk.clone())).reshape(N, N, 2 * (N // 2 + 1))
    assert np.array_equal(padded[:, :, :N].ravel(), out_of_place)
    ref = np.fft.irfftn(np.fft.rfftn(host(f).reshape(N, N, N)), s=(N, N, N), axes=(0, 1, 2)) * N ** 3
    assert np.max(np.abs(padded[:, :, :N] - ref)) <= 1e-12 * np.abs(ref).max()
    plan.close()
