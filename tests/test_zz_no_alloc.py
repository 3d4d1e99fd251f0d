"""pfc_smooth_noalloc (include/pfc_relax.h): the Gaussian smoothing of
pfc_smooth on device buffers with all work memory in the library's static
device array.  Called through the C ABI on device buffers (torch only holds
the memory) and checked against pfc_smooth on the same fields: two calls in
a row, the second on a smaller grid, in both precisions, plus an in-place
call and the refusal of a grid beyond the array's limit.  The smoothing
itself (src/smooth.c:73-119) is pinned against numpy and a closed form in
tests/test_smooth.py."""
import numpy as np
import pytest

from pfc_amd import FP64, MIXED, LATTICE_A
from pfc_amd.engine import smooth, smooth_noalloc
from pfc_amd.inputs import crystal

pytestmark = pytest.mark.gpu

ESIZE = -4


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("prec,tol", [(FP64, 1e-13), (MIXED, 2e-6)])
def test_noalloc_smooth_matches_smooth(gpu, prec, tol):
    import torch
    dx = LATTICE_A / 8
    sigma = 0.35
    for Nx, Ny, seed in ((512, 256, 1), (128, 64, 2)):    # the second call is smaller
        f = crystal(Nx, Ny, dx, dx * 1.02, -0.2, 0.2, LATTICE_A, sym=3, theta=0.1, noise=0.05,
                    seed=seed)
        ref = smooth(f, dx, dx * 1.02, sigma, prec)
        d_in = torch.from_numpy(np.ascontiguousarray(f)).cuda()
        d_out = torch.full_like(d_in, float("nan"))
        torch.cuda.synchronize()
        rc = smooth_noalloc(d_in.data_ptr(), d_out.data_ptr(), Nx, Ny, dx, dx * 1.02, sigma, prec)
        assert rc == 0
        torch.cuda.synchronize()
        got = d_out.cpu().numpy()
        print(f"{Nx}x{Ny} precision {prec}: rel L2 vs pfc_smooth {rel_l2(got, ref):.2e}")
        assert np.isfinite(got).all()
        assert rel_l2(got, ref) < tol
        # in place (in == out)
        rc = smooth_noalloc(d_in.data_ptr(), d_in.data_ptr(), Nx, Ny, dx, dx * 1.02, sigma, prec)
        assert rc == 0
        torch.cuda.synchronize()
        assert rel_l2(d_in.cpu().numpy(), ref) < tol


def test_noalloc_refuses_beyond_limit(gpu):
    import torch
    d = torch.zeros(8, dtype=torch.float64, device="cuda")   # never touched
    # 4096 x 2048 points exceed the 2048 x 2048 limit; 1000 has no kernel
    assert smooth_noalloc(d.data_ptr(), d.data_ptr(), 4096, 2048, 0.7, 0.7, 0.3, FP64) == ESIZE
    assert smooth_noalloc(d.data_ptr(), d.data_ptr(), 1000, 64, 0.7, 0.7, 0.3, FP64) == ESIZE
    assert smooth_noalloc(d.data_ptr(), d.data_ptr(), 64, 64, 0.7, 0.7, -1.0, FP64) == -1
    torch.cuda.synchronize()
