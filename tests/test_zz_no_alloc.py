"""The allocation-free surrogate forward pass (dgs_forward_noalloc) against dgs_forward.

Both run the same kernels on the same padded rows, so the outputs are bitwise equal. Two calls in
a row, the second with fewer rows, show that the shared scratch array keeps no state that
matters between calls; a network wide enough to need several passes per call is checked too.
"""
import numpy as np
import pytest
import torch

import surrogate

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape", [dict(n_freq=64, width=256, depth=3, out_dim=512, max_batch=256),
                                   dict(n_freq=128, width=512, depth=5, out_dim=2048, max_batch=1024)])
def test_forward_noalloc_equals_forward(shape):
    print(f"dgs_forward_noalloc {shape}", flush=True)
    net = surrogate.Net(in_dim=6, freq_sigma=0.5, seed=3, **shape)
    rng = np.random.default_rng(11)
    p, B = net.get_params()
    net.set_params(p + 0.01 * rng.standard_normal(p.shape).astype(np.float32))
    for rows in (1000, 77):  # the second call is smaller than the first
        x = rng.random((rows, 6)).astype(np.float32)
        ref = net.forward(x)
        xd = torch.from_numpy(x).to("cuda:0")
        yd = torch.full((rows, shape["out_dim"]), float("nan"), dtype=torch.float32, device="cuda:0")
        torch.cuda.synchronize()
        net.forward_device(rows, xd.data_ptr(), yd.data_ptr())
        got = yd.cpu().numpy()
        print(f"  rows {rows}: max |diff| {np.abs(got - ref).max():.3e}", flush=True)
        assert np.array_equal(got, ref)
    net.close()


def test_forward_noalloc_rejects_what_does_not_fit():
    # 64 output rows of 66,560 FP32 samples alone take 17 MB, more than the 16 MiB array
    net = surrogate.Net(in_dim=6, n_freq=64, width=64, depth=1, out_dim=66560, max_batch=64, seed=1)
    xd = torch.zeros((64, 6), dtype=torch.float32, device="cuda:0")
    yd = torch.zeros((64, 66560), dtype=torch.float32, device="cuda:0")
    with pytest.raises(RuntimeError, match="scratch"):
        net.forward_device(64, xd.data_ptr(), yd.data_ptr())
    net.close()
