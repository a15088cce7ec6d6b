"""fcg_mlp_gradient_static, the entry point that allocates no device memory
(its partial sums live in a file-scope device array of the library), against
fcg_mlp_gradient on the same network and batch, through the C-ABI."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _batch(n, seed):
    import torch
    g = torch.Generator().manual_seed(seed)
    X = torch.randn(n, 5, generator=g, dtype=torch.float32).cuda()
    T = torch.randn(n, 2, generator=g, dtype=torch.float32).cuda()
    return X, T


def test_static_gradient_equals_gradient_on_consecutive_calls():
    """Two calls in a row, the second with a smaller batch (fewer partial sums
    than the first left in the array): each equals fcg_mlp_gradient bit for
    bit, on the network's stream and on a caller's stream."""
    import torch
    from fesomc_amd import gpu
    net = gpu.Mlp(seed=11)
    G = torch.full((gpu.MLP_PARAMS,), float("nan"), dtype=torch.float32, device="cuda")
    side = torch.cuda.Stream()          # default priority
    stream = side.cuda_stream
    for n, st in ((50_000, None), (1_234, stream), (37, None)):
        X, T = _batch(n, seed=n)
        want = net.gradient(X.data_ptr(), T.data_ptr(), n)
        G.fill_(float("nan"))
        torch.cuda.synchronize()
        net.gradient_static(X.data_ptr(), T.data_ptr(), n, G.data_ptr(), stream=st)
        torch.cuda.synchronize()
        got = G.cpu().numpy()
        assert np.isfinite(got).all(), n
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), n
    net.close()


def test_static_gradient_rejects_an_empty_batch():
    import torch
    from fesomc_amd import gpu
    net = gpu.Mlp(seed=3)
    X, T = _batch(16, seed=1)
    G = torch.zeros(gpu.MLP_PARAMS, dtype=torch.float32, device="cuda")
    with pytest.raises(gpu.GpuError, match="bad argument"):
        net.gradient_static(X.data_ptr(), T.data_ptr(), 0, G.data_ptr())
    net.close()
