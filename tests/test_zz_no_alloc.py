"""The no-allocation entry point nce_ce_energy_noalloc against nce_ce_energy, through the C ABI: two calls in a row,
the second on a smaller input, each bit-identical to the allocating entry point; sizes beyond the fixed scratch are
refused before any GPU work."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda", 0)


def _model(n, x, seed):
    from neuralce2_amd.synthetic import synthetic_ce, fcc_cell
    from neuralce2_amd.ce.classical_ce_gpu import GPUClusterExpansion
    at = fcc_cell(n, x, seed)
    g = GPUClusterExpansion(synthetic_ce(0), at, energy_scale=float(len(at)))
    return g, g.encode(at).astype(np.int8)


def test_noalloc_energy_equals_the_allocating_entry_point(torch_dev):
    import torch
    from neuralce2_amd import _lib
    rng = np.random.default_rng(3)
    for n, R in ((32, 64), (8, 5)):                  # 131,072 sites x 64 replicas, then 2,048 sites x 5
        g, sig = _model(n, 0.3, n)
        occ = torch.tensor(np.stack([rng.permutation(sig) for _ in range(R)]), device=torch_dev)
        E_ref = torch.empty(R, dtype=torch.float64, device=torch_dev)
        E_na = torch.full((R,), np.nan, dtype=torch.float64, device=torch_dev)
        s = torch.cuda.current_stream().cuda_stream
        rc = _lib.lib.nce_ce_energy(g.ctx.h, g.model, occ.data_ptr(), R, occ.shape[1], E_ref.data_ptr(), s)
        assert rc == 0, _lib.lib.nce_last_error()
        rc = _lib.lib.nce_ce_energy_noalloc(g.ctx.h, g.model, occ.data_ptr(), R, occ.shape[1], E_na.data_ptr(), s)
        assert rc == 0, _lib.lib.nce_last_error()
        torch.cuda.synchronize()
        assert torch.equal(E_ref, E_na), (n, R)


def test_noalloc_refuses_more_partials_than_its_scratch(torch_dev):
    import torch
    from neuralce2_amd import _lib
    g, sig = _model(32, 0.3, 0)                       # 131,072 cells -> 512 partial sums per replica
    R = (1 << 20) // 512 + 1
    occ = torch.zeros((1, sig.size), dtype=torch.int8, device=torch_dev)
    E = torch.zeros(1, dtype=torch.float64, device=torch_dev)
    # the size check comes first: the buffers are never read for a refused call
    rc = _lib.lib.nce_ce_energy_noalloc(g.ctx.h, g.model, occ.data_ptr(), R, sig.size, E.data_ptr(),
                                        torch.cuda.current_stream().cuda_stream)
    assert rc == -1 and b"no-allocation scratch" in _lib.lib.nce_last_error()
