"""ndipc_pcg_solve_noalloc (device arrays in and out, the solver's work vectors in the
library's fixed per-GPU array) against ndipc_pcg_solve (host arrays, device work space
allocated per call): the same kernel, so the same bits, on two calls in a row, the second
with a smaller system; a system above the maximum is rejected."""
import numpy as np
import pytest
import torch

from nexdynipc_amd import scenes

pytestmark = pytest.mark.gpu

MAX_BLOCKS = 4096  # NDIPC_PCG_NOALLOC_MAX_BLOCKS (include/ndipc.h)


def _system(nb, seed):
    rng = np.random.default_rng(seed)
    diag = []
    for _ in range(nb):
        A = rng.normal(size=(6, 6))
        diag.append(A @ A.T + 6 * np.eye(6))
    offij = np.array([(i, i + 1) for i in range(nb - 1)] + [(0, nb - 1)], dtype=np.int32)
    offb = np.array([0.1 * rng.normal(size=(6, 6)) for _ in range(len(offij))])
    b = rng.normal(size=6 * nb)
    return np.array(diag), offij, offb, b


def _noalloc(e, diag, offij, offb, b, tol, max_it):
    dev = torch.device("cuda", 0)
    t = {
        "diag": torch.tensor(diag, dtype=torch.float64, device=dev).contiguous(),
        "offij": torch.tensor(offij, dtype=torch.int32, device=dev).contiguous(),
        "offb": torch.tensor(offb, dtype=torch.float64, device=dev).contiguous(),
        "b": torch.tensor(b, dtype=torch.float64, device=dev).contiguous(),
        "x": torch.full((len(b),), np.nan, dtype=torch.float64, device=dev),
        "it": torch.full((1,), -1, dtype=torch.int32, device=dev),
        "rel": torch.full((1,), np.nan, dtype=torch.float64, device=dev),
    }
    torch.cuda.synchronize()  # inputs written by torch's stream are complete
    e.pcg_solve_noalloc(len(diag), t["diag"].data_ptr(), len(offij), t["offij"].data_ptr(), t["offb"].data_ptr(),
                        t["b"].data_ptr(), tol, max_it, t["x"].data_ptr(), t["it"].data_ptr(), t["rel"].data_ptr())
    torch.cuda.synchronize()  # the context stream's work is complete
    return t["x"].cpu().numpy(), int(t["it"].cpu()[0]), float(t["rel"].cpu()[0])


def test_noalloc_pcg_matches_allocating_entry_point(engine_factory):
    sc = scenes.make_config("c1")
    e = engine_factory(sc)
    for nb, seed in ((300, 1), (37, 2)):  # the second call smaller than the first
        diag, offij, offb, b = _system(nb, seed)
        x0, it0, rel0 = e.pcg_solve(diag, offij, offb, b, tol=1e-14, max_it=2000)
        x1, it1, rel1 = _noalloc(e, diag, offij, offb, b, 1e-14, 2000)
        assert it1 == it0 and rel1 == rel0
        assert np.array_equal(x1, x0)
        assert rel0 <= 1e-13
    with pytest.raises(RuntimeError, match="exceed the maximum"):
        diag, offij, offb, b = _system(MAX_BLOCKS + 1, 3)
        _noalloc(e, diag, offij, offb, b, 1e-14, 10)
