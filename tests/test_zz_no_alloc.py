"""neat_raysample_count_noalloc (the count pass with its scan scratch in a static device array, for callers that
cannot allocate) against neat_raysample_count with a caller workspace, through the C ABI: offsets and the largest
per-ray sample count bit for bit, on two calls in a row, the second with fewer rays; and the size limit."""
import ctypes

import numpy as np
import pytest
import torch

from neat_amd import _abi, ops, workloads as W
from neat_amd._abi import NEAT_E_OVERFLOW, call, ptr, stream_ptr

pytestmark = pytest.mark.gpu
DEV = "cuda"


def count(sc, o, d, max_spn, noalloc):
    t = sc["tree"]
    pmin = torch.from_numpy(t["pos_min"]).to(DEV)
    pmax = torch.from_numpy(t["pos_max"]).to(DEV)
    flags = ops.tree_flags(sc["active"], t["max_depth"]).to(DEV)
    to, td = torch.from_numpy(o).to(DEV), torch.from_numpy(d).to(DEV)
    R = o.shape[0]
    seg = torch.full((R + 1,), -7, dtype=torch.int64, device=DEV)
    smp = torch.full((R + 1,), -7, dtype=torch.int64, device=DEV)
    mx = torch.full((1,), -7, dtype=torch.int32, device=DEV)
    common = (stream_ptr(), ptr(pmin), ptr(pmax), ptr(flags), pmin.shape[0], t["max_depth"], ptr(to), ptr(td), R,
              max_spn, ptr(seg), ptr(smp), ptr(mx))
    if noalloc:
        call("neat_raysample_count_noalloc", *common)
    else:
        nb = _abi.lib().neat_raysample_workspace_bytes(R)
        ws = torch.empty(max(nb, 8), dtype=torch.uint8, device=DEV)
        call("neat_raysample_count", *common, ptr(ws), nb)
    torch.cuda.synchronize()
    return seg.cpu().numpy(), smp.cpu().numpy(), int(mx.item())


def test_count_noalloc_matches_count_twice():
    print("case: neat_raysample_count_noalloc vs neat_raysample_count, C2 geometry, 300000 then 5000 rays",
          flush=True)
    sc = W.make_scene(W.C2)
    for n, seed in ((300000, 1), (5000, 2)):
        o, d = W.batch_rays(W.C2, n, seed=seed)
        a = count(sc, o, d, 64, noalloc=True)
        b = count(sc, o, d, 64, noalloc=False)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and a[2] == b[2], n
        assert a[0][-1] > 0 and a[1][-1] > a[0][-1]


def test_count_noalloc_refuses_more_rays_than_its_scratch():
    L = _abi.lib()
    dummy = ctypes.c_void_p(8)
    rc = L.neat_raysample_count_noalloc(None, dummy, dummy, dummy, 73, 2, dummy, dummy, (1 << 24) + 1, 32, dummy,
                                        dummy, None)
    assert rc == NEAT_E_OVERFLOW and "2^24" not in _abi.last_error() and "16777216" in _abi.last_error()
