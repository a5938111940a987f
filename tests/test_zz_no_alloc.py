"""pmp_is_valid_batch_noalloc (GPU): the entry point that allocates nothing
gives the same booleans as pmp_is_valid_batch on two calls in a row (the
second smaller, so stale scratch contents from the first would show), for a
scene without clouds (unit pass, flags in the static scratch; and a batch
above the unit-pass limit) and a scene with a point cloud (cloud queries in
the static scratch); a batch whose workspaces exceed the 32 MiB scratch is
refused with ValueError."""

import numpy as np
import pytest

from plainmp_amd import synth
from support import oracle_cst, product_cst

pytestmark = pytest.mark.gpu


def _run(pc, Q, noalloc):
    import torch

    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        q = torch.from_numpy(np.ascontiguousarray(Q.T)).to("cuda")
        v = torch.full((len(Q),), 9, dtype=torch.uint8, device="cuda")
    stream.synchronize()
    f = pc._dev_is_valid_noalloc if noalloc else pc._dev_is_valid
    f(q.data_ptr(), len(Q), len(Q), v.data_ptr(), 0, stream.cuda_stream)
    stream.synchronize()
    return v.cpu().numpy()


@pytest.mark.parametrize("name,sizes", [("plain_two_bars", (2048, 777)), ("plain_two_bars", (20000, 3000)),
                                        ("cloud_shelves", (4096, 1000))])
def test_noalloc_matches_is_valid_batch(name, sizes):
    base = {"plain_two_bars": 1024, "cloud_shelves": 16384}[name]  # the workloads' full batch sizes
    wl = synth.workload(name, scale=max(sizes) / base)
    oc, pc = oracle_cst(wl), product_cst(wl)
    rng = np.random.default_rng(9)
    for n in sizes:
        Q = wl.Q[rng.permutation(len(wl.Q))[:n]]
        got = _run(pc, Q, True)
        ref = _run(pc, Q, False)
        assert set(np.unique(got)) <= {0, 1}
        assert np.array_equal(got, ref), (name, n)
        assert np.array_equal(got.astype(bool), oc.is_valid_batch(Q, 8).astype(bool)), (name, n)
        print(f"[noalloc] {name} B={n} valid_frac={got.mean():.3f}", flush=True)


def test_noalloc_refuses_oversized_batch():
    wl = synth.workload("cloud_shelves", scale=1.0)  # 16,384 configs x 90 spheres: 41 MB of cloud queries
    pc = product_cst(wl)
    with pytest.raises(ValueError, match="bytes of scratch"):
        _run(pc, wl.Q, True)
