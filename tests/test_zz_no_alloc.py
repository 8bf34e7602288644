"""GPU: the no-allocation search entry point (spf_search_batch_device_noalloc).

Bar: called through the C ABI twice in a row -- the second call with a
smaller batch, reusing the same fixed scratch -- it returns exactly what
spf_search_batch_device returns for the same queries (ids, counts, distance
bits), exact and mixed, filtered and unfiltered; a batch whose working-memory
bound exceeds the scratch is refused with SPF_ENOMEM and a message, before
anything is enqueued.
"""
from __future__ import annotations

import numpy as np
import pytest

from pydiskann_amd import DeviceIndex, build_index, PREC_EXACT, PREC_MIXED
from pydiskann_amd._lib import SpfError, SPF_ENOMEM

pytestmark = pytest.mark.gpu


def clustered(n, d, n_centers, seed, spread=0.12):
    centers = np.random.default_rng(99).uniform(-1, 1, size=(n_centers, d)).astype(np.float32)
    rng = np.random.default_rng(seed)
    return (centers[np.arange(n) % n_centers] + rng.normal(0, spread, size=(n, d))).astype(np.float32)


def run(dev, fn, q, k, n_probe, precision, rr, masks, torch, s):
    nq, dim = q.shape
    q_d = torch.from_numpy(q).cuda()
    ids = torch.empty((nq, k), dtype=torch.int64, device="cuda")
    dd = torch.empty((nq, k), dtype=torch.float32, device="cuda")
    cnt = torch.empty((nq,), dtype=torch.int32, device="cuda")
    m_d = torch.from_numpy(masks.view(np.int64)).cuda() if masks is not None else None
    torch.cuda.synchronize()
    fn(q_d.data_ptr(), nq, dim, k, n_probe, ids.data_ptr(), dd.data_ptr(), cnt.data_ptr(),
       stream=s.cuda_stream, precision=precision, rerank_size=rr,
       masks_ptr=m_d.data_ptr() if m_d is not None else 0)
    s.synchronize()
    return (ids.cpu().numpy().view(np.uint64), dd.cpu().numpy(), cnt.cpu().numpy().astype(np.uint32))


@pytest.mark.parametrize("precision,filtered", [(PREC_EXACT, False), (PREC_MIXED, False),
                                                (PREC_EXACT, True), (PREC_MIXED, True)])
def test_noalloc_matches_device_search_twice(precision, filtered):
    import torch
    x = clustered(30000, 64, 40, seed=1)
    ix = build_index(x, "l2", None, n_postings=128, seed=0)
    if filtered:
        ix.labels = (np.arange(ix.n_raw) % 11).astype(np.uint8)
    dev = DeviceIndex(ix)
    s = torch.cuda.Stream()
    k, n_probe, rr = 10, 12, (200 if precision else 0)
    for call, nq in enumerate((400, 37)):   # the second call is smaller
        q = clustered(nq, 64, 40, seed=20 + call)
        masks = (np.random.default_rng(call).integers(1, 2047, size=nq).astype(np.uint64)
                 if filtered else None)
        print(f"[case] noalloc precision={precision} filtered={filtered} call={call} nq={nq}",
              flush=True)
        want = run(dev, dev.search_device, q, k, n_probe, precision, rr, masks, torch, s)
        got = run(dev, dev.search_device_noalloc, q, k, n_probe, precision, rr, masks, torch, s)
        np.testing.assert_array_equal(got[2], want[2], err_msg=f"call {call}")
        np.testing.assert_array_equal(got[0], want[0], err_msg=f"call {call}")
        np.testing.assert_array_equal(got[1].view(np.uint32), want[1].view(np.uint32),
                                      err_msg=f"call {call}")


def test_noalloc_refuses_a_batch_beyond_the_scratch():
    import torch
    x = clustered(30000, 64, 40, seed=2)
    ix = build_index(x, "l2", None, n_postings=128, seed=0)
    dev = DeviceIndex(ix)
    nq = 200_000   # working-memory bound far above 64 MiB
    q_d = torch.zeros((nq, 64), dtype=torch.float32, device="cuda")
    ids = torch.empty((nq, 10), dtype=torch.int64, device="cuda")
    dd = torch.empty((nq, 10), dtype=torch.float32, device="cuda")
    cnt = torch.empty((nq,), dtype=torch.int32, device="cuda")
    with pytest.raises(SpfError) as ei:
        dev.search_device_noalloc(q_d.data_ptr(), nq, 64, 10, 12, ids.data_ptr(), dd.data_ptr(),
                                  cnt.data_ptr())
    assert ei.value.code == SPF_ENOMEM and "no-allocation scratch" in str(ei.value)
