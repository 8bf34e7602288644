"""GPU: the search server (spf_server_*): continuous batching of mixed
requests under a device-memory budget.

Bar: every request's result in the mix (submitted from several threads,
merged into shared batches, run on several streams at once) is bit-identical
to the same request searched alone; a sample is also checked against the
oracle; the working memory never exceeds a batch's reservation; requests above
the budget are refused at submit; a tight budget still completes every request.
"""
from __future__ import annotations

import threading

import numpy as np
import pytest

import oracle
from pydiskann_amd import DeviceIndex, Server, build_index, PREC_EXACT, PREC_MIXED
from pydiskann_amd._lib import SpfError, SPF_EDIM, SPF_ENOMEM

pytestmark = pytest.mark.gpu


def labelled_index(quant, metric, n=20000, d=64, n_post=200, rerank=0, seed=0):
    rng = np.random.default_rng(seed)
    c = rng.normal(size=(64, d))
    x = (c[rng.integers(0, 64, n)] + 0.3 * rng.normal(size=(n, d))).astype(np.float32)
    if metric in ("dot", "cosine"):
        x /= np.linalg.norm(x, axis=1, keepdims=True)
    ix = build_index(x, metric, quant, n_postings=n_post, rerank_size=rerank, seed=0)
    ix.labels = rng.integers(0, 64, size=ix.n_raw).astype(np.uint8)
    return ix, c


def requests(c, d, n_req, seed, metric, filtered_share=0.75, nq_max=256):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_req):
        nq = int(rng.integers(1, nq_max + 1))
        q = (c[rng.integers(0, c.shape[0], nq)] + 0.3 * rng.normal(size=(nq, d))).astype(np.float32)
        if metric in ("dot", "cosine"):
            q /= np.linalg.norm(q, axis=1, keepdims=True)
        masks = None
        if rng.random() < filtered_share:
            masks = np.zeros(nq, dtype=np.uint64)
            for i in range(nq):
                for b in rng.choice(64, size=rng.integers(1, 9), replace=False):
                    masks[i] |= np.uint64(1) << np.uint64(b)
        out.append(dict(q=q, k=int(rng.choice([10, 100])), n_probe=int(rng.choice([16, 64])),
                        masks=masks))
    return out


def run_mix(srv, reqs, precision, threads=4):
    res = [None] * len(reqs)
    err = []

    def worker(t):
        try:
            mine = list(range(t, len(reqs), threads))
            tickets = [srv.submit(reqs[i]["q"], reqs[i]["k"], reqs[i]["n_probe"],
                                  precision=precision, masks=reqs[i]["masks"]) for i in mine]
            for i, tk in zip(mine, tickets):
                res[i] = srv.wait(tk)
        except Exception as e:   # surfaced in the main thread
            err.append(e)

    th = [threading.Thread(target=worker, args=(t,)) for t in range(threads)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not err, err
    return res


def check_same_alone(dev, reqs, res, precision):
    for i, r in enumerate(reqs):
        ids, dd, cnt = dev.search(r["q"], r["k"], r["n_probe"], precision=precision,
                                  masks=r["masks"])
        m_ids, m_dd, m_cnt, info = res[i]
        np.testing.assert_array_equal(m_cnt, cnt, err_msg=f"request {i}")
        np.testing.assert_array_equal(m_ids, ids, err_msg=f"request {i}")
        np.testing.assert_array_equal(m_dd.view(np.uint32), dd.view(np.uint32),
                                      err_msg=f"request {i}")
        assert info["done_ns"] >= info["start_ns"] >= info["submit_ns"] > 0


@pytest.mark.parametrize("quant,metric,precision,rerank", [
    ("int8", "dot", PREC_EXACT, 100),      # the int8 inner-product index of the serving config
    (None, "l2", PREC_MIXED, 0),
    (None, "l2", PREC_EXACT, 0)])
def test_same_result_alone_and_in_the_mix(quant, metric, precision, rerank):
    ix, c = labelled_index(quant, metric, rerank=rerank)
    dev = DeviceIndex(ix)
    reqs = requests(c, ix.dim, 48, seed=1, metric=metric)
    print(f"[case] serving quant={quant} metric={metric} precision={precision} "
          f"rows={sum(r['q'].shape[0] for r in reqs)}", flush=True)
    srv = Server(dev, n_slots=3, max_batch=512)
    res = run_mix(srv, reqs, precision)
    st = srv.stats()
    srv.close()
    print(st)
    assert st["requests_done"] == len(reqs)
    assert st["rows_done"] == sum(r["q"].shape[0] for r in reqs)
    assert st["budget_overruns"] == 0
    assert st["batches"] < len(reqs)          # requests shared batches
    assert st["max_concurrent"] >= 2          # batches ran side by side
    check_same_alone(dev, reqs, res, precision)
    if precision == PREC_EXACT:               # and the oracle agrees on a sample
        o = oracle.Oracle(ix)
        for i in (0, 7):
            r = reqs[i]
            rid, rd, rc = o.search_batch(r["q"], r["k"], r["n_probe"], masks=r["masks"], threads=8)
            m_ids, m_dd, m_cnt, _ = res[i]
            np.testing.assert_array_equal(m_cnt, rc)
            for j in range(r["q"].shape[0]):
                c_ = int(rc[j])
                np.testing.assert_array_equal(m_ids[j, :c_], rid[j, :c_])
                np.testing.assert_array_equal(m_dd[j, :c_].view(np.uint32),
                                              rd[j, :c_].view(np.uint32))


def test_tight_budget_completes_and_refuses_oversized():
    ix, c = labelled_index("int8", "dot", rerank=100, seed=2)
    dev = DeviceIndex(ix)
    reqs = requests(c, ix.dim, 24, seed=3, metric="dot")
    budget = 24 << 20
    srv = Server(dev, n_slots=3, max_batch=1024, hbm_budget=budget)
    res = run_mix(srv, reqs, PREC_EXACT, threads=3)
    st = srv.stats()
    print(st)
    assert st["requests_done"] == len(reqs)
    assert st["peak_reserved_bytes"] <= budget
    assert st["budget_overruns"] == 0
    check_same_alone(dev, reqs, res, PREC_EXACT)
    tiny = Server(dev, n_slots=1, hbm_budget=4096)   # below any single query's bound
    with pytest.raises(SpfError) as e:
        tiny.submit(reqs[0]["q"], 10, 16)
    assert e.value.code == SPF_ENOMEM
    assert tiny.stats()["rejected"] == 1
    tiny.close()
    srv.close()


def test_server_argument_errors():
    ix, c = labelled_index(None, "l2", n=3000, n_post=30, seed=4)
    dev = DeviceIndex(ix)
    srv = Server(dev, n_slots=2)
    q = np.zeros((3, ix.dim + 1), np.float32)
    with pytest.raises(SpfError) as e:
        srv.submit(q, 10, 8)
    assert e.value.code == SPF_EDIM
    ix.labels = None
    plain = DeviceIndex(ix)
    psrv = Server(plain, n_slots=1)
    with pytest.raises(ValueError, match="without labels"):
        psrv.submit(np.zeros((2, ix.dim), np.float32), 10, 8, masks=np.ones(2, np.uint64))
    empty = psrv.submit(np.zeros((0, ix.dim), np.float32), 10, 8)
    ids, dd, cnt, info = psrv.wait(empty)
    assert cnt.shape == (0,)
    psrv.close()
    srv.close()
