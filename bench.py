#!/usr/bin/env python3
"""SPFresh batch-search benchmark on one MI355X.

Workload (default): 1M x 128 integer-valued f32 base (mixture of 1,024
Gaussians, centres U[0,200], sigma 20, clamped to [0,255]), 10k queries from the
same mixture, L2, 4,096 postings, n_probe 64, k 100.  One step = one batch
search of all 10k queries with queries and outputs resident in HBM.

The headline precision is mixed (bf16 matrix-core probe, f16 matrix-core scan
with f32 accumulate, exact f32 rerank of a 1,000-candidate pool); the plain f32
GPU path is timed in the same run on the same index and queries, and the line
carries both QPS values and their ratio.  Prints one JSON line: QPS, per-batch
p50/p99, recall@10 against an FP64 brute-force oracle on a query sample, the
roofline of the dominant kernel measured with HIP events inside the timed
region (HBM traffic per launch from the committed rocprofv3 counter summary of
the same kernel and workload, when there is one), and the CPU oracle timed on
this host.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "QPS at recall@10>=0.95, 1M x 128, mixed-precision vs fp32 GPU; p99 ms"
HBM_PEAK_GBS = 8000.0
RECALL_FLOOR = 0.95
# stage -> (kernel, algorithmic bytes per launch) for the roofline
STAGE_KERNELS = {
    "exact": {"scan": "k_scan", "select": "k_select", "rerank": "k_rerank"},
    "mixed": {"scan": "k_scan_mfma", "select": "k_pool_select", "rerank": "k_rerank_pruned"},
}


METRIC_NAMES = {0: "L2 (root)", 1: "L2", 2: "cosine", 3: "dot"}
QUANT_NAMES = {0: "f32", 1: "f16", 2: "int8", 3: "rabitq"}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def knn_routing_graph(ix, torch, r: int):
    """Routing graph for the graph-probe path: each centroid linked to its r
    nearest centroids (closest first), entry point the centroid nearest the
    mean.  The reference builds a Vamana graph instead (src/lib.rs:436-489);
    any graph exercises the same walk."""
    import numpy as np
    from pydiskann_amd.index import CentroidGraph
    c = torch.from_numpy(ix.centroids).cuda()
    n = c.shape[0]
    nbr = torch.empty((n, r), dtype=torch.int64)
    cn = (c * c).sum(1)
    for s in range(0, n, 4096):
        d = cn[s:s + 4096, None] + cn[None, :] - 2 * c[s:s + 4096] @ c.T
        d[torch.arange(d.shape[0]), torch.arange(s, s + d.shape[0])] = float("inf")
        nbr[s:s + 4096] = torch.topk(d, r, dim=1, largest=False).indices.cpu()
    medoid = int(((c - c.mean(0)) ** 2).sum(1).argmin())
    adj_off = np.arange(0, (n + 1) * r, r, dtype=np.uint32)
    return CentroidGraph(vectors=ix.centroids.copy(), adj_off=adj_off,
                         adj=nbr.numpy().astype(np.uint32).ravel(), live=np.ones(n, np.uint8),
                         seeds=[medoid], node_post=np.arange(n, dtype=np.int64), n_tombstones=0,
                         centroid_beam=64, base_n=n)


def rabitq_encode_device(torch, x, order, cent_of, padded, signs, chunk=131072):
    """RaBitQ codes of rows x[order] against their posting centroids, on the
    GPU, with the operation order of codecs.rabitq_encode (src/rabitq.rs:
    136-149, 218-256): residual, sequential norm fold, normalise, 3 x (sign
    flip, unnormalised FHT with the reference's butterfly pairing), sign bits,
    |r|^2 and the estimator factor.  Square roots and divisions run in f64 and
    round once to f32, which is the correctly rounded f32 result."""
    import numpy as np
    n, d = order.shape[0], x.shape[1]
    out = np.zeros((n, padded // 8 + 8), dtype=np.uint8)
    sg = torch.from_numpy(np.ascontiguousarray(signs)).cuda()
    sd = torch.sqrt(torch.tensor(float(padded), dtype=torch.float32, device="cuda"))
    bitw = (2 ** torch.arange(8, device="cuda", dtype=torch.int32)).view(1, 1, 8)
    for s0 in range(0, n, chunk):
        e0 = min(n, s0 + chunk)
        rows = torch.from_numpy(x[order[s0:e0]]).cuda()
        r = torch.zeros((e0 - s0, padded), dtype=torch.float32, device="cuda")
        r[:, :d] = rows - torch.from_numpy(cent_of[s0:e0]).cuda()
        sq = r * r
        acc = torch.zeros(e0 - s0, dtype=torch.float32, device="cuda")
        for i in range(padded):   # Iterator::sum: a left fold
            acc = acc + sq[:, i]
        norm = torch.sqrt(acc.double()).float()
        nz = norm > 0
        r[nz] = (r[nz].double() / norm[nz, None].double()).float()
        for rr in range(3):
            r = r * sg[rr][None, :]
            h = 1
            while h < padded:
                a = r.view(e0 - s0, padded // (2 * h), 2, h)
                xa, ya = a[:, :, 0, :].clone(), a[:, :, 1, :].clone()
                a[:, :, 0, :] = xa + ya
                a[:, :, 1, :] = xa - ya
                h *= 2
        ab = torch.abs(r)
        ip = torch.zeros(e0 - s0, dtype=torch.float32, device="cuda")
        for i in range(padded):
            ip = ip + ab[:, i]
        ip = (ip.double() / sd.double()).float()
        fac = torch.where(ip > 0, (norm.double() / ip.double()).float(), torch.zeros_like(ip))
        bits = ((r > 0).to(torch.int32).view(e0 - s0, padded // 8, 8) * bitw).sum(2)
        out[s0:e0, :padded // 8] = bits.to(torch.uint8).cpu().numpy()
        out[s0:e0, padded // 8:padded // 8 + 4] = (norm * norm).cpu().numpy().view(np.uint8).reshape(-1, 4)
        out[s0:e0, padded // 8 + 4:] = fac.cpu().numpy().view(np.uint8).reshape(-1, 4)
    return out


def build_c4(args, torch):
    """5M x 768 unit vectors (anisotropic spectrum), 8,192 postings by seeded
    k-means, RaBitQ codes padded to 1,024 dims (136-byte codes), cosine."""
    import numpy as np
    from pydiskann_amd import codecs
    from pydiskann_amd.datasets import config4
    from pydiskann_amd.index import IndexArrays, assign_nearest, kmeans
    t0 = time.time()
    base, queries = config4(n=args.n_base or 5_000_000, nq=args.nq)
    log(f"[bench] data {base.shape} queries {queries.shape} in {time.time() - t0:.1f}s")
    t0 = time.time()
    n, d = base.shape
    cent = kmeans(base, 8192, iters=6, seed=0, device="cuda")
    assign = assign_nearest(base, cent, device="cuda")
    cnt = np.bincount(assign, minlength=cent.shape[0])
    keep = np.nonzero(cnt > 0)[0]
    remap = np.full(cent.shape[0], -1, dtype=np.int64)
    remap[keep] = np.arange(keep.shape[0])
    assign = remap[assign]
    cent = np.ascontiguousarray(cent[keep])
    order = np.argsort(assign, kind="stable")
    post_off = np.zeros(cent.shape[0] + 1, dtype=np.uint64)
    post_off[1:] = np.cumsum(np.bincount(assign, minlength=cent.shape[0]))
    padded = codecs.rabitq_padded(d)
    signs = codecs.rabitq_signs(padded)
    log(f"[bench] partition {cent.shape[0]} postings in {time.time() - t0:.1f}s")
    t0 = time.time()
    cent_of = cent[assign[order]]
    codes = rabitq_encode_device(torch, base, order, cent_of, padded, signs)
    ref = codecs.rabitq_encode(base[order[:4096]], cent_of[:4096], padded, signs)
    if not np.array_equal(ref, codes[:4096]):
        raise RuntimeError("GPU RaBitQ encoder differs from the host codec")
    del cent_of
    log(f"[bench] RaBitQ codes ({codes.shape[1]} B, checked against the host codec on 4,096 "
        f"rows) in {time.time() - t0:.1f}s")
    ix = IndexArrays(dim=d, metric=2, quant=3, block_ids=np.arange(cent.shape[0], dtype=np.uint32),
                     centroids=cent, post_off=post_off, ids=order.astype(np.uint64), codes=codes,
                     raw=base, rq_padded=padded, rq_signs=signs)
    return base, queries, ix, 48, 10, np


def fp64_topk(torch, base, queries, k, metric):
    """FP64 brute-force top k on the GPU (recall reference)."""
    import numpy as np
    x = torch.from_numpy(base).cuda()
    out = []
    xn = (x.double() * x.double()).sum(1)
    for s in range(0, queries.shape[0], 50):
        q = torch.from_numpy(queries[s:s + 50]).cuda().double()
        ip = torch.cat([q @ x[c:c + 1_000_000].double().T for c in range(0, x.shape[0], 1_000_000)], 1)
        if metric in (0, 1):
            dd = xn[None, :] - 2 * ip + (q * q).sum(1)[:, None]
        elif metric == 3:
            dd = 1.0 - ip
        else:
            dd = 1.0 - ip / ((q * q).sum(1).sqrt()[:, None] * xn.sqrt()[None, :])
        out.append(torch.topk(dd, k, dim=1, largest=False).indices.cpu().numpy())
    return np.concatenate(out)


def build_workload(args, torch):
    import numpy as np
    from pydiskann_amd import build_index
    from pydiskann_amd.datasets import config1, config2, config3, labels

    if args.workload == "c4":
        return build_c4(args, torch)
    t0 = time.time()
    if args.workload in ("c3", "c5"):
        # 10M x 96 unit vectors, one posting per mixture component (Zipf
        # sizes), int8 codes, inner product, exact rerank of 400
        base, queries, mu = config3(n=args.n_base or 10_000_000, nq=args.nq, return_means=True)
        log(f"[bench] data {base.shape} queries {queries.shape} in {time.time() - t0:.1f}s")
        t0 = time.time()
        ix = build_index(base, "dot", "int8", centroids=mu, rerank_size=400, device="cuda")
        if args.workload == "c5":
            ix.labels = labels(ix.n_raw)
        if args.graph:
            ix.graph = knn_routing_graph(ix, torch, args.graph)
            log(f"[bench] routing graph: {ix.graph.n_nodes} nodes, {args.graph} neighbours each")
        log(f"[bench] index {ix.n_post} postings, max len {int(ix.posting_lengths().max())} "
            f"in {time.time() - t0:.1f}s")
        return base, queries, ix, 32, 10, np
    if args.workload == "c1":
        base, queries = config1()
        n_post, n_probe, k = 256, 16, 10
    else:
        base, queries = config2(n=args.n_base or 1_000_000, nq=args.nq)
        n_post, n_probe, k = 4096, 64, 100
    log(f"[bench] data {base.shape} queries {queries.shape} in {time.time() - t0:.1f}s")
    t0 = time.time()
    # one f32 index serves both precisions: exact scans the f32 codes, mixed
    # scans the library's f16 copy and reranks a pool from the f32 rows
    ix = build_index(base, "l2", None, n_postings=n_post, seed=0, kmeans_iters=8,
                     device="cuda")
    log(f"[bench] index {ix.n_post} postings, max len {int(ix.posting_lengths().max())} "
        f"in {time.time() - t0:.1f}s")
    return base, queries, ix, n_probe, k, np


def read_traffic(kernel: str, workload: str):
    """HBM bytes per launch of `kernel` on `workload` from the committed
    rocprofv3 counter summary profiles/<kernel>_<workload>.md (a line
    `traffic_bytes_per_launch: N`), or None when there is none."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload}.md")
    if not os.path.exists(path):
        return None, None
    with open(path) as f:
        for line in f:
            if line.strip().startswith("traffic_bytes_per_launch:"):
                return int(line.split(":", 1)[1].strip().split()[0]), os.path.relpath(path, ROOT)
    return None, None


def serve(args, queries, ix, np) -> int:
    """Filtered serving over the c3 index: requests of U{1..256} queries with
    k in {10, 100}, n_probe in {16, 64} and a 1-8 label allow mask per query,
    through the continuous-batching server.  First a closed loop (every
    request submitted at once) measures capacity; then open-loop Poisson
    arrivals at the target query rate(s) give per-request latency (server
    clock, submit to last row back) and the achieved rate."""
    import queue
    import threading

    from pydiskann_amd import DeviceIndex, Server, PREC_EXACT
    from pydiskann_amd._lib import FLAG_NO_GRAPH
    from pydiskann_amd.datasets import label_masks

    rng = np.random.default_rng(5)
    nq_all = queries.shape[0]
    masks_all = label_masks(nq_all)
    dev = DeviceIndex(ix, device=0)

    def make_requests(n, seed):
        r = np.random.default_rng(seed)
        out = []
        for _ in range(n):
            m = int(r.integers(1, args.req_rows + 1))
            rows = r.integers(0, nq_all, size=m)
            out.append((np.ascontiguousarray(queries[rows]), np.ascontiguousarray(masks_all[rows]),
                        int(r.choice([10, 100])), int(r.choice([16, 64]))))
        return out

    def run(reqs, rate_qps):
        """Submit reqs (all at once if rate_qps is 0, else Poisson arrivals at
        rate_qps queries/s); returns latencies (ms), wall seconds, stats."""
        srv = Server(dev, n_slots=args.slots, max_batch=args.max_batch,
                     hbm_budget=int(args.budget_gib * (1 << 30)))
        tickets = queue.Queue()
        lat = []

        def waiter():
            while True:
                t = tickets.get()
                if t is None:
                    return
                _, _, _, info = srv.wait(t)
                lat.append((info["done_ns"] - info["submit_ns"]) * 1e-6)

        wt = threading.Thread(target=waiter)
        wt.start()
        mean_rows = (args.req_rows + 1) / 2
        gaps = rng.exponential(mean_rows / rate_qps, size=len(reqs)) if rate_qps else None
        t0 = time.perf_counter()
        due = t0
        for i, (q, m, k, npr) in enumerate(reqs):
            if gaps is not None:
                due += gaps[i]
                d = due - time.perf_counter()
                if d > 0:
                    time.sleep(d)
            tickets.put(srv.submit(q, k, npr, precision=PREC_EXACT, masks=m,
                                   flags=FLAG_NO_GRAPH if args.no_graph else 0))
        tickets.put(None)
        wt.join()
        wall = time.perf_counter() - t0
        st = srv.stats()
        srv.close()
        return lat, wall, st

    def summary(lat, wall, rows):
        s = sorted(lat)
        pick = lambda p: s[min(len(s) - 1, int(round(p * (len(s) - 1))))]
        return {"qps": round(rows / wall, 1), "p50_ms": round(pick(0.5), 3),
                "p99_ms": round(pick(0.99), 3), "mean_ms": round(float(np.mean(s)), 3)}

    warm = make_requests(64, 1)
    run(warm, 0)   # grows every slot's working memory once
    log("[bench] c5: warmup done")
    reqs = make_requests(args.requests, 2)
    rows = sum(q.shape[0] for q, _, _, _ in reqs)
    lat, wall, st = run(reqs, 0)
    cap = summary(lat, wall, rows)
    log(f"[bench] c5 closed loop: {rows} queries in {args.requests} requests, {cap}, {st}")
    rates = [args.target_qps] if args.target_qps else [0.5 * cap["qps"], 0.8 * cap["qps"]]
    points = []
    for rate in rates:
        lat, wall, st2 = run(reqs, rate)
        p = summary(lat, wall, rows)
        p["target_qps"] = round(rate, 1)
        p["batches"], p["max_concurrent"] = st2["batches"], st2["max_concurrent"]
        points.append(p)
        log(f"[bench] c5 open loop at {rate:,.0f} QPS: {p}")
    head = points[-1]
    out = {
        "metric": "filtered serving: throughput and per-request p50/p99 latency under Poisson "
                  "arrivals (config 5)",
        "value": head["qps"], "unit": "queries/s", "n_gpus": 1, "higher_is_better": True,
        "p50_ms": head["p50_ms"], "p99_ms": head["p99_ms"], "target_qps": head["target_qps"],
        "load_points": points, "closed_loop": cap, "server_stats_closed_loop": st,
        "dtype": "int8 scan (int32 sums), f32 rerank",
        "data": "synthetic (seeded vMF-like mixture, labels U{0..63}, masks of 1-8 labels)",
        "config": {"workload": f"c5: {ix.n_raw}x{ix.dim} dot int8, {ix.n_post} postings, "
                               f"rerank 400, requests of U{{1..{args.req_rows}}} queries, "
                               f"k in {{10,100}}, n_probe in {{16,64}}, {args.requests} requests "
                               f"per point",
                   "slots": args.slots, "max_batch": args.max_batch,
                   "budget_gib": args.budget_gib,
                   "graph_replay": not args.no_graph},
        "graph_replays": int(dev.info().graph_replays),
    }
    print(json.dumps(out), flush=True)
    return 0

def sweep_point(args, dev, queries, q_d, ids_d, dd_d, cnt_d, stream, gt, ns, k, n_probe, pool, np):
    """QPS (10 timed batches after 2 untimed) and recall@10 on the recall
    sample at one (n_probe, pool) point of the headline precision."""
    import torch
    from pydiskann_amd import PREC_EXACT, PREC_MIXED
    from pydiskann_amd._lib import FLAG_NO_GRAPH
    prec = PREC_EXACT if args.precision == "exact" else PREC_MIXED
    nq, dim = queries.shape

    def step():
        dev.search_device(q_d.data_ptr(), nq, dim, k, n_probe, ids_d.data_ptr(), dd_d.data_ptr(),
                          cnt_d.data_ptr(), stream=stream.cuda_stream, precision=prec,
                          rerank_size=pool if prec else 0,
                          flags=FLAG_NO_GRAPH if args.no_graph else 0)

    for _ in range(2):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(10):
        step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / 10
    ids = ids_d.cpu().numpy().view(np.uint64)
    rec = float(np.mean([len(set(ids[i, :10].tolist()) & set(gt[i].tolist())) / 10.0
                         for i in range(ns)]))
    log(f"[bench] sweep n_probe {n_probe} pool {pool}: {nq / dt:,.0f} QPS, recall@10 {rec:.4f}")
    return {"n_probe": n_probe, "pool": pool if prec else None, "qps": round(nq / dt, 1),
            "ms_per_batch": round(dt * 1e3, 3), "recall_at_10": round(rec, 4)}


# MI355X peaks (dense): HBM3E bandwidth, f16/bf16 matrix cores
MFMA_F16_TFLOPS = 2500.0


def roofline_of(args, head, dim, nq, k, ix):
    """Roofline of the headline run's posting scan (its dominant kernel).

    `achieved` = the scan's minimal HBM bytes per launch (the library's
    profile: each probed posting's entries -- codes, norms, tombstone bits,
    labels -- read once, each query's row read once, one key per candidate
    slot written) over the scan's HIP-event time on the call's stream; `bound`
    follows from the kernel's arithmetic intensity against the ridge point.
    Beside it: the bytes under the work items (every item rereads its
    posting's entries and its queries' rows: the L2-side traffic), the
    matrix-core utilisation of the scan and of the probe, and, from the
    committed rocprofv3 counter summary, the DRAM bytes actually moved."""
    stage, prof = head["stage"], head["prof"]
    kern = STAGE_KERNELS[args.precision]
    if ix.quant == 3:
        kern = {"scan": "k_rq_scan_mfma" if args.precision == "mixed" else "k_rq_search_exact",
                "select": "k_pool_select", "rerank": "k_rerank_f16"}
    ms = stage["scan"]
    alg = prof["scan_hbm_bytes"]
    ach = alg / (ms * 1e-3) / 1e9
    frac = ach / HBM_PEAK_GBS
    assert 0.0 < frac <= 1.0, f"roofline fraction {frac} outside (0, 1]"
    item = prof["scan_bytes"]
    flops = prof["scan_flops"]
    ai = flops / alg if alg else 0.0
    # int8 matrix cores (5.0 POP/s dense) for int8 and 1-bit postings, else f16 (2.5 PFLOP/s)
    peak = 5000.0 if (ix.quant in (2, 3) and args.precision == "mixed") else MFMA_F16_TFLOPS
    if args.precision == "exact":
        peak = 78.6 if ix.quant != 3 else peak   # the f32 VALU scan
    ridge = peak * 1e12 / (HBM_PEAK_GBS * 1e9)
    wl = f"{args.workload}_{args.precision}_{ix.n_raw // 1000}kx{dim}"
    traffic, traffic_src = read_traffic(kern["scan"], wl)
    out = {"bound": "hbm" if ai < ridge else ("mfma" if args.precision == "mixed" else "valu"),
           "achieved": round(ach, 1),
           "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(frac, 4), "traffic": traffic,
           "kernel": kern["scan"], "kernel_ms": round(ms, 4),
           "algorithmic_bytes_per_launch": int(alg),
           "item_bytes_per_launch": int(item),
           "item_GBs": round(item / (ms * 1e-3) / 1e9, 1),
           "arithmetic_intensity_flop_per_byte": round(ai, 2),
           "ridge_flop_per_byte": round(ridge, 1),
           "candidates_per_launch": prof["candidates"], "traffic_source": traffic_src}
    if flops:
        tf = flops / (ms * 1e-3) / 1e12
        out["scan_compute"] = {"tflops": round(tf, 1), "peak_tflops": peak,
                               "frac": round(tf / peak, 4), "flops": int(flops)}
        if out["bound"] != "hbm":   # compute-bound: the roofline is the compute one
            assert 0.0 < tf / peak <= 1.0, f"compute fraction {tf / peak} outside (0, 1]"
            out.update(achieved=round(tf, 1), peak=peak, unit="TFLOP/s", frac=round(tf / peak, 4),
                       hbm_GBs=round(ach, 1), hbm_frac=round(frac, 4))
    if prof["probe_flops"] and stage["probe"] > 0:
        tf = prof["probe_flops"] / (stage["probe"] * 1e-3) / 1e12
        out["probe_mfma" if args.precision == "mixed" else "probe_valu"] = {
            "kernel": "k_probe_mfma" if args.precision == "mixed" else "k_scan (dense)",
            "ms": round(stage["probe"], 4), "tflops": round(tf, 1),
            "peak_tflops": MFMA_F16_TFLOPS if args.precision == "mixed" else 78.6,
            "frac": round(tf / (MFMA_F16_TFLOPS if args.precision == "mixed" else 78.6), 4)}
    if traffic:
        # DRAM bytes the kernel actually moved (rocprofv3 counters) over its live time
        tgb = traffic / (ms * 1e-3) / 1e9
        out["traffic_GBs"] = round(tgb, 1)
        out["traffic_frac"] = round(tgb / HBM_PEAK_GBS, 4)
    return out


def dtype_of(args, ix) -> str:
    if args.precision == "exact":
        return {0: "f32", 1: "f16 codes, f32", 2: "int8 codes (int32 L2 sums / f32 decode), f32 rerank",
                3: "1-bit codes (popcount estimator or f32 decode), f32"}[ix.quant]
    if ix.quant == 2:
        return "int8 scan (int32 sums), bf16 probe, f32 rerank"
    if ix.quant == 3:
        return "1-bit codes (int8 matrix-core estimator), bf16 probe, f16 then f32 rerank"
    return "f16/bf16 scan+probe, f32 rerank"


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(args, ix, queries, k, n_probe, np):
    """The oracle (a C restatement of Inner::search) on this host's cores: one
    task per query like the reference's rayon par_iter, all usable cores (at
    most 16, the host share of one GPU) and 1 core, one warm-up then the median
    of 5 runs over a bounded query sample.  Legs: the exact f32 search of the
    same index, and (L2 f32 index) the reference's like-for-like form of the
    mixed path -- F16 postings of the same partition with an exact rerank of
    the `--rerank` best (src/spfresh.rs:695-723)."""
    import oracle
    from pydiskann_amd import codecs
    from pydiskann_amd.index import IndexArrays
    import dataclasses
    nproc = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    cores = max(1, min(16, nproc or 1))
    nq = queries.shape[0]
    # (leg, index, queries on all cores, queries on 1 core)
    legs = [("exact", ix, args.cpu_sample, args.cpu_sample // 10)]
    if ix.quant == 0 and args.workload != "c1":
        legs.append((f"f16_rerank{args.rerank}", IndexArrays(
            dim=ix.dim, metric=ix.metric, quant=1, block_ids=ix.block_ids,
            centroids=ix.centroids, post_off=ix.post_off, ids=ix.ids,
            codes=codecs.f16_encode(ix.raw[ix.ids.astype(np.int64)]), raw=ix.raw,
            deleted=ix.deleted, rerank_size=args.rerank, probe_ratio=ix.probe_ratio),
            args.cpu_sample, args.cpu_sample // 10))
    if ix.quant == 3 and ix.metric == 2:
        # cosine decodes every probed code (src/spfresh.rs:346-348): a small
        # sample; the L2 bound path on the same unit vectors beside it
        legs = [("exact_cosine_decode", ix, 32, 4),
                ("l2_bound_path", dataclasses.replace(ix, metric=1), 200, 20)]
    out = {}
    for name, lix, n_all, n_one in legs:
        o = oracle.Oracle(lix)
        for threads, ns in ((cores, min(n_all, nq)), (1, min(n_one, nq))):
            qs = np.ascontiguousarray(queries[:ns])
            o.search_batch(qs, k, n_probe, threads=threads)   # warm-up
            ts = []
            for _ in range(5):
                t0 = time.perf_counter()
                o.search_batch(qs, k, n_probe, threads=threads)
                ts.append(time.perf_counter() - t0)
            t = float(np.median(ts))
            out[f"{name}_{threads}t"] = {"qps": round(ns / t, 1), "threads": threads,
                                        "queries": ns, "median_s": round(t, 4)}
            log(f"[bench] cpu {name} {threads} threads: {ns / t:,.1f} QPS ({ns} queries)")
    head = next(k for k in (f"f16_rerank{args.rerank}_{cores}t", f"exact_{cores}t",
                            f"exact_cosine_decode_{cores}t") if k in out)
    like = ("F16 postings + exact rerank of " + str(args.rerank) if head.startswith("f16")
            else "the cosine decode path" if "cosine" in head else "the exact path")
    return {"value": out[head]["qps"], "unit": "queries/s", "cores": cores, "kind": "port",
            "sample": f"first {out[head]['queries']} of the {nq} queries ({like}, the same "
                      f"partition), C restatement of Inner::search, one warm-up then the "
                      f"median of 5 runs; 1-thread legs over a tenth or less of that",
            "nproc": nproc, "cpu_model": cpu_model(), "legs": out}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", choices=["c1", "c2", "c3", "c4", "c5"], default="c2",
                    help="c2: the headline batch search; c3: int8 inner-product batch search "
                         "over 10M x 96; c4: RaBitQ cosine over 5M x 768; c5: filtered serving "
                         "over the c3 index")
    ap.add_argument("--precision", choices=["exact", "mixed"], default="mixed")
    ap.add_argument("--no-compare", action="store_true",
                    help="skip timing the other precision")
    ap.add_argument("--rerank", type=int, default=0,
                    help="mixed-precision rerank pool (0 = the workload's: c2 1,000, c3 400, "
                         "c4 300)")
    ap.add_argument("--n-base", type=int, default=0, help="base size (0 = the workload's)")
    ap.add_argument("--nq", type=int, default=10_000)
    ap.add_argument("--cpu-sample", type=int, default=1000)
    ap.add_argument("--recall-sample", type=int, default=500)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--graph", type=int, default=0,
                    help="c3/c5: probe through a routing graph of this many neighbours per "
                         "centroid (the reference's path above 4,096 postings)")
    ap.add_argument("--admit", action="store_true",
                    help="mixed: two-phase scan with admission (SPF_FLAG_ADMIT)")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the default launch shapes (by default each precision is first "
                         "autotuned on the batch itself, untimed)")
    ap.add_argument("--tune-cache", default="",
                    help="reuse / record the autotuner's choice in this file")
    ap.add_argument("--requests", type=int, default=4000, help="c5: requests per load point")
    ap.add_argument("--target-qps", type=float, default=0,
                    help="c5: open-loop query rate (0 = 50%% and 80%% of measured capacity)")
    ap.add_argument("--slots", type=int, default=3, help="c5: concurrent device batches")
    ap.add_argument("--max-batch", type=int, default=1024, help="c5: rows per device batch")
    ap.add_argument("--budget-gib", type=float, default=16.0, help="c5: working-memory budget")
    ap.add_argument("--req-rows", type=int, default=256,
                    help="c5: request sizes uniform in 1..this many query rows")
    ap.add_argument("--no-graph", action="store_true",
                    help="never replay captured launch chains (SPF_FLAG_NO_GRAPH)")
    ap.add_argument("--sweep", default="",
                    help="batch workloads: comma-separated n_probe[:pool] points, each timed "
                         "and scored for recall@10 after the headline run")
    args = ap.parse_args()
    if not args.rerank:
        args.rerank = {"c3": 400, "c5": 400, "c4": 300}.get(args.workload, 1000)
    if args.gpus != 1:
        print(json.dumps({"error": "runs on one GPU; --gpus above 1 is not supported "
                          "(sharding queries or postings over GPUs is a design, not built)"}))
        return 2

    import torch
    from pydiskann_amd import DeviceIndex, PREC_EXACT, PREC_MIXED
    from pydiskann_amd._lib import last_profile, FLAG_ADMIT, FLAG_NO_GRAPH
    from pydiskann_amd.index import brute_force_f64

    base, queries, ix, n_probe, k, np = build_workload(args, torch)
    if args.workload == "c5":
        return serve(args, queries, ix, np)
    nq, dim = queries.shape
    dev = DeviceIndex(ix, device=0)
    q_d = torch.from_numpy(queries).cuda()
    ids_d = torch.empty((nq, k), dtype=torch.int64, device="cuda")
    dd_d = torch.empty((nq, k), dtype=torch.float32, device="cuda")
    cnt_d = torch.empty((nq,), dtype=torch.int32, device="cuda")
    stream = torch.cuda.current_stream()
    ns = min(args.recall_sample, nq)
    gt = fp64_topk(torch, base, queries[:ns], 10, ix.metric)

    def run(precision: str):
        prec = PREC_EXACT if precision == "exact" else PREC_MIXED
        tune = None
        if not args.no_tune:
            tune = dev.autotune(queries, k, n_probe, precision=prec,
                                rerank_size=args.rerank if prec else 0,
                                cache_path=args.tune_cache or None)
            log(f"[bench] {precision}: autotune {tune}")

        def step(profile=False):
            dev.search_device(q_d.data_ptr(), nq, dim, k, n_probe, ids_d.data_ptr(),
                              dd_d.data_ptr(), cnt_d.data_ptr(), stream=stream.cuda_stream,
                              precision=prec, rerank_size=args.rerank if prec else 0,
                              profile=profile, flags=(FLAG_ADMIT if args.admit else 0) |
                              (FLAG_NO_GRAPH if args.no_graph else 0))

        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        log(f"[bench] {precision}: warmup done ({args.warmup} steps)")
        lat = []
        torch.cuda.synchronize()
        t_start = time.perf_counter()
        for _ in range(args.steps):
            t0 = time.perf_counter()
            step()
            torch.cuda.synchronize()
            lat.append((time.perf_counter() - t0) * 1e3)
        torch.cuda.synchronize()
        total = time.perf_counter() - t_start
        # stage times (HIP events on the stream) from separate profiled steps,
        # outside the timed region
        profs = []
        for _ in range(max(3, min(args.steps, 10))):
            step(profile=True)
            profs.append(last_profile())
        torch.cuda.synchronize()
        qps = nq * args.steps / total
        ids = ids_d.cpu().numpy().view(np.uint64)
        rec = float(np.mean([len(set(ids[i, :10].tolist()) & set(gt[i].tolist())) / 10.0
                             for i in range(ns)]))
        log(f"[bench] {precision}: {args.steps} steps in {total:.3f}s -> {qps:,.0f} QPS, "
            f"recall@10 {rec:.4f} on {ns} queries")
        stage = {s: float(np.mean([p["stage_ms"][s] for p in profs]))
                 for s in profs[0]["stage_ms"]}
        lat_sorted = sorted(lat)
        if precision == "mixed":
            log(f"[bench] mixed: rerank gathered {profs[-1]['rerank_rows']:,} rows of "
                f"{profs[-1]['pool_members']:,} pool members; second scan phase listed "
                f"{profs[-1]['admitted']:,} of {profs[-1]['candidates']:,} candidates")
        return {"qps": qps, "total": total, "recall": rec, "stage": stage, "prof": profs[-1],
                "tune": tune,
                "p50": lat_sorted[len(lat) // 2],
                "p99": lat_sorted[min(len(lat) - 1, int(round(0.99 * (len(lat) - 1))))]}

    head = run(args.precision)
    other = None
    if not args.no_compare:
        other = run("exact" if args.precision == "mixed" else "mixed")
    if head["recall"] < RECALL_FLOOR:
        log(f"[bench] WARNING: recall@10 {head['recall']:.4f} below {RECALL_FLOOR}")

    sweep = []
    for pt in [v for v in args.sweep.split(",") if v]:
        npv, _, pool = pt.partition(":")
        sweep.append(sweep_point(args, dev, queries, q_d, ids_d, dd_d, cnt_d, stream, gt, ns, k,
                                 int(npv), int(pool) if pool else args.rerank, np))
    roofline = roofline_of(args, head, dim, nq, k, ix)
    cpu = None if args.no_cpu else cpu_baseline(args, ix, queries, k, n_probe, np)

    qps, total = head["qps"], head["total"]
    out = {
        "metric": METRIC, "value": round(qps, 1), "unit": "queries/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(total * 1e3 / args.steps, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": dtype_of(args, ix),
        "data": "synthetic (seeded numpy mixture; index by seeded k-means or, c3, one "
                "posting per mixture component)",
        "config": {"workload": f"{args.workload}: {ix.n_raw}x{dim} {METRIC_NAMES[ix.metric]} "
                               f"{QUANT_NAMES[ix.quant]}, {ix.n_post} postings, "
                               f"n_probe {n_probe}, k {k}, {nq} queries per step",
                   "precision": args.precision,
                   "probe": f"routing graph ({args.graph} neighbours)" if args.graph else "exact",
                   "rerank_pool": args.rerank if args.precision == "mixed" else
                   (ix.rerank_size if ix.quant else 0),
                   "launch_shapes": head["tune"] or "defaults"},
        "recall_at_10": round(head["recall"], 4), "recall_floor_met": head["recall"] >= RECALL_FLOOR,
        "p50_ms": round(head["p50"], 3), "p99_ms": round(head["p99"], 3),
        "stage_ms": {s: round(v, 4) for s, v in head["stage"].items()},
        "roofline": roofline, "cpu_baseline": cpu,
    }
    if sweep:
        out["sweep"] = sweep
    if other is not None:
        oname = "exact" if args.precision == "mixed" else "mixed"
        out[f"{oname}_gpu"] = {"value": round(other["qps"], 1), "unit": "queries/s",
                               "recall_at_10": round(other["recall"], 4),
                               "p50_ms": round(other["p50"], 3), "p99_ms": round(other["p99"], 3),
                               "launch_shapes": other["tune"] or "defaults",
                               "stage_ms": {s: round(v, 4) for s, v in other["stage"].items()}}
        if args.precision == "mixed":
            out["speedup_vs_fp32_gpu"] = round(qps / other["qps"], 3)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
