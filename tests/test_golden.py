"""Committed golden fixtures (tests/golden/*.npz, written by make_golden.py).

CPU: the oracle reproduces every stored result bit for bit, and the stored
results still satisfy the FP64 cross-checks they were written under.  GPU: the
HIP path returns the stored results (ids in order, bit-identical distances).
"""
from __future__ import annotations

import glob
import importlib.util
import os

import numpy as np
import pytest

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = sorted(glob.glob(os.path.join(HERE, "golden", "*.npz")))


def _mg():
    spec = importlib.util.spec_from_file_location("make_golden",
                                                  os.path.join(HERE, "golden", "make_golden.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


MG = _mg()


def test_fixtures_present():
    names = {os.path.basename(f)[:-4] for f in FIXTURES}
    assert names == set(MG.CASES), names


def assert_rows(ids, dd, cnt, z, tag):
    np.testing.assert_array_equal(cnt, z[f"{tag}_counts"])
    for i in range(cnt.shape[0]):
        c = int(cnt[i])
        np.testing.assert_array_equal(ids[i, :c], z[f"{tag}_ids"][i, :c], err_msg=f"{tag} q{i}")
        assert np.array_equal(dd[i, :c].view(np.uint32), z[f"{tag}_dists"][i, :c].view(np.uint32)), \
            f"{tag} q{i}: distances differ"


@pytest.mark.parametrize("path", FIXTURES, ids=lambda p: os.path.basename(p)[:-4])
def test_oracle_reproduces_fixture(path):
    ix, z = MG.load(path)
    o = oracle.Oracle(ix)
    q, k = z["queries"], int(z["k"])
    for tag, npr in (("shallow", int(z["n_probe"])), ("full", ix.n_post)):
        ids, dd, cnt = o.search_batch(q, k, npr, threads=4)
        assert_rows(ids, dd, cnt, z, tag)


@pytest.mark.parametrize("path", FIXTURES, ids=lambda p: os.path.basename(p)[:-4])
def test_fixture_fp64_crosschecks(path):
    ix, z = MG.load(path)
    q, k = z["queries"], int(z["k"])
    name = os.path.basename(path)[:-4]
    _, _, _, metric, quant, rerank, _, _, _ = MG.CASES[name]
    if quant is None:
        MG.check_against_fp64(ix, q, k, z["full_ids"], z["full_dists"], z["full_counts"])
    elif rerank or (quant == "rabitq" and metric in ("l2", "l2_root")):
        MG.check_exact_distances(ix, q, z["full_ids"], z["full_dists"], z["full_counts"])
    MG.check_recall(ix, q, k, z["full_ids"], 0.9)


@pytest.mark.gpu
@pytest.mark.parametrize("path", FIXTURES, ids=lambda p: os.path.basename(p)[:-4])
def test_gpu_matches_fixture(path):
    from pydiskann_amd import DeviceIndex
    ix, z = MG.load(path)
    print(f"[case] golden {os.path.basename(path)}", flush=True)
    dev = DeviceIndex(ix)
    q, k = z["queries"], int(z["k"])
    for tag, npr in (("shallow", int(z["n_probe"])), ("full", ix.n_post)):
        ids, dd, cnt = dev.search(q, k, npr)
        assert_rows(ids, dd, cnt, z, tag)


@pytest.mark.parametrize("fault,name", [(1, "c1_small_f32_l2"), (2, "f32_l2root_deletes"),
                                        (1, "rabitq_l2"), (1, "int8_l2_rerank")])
def test_planted_oracle_mistake_breaks_fixture(fault, name):
    ix, z = MG.load(os.path.join(HERE, "golden", f"{name}.npz"))
    oracle.set_fault(fault)
    ids, dd, cnt = oracle.Oracle(ix).search_batch(z["queries"], int(z["k"]), ix.n_post, threads=4)
    with pytest.raises(AssertionError):
        assert_rows(ids, dd, cnt, z, "full")
