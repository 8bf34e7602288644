"""The C-ABI library loads and exports what include/spf.h declares (CPU only:
no call here reaches the GPU)."""
from __future__ import annotations

import ctypes as C
import os
import re
import subprocess

import pytest

from pydiskann_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "spf.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(spf_[a-z_]+)\s*\(", text)))


def test_library_is_built_and_loads():
    assert os.path.exists(_lib.LIB_PATH), "run __graft_entry__.build() first"
    L = _lib.lib()
    assert b"gfx950" in L.spf_version()


def test_every_declared_symbol_is_exported():
    names = declared_functions()
    assert set(names) == set(_lib.EXPORTS)
    L = C.CDLL(_lib.LIB_PATH)
    for n in names:
        assert hasattr(L, n), n
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    for n in names:
        assert re.search(rf"\bT {n}$", out, flags=re.M), n


def test_struct_layouts_match_the_header(tmp_path):
    src = tmp_path / "sz.c"
    src.write_text('#include <stddef.h>\n#include <stdio.h>\n#include "spf.h"\n'
                   "int main(void){printf(\"%zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu\\n\","
                   "sizeof(spf_index_desc), offsetof(spf_index_desc, probe_ratio),"
                   "sizeof(spf_search_opts), sizeof(spf_profile), sizeof(spf_index_info),"
                   "offsetof(spf_index_desc, rabitq_signs),"
                   "offsetof(spf_index_desc, labels), sizeof(spf_server_opts),"
                   "sizeof(spf_request_info), sizeof(spf_server_stats),"
                   "sizeof(spf_graph_desc), offsetof(spf_index_desc, graph),"
                   "sizeof(spf_tune_result));return 0;}\n")
    exe = tmp_path / "sz"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True,
                                          check=True).stdout.split()]
    want = [C.sizeof(_lib.IndexDesc), _lib.IndexDesc.probe_ratio.offset,
            C.sizeof(_lib.SearchOpts), C.sizeof(_lib.Profile), C.sizeof(_lib.IndexInfo),
            _lib.IndexDesc.rabitq_signs.offset, _lib.IndexDesc.labels.offset,
            C.sizeof(_lib.ServerOpts), C.sizeof(_lib.RequestInfo), C.sizeof(_lib.ServerStats),
            C.sizeof(_lib.GraphDesc), _lib.IndexDesc.graph.offset, C.sizeof(_lib.TuneResult)]
    assert got == want


def test_argument_errors_without_a_device():
    L = _lib.lib()
    h = C.c_void_p()
    assert L.spf_index_create(None, C.byref(h)) == _lib.SPF_EINVAL
    assert L.spf_last_error() == b"null argument"
    d = _lib.IndexDesc()
    d.quantizer = 4
    assert L.spf_index_create(C.byref(d), C.byref(h)) == _lib.SPF_EUNSUPPORTED
    # RaBitQ: padded width must be max(dim, 64) rounded up to a power of two
    d.quantizer, d.metric, d.dim, d.rabitq_padded = 3, 1, 100, 100
    assert L.spf_index_create(C.byref(d), C.byref(h)) == _lib.SPF_EINVAL
    assert L.spf_last_error().startswith(b"rabitq_padded")
    d.rabitq_padded = 128
    assert L.spf_index_create(C.byref(d), C.byref(h)) == _lib.SPF_EINVAL
    assert L.spf_last_error() == b"RaBitQ index needs its rotation signs"
    assert L.spf_search_batch(None, None, 1, 4, 10, 8, None, None, None, None) == _lib.SPF_EINVAL
    assert L.spf_delete(None, None, 0) == _lib.SPF_EINVAL
    assert L.spf_nearest_postings(None, None, 1, 4, None, None) == _lib.SPF_EINVAL
    assert L.spf_index_get_info(None, None) == _lib.SPF_EINVAL
    assert L.spf_search_batch_filtered(None, None, 1, 4, 10, 8, None, None, None, None,
                                       None) == _lib.SPF_EINVAL
    assert L.spf_server_create(None, None, C.byref(h)) == _lib.SPF_EINVAL
    t = C.c_uint64()
    assert L.spf_server_submit(None, None, 1, 4, 10, 8, None, None, None, None, None,
                               C.byref(t)) == _lib.SPF_EINVAL
    assert L.spf_server_wait(None, 1, 0, None) == _lib.SPF_EINVAL
    assert L.spf_server_get_stats(None, None) == _lib.SPF_EINVAL
    L.spf_server_destroy(None)
    r = _lib.TuneResult()
    assert L.spf_autotune(None, None, 1, 4, 10, 8, None, None, C.byref(r)) == _lib.SPF_EINVAL


def test_python_mirror_argument_errors():
    import numpy as np
    from pydiskann_amd import SPFresh
    vs = np.zeros((10, 4), np.float32)
    with pytest.raises(ValueError):
        SPFresh.build(vs, "unused", quantizer="nope")
    with pytest.raises(ValueError):
        SPFresh.build(vs, "unused", metric="nope")
