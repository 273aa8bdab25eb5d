"""Committed fixtures (tests/golden, written by make_golden.py): the oracle
reproduces them on the CPU and the library reproduces them on the GPU."""
import json
from pathlib import Path

import numpy as np
import pytest

import oracle
from realtimeraytracing_amd import RTQ_TRACE_LEAF_REFMT, devmem, scenes
from conftest import first_diff, hits_equal

GOLD = Path(__file__).resolve().parent / "golden"
SIZES = [1, 2, 3, 8, 64, 300]


def load(n):
    with np.load(GOLD / f"soup_{n}.npz", allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def test_reference_scan_kat_oracle():
    kat = json.loads((GOLD / "reference_scan_kat.json").read_text())
    for g, e in zip(kat["groups"], kat["exclusive_scan"]):
        assert oracle.scan_exclusive(np.array(g, np.uint32)).tolist() == e


@pytest.mark.parametrize("n", SIZES)
def test_oracle_reproduces_fixtures(n):
    g = load(n)
    models = scenes.identity_models()
    assert np.array_equal(g["tris"].view(np.uint8), scenes.soup(n, seed=n).view(np.uint8))
    assert np.array_equal(oracle.morton_codes(g["tris"], models), g["codes"])
    c = oracle.ploc_build(g["tris"], models)
    assert np.array_equal(c["sorted_codes"], g["sorted_codes"])
    assert np.array_equal(c["sorted_indices"], g["sorted_indices"])
    assert hits_equal(oracle.flatten_preorder(n, c), g["nodes"])
    v9 = oracle.world_vertices(g["tris"], models)
    assert hits_equal(oracle.trace_brute(v9, g["rays"]), g["hits_wt"])
    assert hits_equal(oracle.trace_brute(v9, g["rays"], leaf=oracle.LEAF_REFMT), g["hits_refmt"])
    got, _, _ = oracle.trace_bvh(g["nodes"], v9, g["rays"])
    assert hits_equal(got, g["hits_wt"])
    assert (g["hits_wt"]["id"] >= 0).sum() >= 40


@pytest.mark.gpu
@pytest.mark.parametrize("n", SIZES)
def test_library_reproduces_fixtures(rtq_ctx, n):
    g = load(n)
    sc = rtq_ctx.upload(g["tris"], scenes.identity_models())
    try:
        assert np.array_equal(sc.morton_codes(), g["codes"])
        sc.build()
        cl = sc.export_clusters()
        assert np.array_equal(cl["sorted_codes"], g["sorted_codes"])
        assert np.array_equal(cl["sorted_indices"], g["sorted_indices"])
        nodes = sc.export_ref()
        assert hits_equal(nodes, g["nodes"]), first_diff(nodes, g["nodes"])
        got = sc.trace_host(g["rays"])
        assert hits_equal(got, g["hits_wt"]), first_diff(got, g["hits_wt"])
        got = sc.trace_host(g["rays"], RTQ_TRACE_LEAF_REFMT)
        assert hits_equal(got, g["hits_refmt"]), first_diff(got, g["hits_refmt"])
    finally:
        sc.release()


@pytest.mark.gpu
def test_library_scan_kat(rtq_ctx):
    kat = json.loads((GOLD / "reference_scan_kat.json").read_text())
    for grp, e in zip(kat["groups"], kat["exclusive_scan"]):
        din, dout = devmem.to_device(np.array(grp, np.uint32)), devmem.empty(4, np.uint32)
        rtq_ctx.scan_exclusive(din.data_ptr(), dout.data_ptr(), 4)
        rtq_ctx.synchronize()
        assert devmem.to_host(dout, np.uint32).tolist() == e
