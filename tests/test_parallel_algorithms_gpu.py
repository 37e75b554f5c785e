"""The reference's operator surface (jaxframes_amd.parallel_algorithms) run on
the reference's own test cases (tests/golden/reference_vectors.json, each case
naming its source), with the reference's call signatures."""

import json
import os

import numpy as np
import pytest
import torch

import oracle
from jaxframes_amd import parallel_algorithms as pa
from tests._util import cuda

pytestmark = pytest.mark.gpu


def _cases(section):
    path = os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.json")
    with open(path) as f:
        return json.load(f)[section]


def _np_of(case, key):
    a = np.asarray(case[key])
    if "dtype" in case and key == "keys":
        a = a.astype(case["dtype"])
    return a


def test_parallel_sort_reference_cases():
    for case in _cases("sort"):
        keys = _np_of(case, "keys")
        asc = case.get("ascending", True)
        if "values" in case:
            k, v = pa.parallel_sort(cuda(keys), values=cuda(np.asarray(case["values"])), ascending=asc)
            if "expected_values" in case:
                assert v.cpu().tolist() == case["expected_values"], case["source"]
        else:
            k = pa.parallel_sort(cuda(keys), ascending=asc)
        np.testing.assert_allclose(k.cpu().numpy(), np.asarray(case["expected_keys"], dtype=keys.dtype),
                                   err_msg=case["source"])


def test_radix_sorter_class_and_multi_column():
    sorter = pa.ParallelRadixSort(None, bits_per_pass=8)
    k, v = sorter.sort(cuda(np.array([3, 1, 2], dtype=np.int32)))
    assert k.cpu().tolist() == [1, 2, 3] and v is None
    for case in _cases("lexsort"):
        keys = [cuda(np.asarray(k, dtype=np.int64)) for k in case["keys"]]
        sk, _ = sorter.sort_multi_column(keys, None, case["ascending"])
        assert [s.cpu().tolist() for s in sk] == case["expected"], case["source"]


def test_groupby_aggregate_reference_cases():
    for case in _cases("groupby"):
        keys = cuda(np.asarray(case["keys"]))
        vals = {c: cuda(np.asarray(v)) for c, v in case["values"].items()}
        uk, res = pa.groupby_aggregate(keys, vals, case["aggs"])
        assert uk.cpu().tolist() == case["expected_keys"], case["source"]
        for c, exp in case["expected"].items():
            np.testing.assert_allclose(res[c].cpu().numpy().astype(np.float64), exp, rtol=1e-6,
                                       err_msg=case["source"])


def test_groupby_unknown_aggregation_raises():
    with pytest.raises(ValueError):
        pa.groupby_aggregate(cuda(np.array([1, 2])), {"v": cuda(np.array([1., 2.]))}, {"v": "median"})


def test_groupby_multi_column_surface():
    k1 = np.array([1, 1, 2, 2, 1], dtype=np.int32)
    k2 = np.array([5, 6, 5, 5, 5], dtype=np.int64)
    v = np.array([1., 2., 3., 4., 5.], dtype=np.float32)
    uk, res = pa.groupby_aggregate_multi_column([cuda(k1), cuda(k2)], ["a", "b"], {"v": cuda(v)},
                                                {"v": "sum"})
    ek, er = oracle.groupby_aggregate_multi_column([k1, k2], ["a", "b"], {"v": v}, {"v": "sum"})
    assert uk["a"].cpu().tolist() == ek["a"].tolist() and uk["b"].cpu().tolist() == ek["b"].tolist()
    np.testing.assert_allclose(res["v"].cpu().numpy(), er["v"], rtol=1e-6)


def test_sort_merge_join_reference_cases():
    for case in _cases("join"):
        k, out = pa.sort_merge_join(
            cuda(np.asarray(case["left_keys"])), {c: cuda(np.asarray(v)) for c, v in case["left_values"].items()},
            cuda(np.asarray(case["right_keys"])), {c: cuda(np.asarray(v)) for c, v in case["right_values"].items()},
            how=case["how"])
        assert k.cpu().tolist() == case["expected_keys"], case["source"]
        for c, exp in case["expected"].items():
            np.testing.assert_array_equal(out[c].cpu().numpy(), np.asarray(exp, dtype=np.float64),
                                          case["source"])


def test_join_multi_column_surface():
    joiner = pa.ParallelSortMergeJoin(None)
    for case in _cases("join_multi"):
        names = case["left_key_names"]
        keys, vals = joiner.join_multi_column(
            [cuda(np.asarray(k)) for k in case["left_keys"]], names,
            {c: cuda(np.asarray(v)) for c, v in case["left_values"].items()},
            [cuda(np.asarray(k)) for k in case["right_keys"]], names,
            {c: cuda(np.asarray(v)) for c, v in case["right_values"].items()}, how=case["how"])
        assert keys[names[0]].shape[0] == case["expected_rows"], case["source"]
        if "expected_right_value_isnan" in case:
            assert torch.isnan(vals["right_value"]).cpu().tolist() == case["expected_right_value_isnan"]


def test_hash_and_rank_surfaces():
    k = np.array([5, -3, 5, 2**40], dtype=np.int64)
    h = pa.hash_multi_columns([cuda(k)])
    assert h.view(torch.int32).cpu().numpy().view(np.uint32).tolist() == \
        oracle.hash_multi_columns([k]).tolist()
    r = pa.combine_keys_for_sorting([cuda(np.array([3, 1, 2], dtype=np.int32)),
                                     cuda(np.array([0, 0, 0], dtype=np.int32))])
    assert r.cpu().tolist() == [2, 0, 1]


def test_repartition_surfaces():
    k = cuda(np.arange(6, dtype=np.int64))
    ks, vs, mask = pa.collective_hash_repartition([k], {"v": k.float()})
    assert bool(mask.all())
    m = cuda(np.array([1, 0, 1, 1, 0, 0], dtype=np.uint8))
    ck, cv = pa.compact_repartitioned_rows(ks, vs, m)
    assert ck[0].cpu().tolist() == [0, 2, 3] and cv["v"].cpu().tolist() == [0., 2., 3.]
