"""The reference's operator surface (src/jaxframes/distributed/
parallel_algorithms.py, written pa:LINE) on the MI355X kernels.

Same names, arguments and result conventions, so code written against the
reference switches by changing one import. Arrays are torch tensors on the
GPU (any DLPack producer, numpy array or list is accepted as input).
`sharding_spec` is accepted and ignored: this layer runs one GPU's share, the
reference's mesh-size-1 behaviour. Differences, all value-equal: counts are
int64 (the reference returns the value dtype on the single-key path and int32
on the multi-key path), and float sums are TwoSum-compensated.
"""

from __future__ import annotations

import torch

from . import compaction, grouping, joining, multikey, partition, sorting

__all__ = [
    "multi_column_lexsort", "hash_multi_columns", "combine_keys_for_sorting",
    "collective_hash_repartition", "compact_repartitioned_rows", "ParallelRadixSort",
    "SortBasedGroupBy", "ParallelSortMergeJoin", "parallel_sort", "groupby_aggregate",
    "groupby_aggregate_multi_column", "sort_merge_join",
]

_AGGS = ("sum", "mean", "count", "max", "min")


def multi_column_lexsort(keys: list, ascending=True) -> torch.Tensor:
    """Row order of a stable lexicographic sort, last key least significant (pa:25-66)."""
    return multikey.lexsort_perm(keys, ascending)


def hash_multi_columns(keys: list) -> torch.Tensor:
    """uint32 row hash (pa:69-136)."""
    return partition.hash_multi_columns(keys)


def combine_keys_for_sorting(keys: list) -> torch.Tensor:
    """int32 rank of every row in lexicographic order (pa:139-147)."""
    idx = multikey.lexsort_perm(keys)
    ranks = torch.empty(idx.shape[0], dtype=torch.int32, device=idx.device)
    ranks[idx] = torch.arange(idx.shape[0], dtype=torch.int32, device=idx.device)
    return ranks


def collective_hash_repartition(keys: list, values: dict, key_dtypes=None, value_dtypes=None,
                                sharding_spec=None):
    """(keys, values, valid mask) (pa:185-274); identity on one device (pa:202-203)."""
    return partition.collective_hash_repartition(keys, values, key_dtypes, value_dtypes,
                                                 sharding_spec)


def compact_repartitioned_rows(keys: list, values: dict, valid_mask):
    """Keep the valid rows of every column, row order kept (pa:277-293)."""
    return compaction.compact_repartitioned_rows(keys, values, valid_mask)


class ParallelRadixSort:
    """pa:296-603."""

    def __init__(self, sharding_spec=None, bits_per_pass: int = 8):
        self.sharding_spec = sharding_spec
        self.bits_per_pass = bits_per_pass
        # the reference's digit width is 8; 0 lets the sort pick 8 or 11 per call
        self._digit_bits = bits_per_pass if bits_per_pass in (8, 11) else 0

    def sort(self, keys, values=None, ascending: bool = True):
        """(sorted keys, values permuted alike or None) (pa:463-581). Stable;
        descending is the reversed ascending order (pa:552-555)."""
        k, v = sorting.radix_sort(keys, values, ascending, digit_bits=self._digit_bits)
        return k, v

    def sort_multi_column(self, keys: list, values: dict | None = None, ascending=True):
        """(sorted key columns, sorted values or None) (pa:426-461)."""
        return multikey.sort_multi_column(keys, values, ascending)


class SortBasedGroupBy:
    """pa:606-803."""

    def __init__(self, sharding_spec=None):
        self.sharding_spec = sharding_spec

    def groupby_aggregate(self, keys, values: dict, agg_funcs: dict):
        """(unique keys ascending, {col: aggregate}) (pa:733-803)."""
        for fn in agg_funcs.values():
            if fn not in _AGGS:
                raise ValueError(f"Unsupported aggregation function: {fn}")
        return grouping.groupby_aggregate(keys, values, agg_funcs)

    def groupby_aggregate_multi_column(self, keys: list, key_names: list, values: dict,
                                       agg_funcs: dict):
        """({key_name: unique}, {col: aggregate}) (pa:662-731)."""
        for fn in agg_funcs.values():
            if fn not in _AGGS:
                raise ValueError(f"Unsupported aggregation: {fn}")
        return grouping.groupby_aggregate_multi_column(keys, key_names, values, agg_funcs)


class ParallelSortMergeJoin:
    """pa:806-1075."""

    def __init__(self, sharding_spec=None):
        self.sharding_spec = sharding_spec
        self.sorter = ParallelRadixSort(sharding_spec)

    def join(self, left_keys, left_values: dict, right_keys, right_values: dict, how: str = "inner"):
        """(keys, {left_<c>, right_<c>}) (pa:938-1075)."""
        return joining.sort_merge_join(left_keys, left_values, right_keys, right_values, how)

    def join_multi_column(self, left_keys: list, left_key_names: list, left_values: dict,
                          right_keys: list, right_key_names: list, right_values: dict,
                          how: str = "inner", left_key_dtypes=None, right_key_dtypes=None,
                          left_value_dtypes=None, right_value_dtypes=None):
        """({key: coalesced}, {left_<c>, right_<c>}) (pa:826-936)."""
        return joining.join_multi_column(left_keys, left_key_names, left_values, right_keys,
                                         right_key_names, right_values, how, left_key_dtypes,
                                         right_key_dtypes, left_value_dtypes, right_value_dtypes)


def parallel_sort(arr, sharding_spec=None, values=None, ascending: bool = True):
    """Sorted array, or (sorted array, sorted values) (pa:1079-1114)."""
    k, v = ParallelRadixSort(sharding_spec).sort(arr, values, ascending)
    return k if values is None else (k, v)


def groupby_aggregate(keys, values: dict, agg_funcs: dict, sharding_spec=None):
    """pa:1117-1148."""
    return SortBasedGroupBy(sharding_spec).groupby_aggregate(keys, values, agg_funcs)


def groupby_aggregate_multi_column(keys: list, key_names: list, values: dict, agg_funcs: dict,
                                   sharding_spec=None):
    """pa:1151-1169."""
    return SortBasedGroupBy(sharding_spec).groupby_aggregate_multi_column(keys, key_names, values,
                                                                          agg_funcs)


def sort_merge_join(left_keys, left_values: dict, right_keys, right_values: dict,
                    how: str = "inner", sharding_spec=None):
    """pa:1172-1209."""
    return ParallelSortMergeJoin(sharding_spec).join(left_keys, left_values, right_keys,
                                                     right_values, how)
