"""JaxFrame: the reference's DataFrame API (core/frame.py) on this package's
kernels, eager and lazy.

The entry points of the algorithms layer behave as the reference's:

- sort_values(by, ascending) (core/frame.py:1120-1223);
- groupby(by).agg(...), .sum/.mean/.max/.min/.count (:1225-1352);
- merge(other, on, how) (:1354-1541): result {key, left_<c>, right_<c>};
- boolean filters frame[series] (:317-339), column / list selection;
- head / tail (:1543-1597);
- lazy=True builds plans (InputPlan, ProjectPlan, FilterPlan, SortPlan,
  GroupByPlan, JoinPlan, LimitPlan) that collect() executes (:567-588), and
  explain() prints (:590-615).

Columns live on the GPU as torch tensors (inputs: numpy arrays, lists, torch
tensors or any DLPack producer). String arrays (str / bytes, or object arrays
of strings and None) are dictionary-encoded as the reference does
(core/string_encoding.py): int32 codes on the GPU into a sorted vocabulary,
with the reference's string paths for comparisons, sort_values, groupby and
merge (core/string_fastpaths.py). Other numpy object columns stay on the host
and are carried through gathers but cannot be keys. 64-bit columns keep their
width.
"""

from __future__ import annotations

import numpy as np
import torch

from . import lazy as L
from . import strings as S
from .series import JaxSeries, _dtype_name


def _coerce_column(arr):
    """(column, dtype name, vocab or None) (core/frame.py:165-181): strings
    are encoded to codes, other object arrays stay on the host, numeric data
    goes to the GPU."""
    if isinstance(arr, JaxSeries):
        if arr._dtype == "string":
            return arr.data, "string", arr._string_vocab
        arr = arr.data
    if isinstance(arr, list):
        arr = np.array(arr)
    if isinstance(arr, np.ndarray):
        if S.is_string_array(arr):
            codes, vocab = S.encode_string_array(arr)
            return codes, "string", vocab
        if arr.dtype.kind in "OUS":
            return arr.astype(object), "object", None
        arr = torch.from_numpy(np.ascontiguousarray(arr))
    elif not isinstance(arr, torch.Tensor):
        arr = torch.from_dlpack(arr)
    if torch.cuda.is_available() and arr.device.type != "cuda":
        arr = arr.to("cuda")
    return arr, _dtype_name(arr), None


class JaxFrame:
    """A frame of equal-length columns; lazy=True builds a plan instead."""

    def __init__(self, data: dict | None = None, index=None, lazy: bool = False, plan=None):
        self._lazy = lazy
        self._plan = plan
        self.index = index
        if lazy and plan is not None:
            self.data = {}
            self._dtypes = {}
            self._string_vocabs = {}
            self._columns = L.output_columns(plan)
            self._length = None
            return
        processed, dtypes, vocabs = {}, {}, {}
        for c, a in (data or {}).items():
            processed[c], dtypes[c], vocab = _coerce_column(a)
            if vocab is not None:
                vocabs[c] = vocab
        self._init_eager(processed, dtypes, vocabs)
        if lazy:
            self._plan = L.InputPlan(data=processed, column_names=list(processed))

    def _init_eager(self, data: dict, dtypes: dict | None = None, vocabs: dict | None = None):
        self.data = data
        self._string_vocabs = dict(vocabs or {})
        self._dtypes = dict(dtypes) if dtypes else {}
        for c, a in data.items():
            if c not in self._dtypes:
                self._dtypes[c] = "string" if c in self._string_vocabs else _dtype_name(a)
        self._columns = list(data)
        lengths = [len(a) for a in data.values()]
        if lengths and any(x != lengths[0] for x in lengths):
            raise ValueError("All arrays must have the same length")
        self._length = lengths[0] if lengths else 0

    @classmethod
    def _wrap(cls, data: dict, index=None, vocabs: dict | None = None) -> "JaxFrame":
        f = cls.__new__(cls)
        f._lazy = False
        f._plan = None
        f.index = index
        f._init_eager(data, None, {c: v for c, v in (vocabs or {}).items() if c in data})
        return f

    # ---- properties -----------------------------------------------------------------

    @property
    def columns(self) -> list:
        return list(self._columns)

    @property
    def shape(self):
        return (self._length, len(self._columns))

    @property
    def is_lazy(self) -> bool:
        return self._lazy

    @property
    def plan(self):
        return self._plan

    def __len__(self) -> int:
        if self._length is None:
            raise TypeError("a lazy frame has no length before collect()")
        return self._length

    def __repr__(self) -> str:
        if self._lazy:
            return f"JaxFrame(lazy=True, columns={self.columns})"
        return f"JaxFrame(shape={self.shape}, columns={self.columns})"

    def to_pandas(self):
        import pandas as pd
        from ._tensors import to_numpy
        cols = {}
        for c, a in self.data.items():
            if c in self._string_vocabs:
                cols[c] = S.decode_string_codes(a, self._string_vocabs[c])
            else:
                cols[c] = a if isinstance(a, np.ndarray) else to_numpy(a)
        return pd.DataFrame(cols, index=self.index)

    # ---- lazy plans -----------------------------------------------------------------

    def collect(self) -> "JaxFrame":
        if not self._lazy:
            return self
        if self._plan is None:
            raise ValueError("Lazy frame has no plan to execute")
        return JaxFrame._wrap(L.execute(self._plan), index=self.index)

    def explain(self, verbose: bool = False) -> str:
        if not self._lazy:
            return "Eager execution (no plan)"
        if self._plan is None:
            return "No plan available"
        s = "Query Plan:\n" + repr(self._plan)
        if verbose:
            s += ("\n\n[execution: columns pruned to those read; a head/tail directly "
                  "above a one-key sort gathers the kept rows only]")
        return s

    def _lazy_child(self, plan) -> "JaxFrame":
        return JaxFrame(data=None, index=self.index, lazy=True, plan=plan)

    # ---- selection and filters -------------------------------------------------------

    def __getitem__(self, key):
        if isinstance(key, JaxSeries):
            if key._lazy:
                if key._expr is None:
                    raise ValueError("Lazy series for filtering must have an expression")
                return self._lazy_child(L.FilterPlan(child=self._plan, condition=key._expr))
            mask = key.data
            if not isinstance(mask, torch.Tensor):
                raise TypeError("Boolean indexing requires array data")
            from .compaction import compact_rows
            if mask.shape[0] != self._length:
                raise ValueError("boolean mask has the wrong length")
            pos = compact_rows(mask)
            return JaxFrame._wrap({c: L.take_column(a, pos) for c, a in self.data.items()},
                                  vocabs=self._string_vocabs)
        if isinstance(key, str):
            if key not in self._columns:
                raise KeyError(key)
            if self._lazy:
                return JaxSeries(None, key, lazy=True, parent_frame=self, expr=L.Column(key))
            return JaxSeries._from_processed_data(self.data[key], name=key, index=self.index,
                                                  dtype=self._dtypes[key],
                                                  string_vocab=self._string_vocabs.get(key))
        if isinstance(key, list):
            for c in key:
                if c not in self._columns:
                    raise KeyError(c)
            if self._lazy:
                return self._lazy_child(L.ProjectPlan(child=self._plan,
                                                      expressions={c: L.Column(c) for c in key}))
            return JaxFrame._wrap({c: self.data[c] for c in key}, index=self.index,
                                  vocabs=self._string_vocabs)
        raise TypeError(f"Column selection requires str, list, or JaxSeries, got {type(key)}")

    def __setitem__(self, key: str, value):
        if self._lazy:
            raise TypeError("column assignment needs an eager frame")
        col, dt, vocab = _coerce_column(value)
        if len(col) != self._length:
            raise ValueError(f"Length of values ({len(col)}) does not match length of "
                             f"DataFrame ({self._length})")
        self.data[key] = col
        self._dtypes[key] = dt
        if vocab is not None:
            self._string_vocabs[key] = vocab
        else:
            self._string_vocabs.pop(key, None)
        if key not in self._columns:
            self._columns.append(key)

    # ---- the algorithms layer ---------------------------------------------------------

    def sort_values(self, by, ascending=True) -> "JaxFrame":
        by = [by] if isinstance(by, str) else list(by)
        for c in by:
            if c not in self._columns:
                raise KeyError(f"Column '{c}' not found")
        if self._lazy:
            return self._lazy_child(L.SortPlan(child=self._plan, sort_columns=by,
                                               ascending=ascending))
        for c in by:
            if self._dtypes[c] == "object":
                raise TypeError(f"Cannot sort by object dtype column '{c}'")
        if len(by) == 1 and self._dtypes[by[0]] == "string":
            # lexical order with nulls last, stable (core/frame.py:1160-1167)
            from . import parallel_algorithms as pa
            asc = ascending if isinstance(ascending, bool) else ascending[0]
            keys = S.sortable_string_codes(self.data[by[0]], self._string_vocabs[by[0]],
                                           descending=not asc)
            rows = torch.arange(len(keys), dtype=torch.int64, device=keys.device)
            _, idx = pa.ParallelRadixSort().sort(keys, values=rows, ascending=True)
            return JaxFrame._wrap({c: L.take_column(a, idx) for c, a in self.data.items()},
                                  vocabs=self._string_vocabs)
        if len(by) > 1 and any(self._dtypes[c] == "string" for c in by):
            raise NotImplementedError("Multi-column sorting with string columns is not yet "
                                      "implemented")
        return JaxFrame._wrap(L.sort_columns(self.data, by, ascending), vocabs=self._string_vocabs)

    def groupby(self, by) -> "GroupBy":
        return GroupBy(self, by)

    def merge(self, other: "JaxFrame", on, how: str = "inner") -> "JaxFrame":
        on = [on] if isinstance(on, str) else list(on)
        for c in on:
            if c not in self._columns:
                raise KeyError(f"Column '{c}' not found in left DataFrame")
            if c not in other._columns:
                raise KeyError(f"Column '{c}' not found in right DataFrame")
        if self._lazy:
            right = other._plan if other._lazy else L.InputPlan(data=other.data,
                                                                 column_names=other.columns)
            return self._lazy_child(L.JoinPlan(left=self._plan, right=right, left_keys=on,
                                               right_keys=on, join_type=how))
        if other._lazy:
            other = other.collect()
        for c in on:
            if self._dtypes[c] == "object" or other._dtypes[c] == "object":
                raise TypeError(f"Cannot join on object dtype column '{c}'")
        lcols = [c for c in self._columns if c not in on]
        rcols = [c for c in other._columns if c not in on]
        ldt = {c: self._dtypes[c] for c in lcols}
        rdt = {c: other._dtypes[c] for c in rcols}
        if len(on) == 1 and S.can_use_string_join_fastpath(self._dtypes[on[0]], ldt, rdt, how):
            return self._merge_strings(other, on[0], how, lcols, rcols)
        ldata, rdata, key_vocabs = dict(self.data), dict(other.data), {}
        for c in on:   # string keys in a merged vocabulary (core/frame.py:1466-1487)
            if self._dtypes[c] == "string":
                ldata[c], rdata[c], key_vocabs[c] = S.align_string_code_arrays(
                    self.data[c], self._string_vocabs[c], other.data[c], other._string_vocabs[c])
        out = L.join_data(ldata, rdata, on, on, how, lcols, rcols)
        return JaxFrame._wrap(out, vocabs=key_vocabs)

    def _merge_strings(self, other: "JaxFrame", key: str, how: str, lcols: list,
                       rcols: list) -> "JaxFrame":
        """Single string-key merge (core/frame.py:1424-1464)."""
        lk, rk, vocab = S.align_string_code_arrays(self.data[key], self._string_vocabs[key],
                                                   other.data[key], other._string_vocabs[key])
        plan = S.join_encoded_strings(lk, rk, len(vocab), how)
        out = {key: plan.joined_codes}
        vocabs = {key: vocab}
        for c in lcols:
            out[f"left_{c}"] = S.gather_join_payload(self.data[c], plan.left_indices,
                                                     self._dtypes[c])
            if c in self._string_vocabs:
                vocabs[f"left_{c}"] = self._string_vocabs[c]
        for c in rcols:
            out[f"right_{c}"] = S.gather_join_payload(other.data[c], plan.right_indices,
                                                      other._dtypes[c])
            if c in other._string_vocabs:
                vocabs[f"right_{c}"] = other._string_vocabs[c]
        return JaxFrame._wrap(out, vocabs=vocabs)

    def head(self, n: int = 5) -> "JaxFrame":
        if self._lazy:
            return self._lazy_child(L.LimitPlan(child=self._plan, limit=n, from_end=False))
        k = max(0, min(n, self._length))
        return JaxFrame._wrap({c: a[:k] for c, a in self.data.items()})

    def tail(self, n: int = 5) -> "JaxFrame":
        if self._lazy:
            return self._lazy_child(L.LimitPlan(child=self._plan, limit=n, from_end=True))
        k = max(0, min(n, self._length))
        return JaxFrame._wrap({c: (a[-k:] if k else a[:0]) for c, a in self.data.items()})


class GroupBy:
    """frame.groupby(by) (core/frame.py:1242-1352)."""

    def __init__(self, frame: JaxFrame, by):
        self.frame = frame
        self.by = [by] if isinstance(by, str) else list(by)
        for c in self.by:
            if c not in frame._columns:
                raise KeyError(f"Column '{c}' not found")
            if not frame._lazy and frame._dtypes[c] == "object":
                raise TypeError(f"Cannot group by object dtype column '{c}'")

    def agg(self, agg_funcs) -> JaxFrame:
        f = self.frame
        if isinstance(agg_funcs, str):
            if f._lazy:   # the columns the plan produces (its schema, core/frame.py:1259-1263)
                agg_dict = {c: agg_funcs for c in f._columns if c not in self.by}
            else:
                agg_dict = {c: agg_funcs for c in f._columns
                            if c not in self.by and f._dtypes[c] not in ("object", "string")}
        else:
            agg_dict = dict(agg_funcs)
        if f._lazy:
            return f._lazy_child(L.GroupByPlan(child=f._plan, by=self.by, agg_dict=agg_dict))
        key_vocabs = {c: f._string_vocabs[c] for c in self.by if c in f._string_vocabs}
        if len(self.by) == 1 and S.can_use_string_groupby_fastpath(
                f._dtypes[self.by[0]], agg_dict, f._dtypes):
            key = self.by[0]
            uk, agg = S.groupby_encoded_strings(f.data[key], {c: f.data[c] for c in agg_dict},
                                                agg_dict, len(f._string_vocabs[key]))
            return JaxFrame._wrap({key: uk, **agg}, vocabs=key_vocabs)
        return JaxFrame._wrap(L.groupby_columns(f.data, self.by, agg_dict), vocabs=key_vocabs)

    def sum(self):
        return self.agg("sum")

    def mean(self):
        return self.agg("mean")

    def max(self):
        return self.agg("max")

    def min(self):
        return self.agg("min")

    def count(self):
        return self.agg("count")
