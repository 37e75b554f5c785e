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
tensors or any DLPack producer); numpy object columns stay on the host and
are carried through gathers but cannot be keys. 64-bit columns keep their
width. Dictionary-encoded string columns are not part of this layer.
"""

from __future__ import annotations

import numpy as np
import torch

from . import lazy as L
from .series import JaxSeries, _dtype_name


def _coerce_column(arr):
    """(column, dtype name): numeric data goes to the GPU, object arrays stay."""
    if isinstance(arr, JaxSeries):
        arr = arr.data
    if isinstance(arr, list):
        arr = np.array(arr)
    if isinstance(arr, np.ndarray):
        if arr.dtype.kind in "OUS":
            return arr.astype(object), "object"
        arr = torch.from_numpy(np.ascontiguousarray(arr))
    elif not isinstance(arr, torch.Tensor):
        arr = torch.from_dlpack(arr)
    if torch.cuda.is_available() and arr.device.type != "cuda":
        arr = arr.to("cuda")
    return arr, _dtype_name(arr)


class JaxFrame:
    """A frame of equal-length columns; lazy=True builds a plan instead."""

    def __init__(self, data: dict | None = None, index=None, lazy: bool = False, plan=None):
        self._lazy = lazy
        self._plan = plan
        self.index = index
        if lazy and plan is not None:
            self.data = {}
            self._dtypes = {}
            self._columns = L.output_columns(plan)
            self._length = None
            return
        processed, dtypes = {}, {}
        for c, a in (data or {}).items():
            processed[c], dtypes[c] = _coerce_column(a)
        self._init_eager(processed, dtypes)
        if lazy:
            self._plan = L.InputPlan(data=processed, column_names=list(processed))

    def _init_eager(self, data: dict, dtypes: dict | None = None):
        self.data = data
        self._dtypes = dtypes or {c: _dtype_name(a) for c, a in data.items()}
        self._columns = list(data)
        lengths = [len(a) for a in data.values()]
        if lengths and any(x != lengths[0] for x in lengths):
            raise ValueError("All arrays must have the same length")
        self._length = lengths[0] if lengths else 0

    @classmethod
    def _wrap(cls, data: dict, index=None) -> "JaxFrame":
        f = cls.__new__(cls)
        f._lazy = False
        f._plan = None
        f.index = index
        f._init_eager(data)
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
        return pd.DataFrame({c: (a if isinstance(a, np.ndarray) else to_numpy(a))
                             for c, a in self.data.items()}, index=self.index)

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
            return JaxFrame._wrap({c: L.take_column(a, pos) for c, a in self.data.items()})
        if isinstance(key, str):
            if key not in self._columns:
                raise KeyError(key)
            if self._lazy:
                return JaxSeries(None, key, lazy=True, parent_frame=self, expr=L.Column(key))
            return JaxSeries._from_processed_data(self.data[key], name=key, index=self.index,
                                                  dtype=self._dtypes[key])
        if isinstance(key, list):
            for c in key:
                if c not in self._columns:
                    raise KeyError(c)
            if self._lazy:
                return self._lazy_child(L.ProjectPlan(child=self._plan,
                                                      expressions={c: L.Column(c) for c in key}))
            return JaxFrame._wrap({c: self.data[c] for c in key}, index=self.index)
        raise TypeError(f"Column selection requires str, list, or JaxSeries, got {type(key)}")

    def __setitem__(self, key: str, value):
        if self._lazy:
            raise TypeError("column assignment needs an eager frame")
        col, dt = _coerce_column(value)
        if len(col) != self._length:
            raise ValueError(f"Length of values ({len(col)}) does not match length of "
                             f"DataFrame ({self._length})")
        self.data[key] = col
        self._dtypes[key] = dt
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
        return JaxFrame._wrap(L.sort_columns(self.data, by, ascending))

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
        return JaxFrame._wrap(L.join_data(self.data, other.data, on, on, how, lcols, rcols))

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
        return JaxFrame._wrap(L.groupby_columns(f.data, self.by, agg_dict))

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
