"""Lazy plans: expression and plan nodes, and their executor on the GPU.

Node names, fields, schemas and printing follow the reference's logical plan
(lazy/plan.py:17-419) and expressions (lazy/expressions.py, ops/comparison.py),
so code that builds or inspects plans works unchanged. collect() runs
execute(), which follows JaxFrame._execute_plan (core/frame.py:617-927) node
by node on this package's kernels, with two changes that leave every result
the same:

- column pruning: each node asks its child only for the columns it reads
  (a filter under a groupby gathers the key and aggregated columns, not the
  whole row; a projection above a join drops payloads before they are
  gathered);
- a LimitPlan directly above a SortPlan sorts the key column with its row
  index and gathers the other columns for the kept rows only.
"""

from __future__ import annotations

import enum
from dataclasses import dataclass
from typing import Any, Union

import numpy as np
import torch

# ---- expressions ---------------------------------------------------------------


class ComparisonOpType(enum.Enum):
    """ops/comparison.py ComparisonOpType."""
    EQ = "=="
    NE = "!="
    LT = "<"
    LE = "<="
    GT = ">"
    GE = ">="


class Expr:
    """Expression node (lazy/expressions.py:15-35)."""

    def columns(self) -> set:
        return set()


@dataclass(frozen=True, eq=False)
class Column(Expr):
    name: str

    def columns(self) -> set:
        return {self.name}

    def __repr__(self) -> str:
        return f"Column({self.name!r})"


@dataclass(frozen=True, eq=False)
class Literal(Expr):
    value: Any
    dtype: Any = None

    def __repr__(self) -> str:
        return f"Literal({self.value!r})"


@dataclass(frozen=True, eq=False)
class BinaryOp(Expr):
    left: Expr
    op: str
    right: Expr

    def columns(self) -> set:
        return self.left.columns() | self.right.columns()

    def __repr__(self) -> str:
        return f"({self.left!r} {self.op} {self.right!r})"


@dataclass(frozen=True, eq=False)
class UnaryOp(Expr):
    op: str
    operand: Expr

    def columns(self) -> set:
        return self.operand.columns()

    def __repr__(self) -> str:
        return f"{self.op}({self.operand!r})"


@dataclass(frozen=True, eq=False)
class ComparisonOp(Expr):
    op: ComparisonOpType
    left: Expr
    right: Expr

    def columns(self) -> set:
        return self.left.columns() | self.right.columns()

    def __repr__(self) -> str:
        return f"({self.left!r} {self.op.value} {self.right!r})"


# ---- plan nodes -------------------------------------------------------------------


class LogicalPlan:
    """Plan node (lazy/plan.py:17-87)."""

    def schema(self) -> dict:
        raise NotImplementedError

    def children(self) -> list:
        raise NotImplementedError

    def _node_details(self) -> str:
        return ""

    def __repr__(self) -> str:
        return self._repr_tree(0)

    def _repr_tree(self, indent: int = 0) -> str:
        s = "  " * indent + type(self).__name__
        d = self._node_details()
        if d:
            s += f"({d})"
        kids = self.children()
        if kids:
            return s + "\n" + "\n".join(k._repr_tree(indent + 1) for k in kids)
        return s


@dataclass(frozen=True, eq=False, repr=False)
class InputPlan(LogicalPlan):
    data: dict
    column_names: list

    def schema(self) -> dict:
        return {c: None for c in self.column_names}

    def children(self) -> list:
        return []

    def _node_details(self) -> str:
        if self.data and self.column_names:
            n = len(self.data[self.column_names[0]])
            return f"{n} rows, {len(self.column_names)} cols: {self.column_names}"
        return "empty"

    @property
    def source_id(self) -> str:
        return f"input_{id(self)}"


@dataclass(frozen=True, eq=False, repr=False)
class SelectPlan(LogicalPlan):
    child: LogicalPlan
    column_name: str

    def schema(self) -> dict:
        return {self.column_name: self.child.schema().get(self.column_name)}

    def children(self) -> list:
        return [self.child]

    def _node_details(self) -> str:
        return f"column={self.column_name!r}"


@dataclass(frozen=True, eq=False, repr=False)
class ProjectPlan(LogicalPlan):
    child: LogicalPlan
    expressions: dict

    def schema(self) -> dict:
        return {c: None for c in self.expressions}

    def children(self) -> list:
        return [self.child]

    def _node_details(self) -> str:
        return f"columns={list(self.expressions)}"


@dataclass(frozen=True, eq=False, repr=False)
class FilterPlan(LogicalPlan):
    child: LogicalPlan
    condition: Expr

    def schema(self) -> dict:
        return self.child.schema()

    def children(self) -> list:
        return [self.child]

    def _node_details(self) -> str:
        return f"condition={self.condition!r}"


@dataclass(frozen=True, eq=False, repr=False)
class SortPlan(LogicalPlan):
    child: LogicalPlan
    sort_columns: list
    ascending: Union[bool, list] = True

    def schema(self) -> dict:
        return self.child.schema()

    def children(self) -> list:
        return [self.child]

    def _node_details(self) -> str:
        return f"by={self.sort_columns}, ascending={self.ascending}"


@dataclass(frozen=True, eq=False, repr=False)
class GroupByPlan(LogicalPlan):
    child: LogicalPlan
    by: Union[str, list]
    agg_dict: dict

    def schema(self) -> dict:
        cs = self.child.schema()
        out = {c: cs.get(c) for c in _as_list(self.by)}
        out.update({c: None for c in self.agg_dict})
        return out

    def children(self) -> list:
        return [self.child]

    def _node_details(self) -> str:
        return f"by={self.by!r}, agg={self.agg_dict}"


@dataclass(frozen=True, eq=False, repr=False)
class JoinPlan(LogicalPlan):
    left: LogicalPlan
    right: LogicalPlan
    left_keys: list
    right_keys: list
    join_type: str
    suffixes: tuple = ("_left", "_right")

    def schema(self) -> dict:
        """The reference's declared schema (lazy/plan.py:353-374). The columns
        collect() produces are those of the join itself: the keys, then
        left_<c> and right_<c> for every other column."""
        out = dict(self.left.schema())
        for c, t in self.right.schema().items():
            if c in self.right_keys:
                continue
            out[f"{c}{self.suffixes[1]}" if c in out else c] = t
        return out

    def children(self) -> list:
        return [self.left, self.right]

    def _node_details(self) -> str:
        return f"left_on={self.left_keys}, right_on={self.right_keys}, how={self.join_type!r}"


@dataclass(frozen=True, eq=False, repr=False)
class LimitPlan(LogicalPlan):
    child: LogicalPlan
    limit: int
    from_end: bool = False

    def schema(self) -> dict:
        return self.child.schema()

    def children(self) -> list:
        return [self.child]

    def _node_details(self) -> str:
        return f"{'tail' if self.from_end else 'head'}({self.limit})"


# reference aliases (lazy/plan.py:411-419)
Scan = InputPlan
Selection = FilterPlan
Projection = ProjectPlan
Join = JoinPlan
Sort = SortPlan
Limit = LimitPlan


def _as_list(x) -> list:
    return [x] if isinstance(x, str) else list(x)


# ---- execution ------------------------------------------------------------------------

_CMP_TORCH = {
    ComparisonOpType.EQ: torch.eq, ComparisonOpType.NE: torch.ne, ComparisonOpType.LT: torch.lt,
    ComparisonOpType.LE: torch.le, ComparisonOpType.GT: torch.gt, ComparisonOpType.GE: torch.ge,
}
_BIN_TORCH = {
    "==": torch.eq, "!=": torch.ne, "<": torch.lt, "<=": torch.le, ">": torch.gt, ">=": torch.ge,
    "&": torch.logical_and, "|": torch.logical_or, "+": torch.add, "-": torch.sub,
    "*": torch.mul, "/": torch.true_divide,
}


def output_columns(plan: LogicalPlan) -> list:
    """The columns collect() produces for plan, in order."""
    if isinstance(plan, InputPlan):
        return list(plan.column_names)
    if isinstance(plan, SelectPlan):
        return [plan.column_name]
    if isinstance(plan, ProjectPlan):
        return list(plan.expressions)
    if isinstance(plan, SortPlan) and len(plan.sort_columns) > 1:
        # sort_multi_column returns the key columns first (core/frame.py:795-803)
        return list(plan.sort_columns) + [c for c in output_columns(plan.child)
                                          if c not in plan.sort_columns]
    if isinstance(plan, (FilterPlan, SortPlan, LimitPlan)):
        return output_columns(plan.child)
    if isinstance(plan, GroupByPlan):
        return _as_list(plan.by) + [c for c in plan.agg_dict if c not in _as_list(plan.by)]
    if isinstance(plan, JoinPlan):
        lk, rk = plan.left_keys, plan.right_keys
        keys = [lk[0]] if len(lk) == 1 else list(lk)
        return (keys + [f"left_{c}" for c in output_columns(plan.left) if c not in lk]
                + [f"right_{c}" for c in output_columns(plan.right) if c not in rk])
    raise ValueError(f"Unknown plan type: {type(plan)}")


def _literal(value, like: torch.Tensor):
    if isinstance(value, (np.ndarray, list, torch.Tensor)):
        from ._tensors import to_device
        return to_device(value, like.device)
    return value


def evaluate(expr: Expr, data: dict):
    """Value of an expression over the columns in data (the elementwise part
    of a filter: ExpressionCodeGen.generate, lazy/codegen.py:67-225)."""
    if isinstance(expr, Column):
        if expr.name not in data:
            raise KeyError(f"Column '{expr.name}' not found in available columns: {list(data)}")
        return data[expr.name]
    if isinstance(expr, Literal):
        return expr.value
    if isinstance(expr, ComparisonOp):
        a, b = evaluate(expr.left, data), evaluate(expr.right, data)
        if isinstance(a, torch.Tensor):
            return _CMP_TORCH[expr.op](a, _literal(b, a))
        return _CMP_TORCH[expr.op](torch.as_tensor(a), b)
    if isinstance(expr, BinaryOp):
        if expr.op not in _BIN_TORCH:
            raise ValueError(f"Unsupported binary operator: {expr.op}")
        a, b = evaluate(expr.left, data), evaluate(expr.right, data)
        if isinstance(a, torch.Tensor):
            b = _literal(b, a)
        return _BIN_TORCH[expr.op](a, b)
    if isinstance(expr, UnaryOp):
        v = evaluate(expr.operand, data)
        if expr.op in ("~", "not"):
            return torch.logical_not(v)
        if expr.op == "-":
            return torch.neg(v)
        if expr.op == "abs":
            return torch.abs(v)
        raise ValueError(f"Unsupported unary operator: {expr.op}")
    raise ValueError(f"Unsupported expression type: {type(expr).__name__}")


def _simple_comparison(expr: Expr):
    """(column, op, scalar) when expr is `column <op> scalar` (either side)."""
    if not isinstance(expr, ComparisonOp):
        return None
    flip = {ComparisonOpType.LT: ">", ComparisonOpType.LE: ">=", ComparisonOpType.GT: "<",
            ComparisonOpType.GE: "<=", ComparisonOpType.EQ: "==", ComparisonOpType.NE: "!="}
    scal = (int, float, np.number)
    if isinstance(expr.left, Column) and isinstance(expr.right, Literal) \
            and isinstance(expr.right.value, scal) and not isinstance(expr.right.value, bool):
        return expr.left.name, expr.op.value, expr.right.value
    if isinstance(expr.right, Column) and isinstance(expr.left, Literal) \
            and isinstance(expr.left.value, scal) and not isinstance(expr.left.value, bool):
        return expr.right.name, flip[expr.op], expr.left.value
    return None


def take_column(col, rows: torch.Tensor):
    """col[rows] for a device column (GPU gather) or a host object column."""
    from .compaction import take
    if isinstance(col, np.ndarray):
        return col[rows.cpu().numpy()]
    return take(col, rows)


def filter_positions(condition: Expr, data: dict, n: int) -> torch.Tensor:
    """Ascending int64 positions of the rows where condition holds: a
    `column <op> scalar` condition runs the fused predicate + compaction
    kernel; other conditions build the mask elementwise and compact it."""
    from .compaction import compact_rows, filter_rows
    simple = _simple_comparison(condition)
    if simple is not None and isinstance(data.get(simple[0]), torch.Tensor) \
            and not data[simple[0]].dtype.is_complex and data[simple[0]].dtype != torch.bool:
        col, op, value = simple
        c = data[col]
        if c.dtype.is_floating_point or isinstance(value, (int, np.integer)) \
                or float(value).is_integer():
            if not c.dtype.is_floating_point:
                value = int(value)
            return filter_rows(c, op, value)
    mask = evaluate(condition, data)
    if not isinstance(mask, torch.Tensor) or mask.dim() == 0:
        keep = bool(mask)
        dev = next((v.device for v in data.values() if isinstance(v, torch.Tensor)), None)
        return (torch.arange(n, device=dev) if keep
                else torch.empty(0, dtype=torch.int64, device=dev))
    if mask.shape[0] != n:
        raise ValueError("filter condition has the wrong length")
    return compact_rows(mask)


def _need(plan: LogicalPlan, need):
    return output_columns(plan) if need is None else list(need)


def execute(plan: LogicalPlan, need=None) -> dict:
    """{column: array} for plan; need (None = all) restricts the columns the
    caller reads."""
    from . import parallel_algorithms as pa
    if isinstance(plan, InputPlan):
        cols = _need(plan, need)
        return {c: plan.data[c] for c in plan.column_names if c in cols}
    if isinstance(plan, SelectPlan):
        return {plan.column_name: execute(plan.child, [plan.column_name])[plan.column_name]}
    if isinstance(plan, ProjectPlan):
        want = [c for c in plan.expressions if need is None or c in need]
        srcs = set()
        for c in want:
            srcs |= plan.expressions[c].columns()
        data = execute(plan.child, sorted(srcs))
        out = {}
        for c in want:
            e = plan.expressions[c]
            if not isinstance(e, Column):
                raise NotImplementedError(f"Expression evaluation not yet implemented for {type(e)}")
            out[c] = data[e.name]
        return out
    if isinstance(plan, FilterPlan):
        cols = _need(plan, need)
        data = execute(plan.child, sorted(set(cols) | plan.condition.columns()))
        n = _length(data)
        pos = filter_positions(plan.condition, data, n)
        if pos.shape[0] == n:
            return {c: data[c] for c in cols if c in data}
        return {c: take_column(data[c], pos) for c in cols if c in data}
    if isinstance(plan, LimitPlan):
        if isinstance(plan.child, SortPlan) and len(plan.child.sort_columns) == 1:
            return _sorted_limit(plan.child, plan, need)
        data = execute(plan.child, need)
        if not data:
            return {}
        n = _length(data)
        k = max(0, min(plan.limit, n))
        if k == 0:
            return {c: v[:0] for c, v in data.items()}
        return {c: (v[-k:] if plan.from_end else v[:k]) for c, v in data.items()}
    if isinstance(plan, SortPlan):
        cols = _need(plan, need)
        data = execute(plan.child, sorted(set(cols) | set(plan.sort_columns)))
        out = sort_columns(data, plan.sort_columns, plan.ascending)
        return {c: v for c, v in out.items() if c in cols}
    if isinstance(plan, GroupByPlan):
        by = _as_list(plan.by)
        data = execute(plan.child, sorted(set(by) | set(plan.agg_dict)))
        res = groupby_columns(data, by, plan.agg_dict)
        return {c: v for c, v in res.items() if need is None or c in need}
    if isinstance(plan, JoinPlan):
        res = join_columns(plan, need)
        return res
    raise ValueError(f"Unknown plan type: {type(plan)}")


def _length(data: dict) -> int:
    for v in data.values():
        return len(v)
    return 0


def sort_columns(data: dict, by: list, ascending) -> dict:
    """Rows of data ordered by the key columns (core/frame.py:763-803 and
    1155-1223): one key sorts with the row index and gathers every column;
    several keys go through sort_multi_column."""
    from . import parallel_algorithms as pa
    for c in by:
        if isinstance(data[c], np.ndarray):
            raise TypeError(f"Cannot sort by object dtype column '{c}'")
    sorter = pa.ParallelRadixSort()
    if len(by) == 1:
        key = by[0]
        asc = ascending if isinstance(ascending, bool) else ascending[0]
        n = len(data[key])
        rows = torch.arange(n, dtype=torch.int64, device=data[key].device)
        _, idx = sorter.sort(data[key], values=rows, ascending=asc)
        return {c: take_column(v, idx) for c, v in data.items()}
    keys = [data[c] for c in by]
    vals = {c: v for c, v in data.items() if c not in by}
    sk, sv = sorter.sort_multi_column(keys, values=vals, ascending=ascending)
    out = {c: sk[i] for i, c in enumerate(by)}
    if sv:
        out.update(sv)
    return out


def _sorted_limit(sort: SortPlan, limit: LimitPlan, need) -> dict:
    """head/tail of a single-key sort: the key column is sorted with its row
    index; the other columns are gathered for the kept rows only."""
    from . import parallel_algorithms as pa
    cols = _need(sort, need)
    key = sort.sort_columns[0]
    data = execute(sort.child, sorted(set(cols) | {key}))
    if isinstance(data[key], np.ndarray):
        raise TypeError(f"Cannot sort by object dtype column '{key}'")
    asc = sort.ascending if isinstance(sort.ascending, bool) else sort.ascending[0]
    n = len(data[key])
    k = max(0, min(limit.limit, n))
    rows = torch.arange(n, dtype=torch.int64, device=data[key].device)
    _, idx = pa.ParallelRadixSort().sort(data[key], values=rows, ascending=asc)
    idx = idx[n - k:] if limit.from_end else idx[:k]
    return {c: take_column(data[c], idx) for c in output_columns(sort) if c in cols and c in data}


def groupby_columns(data: dict, by: list, agg_dict: dict) -> dict:
    """{by..., aggregated columns} (core/frame.py:805-842, 1283-1322)."""
    from . import parallel_algorithms as pa
    for c in by:
        if isinstance(data[c], np.ndarray):
            raise TypeError(f"Cannot group by object dtype column '{c}'")
    values = {c: data[c] for c in agg_dict}
    if len(by) == 1:
        keys, agg = pa.groupby_aggregate(data[by[0]], values, agg_dict)
        out = {by[0]: keys}
    else:
        out, agg = pa.groupby_aggregate_multi_column([data[c] for c in by], by, values, agg_dict)
        out = dict(out)
    out.update(agg)
    return out


def join_columns(plan: JoinPlan, need=None) -> dict:
    """{key, left_<c>, right_<c>} of a join (core/frame.py:844-893); only the
    payload columns in need are carried."""
    from . import parallel_algorithms as pa
    lk, rk = list(plan.left_keys), list(plan.right_keys)
    lcols = [c for c in output_columns(plan.left) if c not in lk]
    rcols = [c for c in output_columns(plan.right) if c not in rk]
    if need is not None:
        lcols = [c for c in lcols if f"left_{c}" in need]
        rcols = [c for c in rcols if f"right_{c}" in need]
    left = execute(plan.left, lk + lcols)
    right = execute(plan.right, rk + rcols)
    return join_data(left, right, lk, rk, plan.join_type, lcols, rcols)


def join_data(left: dict, right: dict, lk: list, rk: list, how: str, lcols: list,
              rcols: list) -> dict:
    from . import parallel_algorithms as pa
    for c in lk:
        if isinstance(left[c], np.ndarray):
            raise TypeError(f"Cannot join on object dtype column '{c}'")
    for c in rk:
        if isinstance(right[c], np.ndarray):
            raise TypeError(f"Cannot join on object dtype column '{c}'")
    lv = {c: left[c] for c in lcols}
    rv = {c: right[c] for c in rcols}
    joiner = pa.ParallelSortMergeJoin()
    if len(lk) == 1:
        keys, vals = joiner.join(left[lk[0]], lv, right[rk[0]], rv, how=how)
        out = {lk[0]: keys}
    else:
        kd, vals = joiner.join_multi_column([left[c] for c in lk], lk, lv, [right[c] for c in rk],
                                            rk, rv, how=how)
        out = dict(kd)
    out.update(vals)
    return out
