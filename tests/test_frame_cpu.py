"""Frame API and lazy plans without a GPU: plan construction, schemas,
printing, the columns collect() produces, and the errors the reference
raises (core/frame.py, lazy/plan.py; cases after tests/test_lazy_operations.py,
tests/test_head_tail.py and tests/test_lazy_integration.py of the reference)."""

import numpy as np
import pytest

from jaxframes_amd import lazy as L
from jaxframes_amd.frame import JaxFrame


def _lazy():
    return JaxFrame({"a": np.array([3, 1, 4, 1, 5]), "b": np.array([30.0, 10.0, 40.0, 15.0, 50.0]),
                     "g": np.array([1, 2, 1, 2, 1])}, lazy=True)


def test_lazy_frame_builds_plans_without_executing():
    df = _lazy()
    assert df.is_lazy and isinstance(df.plan, L.InputPlan)
    assert isinstance(df.sort_values("a").plan, L.SortPlan)
    assert isinstance(df.groupby("g").sum().plan, L.GroupByPlan)
    assert isinstance(df[df["a"] > 3].plan, L.FilterPlan)
    assert isinstance(df[["a", "b"]].plan, L.ProjectPlan)
    assert isinstance(df.head(2).plan, L.LimitPlan) and not df.head(2).plan.from_end
    assert df.tail(2).plan.from_end
    m = df.merge(JaxFrame({"a": np.array([1, 2]), "c": np.array([1.0, 2.0])}), on="a")
    assert isinstance(m.plan, L.JoinPlan) and isinstance(m.plan.right, L.InputPlan)


def test_filter_expressions():
    df = _lazy()
    c = (df["a"] > 3)._expr
    assert isinstance(c, L.ComparisonOp) and c.op is L.ComparisonOpType.GT
    assert isinstance(c.left, L.Column) and c.left.name == "a" and c.right.value == 3
    both = ((df["a"] > 1) & (df["b"] < 45.0))._expr
    assert isinstance(both, L.BinaryOp) and both.op == "&"
    assert both.columns() == {"a", "b"}
    neg = (~(df["a"] == 1))._expr
    assert isinstance(neg, L.UnaryOp) and neg.op == "~"


def test_schemas_and_output_columns():
    df = _lazy()
    gp = df.groupby("g").agg({"b": "mean"})
    assert list(gp.plan.schema()) == ["g", "b"] and gp.columns == ["g", "b"]
    srt = df.sort_values(["g", "a"])
    assert srt.columns == ["g", "a", "b"]          # multi-key sort: keys first
    assert df.sort_values("b").columns == ["a", "b", "g"]
    m = df.merge(JaxFrame({"a": np.array([1]), "c": np.array([1.0])}, lazy=True), on="a")
    assert m.columns == ["a", "left_b", "left_g", "right_c"]
    # the reference's declared JoinPlan schema (lazy/plan.py:353-374)
    assert list(m.plan.schema()) == ["a", "b", "g", "c"]
    assert df[["b"]].head(3).columns == ["b"]
    assert df.groupby("g").sum().plan.agg_dict == {"a": "sum", "b": "sum"}


def test_explain_and_repr():
    df = _lazy()
    q = df[df["a"] > 3].sort_values("a").head(2)
    text = q.explain()
    assert text.startswith("Query Plan:\nLimitPlan(head(2))")
    assert "SortPlan(by=['a'], ascending=True)" in text
    assert "FilterPlan(condition=(Column('a') > Literal(3)))" in text
    assert "InputPlan(5 rows, 2 cols" not in text and "InputPlan(5 rows, 3 cols" in text
    assert JaxFrame({"a": np.array([1])}).explain() == "Eager execution (no plan)"
    assert repr(q) == "JaxFrame(lazy=True, columns=['a', 'b', 'g'])"


def test_errors():
    df = _lazy()
    with pytest.raises(KeyError):
        df.sort_values("zz")
    with pytest.raises(KeyError):
        df.groupby("zz")
    with pytest.raises(KeyError):
        df.merge(JaxFrame({"b": np.array([1])}), on="a")
    with pytest.raises(KeyError):
        df["zz"]
    with pytest.raises(TypeError):
        df[3.5]
    eager = JaxFrame({"s": np.array(["x", "y"], dtype=object),
                      "o": np.array([{"a": 1}, [2]], dtype=object), "v": np.array([1.0, 2.0])})
    assert eager._dtypes["s"] == "string" and eager._string_vocabs["s"] == ("x", "y")
    assert eager._dtypes["o"] == "object"
    with pytest.raises(TypeError):
        eager.sort_values("o")
    with pytest.raises(TypeError):
        eager.groupby("o")
    with pytest.raises(ValueError):
        JaxFrame({"a": np.array([1, 2]), "b": np.array([1.0])})


def test_simple_comparison_detection():
    assert L._simple_comparison(L.ComparisonOp(L.ComparisonOpType.GT, L.Column("v"), L.Literal(2.5))) \
        == ("v", ">", 2.5)
    # scalar on the left flips the operator
    assert L._simple_comparison(L.ComparisonOp(L.ComparisonOpType.GT, L.Literal(2), L.Column("v"))) \
        == ("v", "<", 2)
    assert L._simple_comparison(L.ComparisonOp(L.ComparisonOpType.GT, L.Column("a"), L.Column("b"))) \
        is None
    assert L._simple_comparison(L.BinaryOp(L.Column("a"), "&", L.Column("b"))) is None
