"""JaxSeries: one column, eager (device array) or lazy (an expression).

Mirrors the parts of core/series.py the frame's filter path uses: the
comparison operators (_comparison_op / _apply_comparison, series.py:402-484)
and &, |, ~ (series.py:510-576). Eager comparisons give a boolean device
mask; lazy ones build ComparisonOp / BinaryOp / UnaryOp expressions for a
FilterPlan.
"""

from __future__ import annotations

import numpy as np
import torch

from .lazy import BinaryOp, Column, ComparisonOp, ComparisonOpType, Literal, UnaryOp

_SCALARS = (int, float, np.number)


class JaxSeries:
    def __init__(self, data=None, name: str | None = None, index=None, lazy: bool = False,
                 parent_frame=None, expr=None, dtype: str | None = None, string_vocab=None):
        self._string_vocab = string_vocab
        if isinstance(data, np.ndarray) and dtype is None and string_vocab is None:
            from .strings import encode_string_array, is_string_array
            if data.dtype.kind in "OUS" and is_string_array(data):
                data, self._string_vocab = encode_string_array(data)
                dtype = "string"
        self._lazy = lazy
        self._parent_frame = parent_frame
        self._expr = expr
        self.name = name
        self.index = index
        self.data = data
        self._dtype = dtype if dtype is not None else (
            None if data is None else _dtype_name(data))

    @classmethod
    def _from_processed_data(cls, data, name=None, index=None, dtype=None, string_vocab=None):
        return cls(data, name=name, index=index, dtype=dtype, string_vocab=string_vocab)

    @property
    def dtype(self):
        return self._dtype

    @property
    def is_lazy(self) -> bool:
        return self._lazy

    def __len__(self) -> int:
        if self._lazy:
            raise TypeError("a lazy series has no length before collect()")
        return len(self.data)

    def __repr__(self) -> str:
        if self._lazy:
            return f"JaxSeries(lazy=True, expr={self._expr!r})"
        return f"JaxSeries(name={self.name!r}, length={len(self.data)}, dtype={self._dtype})"

    def to_numpy(self) -> np.ndarray:
        from ._tensors import to_numpy
        if self._dtype == "string":
            from .strings import decode_string_codes
            return decode_string_codes(self.data, self._string_vocab or ())
        if isinstance(self.data, np.ndarray):
            return self.data
        return to_numpy(self.data)

    def tolist(self) -> list:
        return self.to_numpy().tolist()

    def __array__(self, dtype=None, copy=None):
        a = self.to_numpy()
        return a if dtype is None else a.astype(dtype)

    # ---- comparisons ----------------------------------------------------------------

    def _expr_of(self):
        return self._expr if self._expr is not None else Column(self.name)

    def _comparison_op(self, other, op: ComparisonOpType):
        if self._lazy:
            if isinstance(other, JaxSeries):
                right = other._expr_of() if other._lazy else Literal(other.data)
            elif isinstance(other, _SCALARS):
                right = Literal(other)
            else:
                return NotImplemented
            return JaxSeries(None, None, lazy=True, parent_frame=self._parent_frame,
                             expr=ComparisonOp(op=op, left=self._expr_of(), right=right))
        if isinstance(self.data, np.ndarray):
            raise TypeError("comparisons on object columns are not supported")
        fn = {ComparisonOpType.GT: torch.gt, ComparisonOpType.LT: torch.lt,
              ComparisonOpType.GE: torch.ge, ComparisonOpType.LE: torch.le,
              ComparisonOpType.EQ: torch.eq, ComparisonOpType.NE: torch.ne}[op]
        lhs = self.data
        if isinstance(other, JaxSeries):
            if other._lazy:
                raise ValueError("Cannot mix lazy and eager series in comparisons")
            rhs = other.data
            if self._dtype == "string" and other._dtype == "string":
                from .strings import align_string_code_arrays
                lhs, rhs, _ = align_string_code_arrays(self.data, self._string_vocab or (),
                                                       other.data, other._string_vocab or ())
        elif self._dtype == "string" and (other is None or isinstance(other, str)):
            # core/series.py:441-460: exact code for == / !=, insertion rank otherwise
            from .strings import string_scalar_code, string_scalar_rank
            vocab = self._string_vocab or ()
            code = string_scalar_code(vocab, other)
            if op in (ComparisonOpType.EQ, ComparisonOpType.NE):
                if code is None:
                    res = torch.full_like(lhs, op is ComparisonOpType.NE, dtype=torch.bool)
                    return JaxSeries._from_processed_data(res, name=self.name, dtype="bool")
                rhs = code
            else:
                if other is None:
                    return NotImplemented
                rhs = string_scalar_rank(vocab, other)
        elif isinstance(other, _SCALARS):
            rhs = other
        else:
            return NotImplemented
        return JaxSeries._from_processed_data(fn(lhs, rhs), name=self.name, dtype="bool")

    def __gt__(self, other):
        return self._comparison_op(other, ComparisonOpType.GT)

    def __lt__(self, other):
        return self._comparison_op(other, ComparisonOpType.LT)

    def __ge__(self, other):
        return self._comparison_op(other, ComparisonOpType.GE)

    def __le__(self, other):
        return self._comparison_op(other, ComparisonOpType.LE)

    def __eq__(self, other):
        return self._comparison_op(other, ComparisonOpType.EQ)

    def __ne__(self, other):
        return self._comparison_op(other, ComparisonOpType.NE)

    __hash__ = object.__hash__

    # ---- logic ------------------------------------------------------------------------

    def _logic(self, other, op: str):
        if self._lazy and isinstance(other, JaxSeries) and other._lazy:
            return JaxSeries(None, None, lazy=True, parent_frame=self._parent_frame,
                             expr=BinaryOp(left=self._expr_of(), op=op, right=other._expr_of()))
        if not self._lazy and not (isinstance(other, JaxSeries) and other._lazy):
            rhs = other.data if isinstance(other, JaxSeries) else other
            fn = torch.logical_and if op == "&" else torch.logical_or
            if self.data.dtype == torch.bool and (not isinstance(rhs, torch.Tensor)
                                                  or rhs.dtype == torch.bool):
                res = fn(self.data, torch.as_tensor(rhs, device=self.data.device))
            else:
                res = (self.data & rhs) if op == "&" else (self.data | rhs)
            return JaxSeries(res, name=self.name)
        raise ValueError("Cannot mix lazy and eager series in logical operations")

    def __and__(self, other):
        return self._logic(other, "&")

    def __or__(self, other):
        return self._logic(other, "|")

    def __invert__(self):
        if self._lazy:
            return JaxSeries(None, None, lazy=True, parent_frame=self._parent_frame,
                             expr=UnaryOp(op="~", operand=self._expr_of()))
        return JaxSeries(~self.data, name=self.name)


def _dtype_name(arr) -> str:
    if isinstance(arr, np.ndarray):
        return "object" if arr.dtype == np.object_ else str(arr.dtype)
    return str(arr.dtype).replace("torch.", "")
