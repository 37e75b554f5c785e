"""jaxframes' sort-centric algorithms layer on MI355X (gfx950) HIP kernels.

- frame.JaxFrame: the DataFrame API (sort_values, groupby().agg, merge,
  boolean filters, head/tail; lazy plans with collect()).
- parallel_algorithms: the reference's operator surface (sort, groupby,
  join, hash partition) with the same names and conventions.
- batched / serving: many groupbys in one launch; a request loop over
  device-resident frames.
"""

from .frame import GroupBy, JaxFrame  # noqa: F401
from .series import JaxSeries  # noqa: F401
