"""CPU oracle for the sort-centric algorithms layer.

TEST INFRASTRUCTURE ONLY. This package restates, in numpy, the algorithms of
jaxframes' ``src/jaxframes/distributed/parallel_algorithms.py`` so that the
HIP kernels of ``jaxframes_amd`` can be checked against it. Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it, and only as the checker. The product path never routes through it.
"""

from .algorithms import *  # noqa: F401,F403
