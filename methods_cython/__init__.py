"""The reference's import path ``methods_cython`` (methods_cython/__init__.py:1-3,
used as ``from methods_cython import global_spai, inner_outer`` in
scripts/utils/methods.py:3 and ``from methods_cython.backward_error import
matvec, pcg, parallel_threads`` in scripts/utils/backward_error.py:9) for this
library's Cython modules, which are built in
parallellocallyoptimalapproximatesparsematrixinverses_amd/methods_cython/.
``methods_cython.methods`` and ``methods_cython.backward_error`` are those
compiled modules themselves."""
import sys

from parallellocallyoptimalapproximatesparsematrixinverses_amd.methods_cython import backward_error, methods
from parallellocallyoptimalapproximatesparsematrixinverses_amd.methods_cython.methods import (OutputResult,
                                                                                               global_spai,
                                                                                               inner_outer)

sys.modules[__name__ + ".methods"] = methods
sys.modules[__name__ + ".backward_error"] = backward_error

__all__ = ["OutputResult", "global_spai", "inner_outer"]
