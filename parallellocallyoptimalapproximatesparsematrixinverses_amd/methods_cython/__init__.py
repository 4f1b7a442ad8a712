"""Device-backed counterpart of the reference package ``methods_cython``
(methods_cython/__init__.py): the SPAI builders and the PCG evaluator."""
from .methods import OutputResult, global_spai, inner_outer

__all__ = ["OutputResult", "global_spai", "inner_outer"]
