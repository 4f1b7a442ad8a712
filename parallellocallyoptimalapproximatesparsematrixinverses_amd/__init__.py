"""MI355X-native Global SPAI / inner-outer preconditioner builders and PCG.

Layout:
  csrc/            HIP kernels (gfx950) and the C ABI of libspai_hip.so
  methods_cython/  Python API of the reference's methods_cython package
  _lib.py          ctypes bindings of the extension entry points
  harness.py       backward-error harness (prepared PCG system, RHS recipe)
  autograd.py      torch.autograd.Function for ||I - A M||_F^2
  datasets.py      sample-dataset study (MatrixMarket loading, construction
                   and backward-error tables of scripts/common.py)
  speedup.py       thread-speedup study format of
                   scripts/methods_thread_speedup.py, with a device column

One HIP runtime per process: libspai_hip.so needs libamdhip64, and the dynamic
loader reuses a copy that is already loaded.  PyTorch ships its own; if
libspai_hip.so were loaded first, torch would load a second runtime and find
no device.  So when torch is installed it is imported here, before anything
loads libspai_hip.so, and both run on torch's runtime.
"""
try:
    import torch  # noqa: F401  (see above: load order of the HIP runtime)
except ImportError:
    pass

__all__ = ["methods_cython"]
