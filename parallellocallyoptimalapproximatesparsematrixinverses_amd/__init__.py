"""MI355X-native Global SPAI / inner-outer preconditioner builders and PCG.

Layout:
  csrc/            HIP kernels (gfx950) and the C ABI of libspai_hip.so
  methods_cython/  Python API of the reference's methods_cython package
  _lib.py          ctypes bindings of the extension entry points
  autograd.py      torch.autograd.Function for ||I - A M||_F^2
"""
__all__ = ["methods_cython"]
