"""rblas_amd — rblas's Level-3 hot path on MI355X (gfx950) behind a C ABI.

Python face of librblas_hip.so, mirroring rblas's modules:
  level3    gemm, try_gemm, syrk, trsm, symm (level3.rs)
  complex3  complex gemm (4M), symm, hemm, trsm (complex3.rs)
  level2    gemv, ger; complex gemv, geru, gerc (level2.rs, complex2.rs)
  level1    dot, nrm2, axpy (level1.rs)
  grouped   grouped / batched gemm and trsm in one persistent launch
  serve     the continuous-batching request loop
  device    HBM buffers, streams and events for resident operands
  types     Layout, Transpose, Uplo, Side, Diag, BlasError
"""
from . import complex3, level1, level2, level3, types  # noqa: F401
from .types import BlasError, BlasPanic, Diag, Layout, Side, Transpose, Uplo  # noqa: F401
