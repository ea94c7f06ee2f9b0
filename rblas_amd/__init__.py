"""rblas_amd — rblas's Level-3 hot path on MI355X (gfx950) behind a C ABI.

Python face of librblas_hip.so, mirroring rblas's modules:
  level3   gemm, try_gemm (level3.rs)
  device   HBM buffers, streams and events for resident operands
  types    Layout, Transpose, Uplo, Side, Diag, BlasError
"""
from . import complex3, level1, level3, types  # noqa: F401
from .types import BlasError, BlasPanic, Diag, Layout, Side, Transpose, Uplo  # noqa: F401
