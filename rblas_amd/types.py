"""rblas's descriptor enums and errors (types.rs:9-72, error.rs:15-60).

Enum values are the C-ABI encodings (declaration order). `Transpose.None`
is spelled `Transpose.NoTrans` because `None` is a Python keyword.
"""
from __future__ import annotations

import enum
from dataclasses import dataclass


class Layout(enum.IntEnum):
    RowMajor = 0
    ColMajor = 1


class Transpose(enum.IntEnum):
    NoTrans = 0
    Trans = 1
    ConjTrans = 2

    def is_transposed(self) -> bool:  # types.rs:42-44
        return self in (Transpose.Trans, Transpose.ConjTrans)


class Uplo(enum.IntEnum):
    Upper = 0
    Lower = 1


class Side(enum.IntEnum):
    Left = 0
    Right = 1


class Diag(enum.IntEnum):
    NonUnit = 0
    Unit = 1


class Compute(enum.IntEnum):
    """SGEMM arithmetic (rblas_sgemm_ex): plain f32, or f32 operands split
    into 2 (BF16X3, 3 products) or 3 (BF16X6, 6 products) bf16 terms."""
    F32 = 0
    BF16X3 = 1
    BF16X6 = 2


# C-ABI status codes (include/rblas_hip.h)
OK = 0
INVALID_LEADING_DIM = 1
BUFFER_TOO_SMALL = 2
ZERO_STRIDE = 3
PANIC = 10
BAD_ENUM = 11
BAD_ARG = 12
HIP_ERROR = 20
ALLOC_FAILED = 21


@dataclass(frozen=True)
class BlasError(Exception):
    """rblas's `BlasError` (error.rs:15-38). `kind` is 'InvalidLeadingDim',
    'BufferTooSmall' or 'ZeroStride'; `bound` is `min` or `need`."""
    kind: str
    which: str
    got: int = 0
    bound: int = 0

    def __str__(self) -> str:  # error.rs:41-57
        if self.kind == "InvalidLeadingDim":
            return (f"invalid leading dimension for `{self.which}`: got {self.got}, "
                    f"need ≥ {self.bound}")
        if self.kind == "BufferTooSmall":
            return (f"buffer `{self.which}` too small: got {self.got} elements, "
                    f"need ≥ {self.bound}")
        return f"stride for `{self.which}` must be positive (got 0)"


class BlasPanic(RuntimeError):
    """Raised where rblas's unchecked routine would panic."""


class DeviceError(RuntimeError):
    """The HIP runtime (or a device allocation) failed."""
