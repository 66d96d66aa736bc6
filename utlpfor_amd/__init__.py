"""MI355X-native UTL-PFOR block codec (gfx950 HIP kernels behind a C-ABI).

`codec` mirrors the Akron/utlpfor Go API per call; `batch` holds the
device-resident batched forms (encode, index scan, decode, Get gather).
"""
from .codec import (Append, Compressor32, Compressor64, Delta, ErrInvalidBlockLength, ErrInvalidBuffer,
                    ErrInvalidFlags, ErrNotImplemented, ErrOverflow, ErrPositionOutOfRange, ErrUnsupportedType,
                    GetUint32, GetUint64, Header, BlockLength, MaxBlockLength32, MaxBlockLength64, NewUint32,
                    NewUint64, NoFOR, NoInPlace, NoPatch, PackUint32, PackUint64, Special, UnpackUint32,
                    UnpackUint64)

__all__ = [
    "Append", "Compressor32", "Compressor64", "Delta", "ErrInvalidBlockLength", "ErrInvalidBuffer",
    "ErrInvalidFlags", "ErrNotImplemented", "ErrOverflow", "ErrPositionOutOfRange", "ErrUnsupportedType",
    "GetUint32", "GetUint64", "Header", "BlockLength", "MaxBlockLength32", "MaxBlockLength64", "NewUint32",
    "NewUint64", "NoFOR", "NoInPlace", "NoPatch", "PackUint32", "PackUint64", "Special", "UnpackUint32",
    "UnpackUint64",
]
