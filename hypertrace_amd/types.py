"""Scene byte-stream configuration, layout arithmetic and writer.

Mirrors the reference's `types` crate: `Config` (types/src/config.rs:38-49),
`upper_multiple` / `aligned_add` / `lcm` (types/src/math.rs:6-16) and the
counting writer with `align` / `skip` (types/src/io.rs:136-180). Only the
little-endian layout is produced (the reference's examples all use
`Endian::Little`, main/examples/hy.rs:89-93).
"""
from __future__ import annotations

import math
import struct
from dataclasses import dataclass


@dataclass(frozen=True)
class Config:
    """Device layout switches: `usize` width and whether `f64` fields are
    stored as double (`real` = double on the device) or narrowed to f32."""

    address_width: int = 32
    double_support: bool = False

    def __post_init__(self):
        if self.address_width not in (32, 64):
            raise ValueError("address_width must be 32 or 64")

    @property
    def usize(self) -> int:
        return self.address_width // 8

    @property
    def real(self) -> int:
        return 8 if self.double_support else 4


X32_F32 = Config(32, False)
X32_F64 = Config(32, True)


def upper_multiple(x: int, m: int) -> int:
    return m * ((x + m - 1) // m)


def aligned_add(a: int, b: int, m: int) -> int:
    return upper_multiple(a, m) + b


def lcm(a: int, b: int) -> int:
    return a // math.gcd(a, b) * b


class Writer:
    """Little-endian byte writer that tracks its position for alignment."""

    def __init__(self, cfg: Config):
        self.cfg = cfg
        self.buf = bytearray()
        self.marks = None  # dict id(entity) -> byte offset, when recording

    @property
    def position(self) -> int:
        return len(self.buf)

    def align(self, m: int) -> None:
        pad = upper_multiple(len(self.buf), m) - len(self.buf)
        self.buf.extend(b"\0" * pad)

    def skip(self, n: int) -> None:
        self.buf.extend(b"\0" * n)

    def u32(self, v: int) -> None:
        self.buf.extend(struct.pack("<I", v & 0xFFFFFFFF))

    def usize(self, v: int) -> None:
        self.buf.extend(struct.pack("<Q" if self.cfg.usize == 8 else "<I", v))

    def f32(self, v: float) -> None:
        self.buf.extend(struct.pack("<f", v))

    def real(self, v: float) -> None:
        """An `f64` field: double with `double_support`, else narrowed to f32
        (types/src/primitive/scalar.rs:169-226)."""
        self.buf.extend(struct.pack("<d" if self.cfg.double_support else "<f", v))

    def bytes(self) -> bytes:
        return bytes(self.buf)
