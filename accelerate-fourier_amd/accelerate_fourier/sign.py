"""Transform direction (Data.Array.Accelerate.Fourier.Sign, Sign.hs:28-54)."""
from __future__ import annotations

from dataclasses import dataclass


@dataclass(frozen=True)
class Sign:
    """``newtype Sign a = Sign {getSign :: a}`` with forward = -1, inverse = +1."""

    value: int

    def __post_init__(self):
        if self.value not in (-1, 1):
            raise ValueError("a Sign is -1 (forward) or +1 (inverse)")

    def negate(self) -> "Sign":
        return Sign(-self.value)


forward = Sign(-1)
inverse = Sign(1)
