"""Letters <-> ncbistdaa codes for test fixtures (no library needed)."""
import numpy as np

STDAA = "-ABCDEFGHIKLMNPQRSTVWXYZU*OJ"
_CODE = {c: i for i, c in enumerate(STDAA)}


def encode(seq: str) -> np.ndarray:
    return np.array([_CODE[c] for c in seq.upper()], dtype=np.uint8)


def decode(codes) -> str:
    return "".join(STDAA[int(c)] for c in codes)
