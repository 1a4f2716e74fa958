"""Device fields: torch tensors on the GPU behind the Taichi field methods the
reference's callers use (``to_numpy``, ``from_numpy``, ``[index]``,
``shape``, ``dtype``, ``fill``).

The tensor is the storage the HIP library reads and writes in place
(``Field.tensor`` / ``Field.ptr``).  Vector and matrix fields keep their
component axes last, so ``shape`` is the grid shape (as for Taichi's
``Vector.field``) while ``to_numpy()`` returns the components too.
"""

from __future__ import annotations

import numpy as np
import torch

_DTYPE_NAMES = {torch.float32: "f32", torch.float64: "f64", torch.int32: "i32", torch.float16: "f16",
                torch.bfloat16: "bf16", torch.uint8: "u8"}


class Field:
    def __init__(self, shape, dtype, *, components=(), device="cuda"):
        self._grid_shape = tuple(int(n) for n in shape)
        self._components = tuple(int(n) for n in components)
        self.tensor = torch.zeros(self._grid_shape + self._components, dtype=dtype, device=device)

    # -- Taichi-like surface ---------------------------------------------
    @property
    def shape(self) -> tuple[int, ...]:
        return self._grid_shape

    @property
    def dtype(self) -> str:
        return _DTYPE_NAMES[self.tensor.dtype]

    @property
    def ptr(self) -> int:
        return self.tensor.data_ptr()

    def to_numpy(self) -> np.ndarray:
        return self.tensor.detach().cpu().numpy()

    def from_numpy(self, array) -> None:
        arr = np.asarray(array)
        full = self._grid_shape + self._components
        if arr.shape != full:
            raise ValueError(f"from_numpy: expected shape {full}, got {arr.shape}")
        self.tensor.copy_(torch.from_numpy(np.ascontiguousarray(arr)).to(self.tensor.dtype))

    def from_torch(self, t: torch.Tensor) -> None:
        self.tensor.copy_(t)

    def fill(self, value) -> None:
        self.tensor.fill_(value)

    def copy_from(self, other: "Field") -> None:
        self.tensor.copy_(other.tensor)

    def _key(self, key):
        if key is None:
            return ()
        return key

    def __getitem__(self, key):
        v = self.tensor[self._key(key)]
        if v.dim() == 0:
            return v.item()
        return v.detach().cpu().numpy()

    def __setitem__(self, key, value):
        self.tensor[self._key(key)] = torch.as_tensor(value, dtype=self.tensor.dtype)

    def __repr__(self) -> str:
        return f"Field(shape={self.shape}, components={self._components}, dtype={self.dtype})"


def scalar_field(shape, dtype) -> Field:
    return Field(shape, dtype)


def vector_field(n: int, shape, dtype) -> Field:
    return Field(shape, dtype, components=(n,))


def matrix_field(n: int, m: int, shape, dtype) -> Field:
    return Field(shape, dtype, components=(n, m))
