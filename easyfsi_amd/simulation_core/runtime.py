"""Runtime configuration (simulation_core/runtime.py:12-59).

The reference initialises Taichi once per process on a CUDA GPU and rejects
``arch="cpu"``.  Here the device is the visible ROCm GPU (torch's device 0) and
``arch`` names it; the float-precision field is kept for interface parity:
device fields keep their declared dtypes either way.
"""

from __future__ import annotations

from dataclasses import dataclass

_SUPPORTED_ARCHS = ("cuda", "gpu", "hip", "rocm")
_SUPPORTED_FPS = ("f32", "f64")


@dataclass(frozen=True)
class RuntimeConfig:
    arch: str = "gpu"
    default_fp: str = "f32"
    random_seed: int = 0


# Drop-in name used by the reference's callers.
TaichiRuntimeConfig = RuntimeConfig

_INITIALIZED_FP: str | None = None


def init_runtime(config: RuntimeConfig | None = None) -> None:
    """Validate the config and make sure the GPU and libhfsi are usable.

    First call wins; a later call with another ``default_fp`` raises, like the
    reference's init_taichi.
    """
    global _INITIALIZED_FP
    cfg = config or RuntimeConfig()
    arch = str(cfg.arch).lower()
    if arch == "cpu":
        raise ValueError("simulation_core is GPU-only; use arch='gpu'")
    if arch not in _SUPPORTED_ARCHS:
        raise ValueError(f"unsupported arch: {cfg.arch!r}")
    fp = str(cfg.default_fp)
    if fp not in _SUPPORTED_FPS:
        raise ValueError(f"unsupported default_fp: {cfg.default_fp!r}; expected one of {_SUPPORTED_FPS}")
    if _INITIALIZED_FP is not None:
        if fp != _INITIALIZED_FP:
            raise ValueError(f"runtime is already initialized with default_fp={_INITIALIZED_FP!r}; "
                             f"cannot re-initialize with default_fp={fp!r}")
        return
    from .. import _lib

    _lib.lib()
    _INITIALIZED_FP = fp


init_taichi = init_runtime
