"""narvalengine_amd -- MI355X (gfx950) path-tracing backend for NarvalEngine.

Python view of the C ABI in include/ne_gpu.h (ctypes).  The compute lives in
libnarval_gpu.so (HIP kernels + C++ host runtime) next to this file; importing
the package does not load it, every GPU entry point does, and fails loudly
when the library is missing.
"""
from __future__ import annotations

from .abi import (  # noqa: F401
    CameraDesc, Counters, NarvalError, RenderParams, SceneDesc, SceneInfo, lib, lib_path,
    NE_FP32, NE_FP64, NE_FLAG_COUNT, NE_MAT_GLOSSY, NE_MAT_EMITTER, NE_ENV_NONE, NE_ENV_CONSTANT, NE_ENV_SKY,
)
from .runtime import Batch, Context, Film, GpuScene, Guide  # noqa: F401
from .scene import Scene, Camera  # noqa: F401
from .engine import OfflineEngine, GpuPathIntegrator, SceneSettings  # noqa: F401
