"""aurora_amd — Python handles over libaurora, the gfx950 path-tracing hot
path (C ABI: include/aurora.h).  Rendering always runs the HIP kernels;
there is no CPU fallback, and a missing library raises RuntimeError."""
from . import abi
from .core import (Camera, Context, RenderSettings, Scene, SceneArrays, checker, image, noise, resolve_rgb8,
                   solid, write_ppm)

__all__ = ["abi", "Camera", "Context", "RenderSettings", "Scene", "SceneArrays", "checker", "image", "noise",
           "resolve_rgb8", "solid", "write_ppm"]
