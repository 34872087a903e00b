"""MI355X (gfx950) ray-tracing hot path of gpugi: traversal of the SAH BVH,
FP32 ray/triangle intersection and the PT/LT/BPT integrators behind a C ABI
(include/gpugi_rt.h).  `gpugi_amd.rt` is the ctypes binding of that ABI."""
from . import rt  # noqa: F401
