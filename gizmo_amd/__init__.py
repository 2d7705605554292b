"""MI355X-native MFM hydro neighbour-loop pipeline (density -> gradients -> force).

The compute path is the HIP library libgizmo_hip.so (built by ``python -m gizmo_amd.build``),
reached through the C ABI in include/gizmo_hydro.h; ``gizmo_amd.hydro.Hydro`` binds it.
"""
from .hydro import GZ_FP32_MIXED, GZ_FP64, GizmoHipError, Hydro, LIB_PATH  # noqa: F401
