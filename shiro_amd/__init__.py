"""shiro_amd: MI355X (gfx950) path-tracing backend for Shiro.

The product is the C-ABI library libshiro_hip.so (include/shiro_hip.h) built
from shiro_amd/csrc; this package holds its ctypes binding (`hip`), the host
data model mirroring Shiro's runtime types (`scene`) and the seeded
workloads (`workloads`).
"""
__all__ = ["abi", "scene", "workloads"]
