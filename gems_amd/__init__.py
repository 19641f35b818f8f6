"""MI355X (gfx950) implementation of the GEMS force-evaluation hot path.

``gems_amd.hip`` binds the C ABI of include/gems_hip.h (library built in-tree
by ``gems_amd/build.py``).
"""
__all__ = ["hip"]
