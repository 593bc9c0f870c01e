"""yastn_amd -- MI355X (gfx950) block-sparse backend for YASTN.

``yastn_amd.backend_hip`` is the drop-in backend module (``yastn.make_config(backend=backend_hip)``);
its hot path runs hand-written HIP kernels from ``libyastn_hip.so`` through a C ABI
(include/yastn_hip.h).  ``yastn_amd.blocksparse`` restates the reference planner and tensor layer
needed to drive the backend where the reference package is not installed.
"""
__version__ = '0.1.0'
