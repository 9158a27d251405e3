"""spinwalk_amd -- MI355X (gfx950) engine for the SpinWalk `sim` hot path.

The compute path is the HIP library ``spinwalk_amd/lib/libspinwalk_engine.so``
behind the C ABI ``include/spinwalk_engine.h``; ``spinwalk_amd.engine`` binds it
with ctypes.  There is no CPU fallback.
"""
__version__ = "0.1.0"
