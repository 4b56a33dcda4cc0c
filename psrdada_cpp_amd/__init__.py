"""MI355X-native MeerKAT FBFUSE beamformer.

The compute path is the native library ``libfbfuse_hip.so`` (hand-written HIP
for gfx950 behind the C ABI in ``include/fbfuse_hip.h``); ``capi`` binds it
for Python, ``workloads`` describes the standard beamformer configurations and
their synthetic inputs.  The C++ operator API (``SplitTranspose``,
``CoherentBeamformer``, ``IncoherentBeamformer``, ``WeightsManager``,
``DoubleDeviceBuffer``, ``Pipeline``) lives in ``cpp/``.
"""

__all__ = ["capi", "workloads"]
