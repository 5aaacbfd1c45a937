"""MI355X-native all-pairs normalized cross-correlation for seismic events.

Drop-in for the reference's GPU path (rasenjop/seismic Correlation_in_CUDA and
Correlation_in_MultiCUDA): the same ctypes C ABI in ``lib/correlation_cuda.so``
and ``lib/correlation_multicuda.so``, backed by hand-written HIP kernels for
gfx950 in ``lib/libseismic_xcorr.so``.

Python entry points
  seismic_amd.correlation_lib   mirror of Correlation_in_CUDA/correlation_lib.py
  seismic_amd.multicuda_lib     mirror of Correlation_in_MultiCUDA/correlation_lib.py
  seismic_amd.engine            device-resident plans (extended C ABI)
  seismic_amd.synth             seeded synthetic events
"""
from ._lib import LIB_DIR, load_engine, library_path  # noqa: F401

__all__ = ["LIB_DIR", "load_engine", "library_path"]
