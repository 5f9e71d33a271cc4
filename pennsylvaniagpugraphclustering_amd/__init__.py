"""MI355X-native Louvain/Leiden clustering hot path (PGGC drop-in).

The compute lives in libpggc.so (hand-written HIP kernels for gfx950, C ABI in
include/pggc.h); ``clustering`` holds the host-side mirror of the reference's
clustering interface and ``graphs`` the synthetic graph generators.
"""
from . import graphs  # noqa: F401

__all__ = ["graphs", "clustering"]
