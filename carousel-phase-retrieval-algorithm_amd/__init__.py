"""MI355X (gfx950) backend of the Carousel Phase Retrieval Algorithm.

The product is libcpra_hip.so (csrc/, C ABI in include/cpra_hip.h and
include/cpra_hip_engine.h).  This package holds its build recipe
(``build``), the ctypes front end mirroring ``CpraImpl<T>`` (``hip``) and
the seeded synthetic workload generators (``synth``).

The directory name is not a Python identifier; load it with
``load_package()`` from ``__graft_entry__`` (it registers the package as
``cpra_amd``).
"""
