"""MI355X-native Flysim simulation core.

The hot path (conductance LIF update, spike delivery, Poisson input) runs as
hand-written HIP kernels for gfx950 behind the C ABI of include/flysim.h;
this package holds the HIP/C++ sources (csrc/), the ctypes front end
(engine.py), synthetic workloads (workloads.py) and the sweep tooling.
"""
__all__ = ["engine", "workloads"]
