"""dvc_amd -- MI355X (gfx950) Wyner-Ziv decoding hot path of the NTU DVC codec.

The compute path is libdvc_accel.so (hand-written HIP kernels behind the C ABI
of include/dvc_accel.h); this package holds its ctypes binding (_lib, which
raises on import when the library is missing -- there is no CPU fallback),
the host-side decoder that mirrors the reference's decodeGOP / ParallelMode
entry points (decoder), the synthetic workloads and WZ encoder that produce
its inputs (workload), and the serving loop (serve).
"""
