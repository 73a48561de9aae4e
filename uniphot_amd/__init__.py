"""uniphot_amd -- MI355X-native forced-photometry fit path of uniphot
(``tract7dt.fit``): source rendering, per-island VarPro-LM and the
multi-band linear flux solve, as hand-written HIP kernels behind a C ABI.

Host modules mirror the reference's ``tract7dt.fit.grid_gpu`` names and
array layouts; device buffers are torch-ROCm tensors.
"""
__version__ = "0.1.0"
