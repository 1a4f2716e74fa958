"""EasyFsi hot path on MI355X: FV pressure projection and neo-Hookean MPM
transfers as hand-written gfx950 HIP kernels behind a C ABI (libhfsi.so),
with the reference's simulation_core class interfaces on top."""

__version__ = "0.1.0"
