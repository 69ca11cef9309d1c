"""MI355X host mirror of HyperCloudCore.jl's explicit-step API (ctypes over libhcc)."""
from . import _lib
from ._lib import Context, Grid, Ensemble, MLP, Stepper, HccError, cloud_generate, rng_uniform, lib
from . import workloads
