"""MI355X-native SMAUG MHD time-step engine (HIP kernels behind a C-ABI)."""
from .params import RunParams, SmaugParams, field_names, nvar, npert  # noqa: F401
