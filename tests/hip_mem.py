"""Device memory for GPU tests through the HIP runtime the library itself
uses (libamdhip64.so.7), so no second runtime (torch's bundled one) is loaded
into the test process.  The package's devmem module provides it."""
from quantkernel_amd.devmem import DevArray, synchronize, to_device  # noqa: F401
