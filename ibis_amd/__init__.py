"""MI355X-native hourly land-surface step of IBIS 3.0 (HIP kernels behind a C ABI)."""
