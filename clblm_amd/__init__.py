"""MI355X-native CLBLM layer optical-depth path (HIP kernels behind a C-ABI)."""
