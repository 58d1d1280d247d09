"""MI355X-native Compact-Kernel MPM substep (HIP kernels behind a C-ABI)."""
