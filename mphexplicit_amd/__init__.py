"""MI355X-native MPH-WC explicit solver hot path (HIP kernels behind a C ABI)."""
