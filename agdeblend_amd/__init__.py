"""MI355X-native AGDeblend deblending path (HIP kernels behind the C API)."""
