"""Flag values from include/sb_status.h, for tests."""
EXACT_JOIN = 1 << 0
SKIP_ZERO_SLICES = 1 << 1
PREC_FP64 = 0
PREC_FP32 = 1 << 8
PREC_BF16X3 = 2 << 8
