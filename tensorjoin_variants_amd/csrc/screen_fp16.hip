// FP16 screen instantiations (one term and hi/lo split).
#include "screen_impl.h"
TJ_SCREEN_DISPATCH(screen_fp16, P_FP16, TJ_CASE(P_FP16, 16) TJ_CASE(P_FP16, 32) TJ_CASE(P_FP16, 64)
                   TJ_CASE(P_FP16, 128) TJ_CASE(P_FP16, 256) TJ_CASE(P_FP16, 512))
TJ_SCREEN_DISPATCH(screen_fp16x3, P_FP16X3, TJ_CASE(P_FP16X3, 16) TJ_CASE(P_FP16X3, 32) TJ_CASE(P_FP16X3, 64)
                   TJ_CASE(P_FP16X3, 128) TJ_CASE(P_FP16X3, 256))
