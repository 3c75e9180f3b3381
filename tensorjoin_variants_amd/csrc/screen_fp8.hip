// FP8 (OCP e4m3, per-row power-of-two E8M0 scales) screen instantiations.
#include "screen_impl.h"
TJ_SCREEN_DISPATCH(screen_fp8, P_FP8, TJ_CASE(P_FP8, 64) TJ_CASE(P_FP8, 128) TJ_CASE(P_FP8, 256) TJ_CASE(P_FP8, 512))
TJ_SCREEN_DISPATCH(screen_fp8x3, P_FP8X3, TJ_CASE(P_FP8X3, 64) TJ_CASE(P_FP8X3, 128) TJ_CASE(P_FP8X3, 256)
                   TJ_CASE(P_FP8X3, 512))
