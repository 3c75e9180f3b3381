// BF16 screen instantiations (one term and hi/lo split).
#include "screen_impl.h"
TJ_SCREEN_DISPATCH(screen_bf16, P_BF16, TJ_CASE(P_BF16, 16) TJ_CASE(P_BF16, 32) TJ_CASE(P_BF16, 64)
                   TJ_CASE(P_BF16, 128) TJ_CASE(P_BF16, 256) TJ_CASE(P_BF16, 512))
TJ_SCREEN_DISPATCH(screen_bf16x3, P_BF16X3, TJ_CASE(P_BF16X3, 16) TJ_CASE(P_BF16X3, 32) TJ_CASE(P_BF16X3, 64)
                   TJ_CASE(P_BF16X3, 128) TJ_CASE(P_BF16X3, 256))
