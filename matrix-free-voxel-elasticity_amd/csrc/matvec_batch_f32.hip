// FP32 instantiations of the batched parity product (kernel and launcher in matvec.hip).
#define VH_MATVEC_BATCH_F32 1
#include "matvec.hip"
