// FP64 instantiations of the batched parity product (kernel and launcher in matvec.hip).
#define VH_MATVEC_BATCH_F64 1
#include "matvec.hip"
