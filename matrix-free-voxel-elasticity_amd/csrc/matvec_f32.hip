// FP32 instantiations of the parity product, FP32 or BF16 storage (kernel and launcher in
// matvec.hip).
#define VH_MATVEC_F32 1
#include "matvec.hip"
