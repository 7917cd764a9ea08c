// FP64 instantiations of the parity product (kernel and launcher in matvec.hip).
#define VH_MATVEC_F64 1
#include "matvec.hip"
