// The search-path kernels for ten uniform 26-bit limbs with Karatsuba products and packed
// 96-byte comb-table entries (the reference's "10k" layout, src/service/gpu/kernel.cu:1102-1316).
#define MK_FIELD 26
#define MK_PACKED 1
#define MK_VARIANT _k26
#include "mk_kernels.hip"
