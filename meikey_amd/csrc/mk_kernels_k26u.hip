// The search-path kernels for ten uniform 26-bit limbs with Karatsuba products and
// 128-byte comb-table entries of limbs (the reference's "10ku" layout,
// src/service/gpu/kernel.cu:1102-1316).
#define MK_FIELD 26
#define MK_PACKED 0
#define MK_VARIANT _k26u
#include "mk_kernels.hip"
