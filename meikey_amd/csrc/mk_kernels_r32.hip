// The search-path kernels for eight 32-bit words (the reference's "8" layout,
// src/service/gpu/kernel.cu:650-848): the field of mk_field8.h with packed 96-byte
// comb-table entries.
#define MK_FIELD 8
#define MK_PACKED 1
#define MK_VARIANT _r32
#include "mk_kernels.hip"
