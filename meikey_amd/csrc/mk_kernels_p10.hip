// The search-path kernels for ten 25.5-bit limbs with packed comb-table entries (the
// reference's "10" layout): each entry holds its three fields as canonical 32-byte values,
// 96 bytes instead of 128, unpacked into limbs after the LDS read (mk_point.h).
#define MK_FIELD 10
#define MK_PACKED 1
#define MK_VARIANT _p10
#include "mk_kernels.hip"
