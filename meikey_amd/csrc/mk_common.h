// Shared definitions for the gfx950 search kernels.
//
// The arithmetic headers (mk_field.h, mk_hash.h, mk_point.h, mk_match.h) are plain
// integer C++ so that the same source compiles for the GPU (hipcc, gfx950) and for
// the host (g++), where the CPU test build checks every function against the oracle
// without a GPU. Only the kernels and the launch code are HIP-specific.
#pragma once

#include <stdint.h>

typedef uint8_t u8;
typedef uint32_t u32;
typedef int32_t i32;
typedef uint64_t u64;
typedef uint16_t u16;

#if defined(__HIP_DEVICE_COMPILE__)
#define MK_CONST __constant__ static const
#define MK_DEV __device__ __forceinline__
#define MK_MEMBER __device__ __forceinline__
#define MK_NOINLINE __device__ __noinline__
#define MK_UNROLL _Pragma("unroll")
#define MK_NO_UNROLL _Pragma("unroll 1")
#elif defined(__HIPCC__)
#define MK_CONST __constant__ static const
#define MK_DEV __host__ __device__ __forceinline__
#define MK_MEMBER __host__ __device__ __forceinline__
#define MK_NOINLINE __host__ __device__ __noinline__
#define MK_UNROLL _Pragma("unroll")
#define MK_NO_UNROLL _Pragma("unroll 1")
#else
#define MK_CONST static const
#define MK_DEV static inline
#define MK_MEMBER inline
#define MK_NOINLINE static
#define MK_UNROLL
#define MK_NO_UNROLL
#endif

// Comb geometry: the clamped scalar is s = 8 * (2^251 + m) with a 251-bit m, split
// into rows = ceil(251 / w) signed base-2^w digits (src/service/derive/comb.rs:1-8,
// 285-287 of the reference).
#define MK_M_BITS 251
#define MK_MIN_W 2
#define MK_MAX_W 28
#define MK_MAX_ROWS ((MK_M_BITS + MK_MIN_W - 1) / MK_MIN_W)

MK_DEV u32 mk_rows(u32 w) { return (MK_M_BITS + w - 1) / w; }
MK_DEV u32 mk_cols(u32 w) { return (1u << (w - 1)) + 1; }

// Field representation and comb-table layout of one build of the kernels (the reference's
// --gpu-limbs codes, src/service/gpu.rs:70-106). Each is compiled into its own set of
// kernels; mk_options.limbs picks the set at run time.
//   MK_FIELD 10: ten 25.5-bit limbs (mk_field.h); 8: eight 32-bit words (mk_field8.h);
//            26: ten 26-bit limbs with Karatsuba products (mk_field26.h)
//   MK_PACKED 0: a table entry holds its three fields as limbs: 30 words + 2 pad words,
//                one 128-byte line ("10u", "10ku");
//             1: a table entry holds them as canonical little-endian 256-bit values,
//                3 x 8 words = 96 bytes ("10", "8", "10k")
#ifndef MK_FIELD
#define MK_FIELD 10
#endif
#ifndef MK_PACKED
#define MK_PACKED 0
#endif
#define MK_FE_LIMBS (MK_FIELD == 8 ? 8 : 10)
// The mk_options.limbs code of this build (include/meikey_hip.h).
#define MK_LIMBS_CODE                                                                                    \
    (MK_FIELD == 8 ? 8u : MK_FIELD == 26 ? (MK_PACKED ? 26u : 27u) : (MK_PACKED ? 10u : 11u))
// One comb-table entry: the affine point as ((y+x)/2, (y-x)/2, dxy) (row 0: xy).
#define MK_ENTRY_WORDS (MK_PACKED ? 24 : 32)

// The 44 base64 characters that depend on the key (blob bytes 18..50), and the
// first of them that a pattern may match (reference src/service/gpu/kernel.cu:112-119).
#define MK_TAIL_B64 44
#define MK_SEARCH_FROM 1
#define MK_MAX_NEEDLE 42
// The tail character whose value carries the x sign in its bit 1 (kernel.cu:1730-1731).
#define MK_SIGN_CHAR 42

#ifndef MK_MODE_CONTAINS
#define MK_MODE_CONTAINS 0
#define MK_MODE_SUFFIX 1
#define MK_MODE_REGEX 2
#endif

#define MK_DFA_ACCEPT 0u
#define MK_DFA_START 1u
