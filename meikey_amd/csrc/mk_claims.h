// How the search kernel's claims map onto key ranges: whole tiles of 64 lanes x K keys
// first, then the tail, and the arithmetic of one claim. Plain integer code shared by the
// kernel and the CPU test build, so the mapping is checked on the host for every shape.
//
// Tail balancing: the last tiles (two per wave of the grid for a wave that claims on its
// own) are cut into U units of K / U keys per lane, claimed one by one from a second
// counter, so the waves finish within a fraction of a tile of each other. Units keep at
// least 64 keys per lane (the inversion is shared by a unit's keys); two tiles per wave
// in units measured 0.5% faster than one, four and eight slower again, and 32-key units
// slower than 64. The last unit of each wave's share (one per wave of the grid, the final
// round of claims) is then claimed as two halves of K / (2U) keys: 0.3-0.6% faster
// (three and two A/B pairs on two boxes); halving the last two rounds, or quartering the
// final halves, gained nothing more. A serving step of one tile per wave is claimed
// entirely in units; there they are 128 keys, which balance the step's mix of matchers
// at half the inversions of 64-key units (0.5% more served keys/s; whole tiles serve 5%
// fewer), and they are not halved (two overlapping steps fill each other's tails). A server
// may ask for 64-key units instead, so that a step of one unit per wave is half as long.
#pragma once

#include "mk_common.h"

// Host and device code both plan a launch's claims (the server sizes its steps with them).
#if defined(__HIPCC__)
#define MK_CLAIMS_FN __host__ __device__ __forceinline__
#else
#define MK_CLAIMS_FN static inline
#endif

// Keys per lane of a serving step's claim unit unless the server asks for another size.
#define MK_SERVE_UNIT_KEYS 128

// Units a tile of K keys per lane is cut into: serve_unit keys per lane in a serving step
// (serve_unit != 0; a unit that does not divide K leaves the tile whole), 64 keys per lane
// in the tail of a search launch.
MK_CLAIMS_FN u32 mk_units_per_tile(u32 K, u32 serve_unit) {
    return serve_unit != 0 ? ((K >= serve_unit && K % serve_unit == 0) ? K / serve_unit : 1)
                           : ((K >= 128 && K % 64 == 0) ? K / 64 : 1);
}

// The claim layout of one launch: claims [0, big_tiles * U) are whole tiles (claim / U,
// taken U at a time from the first counter), then `units` tail units of K / U keys, the
// last `split` of them as 2 * split halves of K / (2U) keys.
struct mk_claim_plan {
    u32 K, U;
    u64 total_tiles, big_tiles, units, split;
};

MK_CLAIMS_FN mk_claim_plan mk_plan_claims(u32 K, u32 serve_unit, u64 total_tiles, u64 waves) {
    const bool serving = serve_unit != 0;
    mk_claim_plan p;
    p.K = K;
    p.U = mk_units_per_tile(K, serve_unit);
    p.total_tiles = total_tiles;
    const u64 tail = p.U > 1 ? (total_tiles < 2 * waves ? total_tiles : 2 * waves) : 0;
    p.big_tiles = total_tiles - tail;
    p.units = tail * p.U;
    p.split = (!serving && p.U > 1 && p.units >= 2 * waves) ? waves : 0;
    return p;
}

// The claim number handed out for the t-th draw of the first counter (t < big_tiles: a
// whole tile) or, past it, the q-th draw of the second counter.
MK_CLAIMS_FN u64 mk_claim_number(const mk_claim_plan& p, u64 t, u64 q) {
    return t < p.big_tiles ? t * p.U : p.big_tiles * p.U + q;
}

// Tile, sub-range and keys per lane of a claim: the claim covers keys
// [tile * 64 * K + sub * 64 * Keff, + 64 * Keff) of the launch. False past the last claim.
MK_CLAIMS_FN bool mk_claim_at(const mk_claim_plan& p, u64 claim, u64& tile, u64& sub, u32& Keff) {
    const u32 U = p.U;
    tile = claim / U;
    sub = claim % U;
    Keff = tile >= p.big_tiles ? p.K / U : p.K;
    if (p.split != 0 && claim >= p.big_tiles * U + p.units - p.split) {
        // half h of tail unit u: sub counts halves of the tile
        const u64 h = claim - (p.big_tiles * U + p.units - p.split);
        if (h >= 2 * p.split) return false;
        const u64 u = p.units - p.split + h / 2;
        tile = p.big_tiles + u / U;
        sub = (u % U) * 2 + (h & 1);
        Keff = p.K / (2 * U);
    }
    return tile < p.total_tiles;
}
