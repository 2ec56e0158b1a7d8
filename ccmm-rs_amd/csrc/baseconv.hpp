// Exact RNS base conversion on the int8 matrix cores (gfx950).
//
// Given the mixed-radix digits c_m (m < A) of an integer X = sum_m c_m M_m,
// M_m = prod_{l<m} s_l over the source moduli s_l, compute X mod p_t for
// every target modulus p_t.  With 7-bit chunks c_m = sum_a c_{m,a} 2^(7a):
//
//   X = sum_{m,a} c_{m,a} (2^(7a) M_m)
//     = sum_{m,a} c_{m,a} sum_b chunk_b((2^(7a) M_m) mod p_t) 2^(7b)   (mod p_t)
//     = sum_b E_{t,b} 2^(7b)                                          (mod p_t)
//
//   E[n][(t,b)] = sum_{(m,a)} C[n][(m,a)] * B[(m,a)][(t,b)]
//
// is one int8 x int8 -> int32 GEMM (C: chunks of the digits, B: constant
// chunks), computed exactly by v_mfma_i32_32x32x32_i8; every entry of E is
// below K * 127^2 < 2^21 (K <= 128), so the int32 sums never overflow.  The
// epilogue recombines sum_b E_b 2^(7b) (< 2^57 for a target of <= 42 bits,
// < 2^78 for 63 bits) and reduces it once modulo p_t.
#pragma once
#include "modarith.hpp"

namespace ck {

constexpr int kBcMaxK = 128;     // chunk rows of the GEMM (sum of source chunk counts)
constexpr int kBcTileCols = 32;  // columns per MFMA column tile
constexpr int kBcMaxTiles = 12;  // column tiles (targets packed whole into each)
constexpr int kBcMaxPerTile = 5;  // targets per column tile
constexpr int kBcRows = 256;     // coefficients per workgroup (8 M-tiles of 32)

CK_HD int chunks7(u64 q) {
    int bits = 64 - __builtin_clzll(q - 1);
    return (bits + 6) / 7;
}

// One target inside a column tile.
struct BcTarget {
    u32 limb;  // index of the target in the destination basis
    u32 col;   // first chunk column within the tile
    u32 nch;   // chunk columns (7-bit chunks of p_t)
    u32 pad;
};

struct BcTile {
    u32 n_targets, pad;
    BcTarget t[kBcMaxPerTile];
};

// Constant side of one conversion (source digits -> targets).
struct BcPlan {
    u32 k_rows;   // sum of source chunk counts, padded to a multiple of 32
    u32 n_tiles;  // column tiles
    u32 src_chunks[16];  // chunk count of each source digit
    BcTile tiles[kBcMaxTiles];
    const signed char *bt;  // [n_tiles * 32 columns][k_rows], column-major B^T
};

}  // namespace ck
