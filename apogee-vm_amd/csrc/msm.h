// Multi-scalar multiplication over BN254 G1 on gfx950: host-side interface.
#pragma once
#include <memory>

#include "curve.cuh"
#include "runtime.h"

namespace apg {

// A table of affine bases held in HBM, optionally with window-shifted copies:
// copy w holds 2^(c*w) * B_i at points + w * stride + i, w < copies. With every
// window of a scalar precomputed (copies == windows) all windows share one
// bucket set and the MSM has no doubling pass at all. Coordinates are held in
// the bucket phase's domain: Montgomery 32-bit words times 2^5 mod p
// (affine_to_table, curve29.cuh); msm_expand_bases and msm_table_convert
// write them so.
struct BaseTable {
    Affine *points = nullptr;  // device, [copies][stride]
    size_t stride = 0;         // points per copy (the prefix length covered)
    int c = 0;                 // window width the copies are shifted for
    int copies = 1;
    bool owned = false;
    ~BaseTable() {
        if (owned && points) (void)hipFree(points);
    }
};

// Window geometry of one MSM call.
struct MsmShape {
    int c;         // window width
    int windows;   // signed digits per scalar: ceil((bits + 1) / c)
    int P;         // precomputed copies used (<= table copies)
    int G;         // bucket groups per column = ceil(windows / P)
    uint32_t B;    // buckets per group = 2^(c-1)
};

// Windows needed for a bits-bit scalar at width c (crates/curve/src/msm.rs:140-142).
constexpr int msm_window_count(int bits, int c) { return (bits + 1 + c - 1) / c; }

// Window width for n points under this library's size classes (by
// floor(log2 n)); msm_set_window_bits overrides one class (c = 0 restores the
// default rule), which is how the autotuner tries widths.
// c <= 17 keeps B / 256 <= 256 row sums per problem, the most k_weighted
// takes in one block.
constexpr int MSM_MIN_C = 3, MSM_MAX_C = 17;
int msm_pick_c(size_t n, size_t columns);
void msm_set_window_bits(int log2n, int c);

// Accumulation slices per resident k_accumulate thread (the bucket tile: the
// slice length is the pair total over rounds x resident threads, at least
// 16 and an eighth of the mean bucket), per size class like the window width;
// msm_set_slice_rounds overrides one class (0 restores the built-in table).
constexpr int MSM_MIN_ROUNDS = 1, MSM_MAX_ROUNDS = 16;
int msm_pick_rounds(size_t n);
void msm_set_slice_rounds(int log2n, int rounds);

// Builds copies = windows(254, c) shifted copies of the first n bases (device
// pointer `src`, affine Montgomery) into a new table.
// msm_table_convert puts n affine Montgomery points (canonical coordinates) in
// place into the table form, for a one-copy table over caller bases.
void msm_table_convert(Context &ctx, Affine *points, size_t n);
std::unique_ptr<BaseTable> msm_expand_bases(Context &ctx, const Affine *src, size_t n, int c,
                                            int copies);

// K column MSMs over the same table prefix of n points. Scalars are either
// Montgomery Fr (fr != nullptr, K*n elements, column-major) or u32 (u32 !=
// nullptr). Results are written as affine Montgomery points to out_host[K].
void msm_run(Context &ctx, const BaseTable &tab, size_t n, size_t K, const Fe *fr_dev,
             const uint32_t *u32_dev, Affine *out_host);

// Host helpers shared by the tails.
Xyzz host_xyzz_mul_pow2(Xyzz p, int k);
void host_batch_to_affine(const Xyzz *in, Affine *out, size_t k);

}  // namespace apg
