// The general GKR layer on gfx950: a resolved gate list evaluated by the
// device, for the forward pass of one list and for the layer sumcheck.
#pragma once
#include <vector>

#include "field.cuh"
#include "runtime.h"
#include "sumcheck.h"

namespace apg {

// constraints::GateDef wire tags (crates/constraints/src/lib.rs:310-360)
enum GateShape : uint32_t {
    GS_LINEAR = 0,
    GS_PRODUCT = 1,
    GS_MASK = 2,
    GS_AFFINE_PRODUCT = 3,
    GS_TREE_PRODUCT = 4,
    GS_QUADRATIC = 5,
    GS_TREE_CROSS = 6,
};

// constraints::VirtualKind in wire order (crates/constraints/src/lib.rs:215-250)
enum VirtualKind : uint32_t {
    VK_ROW_INDEX = 0,
    VK_RAM_LIVE = 1,
    VK_RANGE19 = 2,
    VK_RANGE16 = 3,
    VK_XOR8_A = 4,
    VK_XOR8_B = 5,
    VK_XOR8_OUT = 6,
};

// Where an operand is read (crates/gkr-verify/src/lib.rs:290-300): a source
// word is (kind << 24) | index.
enum SourceKind : uint32_t { SRC_LOWER = 0, SRC_UPPER = 1, SRC_VIRTUAL = 2, SRC_CACHED = 3 };
constexpr uint32_t SRC_SHIFT = 24, SRC_MASK = 0xffffffu;

// constants::memory::RAM_LIVE_BIT
constexpr int RAM_LIVE_BIT = 14;

struct GkrGate {
    uint32_t shape, src_begin, src_end, coef_begin, split;
};

// A gate list flattened as ResolvedList::new resolves it
// (crates/gkr-verify/src/lib.rs:330-395): cached entries, producing gates,
// enforcing gates; gate i reads sources[src_begin, src_end) in eval_gate's
// value order; coefficients in GateDef::coefficients order with challenges
// substituted.
struct GkrList {
    std::vector<GkrGate> gates;
    uint32_t ncached = 0, nproducing = 0, nenforcing = 0;
    std::vector<uint32_t> sources;
    std::vector<Fe> coeffs;
    std::vector<uint32_t> virtual_kinds;
    bool halving = false;
};

// Checks the list against ncols layer columns: shapes, operand and
// coefficient counts, sources in range (an upper source only in a halving
// list, a virtual or cached one only in a row-wise list, no cached entry
// reading another). Returns the largest degree of a weighted gate after
// substitution (a column or virtual table 1). Throws ApgError(-2).
int gkr_check_list(const GkrList &l, size_t ncols);

// gkr::produce (crates/gkr/src/lib.rs:300-336) for one list: cols are the
// layer's ncols columns of 2^num_vars rows (device); out receives one column
// per producing gate, 2^num_vars rows (row-wise) or 2^(num_vars - 1)
// (halving: row y reads rows y and y + half).
void gkr_forward_device(Context &ctx, const GkrList &l, const Fe *const *cols, size_t ncols, int num_vars,
                        Fe *const *out);

// gkr::prove_sumcheck (crates/gkr/src/lib.rs:792-899) for the summand
// sum_j weights[j] gate_j (producing then enforcing): lower, and for a
// halving list upper, hold ncols device tables of 2^n entries each (read,
// not modified). rounds_out: n x 4 coefficients; point_out: n; finals_out: the
// tables bound to the point, lower then upper.
void gkr_prove_sumcheck_device(Context &ctx, const GkrList &l, const Fe *weights, Fe *const *lower,
                               Fe *const *upper, size_t ncols, int n, const Fe *eq_point,
                               const RoundFn &round_fn, Fe *rounds_out, Fe *point_out, Fe *finals_out);

}  // namespace apg
