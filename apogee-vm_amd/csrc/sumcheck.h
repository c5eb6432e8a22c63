// Sumcheck round evaluation and fold over multilinear Fr tables on gfx950.
#pragma once
#include <functional>
#include <type_traits>
#include <vector>

#include "field.cuh"
#include "runtime.h"
#include "transcript.h"

namespace apg {

constexpr int SC_MAX_TABLES = 24;
constexpr int SC_MAX_TERMS = 32;
constexpr int SC_MAX_NODES = 6;  // degree <= 5

// One gate term: coef * x[f0] * x[f1] * x[f2]; an index < 0 is absent. The
// reference's Gate (crates/sumcheck/src/lib.rs:88-112) has at most two
// factors; a third factor is the degree-4 extension.
struct Term {
    Fe coef;
    int32_t f[3];
};

struct Gate {
    int nterms = 0;
    Term terms[SC_MAX_TERMS];
    // degree of a round polynomial: eq times the largest term, never below a
    // cubic (SumcheckProof.rounds are always 4 coefficients for the reference).
    int degree() const;
    // largest number of factors in a term: G's degree along one variable
    int x_degree() const;
    bool is_plain_product(int k) const;  // one term, coef 1, factors 0..k-1
};

// Backings of a table (crates/poly/src/lib.rs:46-58): Fr entries (32 B), or
// narrow entries lifted to Fr on read: one bit (u64 limbs, entry i = bit i %
// 64 of limb i / 64), u8, u16, u32.
enum : uint8_t { SC_FR = 0, SC_U1 = 1, SC_U8 = 2, SC_U16 = 3, SC_U32 = 4 };

// Per-round hook: receives the d+1 ascending coefficients, returns the
// challenge. A hook may split off a prefix, the part of its absorption that
// comes before the coefficients (a transcript's message tag and length): the
// round loops call begin() once the round's kernel is queued, so the prefix
// runs on the host while the kernel runs on the device; a call without a
// preceding begin() runs the prefix itself first.
struct RoundFn {
    std::function<Fe(const Fe *coeffs, int ncoeffs)> fn;
    std::function<void(int ncoeffs)> prefix;
    mutable bool begun = false;

    RoundFn() = default;
    template <class F, class = std::enable_if_t<std::is_invocable_r_v<Fe, F &, const Fe *, int>>>
    RoundFn(F f) : fn(std::move(f)) {}
    RoundFn(std::function<Fe(const Fe *, int)> f, std::function<void(int)> p)
        : fn(std::move(f)), prefix(std::move(p)) {}

    void begin(int ncoeffs) const {
        if (prefix && !begun) prefix(ncoeffs);
        begun = true;
    }
    Fe operator()(const Fe *coeffs, int ncoeffs) const {
        begin(ncoeffs);
        begun = false;
        return fn(coeffs, ncoeffs);
    }
};

// gkr::prove_sumcheck (crates/gkr/src/lib.rs:792-899) with a gate summand.
// Tables are device buffers of 2^num_vars elements each; they are consumed
// (ping-pong scratch is internal). rounds_out receives num_vars * (d + 1)
// coefficients, point_out num_vars challenges, finals_out the k fully bound
// values.
// kinds (nullable: all Fr) gives each table's backing; narrow tables are read
// through the lift in the first round and folded into Fr in the second.
void sumcheck_prove_device(Context &ctx, const Gate &gate, Fe *const *tables_dev, int k,
                           int num_vars, const Fe *eq_point_host, const RoundFn &round_fn,
                           Fe *rounds_out, Fe *point_out, Fe *finals_out, const uint8_t *kinds = nullptr);

// The same proof with all device scratch in one file-scope array per device
// (sumcheck_noalloc_limit() bytes), so the call itself allocates no device
// memory. sumcheck_noalloc_init(device) takes the array's address on that
// device (once, with the device current; the library does it when a context
// is created). Sizes whose scratch (sumcheck_noalloc_bytes) exceeds the
// limit are refused. Calls on one device are serialised by a lock.
// The product summand eq * t_0 * ... * t_(k-1) (1 <= k <= 3) over device
// tables, every round and its transcript step on the device, captured as one
// graph and launched once: no host step between rounds. The transcript is
// the library's (it moves to the device and back); the bytes equal
// sumcheck_prove_device's with the same transcript. Tables are read only.
void sumcheck_prove_graph(Context &ctx, Fe *const *tables_dev, int k, int num_vars, const Fe *eq_point_host,
                          Transcript &t, Fe *rounds_out, Fe *point_out, Fe *finals_out);

void sumcheck_prove_device_noalloc(Context &ctx, const Gate &gate, Fe *const *tables_dev, int k,
                                   int num_vars, const Fe *eq_point_host, const RoundFn &round_fn,
                                   Fe *rounds_out, Fe *point_out, Fe *finals_out, const uint8_t *kinds = nullptr);
void sumcheck_noalloc_init(int device);
size_t sumcheck_noalloc_limit();
size_t sumcheck_noalloc_bytes(int k, int num_vars);

// Host interpolation: the d+1 values at nodes 0..d to ascending coefficients
// (crates/sumcheck/src/lib.rs:294-305 for d = 3, Newton form beyond).
void interpolate_host(const Fe *vals, int d, Fe *coeffs);

// out[o][i] = scale[o] * sum_c w[c] * cols[c][i] for o < outs; cols are K
// device columns of n elements, w (K) and scale (outs) host values.
void combine_columns_device(Context &ctx, const Fe *cols_dev, size_t n, int K, const Fe *w_host,
                            const Fe *scale_host, int outs, Fe *const *out_dev);

}  // namespace apg
