// The Poseidon2 duplex transcript on the device (crates/transcript/src/lib.rs:
// 121-146 the permutation, 241-346 the sponge and typed layer), one
// transcript per thread. Its state has the host Transcript's layout
// (transcript.h), so a transcript moves between host and device as bytes.
//
// A permutation is a chain of about 240 dependent Fr multiplies (the x^5
// S-boxes), so one thread runs it far slower than a host core does; the
// device form exists for work that has to stay on the device (a sumcheck
// whose rounds run as one captured graph) and for many transcripts at once,
// one per thread.
#pragma once
#include <cstddef>

#include "field.cuh"
#include "transcript.h"

namespace apg {

struct DevTranscript {
    Fe state[3];
    Fe input[2];
    uint32_t input_len;
    Fe output[2];
    uint32_t output_len;
};
static_assert(sizeof(DevTranscript) == sizeof(Transcript), "device transcript layout");
static_assert(offsetof(DevTranscript, input_len) == offsetof(Transcript, input_len), "device transcript layout");
static_assert(offsetof(DevTranscript, output) == offsetof(Transcript, output), "device transcript layout");
static_assert(offsetof(DevTranscript, output_len) == offsetof(Transcript, output_len), "device transcript layout");

__device__ __forceinline__ Fe p2_sbox(const Fe &x) {
    const Fe x2 = f_mul<Fr>(x, x);
    const Fe x4 = f_mul<Fr>(x2, x2);
    return f_mul<Fr>(x4, x);
}

// external matrix circ(2, 1, 1): s_i + (s_0 + s_1 + s_2)
__device__ __forceinline__ void p2_external(Fe s[3]) {
    const Fe sum = f_add<Fr>(f_add<Fr>(s[0], s[1]), s[2]);
    for (int i = 0; i < 3; i++) s[i] = f_add<Fr>(s[i], sum);
}

// rc: P2_RC_COUNT Montgomery constants (poseidon2_rc_mont's order)
__device__ inline void p2_permute(Fe s[3], const Fe *__restrict__ rc) {
    p2_external(s);
    for (int r = 0; r < 4; r++) {
        for (int l = 0; l < 3; l++) s[l] = p2_sbox(f_add<Fr>(s[l], rc[3 * r + l]));
        p2_external(s);
    }
    for (int i = 0; i < 56; i++) {
        // internal matrix J + diag(1, 1, 2)
        s[0] = p2_sbox(f_add<Fr>(s[0], rc[P2_RC_INTERNAL + i]));
        const Fe sum = f_add<Fr>(f_add<Fr>(s[0], s[1]), s[2]);
        s[0] = f_add<Fr>(s[0], sum);
        s[1] = f_add<Fr>(s[1], sum);
        s[2] = f_add<Fr>(f_add<Fr>(s[2], s[2]), sum);
    }
    for (int r = 0; r < 4; r++) {
        for (int l = 0; l < 3; l++) s[l] = p2_sbox(f_add<Fr>(s[l], rc[P2_RC_TERMINAL + 3 * r + l]));
        p2_external(s);
    }
}

// lib.rs:273-292: overwrite the rate with the pending input (unfilled lanes
// zero), add the absorbed count to the capacity, permute, refill the output
__device__ inline void dt_duplex(DevTranscript &t, const Fe *__restrict__ rc) {
    const uint32_t n = t.input_len;
    for (uint32_t i = 0; i < n; i++) {
        t.state[i] = t.input[i];
        t.input[i] = f_zero<Fr>();
    }
    t.input_len = 0;
    if (n > 0) {
        for (uint32_t l = n; l < 2; l++) t.state[l] = f_zero<Fr>();
        t.state[2] = f_add<Fr>(t.state[2], f_from_u64<Fr>(n));
    }
    p2_permute(t.state, rc);
    t.output[0] = t.state[0];
    t.output[1] = t.state[1];
    t.output_len = 2;
}

// lib.rs:241-250
__device__ inline void dt_observe(DevTranscript &t, const Fe &x, const Fe *__restrict__ rc) {
    t.output[0] = t.output[1] = f_zero<Fr>();
    t.output_len = 0;
    t.input[t.input_len++] = x;
    if (t.input_len == 2) dt_duplex(t, rc);
}

// lib.rs:253-262
__device__ inline Fe dt_sample(DevTranscript &t, const Fe *__restrict__ rc) {
    if (t.input_len > 0 || t.output_len == 0) dt_duplex(t, rc);
    t.output_len--;
    const Fe out = t.output[t.output_len];
    t.output[t.output_len] = f_zero<Fr>();
    return out;
}

// lib.rs:306-316, 341-346
__device__ inline void dt_append_scalars(DevTranscript &t, uint64_t tag, const Fe *xs, int n,
                                         const Fe *__restrict__ rc) {
    dt_observe(t, f_from_u64<Fr>(tag), rc);
    dt_observe(t, f_from_u64<Fr>((uint64_t)n), rc);
    for (int i = 0; i < n; i++) dt_observe(t, xs[i], rc);
}

__device__ inline Fe dt_challenge(DevTranscript &t, uint64_t tag, const Fe *__restrict__ rc) {
    dt_observe(t, f_from_u64<Fr>(tag), rc);
    return dt_sample(t, rc);
}

}  // namespace apg
