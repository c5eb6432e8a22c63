// Seeded synthetic inputs for the benchmark workloads, on the host or the
// device (both compute the same elements).
//
// Counter-mode ChaCha20 (RFC 8439 block function), so element i of a stream
// is drawn independently of every other element:
//   key     = the 64-bit seed in key words 0-1 (little-endian), zero elsewhere
//   counter = i
//   nonce   = (attempt, domain, 0), domain "fr" (0x7266) or "wi" (0x6977)
// uniform Fr  : each block offers two 32-byte candidates (bytes 0-31, 32-63);
//               a candidate has its top two bits cleared and is accepted if
//               below p (the draw rule of crates/pcs/tests/common/mod.rs:114-123);
//               attempts 0, 1, 2, ... until one is accepted.
// witness mix : attempt-0 block: u32 at bytes 0-3 picks the category c % 10
//               (0-1 zero, 2 one, 3-7 the u32 at bytes 4-7, 8-9 uniform Fr);
//               the Fr candidate order is bytes 32-63 of attempt 0, then both
//               halves of attempts 1, 2, ...
#include <cstring>

#include "../../include/apogee_gpu.h"
#include "runtime.h"
#include "transcript.h"

namespace apg {
namespace {

constexpr uint32_t DOM_FR = 0x7266, DOM_WI = 0x6977;

APG_HD uint32_t rotl32(uint32_t a, int b) { return (a << b) | (a >> (32 - b)); }

APG_HD void qr(uint32_t &a, uint32_t &b, uint32_t &c, uint32_t &d) {
    a += b; d ^= a; d = rotl32(d, 16);
    c += d; b ^= c; b = rotl32(b, 12);
    a += b; d ^= a; d = rotl32(d, 8);
    c += d; b ^= c; b = rotl32(b, 7);
}

APG_HD void chacha_block(uint64_t seed, uint32_t counter, uint32_t n0, uint32_t n1, uint32_t out[16]) {
    uint32_t s[16] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u,
                      (uint32_t)seed, (uint32_t)(seed >> 32), 0, 0, 0, 0, 0, 0,
                      counter, n0, n1, 0};
    uint32_t x[16];
    for (int i = 0; i < 16; i++) x[i] = s[i];
    for (int i = 0; i < 10; i++) {
        qr(x[0], x[4], x[8], x[12]);
        qr(x[1], x[5], x[9], x[13]);
        qr(x[2], x[6], x[10], x[14]);
        qr(x[3], x[7], x[11], x[15]);
        qr(x[0], x[5], x[10], x[15]);
        qr(x[1], x[6], x[11], x[12]);
        qr(x[2], x[7], x[8], x[13]);
        qr(x[3], x[4], x[9], x[14]);
    }
    for (int i = 0; i < 16; i++) out[i] = x[i] + s[i];
}

// candidate from 8 keystream words: top two bits cleared, accepted if < p
APG_HD bool fr_candidate(const uint32_t *w, Fe &out) {
    Fe c;
    for (int i = 0; i < 8; i++) c.l[i] = w[i];
    c.l[7] &= 0x3fffffffu;
    if (!f_is_canonical<Fr>(c)) return false;
    out = f_to_mont<Fr>(c);
    return true;
}

APG_HD Fe draw_fr(uint64_t seed, uint32_t i, uint32_t dom, uint32_t first_attempt) {
    uint32_t blk[16];
    Fe v;
    for (uint32_t a = first_attempt;; a++) {
        chacha_block(seed, i, a, dom, blk);
        if (fr_candidate(blk, v) || fr_candidate(blk + 8, v)) return v;
    }
}

APG_HD Fe draw(uint64_t seed, int kind, uint32_t i) {
    if (kind == 0) return draw_fr(seed, i, DOM_FR, 0);
    uint32_t blk[16];
    chacha_block(seed, i, 0, DOM_WI, blk);
    const uint32_t cat = blk[0] % 10;
    if (cat < 2) return f_zero<Fr>();
    if (cat == 2) return f_one<Fr>();
    if (cat < 8) return f_from_u64<Fr>(blk[1]);
    Fe v;
    if (fr_candidate(blk + 8, v)) return v;
    return draw_fr(seed, i, DOM_WI, 1);
}

__global__ void k_synth(uint64_t seed, int kind, size_t n, Fe *out) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = draw(seed, kind, (uint32_t)i);
}

}  // namespace
}  // namespace apg

using namespace apg;

extern "C" int apg_synthetic_fill(uint64_t seed, int kind, size_t n, uint8_t *out_mont) {
    if ((kind != 0 && kind != 1) || n > 0xffffffffull) return APG_ERR_ARG;
    for (size_t i = 0; i < n; i++) {
        const Fe v = draw(seed, kind, (uint32_t)i);
        std::memcpy(out_mont + 32 * i, v.l, 32);
    }
    return APG_OK;
}

// launched on the caller's stream by apg_synthetic_fill_dev (api.hip)
extern "C" int apg_synthetic_fill_dev_impl(void *stream, uint64_t seed, int kind, size_t n, void *out_dev) {
    if ((kind != 0 && kind != 1) || n > 0xffffffffull) return APG_ERR_ARG;
    if (!n) return APG_OK;
    hipLaunchKernelGGL(k_synth, dim3(grid_for(n, 256)), dim3(256), 0, (hipStream_t)stream, seed, kind,
                       n, static_cast<Fe *>(out_dev));
    return hipGetLastError() == hipSuccess ? APG_OK : APG_ERR_HIP;
}
