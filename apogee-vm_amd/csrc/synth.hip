// Seeded synthetic inputs for the benchmark workloads (host code).
//
// Stream: ChaCha20 (RFC 8439 block function) keyed by the 64-bit seed in key
// words 0-1 (little-endian), nonce 0, block counter from 0, bytes consumed in
// order. Draws:
//   uniform Fr : 32 bytes, top two bits cleared, rejected if >= p (the draw
//                rule of crates/pcs/tests/common/mod.rs:114-123)
//   witness mix: a u32 category c % 10: 0-1 zero, 2 one, 3-7 a u32 value,
//                8-9 a uniform Fr (20% / 10% / 50% / 20%).
#include <cstring>

#include "../../include/apogee_gpu.h"
#include "transcript.h"

namespace {

struct ChaCha {
    uint32_t key[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    uint32_t counter = 0;
    uint8_t buf[64];
    int pos = 64;

    static uint32_t rotl(uint32_t a, int b) { return (a << b) | (a >> (32 - b)); }
    static void qr(uint32_t &a, uint32_t &b, uint32_t &c, uint32_t &d) {
        a += b; d ^= a; d = rotl(d, 16);
        c += d; b ^= c; b = rotl(b, 12);
        a += b; d ^= a; d = rotl(d, 8);
        c += d; b ^= c; b = rotl(b, 7);
    }
    void block() {
        uint32_t s[16] = {0x61707865, 0x3320646e, 0x79622d32, 0x6b206574};
        for (int i = 0; i < 8; i++) s[4 + i] = key[i];
        s[12] = counter++;
        s[13] = s[14] = s[15] = 0;
        uint32_t x[16];
        std::memcpy(x, s, 64);
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
        for (int i = 0; i < 16; i++) {
            uint32_t v = x[i] + s[i];
            for (int k = 0; k < 4; k++) buf[4 * i + k] = (uint8_t)(v >> (8 * k));
        }
        pos = 0;
    }
    uint8_t byte() {
        if (pos == 64) block();
        return buf[pos++];
    }
    uint32_t u32() {
        uint32_t v = 0;
        for (int k = 0; k < 4; k++) v |= (uint32_t)byte() << (8 * k);
        return v;
    }
    apg::Fe fr() {
        for (;;) {
            uint8_t b[32];
            for (int i = 0; i < 32; i++) b[i] = byte();
            b[31] &= 0x3f;
            apg::Fe x;
            if (apg::fe_from_bytes_fr(b, x)) return x;
        }
    }
};

}  // namespace

extern "C" int apg_synthetic_fill(uint64_t seed, int kind, size_t n, uint8_t *out_mont) {
    if (kind != 0 && kind != 1) return APG_ERR_ARG;
    ChaCha c;
    c.key[0] = (uint32_t)seed;
    c.key[1] = (uint32_t)(seed >> 32);
    for (size_t i = 0; i < n; i++) {
        apg::Fe v;
        if (kind == 0) {
            v = c.fr();
        } else {
            const uint32_t cat = c.u32() % 10;
            if (cat < 2)
                v = apg::f_zero<apg::Fr>();
            else if (cat == 2)
                v = apg::f_one<apg::Fr>();
            else if (cat < 8)
                v = apg::f_from_u64<apg::Fr>(c.u32());
            else
                v = c.fr();
        }
        std::memcpy(out_mont + 32 * i, v.l, 32);
    }
    return APG_OK;
}
