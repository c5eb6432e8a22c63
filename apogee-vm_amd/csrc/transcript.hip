// Host Poseidon2 duplex transcript: crates/transcript/src/lib.rs:121-346.
#include "transcript.h"

#include <cstring>
#include <mutex>

#include "host_fr.h"
#include "poseidon2_rc3.h"

namespace apg {

namespace {

using namespace hfr;

Fe from_hex_be(const char *s) {
    uint8_t le[32];
    auto hv = [](char c) -> int { return c <= '9' ? c - '0' : c - 'a' + 10; };
    for (int i = 0; i < 32; i++) le[31 - i] = (uint8_t)((hv(s[2 * i]) << 4) | hv(s[2 * i + 1]));
    Fe r;
    fe_from_bytes_fr(le, r);
    return r;
}

struct RoundConstants {
    W4 ini[4][3], mid[56], ter[4][3];
    RoundConstants() {
        for (int r = 0; r < 4; r++)
            for (int l = 0; l < 3; l++) {
                ini[r][l] = w4_of(from_hex_be(APG_RC3_INITIAL[r][l]));
                ter[r][l] = w4_of(from_hex_be(APG_RC3_TERMINAL[r][l]));
            }
        for (int i = 0; i < 56; i++) mid[i] = w4_of(from_hex_be(APG_RC3_INTERNAL[i]));
    }
};

const RoundConstants &rc() {
    static RoundConstants k;
    return k;
}

template <W4 (*MUL)(const W4 &, const W4 &)>
inline W4 sbox(const W4 &x) {
    const W4 x2 = MUL(x, x);
    const W4 x4 = MUL(x2, x2);
    return MUL(x4, x);
}

inline void ext_matrix(W4 s[3]) {
    const W4 sum = w4_add_lazy(w4_add_lazy(s[0], s[1]), s[2]);
    for (int i = 0; i < 3; i++) s[i] = w4_add_lazy(s[i], sum);
}

// s12 = s[1] + s[2], formed before s[0]'s S-box so that only two additions
// follow the S-box on the chain from one partial round to the next
inline void int_matrix(W4 s[3], const W4 &s12) {
    const W4 sum = w4_add_lazy(s[0], s12);
    s[0] = w4_add_lazy(s[0], sum);
    s[1] = w4_add_lazy(s[1], sum);
    s[2] = w4_add_lazy(w4_add_lazy(s[2], s[2]), sum);
}

void to_bytes(const Fe &a, uint8_t out[32], bool fq) {
    Fe c = fq ? f_from_mont<Fq>(a) : f_from_mont<Fr>(a);
    for (int i = 0; i < 8; i++)
        for (int k = 0; k < 4; k++) out[4 * i + k] = (uint8_t)(c.l[i] >> (8 * k));
}

template <class P>
bool from_bytes(const uint8_t in[32], Fe &out) {
    Fe c;
    for (int i = 0; i < 8; i++) {
        c.l[i] = 0;
        for (int k = 0; k < 4; k++) c.l[i] |= (uint32_t)in[4 * i + k] << (8 * k);
    }
    if (!f_is_canonical<P>(c)) return false;
    out = f_to_mont<P>(c);
    return true;
}

}  // namespace

void fe_to_bytes_fr(const Fe &a, uint8_t out[32]) { to_bytes(a, out, false); }
void fe_to_bytes_fq(const Fe &a, uint8_t out[32]) { to_bytes(a, out, true); }
bool fe_from_bytes_fr(const uint8_t in[32], Fe &out) { return from_bytes<Fr>(in, out); }
bool fe_from_bytes_fq(const uint8_t in[32], Fe &out) { return from_bytes<Fq>(in, out); }

namespace {

template <W4 (*MUL)(const W4 &, const W4 &)>
void permute(Fe st[3]) {
    const RoundConstants &k = rc();
    W4 s[3] = {w4_of(st[0]), w4_of(st[1]), w4_of(st[2])};
    ext_matrix(s);
    for (int r = 0; r < 4; r++) {
        for (int l = 0; l < 3; l++) s[l] = sbox<MUL>(w4_add_lazy(s[l], k.ini[r][l]));
        ext_matrix(s);
    }
    for (int i = 0; i < 56; i++) {
        const W4 s12 = w4_add_lazy(s[1], s[2]);
        s[0] = sbox<MUL>(w4_add_lazy(s[0], k.mid[i]));
        int_matrix(s, s12);
    }
    for (int r = 0; r < 4; r++) {
        for (int l = 0; l < 3; l++) s[l] = sbox<MUL>(w4_add_lazy(s[l], k.ter[r][l]));
        ext_matrix(s);
    }
    // below 2p: one subtraction of p makes each word canonical
    for (int l = 0; l < 3; l++) st[l] = fe_of(w4_reduce(s[l].v[0], s[l].v[1], s[l].v[2], s[l].v[3]));
}

}  // namespace

// The permutation is a chain of 240 dependent multiplies, so the multiply's
// latency sets its speed: the mulx/adcx/adox form where the CPU has it.
void poseidon2_permute_host(Fe st[3]) {
#ifdef APG_HOST_ADX
    static const bool adx = w4_have_adx();
    if (adx) return permute<w4_mul_lazy_adx>(st);
#endif
    permute<w4_mul_lazy>(st);
}

void poseidon2_rc_mont(Fe out[P2_RC_COUNT]) {
    const RoundConstants &k = rc();
    for (int r = 0; r < 4; r++)
        for (int l = 0; l < 3; l++) {
            out[3 * r + l] = fe_of(k.ini[r][l]);
            out[P2_RC_TERMINAL + 3 * r + l] = fe_of(k.ter[r][l]);
        }
    for (int i = 0; i < 56; i++) out[P2_RC_INTERNAL + i] = fe_of(k.mid[i]);
}

bool poseidon2_permute_form(int form, Fe st[3]) {
    switch (form) {
        case 0: poseidon2_permute_host(st); return true;
        case 1: permute<w4_mul_lazy>(st); return true;
#ifdef APG_HOST_ADX
        case 2:
            if (!w4_have_adx()) return false;
            permute<w4_mul_lazy_adx>(st);
            return true;
#endif
        default: return false;
    }
}

Transcript::Transcript() {
    for (int i = 0; i < 3; i++) state[i] = f_zero<Fr>();
    for (int i = 0; i < 2; i++) input[i] = output[i] = f_zero<Fr>();
}

// lib.rs:273-292
void Transcript::duplex() {
    const uint32_t n = input_len;
    for (uint32_t i = 0; i < n; i++) {
        state[i] = input[i];
        input[i] = f_zero<Fr>();
    }
    input_len = 0;
    if (n > 0) {
        for (uint32_t l = n; l < 2; l++) state[l] = f_zero<Fr>();
        state[2] = f_add<Fr>(state[2], f_from_u64<Fr>(n));
    }
    poseidon2_permute_host(state);
    output[0] = state[0];
    output[1] = state[1];
    output_len = 2;
}

// lib.rs:241-250
void Transcript::observe(const Fe &x) {
    output[0] = output[1] = f_zero<Fr>();
    output_len = 0;
    input[input_len++] = x;
    if (input_len == 2) duplex();
}

// lib.rs:253-262
Fe Transcript::sample() {
    if (input_len > 0 || output_len == 0) duplex();
    output_len--;
    Fe out = output[output_len];
    output[output_len] = f_zero<Fr>();
    return out;
}

// lib.rs:306-316
void Transcript::append_scalars(uint64_t tag, const Fe *xs, size_t n) {
    observe(f_from_u64<Fr>(tag));
    observe(f_from_u64<Fr>(n));
    for (size_t i = 0; i < n; i++) observe(xs[i]);
}

// lib.rs:323-338
void Transcript::append_bytes(uint64_t tag, const uint8_t *b, size_t len) {
    observe(f_from_u64<Fr>(tag));
    observe(f_from_u64<Fr>(len));
    for (size_t off = 0; off < len; off += 31) {
        uint8_t limb[32];
        std::memset(limb, 0, 32);
        size_t c = len - off < 31 ? len - off : 31;
        std::memcpy(limb, b + off, c);
        Fe x;
        fe_from_bytes_fr(limb, x);
        observe(x);
    }
}

// lib.rs:341-346
Fe Transcript::challenge(uint64_t tag) {
    observe(f_from_u64<Fr>(tag));
    return sample();
}

// lib.rs:493-521: all-zero = infinity absorbs 2^128 in every lane; otherwise
// x[0..16], x[16..32], y[0..16], y[16..32] as little-endian integers.
void Transcript::append_g1_wire(uint64_t tag, const uint8_t *p, size_t k) {
    observe(f_from_u64<Fr>(tag));
    observe(f_from_u64<Fr>(4 * k));
    for (size_t i = 0; i < k; i++) {
        const uint8_t *w = p + 64 * i;
        bool zero = true;
        for (int q = 0; q < 64; q++) zero &= w[q] == 0;
        for (int q = 0; q < 4; q++) {
            uint8_t b[32];
            std::memset(b, 0, 32);
            if (zero)
                b[16] = 1;
            else
                std::memcpy(b, w + 16 * q, 16);
            Fe x;
            fe_from_bytes_fr(b, x);
            observe(x);
        }
    }
}

}  // namespace apg
