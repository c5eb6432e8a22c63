// The Poseidon2 duplex transcript, host side (crates/transcript/src/lib.rs).
//
// Width-3 BN254 Poseidon2 (x^5, 4 + 4 full rounds around 56 partial rounds),
// rate 2, overwrite absorption with the absorbed count added to the capacity,
// challenges popped from the end of the rate, and the typed `tag, len,
// payload` framing. This lets a caller without its own transcript drive the
// sumcheck natively; a caller that keeps the reference's transcript passes its
// own round hook instead.
#pragma once
#include <stdint.h>

#include "field.cuh"

namespace apg {

void poseidon2_permute_host(Fe s[3]);
// The round constants in Montgomery form, in the order the rounds use them:
// 4 x 3 initial full rounds, 56 internal, 4 x 3 terminal full rounds.
constexpr int P2_RC_INTERNAL = 12, P2_RC_TERMINAL = 68, P2_RC_COUNT = 80;
void poseidon2_rc_mont(Fe out[P2_RC_COUNT]);
// form 0 as poseidon2_permute_host, 1 with the portable multiply, 2 with the
// mulx/adcx/adox one; false if the form is unknown or the CPU lacks it
bool poseidon2_permute_form(int form, Fe s[3]);

struct Transcript {
    Fe state[3];
    Fe input[2];
    uint32_t input_len = 0;
    Fe output[2];
    uint32_t output_len = 0;

    Transcript();
    void observe(const Fe &x);
    Fe sample();
    void append_scalars(uint64_t tag, const Fe *xs, size_t n);
    void append_bytes(uint64_t tag, const uint8_t *b, size_t len);
    Fe challenge(uint64_t tag);
    // one typed message of 4k limbs (crates/transcript/src/lib.rs:493-521)
    void append_g1_wire(uint64_t tag, const uint8_t *wire64s, size_t k);

   private:
    void duplex();
};

// Tags (crates/constants/src/lib.rs:579-770) used on this path.
enum : uint64_t {
    TAG_COMMITMENT = 3,
    TAG_SUMCHECK_ROUND = 4,
    TAG_SUMCHECK_CHALLENGE = 5,
    TAG_EVALUATION_CLAIM = 6,
    TAG_WITNESS_DIGEST = 8,
    TAG_SUMCHECK_FINAL_EVALS = 9,
    TAG_MERCURY_BATCH = 17,
};

// canonical little-endian bytes <-> Montgomery (crates/field/src/lib.rs:252-307)
void fe_to_bytes_fr(const Fe &a, uint8_t out[32]);
bool fe_from_bytes_fr(const uint8_t in[32], Fe &out);
void fe_to_bytes_fq(const Fe &a, uint8_t out[32]);
bool fe_from_bytes_fq(const uint8_t in[32], Fe &out);

}  // namespace apg
