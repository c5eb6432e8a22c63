// The Mercury opening, pcs::open (crates/pcs/src/lib.rs:313-492), with the
// polynomial of the opening resident in HBM.
#pragma once
#include <functional>

#include "runtime.h"
#include "transcript.h"

namespace apg {

// crates/constants/src/lib.rs:606-646
enum : uint64_t {
    TAG_PCS_OPENING = 7,
    TAG_MERCURY_INSTANCE = 10,
    TAG_MERCURY_ALPHA = 11,
    TAG_MERCURY_GAMMA = 12,
    TAG_MERCURY_Z = 13,
    TAG_BDFG_BATCH = 14,
    TAG_BDFG_POINT = 15,
    TAG_PAIRING_MERGE = 16,
};

constexpr size_t MERCURY_PROOF_BYTES = 8 * 64 + 6 * 32;

// [c(x)]_1 = msm(srs[..len], c) in the 64-byte wire form, for device scalars.
using CommitFn = std::function<void(const Fe *dev, size_t len, uint8_t wire[64])>;

// Opens cm at u for the 2^num_vars evaluation table f_dev (read as
// coefficients), advancing tr exactly as the reference does. Writes v and
// the 704-byte proof (8 wire points, then 6 canonical Fr). Throws ApgError on
// a bad argument or a broken identity (APG_ERR_INTERNAL) and on the
// reference's DegenerateChallenge (APG_ERR_DEGENERATE).
void mercury_open(Context &ctx, const CommitFn &commit, const Fe *f_dev, int num_vars,
                  const uint8_t cm_wire[64], const Fe *u, Transcript &tr, Fe &v,
                  uint8_t proof[MERCURY_PROOF_BYTES]);

// batch_open_stacked (crates/pcs/src/lib.rs:701-789): cols_dev holds the
// stacks' columns back to back in slot order, 2^width each; stack s has
// stack_sizes[s] columns (at most 2^nr); one commitment per stack. Writes each
// stack's value at u || r and the proof of cm* = sum rho^i cm_i. Throws
// ApgError(APG_ERR_ARG) for EmptyBatch, BatchLengthMismatch and
// UnsupportedNumVars, as the reference refuses them.
void mercury_batch_open_stacked(Context &ctx, const CommitFn &commit, const Fe *cols_dev,
                                const uint32_t *stack_sizes, size_t nstacks, int width, const uint8_t *cms_wire,
                                const Fe *u, const Fe *r, int nr, Transcript &tr, Fe *vs_out,
                                uint8_t proof[MERCURY_PROOF_BYTES]);

// In place over ncols independent sequences, element j of column i at
// v[j * stride + i]: A_j = v_j + x A_{j+1} from j = count - 1 down to 0
// (A_count = 0). Exposed for tests of the scan.
void horner_scan(Context &ctx, Fe *v, size_t count, size_t stride, size_t ncols, const Fe &x,
                 int level = 0);

}  // namespace apg
