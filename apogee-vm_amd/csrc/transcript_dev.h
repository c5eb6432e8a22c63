// Host entry points of the device transcript (transcript_dev.hip).
#pragma once
#include "runtime.h"
#include "transcript.h"

namespace apg {

// ops of a transcript program, four words each: {op, a, b, 0}
enum : uint32_t {
    TOP_OBSERVE = 0,      // observe operands[a]
    TOP_OBSERVE_U64 = 1,  // observe the field element a + 2^32 b
    TOP_SAMPLE = 2,       // outputs[a] = sample()
};

// the round constants on the context's device (uploaded on first use)
const Fe *poseidon2_rc_device(Context &ctx);
// permutes n states of three Montgomery elements in device memory
void poseidon2_permute_device(Context &ctx, Fe *states_dev, size_t n);
// runs program i (ops[offs[i] .. offs[i + 1]) on transcript i, all on the
// device, one thread each; operands and outputs are host arrays
void transcript_run_device(Context &ctx, Transcript *const *ts, size_t count, const uint32_t *ops, size_t n_ops,
                           const uint32_t *offs, const Fe *operands, size_t n_operands, Fe *outputs,
                           size_t n_outputs);

}  // namespace apg
