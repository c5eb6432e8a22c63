// Device Poseidon2 and transcripts in bulk: permutations of many states, and
// transcript programs, one transcript per thread (transcript_dev.cuh).
#include "transcript_dev.h"

#include <algorithm>
#include <cstddef>
#include <cstring>
#include <vector>

#include "transcript_dev.cuh"

namespace apg {

namespace {

__global__ void k_p2_batch(Fe *__restrict__ states, size_t n, const Fe *__restrict__ rc) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Fe s[3] = {states[3 * i], states[3 * i + 1], states[3 * i + 2]};
    p2_permute(s, rc);
    for (int l = 0; l < 3; l++) states[3 * i + l] = s[l];
}

// Thread i runs ops [offs[i], offs[i + 1]) on transcript i.
__global__ void k_transcript_run(DevTranscript *__restrict__ ts, uint32_t count, const uint4 *__restrict__ ops,
                                 const uint32_t *__restrict__ offs, const Fe *__restrict__ operands,
                                 uint32_t n_operands, Fe *__restrict__ outputs, uint32_t n_outputs,
                                 const Fe *__restrict__ rc) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    DevTranscript t = ts[i];
    for (uint32_t q = offs[i]; q < offs[i + 1]; q++) {
        const uint4 op = ops[q];
        switch (op.x) {
            case TOP_OBSERVE:  // the host checked every index against n_operands
                dt_observe(t, operands[op.y < n_operands ? op.y : 0], rc);
                break;
            case TOP_OBSERVE_U64:
                dt_observe(t, f_from_u64<Fr>((uint64_t)op.y | ((uint64_t)op.z << 32)), rc);
                break;
            default: {  // TOP_SAMPLE
                const Fe v = dt_sample(t, rc);
                if (op.y < n_outputs) outputs[op.y] = v;
                break;
            }
        }
    }
    ts[i] = t;
}

}  // namespace

const Fe *poseidon2_rc_device(Context &ctx) {
    Fe *rc = ctx.d<Fe>("p2.rc", P2_RC_COUNT);
    if (!ctx.p2_rc_ready) {
        Fe h[P2_RC_COUNT];
        poseidon2_rc_mont(h);
        HIP_CHECK(hipMemcpyAsync(rc, h, sizeof(h), hipMemcpyHostToDevice, ctx.stream));
        ctx.sync();  // h is on this stack frame
        ctx.p2_rc_ready = true;
    }
    return rc;
}

void poseidon2_permute_device(Context &ctx, Fe *states_dev, size_t n) {
    if (n == 0) return;
    const Fe *rc = poseidon2_rc_device(ctx);
    hipLaunchKernelGGL(k_p2_batch, dim3(grid_for(n, 64)), dim3(64), 0, ctx.stream, states_dev, n, rc);
    HIP_CHECK(hipGetLastError());
}

void transcript_run_device(Context &ctx, Transcript *const *ts, size_t count, const uint32_t *ops, size_t n_ops,
                           const uint32_t *offs, const Fe *operands, size_t n_operands, Fe *outputs,
                           size_t n_outputs) {
    if (count == 0) return;
    if (count >= (size_t(1) << 31) || n_ops >= (size_t(1) << 31) || n_operands >= (size_t(1) << 31) ||
        n_outputs >= (size_t(1) << 31))
        throw ApgError(-2, "transcript program too large");
    if (offs[0] != 0 || offs[count] != n_ops) throw ApgError(-2, "transcript program offsets do not cover the ops");
    for (size_t i = 0; i < count; i++)
        if (offs[i + 1] < offs[i]) throw ApgError(-2, "transcript program offsets decrease");
    for (size_t q = 0; q < n_ops; q++) {
        const uint32_t *op = ops + 4 * q;
        if (op[0] == TOP_OBSERVE && op[1] >= n_operands) throw ApgError(-2, "transcript program: operand out of range");
        if (op[0] == TOP_SAMPLE && op[1] >= n_outputs) throw ApgError(-2, "transcript program: output out of range");
        if (op[0] != TOP_OBSERVE && op[0] != TOP_OBSERVE_U64 && op[0] != TOP_SAMPLE)
            throw ApgError(-2, "transcript program: unknown op");
    }
    const Fe *rc = poseidon2_rc_device(ctx);
    DevTranscript *t_dev = ctx.d<DevTranscript>("tr.states", count);
    uint4 *op_dev = ctx.d<uint4>("tr.ops", std::max<size_t>(n_ops, 1));
    uint32_t *off_dev = ctx.d<uint32_t>("tr.offs", count + 1);
    Fe *in_dev = ctx.d<Fe>("tr.operands", std::max<size_t>(n_operands, 1));
    Fe *out_dev = ctx.d<Fe>("tr.outputs", std::max<size_t>(n_outputs, 1));
    std::vector<DevTranscript> host(count);
    for (size_t i = 0; i < count; i++) std::memcpy(&host[i], ts[i], sizeof(DevTranscript));
    hipStream_t st = ctx.stream;
    HIP_CHECK(hipMemcpyAsync(t_dev, host.data(), count * sizeof(DevTranscript), hipMemcpyHostToDevice, st));
    if (n_ops) HIP_CHECK(hipMemcpyAsync(op_dev, ops, n_ops * sizeof(uint4), hipMemcpyHostToDevice, st));
    HIP_CHECK(hipMemcpyAsync(off_dev, offs, (count + 1) * sizeof(uint32_t), hipMemcpyHostToDevice, st));
    if (n_operands) HIP_CHECK(hipMemcpyAsync(in_dev, operands, n_operands * sizeof(Fe), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_transcript_run, dim3(grid_for(count, 64)), dim3(64), 0, st, t_dev, (uint32_t)count, op_dev,
                       off_dev, in_dev, (uint32_t)n_operands, out_dev, (uint32_t)n_outputs, rc);
    HIP_CHECK(hipGetLastError());
    if (n_outputs) HIP_CHECK(hipMemcpyAsync(outputs, out_dev, n_outputs * sizeof(Fe), hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipMemcpyAsync(host.data(), t_dev, count * sizeof(DevTranscript), hipMemcpyDeviceToHost, st));
    ctx.sync();
    for (size_t i = 0; i < count; i++) std::memcpy(ts[i], &host[i], sizeof(DevTranscript));
}

}  // namespace apg
