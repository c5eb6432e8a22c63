// The C ABI (include/apogee_gpu.h): contexts, SRS residency, MSM/commit,
// sumcheck, shard composition, transcript and encodings. Every entry catches
// library exceptions and returns a negative code.
#include <chrono>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/apogee_gpu.h"
#include "engine.h"
#include "gkr.h"
#include "srs_io.h"
#include "mercury.h"
#include "curve29.cuh"
#include "msm.h"
#include "sumcheck.h"
#include "transcript.h"
#include "transcript_dev.h"

using namespace apg;

struct apg_ctx {
    Context c;
};

struct apg_srs {
    int device = 0;
    Affine *base = nullptr;  // copy 0: [tau^i]G, i < n
    size_t n = 0;
    mutable std::mutex mu;
    // window width -> shifted copies. Shared: an MSM holds its table until it
    // returns, so a table outgrown by a longer prefix of the same width (and
    // replaced here) is freed only when its last user is done.
    std::map<int, std::shared_ptr<BaseTable>> tables;
    ~apg_srs() {
        tables.clear();
        if (base) (void)hipFree(base);
    }
};

struct apg_transcript {
    Transcript t;
};

namespace apg {
static thread_local std::string g_last_error;
void set_last_error(const std::string &msg) { g_last_error = msg; }
}  // namespace apg

template <class F>
static int guarded(F &&f) {
    try {
        f();
        return APG_OK;
    } catch (const ApgError &e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::bad_alloc &) {
        set_last_error("host out of memory");
        return APG_ERR_OOM;
    } catch (const std::exception &e) {
        set_last_error(e.what());
        return APG_ERR_INTERNAL;
    }
}

// A null handle is an argument error, never a dereference.
static inline void need(const void *handle) {
    if (!handle) throw ApgError(APG_ERR_ARG, "null handle");
}

static inline Fe load_fe(const uint8_t *p) {
    Fe r;
    std::memcpy(r.l, p, 32);
    return r;
}
static inline void store_fe(uint8_t *p, const Fe &v) { std::memcpy(p, v.l, 32); }

namespace {

// Grid-stride 16-byte copy: each byte read once and written once.
__global__ void __launch_bounds__(256) k_copy16(uint4 *__restrict__ dst, const uint4 *__restrict__ src, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

__global__ void k_srs_gen(Fe tau, size_t n, Affine *out) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Fe p = f_one<Fr>(), b = tau;
    for (size_t e = i; e; e >>= 1) {
        if (e & 1) p = f_mul<Fr>(p, b);
        b = f_sqr<Fr>(b);
    }
    const Fe s = f_from_mont<Fr>(p);
    Affine g;
    g.x = f_one<Fq>();
    g.y = f_from_u64<Fq>(2);
    Xyzz acc = xyzz_identity();
    for (int bit = 253; bit >= 0; bit--) {
        acc = xyzz_dbl(acc);
        if ((s.l[bit >> 5] >> (bit & 31)) & 1) acc = xyzz_add_affine(acc, g);
    }
    out[i] = xyzz_to_affine(acc);
}

std::shared_ptr<BaseTable> srs_table(Context &ctx, apg_srs *srs, size_t n) {
    if (n > srs->n) throw ApgError(APG_ERR_ARG, "SRS too small for this commit");
    const int c = msm_pick_c(n, 1);
    std::lock_guard<std::mutex> lk(srs->mu);
    auto it = srs->tables.find(c);
    if (it == srs->tables.end() || it->second->stride < n) {
        const int copies = msm_window_count(254, c);
        std::shared_ptr<BaseTable> t = msm_expand_bases(ctx, srs->base, n, c, copies);
        ctx.sync();
        srs->tables[c] = t;
        return t;
    }
    return it->second;
}

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

void commit_device(Context &ctx, apg_srs *srs, const Fe *fr_dev, const uint32_t *u32_dev, size_t n,
                   size_t k, uint8_t *out_mont) {
    std::vector<Affine> res(k);
    if (n == 0) {
        std::memset(out_mont, 0, 64 * k);
        return;
    }
    need(srs);
    const std::shared_ptr<BaseTable> tab = srs_table(ctx, srs, n);
    const double t0 = now_ms();
    msm_run(ctx, *tab, n, k, fr_dev, u32_dev, res.data());
    ctx.stats.msm_ms += now_ms() - t0;
    ctx.stats.msm_calls += 1;
    ctx.stats.msm_points += n * k;
    for (size_t i = 0; i < k; i++) std::memcpy(out_mont + 64 * i, &res[i], 64);
}

Gate make_gate(const apg_term *terms, size_t n_terms) {
    if (n_terms > (size_t)SC_MAX_TERMS) throw ApgError(APG_ERR_ARG, "too many gate terms");
    Gate g;
    g.nterms = (int)n_terms;
    for (size_t q = 0; q < n_terms; q++) {
        g.terms[q].coef = load_fe(terms[q].coef_mont);
        for (int f = 0; f < 3; f++) g.terms[q].f[f] = terms[q].factor[f];
    }
    return g;
}

struct CbError {};

// append_scalars(SUMCHECK_ROUND, coeffs) then challenge(SUMCHECK_CHALLENGE)
// (crates/sumcheck/src/lib.rs:392-395), with the message's tag and length
// split off as the prefix the round loops run while a round's kernel runs.
RoundFn transcript_round(Transcript &t) {
    return RoundFn(
        [&t](const Fe *coeffs, int nc) -> Fe {
            for (int i = 0; i < nc; i++) t.observe(coeffs[i]);
            return t.challenge(TAG_SUMCHECK_CHALLENGE);
        },
        [&t](int nc) {
            t.observe(f_from_u64<Fr>(TAG_SUMCHECK_ROUND));
            t.observe(f_from_u64<Fr>((uint64_t)nc));
        });
}

RoundFn wrap_round(apg_round_fn fn, void *user) {
    // the library's own transcript as the callback: call it natively
    if (fn == &apg_transcript_round && user) return transcript_round(static_cast<apg_transcript *>(user)->t);
    return [fn, user](const Fe *coeffs, int nc) -> Fe {
        std::vector<uint8_t> buf(32 * (size_t)nc);
        for (int i = 0; i < nc; i++) store_fe(&buf[32 * (size_t)i], coeffs[i]);
        uint8_t ch[32];
        if (fn(user, buf.data(), (size_t)nc, ch) != 0)
            throw ApgError(APG_ERR_CALLBACK, "round callback aborted the proof");
        return load_fe(ch);
    };
}

void sumcheck_common(Context &ctx, const Gate &gate, Fe *const *tabs, size_t k, size_t num_vars,
                     const uint8_t *eq_point_mont, const RoundFn &fn, uint8_t *rounds_out,
                     uint8_t *point_out, uint8_t *final_evals_out, bool noalloc = false,
                     const uint8_t *kinds = nullptr) {
    const int d = gate.degree();
    std::vector<Fe> eqp(num_vars + 1), rounds((size_t)num_vars * (d + 1) + 1), point(num_vars + 1),
        finals(k);
    for (size_t i = 0; i < num_vars; i++) eqp[i] = load_fe(eq_point_mont + 32 * i);
    const double t0 = now_ms();
    if (noalloc)
        sumcheck_prove_device_noalloc(ctx, gate, tabs, (int)k, (int)num_vars, eqp.data(), fn, rounds.data(),
                                      point.data(), finals.data());
    else
        sumcheck_prove_device(ctx, gate, tabs, (int)k, (int)num_vars, eqp.data(), fn, rounds.data(),
                              point.data(), finals.data(), kinds);
    ctx.stats.sumcheck_ms += now_ms() - t0;
    for (size_t i = 0; i < (size_t)num_vars * (d + 1); i++) store_fe(rounds_out + 32 * i, rounds[i]);
    if (point_out)
        for (size_t i = 0; i < num_vars; i++) store_fe(point_out + 32 * i, point[i]);
    for (size_t i = 0; i < k; i++) store_fe(final_evals_out + 32 * i, finals[i]);
}

// Upload host tables into the context arena; returns device pointers.
std::vector<Fe *> upload_tables(Context &ctx, const uint8_t *const *tables, size_t k, size_t len) {
    Fe *pool = ctx.d<Fe>("sc.inputs", k * len);
    std::vector<Fe *> p(k);
    for (size_t t = 0; t < k; t++) {
        p[t] = pool + t * len;
        HIP_CHECK(hipMemcpyAsync(p[t], tables[t], len * 32, hipMemcpyHostToDevice, ctx.stream));
    }
    return p;
}

// A shard's summand is eq times degree - 1 combined tables, one gate term of
// at most three factors, so degree 3 or 4.
void check_shard_args(size_t num_vars, size_t k, int degree) {
    if (num_vars > 30 || num_vars < 1) throw ApgError(APG_ERR_ARG, "num_vars out of range");
    if (degree < 3 || degree > 4) throw ApgError(APG_ERR_ARG, "degree must be 3 or 4");
    if (k < 1 || k > (size_t)SC_MAX_TABLES) throw ApgError(APG_ERR_ARG, "column count out of range");
}

// HBM a shard holds while it runs: the MSM's digit pairs and their sorted
// copies (16 B per digit slot, in column groups of at most 2^28 slots), chunk
// and bucket sums, and the sumcheck's combined tables, eq and fold buffers.
size_t shard_hbm_estimate(size_t num_vars, size_t k, int degree) {
    const size_t n = size_t(1) << num_vars;
    const int c = msm_pick_c(n, k);
    const size_t windows = (size_t)msm_window_count(254, c);
    const size_t slots = std::min(n * k * windows, size_t(1) << 28);
    const size_t msm = slots * 16 + slots / 4 * 128 + k * (size_t(1) << (c - 1)) * 144;
    const size_t f = (size_t)degree - 1;
    const size_t sc = f * n * 32 + (f + 1) * n * 32 * 2;
    return msm + sc + (size_t(64) << 20);
}

// A shard's time alone on an MI355X, for urgency preemption: a fixed part,
// ~45 us per sumcheck round and ~0.45 ns per scalar, fitted to the plain
// path's times over the mixed batch (within ~25% from 2^8 x 13 to 2^22 x 24;
// the scheduler needs only their order and rough scale).
double shard_time_estimate_ms(size_t num_vars, size_t k) {
    return 0.5 + 0.045 * (double)num_vars + 4.5e-7 * std::ldexp((double)k, (int)num_vars);
}

// The shard composition: commit the k columns, absorb the commitments as one
// COMMITMENT message (pcs::append_g1_list, crates/pcs/src/lib.rs:220), draw
// rho (MERCURY_BATCH), fold the columns into degree - 1 tables with weights
// rho^(o k + c), draw the eq point (num_vars SUMCHECK_CHALLENGE), prove the
// rounds of eq * t_0 * ... * t_(degree-2), absorb the final evaluations.
void prove_shard_impl(Context &c, apg_srs *srs, const Fe *cols, size_t num_vars, size_t k, int degree,
                      Transcript &t, uint8_t *commitments_wire, uint8_t *rounds_out,
                      uint8_t *final_evals_out) {
    const size_t n = size_t(1) << num_vars;
    std::vector<uint8_t> cm_mont(64 * k);
    // the phases' shares of the shard's time (progress for urgency
    // preemption): the commit MSM ~85%, the combination ~3%, the rounds
    {
        ProgressSpan span(c, 0.0, 0.85);
        commit_device(c, srs, cols, nullptr, n, k, cm_mont.data());
    }
    for (size_t i = 0; i < k; i++) apg_g1_to_wire(&cm_mont[64 * i], commitments_wire + 64 * i);
    t.append_g1_wire(TAG_COMMITMENT, commitments_wire, k);
    const Fe rho = t.challenge(TAG_MERCURY_BATCH);
    // table o has weights rho^(o k + c): rho^(o k) times sum_c rho^c col_c
    const int f = degree - 1;
    std::vector<Fe> w(k), scale(f);
    Fe pw = f_one<Fr>();
    for (size_t col = 0; col < k; col++) {
        w[col] = pw;
        pw = f_mul<Fr>(pw, rho);
    }
    scale[0] = f_one<Fr>();
    for (int o = 1; o < f; o++) scale[o] = f_mul<Fr>(scale[o - 1], pw);
    Fe *comb = c.d<Fe>("shard.comb", (size_t)f * n);
    std::vector<Fe *> outs(f);
    for (int o = 0; o < f; o++) outs[o] = comb + (size_t)o * n;
    combine_columns_device(c, cols, n, (int)k, w.data(), scale.data(), f, outs.data());
    c.yield_point(0.88);
    ProgressSpan span(c, 0.88, 1.0);
    std::vector<uint8_t> r(32 * num_vars);
    for (size_t i = 0; i < num_vars; i++) store_fe(&r[32 * i], t.challenge(TAG_SUMCHECK_CHALLENGE));
    Gate g;
    g.nterms = 1;
    g.terms[0].coef = f_one<Fr>();
    for (int q = 0; q < 3; q++) g.terms[0].f[q] = q < f ? q : -1;
    const RoundFn fn = transcript_round(t);
    sumcheck_common(c, g, outs.data(), (size_t)f, num_vars, r.data(), fn, rounds_out, nullptr,
                    final_evals_out);
    std::vector<Fe> fin(f);
    for (int i = 0; i < f; i++) fin[i] = load_fe(final_evals_out + 32 * i);
    t.append_scalars(TAG_SUMCHECK_FINAL_EVALS, fin.data(), f);
}

}  // namespace

struct apg_engine {
    std::unique_ptr<Engine> e;
};

extern "C" {

const char *apg_last_error(void) { return g_last_error.c_str(); }
int apg_version(void) { return 1; }

int apg_ctx_create(int device, apg_ctx **out) {
    return guarded([&] {
        int ndev = 0;
        HIP_CHECK(hipGetDeviceCount(&ndev));
        if (device < 0 || device >= ndev) throw ApgError(APG_ERR_ARG, "no such device");
        HIP_CHECK(hipSetDevice(device));
        auto c = std::make_unique<apg_ctx>();
        context_init(c->c, device);
        sumcheck_noalloc_init(device);  // the device's no-allocation scratch address
        *out = c.release();
    });
}

void apg_ctx_destroy(apg_ctx *ctx) {
    if (!ctx) return;
    context_release(ctx->c);
    delete ctx;
}

int apg_ctx_synchronize(apg_ctx *ctx) {
    return guarded([&] {
        need(ctx);
        ctx->c.sync();
    });
}

int apg_ctx_set_profiling(apg_ctx *ctx, int on) {
    return guarded([&] {
        need(ctx);
        std::lock_guard<std::mutex> lk(ctx->c.mu);
        ctx->c.profiling = on != 0;
    });
}

static void export_stats(const Stats &s, apg_stats *out) {
    out->accumulate_ms = s.accumulate_ms;
    out->accumulate_launches = s.accumulate_launches;
    out->accumulate_entries = s.accumulate_entries;
    out->msm_ms = s.msm_ms;
    out->msm_calls = s.msm_calls;
    out->msm_points = s.msm_points;
    out->sumcheck_kernel_ms = s.sumcheck_kernel_ms;
    out->sumcheck_launches = s.sumcheck_launches;
    out->sumcheck_bytes = s.sumcheck_bytes;
    out->sumcheck_ms = s.sumcheck_ms;
}

int apg_ctx_stats(apg_ctx *ctx, apg_stats *out) {
    return guarded([&] {
        need(ctx);
        std::lock_guard<std::mutex> lk(ctx->c.mu);
        export_stats(ctx->c.stats, out);
    });
}

int apg_ctx_reset_stats(apg_ctx *ctx) {
    return guarded([&] {
        need(ctx);
        std::lock_guard<std::mutex> lk(ctx->c.mu);
        ctx->c.stats = Stats();
    });
}

extern "C" int apg_synthetic_fill_dev_impl(void *stream, uint64_t seed, int kind, size_t n, void *out_dev);

int apg_synthetic_fill_dev(apg_ctx *ctx, uint64_t seed, int kind, size_t n, void *out_dev) {
    return guarded([&] {
        need(ctx);
        std::lock_guard<std::mutex> lk(ctx->c.mu);
        HIP_CHECK(hipSetDevice(ctx->c.device));
        const int rc = apg_synthetic_fill_dev_impl(ctx->c.stream, seed, kind, n, out_dev);
        if (rc != APG_OK) throw ApgError(rc, "synthetic fill failed");
        ctx->c.sync();
    });
}

int apg_dev_alloc(apg_ctx *ctx, size_t bytes, void **out) {
    return guarded([&] {
        need(ctx);
        HIP_CHECK(hipSetDevice(ctx->c.device));
        if (hipMalloc(out, bytes ? bytes : 1) != hipSuccess) throw ApgError(APG_ERR_OOM, "device out of memory");
    });
}
int apg_dev_free(apg_ctx *ctx, void *p) {
    return guarded([&] {
        need(ctx);
        HIP_CHECK(hipSetDevice(ctx->c.device));
        HIP_CHECK(hipFree(p));
    });
}
int apg_dev_upload(apg_ctx *ctx, void *dst, const void *src, size_t bytes) {
    return guarded([&] {
        need(ctx);
        std::lock_guard<std::mutex> lk(ctx->c.mu);
        HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->c.stream));
        ctx->c.sync();
    });
}
int apg_dev_copy(apg_ctx *ctx, void *dst, const void *src, size_t bytes) {
    return guarded([&] {
        need(ctx);
        if (bytes % 16) throw ApgError(APG_ERR_ARG, "apg_dev_copy: bytes must be a multiple of 16");
        if (!bytes) return;
        std::lock_guard<std::mutex> lk(ctx->c.mu);
        HIP_CHECK(hipSetDevice(ctx->c.device));
        const size_t n = bytes / 16;
        const unsigned grid = (unsigned)std::min<size_t>(grid_for(n, 256), 8u * (unsigned)ctx->c.cu_count);
        hipLaunchKernelGGL(k_copy16, dim3(grid), dim3(256), 0, ctx->c.stream, static_cast<uint4 *>(dst),
                           static_cast<const uint4 *>(src), n);
        HIP_CHECK(hipGetLastError());
        ctx->c.sync();
    });
}
int apg_dev_download(apg_ctx *ctx, void *dst, const void *src, size_t bytes) {
    return guarded([&] {
        need(ctx);
        std::lock_guard<std::mutex> lk(ctx->c.mu);
        HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->c.stream));
        ctx->c.sync();
    });
}

int apg_srs_upload(apg_ctx *ctx, const uint8_t *bases_mont, size_t n, apg_srs **out) {
    return guarded([&] {
        need(ctx);
        std::lock_guard<std::mutex> lk(ctx->c.mu);
        HIP_CHECK(hipSetDevice(ctx->c.device));
        auto s = std::make_unique<apg_srs>();
        s->device = ctx->c.device;
        s->n = n;
        if (hipMalloc(&s->base, std::max<size_t>(n, 1) * sizeof(Affine)) != hipSuccess)
            throw ApgError(APG_ERR_OOM, "device out of memory for the SRS");
        if (n) HIP_CHECK(hipMemcpyAsync(s->base, bases_mont, n * 64, hipMemcpyHostToDevice, ctx->c.stream));
        ctx->c.sync();
        *out = s.release();
    });
}

static const char *srs_error_text(int e) {
    switch (e) {
        case SRS_IO: return "srs: i/o error";
        case SRS_TRUNCATED: return "srs: truncated archive";
        case SRS_BAD_MAGIC: return "srs: bad magic";
        case SRS_BAD_VERSION: return "srs: bad version";
        case SRS_BAD_SECTION: return "srs: archive G1 count is not 2^power";
        default: return "srs: invalid point";
    }
}

int apg_srs_load(apg_ctx *ctx, const char *path, apg_srs **out, uint8_t g2_out[256], int *srs_error,
                 uint64_t *detail) {
    return guarded([&] {
        need(ctx);
        if (!path || !out) throw ApgError(APG_ERR_ARG, "srs load: null argument");
        std::lock_guard<std::mutex> lk(ctx->c.mu);
        HIP_CHECK(hipSetDevice(ctx->c.device));
        Affine *base = nullptr;
        size_t count = 0;
        uint8_t g2[256];
        uint64_t det = 0;
        const int e = srs_load_archive(ctx->c, path, &base, &count, g2, &det);
        if (srs_error) *srs_error = e;
        if (detail) *detail = det;
        if (e != SRS_OK) throw ApgError(APG_ERR_ARG, srs_error_text(e));
        auto s = std::make_unique<apg_srs>();
        s->device = ctx->c.device;
        s->n = count;
        s->base = base;
        if (g2_out) std::memcpy(g2_out, g2, 256);
        *out = s.release();
    });
}

int apg_srs_archive_check(const char *path, uint32_t *power, uint8_t g2_out[256], int *srs_error,
                          uint64_t *detail) {
    return guarded([&] {
        if (!path) throw ApgError(APG_ERR_ARG, "srs check: null path");
        uint64_t det = 0;
        uint32_t pw = 0;
        uint8_t g2[256];
        const int e = srs_check_archive_header(path, &pw, g2, &det);
        if (srs_error) *srs_error = e;
        if (detail) *detail = det;
        if (e != SRS_OK) throw ApgError(APG_ERR_ARG, srs_error_text(e));
        if (power) *power = pw;
        if (g2_out) std::memcpy(g2_out, g2, 256);
    });
}

int apg_srs_save(apg_ctx *ctx, const apg_srs *srs, const uint8_t g2[256], const char *path) {
    return guarded([&] {
        need(ctx);
        if (!srs || !g2 || !path) throw ApgError(APG_ERR_ARG, "srs save: null argument");
        std::lock_guard<std::mutex> lk(ctx->c.mu);
        HIP_CHECK(hipSetDevice(ctx->c.device));
        srs_save_archive(ctx->c, srs->base, srs->n, g2, path);
    });
}

int apg_srs_upload_wire(apg_ctx *ctx, const uint8_t *wire, size_t n, apg_srs **out, uint64_t *bad_index) {
    return guarded([&] {
        need(ctx);
        if (!out || (n && !wire)) throw ApgError(APG_ERR_ARG, "srs upload: null argument");
        std::lock_guard<std::mutex> lk(ctx->c.mu);
        HIP_CHECK(hipSetDevice(ctx->c.device));
        auto s = std::make_unique<apg_srs>();
        s->device = ctx->c.device;
        s->n = n;
        if (hipMalloc(&s->base, std::max<size_t>(n, 1) * sizeof(Affine)) != hipSuccess)
            throw ApgError(APG_ERR_OOM, "device out of memory for the SRS");
        if (n) HIP_CHECK(hipMemcpyAsync(s->base, wire, n * 64, hipMemcpyHostToDevice, ctx->c.stream));
        const uint64_t bad = g1_decode_wire_device(ctx->c, s->base, n);
        if (bad != UINT64_MAX) {
            if (bad_index) *bad_index = bad;
            throw ApgError(APG_ERR_ARG, "srs upload: invalid point");
        }
        *out = s.release();
    });
}

int apg_g2_wire_valid(const uint8_t wire[128]) {
    if (!wire) return 0;
    try {
        return g2_wire_valid(wire) ? 1 : 0;
    } catch (...) {
        return 0;
    }
}

int apg_g1_batch_to_affine(apg_ctx *ctx, const uint8_t *jac_mont, size_t n, uint8_t *out_affine) {
    return guarded([&] {
        need(ctx);
        if (n && (!jac_mont || !out_affine)) throw ApgError(APG_ERR_ARG, "batch_to_affine: null argument");
        if (!n) return;
        std::lock_guard<std::mutex> lk(ctx->c.mu);
        HIP_CHECK(hipSetDevice(ctx->c.device));
        Fe *jac = ctx->c.d<Fe>("b2a.jac", 3 * n);
        Affine *aff = ctx->c.d<Affine>("b2a.aff", n);
        HIP_CHECK(hipMemcpyAsync(jac, jac_mont, n * 96, hipMemcpyHostToDevice, ctx->c.stream));
        g1_jac_to_affine_device(ctx->c, jac, n, aff);
        HIP_CHECK(hipMemcpyAsync(out_affine, aff, n * 64, hipMemcpyDeviceToHost, ctx->c.stream));
        ctx->c.sync();
    });
}

int apg_srs_generate(apg_ctx *ctx, const uint8_t tau_mont[32], size_t n, apg_srs **out) {
    return guarded([&] {
        need(ctx);
        std::lock_guard<std::mutex> lk(ctx->c.mu);
        HIP_CHECK(hipSetDevice(ctx->c.device));
        auto s = std::make_unique<apg_srs>();
        s->device = ctx->c.device;
        s->n = n;
        if (hipMalloc(&s->base, std::max<size_t>(n, 1) * sizeof(Affine)) != hipSuccess)
            throw ApgError(APG_ERR_OOM, "device out of memory for the SRS");
        if (n) {
            hipLaunchKernelGGL(k_srs_gen, dim3(grid_for(n, 128)), dim3(128), 0, ctx->c.stream,
                               load_fe(tau_mont), n, s->base);
            HIP_CHECK(hipGetLastError());
        }
        ctx->c.sync();
        *out = s.release();
    });
}

int apg_srs_download(apg_ctx *ctx, const apg_srs *srs, size_t offset, size_t n, uint8_t *out) {
    return guarded([&] {
        need(ctx);
        if (offset + n > srs->n) throw ApgError(APG_ERR_ARG, "SRS range out of bounds");
        std::lock_guard<std::mutex> lk(ctx->c.mu);
        HIP_CHECK(hipMemcpyAsync(out, srs->base + offset, n * 64, hipMemcpyDeviceToHost, ctx->c.stream));
        ctx->c.sync();
    });
}

size_t apg_srs_len(const apg_srs *srs) { return srs ? srs->n : 0; }

size_t apg_srs_device_bytes(const apg_srs *srs) {
    if (!srs) return 0;
    std::lock_guard<std::mutex> lk(srs->mu);
    size_t b = srs->n * 64;
    for (auto &kv : srs->tables) b += kv.second->stride * kv.second->copies * 64;
    return b;
}

int apg_msm_set_window_bits(int log2n, int c) {
    return guarded([&] { msm_set_window_bits(log2n, c); });
}

int apg_msm_window_bits(size_t n) { return msm_pick_c(n, 1); }

int apg_msm_set_slice_rounds(int log2n, int rounds) {
    return guarded([&] { msm_set_slice_rounds(log2n, rounds); });
}

int apg_msm_slice_rounds(size_t n) { return msm_pick_rounds(n); }

int apg_srs_prepare(apg_ctx *ctx, apg_srs *srs, size_t n) {
    return guarded([&] {
        need(ctx);
        std::lock_guard<std::mutex> lk(ctx->c.mu);
        HIP_CHECK(hipSetDevice(ctx->c.device));
        if (n) srs_table(ctx->c, srs, n);
    });
}

void apg_srs_release(apg_srs *srs) { delete srs; }

int apg_msm_g1_fr(apg_ctx *ctx, const uint8_t *bases_mont, const uint8_t *scalars_mont, size_t n,
                  uint8_t out_mont[64]) {
    return guarded([&] {
        need(ctx);
        std::lock_guard<std::mutex> lk(ctx->c.mu);
        HIP_CHECK(hipSetDevice(ctx->c.device));
        if (n == 0) {
            std::memset(out_mont, 0, 64);
            return;
        }
        Context &c = ctx->c;
        BaseTable tab;
        tab.points = c.d<Affine>("api.bases", n);
        tab.stride = n;
        tab.c = msm_pick_c(n, 1);
        tab.copies = 1;
        Fe *sc = c.d<Fe>("api.scalars", n);
        HIP_CHECK(hipMemcpyAsync(tab.points, bases_mont, n * 64, hipMemcpyHostToDevice, c.stream));
        msm_table_convert(c, tab.points, n);
        HIP_CHECK(hipMemcpyAsync(sc, scalars_mont, n * 32, hipMemcpyHostToDevice, c.stream));
        Affine res;
        msm_run(c, tab, n, 1, sc, nullptr, &res);
        std::memcpy(out_mont, &res, 64);
    });
}

static void pcs_open_common(Context &c, apg_srs *srs, const Fe *f_dev, size_t num_vars, const uint8_t *cm_wire,
                            const uint8_t *u_mont, apg_transcript *tr, uint8_t *v_mont, uint8_t *proof) {
    if (num_vars < 2 || num_vars > 30 || (num_vars & 1))
        throw ApgError(APG_ERR_ARG, "pcs open: num_vars must be even and in [2, 30]");
    if (srs->n < (size_t(1) << num_vars)) throw ApgError(APG_ERR_ARG, "pcs open: SRS too small");
    std::vector<Fe> u(num_vars);
    for (size_t i = 0; i < num_vars; i++) u[i] = load_fe(u_mont + 32 * i);
    CommitFn commit = [&](const Fe *dev, size_t len, uint8_t wire[64]) {
        uint8_t m[64];
        commit_device(c, srs, dev, nullptr, len, 1, m);
        if (apg_g1_to_wire(m, wire) != APG_OK) throw ApgError(APG_ERR_INTERNAL, "pcs open: bad point");
    };
    Fe v;
    mercury_open(c, commit, f_dev, (int)num_vars, cm_wire, u.data(), tr->t, v, proof);
    store_fe(v_mont, v);
}

int apg_pcs_open(apg_ctx *ctx, apg_srs *srs, const uint8_t *table_mont, size_t num_vars, const uint8_t cm_wire[64],
                 const uint8_t *u_mont, apg_transcript *tr, uint8_t v_mont[32], uint8_t proof[704]) {
    return guarded([&] {
        need(ctx);
        need(srs);
        need(tr);
        Context &c = ctx->c;
        std::lock_guard<std::mutex> lk(c.mu);
        HIP_CHECK(hipSetDevice(c.device));
        if (num_vars > 30) throw ApgError(APG_ERR_ARG, "pcs open: num_vars out of range");
        const size_t n = size_t(1) << num_vars;
        Fe *f = c.d<Fe>("pcs.table", n);
        HIP_CHECK(hipMemcpyAsync(f, table_mont, n * 32, hipMemcpyHostToDevice, c.stream));
        pcs_open_common(c, srs, f, num_vars, cm_wire, u_mont, tr, v_mont, proof);
    });
}

int apg_pcs_open_dev(apg_ctx *ctx, apg_srs *srs, const void *table_dev, size_t num_vars, const uint8_t cm_wire[64],
                     const uint8_t *u_mont, apg_transcript *tr, uint8_t v_mont[32], uint8_t proof[704]) {
    return guarded([&] {
        need(ctx);
        need(srs);
        need(tr);
        need(table_dev);
        Context &c = ctx->c;
        std::lock_guard<std::mutex> lk(c.mu);
        HIP_CHECK(hipSetDevice(c.device));
        pcs_open_common(c, srs, static_cast<const Fe *>(table_dev), num_vars, cm_wire, u_mont, tr, v_mont, proof);
    });
}

static void batch_open_common(Context &c, apg_srs *srs, const Fe *cols_dev, const uint32_t *stack_sizes,
                              size_t nstacks, size_t width, const uint8_t *cms_wire, const uint8_t *u_mont,
                              const uint8_t *r_mont, size_t nr, apg_transcript *tr, uint8_t *vs_mont,
                              uint8_t *proof) {
    if (width > 30 || nr > 30 || width + nr > 30) throw ApgError(APG_ERR_ARG, "UnsupportedNumVars");
    if (srs->n < (size_t(1) << (width + nr))) throw ApgError(APG_ERR_ARG, "SrsTooSmall");
    std::vector<Fe> u(width), r(nr), vs(nstacks);
    for (size_t i = 0; i < width; i++) u[i] = load_fe(u_mont + 32 * i);
    for (size_t i = 0; i < nr; i++) r[i] = load_fe(r_mont + 32 * i);
    CommitFn commit = [&](const Fe *dev, size_t len, uint8_t wire[64]) {
        uint8_t m[64];
        commit_device(c, srs, dev, nullptr, len, 1, m);
        if (apg_g1_to_wire(m, wire) != APG_OK) throw ApgError(APG_ERR_INTERNAL, "pcs open: bad point");
    };
    mercury_batch_open_stacked(c, commit, cols_dev, stack_sizes, nstacks, (int)width, cms_wire, u.data(), r.data(),
                               (int)nr, tr->t, vs.data(), proof);
    for (size_t s = 0; s < nstacks; s++) store_fe(vs_mont + 32 * s, vs[s]);
}

int apg_pcs_batch_open_stacked(apg_ctx *ctx, apg_srs *srs, const uint8_t *cols_mont, const uint32_t *stack_sizes,
                               size_t nstacks, size_t width, const uint8_t *cms_wire, const uint8_t *u_mont,
                               const uint8_t *r_mont, size_t nr, apg_transcript *tr, uint8_t *vs_mont,
                               uint8_t proof[704]) {
    return guarded([&] {
        need(ctx);
        need(srs);
        need(tr);
        if (nstacks == 0) throw ApgError(APG_ERR_ARG, "EmptyBatch");
        need(stack_sizes);
        if (width > 30) throw ApgError(APG_ERR_ARG, "UnsupportedNumVars");
        size_t ncols = 0;
        for (size_t s = 0; s < nstacks; s++) ncols += stack_sizes[s];
        Context &c = ctx->c;
        std::lock_guard<std::mutex> lk(c.mu);
        HIP_CHECK(hipSetDevice(c.device));
        const size_t bytes = ncols * (size_t(32) << width);
        Fe *cols = c.d<Fe>("pcs.cols", std::max<size_t>(ncols << width, 1));
        if (bytes) HIP_CHECK(hipMemcpyAsync(cols, cols_mont, bytes, hipMemcpyHostToDevice, c.stream));
        batch_open_common(c, srs, cols, stack_sizes, nstacks, width, cms_wire, u_mont, r_mont, nr, tr, vs_mont, proof);
    });
}

int apg_pcs_batch_open_stacked_dev(apg_ctx *ctx, apg_srs *srs, const void *cols_dev, const uint32_t *stack_sizes,
                                   size_t nstacks, size_t width, const uint8_t *cms_wire, const uint8_t *u_mont,
                                   const uint8_t *r_mont, size_t nr, apg_transcript *tr, uint8_t *vs_mont,
                                   uint8_t proof[704]) {
    return guarded([&] {
        need(ctx);
        need(srs);
        need(tr);
        if (nstacks == 0) throw ApgError(APG_ERR_ARG, "EmptyBatch");
        need(stack_sizes);
        need(cols_dev);
        Context &c = ctx->c;
        std::lock_guard<std::mutex> lk(c.mu);
        HIP_CHECK(hipSetDevice(c.device));
        batch_open_common(c, srs, static_cast<const Fe *>(cols_dev), stack_sizes, nstacks, width, cms_wire, u_mont,
                          r_mont, nr, tr, vs_mont, proof);
    });
}

int apg_msm_g1_u32(apg_ctx *ctx, const uint8_t *bases_mont, const uint32_t *scalars, size_t n,
                   uint8_t out_mont[64]) {
    return guarded([&] {
        need(ctx);
        std::lock_guard<std::mutex> lk(ctx->c.mu);
        HIP_CHECK(hipSetDevice(ctx->c.device));
        if (n == 0) {
            std::memset(out_mont, 0, 64);
            return;
        }
        Context &c = ctx->c;
        BaseTable tab;
        tab.points = c.d<Affine>("api.bases", n);
        tab.stride = n;
        tab.c = msm_pick_c(n, 1);
        tab.copies = 1;
        uint32_t *sc = c.d<uint32_t>("api.u32", n);
        HIP_CHECK(hipMemcpyAsync(tab.points, bases_mont, n * 64, hipMemcpyHostToDevice, c.stream));
        msm_table_convert(c, tab.points, n);
        HIP_CHECK(hipMemcpyAsync(sc, scalars, n * 4, hipMemcpyHostToDevice, c.stream));
        Affine res;
        msm_run(c, tab, n, 1, nullptr, sc, &res);
        std::memcpy(out_mont, &res, 64);
    });
}

int apg_commit_fr(apg_ctx *ctx, apg_srs *srs, const uint8_t *scalars_mont, size_t n, size_t k,
                  uint8_t *out_mont) {
    return guarded([&] {
        need(ctx);
        std::lock_guard<std::mutex> lk(ctx->c.mu);
        HIP_CHECK(hipSetDevice(ctx->c.device));
        Context &c = ctx->c;
        Fe *sc = c.d<Fe>("api.scalars", std::max<size_t>(n * k, 1));
        if (n * k) HIP_CHECK(hipMemcpyAsync(sc, scalars_mont, n * k * 32, hipMemcpyHostToDevice, c.stream));
        commit_device(c, srs, sc, nullptr, n, k, out_mont);
    });
}

int apg_commit_u32(apg_ctx *ctx, apg_srs *srs, const uint32_t *scalars, size_t n, size_t k,
                   uint8_t *out_mont) {
    return guarded([&] {
        need(ctx);
        std::lock_guard<std::mutex> lk(ctx->c.mu);
        HIP_CHECK(hipSetDevice(ctx->c.device));
        Context &c = ctx->c;
        uint32_t *sc = c.d<uint32_t>("api.u32", std::max<size_t>(n * k, 1));
        if (n * k) HIP_CHECK(hipMemcpyAsync(sc, scalars, n * k * 4, hipMemcpyHostToDevice, c.stream));
        commit_device(c, srs, nullptr, sc, n, k, out_mont);
    });
}

int apg_commit_fr_dev(apg_ctx *ctx, apg_srs *srs, const void *scalars_dev, size_t n, size_t k,
                      uint8_t *out_mont) {
    return guarded([&] {
        need(ctx);
        std::lock_guard<std::mutex> lk(ctx->c.mu);
        HIP_CHECK(hipSetDevice(ctx->c.device));
        commit_device(ctx->c, srs, static_cast<const Fe *>(scalars_dev), nullptr, n, k, out_mont);
    });
}

int apg_sumcheck_prove(apg_ctx *ctx, const apg_term *terms, size_t n_terms,
                       const uint8_t *const *tables_mont, size_t k, size_t num_vars,
                       const uint8_t *eq_point_mont, apg_round_fn round_fn, void *user,
                       uint8_t *rounds_out, uint8_t *point_out, uint8_t *final_evals_out) {
    return guarded([&] {
        need(ctx);
        if (num_vars > 40) throw ApgError(APG_ERR_ARG, "num_vars out of range");
        std::lock_guard<std::mutex> lk(ctx->c.mu);
        HIP_CHECK(hipSetDevice(ctx->c.device));
        Gate g = make_gate(terms, n_terms);
        auto tabs = upload_tables(ctx->c, tables_mont, k, size_t(1) << num_vars);
        sumcheck_common(ctx->c, g, tabs.data(), k, num_vars, eq_point_mont, wrap_round(round_fn, user),
                        rounds_out, point_out, final_evals_out);
    });
}

int apg_sumcheck_prove_dev(apg_ctx *ctx, const apg_term *terms, size_t n_terms,
                           void *const *tables_dev, size_t k, size_t num_vars,
                           const uint8_t *eq_point_mont, apg_round_fn round_fn, void *user,
                           uint8_t *rounds_out, uint8_t *point_out, uint8_t *final_evals_out) {
    return guarded([&] {
        need(ctx);
        if (num_vars > 40) throw ApgError(APG_ERR_ARG, "num_vars out of range");
        std::lock_guard<std::mutex> lk(ctx->c.mu);
        HIP_CHECK(hipSetDevice(ctx->c.device));
        Gate g = make_gate(terms, n_terms);
        std::vector<Fe *> tabs(k);
        for (size_t t = 0; t < k; t++) tabs[t] = static_cast<Fe *>(tables_dev[t]);
        sumcheck_common(ctx->c, g, tabs.data(), k, num_vars, eq_point_mont, wrap_round(round_fn, user),
                        rounds_out, point_out, final_evals_out);
    });
}

int apg_sumcheck_prove_graph_dev(apg_ctx *ctx, void *const *tables_dev, size_t k, size_t num_vars,
                                 const uint8_t *eq_point_mont, apg_transcript *t, uint8_t *rounds_out,
                                 uint8_t *point_out, uint8_t *final_evals_out) {
    return guarded([&] {
        need(ctx);
        need(t);
        need(tables_dev);
        if (k < 1 || k > 3) throw ApgError(APG_ERR_ARG, "graph sumcheck: k must be 1, 2 or 3");
        if (num_vars < 1 || num_vars > 30) throw ApgError(APG_ERR_ARG, "graph sumcheck: num_vars out of range");
        need(eq_point_mont);
        need(rounds_out);
        need(point_out);
        need(final_evals_out);
        const size_t nc = std::max<size_t>(4, k + 2);  // coefficients per round
        std::vector<Fe> eqp(num_vars), rounds(num_vars * nc), point(num_vars), finals(k);
        for (size_t i = 0; i < num_vars; i++) eqp[i] = load_fe(eq_point_mont + 32 * i);
        std::vector<Fe *> tabs(k);
        for (size_t i = 0; i < k; i++) {
            need(tables_dev[i]);
            tabs[i] = static_cast<Fe *>(tables_dev[i]);
        }
        std::lock_guard<std::mutex> lk(ctx->c.mu);
        HIP_CHECK(hipSetDevice(ctx->c.device));
        const double t0 = now_ms();
        sumcheck_prove_graph(ctx->c, tabs.data(), (int)k, (int)num_vars, eqp.data(), t->t, rounds.data(),
                             point.data(), finals.data());
        ctx->c.stats.sumcheck_ms += now_ms() - t0;
        for (size_t i = 0; i < rounds.size(); i++) store_fe(rounds_out + 32 * i, rounds[i]);
        for (size_t i = 0; i < num_vars; i++) store_fe(point_out + 32 * i, point[i]);
        for (size_t i = 0; i < k; i++) store_fe(final_evals_out + 32 * i, finals[i]);
    });
}

int apg_sumcheck_prove_dev_lifted(apg_ctx *ctx, const apg_term *terms, size_t n_terms, void *const *tables_dev,
                                  const int *backings, size_t k, size_t num_vars, const uint8_t *eq_point_mont,
                                  apg_round_fn round_fn, void *user, uint8_t *rounds_out, uint8_t *point_out,
                                  uint8_t *final_evals_out) {
    return guarded([&] {
        need(ctx);
        if (num_vars > 40) throw ApgError(APG_ERR_ARG, "num_vars out of range");
        if (k < 1 || k > (size_t)SC_MAX_TABLES) throw ApgError(APG_ERR_ARG, "sumcheck: table count out of range");
        if (!tables_dev || !backings) throw ApgError(APG_ERR_ARG, "sumcheck: no tables or backings");
        uint8_t kinds[SC_MAX_TABLES];
        for (size_t t = 0; t < k; t++) {
            if (backings[t] < APG_BACKING_FR || backings[t] > APG_BACKING_U32)
                throw ApgError(APG_ERR_ARG, "sumcheck: unknown table backing");
            kinds[t] = (uint8_t)backings[t];
        }
        static_assert(APG_BACKING_FR == SC_FR && APG_BACKING_U1 == SC_U1 && APG_BACKING_U8 == SC_U8 &&
                          APG_BACKING_U16 == SC_U16 && APG_BACKING_U32 == SC_U32,
                      "backing codes");
        std::lock_guard<std::mutex> lk(ctx->c.mu);
        HIP_CHECK(hipSetDevice(ctx->c.device));
        Gate g = make_gate(terms, n_terms);
        Fe *tabs[SC_MAX_TABLES];
        for (size_t t = 0; t < k; t++) tabs[t] = static_cast<Fe *>(tables_dev[t]);
        sumcheck_common(ctx->c, g, tabs, k, num_vars, eq_point_mont, wrap_round(round_fn, user), rounds_out,
                        point_out, final_evals_out, false, kinds);
    });
}

int apg_sumcheck_prove_dev_noalloc(apg_ctx *ctx, const apg_term *terms, size_t n_terms,
                                   void *const *tables_dev, size_t k, size_t num_vars,
                                   const uint8_t *eq_point_mont, apg_round_fn round_fn, void *user,
                                   uint8_t *rounds_out, uint8_t *point_out, uint8_t *final_evals_out) {
    return guarded([&] {
        need(ctx);
        if (num_vars > 40) throw ApgError(APG_ERR_ARG, "num_vars out of range");
        if (k < 1 || k > (size_t)SC_MAX_TABLES) throw ApgError(APG_ERR_ARG, "sumcheck: table count out of range");
        if (!tables_dev) throw ApgError(APG_ERR_ARG, "sumcheck: no tables");
        std::lock_guard<std::mutex> lk(ctx->c.mu);
        HIP_CHECK(hipSetDevice(ctx->c.device));
        Gate g = make_gate(terms, n_terms);
        Fe *tabs[SC_MAX_TABLES];
        for (size_t t = 0; t < k; t++) tabs[t] = static_cast<Fe *>(tables_dev[t]);
        sumcheck_common(ctx->c, g, tabs, k, num_vars, eq_point_mont, wrap_round(round_fn, user), rounds_out,
                        point_out, final_evals_out, /*noalloc=*/true);
    });
}

// The flat gate list of the C ABI as the device code holds it.
static GkrList load_gkr_list(const apg_gkr_list *list) {
    if (!list) throw ApgError(APG_ERR_ARG, "gkr: no gate list");
    GkrList l;
    const size_t all = (size_t)list->ncached + list->nproducing + list->nenforcing;
    if (all && !list->gates) throw ApgError(APG_ERR_ARG, "gkr: no gates");
    if ((list->nsources && !list->sources) || (list->ncoeffs && !list->coeffs_mont) ||
        (list->nvirtual && !list->virtual_kinds))
        throw ApgError(APG_ERR_ARG, "gkr: a null array");
    l.ncached = list->ncached;
    l.nproducing = list->nproducing;
    l.nenforcing = list->nenforcing;
    l.halving = list->halving != 0;
    l.gates.resize(all);
    for (size_t i = 0; i < all; i++) {
        const apg_gkr_gate &g = list->gates[i];
        l.gates[i] = GkrGate{g.shape, g.src_begin, g.src_end, g.coef_begin, g.split};
    }
    l.sources.assign(list->sources, list->sources + list->nsources);
    l.virtual_kinds.assign(list->virtual_kinds, list->virtual_kinds + list->nvirtual);
    l.coeffs.resize(list->ncoeffs);
    for (uint32_t i = 0; i < list->ncoeffs; i++) l.coeffs[i] = load_fe(list->coeffs_mont + 32 * (size_t)i);
    return l;
}

static void gkr_forward_common(Context &ctx, const apg_gkr_list *list, const Fe *const *cols, size_t ncols,
                               size_t num_vars, Fe *const *out) {
    if (num_vars > 40) throw ApgError(APG_ERR_ARG, "num_vars out of range");
    const GkrList l = load_gkr_list(list);
    gkr_forward_device(ctx, l, cols, ncols, (int)num_vars, out);
}

int apg_gkr_forward_layer_dev(apg_ctx *ctx, const apg_gkr_list *list, const void *const *cols_dev, size_t ncols,
                              size_t num_vars, void *const *out_dev) {
    return guarded([&] {
        need(ctx);
        std::lock_guard<std::mutex> lk(ctx->c.mu);
        HIP_CHECK(hipSetDevice(ctx->c.device));
        std::vector<const Fe *> cols(ncols);
        for (size_t c = 0; c < ncols; c++) cols[c] = static_cast<const Fe *>(cols_dev[c]);
        const uint32_t np = list ? list->nproducing : 0;
        std::vector<Fe *> out(np);
        for (uint32_t g = 0; g < np; g++) out[g] = static_cast<Fe *>(out_dev[g]);
        gkr_forward_common(ctx->c, list, cols.data(), ncols, num_vars, out.data());
    });
}

int apg_gkr_forward_layer(apg_ctx *ctx, const apg_gkr_list *list, const uint8_t *const *cols_mont, size_t ncols,
                          size_t num_vars, uint8_t *const *out_mont) {
    return guarded([&] {
        need(ctx);
        if (num_vars > 40) throw ApgError(APG_ERR_ARG, "num_vars out of range");
        if (list && list->halving && num_vars < 1) throw ApgError(APG_ERR_ARG, "gkr: a halving list needs num_vars >= 1");
        std::lock_guard<std::mutex> lk(ctx->c.mu);
        HIP_CHECK(hipSetDevice(ctx->c.device));
        const size_t len = size_t(1) << num_vars;
        const size_t rows = list && list->halving ? len / 2 : len;
        const uint32_t np = list ? list->nproducing : 0;
        std::vector<Fe *> cols = upload_tables(ctx->c, cols_mont, ncols, len);
        Fe *pool = ctx->c.d<Fe>("gkr.fwd_out", std::max<size_t>((size_t)np * rows, 1));
        std::vector<Fe *> out(np);
        for (uint32_t g = 0; g < np; g++) out[g] = pool + (size_t)g * rows;
        std::vector<const Fe *> ccols(cols.begin(), cols.end());
        gkr_forward_common(ctx->c, list, ccols.data(), ncols, num_vars, out.data());
        for (uint32_t g = 0; g < np; g++)
            HIP_CHECK(hipMemcpyAsync(out_mont[g], out[g], rows * 32, hipMemcpyDeviceToHost, ctx->c.stream));
        ctx->c.sync();
    });
}

static void gkr_sumcheck_common(Context &ctx, const apg_gkr_list *list, const uint8_t *weights_mont,
                                const uint8_t *eq_point_mont, size_t num_vars, Fe *const *lower,
                                Fe *const *upper, size_t ncols, const RoundFn &fn, uint8_t *rounds_out,
                                uint8_t *point_out, uint8_t *finals_out) {
    if (num_vars > 40) throw ApgError(APG_ERR_ARG, "num_vars out of range");
    const GkrList l = load_gkr_list(list);
    const size_t nw = (size_t)l.nproducing + l.nenforcing;
    if (nw && !weights_mont) throw ApgError(APG_ERR_ARG, "gkr: no weights");
    std::vector<Fe> w(nw + 1), eqp(num_vars + 1), rounds(4 * num_vars + 1), point(num_vars + 1);
    for (size_t i = 0; i < nw; i++) w[i] = load_fe(weights_mont + 32 * i);
    for (size_t i = 0; i < num_vars; i++) eqp[i] = load_fe(eq_point_mont + 32 * i);
    const size_t nt = ncols * (l.halving ? 2 : 1);
    std::vector<Fe> finals(nt + 1);
    const double t0 = now_ms();
    gkr_prove_sumcheck_device(ctx, l, w.data(), lower, upper, ncols, (int)num_vars, eqp.data(), fn, rounds.data(),
                              point.data(), finals.data());
    ctx.stats.sumcheck_ms += now_ms() - t0;
    for (size_t i = 0; i < 4 * num_vars; i++) store_fe(rounds_out + 32 * i, rounds[i]);
    if (point_out)
        for (size_t i = 0; i < num_vars; i++) store_fe(point_out + 32 * i, point[i]);
    if (finals_out)
        for (size_t i = 0; i < nt; i++) store_fe(finals_out + 32 * i, finals[i]);
}

int apg_gkr_prove_sumcheck_dev(apg_ctx *ctx, const apg_gkr_list *list, const uint8_t *weights_mont,
                               const uint8_t *eq_point_mont, size_t num_vars, void *const *lower_dev,
                               void *const *upper_dev, size_t ncols, apg_round_fn round_fn, void *user,
                               uint8_t *rounds_out, uint8_t *point_out, uint8_t *finals_out) {
    return guarded([&] {
        need(ctx);
        if (list && list->halving && ncols && !upper_dev) throw ApgError(APG_ERR_ARG, "gkr: a halving list needs child-1 tables");
        std::lock_guard<std::mutex> lk(ctx->c.mu);
        HIP_CHECK(hipSetDevice(ctx->c.device));
        std::vector<Fe *> lo(ncols + 1), up(ncols + 1);
        for (size_t c = 0; c < ncols; c++) {
            lo[c] = static_cast<Fe *>(lower_dev[c]);
            up[c] = list && list->halving ? static_cast<Fe *>(upper_dev[c]) : nullptr;
        }
        gkr_sumcheck_common(ctx->c, list, weights_mont, eq_point_mont, num_vars, lo.data(), up.data(), ncols,
                            wrap_round(round_fn, user), rounds_out, point_out, finals_out);
    });
}

int apg_gkr_prove_sumcheck(apg_ctx *ctx, const apg_gkr_list *list, const uint8_t *weights_mont,
                           const uint8_t *eq_point_mont, size_t num_vars, const uint8_t *const *lower_mont,
                           const uint8_t *const *upper_mont, size_t ncols, apg_round_fn round_fn, void *user,
                           uint8_t *rounds_out, uint8_t *point_out, uint8_t *finals_out) {
    return guarded([&] {
        need(ctx);
        if (num_vars > 40) throw ApgError(APG_ERR_ARG, "num_vars out of range");
        const bool halving = list && list->halving;
        if (halving && ncols && !upper_mont) throw ApgError(APG_ERR_ARG, "gkr: a halving list needs child-1 tables");
        std::lock_guard<std::mutex> lk(ctx->c.mu);
        HIP_CHECK(hipSetDevice(ctx->c.device));
        const size_t len = size_t(1) << num_vars;
        std::vector<const uint8_t *> all(lower_mont, lower_mont + ncols);
        if (halving) all.insert(all.end(), upper_mont, upper_mont + ncols);
        std::vector<Fe *> t = upload_tables(ctx->c, all.data(), all.size(), len);
        std::vector<Fe *> lo(t.begin(), t.begin() + ncols), up(ncols + 1, nullptr);
        if (halving)
            for (size_t c = 0; c < ncols; c++) up[c] = t[ncols + c];
        lo.push_back(nullptr);
        gkr_sumcheck_common(ctx->c, list, weights_mont, eq_point_mont, num_vars, lo.data(), up.data(), ncols,
                            wrap_round(round_fn, user), rounds_out, point_out, finals_out);
    });
}

static void witness_digest_host(const Fe *const *cols, size_t k, size_t num_vars, uint8_t out[32]) {
    if (k == 0) throw ApgError(APG_ERR_ARG, "witness_digest: a witness has at least one column");
    if (num_vars > 40) throw ApgError(APG_ERR_ARG, "num_vars out of range");
    Transcript s;
    const Fe hdr[2] = {f_from_u64<Fr>(k), f_from_u64<Fr>(num_vars)};
    s.append_scalars(TAG_WITNESS_DIGEST, hdr, 2);
    for (size_t c = 0; c < k; c++) s.append_scalars(TAG_WITNESS_DIGEST, cols[c], size_t(1) << num_vars);
    store_fe(out, s.sample());
}

int apg_witness_digest(const uint8_t *const *columns_mont, size_t k, size_t num_vars, uint8_t out_mont[32]) {
    return guarded([&] {
        std::vector<const Fe *> cols(k);
        for (size_t c = 0; c < k; c++) cols[c] = reinterpret_cast<const Fe *>(columns_mont[c]);
        witness_digest_host(cols.data(), k, num_vars, out_mont);
    });
}

int apg_witness_digest_dev(apg_ctx *ctx, const void *columns_dev, size_t k, size_t num_vars,
                           uint8_t out_mont[32]) {
    return guarded([&] {
        need(ctx);
        if (k == 0) throw ApgError(APG_ERR_ARG, "witness_digest: a witness has at least one column");
        if (num_vars > 40) throw ApgError(APG_ERR_ARG, "num_vars out of range");
        std::lock_guard<std::mutex> lk(ctx->c.mu);
        HIP_CHECK(hipSetDevice(ctx->c.device));
        const size_t n = size_t(1) << num_vars;
        std::vector<Fe> host(k * n);
        HIP_CHECK(hipMemcpyAsync(host.data(), columns_dev, k * n * sizeof(Fe), hipMemcpyDeviceToHost,
                                 ctx->c.stream));
        ctx->c.sync();
        std::vector<const Fe *> cols(k);
        for (size_t c = 0; c < k; c++) cols[c] = host.data() + c * n;
        witness_digest_host(cols.data(), k, num_vars, out_mont);
    });
}

int apg_zerocheck_prove(apg_ctx *ctx, const apg_term *terms, size_t n_terms,
                        const uint8_t *const *tables_mont, size_t k, size_t num_vars,
                        apg_transcript *t, uint8_t *rounds_out, uint8_t *final_evals_out) {
    return guarded([&] {
        need(ctx);
        if (num_vars > 40) throw ApgError(APG_ERR_ARG, "num_vars out of range");
        std::lock_guard<std::mutex> lk(ctx->c.mu);
        HIP_CHECK(hipSetDevice(ctx->c.device));
        Gate g = make_gate(terms, n_terms);
        // crates/sumcheck/src/lib.rs:381-383: n eq-randomizers first
        std::vector<uint8_t> r(32 * (num_vars + 1));
        for (size_t i = 0; i < num_vars; i++) store_fe(&r[32 * i], t->t.challenge(TAG_SUMCHECK_CHALLENGE));
        auto tabs = upload_tables(ctx->c, tables_mont, k, size_t(1) << num_vars);
        const RoundFn fn = transcript_round(t->t);
        sumcheck_common(ctx->c, g, tabs.data(), k, num_vars, r.data(), fn, rounds_out, nullptr,
                        final_evals_out);
        std::vector<Fe> fin(k);
        for (size_t i = 0; i < k; i++) fin[i] = load_fe(final_evals_out + 32 * i);
        t->t.append_scalars(TAG_SUMCHECK_FINAL_EVALS, fin.data(), k);
    });
}

int apg_prove_shard_dev(apg_ctx *ctx, apg_srs *srs, const void *columns_dev, size_t num_vars,
                        size_t k, int degree, apg_transcript *t, uint8_t *commitments_wire,
                        uint8_t *rounds_out, uint8_t *final_evals_out) {
    return guarded([&] {
        check_shard_args(num_vars, k, degree);
        need(ctx);
        need(srs);
        need(t);
        need(columns_dev);
        std::lock_guard<std::mutex> lk(ctx->c.mu);
        HIP_CHECK(hipSetDevice(ctx->c.device));
        prove_shard_impl(ctx->c, srs, static_cast<const Fe *>(columns_dev), num_vars, k, degree, t->t,
                         commitments_wire, rounds_out, final_evals_out);
    });
}

int apg_engine_create(int device, int workers, size_t hbm_budget_bytes, apg_engine **out) {
    return guarded([&] {
        int ndev = 0;
        HIP_CHECK(hipGetDeviceCount(&ndev));
        if (device < 0 || device >= ndev) throw ApgError(APG_ERR_ARG, "no such device");
        if (workers < 1 || workers > 64) throw ApgError(APG_ERR_ARG, "workers must be 1..64");
        auto *e = new apg_engine();
        e->e = std::make_unique<Engine>(device, workers, hbm_budget_bytes);
        *out = e;
    });
}

int apg_engine_create_lanes(int device, int workers, size_t hbm_budget_bytes, int lane_workers, int lane_cus,
                            double small_scalars, apg_engine **out) {
    return guarded([&] {
        int ndev = 0;
        HIP_CHECK(hipGetDeviceCount(&ndev));
        if (device < 0 || device >= ndev) throw ApgError(APG_ERR_ARG, "no such device");
        if (workers < 2 || workers > 64) throw ApgError(APG_ERR_ARG, "workers must be 2..64");
        if (lane_workers < 1 || lane_workers >= workers) throw ApgError(APG_ERR_ARG, "lane_workers must be 1..workers-1");
        if (lane_cus < 0) throw ApgError(APG_ERR_ARG, "lane_cus must be >= 0");
        auto *e = new apg_engine();
        e->e = std::make_unique<Engine>(device, workers, hbm_budget_bytes, lane_workers, lane_cus);
        e->e->set_policy(Policy::kShortest, lane_workers, small_scalars);
        *out = e;
    });
}

void apg_engine_destroy(apg_engine *e) { delete e; }

int apg_engine_submit_shard(apg_engine *e, apg_srs *srs, const void *columns_dev, size_t num_vars,
                            size_t k, int degree, const apg_transcript *seed, uint64_t *job_id) {
    return guarded([&] {
        check_shard_args(num_vars, k, degree);
        need(e);
        need(srs);
        need(columns_dev);
        need(job_id);
        auto job = std::make_unique<Job>();
        job->mem_bytes = shard_hbm_estimate(num_vars, k, degree);
        job->work = std::ldexp((double)k, (int)num_vars);
        job->iso_ms = shard_time_estimate_ms(num_vars, k);
        const Transcript t0 = seed ? seed->t : Transcript();
        const Fe *cols = static_cast<const Fe *>(columns_dev);
        Engine *eng = e->e.get();
        const double work = job->work;
        job->fn = [=](Context &c, Job &j) {
            c.low_acc_prio = eng->small_work() > 0 && work > eng->small_work();
            Transcript t = t0;
            const int d = degree;
            j.commitments.assign(64 * k, 0);
            j.rounds.assign(32 * num_vars * (d + 1), 0);
            j.finals.assign(32 * (d - 1), 0);
            prove_shard_impl(c, srs, cols, num_vars, k, degree, t, j.commitments.data(), j.rounds.data(),
                             j.finals.data());
            j.transcript_blob.assign(reinterpret_cast<const uint8_t *>(&t),
                                     reinterpret_cast<const uint8_t *>(&t) + sizeof(Transcript));
        };
        *job_id = e->e->submit(std::move(job));
    });
}

int apg_engine_poll(apg_engine *e, uint64_t job_id) {
    if (!e) return APG_ERR_ARG;
    const int r = e->e->poll(job_id);
    if (r < 0) {
        set_last_error("engine: unknown or already collected job id");
        return APG_ERR_ARG;
    }
    return r;
}

int apg_engine_wait(apg_engine *e, uint64_t job_id, uint8_t *commitments_wire, uint8_t *rounds_out,
                    uint8_t *final_evals_out, apg_transcript *transcript_out, apg_job_stats *stats) {
    return guarded([&] {
        need(e);
        auto j = e->e->wait(job_id);
        if (stats) {
            stats->queue_ms = j->t_start - j->t_submit;
            stats->run_ms = j->t_end - j->t_start;
            stats->total_ms = j->t_end - j->t_submit;
            stats->worker = j->worker;
        }
        if (j->status != 0) throw ApgError(j->status < 0 ? j->status : APG_ERR_INTERNAL, j->error);
        if (commitments_wire) std::memcpy(commitments_wire, j->commitments.data(), j->commitments.size());
        if (rounds_out) std::memcpy(rounds_out, j->rounds.data(), j->rounds.size());
        if (final_evals_out) std::memcpy(final_evals_out, j->finals.data(), j->finals.size());
        if (transcript_out) std::memcpy(&transcript_out->t, j->transcript_blob.data(), sizeof(Transcript));
    });
}

int apg_engine_stats(apg_engine *e, apg_engine_info *out) {
    return guarded([&] {
        need(e);
        EngineStats s = e->e->stats();
        out->queued = s.queued;
        out->running = s.running;
        out->completed = s.completed;
        out->hbm_in_use = s.hbm_in_use;
        out->hbm_peak = s.hbm_peak;
        out->hbm_budget = s.hbm_budget;
        out->arena_bytes = s.arena_bytes;
        out->workers = (size_t)e->e->workers();
    });
}

int apg_engine_set_policy(apg_engine *e, int policy, int reserve, double small_scalars) {
    return guarded([&] {
        need(e);
        if (policy != APG_POLICY_FIFO && policy != APG_POLICY_SHORTEST && policy != APG_POLICY_LONGEST)
            throw ApgError(APG_ERR_ARG, "unknown engine policy");
        if (reserve < 0) throw ApgError(APG_ERR_ARG, "reserve must be >= 0");
        e->e->set_policy(static_cast<Policy>(policy), reserve, small_scalars);
    });
}

int apg_engine_set_preemption(apg_engine *e, double ratio, int mode, uint64_t *pauses) {
    return guarded([&] {
        need(e);
        if (!(ratio >= 0) || (ratio > 0 && ratio < 1)) throw ApgError(APG_ERR_ARG, "preemption ratio must be 0 or >= 1");
        if (mode != APG_PREEMPT_SIZE && mode != APG_PREEMPT_URGENCY)
            throw ApgError(APG_ERR_ARG, "unknown preemption mode");
        e->e->set_preemption(ratio, static_cast<Preempt>(mode));
        if (pauses) *pauses = e->e->pauses();
    });
}

int apg_engine_set_profiling(apg_engine *e, int on) {
    return guarded([&] {
        need(e);
        e->e->set_profiling(on != 0);
    });
}

int apg_engine_kernel_stats(apg_engine *e, apg_stats *out) {
    return guarded([&] {
        need(e);
        need(out);
        export_stats(e->e->kernel_stats(), out);
    });
}

int apg_engine_reset_kernel_stats(apg_engine *e) {
    return guarded([&] {
        need(e);
        e->e->reset_kernel_stats();
    });
}

int apg_transcript_new(apg_transcript **out) {
    return guarded([&] { *out = new apg_transcript(); });
}
void apg_transcript_free(apg_transcript *t) { delete t; }
int apg_transcript_clone(const apg_transcript *t, apg_transcript **out) {
    return guarded([&] { *out = new apg_transcript(*t); });
}
void apg_transcript_observe(apg_transcript *t, const uint8_t x[32]) { t->t.observe(load_fe(x)); }
void apg_transcript_sample(apg_transcript *t, uint8_t out[32]) { store_fe(out, t->t.sample()); }
void apg_transcript_append_scalars(apg_transcript *t, uint64_t tag, const uint8_t *xs, size_t n) {
    std::vector<Fe> v(n + 1);
    for (size_t i = 0; i < n; i++) v[i] = load_fe(xs + 32 * i);
    t->t.append_scalars(tag, v.data(), n);
}
void apg_transcript_append_bytes(apg_transcript *t, uint64_t tag, const uint8_t *b, size_t len) {
    t->t.append_bytes(tag, b, len);
}
void apg_transcript_challenge(apg_transcript *t, uint64_t tag, uint8_t out[32]) {
    store_fe(out, t->t.challenge(tag));
}
void apg_transcript_append_g1(apg_transcript *t, uint64_t tag, const uint8_t *w, size_t k) {
    t->t.append_g1_wire(tag, w, k);
}
int apg_transcript_round(void *tr, const uint8_t *coeffs, size_t n, uint8_t ch[32]) {
    auto *t = static_cast<apg_transcript *>(tr);
    apg_transcript_append_scalars(t, TAG_SUMCHECK_ROUND, coeffs, n);
    apg_transcript_challenge(t, TAG_SUMCHECK_CHALLENGE, ch);
    return 0;
}
int apg_transcript_state(const apg_transcript *t, uint8_t out[232]) {
    for (int i = 0; i < 3; i++) store_fe(out + 32 * i, t->t.state[i]);
    for (int i = 0; i < 2; i++) store_fe(out + 96 + 32 * i, t->t.input[i]);
    for (int i = 0; i < 2; i++) store_fe(out + 160 + 32 * i, t->t.output[i]);
    std::memcpy(out + 224, &t->t.input_len, 4);
    std::memcpy(out + 228, &t->t.output_len, 4);
    return APG_OK;
}

int apg_transcript_set_state(apg_transcript *t, const uint8_t in[232]) {
    if (!t || !in) return APG_ERR_ARG;
    Fe st[3], ip[2], op[2];
    for (int i = 0; i < 3; i++) st[i] = load_fe(in + 32 * i);
    for (int i = 0; i < 2; i++) ip[i] = load_fe(in + 96 + 32 * i);
    for (int i = 0; i < 2; i++) op[i] = load_fe(in + 160 + 32 * i);
    uint32_t il, ol;
    std::memcpy(&il, in + 224, 4);
    std::memcpy(&ol, in + 228, 4);
    if (il > 2 || ol > 2) return APG_ERR_ARG;
    for (const Fe &x : st)
        if (!f_is_canonical<Fr>(x)) return APG_ERR_ARG;
    for (const Fe &x : ip)
        if (!f_is_canonical<Fr>(x)) return APG_ERR_ARG;
    for (const Fe &x : op)
        if (!f_is_canonical<Fr>(x)) return APG_ERR_ARG;
    for (int i = 0; i < 3; i++) t->t.state[i] = st[i];
    for (int i = 0; i < 2; i++) {
        t->t.input[i] = ip[i];
        t->t.output[i] = op[i];
    }
    t->t.input_len = il;
    t->t.output_len = ol;
    return APG_OK;
}

int apg_poseidon2_permute_dev(apg_ctx *ctx, uint8_t *states_mont, size_t n) {
    return guarded([&] {
        need(ctx);
        if (n == 0) return;
        need(states_mont);
        for (size_t i = 0; i < 3 * n; i++)
            if (!f_is_canonical<Fr>(load_fe(states_mont + 32 * i)))
                throw ApgError(APG_ERR_ARG, "poseidon2: state element not canonical");
        std::lock_guard<std::mutex> lk(ctx->c.mu);
        Context &c = ctx->c;
        Fe *dev = c.d<Fe>("p2.states", 3 * n);
        HIP_CHECK(hipMemcpyAsync(dev, states_mont, 96 * n, hipMemcpyHostToDevice, c.stream));
        poseidon2_permute_device(c, dev, n);
        HIP_CHECK(hipMemcpyAsync(states_mont, dev, 96 * n, hipMemcpyDeviceToHost, c.stream));
        c.sync();
    });
}

int apg_transcript_run_dev(apg_ctx *ctx, apg_transcript *const *ts, size_t count, const uint32_t *ops, size_t n_ops,
                           const uint32_t *offsets, const uint8_t *operands_mont, size_t n_operands,
                           uint8_t *outputs_mont, size_t n_outputs) {
    return guarded([&] {
        need(ctx);
        if (count == 0) return;
        need(ts);
        need(offsets);
        if (n_ops) need(ops);
        if (n_operands) need(operands_mont);
        if (n_outputs) need(outputs_mont);
        std::vector<Transcript *> tp(count);
        for (size_t i = 0; i < count; i++) {
            need(ts[i]);
            tp[i] = &ts[i]->t;
        }
        std::vector<Fe> in(n_operands), out(n_outputs);
        for (size_t i = 0; i < n_operands; i++) {
            in[i] = load_fe(operands_mont + 32 * i);
            if (!f_is_canonical<Fr>(in[i])) throw ApgError(APG_ERR_ARG, "transcript program: operand not canonical");
        }
        std::lock_guard<std::mutex> lk(ctx->c.mu);
        transcript_run_device(ctx->c, tp.data(), count, ops, n_ops, offsets, in.data(), n_operands, out.data(),
                              n_outputs);
        for (size_t i = 0; i < n_outputs; i++) store_fe(outputs_mont + 32 * i, out[i]);
    });
}

int apg_fr_to_bytes(const uint8_t mont[32], uint8_t canonical[32]) {
    Fe a = load_fe(mont);
    if (!f_is_canonical<Fr>(a)) return APG_ERR_ARG;
    fe_to_bytes_fr(a, canonical);
    return APG_OK;
}
int apg_fr_from_bytes(const uint8_t canonical[32], uint8_t mont[32]) {
    Fe a;
    if (!fe_from_bytes_fr(canonical, a)) return APG_ERR_ARG;
    store_fe(mont, a);
    return APG_OK;
}
int apg_fq_to_bytes(const uint8_t mont[32], uint8_t canonical[32]) {
    Fe a = load_fe(mont);
    if (!f_is_canonical<Fq>(a)) return APG_ERR_ARG;
    fe_to_bytes_fq(a, canonical);
    return APG_OK;
}
int apg_fq_from_bytes(const uint8_t canonical[32], uint8_t mont[32]) {
    Fe a;
    if (!fe_from_bytes_fq(canonical, a)) return APG_ERR_ARG;
    store_fe(mont, a);
    return APG_OK;
}
void apg_fr_from_u64(uint64_t x, uint8_t mont[32]) { store_fe(mont, f_from_u64<Fr>(x)); }

namespace {

__global__ void k_debug_fq29(int op, const Fe *a, const Fe *b, size_t n, Fe *out) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = f29_test_chain<Fq>(a[i], b[i], op);
}

// One chain: points [lo, hi) summed in the lazy 29-bit form with mixed adds
// (form 0), in the 32-bit form with mixed adds (form 1), or in the lazy form
// with general adds of the points scaled to Z = index + 2 (form 2).
APG_HD Xyzz debug_chain(int form, const Affine *pts, const uint8_t *neg, uint64_t lo, uint64_t hi) {
    if (form == 0) {
        Xyzz29 acc{};
        bool id = true;
        for (uint64_t i = lo; i < hi; i++) xyzz29_accumulate(acc, id, affine_to_table(pts[i]), neg[i] != 0);
        return xyzz29_to32(acc, id);
    }
    if (form == 1) {
        Xyzz acc = xyzz_identity();
        for (uint64_t i = lo; i < hi; i++) {
            Affine p = pts[i];
            if (neg[i]) p.y = f_neg<Fq>(p.y);
            acc = xyzz_add_affine(acc, p);
        }
        return acc;
    }
    Xyzz29 acc = xyzz29_identity();
    for (uint64_t i = lo; i < hi; i++) {
        if (aff_is_inf(pts[i])) continue;
        const Fe z = f_from_u64<Fq>(i + 2), z2 = f_mul<Fq>(z, z), z3 = f_mul<Fq>(z2, z);
        Xyzz q;
        q.x = f_mul<Fq>(pts[i].x, z2);
        q.y = f_mul<Fq>(neg[i] ? f_neg<Fq>(pts[i].y) : pts[i].y, z3);
        q.zz = z2;
        q.zzz = z3;
        acc = xyzz29_add(acc, xyzz29_from32(q));
    }
    return xyzz29_store(acc);
}

__global__ void k_debug_accumulate(int form, const Affine *pts, const uint8_t *neg, const uint64_t *off,
                                   size_t ncases, Xyzz *out) {
    const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c < ncases) out[c] = debug_chain(form, pts, neg, off[c], off[c + 1]);
}

// Issue-rate probe of the 32 x 32 -> 64 multiply-add the Fq/Fr multiplies
// are built from: every lane runs 32 independent chains acc = lo(acc) * m +
// acc (one v_mad_u64_u32 each), so the rate measured is the instruction's
// issue rate, not its latency.
__global__ void __launch_bounds__(256) k_imul_peak(uint32_t seed, int iters, uint64_t *out) {
    uint64_t acc[32];
    const uint32_t m = seed * 2654435761u + threadIdx.x + 1u;
#pragma unroll
    for (int c = 0; c < 32; c++) acc[c] = ((uint64_t)(m + (uint32_t)c) << 7) | (uint64_t)c;
    for (int i = 0; i < iters; i++) {
#pragma unroll
        for (int c = 0; c < 32; c++) acc[c] = (uint64_t)(uint32_t)acc[c] * m + acc[c];
    }
    uint64_t x = 0;
#pragma unroll
    for (int c = 0; c < 32; c++) x ^= acc[c];
    out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = x;
}

// Device scratch for the debug entries: copies in, one launch, copies out.
struct DebugBufs {
    std::vector<void *> ptrs;
    void *get(size_t bytes) {
        void *p = nullptr;
        HIP_CHECK(hipMalloc(&p, bytes ? bytes : 1));
        ptrs.push_back(p);
        return p;
    }
    ~DebugBufs() {
        for (void *p : ptrs) (void)hipFree(p);
    }
};

}  // namespace

int apg_debug_imul_peak(apg_ctx *ctx, int iters, int reps, double *best_ms, double *imul_per_s) {
    return guarded([&] {
        need(ctx);
        if (iters < 1 || reps < 1 || !best_ms || !imul_per_s) throw ApgError(APG_ERR_ARG, "imul peak: bad arguments");
        std::lock_guard<std::mutex> lk(ctx->c.mu);
        HIP_CHECK(hipSetDevice(ctx->c.device));
        const unsigned blocks = 8u * (unsigned)ctx->c.cu_count;
        uint64_t *out = ctx->c.d<uint64_t>("peak.out", (size_t)blocks * 256);
        hipEvent_t e0, e1;
        HIP_CHECK(hipEventCreate(&e0));
        HIP_CHECK(hipEventCreate(&e1));
        float best = 1e30f;
        for (int r = 0; r <= reps; r++) {  // the first launch warms up
            HIP_CHECK(hipEventRecord(e0, ctx->c.stream));
            hipLaunchKernelGGL(k_imul_peak, dim3(blocks), dim3(256), 0, ctx->c.stream, 12345u + (uint32_t)r, iters, out);
            HIP_CHECK(hipGetLastError());
            HIP_CHECK(hipEventRecord(e1, ctx->c.stream));
            HIP_CHECK(hipEventSynchronize(e1));
            float ms = 0;
            HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
            if (r > 0) best = std::min(best, ms);
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        *best_ms = best;
        *imul_per_s = (double)blocks * 256.0 * 32.0 * (double)iters / (best * 1e-3);
    });
}

int apg_debug_fq29_batch(apg_ctx *ctx, int op, const uint8_t *a, const uint8_t *b, size_t n, uint8_t *out) {
    if (op < 0 || op > 5) return APG_ERR_ARG;
    if (!ctx) {
        for (size_t i = 0; i < n; i++)
            store_fe(out + 32 * i, f29_test_chain<Fq>(load_fe(a + 32 * i), load_fe(b + 32 * i), op));
        return APG_OK;
    }
    return guarded([&] {
        std::lock_guard<std::mutex> lk(ctx->c.mu);
        HIP_CHECK(hipSetDevice(ctx->c.device));
        DebugBufs d;
        Fe *da = (Fe *)d.get(32 * n), *db = (Fe *)d.get(32 * n), *dout = (Fe *)d.get(32 * n);
        HIP_CHECK(hipMemcpy(da, a, 32 * n, hipMemcpyHostToDevice));
        HIP_CHECK(hipMemcpy(db, b, 32 * n, hipMemcpyHostToDevice));
        if (n) hipLaunchKernelGGL(k_debug_fq29, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->c.stream, op, da, db, n, dout);
        HIP_CHECK(hipGetLastError());
        ctx->c.sync();
        HIP_CHECK(hipMemcpy(out, dout, 32 * n, hipMemcpyDeviceToHost));
    });
}

int apg_debug_accumulate29(apg_ctx *ctx, int form, const uint8_t *points, const uint8_t *neg,
                           const uint64_t *offsets, size_t ncases, uint8_t *out_affine) {
    if (form < 0 || form > 2) return APG_ERR_ARG;
    for (size_t c = 0; c < ncases; c++)
        if (offsets[c + 1] < offsets[c]) return APG_ERR_ARG;
    const size_t npts = ncases ? (size_t)offsets[ncases] : 0;
    std::vector<Xyzz> sums(ncases);
    if (!ctx) {
        std::vector<Affine> pts(npts);
        if (npts) std::memcpy(pts.data(), points, 64 * npts);
        for (size_t c = 0; c < ncases; c++) sums[c] = debug_chain(form, pts.data(), neg, offsets[c], offsets[c + 1]);
    } else {
        const int rc = guarded([&] {
            std::lock_guard<std::mutex> lk(ctx->c.mu);
            HIP_CHECK(hipSetDevice(ctx->c.device));
            DebugBufs d;
            Affine *dp = (Affine *)d.get(64 * npts);
            uint8_t *dn = (uint8_t *)d.get(npts);
            uint64_t *doff = (uint64_t *)d.get(8 * (ncases + 1));
            Xyzz *dout = (Xyzz *)d.get(sizeof(Xyzz) * ncases);
            HIP_CHECK(hipMemcpy(dp, points, 64 * npts, hipMemcpyHostToDevice));
            HIP_CHECK(hipMemcpy(dn, neg, npts, hipMemcpyHostToDevice));
            HIP_CHECK(hipMemcpy(doff, offsets, 8 * (ncases + 1), hipMemcpyHostToDevice));
            if (ncases)
                hipLaunchKernelGGL(k_debug_accumulate, dim3((unsigned)((ncases + 63) / 64)), dim3(64), 0,
                                   ctx->c.stream, form, dp, dn, doff, ncases, dout);
            HIP_CHECK(hipGetLastError());
            ctx->c.sync();
            HIP_CHECK(hipMemcpy(sums.data(), dout, sizeof(Xyzz) * ncases, hipMemcpyDeviceToHost));
        });
        if (rc != APG_OK) return rc;
    }
    for (size_t c = 0; c < ncases; c++) {
        const Affine r = xyzz_to_affine(sums[c]);
        std::memcpy(out_affine + 64 * c, &r, 64);
    }
    return APG_OK;
}

int apg_debug_mul32_batch(int which, const uint8_t *a, const uint8_t *b, size_t n, uint8_t *out) {
    if (which != 0 && which != 1) return APG_ERR_ARG;
    for (size_t i = 0; i < n; i++) {
        const Fe x = load_fe(a + 32 * i), y = load_fe(b + 32 * i);
        store_fe(out + 32 * i, which ? f_mul32<Fq>(x, y) : f_mul32<Fr>(x, y));
    }
    return APG_OK;
}

int apg_debug_poseidon2_batch(int form, const uint8_t *states, size_t n, uint8_t *out) {
    return guarded([&] {
        if (form < 0 || form > 2 || (n && (!states || !out))) throw ApgError(APG_ERR_ARG, "bad arguments");
        for (size_t i = 0; i < n; i++) {
            Fe st[3];
            for (int l = 0; l < 3; l++) st[l] = load_fe(states + 96 * i + 32 * l);
            if (!poseidon2_permute_form(form, st))
                throw ApgError(APG_ERR_ARG, "this CPU lacks BMI2/ADX (form 2)");
            for (int l = 0; l < 3; l++) store_fe(out + 96 * i + 32 * l, st[l]);
        }
    });
}

int apg_g1_to_wire(const uint8_t mont[64], uint8_t wire[64]) {
    Affine a;
    std::memcpy(&a, mont, 64);
    if (aff_is_inf(a)) {
        std::memset(wire, 0, 64);
        return APG_OK;
    }
    if (!f_is_canonical<Fq>(a.x) || !f_is_canonical<Fq>(a.y)) return APG_ERR_ARG;
    fe_to_bytes_fq(a.x, wire);
    fe_to_bytes_fq(a.y, wire + 32);
    return APG_OK;
}

int apg_g1_from_wire(const uint8_t wire[64], uint8_t mont[64]) {
    bool zero = true;
    for (int i = 0; i < 64; i++) zero &= wire[i] == 0;
    if (zero) {
        std::memset(mont, 0, 64);
        return APG_OK;
    }
    Affine a;
    if (!fe_from_bytes_fq(wire, a.x) || !fe_from_bytes_fq(wire + 32, a.y)) return APG_ERR_ARG;
    Fe y2 = f_sqr<Fq>(a.y);
    Fe rhs = f_add<Fq>(f_mul<Fq>(f_sqr<Fq>(a.x), a.x), f_from_u64<Fq>(3));
    if (!f_eq(y2, rhs)) return APG_ERR_ARG;
    std::memcpy(mont, &a, 64);
    return APG_OK;
}

}  // extern "C"
