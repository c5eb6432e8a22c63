// SRS ingestion for gfx950: the archive (crates/srs/src/lib.rs:175-240,
// docs/spec/srs.md §4) with every G1 power decoded and validated on the
// device, the G2 points of its header validated on the host, and
// batch_to_affine on the device.
//
// G1 has cofactor 1, so G1Affine::from_bytes's subgroup test is the curve
// equation (crates/curve/src/g1.rs:80-88): a point is valid when both
// coordinates are canonical (< q) and y^2 = x^3 + 3, and 64 zero bytes are
// infinity. That is one thread per point: two canonical compares, two
// multiplies into the Montgomery domain and three more for the equation, so
// the decode runs at HBM speed and the archive's read from storage is the
// cost. The G2 check is the reference's direct one (on the twist, then
// r P = O by a double-and-add over Fq2), two points per archive, on the host.
#include <cstdio>
#include <cstring>
#include <vector>

#include "srs_io.h"
#include "transcript.h"

namespace apg {

namespace {

__global__ void k_g1_decode(Affine *buf, size_t n, unsigned long long *bad) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint4 *q = reinterpret_cast<uint4 *>(buf + i);
    const uint4 a = q[0], b = q[1], c = q[2], d = q[3];
    Fe x, y;
    x.l[0] = a.x; x.l[1] = a.y; x.l[2] = a.z; x.l[3] = a.w;
    x.l[4] = b.x; x.l[5] = b.y; x.l[6] = b.z; x.l[7] = b.w;
    y.l[0] = c.x; y.l[1] = c.y; y.l[2] = c.z; y.l[3] = c.w;
    y.l[4] = d.x; y.l[5] = d.y; y.l[6] = d.z; y.l[7] = d.w;
    if (f_is_zero(x) && f_is_zero(y)) return;  // infinity: stays all-zero
    bool ok = f_is_canonical<Fq>(x) && f_is_canonical<Fq>(y);
    Fe xm = f_zero<Fq>(), ym = xm;
    if (ok) {
        xm = f_to_mont<Fq>(x);
        ym = f_to_mont<Fq>(y);
        const Fe rhs = f_add<Fq>(f_mul<Fq>(f_sqr<Fq>(xm), xm), f_from_u64<Fq>(3));
        ok = f_eq(f_sqr<Fq>(ym), rhs);
    }
    if (!ok) {
        atomicMin(bad, (unsigned long long)i);
        xm = ym = f_zero<Fq>();
    }
    q[0] = make_uint4(xm.l[0], xm.l[1], xm.l[2], xm.l[3]);
    q[1] = make_uint4(xm.l[4], xm.l[5], xm.l[6], xm.l[7]);
    q[2] = make_uint4(ym.l[0], ym.l[1], ym.l[2], ym.l[3]);
    q[3] = make_uint4(ym.l[4], ym.l[5], ym.l[6], ym.l[7]);
}

__global__ void k_g1_encode(const Affine *in, size_t n, uint8_t *out) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Affine p = in[i];
    Fe x = f_zero<Fq>(), y = x;
    if (!aff_is_inf(p)) {
        x = f_from_mont<Fq>(p.x);
        y = f_from_mont<Fq>(p.y);
    }
    uint4 *q = reinterpret_cast<uint4 *>(out + 64 * i);
    q[0] = make_uint4(x.l[0], x.l[1], x.l[2], x.l[3]);
    q[1] = make_uint4(x.l[4], x.l[5], x.l[6], x.l[7]);
    q[2] = make_uint4(y.l[0], y.l[1], y.l[2], y.l[3]);
    q[3] = make_uint4(y.l[4], y.l[5], y.l[6], y.l[7]);
}

// (x / z^2, y / z^3), the identity for z = 0 (crates/curve/src/g1.rs:328-347)
__global__ void k_g1_jac_to_affine(const Fe *jac, size_t n, Affine *out) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Fe x = jac[3 * i], y = jac[3 * i + 1], z = jac[3 * i + 2];
    Affine r;
    if (f_is_zero(z)) {
        r.x = r.y = f_zero<Fq>();
    } else {
        const Fe zi = f_inv<Fq>(z), zi2 = f_sqr<Fq>(zi);
        r.x = f_mul<Fq>(x, zi2);
        r.y = f_mul<Fq>(y, f_mul<Fq>(zi2, zi));
    }
    out[i] = r;
}

// ---- Fq2 = Fq[u] / (u^2 + 1) and the twist E': y^2 = x^3 + 3 / (9 + u), host

struct Fq2 {
    Fe c0, c1;
};

Fq2 q2_add(const Fq2 &a, const Fq2 &b) { return {f_add<Fq>(a.c0, b.c0), f_add<Fq>(a.c1, b.c1)}; }
Fq2 q2_sub(const Fq2 &a, const Fq2 &b) { return {f_sub<Fq>(a.c0, b.c0), f_sub<Fq>(a.c1, b.c1)}; }
Fq2 q2_mul(const Fq2 &a, const Fq2 &b) {
    const Fe t0 = f_mul<Fq>(a.c0, b.c0), t1 = f_mul<Fq>(a.c1, b.c1);
    const Fe m = f_mul<Fq>(f_add<Fq>(a.c0, a.c1), f_add<Fq>(b.c0, b.c1));
    return {f_sub<Fq>(t0, t1), f_sub<Fq>(f_sub<Fq>(m, t0), t1)};
}
Fq2 q2_sqr(const Fq2 &a) { return q2_mul(a, a); }
Fq2 q2_dbl(const Fq2 &a) { return q2_add(a, a); }
bool q2_is_zero(const Fq2 &a) { return f_is_zero(a.c0) && f_is_zero(a.c1); }
bool q2_eq(const Fq2 &a, const Fq2 &b) { return f_eq(a.c0, b.c0) && f_eq(a.c1, b.c1); }
Fq2 q2_inv(const Fq2 &a) {  // conj(a) / (a0^2 + a1^2)
    const Fe ni = f_inv<Fq>(f_add<Fq>(f_sqr<Fq>(a.c0), f_sqr<Fq>(a.c1)));
    return {f_mul<Fq>(a.c0, ni), f_sub<Fq>(f_zero<Fq>(), f_mul<Fq>(a.c1, ni))};
}

// b' = 3 / (9 + u), derived from its definition
Fq2 twist_b() {
    static const Fq2 b = [] {
        const Fq2 nine_u{f_from_u64<Fq>(9), f_one<Fq>()};
        const Fq2 three{f_from_u64<Fq>(3), f_zero<Fq>()};
        return q2_mul(three, q2_inv(nine_u));
    }();
    return b;
}

struct G2J {
    Fq2 x, y, z;
};

G2J g2_dbl(const G2J &p) {  // dbl-2009-l, a = 0
    if (q2_is_zero(p.z)) return p;
    const Fq2 A = q2_sqr(p.x), B = q2_sqr(p.y), C = q2_sqr(B);
    const Fq2 D = q2_dbl(q2_sub(q2_sub(q2_sqr(q2_add(p.x, B)), A), C));
    const Fq2 E = q2_add(q2_dbl(A), A), F = q2_sqr(E);
    G2J r;
    r.x = q2_sub(F, q2_dbl(D));
    r.y = q2_sub(q2_mul(E, q2_sub(D, r.x)), q2_dbl(q2_dbl(q2_dbl(C))));
    r.z = q2_dbl(q2_mul(p.y, p.z));
    return r;
}

G2J g2_add_affine(const G2J &p, const Fq2 &x2, const Fq2 &y2) {  // madd-2007-bl
    if (q2_is_zero(p.z)) return {x2, y2, {f_one<Fq>(), f_zero<Fq>()}};
    const Fq2 z1z1 = q2_sqr(p.z);
    const Fq2 u2 = q2_mul(x2, z1z1), s2 = q2_mul(y2, q2_mul(p.z, z1z1));
    const Fq2 h = q2_sub(u2, p.x), rr = q2_dbl(q2_sub(s2, p.y));
    if (q2_is_zero(h)) {
        if (q2_is_zero(rr)) return g2_dbl(p);
        return {{f_one<Fq>(), f_zero<Fq>()}, {f_one<Fq>(), f_zero<Fq>()}, {f_zero<Fq>(), f_zero<Fq>()}};
    }
    const Fq2 hh = q2_sqr(h), i = q2_dbl(q2_dbl(hh)), j = q2_mul(h, i), v = q2_mul(p.x, i);
    G2J r;
    r.x = q2_sub(q2_sub(q2_sqr(rr), j), q2_dbl(v));
    r.y = q2_sub(q2_mul(rr, q2_sub(v, r.x)), q2_dbl(q2_mul(p.y, j)));
    r.z = q2_sub(q2_sub(q2_sqr(q2_add(p.z, h)), z1z1), hh);
    return r;
}

bool fq2_from_wire(const uint8_t b[64], Fq2 &out) {
    return fe_from_bytes_fq(b, out.c0) && fe_from_bytes_fq(b + 32, out.c1);
}

}  // namespace

bool g2_wire_valid(const uint8_t b[128]) {
    bool zero = true;
    for (int i = 0; i < 128; i++) zero = zero && b[i] == 0;
    if (zero) return true;
    Fq2 x, y;
    if (!fq2_from_wire(b, x) || !fq2_from_wire(b + 64, y)) return false;
    // on the twist (crates/curve/src/g2.rs:117-122)
    if (!q2_eq(q2_sqr(y), q2_add(q2_mul(q2_sqr(x), x), twist_b()))) return false;
    // r P = O (g2.rs:140-145): double-and-add over the bits of r
    G2J acc{{f_one<Fq>(), f_zero<Fq>()}, {f_one<Fq>(), f_zero<Fq>()}, {f_zero<Fq>(), f_zero<Fq>()}};
    for (int bit = 255; bit >= 0; bit--) {
        acc = g2_dbl(acc);
        if ((Fr::mod(bit >> 5) >> (bit & 31)) & 1) acc = g2_add_affine(acc, x, y);
    }
    return q2_is_zero(acc.z);
}

uint64_t g1_decode_wire_device(Context &ctx, Affine *buf, size_t n) {
    unsigned long long *bad = ctx.d<unsigned long long>("srs.bad", 1);
    HIP_CHECK(hipMemsetAsync(bad, 0xff, sizeof(unsigned long long), ctx.stream));
    if (n) {
        hipLaunchKernelGGL(k_g1_decode, dim3(grid_for(n, 256)), dim3(256), 0, ctx.stream, buf, n, bad);
        HIP_CHECK(hipGetLastError());
    }
    unsigned long long *h = ctx.h<unsigned long long>("srs.bad", 1);
    HIP_CHECK(hipMemcpyAsync(h, bad, sizeof(unsigned long long), hipMemcpyDeviceToHost, ctx.stream));
    ctx.sync();
    return (uint64_t)*h;
}

void g1_encode_wire_device(Context &ctx, const Affine *in, size_t n, uint8_t *out_dev) {
    if (!n) return;
    hipLaunchKernelGGL(k_g1_encode, dim3(grid_for(n, 256)), dim3(256), 0, ctx.stream, in, n, out_dev);
    HIP_CHECK(hipGetLastError());
}

void g1_jac_to_affine_device(Context &ctx, const Fe *jac, size_t n, Affine *out) {
    if (!n) return;
    hipLaunchKernelGGL(k_g1_jac_to_affine, dim3(grid_for(n, 128)), dim3(128), 0, ctx.stream, jac, n, out);
    HIP_CHECK(hipGetLastError());
}

namespace {
constexpr char SRS_MAGIC[8] = {'A', 'P', 'O', 'G', 'E', 'S', 'R', 'S'};
constexpr size_t SRS_CHUNK = size_t(1) << 20;  // points per host-to-device chunk (64 MiB)

struct File {
    FILE *f = nullptr;
    ~File() {
        if (f) fclose(f);
    }
};
}  // namespace

namespace {
// The header rules and both G2 points (crates/srs/src/lib.rs:200-234, in that
// order); on SRS_OK the file is positioned at the G1 block.
int check_header(FILE *f, uint32_t *power_out, uint64_t *count_out, uint8_t g2[256], uint64_t *detail) {
    if (fseeko(f, 0, SEEK_END) != 0) return SRS_IO;
    const off_t len = ftello(f);
    if (len < 0 || fseeko(f, 0, SEEK_SET) != 0) return SRS_IO;
    if ((uint64_t)len < SRS_ARCHIVE_HEADER) return SRS_TRUNCATED;
    uint8_t head[SRS_ARCHIVE_HEADER];
    if (fread(head, 1, SRS_ARCHIVE_HEADER, f) != SRS_ARCHIVE_HEADER) return SRS_IO;
    if (std::memcmp(head, SRS_MAGIC, 8) != 0) return SRS_BAD_MAGIC;
    uint32_t version, power;
    uint64_t count;
    std::memcpy(&version, head + 8, 4);
    std::memcpy(&power, head + 12, 4);
    std::memcpy(&count, head + 16, 8);
    if (version != SRS_ARCHIVE_VERSION) {
        *detail = version;
        return SRS_BAD_VERSION;
    }
    // the power is capped before it is shifted
    if (power > SRS_MAX_POWER || count != (uint64_t(1) << power)) return SRS_BAD_SECTION;
    if ((uint64_t)len != SRS_ARCHIVE_HEADER + count * 64) return SRS_TRUNCATED;
    if (!g2_wire_valid(head + 24)) {
        *detail = count;
        return SRS_INVALID_POINT;
    }
    if (!g2_wire_valid(head + 152)) {
        *detail = count + 1;
        return SRS_INVALID_POINT;
    }
    std::memcpy(g2, head + 24, 256);
    *power_out = power;
    *count_out = count;
    return SRS_OK;
}
}  // namespace

int srs_check_archive_header(const char *path, uint32_t *power, uint8_t g2[256], uint64_t *detail) {
    File fh;
    fh.f = fopen(path, "rb");
    if (!fh.f) return SRS_IO;
    uint64_t count = 0;
    return check_header(fh.f, power, &count, g2, detail);
}

int srs_load_archive(Context &ctx, const char *path, Affine **base, size_t *count_out, uint8_t g2[256],
                     uint64_t *detail) {
    File fh;
    fh.f = fopen(path, "rb");
    if (!fh.f) return SRS_IO;
    uint32_t power = 0;
    uint64_t count = 0;
    const int e = check_header(fh.f, &power, &count, g2, detail);
    if (e != SRS_OK) return e;

    Affine *d = nullptr;
    if (hipMalloc(&d, count * sizeof(Affine)) != hipSuccess)
        throw ApgError(-4, "device out of memory for the SRS");
    struct Guard {
        Affine *p;
        ~Guard() {
            if (p) (void)hipFree(p);
        }
    } guard{d};
    uint8_t *stage = ctx.h<uint8_t>("srs.stage", std::min<size_t>(count, SRS_CHUNK) * 64);
    for (size_t done = 0; done < count;) {
        const size_t take = std::min<size_t>(SRS_CHUNK, count - done);
        if (fread(stage, 64, take, fh.f) != take) return SRS_IO;
        HIP_CHECK(hipMemcpyAsync(reinterpret_cast<uint8_t *>(d) + 64 * done, stage, 64 * take,
                                 hipMemcpyHostToDevice, ctx.stream));
        ctx.sync();  // the staging buffer is reused
        done += take;
    }
    const uint64_t bad = g1_decode_wire_device(ctx, d, count);
    if (bad != UINT64_MAX) {
        *detail = bad;
        return SRS_INVALID_POINT;
    }
    guard.p = nullptr;
    *base = d;
    *count_out = count;
    return SRS_OK;
}

void srs_save_archive(Context &ctx, const Affine *base, size_t n, const uint8_t g2[256], const char *path) {
    if (n == 0 || (n & (n - 1)) != 0) throw ApgError(-2, "srs save: the power count is a power of two");
    uint32_t power = 0;
    while ((size_t(1) << power) < n) power++;
    File fh;
    fh.f = fopen(path, "wb");
    if (!fh.f) throw ApgError(-2, "srs save: cannot open the archive for writing");
    uint8_t head[SRS_ARCHIVE_HEADER];
    std::memcpy(head, SRS_MAGIC, 8);
    const uint32_t version = SRS_ARCHIVE_VERSION;
    const uint64_t count = n;
    std::memcpy(head + 8, &version, 4);
    std::memcpy(head + 12, &power, 4);
    std::memcpy(head + 16, &count, 8);
    std::memcpy(head + 24, g2, 256);
    if (fwrite(head, 1, SRS_ARCHIVE_HEADER, fh.f) != SRS_ARCHIVE_HEADER)
        throw ApgError(-2, "srs save: write failed");
    const size_t chunk = std::min<size_t>(n, SRS_CHUNK);
    uint8_t *dwire = ctx.d<uint8_t>("srs.wire", chunk * 64);
    uint8_t *stage = ctx.h<uint8_t>("srs.stage", chunk * 64);
    for (size_t done = 0; done < n;) {
        const size_t take = std::min<size_t>(chunk, n - done);
        g1_encode_wire_device(ctx, base + done, take, dwire);
        HIP_CHECK(hipMemcpyAsync(stage, dwire, 64 * take, hipMemcpyDeviceToHost, ctx.stream));
        ctx.sync();
        if (fwrite(stage, 64, take, fh.f) != take) throw ApgError(-2, "srs save: write failed");
        done += take;
    }
    if (fflush(fh.f) != 0) throw ApgError(-2, "srs save: write failed");
}

}  // namespace apg
