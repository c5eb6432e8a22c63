// SRS ingestion with the point validation on the device: the archive of
// Srs::save / Srs::load (crates/srs/src/lib.rs:175-240, docs/spec/srs.md §4),
// G1Affine::from_bytes for every power (crates/curve/src/g1.rs:102-127) in a
// kernel, the two G2 points checked on the host (G2Affine::from_bytes,
// crates/curve/src/g2.rs:117-199), and G1Projective::batch_to_affine
// (crates/curve/src/g1.rs:328-347) in a kernel.
#pragma once
#include <stdint.h>

#include "curve.cuh"
#include "runtime.h"

namespace apg {

// SrsError (crates/srs/src/lib.rs) as codes
enum SrsErr : int {
    SRS_OK = 0,
    SRS_IO = 1,
    SRS_TRUNCATED = 2,
    SRS_BAD_MAGIC = 3,
    SRS_BAD_VERSION = 4,
    SRS_BAD_SECTION = 5,
    SRS_INVALID_POINT = 6,
};

constexpr size_t SRS_ARCHIVE_HEADER = 8 + 4 + 4 + 8 + 128 + 128;
constexpr uint32_t SRS_ARCHIVE_VERSION = 1;
constexpr uint32_t SRS_MAX_POWER = 30;

// G2Affine::from_bytes on the host: canonical coordinate halves, on the
// twist, in the order-r subgroup; 128 zero bytes are the point at infinity.
bool g2_wire_valid(const uint8_t b[128]);

// G1Affine::from_bytes for n wire points held in buf (64 B each), decoded in
// place into Montgomery affine points (infinity all-zero) on ctx's stream.
// Returns the lowest index of an invalid point, or UINT64_MAX.
uint64_t g1_decode_wire_device(Context &ctx, Affine *buf, size_t n);

// The canonical wire form of n device points into out_dev (64 B each).
void g1_encode_wire_device(Context &ctx, const Affine *in, size_t n, uint8_t *out_dev);

// Jacobian (x, y, z Montgomery, z = 0 the identity) to affine, one Fermat
// inversion per point.
void g1_jac_to_affine_device(Context &ctx, const Fe *jac, size_t n, Affine *out);

// Srs::load: the header rules and both G2 points on the host, the G1 block
// streamed to the device and decoded there. On success *base (hipMalloc'd,
// the caller's) holds count points. detail: the version (SRS_BAD_VERSION) or
// the point index (SRS_INVALID_POINT: count and count + 1 name the G2
// points). g2 receives the header's two G2 points as stored.
int srs_load_archive(Context &ctx, const char *path, Affine **base, size_t *count, uint8_t g2[256],
                     uint64_t *detail);

// The header rules and both G2 points alone (host only).
int srs_check_archive_header(const char *path, uint32_t *power, uint8_t g2[256], uint64_t *detail);

// Srs::save of n (a power of two) device points with the caller's G2 bytes.
void srs_save_archive(Context &ctx, const Affine *base, size_t n, const uint8_t g2[256], const char *path);

}  // namespace apg
