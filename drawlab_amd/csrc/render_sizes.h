// Device bytes of a render's buffers, shared by the code that allocates them
// (hip/integrator.hip, hip/guide.hip: each checks these sums against its own
// buffer types with static_asserts) and the render service, which reserves
// them before a job starts (host/service.cpp).
#pragma once
#include <stdint.h>

namespace dl_sizes {

// Path state, per path of a pass: ray origin and direction, hit record,
// shadow-ray direction (a shadow ray starts at the ray origin), throughput,
// radiance, pending NEE term and the aux word (one 16-byte vector each), the pcg32 state (8), the
// sample index (4), the film position (8), the occlusion byte (1), the
// scene index (1), two double-buffered material queues of kMatQueueSegs
// segments (4 bytes per path per segment each), and the ray, shadow and
// emission queues (4 each).
constexpr int kMatQueueSegs = 8;
constexpr uint64_t kStateBytesPerPath = 8 * 16 + 8 + 4 + 8 + 1 + 1 + 2 * kMatQueueSegs * 4 + 3 * 4;  // 226

// Film: float4 accumulation (sum w L, sum w) + resolved RGB per pixel.
constexpr uint64_t kFilmBytesPerPixel = 16 + 12;

// Guided rendering (the learned proposal).
constexpr int kGuideMixFloats = 40;          // decoded mixture per queued shading point
constexpr int kGuideRecordFloats = 24;       // training record
constexpr uint32_t kGuideRecordSlots = 1u << 23;
constexpr uint32_t kGuideBatchRows = 1u << 18;  // records per Adam step
constexpr uint32_t kGuideChunkRows = 1024;      // rows per partial weight gradient
constexpr uint32_t kGuideScanTile = 1024;       // records per prefix-sum tile
constexpr int kGuideParams = 21808;
constexpr int kGuideWSlot = 64 * 96;            // bfloat16 elements per layer copy (W or W^T)

// network parameters, Adam moments, gradient, biases, bfloat16 W and W^T,
// scalars
constexpr uint64_t kGuideNetBytes = 4ull * kGuideParams * 4 + 5 * 64 * 4 + 2ull * 5 * kGuideWSlot * 2 + 4 * 4;
// one training batch: bfloat16 layer inputs (96 + 4 x 64 per row), outputs
// (48), two gradient rows (64 each), per-chunk partial gradients of the
// largest layer; the batch's features (96), directions (3) and weights (1)
constexpr uint64_t kGuideTrainBytes =
    (uint64_t)kGuideBatchRows * ((96 + 4 * 64) * 2 + 48 * 4 + 2 * 64 * 4 + (96 + 3 + 1) * 4) +
    (uint64_t)(kGuideBatchRows / kGuideChunkRows) * (64 * 96 + 64) * 4;
// record weights and the prefix-sum tiles over every record slot
constexpr uint64_t kGuideResampleBytes =
    (uint64_t)kGuideRecordSlots * 4 + 2ull * (kGuideRecordSlots / kGuideScanTile + 1) * 4;

// Everything a guided render adds to a plain one whose passes hold `paths`
// paths.
constexpr uint64_t guide_bytes(uint64_t paths) {
    return paths * kGuideMixFloats * 4 + (uint64_t)kGuideRecordSlots * kGuideRecordFloats * 4 + kGuideNetBytes +
           kGuideTrainBytes + kGuideResampleBytes;
}

}  // namespace dl_sizes
