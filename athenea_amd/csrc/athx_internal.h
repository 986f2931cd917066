// athx_internal.h -- what the HIP sources of the rasterizer share: the view
// as the kernels read it, wavefront primitives, and the launchers each source
// file exports to the C ABI (capi.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/athx.h"
#include "../../include/athx_detmath.h"

namespace athx {

// The splat rasterizer's constants (shaders/athenea/splat/frame.slang:42-49).
constexpr uint32_t kCulledKey = 0xFF000000u;
// Floats in a projected record: screen (x, y, z, cutoff), conic (a, b, c,
// alpha), colour (r, g, b, branch bits) -- frame.slang:269-281's ProjRecord
// without the surfel words, which a plain Gaussian leaves zero. The last word
// holds, as uint bits, the branches the projection took (kFlag*), which the
// adjoint replays.
constexpr uint32_t kRecordFloats = 12;
constexpr uint32_t kDepthBits = 24;
constexpr uint32_t kFlagClampXLo = 1u;   // Jacobian point clamped to the frustum's left
constexpr uint32_t kFlagClampXHi = 2u;   // ... right
constexpr uint32_t kFlagClampYLo = 4u;   // ... bottom
constexpr uint32_t kFlagClampYHi = 8u;   // ... top
constexpr uint32_t kFlagPaidBack = 16u;  // paidOpacity's pay-back branch (frame.slang:316-319)
constexpr uint32_t kFlagColourR = 32u;   // channel's harmonics sum > 0 (R; G and B the next bits)
constexpr uint32_t kFlagSrgbR = 256u;    // sRGB decode took the power branch (R; G, B next)
constexpr uint32_t kFlagRelit = 2048u;   // the colour is the relit one

// One view as every kernel reads it: the camera (athx_view), where its tiles
// and slots start in the frame's global numbering, and its targets.
struct DView {
    uint32_t width, height, tiles_x, tiles_y;
    float focal_x, focal_y, centre_x, centre_y;
    float near_z, far_z;
    uint32_t orthographic;
    uint32_t tile_base;  // first global tile of this view
    float m[12];
    float eye[3];
    float bg[4];
    float* colour;
    float* depth;
    float* final_t;
    uint32_t* last;
    const float* dl_dcolour;
};

// What the relit colour reads besides the Gaussian's own geometry
// (athx_relight; the sky is one prepared dome).
struct RelightArgs {
    const float* normals;    // [n][3] stored shading normals (cloud space), or null
    const float* albedo;     // [n][3], or null: the evaluated harmonics
    const float* metallic;   // [n]
    const float* roughness;  // [n]
    const uint32_t* normals_packed;  // [n] octahedral 2 x 16 bits, or null (then normals)
    const uint32_t* pbr;     // [n] metallic byte 0, roughness byte 1, or null (then metallic, roughness)
    float o2w[12];           // object -> world rows
    const uint32_t* texels;  // prepared dome: 8 octahedral levels, 2 words a texel
    const float* sky_sh;     // 16 float4
    const float* sky_sun;    // 2 float4
    uint32_t base_side;
};

// What the projection is told besides the views.
struct ProjectArgs {
    const float* means;
    const float* log_scales;
    const float* quats;
    const float* logits;
    const float* sh;
    const uint16_t* sh_half;
    uint32_t n;
    uint32_t sh_coeffs;
    uint32_t sh_degree;  // min(degree the coefficients hold, the frame's limit)
    uint32_t linear;
    uint32_t antialias;
    uint32_t tile;
    uint32_t relit;  // 1: the colour is the relit one (rl), not the harmonics'
    RelightArgs rl;
    // athenea's packed layout (athx_packed_splats) instead of the arrays above.
    uint32_t packed;
    const float* positions;     // float4 a Gaussian: xyz, opacity
    const uint32_t* shape;      // 4 words a Gaussian
    const uint32_t* packed_sh;  // packed_words a Gaussian: rest coefficients as halves
    uint32_t packed_words;
};

__host__ __device__ inline uint32_t div_up(uint32_t a, uint32_t b) { return (a + b - 1) / b; }

// Which tiles each footprint touches, kept from the count for the emission:
// the reference tests every tile twice (splat_project.slang:764-776 and
// splat_emit.slang:37-49); here the second pass reads the first one's answer.
//
// A footprint over kBigRect tiles is walked in steps of one hit word each
// (big_steps); the big footprints are listed, and their words placed, by a
// scan over the slots (launch_big_list), so no kernel takes a place with an
// atomic on one shared address -- those serialise in the memory system.
struct Footprints {
    uint2* rects;             // [slots] tile rect: x0 | y0 << 16, w | h << 16 (0, 0 when culled before)
    uint64_t* masks;          // [slots] a small footprint's hits, bit k = tile k of its rect, row-major
    uint32_t* big_list;       // [slots] slots whose rect holds over kBigRect tiles, in slot order
    uint32_t* big_mask_base;  // [slots] a big footprint's first hit word (its first step)
    uint64_t* big_masks;      // [big_mask_cap] a word a step
    uint32_t big_mask_cap;
    uint64_t* big_total;      // big footprints << 32 | their steps, written by the scan
};
constexpr uint32_t kBigRect = 64;

// The steps a big footprint's w x h rectangle is walked in, a hit word each:
// whole rows, 64 / w a step, when it is at most 64 tiles wide, else 64
// columns of one row a step.
__host__ __device__ inline uint32_t big_steps(uint32_t w, uint32_t h) {
    return w <= 64 ? div_up(h, 64 / w) : h * div_up(w, 64);
}

// Launchers (each returns the hipError_t of its launches).
// The projection, then the exact tile counts: small rectangles spread over
// the lanes of a wavefront, big ones listed by a scan and walked in steps.
hipError_t launch_project(const ProjectArgs& a, const DView* views_dev, uint32_t nviews, float* records,
                          uint32_t* touched, uint32_t* depth_keys, const Footprints& fp, uint32_t* status_words,
                          hipStream_t s);
// Lists the big footprints (rectangles over kBigRect tiles) in slot order
// with each one's first step; the count and the steps land in *fp.big_total.
hipError_t launch_big_list(const Footprints& fp, uint32_t slots, uint32_t* status_words, hipStream_t s);
hipError_t launch_fill_u32(uint32_t* p, uint32_t value, uint64_t n, hipStream_t s);
size_t scan_status_words(uint32_t n);
hipError_t launch_scan(const uint32_t* in, uint32_t* out, uint64_t* total, uint32_t n, uint32_t* status,
                       hipStream_t s);
// Pairs land at their scanned offsets; those at or past cap are not written
// (the frame is then incomplete, as the reference's emit past its pair
// capacity is, splat_emit.slang:43-46). A pair's key is its global tile:
// the pairs come out in depth order, so the depth half of the reference's
// 64-bit key (global tile << 24 | depth key) is already sorted and is not
// stored.
hipError_t launch_emit(const DView* views_dev, uint32_t nviews, uint32_t n, uint32_t tile, const float* records,
                       const uint32_t* touched, const uint32_t* offsets, const Footprints& fp, uint32_t cap,
                       uint32_t* tiles_out, uint32_t* values, hipStream_t s);
hipError_t launch_ranges(const uint32_t* sorted_tiles, const uint64_t* n_dev, uint32_t cap, uint32_t tiles,
                         uint32_t* ranges, hipStream_t s);
// The reference's 64-bit pair keys of the sorted pairs: tile << 24 | the
// depth key of the pair's slot.
hipError_t launch_widen_keys(const uint32_t* sorted_tiles, const uint32_t* values, const uint32_t* depth_keys,
                             uint64_t pairs, uint64_t* keys, hipStream_t s);
hipError_t launch_blend(const DView* views_dev, const uint32_t* view_of_tile_base, uint32_t nviews,
                        uint32_t total_tiles, uint32_t tile, uint32_t batch, const uint32_t* ranges,
                        const uint32_t* pair_slots, const float* records, uint32_t depth_mode,
                        float depth_threshold, hipStream_t s);
hipError_t launch_blend_backward(const DView* views_dev, const uint32_t* tile_bases, uint32_t nviews,
                                 uint32_t total_tiles, uint32_t tile, uint32_t batch, const uint32_t* ranges,
                                 const uint32_t* pair_slots, const float* records, float* grad2d, hipStream_t s);
hipError_t launch_project_backward(const ProjectArgs& a, const DView* views_dev, uint32_t nviews,
                                   const uint32_t* touched, const float* records, const float* grad2d,
                                   const athx_grads& g, hipStream_t s);
hipError_t launch_counters(const uint32_t* depth_keys, const uint32_t* touched, uint64_t slots,
                           uint32_t* words, hipStream_t s);

// Onesweep radix sort of (uint32 or uint64 key, uint32 value) pairs.
struct SortScratch {
    void* keys_alt;  // keys of the sort's width
    uint32_t* values_alt;
    uint32_t* hist;      // passes * 256
    uint32_t* lookback;  // passes * partitions * 256
    uint32_t* counters;  // passes
    uint32_t* partials;  // passes * 256 * sort_hist_blocks: each histogram block's counts
    uint32_t* counts;    // 256 * partitions: a pass's digit counts, digit-major
    uint32_t* offsets;   // 256 * partitions: their exclusive scan
    uint32_t* scan_status;  // scan_status_words(256 * partitions)
    uint64_t* scan_total;
};
uint32_t sort_partitions(uint32_t n);
uint32_t sort_hist_blocks(uint32_t n);
// Sorts the first min(*n_dev, cap) pairs stably by key bits [lo_bit,
// key_bits); scratch is sized for cap and (key_bits - lo_bit) / 8 passes.
// in_alt non-null: an odd number of passes leaves the result in the alternate
// buffers (sc.keys_alt, sc.values_alt) and *in_alt says so, where it would
// otherwise be copied back.
hipError_t launch_sort(uint64_t* keys, uint32_t* values, const uint64_t* n_dev, uint32_t cap, uint32_t lo_bit,
                       uint32_t key_bits, const SortScratch& sc, hipStream_t s, bool* in_alt = nullptr);
hipError_t launch_sort(uint32_t* keys, uint32_t* values, const uint64_t* n_dev, uint32_t cap, uint32_t lo_bit,
                       uint32_t key_bits, const SortScratch& sc, hipStream_t s, bool* in_alt = nullptr);
// The visible slots in slot order with their depth keys; the count lands in
// *visible_total (device).
hipError_t launch_compact(const uint32_t* touched, const uint32_t* depth_keys, uint32_t slots, uint32_t* keys_out,
                          uint32_t* slots_out, uint64_t* visible_total, uint32_t* status_words, hipStream_t s);
// Tile counts of the depth-sorted slots, scanned: slot_offsets[slot] is its
// first pair; the pair total lands in *pair_total (device).
hipError_t launch_gather_scan(const uint32_t* touched, const uint32_t* sorted_slots, const uint64_t* visible_dev,
                              uint32_t max_visible, uint32_t* slot_offsets, uint64_t* pair_total,
                              uint32_t* status_words, hipStream_t s);

// Per-pair 2D gradient words the blend's adjoint accumulates a slot:
// centre x, y, conic a, b, c, alpha, colour r, g, b.
constexpr uint32_t kGrad2d = 9;

}  // namespace athx
