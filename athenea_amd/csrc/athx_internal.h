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
    uint32_t tile_base;   // first global tile of this view
    uint32_t strips_x;    // strips a tile row: div_up(tiles_x, kStripTiles)
    uint32_t strip_base;  // first global strip of this view
    float m[12];
    float eye[3];
    float bg[4];
    // The projection's Jacobian clamp, 1.3x the frustum in screen units:
    // (-0.15 W - cx) / fx, (1.15 W - cx) / fx, then the same for y with H,
    // cy, fy (the host's float operations are the kernels' own).
    float lim[4];
    float* colour;
    float* depth;
    float* final_t;
    uint32_t* last;
    const float* dl_dcolour;
    // The walk's variants (athx_targets): null when off.
    float* crypto;
    uint32_t* trace;
    const float* under_colour;
    const float* under_depth;
    const uint32_t* under_crypto;
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
    // Where this cloud's Gaussians sit among the frame's slots: Gaussian i of
    // view v is slot v * slot_stride + slot_base + i. One cloud: stride n,
    // base 0; several instances (athx_render_instances) share each view's
    // slot range, each from its own base (TileRasterizer.cpp:896-898).
    uint32_t slot_stride;
    uint32_t slot_base;
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
    // A matte frame: each drawn slot's Cryptomatte id (crypto_ids[i], or 0
    // without them) into slot_ids.
    const uint32_t* crypto_ids;
    uint32_t* slot_ids;
};

__host__ __device__ inline uint32_t div_up(uint32_t a, uint32_t b) { return (a + b - 1) / b; }
__host__ __device__ inline uint64_t div_up64(uint64_t a, uint64_t b) { return (a + b - 1) / b; }

__device__ inline float fmin_(float a, float b) { return a < b ? a : b; }
__device__ inline float fmax_(float a, float b) { return a > b ? a : b; }
__device__ inline float clamp_(float x, float lo, float hi) { return fmin_(fmax_(x, lo), hi); }

// frame.slang:354-371: the least power a * x^2 + 2 b x y + c y^2 takes
// over the box [lox, hix] x [loy, hiy]: zero when the box holds the centre,
// else the least over its four edges, each at the edge's clamped 1D
// minimiser ((-b) x) / c, an IEEE divide as the reference writes it
// (-b * x / c; the library is built with correctly rounded division).
__device__ inline float min_power_over_box(float a, float b, float c, float lox, float loy, float hix, float hiy) {
    if (lox <= 0.0f && hix >= 0.0f && loy <= 0.0f && hiy >= 0.0f) {
        return 0.0f;
    }
    float best = 3.4e38f;
    for (int e = 0; e < 2; ++e) {
        const float x = e == 0 ? lox : hix;
        const float y = c > 1e-12f ? clamp_(((-b) * x) / c, loy, hiy) : loy;
        best = fmin_(best, a * x * x + 2.0f * b * x * y + c * y * y);
    }
    for (int e = 0; e < 2; ++e) {
        const float y = e == 0 ? loy : hiy;
        const float x = a > 1e-12f ? clamp_(((-b) * y) / a, lox, hix) : lox;
        best = fmin_(best, a * x * x + 2.0f * b * x * y + c * y * y);
    }
    return best;
}

// The reciprocals the strip classifier's slice extents take, formed once a
// footprint with the hardware reciprocal (within an ulp of 1 / v): they only
// place the extents, whose margin is hundreds of ulps, never a tile test.
__device__ inline float recip_of(float v) { return v > 1e-12f ? __builtin_amdgcn_rcpf(v) : 0.0f; }

// MI355X deals a launch's workgroups to its 8 XCDs round robin (workgroup b
// to XCD b % 8), and each XCD has its own L2. Work item xcd_band(b, n) gives
// XCD x the x-th eighth of n items, in order, so items that read the same
// data (neighbouring tiles, a strip's neighbours) share an L2.
constexpr uint32_t kXcds = 8;
__device__ inline uint32_t xcd_band(uint32_t b, uint32_t n) {
    const uint32_t xcd = b % kXcds, local = b / kXcds;
    const uint32_t per = n / kXcds, extra = n % kXcds;
    return xcd * per + (xcd < extra ? xcd : extra) + local;
}

// The exact tile counts' working set (launch_exact_counts; the frame itself
// bins its pairs instead): each footprint's rectangle, and which of its
// tiles it touches.
//
// A footprint over kBigRect tiles is walked in steps of one hit word each
// (big_steps); the big footprints are listed, and their words placed, by a
// scan over the slots (launch_big_list), so no kernel takes a place with an
// atomic on one shared address -- those serialise in the memory system.
struct Footprints {
    uint2* rects;             // [slots] tile rect: x0 | y0 << 16, w | h << 16 (0, 0 when culled before),
                              // recomputed from the records for the exact counts
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

// The frame bins its pairs in two levels. A strip is kStripTiles tiles of one
// tile row of one view; strips are numbered view by view, row by row, left
// to right, so a strip's tiles are consecutive global tiles and the strips'
// order is the tiles' order. Each drawn footprint lists one strip entry for
// every (row, strip) its tile rectangle covers, in depth order; a stable
// sort on the strip puts each strip's entries in depth order; and each run
// of kStripSeg entries of one strip (a segment) tests its entries against
// the strip's tiles, counts the hits of every tile, and -- after one scan
// of those counts, laid out strip by strip, tile by tile, segment by
// segment -- writes each hit straight to its place in the tile-sorted pair
// list. That list is the reference's sort of (tile << 24 | depth) keys
// (TileRasterizer.cpp:1138-1150), without the sort over the pairs.
constexpr uint32_t kStripTiles = 16;
constexpr uint32_t kStripSeg = 1024;  // entries a segment: a workgroup of 256 lanes, 4 entries a lane
// A slot's word in `touched` after the projection: its strip entries, with
// kUnsettled set while the tile under its centre is not touched (whether
// any tile is, is settled by launch_project's second kernel).
constexpr uint32_t kUnsettled = 1u << 31;

// Launchers (each returns the hipError_t of its launches).
// The projection: records, depth keys and each slot's strip entries (0:
// not drawn, with the cull in its depth key); a footprint whose centre tile
// is not touched is left unsettled (kUnsettled) for launch_settle_count, a
// wavefront each walking its rectangle until a tile is
// (splat_project.slang:764-776's zero count is the cull no-tile).
hipError_t launch_project(const ProjectArgs& a, const DView* views_dev, uint32_t nviews, float* records,
                          uint32_t* touched, uint32_t* depth_keys, hipStream_t s);
// The exact tile count of every slot (splat_project.slang:764-776), for the
// readers of the frame (copy_frame, the counters) -- not on the frame's
// path: small rectangles spread over the lanes of a wavefront, big ones
// listed by a scan and walked in steps.
hipError_t launch_exact_counts(const DView* views_dev, uint32_t nviews, uint32_t n, uint32_t tile,
                               const float* records, uint32_t* counts, uint32_t* depth_keys, const Footprints& fp,
                               uint32_t* status_words, hipStream_t s);
// The strip entries of the depth-sorted slots (key: strip, value: depth
// rank), rank k's from offsets[k / 64] on (every 64th rank's first entry),
// entries at or past cap not written; and each rank's footprint, 32 bytes
// (fp[2k], fp[2k + 1]) the binning reads.
hipError_t launch_strip_emit(const DView* views_dev, uint32_t n, uint32_t tile, const uint32_t* sorted_slots,
                             const uint32_t* offsets, const uint64_t* visible_dev, uint32_t max_visible, uint32_t cap,
                             const float* records, uint4* fp, uint32_t* strip_keys, uint32_t* strip_ranks,
                             hipStream_t s);
// The strip-sorted entries binned into tile-sorted pairs.
struct StripBins {
    const DView* views;
    const uint32_t* strip_bases;  // [nviews + 1] first strip of each view, then the strips
    uint32_t nviews;
    uint32_t n;                   // Gaussians (slot = view * n + Gaussian)
    uint32_t tile;
    uint32_t strips;
    const float* records;
    uint32_t* strip_slots;         // [entries] strip-sorted entries' depth ranks, made slots by the count
                                   // (slots already when strip_keys is set)
    const uint32_t* strip_keys;    // [entries] sorted keys carrying the masks (strip_mask_emit), or null
    const uint4* fp;               // [visible][2] each rank's footprint (launch_strip_emit)
    const uint32_t* strip_ranges;  // [strips][2] each strip's entries (0, 0 when none)
    const uint32_t* segbase;       // [strips] first segment of each strip
    const uint4* seg_desc;         // [seg_cap][2] each segment's entries, table column, strip place
    const uint64_t* seg_total;     // segments (device)
    uint32_t seg_cap;              // segments the table holds
    uint16_t* masks;               // [entries] tiles of its strip an entry touches
    uint32_t masks_sorted;         // the strip sort wrote masks (strip_keys unused by the count)
    uint32_t* table;               // [16 * seg_cap] hit counts, then their exclusive scan
    const uint64_t* pair_total;    // pairs (device; the table's total)
    uint32_t pair_cap;
    uint32_t* pair_slots;          // [pair_cap] the tile-sorted pairs' slots
};
// The strip entries with their tiles decided (frames of at most 65 536
// strips): key = strip | 16-bit mask of its touched tiles << 16, value = the
// slot; the binning then reads only these. view_entry, given when the ranks
// are grouped by view, gets each view's first entry (from its first rank).
hipError_t launch_strip_mask_emit(const DView* views_dev, uint32_t n, uint32_t tile, const uint32_t* sorted_slots,
                                  const uint32_t* offsets, const uint64_t* visible_dev, uint32_t max_visible,
                                  uint32_t cap, const float* records, uint32_t* strip_keys, uint32_t* strip_slots,
                                  uint32_t* rank_tiles, hipStream_t s, uint32_t* view_entry = nullptr);
// Each strip's segments, scanned: segbase[s] and the total in *seg_total.
hipError_t launch_seg_scan(const uint32_t* strip_ranges, uint32_t strips, uint32_t* segbase, uint64_t* seg_total,
                           uint32_t* status_words, hipStream_t s);
// Each segment's descriptor (its entries, its first table cell, its strip's
// view, row and first column), a wavefront a strip; the segments past the
// total up to the capacity get empty ones, so a workgroup of the binning
// starts from one load.
hipError_t launch_seg_desc(const StripBins& b, uint4* desc, hipStream_t s);
// Tests and counts every segment's entries (masks, table).
hipError_t launch_strip_count(const StripBins& b, hipStream_t s);
// The table's exclusive scan in place; the pair total lands in *pair_total.
hipError_t launch_table_scan(uint32_t* table, const uint64_t* seg_total, uint32_t seg_cap, uint64_t* pair_total,
                             uint32_t* status_words, hipStream_t s);
// Every tile's pair range (0, 0 when empty), from the scanned table.
hipError_t launch_tile_ranges(const StripBins& b, const uint32_t* tile_bases, uint32_t tiles, uint32_t* ranges,
                              hipStream_t s);
// The pairs' slots at their places.
hipError_t launch_strip_pairs(const StripBins& b, hipStream_t s);
// Lists the big footprints (rectangles over kBigRect tiles) in slot order
// with each one's first step; the count and the steps land in *fp.big_total.
hipError_t launch_big_list(const Footprints& fp, uint32_t slots, uint32_t* status_words, hipStream_t s);
hipError_t launch_fill_u32(uint32_t* p, uint32_t value, uint64_t n, hipStream_t s);
size_t scan_status_words(uint32_t n);
hipError_t launch_scan(const uint32_t* in, uint32_t* out, uint64_t* total, uint32_t n, uint32_t* status,
                       hipStream_t s);
// [start, end) of every key in 0..keys-1 among the first min(*n_dev, cap)
// sorted keys (compared under key_mask); (0, 0) for a key that is absent.
hipError_t launch_ranges(const uint32_t* sorted_keys, const uint64_t* n_dev, uint32_t cap, uint32_t keys,
                         uint32_t* ranges, hipStream_t s, uint32_t key_mask = ~0u);
// The reference's 64-bit pair keys of the tile-sorted pairs: tile << 24 |
// the depth key of the pair's slot, a workgroup a tile over its range.
hipError_t launch_pair_keys(const uint32_t* ranges, uint32_t tiles, const uint32_t* values,
                            const uint32_t* depth_keys, uint64_t* keys, hipStream_t s);
hipError_t launch_blend(const DView* views_dev, const uint32_t* view_of_tile_base, uint32_t nviews,
                        uint32_t total_tiles, uint32_t tile, uint32_t batch, const uint32_t* ranges,
                        const uint32_t* pair_slots, const float* records, uint32_t depth_mode, float depth_threshold,
                        hipStream_t s, bool variants = false, const uint32_t* slot_ids = nullptr);
// What each splat gave the last frame's walk (the frame's ranges, pair
// slots and records): mode 1 the blend weight in 1/4096, one word a slot;
// mode 4 the colour gradient against each view's dl_dcolour in 1/256 (three
// signed words) and the weight in 1/256, four words a slot. Adds into out.
hipError_t launch_contribution(const DView* views_dev, const uint32_t* tile_bases, uint32_t nviews,
                               uint32_t total_tiles, uint32_t tile, uint32_t batch, const uint32_t* ranges,
                               const uint32_t* pair_slots, const float* records, uint32_t mode, uint32_t* out,
                               hipStream_t s);
hipError_t launch_blend_backward(const DView* views_dev, const uint32_t* tile_bases, uint32_t nviews,
                                 uint32_t total_tiles, uint32_t tile, uint32_t batch, const uint32_t* ranges,
                                 const uint32_t* pair_slots, const float* records, float* grad2d, hipStream_t s);
hipError_t launch_project_backward(const ProjectArgs& a, const DView* views_dev, uint32_t nviews,
                                   const uint32_t* touched, const float* records, const float* grad2d,
                                   const athx_grads& g, hipStream_t s);
// SplatCounters words (TileRasterizer.h:370-401; splat_frame_counters.slang):
// [0] slots (the host's), [1] visible, [2] pairs low, [3] pairs high, [4]
// most tiles of one footprint, [8 + r] culls of reason r, and for cloud j <
// ATHX_MAX_COUNTED_CLOUDS [16 + 2j] its drawn footprints, [17 + 2j] their pairs.
constexpr uint32_t kCounterWords = 16 + 2 * ATHX_MAX_COUNTED_CLOUDS;

// What the counters read: every slot's depth key (a culled one holds its
// reason), the depth-ordered drawn slots, and each drawn footprint's tiles
// -- by rank (rank_tiles) or by slot (slot_tiles); the instances' bases in a
// view's slots, base[clouds] the last counted one's end.
struct CounterArgs {
    const uint32_t* depth_keys;
    const uint32_t* sorted_slots;
    const uint32_t* rank_tiles;
    const uint32_t* slot_tiles;
    const uint64_t* visible_dev;
    uint32_t slots, count, clouds;
    uint32_t base[ATHX_MAX_COUNTED_CLOUDS + 1];
};
hipError_t launch_frame_counters(const CounterArgs& a, uint32_t* words, hipStream_t s);

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
// A multi-view frame's strip entries come out of the emission view after
// view (the ranks are grouped by view), so a view's entries are one run of
// the entry array: a segment. The segmented strip sort cuts each run into
// partitions of its own (a partition holds one view's entries) and orders
// each run by the view's own strip number (key & 0xFFFF minus seg_base) -- a
// pass or two where the frame's whole strip number would take more. Filled
// on the device by launch_strip_segments.
constexpr uint32_t kSegmentPart = 8192;
struct SegmentArgs {
    const uint32_t* part_seg;    // partition -> view
    const uint32_t* seg_start;   // a view's first entry
    const uint32_t* seg_len;     // its entries (those the capacity holds)
    const uint32_t* seg_part;    // its first partition
    const uint32_t* seg_base;    // its first strip
    const uint32_t* parts;       // partitions in use
};
// Sorts the segments' (strip key, slot) pairs stably by the low `bits` bits
// of the key's strip relative to its view's first: view after view, each by
// strip, in the input's order within a strip; in the alternate buffers when
// in_alt ends true. parts_cap bounds the partitions (div_up(cap, part) + a
// partial one a view).
// strips false: the digits are the key's own low `bits` bits (the depth
// sort's segments, the frame's views' visible slots), seg_base unused.
// mask_out (strips only): the last pass writes each key's upper half (its
// tile mask) there instead of the key itself.
hipError_t launch_segment_sort(uint32_t* keys, uint32_t* values, const SegmentArgs& g, uint32_t parts_cap,
                               uint32_t bits, const SortScratch& sc, hipStream_t s, bool& in_alt,
                               bool strips = true, uint16_t* mask_out = nullptr);
// A one-pass segmented strip sort's strip ranges ([strips][2], (0, 0) when
// empty) from its scanned counts (radix digits a partition).
hipError_t launch_strip_ranges_seg(const uint32_t* offsets, const SegmentArgs& g, uint32_t nviews,
                                   const uint32_t* strip_bases, uint32_t strips, uint32_t radix,
                                   const uint64_t* total_dev, uint32_t* ranges, hipStream_t s);
constexpr uint32_t kMaxSegmentViews = 256;
struct SegmentBuild {
    uint32_t nviews;
    // view_first null: view_entry holds every view's start. Else view v's
    // first rank is view_first[v * view_stride], and view_entry holds the
    // first entry of each view with ranks (written by the emission).
    const uint32_t* view_first;
    uint32_t view_stride;
    const uint32_t* view_entry;
    const uint64_t* visible_dev;
    uint32_t max_visible;
    const uint64_t* entry_total_dev;
    uint32_t cap;        // entries the strip arrays hold
    uint32_t parts_cap;
    uint32_t* part_seg;
    uint32_t* seg_start;
    uint32_t* seg_len;
    uint32_t* seg_part;
    uint32_t* parts;
    uint64_t* binned_total;
};
hipError_t launch_strip_segments(const SegmentBuild& g, hipStream_t s);
// The compaction's block of slots: its first step counts each block's
// drawn slots into a u64 a block at the head of its status words.
constexpr uint32_t kCompactTile = 4096;
// The unsettled footprints settled (launch_project leaves them), and each
// block's drawn slots counted into sums (the compaction's status words).
hipError_t launch_settle_count(const DView* views_dev, uint32_t n, uint32_t tile, uint32_t slots, const float* records,
                               uint32_t* touched, uint32_t* depth_keys, uint64_t* sums, hipStream_t s);
// The visible slots in slot order with their depth keys (each key's top byte
// holding the slot's strip-entry count, capped at 255); the count lands in
// *visible_total (device). counted: the blocks' counts are already in the
// status words (launch_settle_count). With view_start (per_view >= 2: slot =
// view * per_view + Gaussian), view v's first compacted position lands in
// view_start[v].
hipError_t launch_compact(const uint32_t* touched, const uint32_t* depth_keys, uint32_t slots, uint32_t* keys_out,
                          uint32_t* slots_out, uint64_t* visible_total, uint32_t* status_words, hipStream_t s,
                          bool counted = false, uint32_t per_view = 0,
                          uint32_t* view_start = nullptr);
// Strip entries of the depth-sorted slots, scanned in that order:
// offsets[k] is rank k's first entry; the total lands in *entry_total.
hipError_t launch_entry_scan(const uint32_t* touched, const uint32_t* sorted_slots, const uint32_t* sorted_keys,
                             const uint64_t* visible_dev, uint32_t max_visible, uint32_t* offsets,
                             uint64_t* entry_total, uint32_t* status_words, hipStream_t s);
// Tile counts of the depth-sorted slots, scanned: slot_offsets[slot] is its
// first pair; the pair total lands in *pair_total (device).
hipError_t launch_gather_scan(const uint32_t* touched, const uint32_t* sorted_slots, const uint64_t* visible_dev,
                              uint32_t max_visible, uint32_t* slot_offsets, uint64_t* pair_total,
                              uint32_t* status_words, hipStream_t s);

// Per-pair 2D gradient words the blend's adjoint accumulates a slot:
// centre x, y, conic a, b, c, alpha, colour r, g, b.
constexpr uint32_t kGrad2d = 9;

// A PLY vertex's fields the ingest reads, each as its byte offset in the
// vertex | its athx_ply_type << 24: x y z, opacity, scale 0..2, rot 0..3
// (w x y z), f_dc 0..2, then from kIngestBase the kept rest coefficients
// in the order they are packed (rgb a basis).
constexpr uint32_t kIngestBase = 14;
constexpr uint32_t kIngestCodes = kIngestBase + 45;
struct IngestFields {
    uint32_t count;       // vertices
    uint32_t stride;      // bytes a vertex
    uint32_t rest;        // rest coefficients kept a splat (3 a basis)
    uint32_t sh_words;    // words a splat in the output's sh
    uint32_t keep_faint;  // keep every vertex with a finite position
    // athx_ply_layout's encodings of opacity, scale, colour and rotation.
    uint32_t opacity_mode, scale_mode, colour_mode, rotation_mode;
    uint32_t code[kIngestCodes];
};
// The ground truth (reference.hip): the reference's own projection into
// records of the rasterizer's layout, keys (the view depth's bits, all ones
// when culled) and values (the index) for a stable sort; then every drawn
// splat blended at every pixel in that order (overflow: pixels with more
// than max_hits contributors); and the image comparisons.
hipError_t launch_ref_project(const float* positions, const uint32_t* shape, const uint32_t* sh, uint32_t n,
                              uint32_t rest_per_colour, uint32_t sh_limit, uint32_t antialias, uint32_t linear,
                              const DView* view, float* records, uint32_t* keys, uint32_t* values, hipStream_t s,
                              uint32_t base = 0);
hipError_t launch_ref_blend(const DView* view, uint32_t width, uint32_t height, const float* records,
                            const uint32_t* sorted_keys, const uint32_t* sorted, uint32_t n, uint32_t max_hits,
                            uint32_t depth_mode, float depth_threshold, uint32_t* overflow, hipStream_t s);
hipError_t launch_compare_codes(const float* a, const float* b, uint32_t width, uint32_t height, uint32_t* rows,
                                uint32_t* bins, hipStream_t s);
hipError_t launch_compare_hdr(const float* a, const float* b, uint32_t width, uint32_t height, uint32_t* rows,
                              float* row_error, uint32_t* bins, float* total, hipStream_t s);

// A sky prepared on the device (sky.hip): texels of the dome for this base
// side (sides and offsets of the eight levels when not null), and the
// launch that writes the texel words and the 16 float4 coefficients.
uint32_t sky_texels(uint32_t base_side, uint32_t* sides, uint32_t* offsets);
// The sun's extraction (sky.hip): a level down, the histogram and peak, the
// sun's sums, and the residual equirect. rows: (sin theta, cos theta) a row;
// cols: (cos phi, sin phi) a column; row_weight: sin theta in 2^-24 units.
hipError_t launch_sky_half(const float* src, uint32_t w, uint32_t h, float* dst, uint32_t dw, uint32_t dh,
                           hipStream_t s);
hipError_t launch_sky_sun_scan(const float* img, uint32_t w, uint32_t h, const uint32_t* row_weight,
                               unsigned long long* hist, unsigned long long* peak, hipStream_t s);
hipError_t launch_sky_sun_sum(const float* img, uint32_t w, uint32_t h, const float* rows, const float* cols,
                              float area, float threshold, const float axis[3], float cos_widest, double* sums,
                              hipStream_t s);
hipError_t launch_sky_residual(const float* img, uint32_t w, uint32_t h, const float* rows, const float* cols,
                               float threshold, const float axis[3], float cos_widest, float* out, hipStream_t s);
hipError_t launch_prepare_sky(const float* equirect, uint32_t width, uint32_t height, uint32_t base_side,
                              uint32_t* texels, float* sh, hipStream_t s);

// Each vertex's keep flag (1 or 0).
hipError_t launch_ingest_flags(const uint8_t* vertices, const IngestFields& f, uint32_t* keep, hipStream_t s);
// The kept vertices decoded and packed at dest[i] (the flags' exclusive scan).
hipError_t launch_ingest_pack(const uint8_t* vertices, const IngestFields& f, const uint32_t* keep,
                              const uint32_t* dest, const athx_ingest_out& out, hipStream_t s);

}  // namespace athx
