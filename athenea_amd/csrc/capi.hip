// capi.hip -- the C ABI of include/athx.h: a context that owns the frame's
// scratch and drives the kernels on the caller's stream.
//
// The frame (modules/render/src/TileRasterizer.cpp:296-1295, reordered for
// one 64-bit key sort):
//   project      every (view, Gaussian) slot: record, tile count, depth key
//   scan         exclusive scan of tile counts -> pair offsets and the total
//   readback     the pair total (the reference's one wait, :1093-1102)
//   emit         (global tile << 24 | depth24, slot) pairs in slot order
//   sort         stable onesweep radix sort over 24 + tile bits
//   ranges       each tile's [start, end) in the sorted pairs
//   blend        per tile, front to back
// Scratch grows to 1.5x what a frame needs and, after athx_release_scratch,
// to exactly what it needs (TileRasterizer.cpp:30-36, 178-191).
#include <algorithm>
#include <array>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <exception>
#include <limits>
#include <cmath>
#include <new>
#include <sstream>
#include <string_view>
#include <string>
#include <vector>

#include "athx_internal.h"

using namespace athx;

namespace {

struct DeviceBuffer {
    void* ptr = nullptr;
    size_t bytes = 0;
};

// Hit words kept for big footprints: one a slot is room for 64 tiles a slot
// on average; a footprint past the room has its tiles tested again.
uint32_t big_mask_words(uint32_t slots) { return slots + 1024; }

constexpr uint32_t kDepthPasses = 3;  // 24-bit depth keys (splat_project.slang:1129; TileRasterizer.cpp:1122)

// The frame's totals, u64 words on the device (and pinned on the host):
// pairs, visible slots, the exact counts' big footprints, strip entries,
// segments, the exact counts' sum.
constexpr uint32_t kTotalPairs = 0, kTotalVisible = 1, kTotalBig = 2, kTotalEntries = 3, kTotalSegs = 4;
constexpr uint32_t kTotalExact = 5;  // the exact counts' total (copy_frame's offsets)
constexpr uint32_t kTotalBinned = 6;  // the entries a segmented frame's strip sort holds
constexpr uint32_t kTotalOrder = 7;   // the visible slots of the whole depth order (copy_frame's offsets)
constexpr uint32_t kTotalDepthBinned = 8;  // the visible slots a grouped frame's depth segments hold
constexpr uint32_t kTotalWords = 9;
constexpr uint64_t kMaxPairs = uint64_t{1} << 30;  // pairs and entries are 32-bit, with room

// The scratch of athx_exclusive_scan_noalloc: the scan's look-back status
// words, or a long scan's partition sums and their runs' tops, for up to
// kNoAllocScanMax elements (two words a partition of 4096, two a run of 1024
// partitions and four more: scan_status_words), held by each device the library runs on
// until the process ends. Host code never names it directly: athx_create
// looks up its device address (hipGetSymbolAddress) and keeps it in the
// context.
constexpr uint32_t kNoAllocScanMax = 1u << 24;
constexpr uint32_t kNoAllocScanWords = 2u * (2u + kNoAllocScanMax / 4096u + kNoAllocScanMax / 4096u / 1024u);
static __device__ __align__(256) uint32_t scratch[kNoAllocScanWords];

uint32_t bits_for(uint32_t values) {
    // modules/render/src/FrameParams.h:18-24
    uint32_t bits = 1;
    while (bits < 32 && (uint64_t{1} << bits) < values) {
        ++bits;
    }
    return bits;
}

}  // namespace

struct athx_ctx {
    int device = 0;
    std::string error;
    bool exact_growth = false;

    DeviceBuffer views, tile_bases, records, touched, depth_keys, offsets, scan_status, totals;
    DeviceBuffer pair_vals, ranges, grad2d, counter_words, big_list, rank_tiles, slot_ids;
    DeviceBuffer sky_levels, sky_tables, sky_words;  // athx_sky_sun / athx_sky_residual
    DeviceBuffer depth_keys_sorted, depth_slots, depth_alt_keys, depth_alt_slots, depth_sort_meta;
    DeviceBuffer segments;  // a multi-view frame's strip segments (SegmentBuild's outputs)
    DeviceBuffer depth_segments;  // and its depth sort's, with each view's first compacted position
    DeviceBuffer order_keys, order_slots, order_alt_keys, order_alt_slots, order_meta;  // the whole depth order
    DeviceBuffer fp_rects, fp_masks, fp_big_base, fp_big_masks;
    // The binning: strip entries (keys, slots, the sort's alternates and
    // metadata), each strip's range and first segment, the entries' hit
    // masks and the per (tile, segment) table.
    DeviceBuffer strip_keys, strip_slots, strip_alt_keys, strip_alt_slots, strip_sort_meta;
    DeviceBuffer strip_ranges, segbase, seg_map, masks, seg_table, rank_fp;
    // For the frame's readers: exact tile counts and their scan in depth order.
    DeviceBuffer exact, exact_offsets, exact_status;
    DeviceBuffer api_sort_meta, api_alt_keys, api_alt_vals, api_scan_status, api_count;
    // The PLY ingest: each vertex's keep flag, their scan, the kept total.
    DeviceBuffer ingest_keep, ingest_dest, ingest_total;
    // The ground truth and the image comparisons.
    DeviceBuffer ref_records, ref_keys, ref_values, ref_alt_keys, ref_alt_values, ref_sort_meta, ref_count;
    DeviceBuffer cmp_rows, cmp_row_error, cmp_bins;
    uint64_t* host_totals = nullptr;  // pinned: the totals words (kTotal*)
    DView* host_views = nullptr;      // pinned staging for the views
    uint32_t* host_tile_bases = nullptr;  // [nviews + 1] tile bases, then [nviews + 1] strip bases
    uint32_t host_view_capacity = 0;

    // What the last render left for the adjoint and the frame's readers.
    uint32_t last_n = 0, last_nviews = 0, last_tiles = 0, last_tile = 0, last_key_bits = 0, last_capacity = 0;
    uint32_t last_strips = 0, last_entry_cap = 0;
    uint32_t last_instances = 0;  // the clouds the last frame projected (athx_render_instances)
    bool last_under = false;      // the last frame's walks ended at an under layer
    bool last_sorted_in_alt = false;  // the ranked slots are in depth_alt_slots
    // The last frame sorted each view's slots apart (its ranks grouped by
    // view): the exact offsets, scanned in the frame's whole depth order,
    // then sort that order on request (order_* buffers).
    bool last_grouped = false;
    uint64_t last_true_pairs = 0;  // the last frame's pair total (an async frame's once athx_frame_pairs ran)
    uint64_t pair_hint = 0;        // the pair capacity a frame takes before it knows: 1.25x the largest seen
    uint64_t entry_hint = 0;       // the same for strip entries
    uint64_t pair_last = 0, entry_last = 0;  // 1.25x the last eager frame's: what the next one is queued with
    hipEvent_t async_done = nullptr;
    bool async_pending = false;
    bool have_frame = false;
    // Numbers the frames rendered on this context (1, 2, ...): the adjoint
    // and the contributions name the frame they follow, so a later render
    // that replaced its buffers is refused rather than walked.
    uint64_t serial = 0;
    // The counted frames' words on their way to the host: a ring of three
    // pinned copies, each made on its frame's stream after the counts and
    // marked by an event (TileRasterizer's AsyncReadback of 3); seq 0 is an
    // empty slot, and the newest finished one is what latestCounters reads.
    struct CounterCopy {
        uint32_t* host = nullptr;  // kCounterWords, pinned
        hipEvent_t done = nullptr;
        uint64_t seq = 0, tag = 0;
        uint32_t slots = 0, clouds = 0;
    };
    CounterCopy counter_ring[3];
    uint64_t counter_seq = 0;

    // Device address of this device's `scratch` (athx_exclusive_scan_noalloc).
    uint32_t* noalloc_scratch = nullptr;

    hipEvent_t ev[8] = {};
    // Recorded after each upload from the pinned staging views: the staging
    // is not written again until that copy has run.
    hipEvent_t upload_done = nullptr;
    bool upload_pending = false;
};

namespace {

std::vector<DeviceBuffer*> all_buffers(athx_ctx* ctx) {
    return {&ctx->views, &ctx->tile_bases, &ctx->records, &ctx->touched, &ctx->depth_keys, &ctx->offsets,
            &ctx->scan_status, &ctx->totals, &ctx->pair_vals, &ctx->ranges, &ctx->grad2d, &ctx->counter_words,
            &ctx->rank_tiles, &ctx->slot_ids, &ctx->sky_levels, &ctx->sky_tables, &ctx->sky_words,
            &ctx->big_list, &ctx->depth_keys_sorted, &ctx->depth_slots, &ctx->depth_alt_keys, &ctx->depth_alt_slots,
            &ctx->depth_sort_meta, &ctx->fp_rects, &ctx->fp_masks, &ctx->fp_big_base, &ctx->fp_big_masks,
            &ctx->strip_keys, &ctx->strip_slots, &ctx->strip_alt_keys, &ctx->strip_alt_slots, &ctx->strip_sort_meta,
            &ctx->segments, &ctx->depth_segments, &ctx->order_keys, &ctx->order_slots, &ctx->order_alt_keys,
            &ctx->order_alt_slots, &ctx->order_meta,
            &ctx->strip_ranges, &ctx->segbase, &ctx->seg_map, &ctx->rank_fp, &ctx->masks, &ctx->seg_table, &ctx->exact, &ctx->exact_offsets,
            &ctx->exact_status, &ctx->api_sort_meta, &ctx->api_alt_keys, &ctx->api_alt_vals, &ctx->api_scan_status,
            &ctx->api_count, &ctx->ingest_keep, &ctx->ingest_dest, &ctx->ingest_total, &ctx->ref_records,
            &ctx->ref_keys, &ctx->ref_values, &ctx->ref_alt_keys, &ctx->ref_alt_values, &ctx->ref_sort_meta,
            &ctx->ref_count, &ctx->cmp_rows, &ctx->cmp_row_error, &ctx->cmp_bins};
}

athx_status fail(athx_ctx* ctx, athx_status code, const std::string& msg) {
    if (ctx) {
        ctx->error = msg;
    }
    return code;
}

athx_status hip_fail(athx_ctx* ctx, hipError_t e, const char* what) {
    return fail(ctx, e == hipErrorOutOfMemory ? ATHX_OUT_OF_MEMORY : ATHX_DEVICE_FAILURE,
                std::string(what) + ": " + hipGetErrorString(e));
}

#define ATHX_HIP(ctx, expr, what)                 \
    do {                                          \
        hipError_t e_ = (expr);                   \
        if (e_ != hipSuccess) {                   \
            return hip_fail((ctx), e_, (what));   \
        }                                         \
    } while (0)

hipError_t reserve(athx_ctx* ctx, DeviceBuffer& b, size_t bytes) {
    if (bytes <= b.bytes) {
        return hipSuccess;
    }
    size_t want = ctx->exact_growth ? bytes : bytes + bytes / 2;
    want = (want + 255) & ~size_t(255);
    if (b.ptr) {
        // Queued work may still read the old buffer.
        hipError_t e = hipDeviceSynchronize();
        if (e != hipSuccess) {
            return e;
        }
        e = hipFree(b.ptr);
        b.ptr = nullptr;
        b.bytes = 0;
        if (e != hipSuccess) {
            return e;
        }
    }
    hipError_t e = hipMalloc(&b.ptr, want);
    if (e != hipSuccess) {
        b.ptr = nullptr;
        return e;
    }
    b.bytes = want;
    return hipSuccess;
}

void release(DeviceBuffer& b) {
    if (b.ptr) {
        (void)hipFree(b.ptr);
    }
    b.ptr = nullptr;
    b.bytes = 0;
}

template <typename T>
T* as(const DeviceBuffer& b) {
    return static_cast<T*>(b.ptr);
}

athx_status check_view(athx_ctx* ctx, const athx_view& v, uint32_t k) {
    if (v.width == 0 || v.height == 0) {
        return fail(ctx, ATHX_INVALID_ARGUMENT, "view " + std::to_string(k) + ": zero size");
    }
    if (v.width > 65535u || v.height > 65535u) {
        return fail(ctx, ATHX_INVALID_ARGUMENT, "view " + std::to_string(k) + ": width and height must be at most 65535");
    }
    if (!(v.focal_x > 0.0f) || !(v.focal_y > 0.0f)) {
        return fail(ctx, ATHX_INVALID_ARGUMENT, "view " + std::to_string(k) + ": focal length must be positive");
    }
    return ATHX_OK;
}

athx_status check_settings(athx_ctx* ctx, const athx_settings* s) {
    if (!s) {
        return fail(ctx, ATHX_INVALID_ARGUMENT, "settings is null");
    }
    if (s->tile != 8 && s->tile != 16 && s->tile != 32) {
        return fail(ctx, ATHX_INVALID_ARGUMENT, "tile must be 8, 16 or 32");
    }
    if (s->batch < 32 || s->batch > 256 || s->batch % 32 != 0) {
        return fail(ctx, ATHX_INVALID_ARGUMENT, "batch must be a multiple of 32 in [32, 256]");
    }
    return ATHX_OK;
}

athx_status check_gaussians(athx_ctx* ctx, const athx_gaussians* g) {
    if (!g) {
        return fail(ctx, ATHX_INVALID_ARGUMENT, "gaussians is null");
    }
    if (const athx_packed_splats* p = g->packed) {
        const uint32_t rest = p->rest_per_colour;
        if (rest != 0 && rest != 3 && rest != 8 && rest != 15) {
            return fail(ctx, ATHX_INVALID_ARGUMENT, "rest_per_colour must be 0, 3, 8 or 15 (packing.slang:454-456)");
        }
        if (g->count != 0 && (!p->positions || !p->shape || (rest != 0 && !p->sh))) {
            return fail(ctx, ATHX_INVALID_ARGUMENT, "a packed buffer is null");
        }
        if ((reinterpret_cast<uintptr_t>(p->positions) | reinterpret_cast<uintptr_t>(p->shape)) & 15u) {
            return fail(ctx, ATHX_INVALID_ARGUMENT, "packed positions and shape must be 16-byte aligned");
        }
    } else {
        if (g->sh_coeffs != 1 && g->sh_coeffs != 4 && g->sh_coeffs != 9 && g->sh_coeffs != 16) {
            return fail(ctx, ATHX_INVALID_ARGUMENT, "sh_coeffs must be 1, 4, 9 or 16");
        }
        if (g->count != 0) {
            if (!g->means || !g->log_scales || !g->quats || !g->opacity_logits) {
                return fail(ctx, ATHX_INVALID_ARGUMENT, "a Gaussian parameter pointer is null");
            }
            if ((g->sh == nullptr) == (g->sh_half == nullptr)) {
                return fail(ctx, ATHX_INVALID_ARGUMENT, "exactly one of sh and sh_half must be given");
            }
        }
    }
    if (g->count != 0) {
        if (const athx_relight* r = g->relight) {
            if (!r->pbr && (!r->metallic || !r->roughness)) {
                return fail(ctx, ATHX_INVALID_ARGUMENT, "relight needs metallic and roughness, or pbr words");
            }
            if (!r->sky.texels || !r->sky.sh || !r->sky.sun || r->sky.base_side == 0) {
                return fail(ctx, ATHX_INVALID_ARGUMENT, "relight needs a prepared sky (texels, sh, sun, base side)");
            }
            if (r->sky.base_side > 16384) {
                return fail(ctx, ATHX_INVALID_ARGUMENT, "sky base side over 16384");
            }
        }
    }
    return ATHX_OK;
}

uint32_t degree_of(uint32_t coeffs) { return coeffs >= 16 ? 3u : coeffs >= 9 ? 2u : coeffs >= 4 ? 1u : 0u; }

ProjectArgs project_args(const athx_gaussians* g, const athx_settings* s) {
    ProjectArgs a{};
    a.means = g->means;
    a.log_scales = g->log_scales;
    a.quats = g->quats;
    a.logits = g->opacity_logits;
    a.sh = g->sh;
    a.sh_half = g->sh_half;
    a.n = g->count;
    a.slot_stride = g->count;
    a.slot_base = 0;
    a.crypto_ids = g->crypto_ids;
    a.slot_ids = nullptr;
    a.sh_coeffs = g->sh_coeffs;
    if (const athx_packed_splats* p = g->packed) {
        a.packed = 1;
        a.positions = p->positions;
        a.shape = p->shape;
        a.packed_sh = p->sh;
        a.packed_words = (p->rest_per_colour * 3 + 1) / 2;  // packing.slang:456 (shWordsFor)
        a.sh_coeffs = p->rest_per_colour + 1;
    }
    const uint32_t held = degree_of(a.sh_coeffs);
    a.sh_degree = held < s->max_sh_degree ? held : s->max_sh_degree;
    a.linear = s->linear ? 1u : 0u;
    a.antialias = s->antialias ? 1u : 0u;
    a.tile = s->tile;
    if (const athx_relight* r = g->relight) {
        a.relit = 1;
        a.rl.normals = r->normals;
        a.rl.albedo = r->albedo;
        a.rl.metallic = r->metallic;
        a.rl.roughness = r->roughness;
        a.rl.normals_packed = r->normals_packed;
        a.rl.pbr = r->pbr;
        for (int k = 0; k < 12; ++k) {
            a.rl.o2w[k] = r->object_to_world[k];
        }
        a.rl.texels = r->sky.texels;
        a.rl.sky_sh = r->sky.sh;
        a.rl.sky_sun = r->sky.sun;
        a.rl.base_side = r->sky.base_side;
    }
    return a;
}

// One cloud of a frame and where it stands in each view (athx_instance).
struct FrameInstance {
    const athx_gaussians* g = nullptr;
    const float* object_to_view = nullptr;  // [nviews][12], or null: the view's own
    const float* eye = nullptr;             // [nviews][3], or null
};

// The projection's arguments of each instance: its own cloud, its slots
// from the counts of the instances before it.
std::vector<ProjectArgs> instance_args(const FrameInstance* inst, uint32_t ninst, const athx_settings* s,
                                       uint32_t count, uint32_t* slot_bases = nullptr) {
    std::vector<ProjectArgs> out;
    out.reserve(ninst);
    uint32_t base = 0;
    for (uint32_t j = 0; j < ninst; ++j) {
        ProjectArgs a = project_args(inst[j].g, s);
        a.slot_stride = count;
        a.slot_base = base;
        if (slot_bases) {
            slot_bases[j] = base;
        }
        base += inst[j].g->count;
        out.push_back(a);
    }
    return out;
}

// Fills the pinned staging views and tile bases; returns the frame's tiles.
// With several instances, the staging holds nviews views an instance, instance
// by instance, each with that instance's object-to-view rows and eye; the
// kernels after the projection read the first nviews (the screens, tiles and
// targets, which every instance shares).
athx_status stage_views(athx_ctx* ctx, const athx_view* views, uint32_t nviews, const athx_settings* s,
                        const athx_targets* targets, const float* const* dl, uint64_t& total_tiles,
                        uint64_t* strips_out = nullptr, const FrameInstance* inst = nullptr, uint32_t ninst = 1) {
    if (ctx->upload_pending) {
        ATHX_HIP(ctx, hipEventSynchronize(ctx->upload_done), "wait for the last upload");
        ctx->upload_pending = false;
    }
    const uint64_t staged64 = (uint64_t)nviews * (ninst ? ninst : 1);
    if (staged64 > (uint64_t{1} << 24)) {
        return fail(ctx, ATHX_UNSUPPORTED, "more than 2^24 (view, instance) pairs in one frame");
    }
    const uint32_t staged = (uint32_t)staged64;
    if (staged > ctx->host_view_capacity) {
        if (ctx->host_views) {
            (void)hipHostFree(ctx->host_views);
            ctx->host_views = nullptr;
        }
        if (ctx->host_tile_bases) {
            (void)hipHostFree(ctx->host_tile_bases);
            ctx->host_tile_bases = nullptr;
        }
        ctx->host_view_capacity = 0;
        ATHX_HIP(ctx, hipHostMalloc(reinterpret_cast<void**>(&ctx->host_views), sizeof(DView) * staged),
                 "pinned views");
        ATHX_HIP(ctx, hipHostMalloc(reinterpret_cast<void**>(&ctx->host_tile_bases),
                                    sizeof(uint32_t) * 2 * (staged + 1)),
                 "pinned tile bases");
        ctx->host_view_capacity = staged;
    }
    total_tiles = 0;
    uint64_t total_strips = 0;
    uint32_t* strip_bases = ctx->host_tile_bases + nviews + 1;
    for (uint32_t k = 0; k < nviews; ++k) {
        const athx_view& v = views[k];
        if (athx_status st = check_view(ctx, v, k); st != ATHX_OK) {
            return st;
        }
        if (!targets[k].colour) {
            return fail(ctx, ATHX_INVALID_ARGUMENT, "targets[" + std::to_string(k) + "].colour is null");
        }
        DView& d = ctx->host_views[k];
        std::memset(&d, 0, sizeof d);
        d.width = v.width;
        d.height = v.height;
        d.tiles_x = div_up(v.width, s->tile);
        d.tiles_y = div_up(v.height, s->tile);
        d.focal_x = v.focal_x;
        d.focal_y = v.focal_y;
        d.centre_x = v.centre_x;
        d.centre_y = v.centre_y;
        d.near_z = v.near_z;
        d.far_z = v.far_z;
        d.orthographic = v.orthographic ? 1u : 0u;
        {
            const float W = (float)v.width, H = (float)v.height;
            d.lim[0] = (-0.15f * W - v.centre_x) / v.focal_x;
            d.lim[1] = (1.15f * W - v.centre_x) / v.focal_x;
            d.lim[2] = (-0.15f * H - v.centre_y) / v.focal_y;
            d.lim[3] = (1.15f * H - v.centre_y) / v.focal_y;
        }
        d.tile_base = (uint32_t)total_tiles;
        d.strips_x = div_up(d.tiles_x, kStripTiles);
        d.strip_base = (uint32_t)total_strips;
        strip_bases[k] = d.strip_base;
        total_strips += (uint64_t)d.strips_x * d.tiles_y;
        std::memcpy(d.m, v.object_to_view, sizeof d.m);
        std::memcpy(d.eye, v.eye, sizeof d.eye);
        std::memcpy(d.bg, v.background, sizeof d.bg);
        d.colour = targets[k].colour;
        d.depth = targets[k].depth;
        d.final_t = targets[k].final_t;
        d.last = targets[k].last;
        d.dl_dcolour = dl ? dl[k] : nullptr;
        d.crypto = targets[k].crypto;
        d.trace = targets[k].trace;
        d.under_colour = targets[k].under_colour;
        d.under_depth = targets[k].under_depth;
        d.under_crypto = targets[k].under_crypto;
        // The walk's variants are the frame's: every view has each or none.
        const auto on = [&](uint32_t j) {
            return std::array<bool, 3>{targets[j].crypto != nullptr, targets[j].trace != nullptr,
                                       targets[j].under_colour != nullptr};
        };
        if (on(k) != on(0)) {
            return fail(ctx, ATHX_INVALID_ARGUMENT,
                        "targets[" + std::to_string(k) + "]: crypto, trace and the under layer are each the "
                        "frame's -- given for every view or for none");
        }
        if ((targets[k].under_colour != nullptr) != (targets[k].under_depth != nullptr)) {
            return fail(ctx, ATHX_INVALID_ARGUMENT, "the under layer needs both its colour and its depth");
        }
        ctx->host_tile_bases[k] = (uint32_t)total_tiles;
        total_tiles += (uint64_t)d.tiles_x * d.tiles_y;
        if (total_tiles >= (uint64_t{1} << 32) / 2) {
            return fail(ctx, ATHX_UNSUPPORTED, "too many tiles in one frame");
        }
    }
    ctx->host_tile_bases[nviews] = (uint32_t)total_tiles;
    strip_bases[nviews] = (uint32_t)total_strips;
    if (strips_out) {
        *strips_out = total_strips;
    }
    // Each instance's copy of the views, its own transform in place of the
    // view's (the last first, so the first block is copied before it is
    // changed).
    for (uint32_t j = ninst; inst && j-- > 0;) {
        for (uint32_t k = 0; k < nviews; ++k) {
            DView& d = ctx->host_views[(size_t)j * nviews + k];
            if (j != 0) {
                d = ctx->host_views[k];
            }
            if (inst[j].object_to_view) {
                std::memcpy(d.m, inst[j].object_to_view + (size_t)k * 12, sizeof d.m);
            }
            if (inst[j].eye) {
                std::memcpy(d.eye, inst[j].eye + (size_t)k * 3, sizeof d.eye);
            }
        }
    }
    return ATHX_OK;
}

// Uploads the staged views (nviews of them an instance) and tile bases.
athx_status upload_views(athx_ctx* ctx, uint32_t nviews, hipStream_t s, uint32_t ninst = 1) {
    const size_t staged = (size_t)nviews * ninst;
    ATHX_HIP(ctx, reserve(ctx, ctx->views, sizeof(DView) * staged), "views");
    ATHX_HIP(ctx, reserve(ctx, ctx->tile_bases, sizeof(uint32_t) * 2 * (nviews + 1)), "tile bases");
    ATHX_HIP(ctx, hipMemcpyAsync(ctx->views.ptr, ctx->host_views, sizeof(DView) * staged, hipMemcpyHostToDevice, s),
             "upload views");
    ATHX_HIP(ctx, hipMemcpyAsync(ctx->tile_bases.ptr, ctx->host_tile_bases, sizeof(uint32_t) * 2 * (nviews + 1),
                                 hipMemcpyHostToDevice, s),
             "upload tile bases");
    ATHX_HIP(ctx, hipEventRecord(ctx->upload_done, s), "record upload");
    ctx->upload_pending = true;
    return ATHX_OK;
}

double elapsed_ms(hipEvent_t a, hipEvent_t b) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) {
        return 0.0;
    }
    return ms;
}

// Sort metadata words for n pairs: hist (passes * 256), look-back
// (passes * partitions * 256), counters (passes), the histogram blocks'
// partial counts (passes * 256 * blocks), back to back.
size_t sort_meta_words(uint32_t n, uint32_t passes) {
    const size_t digit_cells = (size_t)256 * sort_partitions(n);
    return (size_t)passes * 256 + (size_t)passes * sort_partitions(n) * 256 + passes +
           (size_t)passes * 256 * sort_hist_blocks(n) + 2 * digit_cells + scan_status_words((uint32_t)digit_cells) +
           4;
}

SortScratch sort_scratch(const DeviceBuffer& meta, void* alt_keys, uint32_t* alt_vals, uint32_t n,
                         uint32_t passes) {
    SortScratch sc{};
    sc.keys_alt = alt_keys;
    sc.values_alt = alt_vals;
    sc.hist = static_cast<uint32_t*>(meta.ptr);
    sc.lookback = sc.hist + (size_t)passes * 256;
    sc.counters = sc.lookback + (size_t)passes * sort_partitions(n) * 256;
    sc.partials = sc.counters + passes;
    const size_t digit_cells = (size_t)256 * sort_partitions(n);
    sc.counts = sc.partials + (size_t)passes * 256 * sort_hist_blocks(n);
    sc.offsets = sc.counts + digit_cells;
    // The scan's status words are 64-bit: an even word offset from the (256-byte
    // aligned) buffer start.
    const size_t status_at = (size_t)((sc.offsets + digit_cells) - sc.hist);
    sc.scan_status = sc.hist + ((status_at + 1) & ~size_t{1});
    // The scan's total, 8-byte aligned.
    uint32_t* after = sc.scan_status + scan_status_words((uint32_t)digit_cells);
    sc.scan_total = reinterpret_cast<uint64_t*>(reinterpret_cast<uintptr_t>(after + 1) & ~uintptr_t{7});
    return sc;
}

// The sizes a frame runs at.
struct FrameShape {
    uint32_t nviews = 0;
    uint32_t instances = 1;    // clouds projected into each view
    uint32_t count = 0;        // Gaussians a view: every instance's, back to back
    uint32_t slots = 0;        // views * count
    uint32_t tiles = 0;        // over all views
    uint32_t strips = 0;       // over all views
    uint32_t key_bits = 0;     // 24 depth bits + the tile bits (the reference's pair keys)
    uint32_t strip_bits = 0;   // bits of a strip number
    uint32_t strip_passes = 0;
    uint32_t entry_cap = 0;    // strip entries the binning holds
    uint32_t seg_cap = 0;      // segments: entry_cap / kStripSeg + a partial one a strip
    // A frame of several views whose ranks are grouped by view (the view
    // pass), so that each view's strip entries are one run: its strip sort
    // orders each run by the view's own strip number (launch_segment_sort),
    // view_strip_bits bits; the sort's counts are sized for entry_pad more
    // entries (a partial partition a view).
    bool segmented = false;
    uint32_t view_strip_bits = 0;
    uint32_t entry_pad = 0;
    uint32_t capacity = 0;     // pairs the pair buffer holds
    std::vector<uint32_t> bases;  // each instance's first Gaussian in a view's slots, then count
    bool matte = false, traced = false, under = false;  // the walk's variants (athx_targets)
    bool variants() const { return matte || traced || under; }
};

bool masks_in_keys(const FrameShape& fs) { return fs.strip_bits <= 16; }
// The segmented strip sort's partitions: the entries' and a partial one a view.
uint32_t segment_parts_cap(const FrameShape& fs) { return div_up(fs.entry_cap, kSegmentPart) + fs.nviews; }
// The depth sort's, over the views' visible slots.
uint32_t depth_parts_cap(const FrameShape& fs) { return div_up(fs.slots, kSegmentPart) + fs.nviews; }
// Several views: the depth sort orders each view's visible slots apart
// (launch_segment_sort over the views' runs), so the ranks come out grouped
// by view, each view's in depth order.
bool grouped_by_view(const FrameShape& fs) { return fs.nviews >= 2 && fs.nviews <= 256 && fs.count >= 2; }

void set_entry_cap(FrameShape& fs, uint32_t entries) {
    fs.entry_cap = entries;
    fs.entry_pad = fs.segmented ? fs.nviews * kSegmentPart : 0u;
    fs.seg_cap = div_up(entries, kStripSeg) + fs.strips;
}

// Validates the call and stages the views (pinned), filling the shape.
athx_status begin_frame(athx_ctx* ctx, const athx_view* views, uint32_t nviews, const FrameInstance* inst,
                        uint32_t ninst, const athx_settings* settings, const athx_targets* targets,
                        const float* const* dl, FrameShape& fs) {
    if (!views || nviews == 0 || !targets) {
        return fail(ctx, ATHX_INVALID_ARGUMENT, "no views or targets");
    }
    if (!inst || ninst == 0) {
        return fail(ctx, ATHX_INVALID_ARGUMENT, "no instances");
    }
    if (athx_status st = check_settings(ctx, settings); st != ATHX_OK) {
        return st;
    }
    uint64_t count64 = 0;
    for (uint32_t j = 0; j < ninst; ++j) {
        if (athx_status st = check_gaussians(ctx, inst[j].g); st != ATHX_OK) {
            if (ninst > 1) {
                ctx->error = "instance " + std::to_string(j) + ": " + ctx->error;
            }
            return st;
        }
        count64 += inst[j].g->count;
    }
    const uint64_t slots64 = count64 * nviews;
    if (slots64 >= (uint64_t{1} << 32)) {
        // TileRasterizer.cpp:323-325: slots are 32-bit.
        return fail(ctx, ATHX_UNSUPPORTED, "more than 2^32 (view, Gaussian) slots in one frame");
    }
    ATHX_HIP(ctx, hipSetDevice(ctx->device), "set device");
    uint64_t tiles64 = 0, strips64 = 0;
    if (athx_status st = stage_views(ctx, views, nviews, settings, targets, dl, tiles64, &strips64, inst, ninst);
        st != ATHX_OK) {
        return st;
    }
    fs.nviews = nviews;
    fs.instances = ninst;
    fs.count = (uint32_t)count64;
    fs.matte = targets[0].crypto != nullptr;
    fs.traced = targets[0].trace != nullptr;
    fs.under = targets[0].under_colour != nullptr;
    fs.bases.assign(1, 0u);
    for (uint32_t j = 0; j < ninst; ++j) {
        fs.bases.push_back(fs.bases.back() + inst[j].g->count);
    }
    fs.slots = (uint32_t)slots64;
    fs.tiles = (uint32_t)tiles64;
    fs.strips = (uint32_t)strips64;
    fs.key_bits = kDepthBits + bits_for(fs.tiles);
    fs.strip_bits = bits_for(fs.strips);
    fs.strip_passes = div_up(fs.strip_bits, 8);
    uint32_t view_strips = 1;
    const uint32_t* strip_bases = ctx->host_tile_bases + nviews + 1;
    for (uint32_t k = 0; k < nviews; ++k) {
        view_strips = std::max(view_strips, strip_bases[k + 1] - strip_bases[k]);
    }
    fs.view_strip_bits = bits_for(view_strips);
    fs.segmented = grouped_by_view(fs) && masks_in_keys(fs) && nviews <= kMaxSegmentViews;
    return ATHX_OK;
}

athx_status reserve_front(athx_ctx* ctx, const FrameShape& fs) {
    const size_t slots = fs.slots;
    ATHX_HIP(ctx, reserve(ctx, ctx->records, slots * kRecordFloats * sizeof(float) + 16), "records");
    ATHX_HIP(ctx, reserve(ctx, ctx->touched, slots * 4 + 16), "touched");
    ATHX_HIP(ctx, reserve(ctx, ctx->depth_keys, slots * 4 + 16), "depth keys");
    ATHX_HIP(ctx, reserve(ctx, ctx->offsets, (slots / 64 + 1) * 4 + 16), "entry offsets");
    ATHX_HIP(ctx, reserve(ctx, ctx->scan_status, scan_status_words(std::max(fs.slots, fs.strips)) * 4),
             "scan status");
    ATHX_HIP(ctx, reserve(ctx, ctx->totals, 8 * kTotalWords), "totals");
    ATHX_HIP(ctx, reserve(ctx, ctx->ranges, (size_t)fs.tiles * 8 + 16), "ranges");
    ATHX_HIP(ctx, reserve(ctx, ctx->counter_words, kCounterWords * 4), "counters");
    ATHX_HIP(ctx, reserve(ctx, ctx->depth_keys_sorted, slots * 4 + 16), "depth sort keys");
    ATHX_HIP(ctx, reserve(ctx, ctx->depth_slots, slots * 4 + 16), "depth sort slots");
    ATHX_HIP(ctx, reserve(ctx, ctx->depth_alt_keys, slots * 4 + 16), "depth sort keys");
    ATHX_HIP(ctx, reserve(ctx, ctx->depth_alt_slots, slots * 4 + 16), "depth sort slots");
    const bool grouped = grouped_by_view(fs);
    ATHX_HIP(ctx, reserve(ctx, ctx->depth_sort_meta,
                          sort_meta_words(fs.slots + (grouped ? fs.nviews * kSegmentPart : 0u), kDepthPasses) * 4 + 16),
             "depth sort metadata");
    if (grouped) {
        ATHX_HIP(ctx, reserve(ctx, ctx->depth_segments, ((size_t)depth_parts_cap(fs) + 5u * fs.nviews + 8) * 4),
                 "depth segments");
    }
    ATHX_HIP(ctx, reserve(ctx, ctx->strip_ranges, (size_t)fs.strips * 8 + 16), "strip ranges");
    ATHX_HIP(ctx, reserve(ctx, ctx->segbase, (size_t)fs.strips * 4 + 16), "strip segments");
    ATHX_HIP(ctx, reserve(ctx, ctx->rank_fp, slots * 32 + 32), "ranked footprints");
    if (fs.matte) {
        ATHX_HIP(ctx, reserve(ctx, ctx->slot_ids, slots * 4 + 16), "Cryptomatte ids");
    }
    return ATHX_OK;
}

// The binning's buffers for fs.entry_cap entries.
athx_status reserve_bins(athx_ctx* ctx, const FrameShape& fs) {
    const size_t cap = fs.entry_cap;
    ATHX_HIP(ctx, reserve(ctx, ctx->strip_keys, cap * 4 + 16), "strip keys");
    ATHX_HIP(ctx, reserve(ctx, ctx->strip_slots, cap * 4 + 16), "strip slots");
    ATHX_HIP(ctx, reserve(ctx, ctx->strip_alt_keys, cap * 4 + 16), "strip sort keys");
    ATHX_HIP(ctx, reserve(ctx, ctx->strip_alt_slots, cap * 4 + 16), "strip sort slots");
    const uint32_t passes = fs.segmented ? div_up(fs.view_strip_bits, 8) : fs.strip_passes;
    ATHX_HIP(ctx, reserve(ctx, ctx->strip_sort_meta, sort_meta_words(fs.entry_cap + fs.entry_pad, passes) * 4 + 16),
             "strip sort metadata");
    if (fs.segmented) {
        ATHX_HIP(ctx, reserve(ctx, ctx->segments, ((size_t)segment_parts_cap(fs) + 4u * fs.nviews + 4) * 4),
                 "strip segments");
    }
    ATHX_HIP(ctx, reserve(ctx, ctx->masks, cap * 2 + 16), "entry masks");
    ATHX_HIP(ctx, reserve(ctx, ctx->seg_map, (size_t)fs.seg_cap * 32 + 32), "segment descriptors");
    const size_t table = (size_t)kStripTiles * fs.seg_cap;
    // The table's scan status follows the table.
    ATHX_HIP(ctx, reserve(ctx, ctx->seg_table, table * 4 + scan_status_words((uint32_t)table) * 4 + 64),
             "segment table");
    return ATHX_OK;
}

athx_status reserve_pairs(athx_ctx* ctx, const FrameShape& fs) {
    ATHX_HIP(ctx, reserve(ctx, ctx->pair_vals, (size_t)fs.capacity * 4 + 16), "pair values");
    return ATHX_OK;
}

// The buffers of the exact counts (copy_frame, the counters).
athx_status reserve_exact(athx_ctx* ctx, uint32_t slots) {
    ATHX_HIP(ctx, reserve(ctx, ctx->exact, (size_t)slots * 4 + 16), "exact counts");
    ATHX_HIP(ctx, reserve(ctx, ctx->exact_offsets, (size_t)slots * 4 + 16), "exact offsets");
    ATHX_HIP(ctx, reserve(ctx, ctx->exact_status, scan_status_words(slots) * 4 + 16), "exact scan status");
    ATHX_HIP(ctx, reserve(ctx, ctx->fp_rects, (size_t)slots * 8 + 16), "footprint rects");
    ATHX_HIP(ctx, reserve(ctx, ctx->big_list, (size_t)slots * 4 + 16), "big footprints");
    ATHX_HIP(ctx, reserve(ctx, ctx->fp_masks, (size_t)slots * 8 + 16), "footprint masks");
    ATHX_HIP(ctx, reserve(ctx, ctx->fp_big_base, (size_t)slots * 4 + 16), "big footprint bases");
    ATHX_HIP(ctx, reserve(ctx, ctx->fp_big_masks, big_mask_words(slots) * 8 + 16), "big footprint masks");
    return ATHX_OK;
}

Footprints footprints_of(athx_ctx* ctx, uint32_t slots) {
    Footprints fp{};
    fp.rects = as<uint2>(ctx->fp_rects);
    fp.masks = as<uint64_t>(ctx->fp_masks);
    fp.big_list = as<uint32_t>(ctx->big_list);
    fp.big_mask_base = as<uint32_t>(ctx->fp_big_base);
    fp.big_masks = as<uint64_t>(ctx->fp_big_masks);
    fp.big_mask_cap = big_mask_words(slots);
    fp.big_total = as<uint64_t>(ctx->totals) + kTotalBig;
    return fp;
}

// Enqueues the exact tile counts into ctx->exact (reserved).
athx_status enqueue_exact(athx_ctx* ctx, uint32_t n, uint32_t nviews, uint32_t tile, hipStream_t s) {
    ATHX_HIP(ctx, launch_exact_counts(as<DView>(ctx->views), nviews, n, tile, as<float>(ctx->records),
                                      as<uint32_t>(ctx->exact), as<uint32_t>(ctx->depth_keys),
                                      footprints_of(ctx, n * nviews), as<uint32_t>(ctx->exact_status), s),
             "exact counts");
    return ATHX_OK;
}

const uint32_t* sorted_slots_of(const athx_ctx* ctx, bool in_alt) {
    return static_cast<const uint32_t*>(in_alt ? ctx->depth_alt_slots.ptr : ctx->depth_slots.ptr);
}

// A segmented frame whose views sort their strips in one pass keeps only the
// masks of the sorted keys (the sort writes them), its strip ranges read off
// the sort's counts.
bool sorts_masks(const FrameShape& fs) { return fs.segmented && masks_in_keys(fs) && fs.view_strip_bits <= 8; }

// Frames of at most 65 536 strips decide their entries' tiles in the
// emission and carry the masks in the keys' upper half.
StripBins bins_of(athx_ctx* ctx, const FrameShape& fs, uint32_t n, uint32_t tile, bool strips_in_alt) {
    StripBins b{};
    b.views = as<DView>(ctx->views);
    b.strip_bases = as<uint32_t>(ctx->tile_bases) + fs.nviews + 1;
    b.nviews = fs.nviews;
    b.n = n;
    b.tile = tile;
    b.strips = fs.strips;
    b.records = as<float>(ctx->records);
    b.strip_slots = as<uint32_t>(strips_in_alt ? ctx->strip_alt_slots : ctx->strip_slots);
    b.fp = as<uint4>(ctx->rank_fp);
    b.masks_sorted = sorts_masks(fs) ? 1u : 0u;
    b.strip_keys = masks_in_keys(fs) && !b.masks_sorted
                       ? as<uint32_t>(strips_in_alt ? ctx->strip_alt_keys : ctx->strip_keys)
                       : nullptr;
    b.strip_ranges = as<uint32_t>(ctx->strip_ranges);
    b.segbase = as<uint32_t>(ctx->segbase);
    b.seg_desc = as<uint4>(ctx->seg_map);
    b.seg_total = as<uint64_t>(ctx->totals) + kTotalSegs;
    b.seg_cap = fs.seg_cap;
    b.masks = as<uint16_t>(ctx->masks);
    b.table = as<uint32_t>(ctx->seg_table);
    b.pair_total = as<uint64_t>(ctx->totals) + kTotalPairs;
    b.pair_cap = fs.capacity;
    b.pair_slots = as<uint32_t>(ctx->pair_vals);
    return b;
}

// A grouped frame's depth segments in ctx->depth_segments: part_seg (a word
// a partition), then seg_start, seg_len, seg_part (a word a view), each
// view's first compacted position (a word a view and the total), and the
// partitions in use.
SegmentArgs depth_segments_of(athx_ctx* ctx, const FrameShape& fs) {
    uint32_t* w = as<uint32_t>(ctx->depth_segments);
    SegmentArgs a{};
    a.part_seg = w;
    a.seg_start = w + depth_parts_cap(fs);
    a.seg_len = a.seg_start + fs.nviews;
    a.seg_part = a.seg_len + fs.nviews;
    a.parts = a.seg_part + fs.nviews + (fs.nviews + 1);  // after the view starts
    a.seg_base = nullptr;
    return a;
}

uint32_t* view_start_of(athx_ctx* ctx, const FrameShape& fs) {
    return const_cast<uint32_t*>(depth_segments_of(ctx, fs).seg_part) + fs.nviews;
}

// Projection, then the reference's depth order (TileRasterizer.cpp:1111-1137):
// the visible slots compacted in slot order, stably sorted by their 24-bit
// depth keys, and their strip entries scanned in that order. The visible
// total lands in totals[kTotalVisible], the entry total in
// totals[kTotalEntries], both on the device.
athx_status enqueue_front(athx_ctx* ctx, const std::vector<ProjectArgs>& pas, const FrameShape& fs, hipStream_t s,
                          bool timed) {
    uint64_t* totals = as<uint64_t>(ctx->totals);
    if (timed) {
        ATHX_HIP(ctx, hipEventRecord(ctx->ev[0], s), "event");
    }
    // A projection launch an instance (TileRasterizer.cpp:896-898), each with
    // its own copy of the views.
    for (size_t j = 0; j < pas.size(); ++j) {
        ProjectArgs pa = pas[j];
        pa.slot_ids = fs.matte ? as<uint32_t>(ctx->slot_ids) : nullptr;
        ATHX_HIP(ctx, launch_project(pa, as<DView>(ctx->views) + j * fs.nviews, fs.nviews, as<float>(ctx->records),
                                     as<uint32_t>(ctx->touched), as<uint32_t>(ctx->depth_keys), s),
                 "project");
    }
    if (timed) {
        ATHX_HIP(ctx, hipEventRecord(ctx->ev[1], s), "event");
    }
    const uint32_t tile = pas.empty() ? 16u : pas[0].tile;
    ATHX_HIP(ctx, launch_settle_count(as<DView>(ctx->views), fs.count, tile, fs.slots, as<float>(ctx->records),
                                      as<uint32_t>(ctx->touched), as<uint32_t>(ctx->depth_keys),
                                      reinterpret_cast<uint64_t*>(ctx->scan_status.ptr), s),
             "settle");
    // Several views: each view's visible slots are sorted apart, so the
    // ranks come out grouped by view, each view's in depth order. A strip is
    // one view's, so its entries -- and the pairs -- keep the reference's
    // order, while the emission reads one view's records at a time (which the
    // caches hold, where the frame's whole depth order would read them
    // scattered over every view's) and a view's entries come out as one run.
    const bool grouped = grouped_by_view(fs);
    const SegmentArgs dseg = grouped ? depth_segments_of(ctx, fs) : SegmentArgs{};
    ATHX_HIP(ctx, launch_compact(as<uint32_t>(ctx->touched), as<uint32_t>(ctx->depth_keys), fs.slots,
                                 as<uint32_t>(ctx->depth_keys_sorted), as<uint32_t>(ctx->depth_slots),
                                 totals + kTotalVisible, as<uint32_t>(ctx->scan_status), s, true,
                                 grouped ? fs.count : 0u, grouped ? view_start_of(ctx, fs) : nullptr),
             "compact");
    const SortScratch dsc =
        sort_scratch(ctx->depth_sort_meta, ctx->depth_alt_keys.ptr, as<uint32_t>(ctx->depth_alt_slots),
                     fs.slots + (grouped ? fs.nviews * kSegmentPart : 0u), kDepthPasses);
    bool in_alt = false;  // three passes end in the alternate buffers: read them there
    if (grouped) {
        SegmentBuild g{};
        g.nviews = fs.nviews;
        g.view_entry = view_start_of(ctx, fs);  // every view's start, from the compaction
        g.visible_dev = totals + kTotalVisible;
        g.max_visible = fs.slots;
        g.entry_total_dev = totals + kTotalVisible;
        g.cap = fs.slots;
        g.parts_cap = depth_parts_cap(fs);
        g.part_seg = const_cast<uint32_t*>(dseg.part_seg);
        g.seg_start = const_cast<uint32_t*>(dseg.seg_start);
        g.seg_len = const_cast<uint32_t*>(dseg.seg_len);
        g.seg_part = const_cast<uint32_t*>(dseg.seg_part);
        g.parts = const_cast<uint32_t*>(dseg.parts);
        g.binned_total = totals + kTotalDepthBinned;
        ATHX_HIP(ctx, launch_strip_segments(g, s), "depth segments");
        ATHX_HIP(ctx, launch_segment_sort(as<uint32_t>(ctx->depth_keys_sorted), as<uint32_t>(ctx->depth_slots), dseg,
                                          depth_parts_cap(fs), kDepthBits, dsc, s, in_alt, false),
                 "depth sort");
    } else {
        ATHX_HIP(ctx, launch_sort(as<uint32_t>(ctx->depth_keys_sorted), as<uint32_t>(ctx->depth_slots),
                                  totals + kTotalVisible, fs.slots, 0, kDepthBits, dsc, s, &in_alt),
                 "depth sort");
    }
    ctx->last_sorted_in_alt = in_alt;
    ctx->last_grouped = grouped;
    if (timed) {
        ATHX_HIP(ctx, hipEventRecord(ctx->ev[2], s), "event");
    }
    ATHX_HIP(ctx, launch_entry_scan(as<uint32_t>(ctx->touched), sorted_slots_of(ctx, in_alt),
                                    as<uint32_t>(in_alt ? ctx->depth_alt_keys : ctx->depth_keys_sorted),
                                    totals + kTotalVisible,
                                    fs.slots, as<uint32_t>(ctx->offsets), totals + kTotalEntries,
                                    as<uint32_t>(ctx->scan_status), s),
             "entries");
    if (timed) {
        ATHX_HIP(ctx, hipEventRecord(ctx->ev[3], s), "event");
    }
    return ATHX_OK;
}

// A segmented frame's segment arrays in ctx->segments: part_seg (a word a
// partition), then seg_start, seg_len, seg_part and view_entry (a word a
// view), and the partitions in use.
SegmentArgs segments_of(athx_ctx* ctx, const FrameShape& fs) {
    uint32_t* w = as<uint32_t>(ctx->segments);
    SegmentArgs a{};
    a.part_seg = w;
    a.seg_start = w + segment_parts_cap(fs);
    a.seg_len = a.seg_start + fs.nviews;
    a.seg_part = a.seg_len + fs.nviews;
    a.parts = a.seg_part + 2 * fs.nviews;  // after view_entry
    a.seg_base = as<uint32_t>(ctx->tile_bases) + fs.nviews + 1;  // each view's first strip
    return a;
}

uint32_t* view_entry_of(athx_ctx* ctx, const FrameShape& fs) {
    return const_cast<uint32_t*>(segments_of(ctx, fs).seg_part) + fs.nviews;
}

// Each view's first rank: its first compacted position (the depth sort keeps
// the views' runs where they were).
const uint32_t* view_first_of(athx_ctx* ctx, const FrameShape& fs) { return view_start_of(ctx, fs); }

// The views' runs of strip entries (launch_strip_segments), from the views'
// first ranks and the first entries the emission recorded.
athx_status enqueue_segments(athx_ctx* ctx, const FrameShape& fs, hipStream_t s) {
    uint64_t* totals = as<uint64_t>(ctx->totals);
    const SegmentArgs a = segments_of(ctx, fs);
    SegmentBuild g{};
    g.nviews = fs.nviews;
    g.view_first = view_first_of(ctx, fs);
    g.view_stride = 1;
    g.view_entry = view_entry_of(ctx, fs);
    g.visible_dev = totals + kTotalVisible;
    g.max_visible = fs.slots;
    g.entry_total_dev = totals + kTotalEntries;
    g.cap = fs.entry_cap;
    g.parts_cap = segment_parts_cap(fs);
    g.part_seg = const_cast<uint32_t*>(a.part_seg);
    g.seg_start = const_cast<uint32_t*>(a.seg_start);
    g.seg_len = const_cast<uint32_t*>(a.seg_len);
    g.seg_part = const_cast<uint32_t*>(a.seg_part);
    g.parts = const_cast<uint32_t*>(a.parts);
    g.binned_total = totals + kTotalBinned;
    ATHX_HIP(ctx, launch_strip_segments(g, s), "strip segments");
    return ATHX_OK;
}

// The strip entries, emitted in depth order, sorted on the strip, their
// segments scanned, tested and counted, and the counts scanned: the pair
// total lands in totals[kTotalPairs]. Every count is read on the device
// and bounded by fs's capacities, so the launches can be captured.
athx_status enqueue_bins(athx_ctx* ctx, const FrameShape& fs, uint32_t n, uint32_t tile, hipStream_t s, bool timed,
                         bool& strips_in_alt, bool count = false) {
    uint64_t* totals = as<uint64_t>(ctx->totals);
    const bool masked = masks_in_keys(fs);
    if (masked) {
        // Counting frames keep each rank's tiles (its entries' mask bits) for the counters.
        uint32_t* rank_tiles = nullptr;
        if (count) {
            ATHX_HIP(ctx, reserve(ctx, ctx->rank_tiles, (size_t)fs.slots * 4 + 16), "rank tiles");
            rank_tiles = as<uint32_t>(ctx->rank_tiles);
        }
        // A segmented frame's emission records each view's first entry.
        ATHX_HIP(ctx, launch_strip_mask_emit(as<DView>(ctx->views), n, tile,
                                             sorted_slots_of(ctx, ctx->last_sorted_in_alt), as<uint32_t>(ctx->offsets),
                                             totals + kTotalVisible, fs.slots, fs.entry_cap, as<float>(ctx->records),
                                             as<uint32_t>(ctx->strip_keys), as<uint32_t>(ctx->strip_slots), rank_tiles,
                                             s, fs.segmented ? view_entry_of(ctx, fs) : nullptr),
                 "strip entries");
        if (fs.segmented) {
            if (athx_status st = enqueue_segments(ctx, fs, s); st != ATHX_OK) {
                return st;
            }
        }
    } else {
        ATHX_HIP(ctx, launch_strip_emit(as<DView>(ctx->views), n, tile,
                                        sorted_slots_of(ctx, ctx->last_sorted_in_alt), as<uint32_t>(ctx->offsets),
                                        totals + kTotalVisible, fs.slots, fs.entry_cap, as<float>(ctx->records),
                                        as<uint4>(ctx->rank_fp), as<uint32_t>(ctx->strip_keys),
                                        as<uint32_t>(ctx->strip_slots), s),
                 "strip entries");
    }
    const bool segmented = masked && fs.segmented;
    // The segmented sort's counts are sized for a partial partition a view
    // more (entry_pad entries).
    const SortScratch sc = sort_scratch(ctx->strip_sort_meta, ctx->strip_alt_keys.ptr,
                                        as<uint32_t>(ctx->strip_alt_slots), fs.entry_cap + fs.entry_pad,
                                        segmented ? div_up(fs.view_strip_bits, 8) : fs.strip_passes);
    strips_in_alt = false;
    const bool mask_sort = sorts_masks(fs);
    if (segmented) {
        ATHX_HIP(ctx, launch_segment_sort(as<uint32_t>(ctx->strip_keys), as<uint32_t>(ctx->strip_slots),
                                          segments_of(ctx, fs), segment_parts_cap(fs), fs.view_strip_bits, sc, s,
                                          strips_in_alt, true, mask_sort ? as<uint16_t>(ctx->masks) : nullptr),
                 "strip sort");
    } else {
        ATHX_HIP(ctx, launch_sort(as<uint32_t>(ctx->strip_keys), as<uint32_t>(ctx->strip_slots),
                                  totals + kTotalEntries, fs.entry_cap, 0, fs.strip_bits, sc, s, &strips_in_alt),
                 "strip sort");
    }
    // The sorted entries are dense from here on; a segmented frame's count is
    // what its segments held.
    const uint32_t* keys = as<uint32_t>(strips_in_alt ? ctx->strip_alt_keys : ctx->strip_keys);
    if (mask_sort) {
        ATHX_HIP(ctx, launch_strip_ranges_seg(sc.offsets, segments_of(ctx, fs), fs.nviews,
                                              as<uint32_t>(ctx->tile_bases) + fs.nviews + 1, fs.strips,
                                              1u << fs.view_strip_bits, totals + kTotalBinned,
                                              as<uint32_t>(ctx->strip_ranges), s),
                 "strip ranges");
    } else {
        ATHX_HIP(ctx, launch_ranges(keys, totals + (segmented ? kTotalBinned : kTotalEntries), fs.entry_cap,
                                    fs.strips, as<uint32_t>(ctx->strip_ranges), s, masked ? 0xFFFFu : ~0u),
                 "strip ranges");
    }
    ATHX_HIP(ctx, launch_seg_scan(as<uint32_t>(ctx->strip_ranges), fs.strips, as<uint32_t>(ctx->segbase),
                                  totals + kTotalSegs, as<uint32_t>(ctx->scan_status), s),
             "segments");
    const StripBins b = bins_of(ctx, fs, n, tile, strips_in_alt);
    ATHX_HIP(ctx, launch_seg_desc(b, as<uint4>(ctx->seg_map), s), "segment descriptors");
    if (timed) {
        ATHX_HIP(ctx, hipEventRecord(ctx->ev[4], s), "event");
    }
    ATHX_HIP(ctx, launch_strip_count(b, s), "strip counts");
    uint32_t* table = as<uint32_t>(ctx->seg_table);
    uint32_t* table_status = table + (size_t)kStripTiles * fs.seg_cap;
    table_status += (reinterpret_cast<uintptr_t>(table_status) & 7u) ? 1 : 0;
    ATHX_HIP(ctx, launch_table_scan(table, totals + kTotalSegs, fs.seg_cap, totals + kTotalPairs, table_status, s),
             "table scan");
    return ATHX_OK;
}

// Ranges, the pairs at their places, the blend and the counters.
athx_status enqueue_draw(athx_ctx* ctx, const FrameShape& fs, const athx_settings* settings, hipStream_t s,
                         bool timed, bool strips_in_alt, bool capturing = false) {
    const StripBins b = bins_of(ctx, fs, fs.count, settings->tile, strips_in_alt);
    ATHX_HIP(ctx, launch_strip_pairs(b, s), "pairs");
    if (timed) {
        ATHX_HIP(ctx, hipEventRecord(ctx->ev[5], s), "event");
    }
    ATHX_HIP(ctx, launch_tile_ranges(b, as<uint32_t>(ctx->tile_bases), fs.tiles, as<uint32_t>(ctx->ranges), s),
             "ranges");
    if (timed) {
        ATHX_HIP(ctx, hipEventRecord(ctx->ev[6], s), "event");
    }
    ATHX_HIP(ctx, launch_blend(as<DView>(ctx->views), as<uint32_t>(ctx->tile_bases), fs.nviews, fs.tiles,
                               settings->tile, settings->batch, as<uint32_t>(ctx->ranges), as<uint32_t>(ctx->pair_vals),
                               as<float>(ctx->records), settings->depth_mode, settings->depth_threshold, s,
                               fs.variants(), fs.matte ? as<uint32_t>(ctx->slot_ids) : nullptr),
             "blend");
    if (timed) {
        ATHX_HIP(ctx, hipEventRecord(ctx->ev[7], s), "event");
    }
    if (settings->count_splats) {
        // SplatCounters from what the frame already holds: each slot's cull
        // reason in its depth key, and each drawn footprint's tiles -- kept
        // by the emission a rank when it decided them, else counted exactly
        // a slot (frames past 65 536 strips).
        CounterArgs ca{};
        ca.depth_keys = as<uint32_t>(ctx->depth_keys);
        ca.sorted_slots = sorted_slots_of(ctx, ctx->last_sorted_in_alt);
        ca.visible_dev = as<uint64_t>(ctx->totals) + kTotalVisible;
        ca.slots = fs.slots;
        ca.count = fs.count;
        ca.clouds = std::min<uint32_t>(fs.instances, ATHX_MAX_COUNTED_CLOUDS);
        for (uint32_t j = 0; j <= ca.clouds; ++j) {
            ca.base[j] = fs.bases[j];
        }
        if (masks_in_keys(fs)) {
            ca.rank_tiles = as<uint32_t>(ctx->rank_tiles);
        } else {
            if (athx_status st = reserve_exact(ctx, fs.slots); st != ATHX_OK) {
                return st;
            }
            if (athx_status st = enqueue_exact(ctx, fs.count, fs.nviews, settings->tile, s); st != ATHX_OK) {
                return st;
            }
            ca.slot_tiles = as<uint32_t>(ctx->exact);
        }
        ATHX_HIP(ctx, launch_fill_u32(as<uint32_t>(ctx->counter_words), 0u, kCounterWords, s), "clear counters");
        ATHX_HIP(ctx, launch_frame_counters(ca, as<uint32_t>(ctx->counter_words), s), "counters");
        if (!capturing) {
            athx_ctx::CounterCopy& c = ctx->counter_ring[ctx->counter_seq % 3];
            ATHX_HIP(ctx, hipMemcpyAsync(c.host, ctx->counter_words.ptr, kCounterWords * 4, hipMemcpyDeviceToHost, s),
                     "read counters back");
            ATHX_HIP(ctx, hipEventRecord(c.done, s), "record counters");
            c.seq = ++ctx->counter_seq;
            c.tag = settings->counters_tag;
            c.slots = fs.slots;
            c.clouds = ca.clouds;
        }
    }
    return ATHX_OK;
}

// The adjoint: clear the per-slot 2D gradients, walk every tile back to
// front, then differentiate the projection. grad2d must be reserved.
athx_status enqueue_backward(athx_ctx* ctx, const ProjectArgs& pa, uint32_t nviews, uint32_t slots,
                             const athx_settings* settings, const athx_grads& grads, hipStream_t s) {
    ATHX_HIP(ctx, launch_fill_u32(as<uint32_t>(ctx->grad2d), 0u, (uint64_t)slots * kGrad2d, s), "clear gradients");
    ATHX_HIP(ctx, launch_blend_backward(as<DView>(ctx->views), as<uint32_t>(ctx->tile_bases), nviews, ctx->last_tiles,
                                        settings->tile, settings->batch, as<uint32_t>(ctx->ranges),
                                        as<uint32_t>(ctx->pair_vals), as<float>(ctx->records),
                                        as<float>(ctx->grad2d), s),
             "blend backward");
    ATHX_HIP(ctx, launch_project_backward(pa, as<DView>(ctx->views), nviews, as<uint32_t>(ctx->touched),
                                          as<float>(ctx->records), as<float>(ctx->grad2d), grads, s),
             "projection backward");
    return ATHX_OK;
}

uint64_t grown(uint64_t seen) { return std::max<uint64_t>(seen + seen / 4, 1024); }

// What the adjoint and the readers of the frame's buffers need to know.
void finish_frame(athx_ctx* ctx, const athx_settings* settings, const FrameShape& fs, uint64_t pairs_known) {
    ctx->last_n = fs.count;
    ctx->last_instances = fs.instances;
    ctx->last_under = fs.under;
    ctx->last_nviews = fs.nviews;
    ctx->last_tiles = fs.tiles;
    ctx->last_strips = fs.strips;
    ctx->last_tile = settings->tile;
    ctx->last_capacity = fs.capacity;
    ctx->last_entry_cap = fs.entry_cap;
    ctx->last_key_bits = fs.key_bits;
    ctx->last_true_pairs = pairs_known;
    ctx->have_frame = true;
    ctx->serial += 1;
}

// The frame a backward pass or a contribution walk names must be the one
// whose buffers the context still holds.
athx_status check_serial(athx_ctx* ctx, uint64_t frame_serial) {
    if (!ctx->have_frame) {
        return fail(ctx, ATHX_INVALID_ARGUMENT, "no frame has been rendered on this context");
    }
    if (frame_serial != ctx->serial) {
        return fail(ctx, ATHX_INVALID_ARGUMENT,
                    "frame " + std::to_string(frame_serial) + " was replaced by frame " + std::to_string(ctx->serial) +
                        ": its pairs and records are gone; render it again before walking it");
    }
    return ATHX_OK;
}



// Reads the frame's totals into the pinned words, waiting for them.
athx_status read_totals(athx_ctx* ctx, hipStream_t s) {
    ATHX_HIP(ctx, hipMemcpyAsync(ctx->host_totals, ctx->totals.ptr, 8 * kTotalWords, hipMemcpyDeviceToHost, s),
             "read totals");
    ATHX_HIP(ctx, hipStreamSynchronize(s), "wait for totals");
    return ATHX_OK;
}

}  // namespace

extern "C" {

ATHX_API const char* athx_version(void) { return "athx 0.2 (gfx950)"; }

ATHX_API int athx_error_code(athx_status status) {
    return status >= ATHX_UNKNOWN && status <= ATHX_INTERNAL_ERROR ? (int)status - 1 : -1;
}

ATHX_API athx_status athx_create(int device, athx_ctx** out) {
    if (!out) {
        return ATHX_INVALID_ARGUMENT;
    }
    *out = nullptr;
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0) {
        return ATHX_DEVICE_FAILURE;
    }
    if (device < 0 || device >= count) {
        return ATHX_INVALID_ARGUMENT;
    }
    athx_ctx* ctx = new (std::nothrow) athx_ctx();
    if (!ctx) {
        return ATHX_OUT_OF_MEMORY;
    }
    ctx->device = device;
    if (hipSetDevice(device) != hipSuccess || hipHostMalloc(reinterpret_cast<void**>(&ctx->host_totals), 8 * kTotalWords) != hipSuccess) {
        delete ctx;
        return ATHX_DEVICE_FAILURE;
    }
    for (hipEvent_t& ev : ctx->ev) {
        if (hipEventCreate(&ev) != hipSuccess) {
            ev = nullptr;
        }
    }
    if (hipEventCreateWithFlags(&ctx->upload_done, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->async_done, hipEventDisableTiming) != hipSuccess) {
        athx_destroy(ctx);
        return ATHX_DEVICE_FAILURE;
    }
    for (athx_ctx::CounterCopy& c : ctx->counter_ring) {
        if (hipHostMalloc(reinterpret_cast<void**>(&c.host), kCounterWords * 4) != hipSuccess ||
            hipEventCreateWithFlags(&c.done, hipEventDisableTiming) != hipSuccess) {
            athx_destroy(ctx);
            return ATHX_DEVICE_FAILURE;
        }
    }
    // This device's copy of the no-allocation scan's scratch, with the device
    // current; and room for error messages, so a failing call of the
    // no-allocation entry point does not grow the string.
    if (hipGetSymbolAddress(reinterpret_cast<void**>(&ctx->noalloc_scratch), HIP_SYMBOL(scratch)) != hipSuccess ||
        !ctx->noalloc_scratch) {
        athx_destroy(ctx);
        return ATHX_DEVICE_FAILURE;
    }
    try {
        ctx->error.reserve(512);
    } catch (const std::bad_alloc&) {
        athx_destroy(ctx);
        return ATHX_OUT_OF_MEMORY;
    }
    *out = ctx;
    return ATHX_OK;
}

ATHX_API void athx_destroy(athx_ctx* ctx) {
    if (!ctx) {
        return;
    }
    (void)hipSetDevice(ctx->device);
    (void)hipDeviceSynchronize();
    for (DeviceBuffer* b : all_buffers(ctx)) {
        release(*b);
    }
    if (ctx->host_totals) {
        (void)hipHostFree(ctx->host_totals);
    }
    for (athx_ctx::CounterCopy& c : ctx->counter_ring) {
        if (c.host) {
            (void)hipHostFree(c.host);
        }
        if (c.done) {
            (void)hipEventDestroy(c.done);
        }
    }
    if (ctx->host_views) {
        (void)hipHostFree(ctx->host_views);
    }
    if (ctx->host_tile_bases) {
        (void)hipHostFree(ctx->host_tile_bases);
    }
    for (hipEvent_t ev : ctx->ev) {
        if (ev) {
            (void)hipEventDestroy(ev);
        }
    }
    if (ctx->upload_done) {
        (void)hipEventDestroy(ctx->upload_done);
    }
    if (ctx->async_done) {
        (void)hipEventDestroy(ctx->async_done);
    }
    delete ctx;
}

ATHX_API const char* athx_last_error(const athx_ctx* ctx) { return ctx ? ctx->error.c_str() : "no context"; }

ATHX_API athx_status athx_release_scratch(athx_ctx* ctx)  try {
    if (!ctx) {
        return ATHX_INVALID_ARGUMENT;
    }
    (void)hipSetDevice(ctx->device);
    ATHX_HIP(ctx, hipDeviceSynchronize(), "synchronize before release");
    for (DeviceBuffer* b : all_buffers(ctx)) {
        if (b != &ctx->views && b != &ctx->tile_bases && b != &ctx->totals && b != &ctx->counter_words &&
            b != &ctx->api_count) {
            release(*b);
        }
    }
    ctx->exact_growth = true;
    ctx->pair_hint = 0;
    ctx->entry_hint = 0;
    ctx->pair_last = 0;
    ctx->entry_last = 0;
    ctx->have_frame = false;
    return ATHX_OK;
} catch (const std::bad_alloc&) {
    return fail(ctx, ATHX_OUT_OF_MEMORY, "host allocation failed");
} catch (const std::exception& e) {
    return fail(ctx, ATHX_INTERNAL_ERROR, e.what());
}

ATHX_API uint64_t athx_scratch_bytes(const athx_ctx* ctx) {
    if (!ctx) {
        return 0;
    }
    uint64_t sum = 0;
    for (const DeviceBuffer* b : all_buffers(const_cast<athx_ctx*>(ctx))) {
        sum += b->bytes;
    }
    return sum;
}

}  // extern "C"

namespace {

// athx_render and athx_render_instances: one frame of every instance in
// every view.
athx_status render_frame(athx_ctx* ctx, const athx_view* views, uint32_t nviews, const FrameInstance* inst,
                         uint32_t ninst, const athx_settings* settings, const athx_targets* targets, void* stream,
                         athx_frame_stats* stats, uint32_t* slot_bases) {
    FrameShape fs{};
    if (athx_status st = begin_frame(ctx, views, nviews, inst, ninst, settings, targets, nullptr, fs); st != ATHX_OK) {
        return st;
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    const bool timed = settings->time_stages != 0 && ctx->ev[7] != nullptr;
    const auto wall0 = std::chrono::steady_clock::now();
    if (athx_status st = upload_views(ctx, nviews, s, ninst); st != ATHX_OK) {
        return st;
    }
    if (athx_status st = reserve_front(ctx, fs); st != ATHX_OK) {
        return st;
    }
    const std::vector<ProjectArgs> pas = instance_args(inst, ninst, settings, fs.count, slot_bases);
    const uint64_t* host = ctx->host_totals;
    // With the sizes of earlier frames known, the whole frame is queued at
    // once and the one wait is at its end, where the totals say whether it
    // fitted; the first frame (and one that outgrows them) waits for its
    // entry and pair totals before sizing what holds them, as the reference
    // does for its pairs (TileRasterizer.cpp:1093-1102).
    const bool ahead = ctx->entry_last != 0 && ctx->pair_last != 0;
    if (ahead) {
        set_entry_cap(fs, (uint32_t)ctx->entry_last);
        fs.capacity = (uint32_t)ctx->pair_last;
        if (athx_status st = reserve_bins(ctx, fs); st != ATHX_OK) {
            return st;
        }
        if (athx_status st = reserve_pairs(ctx, fs); st != ATHX_OK) {
            return st;
        }
    }
    if (athx_status st = enqueue_front(ctx, pas, fs, s, timed); st != ATHX_OK) {
        return st;
    }
    bool strips_in_alt = false;
    bool fitted = ahead;
    if (ahead) {
        if (athx_status st = enqueue_bins(ctx, fs, fs.count, settings->tile, s, timed, strips_in_alt, settings->count_splats != 0); st != ATHX_OK) {
            return st;
        }
        if (athx_status st = enqueue_draw(ctx, fs, settings, s, timed, strips_in_alt); st != ATHX_OK) {
            return st;
        }
        if (athx_status st = read_totals(ctx, s); st != ATHX_OK) {
            return st;
        }
        fitted = host[kTotalEntries] <= fs.entry_cap && host[kTotalPairs] <= fs.capacity;
    } else if (athx_status st = read_totals(ctx, s); st != ATHX_OK) {
        return st;
    }
    if (!fitted) {
        // The front half's outputs stand; bin and draw again at the exact sizes.
        const uint64_t entries = host[kTotalEntries];
        if (entries >= kMaxPairs) {
            return fail(ctx, ATHX_UNSUPPORTED, "more than 2^30 strip entries in one frame");
        }
        if (!ahead || entries > fs.entry_cap) {
            set_entry_cap(fs, std::max<uint32_t>((uint32_t)entries, 1u));
            if (athx_status st = reserve_bins(ctx, fs); st != ATHX_OK) {
                return st;
            }
            if (athx_status st = enqueue_bins(ctx, fs, fs.count, settings->tile, s, timed, strips_in_alt, settings->count_splats != 0); st != ATHX_OK) {
                return st;
            }
            if (athx_status st = read_totals(ctx, s); st != ATHX_OK) {
                return st;
            }
        }
        const uint64_t pairs = host[kTotalPairs];
        if (pairs >= kMaxPairs) {
            return fail(ctx, ATHX_UNSUPPORTED, "more than 2^30 (tile, Gaussian) pairs in one frame");
        }
        fs.capacity = std::max<uint32_t>(std::max<uint32_t>((uint32_t)pairs, fs.capacity), 1u);
        if (athx_status st = reserve_pairs(ctx, fs); st != ATHX_OK) {
            return st;
        }
        if (athx_status st = enqueue_draw(ctx, fs, settings, s, timed, strips_in_alt); st != ATHX_OK) {
            return st;
        }
        if (timed) {
            ATHX_HIP(ctx, hipStreamSynchronize(s), "wait for frame");
        }
    }
    if (settings->count_splats) {
        // A counting frame waits for itself (TileRasterizer.cpp:1269-1282), so
        // its counters are the newest finished when render returns.
        ATHX_HIP(ctx, hipStreamSynchronize(s), "wait for the counted frame");
    }
    const uint64_t pairs64 = host[kTotalPairs];
    ctx->entry_last = std::min<uint64_t>(grown(host[kTotalEntries]), kMaxPairs - 1);
    ctx->pair_last = std::min<uint64_t>(grown(pairs64), kMaxPairs - 1);
    ctx->entry_hint = std::max(ctx->entry_hint, ctx->entry_last);
    ctx->pair_hint = std::max(ctx->pair_hint, ctx->pair_last);
    finish_frame(ctx, settings, fs, pairs64);
    if (stats) {
        std::memset(stats, 0, sizeof *stats);
        stats->splats = fs.count;
        stats->visible = (uint32_t)host[kTotalVisible];
        stats->pairs = pairs64;
        stats->frame_serial = ctx->serial;
        if (timed) {
            ATHX_HIP(ctx, hipEventSynchronize(ctx->ev[7]), "wait for frame");
            stats->project_ms = elapsed_ms(ctx->ev[0], ctx->ev[1]);
            stats->depth_sort_ms = elapsed_ms(ctx->ev[1], ctx->ev[2]);
            stats->counts_ms = elapsed_ms(ctx->ev[2], ctx->ev[3]);
            stats->emit_ms = elapsed_ms(ctx->ev[3], ctx->ev[4]);
            stats->tile_sort_ms = elapsed_ms(ctx->ev[4], ctx->ev[5]);
            stats->ranges_ms = elapsed_ms(ctx->ev[5], ctx->ev[6]);
            stats->blend_ms = elapsed_ms(ctx->ev[6], ctx->ev[7]);
            stats->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - wall0).count();
        }
    }
    return ATHX_OK;
}

}  // namespace

extern "C" {

ATHX_API athx_status athx_render(athx_ctx* ctx, const athx_view* views, uint32_t nviews,
                                 const athx_gaussians* gaussians, const athx_settings* settings,
                                 const athx_targets* targets, void* stream, athx_frame_stats* stats) try {
    if (!ctx) {
        return ATHX_INVALID_ARGUMENT;
    }
    ctx->error.clear();
    FrameInstance one;
    one.g = gaussians;
    return render_frame(ctx, views, nviews, &one, 1, settings, targets, stream, stats, nullptr);
} catch (const std::bad_alloc&) {
    return fail(ctx, ATHX_OUT_OF_MEMORY, "host allocation failed");
} catch (const std::exception& e) {
    return fail(ctx, ATHX_INTERNAL_ERROR, e.what());
}

ATHX_API athx_status athx_render_instances(athx_ctx* ctx, const athx_view* views, uint32_t nviews,
                                           const athx_instance* instances, uint32_t ninstances,
                                           const athx_settings* settings, const athx_targets* targets, void* stream,
                                           athx_frame_stats* stats, uint32_t* slot_bases) try {
    if (!ctx) {
        return ATHX_INVALID_ARGUMENT;
    }
    ctx->error.clear();
    if (!instances || ninstances == 0) {
        return fail(ctx, ATHX_INVALID_ARGUMENT, "no instances");
    }
    std::vector<FrameInstance> inst(ninstances);
    for (uint32_t j = 0; j < ninstances; ++j) {
        inst[j].g = instances[j].gaussians;
        inst[j].object_to_view = instances[j].object_to_view;
        inst[j].eye = instances[j].eye;
    }
    return render_frame(ctx, views, nviews, inst.data(), ninstances, settings, targets, stream, stats, slot_bases);
} catch (const std::bad_alloc&) {
    return fail(ctx, ATHX_OUT_OF_MEMORY, "host allocation failed");
} catch (const std::exception& e) {
    return fail(ctx, ATHX_INTERNAL_ERROR, e.what());
}

ATHX_API athx_status athx_render_async(athx_ctx* ctx, const athx_view* views, uint32_t nviews,
                                       const athx_gaussians* gaussians, const athx_settings* settings,
                                       const athx_targets* targets, uint64_t pair_capacity, void* stream) try {
    if (!ctx) {
        return ATHX_INVALID_ARGUMENT;
    }
    ctx->error.clear();
    FrameShape fs{};
    FrameInstance one;
    one.g = gaussians;
    if (athx_status st = begin_frame(ctx, views, nviews, &one, 1, settings, targets, nullptr, fs); st != ATHX_OK) {
        return st;
    }
    const uint64_t cap = pair_capacity ? pair_capacity : ctx->pair_hint;
    if (cap == 0) {
        return fail(ctx, ATHX_INVALID_ARGUMENT, "no pair capacity given and no frame seen to size one from");
    }
    // Strip entries are fewer than pairs in any scene worth drawing; a frame
    // seen before gives their size.
    const uint64_t entry_cap = ctx->entry_hint ? ctx->entry_hint : cap;
    if (cap >= kMaxPairs || entry_cap >= kMaxPairs) {
        return fail(ctx, ATHX_UNSUPPORTED, "a pair capacity of 2^30 or more");
    }
    fs.capacity = (uint32_t)cap;
    set_entry_cap(fs, (uint32_t)entry_cap);
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (athx_status st = upload_views(ctx, nviews, s); st != ATHX_OK) {
        return st;
    }
    if (athx_status st = reserve_front(ctx, fs); st != ATHX_OK) {
        return st;
    }
    if (athx_status st = reserve_bins(ctx, fs); st != ATHX_OK) {
        return st;
    }
    if (athx_status st = reserve_pairs(ctx, fs); st != ATHX_OK) {
        return st;
    }
    const std::vector<ProjectArgs> pas = instance_args(&one, 1, settings, fs.count);
    if (athx_status st = enqueue_front(ctx, pas, fs, s, false); st != ATHX_OK) {
        return st;
    }
    bool strips_in_alt = false;
    if (athx_status st = enqueue_bins(ctx, fs, fs.count, settings->tile, s, false, strips_in_alt, settings->count_splats != 0); st != ATHX_OK) {
        return st;
    }
    if (athx_status st = enqueue_draw(ctx, fs, settings, s, false, strips_in_alt); st != ATHX_OK) {
        return st;
    }
    ATHX_HIP(ctx, hipMemcpyAsync(ctx->host_totals, ctx->totals.ptr, 8 * kTotalWords, hipMemcpyDeviceToHost, s),
             "read totals");
    ATHX_HIP(ctx, hipEventRecord(ctx->async_done, s), "record frame");
    ctx->async_pending = true;
    finish_frame(ctx, settings, fs, cap);
    return ATHX_OK;
} catch (const std::bad_alloc&) {
    return fail(ctx, ATHX_OUT_OF_MEMORY, "host allocation failed");
} catch (const std::exception& e) {
    return fail(ctx, ATHX_INTERNAL_ERROR, e.what());
}

ATHX_API athx_status athx_frame_pairs(athx_ctx* ctx, uint64_t* pairs, uint64_t* capacity)  try {
    if (!ctx || !pairs) {
        return ATHX_INVALID_ARGUMENT;
    }
    if (!ctx->have_frame) {
        return fail(ctx, ATHX_NOT_FOUND, "no frame has been rendered");
    }
    if (ctx->async_pending) {
        ATHX_HIP(ctx, hipEventSynchronize(ctx->async_done), "wait for the frame");
        ctx->async_pending = false;
        const uint64_t p = ctx->host_totals[kTotalPairs], e = ctx->host_totals[kTotalEntries];
        // A frame whose strip entries passed their room is incomplete even
        // when its (undercounted) pairs fit: it reports more than its capacity.
        ctx->last_true_pairs = e > ctx->last_entry_cap ? std::max<uint64_t>(p, (uint64_t)ctx->last_capacity + 1) : p;
        ctx->pair_hint = std::max<uint64_t>(ctx->pair_hint, std::min<uint64_t>(grown(p), kMaxPairs - 1));
        ctx->entry_hint = std::max<uint64_t>(ctx->entry_hint, std::min<uint64_t>(grown(e), kMaxPairs - 1));
    }
    *pairs = ctx->last_true_pairs;
    if (capacity) {
        *capacity = ctx->last_capacity;
    }
    return ATHX_OK;
} catch (const std::bad_alloc&) {
    return fail(ctx, ATHX_OUT_OF_MEMORY, "host allocation failed");
} catch (const std::exception& e) {
    return fail(ctx, ATHX_INTERNAL_ERROR, e.what());
}

ATHX_API athx_status athx_render_backward(athx_ctx* ctx, const athx_view* views, uint32_t nviews,
                                          const athx_gaussians* gaussians, const athx_settings* settings,
                                          const athx_targets* targets, uint64_t frame_serial,
                                          const float* const* dl_dcolour, const athx_grads* grads, void* stream) try {
    if (!ctx) {
        return ATHX_INVALID_ARGUMENT;
    }
    ctx->error.clear();
    if (!views || nviews == 0 || !targets || !dl_dcolour || !grads) {
        return fail(ctx, ATHX_INVALID_ARGUMENT, "no views, targets, dL/dcolour or gradients");
    }
    if (athx_status st = check_settings(ctx, settings); st != ATHX_OK) {
        return st;
    }
    if (athx_status st = check_gaussians(ctx, gaussians); st != ATHX_OK) {
        return st;
    }
    if (athx_status st = check_serial(ctx, frame_serial); st != ATHX_OK) {
        return st;
    }
    if (ctx->last_instances != 1) {
        return fail(ctx, ATHX_UNSUPPORTED, "the adjoint follows a one-cloud frame, not a multi-instance one");
    }
    if (ctx->last_under) {
        return fail(ctx, ATHX_UNSUPPORTED, "the adjoint of a frame over an under layer is not built");
    }
    if (ctx->last_n != gaussians->count || ctx->last_nviews != nviews || ctx->last_tile != settings->tile) {
        return fail(ctx, ATHX_INVALID_ARGUMENT, "the backward pass must follow a render of the same frame");
    }
    if (gaussians->relight) {
        return fail(ctx, ATHX_UNSUPPORTED, "the adjoint of the relit colour is not built: backward is SH-coloured only");
    }
    if (gaussians->packed) {
        return fail(ctx, ATHX_UNSUPPORTED, "the adjoint is taken with respect to the FP32 arrays, not a packed cloud");
    }
    for (uint32_t k = 0; k < nviews; ++k) {
        if (!targets[k].final_t || !targets[k].last || !dl_dcolour[k]) {
            return fail(ctx, ATHX_INVALID_ARGUMENT,
                        "view " + std::to_string(k) + ": final_t, last and dL/dcolour are required");
        }
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    ATHX_HIP(ctx, hipSetDevice(ctx->device), "set device");
    uint64_t tiles64 = 0;
    if (athx_status st = stage_views(ctx, views, nviews, settings, targets, dl_dcolour, tiles64); st != ATHX_OK) {
        return st;
    }
    if (tiles64 != ctx->last_tiles) {
        return fail(ctx, ATHX_INVALID_ARGUMENT, "the views differ from the rendered frame's");
    }
    if (athx_status st = upload_views(ctx, nviews, s); st != ATHX_OK) {
        return st;
    }
    const uint32_t slots = gaussians->count * nviews;
    ATHX_HIP(ctx, reserve(ctx, ctx->grad2d, (size_t)slots * kGrad2d * sizeof(float) + 16), "2D gradients");
    return enqueue_backward(ctx, project_args(gaussians, settings), nviews, slots, settings, *grads, s);
} catch (const std::bad_alloc&) {
    return fail(ctx, ATHX_OUT_OF_MEMORY, "host allocation failed");
} catch (const std::exception& e) {
    return fail(ctx, ATHX_INTERNAL_ERROR, e.what());
}

ATHX_API athx_status athx_contribution(athx_ctx* ctx, const athx_view* views, uint32_t nviews,
                                      const athx_gaussians* gaussians, const athx_settings* settings,
                                      const athx_targets* targets, uint64_t frame_serial, uint32_t mode,
                                      const float* const* dl_dcolour, uint32_t* out, void* stream) try {
    if (!ctx) {
        return ATHX_INVALID_ARGUMENT;
    }
    ctx->error.clear();
    if (mode < 1 || mode > 4) {
        return fail(ctx, ATHX_INVALID_ARGUMENT, "contribution modes are 1 to 4");
    }
    if (!views || nviews == 0 || !targets || !out || (mode != 1 && !dl_dcolour)) {
        return fail(ctx, ATHX_INVALID_ARGUMENT, "no views, targets or output (or no per-pixel input for modes 2-4)");
    }
    if (athx_status st = check_settings(ctx, settings); st != ATHX_OK) {
        return st;
    }
    if (athx_status st = check_serial(ctx, frame_serial); st != ATHX_OK) {
        return st;
    }
    if ((gaussians && ctx->last_n != gaussians->count) || ctx->last_nviews != nviews ||
        ctx->last_tile != settings->tile) {
        return fail(ctx, ATHX_INVALID_ARGUMENT, "contributions must follow a render of the same frame");
    }
    if (ctx->last_under) {
        return fail(ctx, ATHX_UNSUPPORTED, "contributions of a frame over an under layer are not built");
    }
    if (mode != 1) {
        for (uint32_t k = 0; k < nviews; ++k) {
            if (!dl_dcolour[k]) {
                return fail(ctx, ATHX_INVALID_ARGUMENT, "view " + std::to_string(k) + ": no per-pixel input");
            }
        }
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    ATHX_HIP(ctx, hipSetDevice(ctx->device), "set device");
    uint64_t tiles64 = 0;
    if (athx_status st = stage_views(ctx, views, nviews, settings, targets, mode != 1 ? dl_dcolour : nullptr, tiles64);
        st != ATHX_OK) {
        return st;
    }
    if (tiles64 != ctx->last_tiles) {
        return fail(ctx, ATHX_INVALID_ARGUMENT, "the views differ from the rendered frame's");
    }
    if (athx_status st = upload_views(ctx, nviews, s); st != ATHX_OK) {
        return st;
    }
    const uint64_t words = (uint64_t)ctx->last_n * nviews * (mode == 4 ? 4u : mode == 1 ? 1u : 2u);
    ATHX_HIP(ctx, launch_fill_u32(out, 0u, words, s), "clear contributions");
    ATHX_HIP(ctx, launch_contribution(as<DView>(ctx->views), as<uint32_t>(ctx->tile_bases), nviews, ctx->last_tiles,
                                      settings->tile, settings->batch, as<uint32_t>(ctx->ranges),
                                      as<uint32_t>(ctx->pair_vals), as<float>(ctx->records), mode, out, s),
             "contribution");
    return ATHX_OK;
} catch (const std::bad_alloc&) {
    return fail(ctx, ATHX_OUT_OF_MEMORY, "host allocation failed");
} catch (const std::exception& e) {
    return fail(ctx, ATHX_INTERNAL_ERROR, e.what());
}

ATHX_API uint64_t athx_frame_serial(const athx_ctx* ctx) { return ctx && ctx->have_frame ? ctx->serial : 0; }

ATHX_API athx_status athx_latest_counters(athx_ctx* ctx, athx_counters* out)  try {
    if (!ctx || !out) {
        return ATHX_INVALID_ARGUMENT;
    }
    ctx->error.clear();
    const athx_ctx::CounterCopy* newest = nullptr;
    for (const athx_ctx::CounterCopy& c : ctx->counter_ring) {
        if (c.seq == 0 || (newest && c.seq < newest->seq)) {
            continue;
        }
        const hipError_t q = hipEventQuery(c.done);
        if (q == hipSuccess) {
            newest = &c;
        } else if (q != hipErrorNotReady) {
            return hip_fail(ctx, q, "counters readback");
        }
    }
    if (!newest) {
        return fail(ctx, ATHX_NOT_FOUND, "no frame counted with count_splats has finished");
    }
    const uint32_t* w = newest->host;
    std::memset(out, 0, sizeof *out);
    out->tag = newest->tag;
    out->slots = newest->slots;
    out->visible = w[1];
    out->pairs = (uint64_t)w[2] | ((uint64_t)w[3] << 32);
    out->max_tiles = w[4];
    for (uint32_t r = 0; r < ATHX_CULL_REASONS; ++r) {
        out->culled[r] = w[8 + r];
    }
    out->clouds = newest->clouds;
    for (uint32_t j = 0; j < newest->clouds; ++j) {
        out->cloud_visible[j] = w[16 + 2 * j];
        out->cloud_pairs[j] = w[17 + 2 * j];
    }
    return ATHX_OK;
} catch (const std::bad_alloc&) {
    return fail(ctx, ATHX_OUT_OF_MEMORY, "host allocation failed");
} catch (const std::exception& e) {
    return fail(ctx, ATHX_INTERNAL_ERROR, e.what());
}

ATHX_API athx_status athx_frame_info_get(athx_ctx* ctx, athx_frame_info* out)  try {
    if (!ctx || !out) {
        return ATHX_INVALID_ARGUMENT;
    }
    uint64_t pairs = 0, cap = 0;
    if (athx_status st = athx_frame_pairs(ctx, &pairs, &cap); st != ATHX_OK) {
        return st;
    }
    out->slots = ctx->last_n * ctx->last_nviews;
    out->tiles = ctx->last_tiles;
    out->key_bits = ctx->last_key_bits;
    out->pairs = std::min(pairs, cap);
    out->strips = ctx->last_strips;
    out->entries = ctx->host_totals[kTotalEntries];
    return ATHX_OK;
} catch (const std::bad_alloc&) {
    return fail(ctx, ATHX_OUT_OF_MEMORY, "host allocation failed");
} catch (const std::exception& e) {
    return fail(ctx, ATHX_INTERNAL_ERROR, e.what());
}

ATHX_API athx_status athx_copy_frame(athx_ctx* ctx, athx_frame_part part, void* dst, uint64_t bytes, void* stream)  try {
    if (!ctx) {
        return ATHX_INVALID_ARGUMENT;
    }
    ctx->error.clear();
    if (!ctx->have_frame) {
        return fail(ctx, ATHX_NOT_FOUND, "no frame has been rendered");
    }
    uint64_t true_pairs = 0, cap = 0;
    if (athx_status st = athx_frame_pairs(ctx, &true_pairs, &cap); st != ATHX_OK) {
        return st;
    }
    const uint64_t pairs = std::min(true_pairs, cap);
    const uint64_t slots = (uint64_t)ctx->last_n * ctx->last_nviews;
    const void* src = nullptr;
    uint64_t size = 0;
    switch (part) {
    case ATHX_PART_RECORDS: src = ctx->records.ptr; size = slots * kRecordFloats * 4; break;
    case ATHX_PART_TOUCHED: size = slots * 4; break;  // the exact counts, taken now
    case ATHX_PART_DEPTH_KEYS: src = ctx->depth_keys.ptr; size = slots * 4; break;
    case ATHX_PART_OFFSETS: size = slots * 4; break;  // their scan in depth order, taken now
    case ATHX_PART_PAIR_KEYS: size = pairs * 8; break;  // rebuilt from the ranges and depth keys
    case ATHX_PART_PAIR_VALUES: src = ctx->pair_vals.ptr; size = pairs * 4; break;
    case ATHX_PART_RANGES: src = ctx->ranges.ptr; size = (uint64_t)ctx->last_tiles * 8; break;
    default: return fail(ctx, ATHX_INVALID_ARGUMENT, "unknown frame part");
    }
    if (bytes != size) {
        return fail(ctx, ATHX_INVALID_ARGUMENT,
                    "part is " + std::to_string(size) + " bytes, asked for " + std::to_string(bytes));
    }
    if (size == 0) {
        return ATHX_OK;
    }
    if (!dst) {
        return fail(ctx, ATHX_INVALID_ARGUMENT, "null destination");
    }
    ATHX_HIP(ctx, hipSetDevice(ctx->device), "set device");
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (part == ATHX_PART_PAIR_KEYS) {
        ATHX_HIP(ctx, launch_pair_keys(as<uint32_t>(ctx->ranges), ctx->last_tiles, as<uint32_t>(ctx->pair_vals),
                                       as<uint32_t>(ctx->depth_keys), static_cast<uint64_t*>(dst), s),
                 "pair keys");
        return ATHX_OK;
    }
    if (part == ATHX_PART_TOUCHED || part == ATHX_PART_OFFSETS) {
        // The frame never counts a slot's tiles (it bins them); the exact
        // counts (splat_project.slang:764-776) and their scan in depth order
        // (splat_gather_counts.slang:11-19) are taken on request. A slot not
        // drawn has offset 0.
        if (athx_status st = reserve_exact(ctx, (uint32_t)slots); st != ATHX_OK) {
            return st;
        }
        if (athx_status st = enqueue_exact(ctx, ctx->last_n, ctx->last_nviews, ctx->last_tile, s); st != ATHX_OK) {
            return st;
        }
        src = ctx->exact.ptr;
        if (part == ATHX_PART_OFFSETS) {
            uint64_t* totals = as<uint64_t>(ctx->totals);
            ATHX_HIP(ctx, launch_fill_u32(as<uint32_t>(ctx->exact_offsets), 0u, slots, s), "clear offsets");
            const uint32_t* order = sorted_slots_of(ctx, ctx->last_sorted_in_alt);
            const uint64_t* order_total = totals + kTotalVisible;
            if (ctx->last_grouped) {
                // The frame sorted each view apart: the whole depth order
                // (TileRasterizer.cpp:1111-1123) is sorted here, from the
                // slots' depth keys.
                const uint32_t n32 = (uint32_t)slots;
                ATHX_HIP(ctx, reserve(ctx, ctx->order_keys, slots * 4 + 16), "depth order");
                ATHX_HIP(ctx, reserve(ctx, ctx->order_slots, slots * 4 + 16), "depth order");
                ATHX_HIP(ctx, reserve(ctx, ctx->order_alt_keys, slots * 4 + 16), "depth order");
                ATHX_HIP(ctx, reserve(ctx, ctx->order_alt_slots, slots * 4 + 16), "depth order");
                ATHX_HIP(ctx, reserve(ctx, ctx->order_meta, sort_meta_words(n32, kDepthPasses) * 4 + 16),
                         "depth order");
                ATHX_HIP(ctx, launch_compact(as<uint32_t>(ctx->touched), as<uint32_t>(ctx->depth_keys), n32,
                                             as<uint32_t>(ctx->order_keys), as<uint32_t>(ctx->order_slots),
                                             totals + kTotalOrder, as<uint32_t>(ctx->exact_status), s),
                         "depth order compaction");
                const SortScratch osc = sort_scratch(ctx->order_meta, ctx->order_alt_keys.ptr,
                                                     as<uint32_t>(ctx->order_alt_slots), n32, kDepthPasses);
                bool oalt = false;
                ATHX_HIP(ctx, launch_sort(as<uint32_t>(ctx->order_keys), as<uint32_t>(ctx->order_slots),
                                          totals + kTotalOrder, n32, 0, kDepthBits, osc, s, &oalt),
                         "depth order sort");
                order = as<uint32_t>(oalt ? ctx->order_alt_slots : ctx->order_slots);
                order_total = totals + kTotalOrder;
            }
            ATHX_HIP(ctx, launch_gather_scan(as<uint32_t>(ctx->exact), order, order_total, (uint32_t)slots,
                                             as<uint32_t>(ctx->exact_offsets), totals + kTotalExact,
                                             as<uint32_t>(ctx->exact_status), s),
                     "exact offsets");
            src = ctx->exact_offsets.ptr;
        }
    }
    ATHX_HIP(ctx, hipMemcpyAsync(dst, src, size, hipMemcpyDeviceToDevice, s), "copy frame part");
    return ATHX_OK;
} catch (const std::bad_alloc&) {
    return fail(ctx, ATHX_OUT_OF_MEMORY, "host allocation failed");
} catch (const std::exception& e) {
    return fail(ctx, ATHX_INTERNAL_ERROR, e.what());
}

ATHX_API athx_status athx_exclusive_scan(athx_ctx* ctx, const uint32_t* in, uint32_t* out, uint64_t* total,
                                         uint32_t n, void* stream)  try {
    if (!ctx) {
        return ATHX_INVALID_ARGUMENT;
    }
    ctx->error.clear();
    if ((n != 0 && (!in || !out)) || !total) {
        return fail(ctx, ATHX_INVALID_ARGUMENT, "null buffer");
    }
    if (in == out) {
        return fail(ctx, ATHX_INVALID_ARGUMENT, "input and output must be different buffers (PrefixSum.h:41-42)");
    }
    ATHX_HIP(ctx, hipSetDevice(ctx->device), "set device");
    ATHX_HIP(ctx, reserve(ctx, ctx->api_scan_status, scan_status_words(n) * 4), "scan status");
    ATHX_HIP(ctx, launch_scan(in, out, total, n, as<uint32_t>(ctx->api_scan_status), static_cast<hipStream_t>(stream)),
             "scan");
    return ATHX_OK;
} catch (const std::bad_alloc&) {
    return fail(ctx, ATHX_OUT_OF_MEMORY, "host allocation failed");
} catch (const std::exception& e) {
    return fail(ctx, ATHX_INTERNAL_ERROR, e.what());
}

ATHX_API athx_status athx_exclusive_scan_noalloc(athx_ctx* ctx, const uint32_t* in, uint32_t* out, uint64_t* total,
                                                 uint32_t n, void* stream) {
    if (!ctx) {
        return ATHX_INVALID_ARGUMENT;
    }
    ctx->error.clear();
    if ((n != 0 && (!in || !out)) || !total) {
        ctx->error.assign("null buffer");
        return ATHX_INVALID_ARGUMENT;
    }
    if (in == out) {
        ctx->error.assign("input and output must be different buffers (PrefixSum.h:41-42)");
        return ATHX_INVALID_ARGUMENT;
    }
    if (n > kNoAllocScanMax || scan_status_words(n) > kNoAllocScanWords) {
        ctx->error.assign("athx_exclusive_scan_noalloc scans at most 2^24 elements (its static scratch)");
        return ATHX_INVALID_ARGUMENT;
    }
    if (hipSetDevice(ctx->device) != hipSuccess) {
        ctx->error.assign("set device");
        return ATHX_DEVICE_FAILURE;
    }
    // The status words are cleared by a fill kernel inside launch_scan (or,
    // for a long scan, written whole as partition sums) before the scan reads
    // them, so whatever the last call left does not matter.
    if (launch_scan(in, out, total, n, ctx->noalloc_scratch, static_cast<hipStream_t>(stream)) != hipSuccess) {
        ctx->error.assign("scan launch failed");
        return ATHX_DEVICE_FAILURE;
    }
    return ATHX_OK;
}

ATHX_API athx_status athx_sort_pairs(athx_ctx* ctx, uint64_t* keys, uint32_t* values, uint32_t n, uint32_t key_bits,
                                     void* stream)  try {
    if (!ctx) {
        return ATHX_INVALID_ARGUMENT;
    }
    ctx->error.clear();
    if (key_bits == 0 || key_bits > 64) {
        return fail(ctx, ATHX_INVALID_ARGUMENT, "key_bits must be in 1..64 (RadixSort.h:61)");
    }
    if (n != 0 && (!keys || !values)) {
        return fail(ctx, ATHX_INVALID_ARGUMENT, "null buffer");
    }
    if (n >= (1u << 30)) {
        return fail(ctx, ATHX_UNSUPPORTED, "at most 2^30 - 1 pairs a sort");
    }
    if (n < 2) {
        return ATHX_OK;
    }
    const uint32_t passes = div_up(key_bits, 8);
    ATHX_HIP(ctx, hipSetDevice(ctx->device), "set device");
    ATHX_HIP(ctx, reserve(ctx, ctx->api_alt_keys, (size_t)n * 8), "sort keys");
    ATHX_HIP(ctx, reserve(ctx, ctx->api_alt_vals, (size_t)n * 4), "sort values");
    ATHX_HIP(ctx, reserve(ctx, ctx->api_sort_meta, sort_meta_words(n, passes) * 4), "sort metadata");
    ATHX_HIP(ctx, reserve(ctx, ctx->api_count, 16), "sort count");
    const SortScratch sc = sort_scratch(ctx->api_sort_meta, ctx->api_alt_keys.ptr,
                                        as<uint32_t>(ctx->api_alt_vals), n, passes);
    hipStream_t s = static_cast<hipStream_t>(stream);
    // The sort reads its count on the device: put n there (a kernel, not a
    // memset, so a captured graph replays it faithfully).
    ATHX_HIP(ctx, launch_fill_u32(as<uint32_t>(ctx->api_count), n, 1, s), "sort count");
    ATHX_HIP(ctx, launch_fill_u32(as<uint32_t>(ctx->api_count) + 1, 0u, 1, s), "sort count");
    ATHX_HIP(ctx, launch_sort(keys, values, as<uint64_t>(ctx->api_count), n, 0, key_bits, sc, s), "sort");
    return ATHX_OK;
} catch (const std::bad_alloc&) {
    return fail(ctx, ATHX_OUT_OF_MEMORY, "host allocation failed");
} catch (const std::exception& e) {
    return fail(ctx, ATHX_INTERNAL_ERROR, e.what());
}

ATHX_API athx_status athx_ingest_ply(athx_ctx* ctx, const void* vertices, const athx_ply_layout* layout,
                                     uint32_t max_degree, uint32_t keep_faint, const athx_ingest_out* out,
                                     uint32_t* kept) try {
    if (!ctx) {
        return ATHX_INVALID_ARGUMENT;
    }
    ctx->error.clear();
    if (!layout || !out || !kept) {
        return fail(ctx, ATHX_INVALID_ARGUMENT, "null layout, output or kept");
    }
    *kept = 0;
    const athx_ply_layout& L = *layout;
    if (L.count > 0xFFFFFFFFull || L.properties > ATHX_PLY_MAX_PROPERTIES || L.stride == 0) {
        return fail(ctx, ATHX_INVALID_ARGUMENT, "layout out of range");
    }
    if (L.count == 0) {
        return ATHX_OK;
    }
    if (!vertices || !out->positions || !out->shape) {
        return fail(ctx, ATHX_INVALID_ARGUMENT, "null vertices, positions or shape");
    }
    IngestFields f{};
    f.count = (uint32_t)L.count;
    f.stride = L.stride;
    f.keep_faint = keep_faint ? 1u : 0u;
    if (L.opacity_mode > 2 || L.scale_mode > 2 || L.colour_mode > 4 || L.rotation_mode > 1) {
        return fail(ctx, ATHX_UNSUPPORTED, "an encoding this ingest does not decode");
    }
    f.opacity_mode = L.opacity_mode;
    f.scale_mode = L.scale_mode;
    f.colour_mode = L.colour_mode;
    f.rotation_mode = L.rotation_mode;
    const auto code_of = [&](uint32_t index, uint32_t& code) {
        if (index >= L.properties || L.type[index] > ATHX_PLY_I32) {
            return false;
        }
        const uint32_t bytes = L.type[index] == ATHX_PLY_F64 ? 8u
                               : (L.type[index] == ATHX_PLY_U8 || L.type[index] == ATHX_PLY_I8) ? 1u
                               : (L.type[index] == ATHX_PLY_U16 || L.type[index] == ATHX_PLY_I16) ? 2u
                                                                                                   : 4u;
        if ((uint64_t)L.offset[index] + bytes > L.stride || L.offset[index] > 0xFFFFFFu) {
            return false;
        }
        code = L.offset[index] | (L.type[index] << 24);
        return true;
    };
    const uint32_t fields[kIngestBase] = {L.x,        L.y,        L.z,        L.opacity, L.scale[0],
                                          L.scale[1], L.scale[2], L.rot[0],   L.rot[1],  L.rot[2],
                                          L.rot[3],   L.dc[0],    L.dc[1],    L.dc[2]};
    for (uint32_t k = 0; k < kIngestBase; ++k) {
        if (!code_of(fields[k], f.code[k])) {
            return fail(ctx, ATHX_INVALID_ARGUMENT, "a decoded field outside the vertex");
        }
    }
    // Rest coefficients kept a colour: the file's, capped by the degree
    // (0, 3, 8, 15 for degrees 0..3); packed rgb a basis.
    static const uint32_t kBasisOfDegree[4] = {0, 3, 8, 15};
    const uint32_t cap = kBasisOfDegree[max_degree < 3 ? max_degree : 3];
    if (L.rest_per_colour != 0 && L.rest_per_colour != 3 && L.rest_per_colour != 8 && L.rest_per_colour != 15) {
        return fail(ctx, ATHX_INVALID_ARGUMENT, "rest_per_colour must be 0, 3, 8 or 15");
    }
    const uint32_t keep = L.rest_per_colour < cap ? L.rest_per_colour : cap;
    f.rest = keep * 3;
    f.sh_words = (f.rest + 1) / 2;
    for (uint32_t basis = 0; basis < keep; ++basis) {
        for (uint32_t colour = 0; colour < 3; ++colour) {
            if (!code_of(L.rest_base + colour * L.rest_per_colour + basis, f.code[kIngestBase + basis * 3 + colour])) {
                return fail(ctx, ATHX_INVALID_ARGUMENT, "a rest coefficient outside the vertex");
            }
        }
    }
    if (f.rest != 0 && !out->sh) {
        return fail(ctx, ATHX_INVALID_ARGUMENT, "sh is null with rest coefficients kept");
    }
    athx_ingest_out o = *out;
    if (f.rest == 0) {
        o.sh = nullptr;
    }
    ATHX_HIP(ctx, hipSetDevice(ctx->device), "set device");
    ATHX_HIP(ctx, reserve(ctx, ctx->ingest_keep, (size_t)f.count * 4 + 16), "ingest flags");
    ATHX_HIP(ctx, reserve(ctx, ctx->ingest_dest, (size_t)f.count * 4 + 16), "ingest places");
    ATHX_HIP(ctx, reserve(ctx, ctx->ingest_total, 16), "ingest total");
    ATHX_HIP(ctx, reserve(ctx, ctx->api_scan_status, scan_status_words(f.count) * 4), "scan status");
    const hipStream_t s = nullptr;
    const uint8_t* bytes = static_cast<const uint8_t*>(vertices);
    ATHX_HIP(ctx, launch_ingest_flags(bytes, f, as<uint32_t>(ctx->ingest_keep), s), "ingest flags");
    ATHX_HIP(ctx, launch_scan(as<uint32_t>(ctx->ingest_keep), as<uint32_t>(ctx->ingest_dest),
                              as<uint64_t>(ctx->ingest_total), f.count, as<uint32_t>(ctx->api_scan_status), s),
             "ingest scan");
    ATHX_HIP(ctx, launch_ingest_pack(bytes, f, as<uint32_t>(ctx->ingest_keep), as<uint32_t>(ctx->ingest_dest), o, s),
             "ingest pack");
    uint64_t total = 0;
    ATHX_HIP(ctx, hipMemcpyAsync(&total, ctx->ingest_total.ptr, sizeof total, hipMemcpyDeviceToHost, s), "kept");
    ATHX_HIP(ctx, hipStreamSynchronize(s), "ingest wait");
    *kept = (uint32_t)total;
    return ATHX_OK;
} catch (const std::bad_alloc&) {
    return fail(ctx, ATHX_OUT_OF_MEMORY, "host allocation failed");
} catch (const std::exception& e) {
    return fail(ctx, ATHX_INTERNAL_ERROR, e.what());
}


struct athx_graph {
    athx_ctx* ctx = nullptr;
    hipStream_t capture = nullptr;
    hipGraph_t forward = nullptr, backward = nullptr;
    hipGraphExec_t forward_exec = nullptr, backward_exec = nullptr;
};

ATHX_API void athx_graph_destroy(athx_graph* g) {
    if (!g) {
        return;
    }
    if (g->ctx) {
        (void)hipSetDevice(g->ctx->device);
        (void)hipDeviceSynchronize();
    }
    if (g->forward_exec) {
        (void)hipGraphExecDestroy(g->forward_exec);
    }
    if (g->backward_exec) {
        (void)hipGraphExecDestroy(g->backward_exec);
    }
    if (g->forward) {
        (void)hipGraphDestroy(g->forward);
    }
    if (g->backward) {
        (void)hipGraphDestroy(g->backward);
    }
    if (g->capture) {
        (void)hipStreamDestroy(g->capture);
    }
    athx_destroy(g->ctx);
    delete g;
}

namespace {

athx_status graph_fail(athx_graph* g, athx_status st, char* error, uint32_t error_bytes) {
    if (error && error_bytes) {
        const std::string msg = g && g->ctx ? g->ctx->error : std::string("graph creation failed");
        std::snprintf(error, error_bytes, "%s", msg.c_str());
    }
    athx_graph_destroy(g);
    return st;
}

athx_status capture_forward(athx_graph* g, const ProjectArgs& pa, const FrameShape& fs, const athx_settings* settings) {
    athx_ctx* ctx = g->ctx;
    ATHX_HIP(ctx, hipStreamBeginCapture(g->capture, hipStreamCaptureModeThreadLocal), "begin capture");
    athx_status st = enqueue_front(ctx, std::vector<ProjectArgs>{pa}, fs, g->capture, false);
    bool strips_in_alt = false;
    if (st == ATHX_OK) {
        st = enqueue_bins(ctx, fs, fs.count, settings->tile, g->capture, false, strips_in_alt, settings->count_splats != 0);
    }
    if (st == ATHX_OK) {
        st = enqueue_draw(ctx, fs, settings, g->capture, false, strips_in_alt, true);
    }
    hipError_t copy = hipSuccess;
    if (st == ATHX_OK) {
        copy = hipMemcpyAsync(ctx->host_totals, ctx->totals.ptr, 8 * kTotalWords, hipMemcpyDeviceToHost, g->capture);
    }
    hipError_t end = hipStreamEndCapture(g->capture, &g->forward);
    if (st != ATHX_OK) {
        return st;
    }
    ATHX_HIP(ctx, copy, "capture the pair readback");
    ATHX_HIP(ctx, end, "end capture");
    ATHX_HIP(ctx, hipGraphInstantiate(&g->forward_exec, g->forward, nullptr, nullptr, 0), "instantiate the frame");
    return ATHX_OK;
}

athx_status capture_backward(athx_graph* g, const ProjectArgs& pa, const FrameShape& fs,
                             const athx_settings* settings, const athx_grads& grads) {
    athx_ctx* ctx = g->ctx;
    ATHX_HIP(ctx, hipStreamBeginCapture(g->capture, hipStreamCaptureModeThreadLocal), "begin capture");
    const athx_status st = enqueue_backward(ctx, pa, fs.nviews, fs.slots, settings, grads, g->capture);
    hipError_t end = hipStreamEndCapture(g->capture, &g->backward);
    if (st != ATHX_OK) {
        return st;
    }
    ATHX_HIP(ctx, end, "end capture");
    ATHX_HIP(ctx, hipGraphInstantiate(&g->backward_exec, g->backward, nullptr, nullptr, 0),
             "instantiate the adjoint");
    return ATHX_OK;
}

}  // namespace

ATHX_API athx_status athx_graph_create(int device, const athx_view* views, uint32_t nviews,
                                       const athx_gaussians* gaussians, const athx_settings* settings,
                                       const athx_targets* targets, uint64_t pair_capacity,
                                       const float* const* dl_dcolour, const athx_grads* grads, athx_graph** out,
                                       char* error, uint32_t error_bytes) try {
    if (!out) {
        return ATHX_INVALID_ARGUMENT;
    }
    *out = nullptr;
    athx_graph* g = new athx_graph();
    if (athx_status st = athx_create(device, &g->ctx); st != ATHX_OK) {
        return graph_fail(g, st, error, error_bytes);
    }
    athx_ctx* ctx = g->ctx;
    ctx->exact_growth = true;
    const bool with_backward = dl_dcolour != nullptr && grads != nullptr;
    if (with_backward && gaussians && (gaussians->relight || gaussians->packed)) {
        return graph_fail(g, fail(ctx, ATHX_UNSUPPORTED, "the adjoint of the relit colour is not built"), error,
                          error_bytes);
    }
    if (with_backward) {
        for (uint32_t k = 0; views && targets && k < nviews; ++k) {
            if (!targets[k].final_t || !targets[k].last || !dl_dcolour[k]) {
                return graph_fail(g, fail(ctx, ATHX_INVALID_ARGUMENT, "the adjoint needs final_t, last and dL/dcolour"),
                                  error, error_bytes);
            }
        }
    }
    if (pair_capacity == 0) {
        // Size from one eager frame of the same inputs.
        athx_frame_stats stats{};
        if (athx_status st = athx_render(ctx, views, nviews, gaussians, settings, targets, nullptr, &stats);
            st != ATHX_OK) {
            return graph_fail(g, st, error, error_bytes);
        }
        pair_capacity = std::max<uint64_t>(stats.pairs + stats.pairs / 2, 1024);
    }
    const uint64_t entry_capacity = ctx->entry_hint ? ctx->entry_hint : pair_capacity;
    if (pair_capacity >= kMaxPairs || entry_capacity >= kMaxPairs) {
        return graph_fail(g, fail(ctx, ATHX_UNSUPPORTED, "a pair capacity of 2^30 or more"), error, error_bytes);
    }
    FrameShape fs{};
    FrameInstance one;
    one.g = gaussians;
    if (athx_status st = begin_frame(ctx, views, nviews, &one, 1, settings, targets,
                                     with_backward ? dl_dcolour : nullptr, fs);
        st != ATHX_OK) {
        return graph_fail(g, st, error, error_bytes);
    }
    fs.capacity = (uint32_t)pair_capacity;
    set_entry_cap(fs, (uint32_t)entry_capacity);
    if (athx_status st = reserve_front(ctx, fs); st != ATHX_OK) {
        return graph_fail(g, st, error, error_bytes);
    }
    if (athx_status st = reserve_bins(ctx, fs); st != ATHX_OK) {
        return graph_fail(g, st, error, error_bytes);
    }
    if (athx_status st = reserve_pairs(ctx, fs); st != ATHX_OK) {
        return graph_fail(g, st, error, error_bytes);
    }
    if (settings && settings->count_splats) {
        if (athx_status st = reserve_exact(ctx, fs.slots); st != ATHX_OK) {
            return graph_fail(g, st, error, error_bytes);
        }
        if (hipError_t e = reserve(ctx, ctx->rank_tiles, (size_t)fs.slots * 4 + 16); e != hipSuccess) {
            return graph_fail(g, hip_fail(ctx, e, "rank tiles"), error, error_bytes);
        }
    }
    if (with_backward) {
        if (hipError_t e = reserve(ctx, ctx->grad2d, (size_t)fs.slots * kGrad2d * sizeof(float) + 16);
            e != hipSuccess) {
            return graph_fail(g, hip_fail(ctx, e, "2D gradients"), error, error_bytes);
        }
    }
    // The views once, before capture: the graph reads them from the device.
    if (hipError_t e = reserve(ctx, ctx->views, sizeof(DView) * nviews); e != hipSuccess) {
        return graph_fail(g, hip_fail(ctx, e, "views"), error, error_bytes);
    }
    if (hipError_t e = reserve(ctx, ctx->tile_bases, sizeof(uint32_t) * 2 * (nviews + 1)); e != hipSuccess) {
        return graph_fail(g, hip_fail(ctx, e, "tile bases"), error, error_bytes);
    }
    if (hipMemcpy(ctx->views.ptr, ctx->host_views, sizeof(DView) * nviews, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(ctx->tile_bases.ptr, ctx->host_tile_bases, sizeof(uint32_t) * 2 * (nviews + 1),
                  hipMemcpyHostToDevice) != hipSuccess) {
        return graph_fail(g, fail(ctx, ATHX_DEVICE_FAILURE, "upload views"), error, error_bytes);
    }
    if (hipStreamCreateWithFlags(&g->capture, hipStreamNonBlocking) != hipSuccess) {
        return graph_fail(g, fail(ctx, ATHX_DEVICE_FAILURE, "capture stream"), error, error_bytes);
    }
    const ProjectArgs pa = project_args(gaussians, settings);
    finish_frame(ctx, settings, fs, 0);
    if (athx_status st = capture_forward(g, pa, fs, settings); st != ATHX_OK) {
        return graph_fail(g, st, error, error_bytes);
    }
    if (with_backward) {
        if (athx_status st = capture_backward(g, pa, fs, settings, *grads); st != ATHX_OK) {
            return graph_fail(g, st, error, error_bytes);
        }
    }
    ctx->have_frame = false;  // nothing has been drawn until the first replay
    *out = g;
    return ATHX_OK;
} catch (const std::bad_alloc&) {
    return ATHX_OUT_OF_MEMORY;
}

ATHX_API athx_status athx_graph_launch(athx_graph* g, void* stream)  try {
    if (!g || !g->forward_exec) {
        return ATHX_INVALID_ARGUMENT;
    }
    athx_ctx* ctx = g->ctx;
    hipStream_t s = static_cast<hipStream_t>(stream);
    ATHX_HIP(ctx, hipSetDevice(ctx->device), "set device");
    ATHX_HIP(ctx, hipGraphLaunch(g->forward_exec, s), "launch the frame");
    ATHX_HIP(ctx, hipEventRecord(ctx->async_done, s), "record frame");
    ctx->async_pending = true;
    ctx->have_frame = true;
    return ATHX_OK;
} catch (const std::bad_alloc&) {
    return fail(g ? g->ctx : nullptr, ATHX_OUT_OF_MEMORY, "host allocation failed");
} catch (const std::exception& e) {
    return fail(g ? g->ctx : nullptr, ATHX_INTERNAL_ERROR, e.what());
}

ATHX_API athx_status athx_graph_launch_backward(athx_graph* g, void* stream)  try {
    if (!g || !g->backward_exec) {
        return ATHX_INVALID_ARGUMENT;
    }
    if (!g->ctx->have_frame) {
        return fail(g->ctx, ATHX_INVALID_ARGUMENT, "replay the frame before its adjoint");
    }
    ATHX_HIP(g->ctx, hipSetDevice(g->ctx->device), "set device");
    ATHX_HIP(g->ctx, hipGraphLaunch(g->backward_exec, static_cast<hipStream_t>(stream)), "launch the adjoint");
    return ATHX_OK;
} catch (const std::bad_alloc&) {
    return fail(g ? g->ctx : nullptr, ATHX_OUT_OF_MEMORY, "host allocation failed");
} catch (const std::exception& e) {
    return fail(g ? g->ctx : nullptr, ATHX_INTERNAL_ERROR, e.what());
}

ATHX_API athx_status athx_graph_pairs(athx_graph* g, uint64_t* pairs, uint64_t* capacity)  try {
    if (!g) {
        return ATHX_INVALID_ARGUMENT;
    }
    return athx_frame_pairs(g->ctx, pairs, capacity);
} catch (const std::bad_alloc&) {
    return fail(g ? g->ctx : nullptr, ATHX_OUT_OF_MEMORY, "host allocation failed");
} catch (const std::exception& e) {
    return fail(g ? g->ctx : nullptr, ATHX_INTERNAL_ERROR, e.what());
}

}  // extern "C"

namespace athx {
namespace {

// The percentile the reference reads off a 256-bin histogram
// (ReferenceRenderer.cpp:292-312, 404-422): the first bin whose running
// count reaches pixels - pixels / 100.
uint32_t p99_bin(const uint32_t bins[256], uint64_t pixels) {
    const uint64_t at = pixels - pixels / 100;
    uint64_t running = 0;
    for (uint32_t k = 0; k < 256; ++k) {
        running += bins[k];
        if (running >= at) {
            return k;
        }
    }
    return 255;
}

athx_status compare_setup(athx_ctx* ctx, const float* a, const float* b, uint32_t width, uint32_t height,
                          const void* out) {
    if (!a || !b || !out) {
        return fail(ctx, ATHX_INVALID_ARGUMENT, "null image or result");
    }
    if (width == 0 || height == 0) {
        return fail(ctx, ATHX_INVALID_ARGUMENT, "an empty image");
    }
    ATHX_HIP(ctx, hipSetDevice(ctx->device), "set device");
    ATHX_HIP(ctx, reserve(ctx, ctx->cmp_rows, (size_t)height * 256 * 4), "comparison rows");
    ATHX_HIP(ctx, reserve(ctx, ctx->cmp_row_error, (size_t)height * 4 + 16), "comparison row errors");
    ATHX_HIP(ctx, reserve(ctx, ctx->cmp_bins, 256 * 4 + 16), "comparison bins");
    return ATHX_OK;
}

}  // namespace
}  // namespace athx

using namespace athx;

extern "C" {

ATHX_API athx_status athx_compare_images(athx_ctx* ctx, const float* a, const float* b, uint32_t width,
                                         uint32_t height, athx_image_difference* out)  try {
    if (!ctx) {
        return ATHX_INVALID_ARGUMENT;
    }
    ctx->error.clear();
    if (athx_status st = compare_setup(ctx, a, b, width, height, out); st != ATHX_OK) {
        return st;
    }
    ATHX_HIP(ctx, launch_compare_codes(a, b, width, height, as<uint32_t>(ctx->cmp_rows), as<uint32_t>(ctx->cmp_bins),
                                       nullptr),
             "compare");
    uint32_t bins[256];
    ATHX_HIP(ctx, hipMemcpy(bins, ctx->cmp_bins.ptr, sizeof bins, hipMemcpyDeviceToHost), "read bins");
    athx_image_difference d{};
    for (uint32_t k = 0; k < 256; ++k) {
        d.pixels += bins[k];
        if (bins[k] > 0) {
            d.max = k;
        }
        if (k > 2) {
            d.over2 += bins[k];
        }
    }
    d.p99 = p99_bin(bins, d.pixels);
    *out = d;
    return ATHX_OK;
} catch (const std::bad_alloc&) {
    return fail(ctx, ATHX_OUT_OF_MEMORY, "host allocation failed");
} catch (const std::exception& e) {
    return fail(ctx, ATHX_INTERNAL_ERROR, e.what());
}

ATHX_API athx_status athx_compare_hdr(athx_ctx* ctx, const float* a, const float* b, uint32_t width, uint32_t height,
                                      athx_hdr_difference* out)  try {
    if (!ctx) {
        return ATHX_INVALID_ARGUMENT;
    }
    ctx->error.clear();
    if (athx_status st = compare_setup(ctx, a, b, width, height, out); st != ATHX_OK) {
        return st;
    }
    float* row_error = as<float>(ctx->cmp_row_error);
    ATHX_HIP(ctx, launch_compare_hdr(a, b, width, height, as<uint32_t>(ctx->cmp_rows), row_error,
                                     as<uint32_t>(ctx->cmp_bins), row_error + height, nullptr),
             "compare hdr");
    uint32_t bins[256];
    float error = 0.0f;
    ATHX_HIP(ctx, hipMemcpy(bins, ctx->cmp_bins.ptr, sizeof bins, hipMemcpyDeviceToHost), "read bins");
    ATHX_HIP(ctx, hipMemcpy(&error, row_error + height, sizeof error, hipMemcpyDeviceToHost), "read error");
    // Bin k > 0 holds relative differences below 2^(k / 8 - 16).
    const auto upper = [](uint32_t bin) { return bin == 0 ? 0.0 : std::pow(2.0, bin / 8.0 - 16.0); };
    athx_hdr_difference d{};
    for (uint32_t k = 0; k < 256; ++k) {
        d.pixels += bins[k];
        if (bins[k] > 0) {
            d.max_relative = upper(k);
        }
    }
    d.p99_relative = upper(p99_bin(bins, d.pixels));
    d.rel_mse = d.pixels > 0 ? (double)error / (double)d.pixels : 0.0;
    *out = d;
    return ATHX_OK;
} catch (const std::bad_alloc&) {
    return fail(ctx, ATHX_OUT_OF_MEMORY, "host allocation failed");
} catch (const std::exception& e) {
    return fail(ctx, ATHX_INTERNAL_ERROR, e.what());
}

ATHX_API uint32_t athx_sky_texels(uint32_t base_side) { return sky_texels(base_side, nullptr, nullptr); }

ATHX_API athx_status athx_prepare_sky(athx_ctx* ctx, const float* equirect, uint32_t width, uint32_t height,
                                      uint32_t base_side, uint32_t* texels, float* sh)  try {
    if (!ctx) {
        return ATHX_INVALID_ARGUMENT;
    }
    ctx->error.clear();
    if (!equirect || !texels || !sh) {
        return fail(ctx, ATHX_INVALID_ARGUMENT, "null equirect, texels or sh");
    }
    // The rough levels integrate every source texel for every output texel:
    // these bounds keep one launch under about 10^11 lobe evaluations.
    if (width == 0 || height == 0 || (uint64_t)width * height > (1u << 18)) {
        return fail(ctx, ATHX_INVALID_ARGUMENT, "equirect size out of range (at most 2^18 texels)");
    }
    if (base_side == 0 || base_side > 1024) {
        return fail(ctx, ATHX_INVALID_ARGUMENT, "base_side out of range (1 to 1024)");
    }
    ATHX_HIP(ctx, hipSetDevice(ctx->device), "set device");
    ATHX_HIP(ctx, launch_prepare_sky(equirect, width, height, base_side, texels, sh, nullptr), "prepare sky");
    ATHX_HIP(ctx, hipStreamSynchronize(nullptr), "prepare sky wait");
    return ATHX_OK;
} catch (const std::bad_alloc&) {
    return fail(ctx, ATHX_OUT_OF_MEMORY, "host allocation failed");
} catch (const std::exception& e) {
    return fail(ctx, ATHX_INTERNAL_ERROR, e.what());
}

}  // extern "C"

namespace {

// env_sun.slang's constants: a texel is the sun at kSunOverMedian times the
// sky's median and kSunOfPeak of its peak, within kSunWidest of the peak; a sky
// whose peak is not kSunPeakOverMedian times its median has none.
constexpr float kSunOverMedian = 16.0f, kSunOfPeak = 0.02f, kSunPeakOverMedian = 32.0f;
constexpr double kSunWidest = 0.175;
constexpr uint32_t kSunLevelWidth = 256;  // the level searched: at most this wide (env_sun.slang:18-20)

// A level's direction tables, computed once in double: rows (sin theta,
// cos theta), theta = pi (j + 0.5) / h from +y; columns (cos phi, sin phi),
// phi = 2 pi (i + 0.5) / w from +x towards +z; and each row's sin theta in
// 2^-24 units.
void sky_tables(uint32_t w, uint32_t h, std::vector<float>& rows, std::vector<float>& cols,
                std::vector<uint32_t>* weights) {
    const double pi = 3.14159265358979323846;
    rows.resize(2 * (size_t)h);
    cols.resize(2 * (size_t)w);
    if (weights) {
        weights->resize(h);
    }
    for (uint32_t j = 0; j < h; ++j) {
        const double theta = pi * ((double)j + 0.5) / (double)h;
        rows[2 * j] = (float)std::sin(theta);
        rows[2 * j + 1] = (float)std::cos(theta);
        if (weights) {
            (*weights)[j] = (uint32_t)std::llround(std::max(std::sin(theta), 0.0) * 16777216.0);
        }
    }
    for (uint32_t i = 0; i < w; ++i) {
        const double phi = 2.0 * pi * ((double)i + 0.5) / (double)w;
        cols[2 * i] = (float)std::cos(phi);
        cols[2 * i + 1] = (float)std::sin(phi);
    }
}

athx_status check_sky_input(athx_ctx* ctx, const float* equirect, uint32_t width, uint32_t height) {
    if (!equirect) {
        return fail(ctx, ATHX_INVALID_ARGUMENT, "null equirect");
    }
    if (width == 0 || height == 0 || (uint64_t)width * height > (1u << 24)) {
        return fail(ctx, ATHX_INVALID_ARGUMENT, "equirect size out of range (at most 2^24 texels)");
    }
    return ATHX_OK;
}

}  // namespace

extern "C" {

ATHX_API athx_status athx_sky_sun(athx_ctx* ctx, const float* equirect, uint32_t width, uint32_t height,
                                  float* sun)  try {
    if (!ctx) {
        return ATHX_INVALID_ARGUMENT;
    }
    ctx->error.clear();
    if (athx_status st = check_sky_input(ctx, equirect, width, height); st != ATHX_OK) {
        return st;
    }
    if (!sun) {
        return fail(ctx, ATHX_INVALID_ARGUMENT, "null sun");
    }
    ATHX_HIP(ctx, hipSetDevice(ctx->device), "set device");
    hipStream_t s = nullptr;
    // The level searched and summed: halved until at most kSunLevelWidth wide.
    uint32_t w = width, h = height, levels = 0;
    size_t level_floats = 0;
    while (w > kSunLevelWidth) {
        w = std::max(w >> 1, 1u);
        h = std::max(h >> 1, 1u);
        level_floats += 3 * (size_t)w * h;
        ++levels;
    }
    ATHX_HIP(ctx, reserve(ctx, ctx->sky_levels, level_floats * 4 + 16), "sky levels");
    const float* img = equirect;
    {
        uint32_t sw = width, sh = height;
        float* dst = as<float>(ctx->sky_levels);
        for (uint32_t k = 0; k < levels; ++k) {
            const uint32_t dw = std::max(sw >> 1, 1u), dh = std::max(sh >> 1, 1u);
            ATHX_HIP(ctx, launch_sky_half(img, sw, sh, dst, dw, dh, s), "sky level");
            img = dst;
            dst += 3 * (size_t)dw * dh;
            sw = dw;
            sh = dh;
        }
    }
    std::vector<float> rows, cols;
    std::vector<uint32_t> weights;
    sky_tables(w, h, rows, cols, &weights);
    const size_t table_bytes = (rows.size() + cols.size()) * 4 + weights.size() * 4;
    ATHX_HIP(ctx, reserve(ctx, ctx->sky_tables, table_bytes + 16), "sky tables");
    float* d_rows = as<float>(ctx->sky_tables);
    float* d_cols = d_rows + rows.size();
    uint32_t* d_weights = reinterpret_cast<uint32_t*>(d_cols + cols.size());
    ATHX_HIP(ctx, hipMemcpy(d_rows, rows.data(), rows.size() * 4, hipMemcpyHostToDevice), "sky tables");
    ATHX_HIP(ctx, hipMemcpy(d_cols, cols.data(), cols.size() * 4, hipMemcpyHostToDevice), "sky tables");
    ATHX_HIP(ctx, hipMemcpy(d_weights, weights.data(), weights.size() * 4, hipMemcpyHostToDevice), "sky tables");
    // Words: 80 histogram bins, the peak, then the sun's 8 sums (doubles).
    ATHX_HIP(ctx, reserve(ctx, ctx->sky_words, (80 + 1 + 8) * 8 + 16), "sky words");
    unsigned long long* d_hist = as<unsigned long long>(ctx->sky_words);
    unsigned long long* d_peak = d_hist + 80;
    double* d_sums = reinterpret_cast<double*>(d_peak + 1);
    ATHX_HIP(ctx, launch_fill_u32(reinterpret_cast<uint32_t*>(d_hist), 0u, 2 * (80 + 1 + 8), s), "sky words clear");
    ATHX_HIP(ctx, launch_sky_sun_scan(img, w, h, d_weights, d_hist, d_peak, s), "sky scan");
    std::vector<unsigned long long> hist(81);
    ATHX_HIP(ctx, hipMemcpy(hist.data(), d_hist, 81 * 8, hipMemcpyDeviceToHost), "sky scan read");
    // The median: the first half stop whose cumulative weight reaches half.
    unsigned long long total = 0;
    for (uint32_t b = 0; b < 80; ++b) {
        total += hist[b];
    }
    float median = 0.0f;
    unsigned long long below = 0;
    for (uint32_t b = 0; b < 80; ++b) {
        below += hist[b];
        if (2 * below >= total && total > 0) {
            median = (float)std::exp2((double)b * 0.5 - 20.0 + 0.25);
            break;
        }
    }
    const unsigned long long peak_word = hist[80];
    const float peak = peak_word ? __builtin_bit_cast(float, (uint32_t)(peak_word >> 32)) : -1.0f;
    const uint32_t at = peak_word ? 0xFFFFFFFFu - (uint32_t)(peak_word & 0xFFFFFFFFull) : 0u;
    const uint32_t py = at / w, px = at - py * w;
    const float axis[3] = {rows[2 * py] * cols[2 * px], rows[2 * py + 1], rows[2 * py] * cols[2 * px + 1]};
    const float threshold = std::max(kSunOverMedian * median, kSunOfPeak * peak);
    const bool found = peak_word != 0 && peak > kSunPeakOverMedian * std::max(median, 1.0e-9f);
    double sums[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const float cos_widest = (float)std::cos(kSunWidest);
    if (found) {
        const double pi = 3.14159265358979323846;
        const float area = (float)(2.0 * pi * pi / ((double)w * (double)h));
        ATHX_HIP(ctx, launch_sky_sun_sum(img, w, h, d_rows, d_cols, area, threshold, axis, cos_widest, d_sums, s),
                 "sky sums");
        ATHX_HIP(ctx, hipMemcpy(sums, d_sums, 8 * 8, hipMemcpyDeviceToHost), "sky sums read");
    }
    // env_sun.slang:236-261: the disc's cone from its solid angle, and a region
    // spread past twice a disc's mean square angle left in the sky.
    const double pi = 3.14159265358979323846;
    const double solid = sums[7], weight = sums[6];
    const double clen = std::sqrt(sums[3] * sums[3] + sums[4] * sums[4] + sums[5] * sums[5]);
    const double disc = solid > 0.0 ? std::acos(std::min(std::max(1.0 - solid / (2.0 * pi), 0.0), 1.0)) : 0.0;
    const double spread = weight > 0.0 ? 2.0 * std::max(1.0 - clen / weight, 0.0) : 0.0;
    const bool compact = spread <= disc * disc + 1.0e-6;
    const bool is_sun = found && solid > 0.0 && clen > 0.0 && compact;
    for (int c = 0; c < 3; ++c) {
        sun[c] = is_sun ? (float)((c == 0 ? sums[3] : c == 1 ? sums[4] : sums[5]) / clen) : axis[c];
        sun[4 + c] = (float)sums[c];
    }
    sun[3] = is_sun ? std::max((float)disc, 1.0e-4f) : 0.0f;
    sun[7] = threshold;
    return ATHX_OK;
} catch (const std::bad_alloc&) {
    return fail(ctx, ATHX_OUT_OF_MEMORY, "host allocation failed");
} catch (const std::exception& e) {
    return fail(ctx, ATHX_INTERNAL_ERROR, e.what());
}

ATHX_API athx_status athx_sky_residual(athx_ctx* ctx, const float* equirect, uint32_t width, uint32_t height,
                                       const float* sun, float* residual)  try {
    if (!ctx) {
        return ATHX_INVALID_ARGUMENT;
    }
    ctx->error.clear();
    if (athx_status st = check_sky_input(ctx, equirect, width, height); st != ATHX_OK) {
        return st;
    }
    if (!sun || !residual) {
        return fail(ctx, ATHX_INVALID_ARGUMENT, "null sun or residual");
    }
    ATHX_HIP(ctx, hipSetDevice(ctx->device), "set device");
    std::vector<float> rows, cols;
    sky_tables(width, height, rows, cols, nullptr);
    ATHX_HIP(ctx, reserve(ctx, ctx->sky_tables, (rows.size() + cols.size()) * 4 + 16), "sky tables");
    float* d_rows = as<float>(ctx->sky_tables);
    float* d_cols = d_rows + rows.size();
    ATHX_HIP(ctx, hipMemcpy(d_rows, rows.data(), rows.size() * 4, hipMemcpyHostToDevice), "sky tables");
    ATHX_HIP(ctx, hipMemcpy(d_cols, cols.data(), cols.size() * 4, hipMemcpyHostToDevice), "sky tables");
    // No sun (a half angle of 0): nothing is cut.
    const float threshold = sun[3] > 0.0f ? sun[7] : std::numeric_limits<float>::infinity();
    ATHX_HIP(ctx, launch_sky_residual(equirect, width, height, d_rows, d_cols, threshold, sun,
                                      (float)std::cos(kSunWidest), residual, nullptr),
             "sky residual");
    ATHX_HIP(ctx, hipStreamSynchronize(nullptr), "sky residual wait");
    return ATHX_OK;
} catch (const std::bad_alloc&) {
    return fail(ctx, ATHX_OUT_OF_MEMORY, "host allocation failed");
} catch (const std::exception& e) {
    return fail(ctx, ATHX_INTERNAL_ERROR, e.what());
}

}  // extern "C"

namespace {

// athx_reference_render(_instances): the ground truth of one view.
athx_status reference_frame(athx_ctx* ctx, const athx_view* view, const FrameInstance* inst, uint32_t ninst,
                            const athx_settings* settings, const athx_targets* targets, float* records,
                            uint32_t* overflow) {
    if (!view || !inst || ninst == 0 || !targets || !overflow) {
        return fail(ctx, ATHX_INVALID_ARGUMENT, "null view, gaussians, targets or overflow");
    }
    if (athx_status st = check_settings(ctx, settings); st != ATHX_OK) {
        return st;
    }
    uint64_t total = 0;
    for (uint32_t j = 0; j < ninst; ++j) {
        const athx_gaussians* g = inst[j].g;
        const athx_packed_splats* pk = g ? g->packed : nullptr;
        if (!pk || !pk->positions || !pk->shape) {
            return fail(ctx, ATHX_UNSUPPORTED, "the reference renderer reads the packed layout (gaussians->packed)");
        }
        if (pk->rest_per_colour != 0 && pk->rest_per_colour != 3 && pk->rest_per_colour != 8 &&
            pk->rest_per_colour != 15) {
            return fail(ctx, ATHX_INVALID_ARGUMENT, "rest_per_colour must be 0, 3, 8 or 15");
        }
        if (pk->rest_per_colour != 0 && !pk->sh) {
            return fail(ctx, ATHX_INVALID_ARGUMENT, "null sh with rest coefficients");
        }
        total += g->count;
    }
    if (total >= (uint64_t{1} << 31)) {
        return fail(ctx, ATHX_UNSUPPORTED, "more than 2^31 splats for the ground truth");
    }
    *overflow = 0;
    ATHX_HIP(ctx, hipSetDevice(ctx->device), "set device");
    uint64_t tiles = 0;
    if (athx_status st = stage_views(ctx, view, 1, settings, targets, nullptr, tiles, nullptr, inst, ninst);
        st != ATHX_OK) {
        return st;
    }
    const hipStream_t s = nullptr;
    if (athx_status st = upload_views(ctx, 1, s, ninst); st != ATHX_OK) {
        return st;
    }
    const uint32_t n = (uint32_t)total;
    const size_t cap = n ? n : 1;
    ATHX_HIP(ctx, reserve(ctx, ctx->ref_records, cap * kRecordFloats * 4), "reference records");
    ATHX_HIP(ctx, reserve(ctx, ctx->ref_keys, cap * 4 + 16), "reference keys");
    ATHX_HIP(ctx, reserve(ctx, ctx->ref_values, cap * 4 + 16), "reference values");
    ATHX_HIP(ctx, reserve(ctx, ctx->ref_alt_keys, cap * 4 + 16), "reference keys");
    ATHX_HIP(ctx, reserve(ctx, ctx->ref_alt_values, cap * 4 + 16), "reference values");
    ATHX_HIP(ctx, reserve(ctx, ctx->ref_sort_meta, sort_meta_words((uint32_t)cap, 4) * 4 + 16), "reference sort");
    ATHX_HIP(ctx, reserve(ctx, ctx->ref_count, 64), "reference counts");
    float* rec = records ? records : as<float>(ctx->ref_records);
    uint64_t* count = as<uint64_t>(ctx->ref_count);
    // The sort's count (a u64: n, 0) and the overflow counter, written on the stream.
    ATHX_HIP(ctx, launch_fill_u32(reinterpret_cast<uint32_t*>(count), n, 1, s), "count");
    ATHX_HIP(ctx, launch_fill_u32(reinterpret_cast<uint32_t*>(count) + 1, 0u, 3, s), "count and overflow");
    uint32_t base = 0;
    for (uint32_t j = 0; j < ninst; ++j) {
        const athx_packed_splats* pk = inst[j].g->packed;
        ATHX_HIP(ctx, launch_ref_project(pk->positions, pk->shape, pk->sh, inst[j].g->count, pk->rest_per_colour,
                                         settings->max_sh_degree, settings->antialias, settings->linear,
                                         as<DView>(ctx->views) + j, rec, as<uint32_t>(ctx->ref_keys),
                                         as<uint32_t>(ctx->ref_values), s, base),
                 "reference projection");
        base += inst[j].g->count;
    }
    // The drawn splats in (depth, index) order: a stable sort of the
    // depth's 32 bits (culled keys, all ones, last).
    const SortScratch sc = sort_scratch(ctx->ref_sort_meta, ctx->ref_alt_keys.ptr, as<uint32_t>(ctx->ref_alt_values),
                                        (uint32_t)cap, 4);
    ATHX_HIP(ctx, launch_sort(as<uint32_t>(ctx->ref_keys), as<uint32_t>(ctx->ref_values), count, (uint32_t)cap, 0, 32,
                              sc, s, nullptr),
             "reference sort");
    uint32_t* over = reinterpret_cast<uint32_t*>(count + 1);
    ATHX_HIP(ctx, launch_ref_blend(as<DView>(ctx->views), view->width, view->height, rec, as<uint32_t>(ctx->ref_keys),
                                   as<uint32_t>(ctx->ref_values), n, 1024, settings->depth_mode,
                                   settings->depth_threshold, over, s),
             "reference blend");
    uint32_t host_over = 0;
    ATHX_HIP(ctx, hipMemcpy(&host_over, over, sizeof host_over, hipMemcpyDeviceToHost), "read overflow");
    *overflow = host_over;
    return ATHX_OK;
}

}  // namespace

extern "C" {

ATHX_API athx_status athx_reference_render(athx_ctx* ctx, const athx_view* view, const athx_gaussians* g,
                                           const athx_settings* settings, const athx_targets* targets,
                                           float* records, uint32_t* overflow)  try {
    if (!ctx) {
        return ATHX_INVALID_ARGUMENT;
    }
    ctx->error.clear();
    FrameInstance one;
    one.g = g;
    return reference_frame(ctx, view, g ? &one : nullptr, 1, settings, targets, records, overflow);
} catch (const std::bad_alloc&) {
    return fail(ctx, ATHX_OUT_OF_MEMORY, "host allocation failed");
} catch (const std::exception& e) {
    return fail(ctx, ATHX_INTERNAL_ERROR, e.what());
}

ATHX_API athx_status athx_reference_render_instances(athx_ctx* ctx, const athx_view* view,
                                                     const athx_instance* instances, uint32_t ninstances,
                                                     const athx_settings* settings, const athx_targets* targets,
                                                     float* records, uint32_t* overflow)  try {
    if (!ctx) {
        return ATHX_INVALID_ARGUMENT;
    }
    ctx->error.clear();
    if (!instances || ninstances == 0) {
        return fail(ctx, ATHX_INVALID_ARGUMENT, "no instances");
    }
    std::vector<FrameInstance> inst(ninstances);
    for (uint32_t j = 0; j < ninstances; ++j) {
        inst[j].g = instances[j].gaussians;
        inst[j].object_to_view = instances[j].object_to_view;
        inst[j].eye = instances[j].eye;
    }
    return reference_frame(ctx, view, inst.data(), ninstances, settings, targets, records, overflow);
} catch (const std::bad_alloc&) {
    return fail(ctx, ATHX_OUT_OF_MEMORY, "host allocation failed");
} catch (const std::exception& e) {
    return fail(ctx, ATHX_INTERNAL_ERROR, e.what());
}

}  // extern "C"
