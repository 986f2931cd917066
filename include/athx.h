/*
 * athx.h -- the C ABI of the MI355X Gaussian-splat tile rasterizer.
 *
 * Plain C: device pointers, sizes and a hipStream_t passed as void*. No
 * exceptions cross it; every call returns an athx_status: 0 for success, else
 * athenea's ErrorCode (modules/core/include/athenea/core/Error.h:14-26) plus
 * one, in the enum's order (athx_error_code() gives the ErrorCode value), and
 * athx_last_error() returns the message of the last failure on a context.
 *
 * What each entry point replaces in the reference:
 *
 *   athx_create / athx_destroy
 *       render::TileRasterizer::create (modules/render/include/athenea/render/
 *       TileRasterizer.h:454) and its destructor.
 *   athx_release_scratch
 *       TileRasterizer::releaseScratch (TileRasterizer.h:487;
 *       modules/render/src/TileRasterizer.cpp:178-191).
 *   athx_render
 *       TileRasterizer::render(const Projection&, span<const SplatInstance>,
 *       const RenderSettings&, RenderTargets&) (TileRasterizer.h:465-477;
 *       TileRasterizer.cpp:296-1295) for one instance at the identity:
 *       plain or relit clouds, an under layer, Cryptomatte and trace targets
 *       (no point sprites), generalised to several views in one call:
 *       views[k] is a camera's Projection and writes targets[k].
 *   athx_render_instances
 *       the same render over a span of SplatInstance (TileRasterizer.h:
 *       44-60, 465-477): several placed clouds, or one cloud placed several
 *       times, in one frame (TileRasterizer.cpp:896-898's slot numbering).
 *   athx_render_async / athx_frame_pairs
 *       the render's asynchronous pair readback (TileRasterizer.cpp:
 *       1240-1290): no host wait, the pair count read a frame later.
 *   athx_graph_create / athx_graph_launch / athx_graph_launch_backward /
 *   athx_graph_pairs / athx_graph_destroy
 *       no counterpart (the reference records a command buffer per frame):
 *       one frame's launches captured in a HIP graph and replayed.
 *   athx_frame_serial / athx_frame_info_get / athx_copy_frame
 *       the frame's buffers as RenderTargets' debug views (TileRasterizer.h:
 *       403-452): pairs, ranges and records of the last frame, for tests.
 *   athx_render_backward
 *       no counterpart in the reference (its only gradient is the blend's
 *       contributionMode 4, shaders/athenea/splat/splat_blend.slang:381-398):
 *       the adjoint of athx_render's colour with respect to every parameter.
 *   athx_contribution
 *       RenderSettings::contributionMode 1 to 4 (splat_blend.slang:364-408).
 *   athx_latest_counters
 *       TileRasterizer::latestCounters (TileRasterizer.h:481) and
 *       SplatCounters (TileRasterizer.h:370-401).
 *   athx_exclusive_scan
 *       gpu::PrefixSum::apply (modules/gpu/include/athenea/gpu/algo/
 *       PrefixSum.h:43-44).
 *   athx_sort_pairs
 *       gpu::RadixSort::sort (modules/gpu/include/athenea/gpu/algo/
 *       RadixSort.h:63-64) for 64-bit keys: the reference's keysLo/keysHi
 *       halves are one uint64 here.
 *   athx_ply_parse / athx_splat_layout / athx_ingest_ply
 *       io::parsePlyHeader + io::readSplatPly (modules/io/src/PlyHeader.cpp,
 *       SplatReaders.cpp:48-195), the .splat record reader (SplatReaders.cpp:
 *       203-241) and scene::CloudLoader::upload for either (modules/scene/
 *       include/athenea/scene/GpuClouds.h:349-516).
 *   athx_reference_render / athx_reference_render_instances /
 *   athx_compare_images / athx_compare_hdr
 *       render::ReferenceRenderer::render, compareImages, compareHdr
 *       (modules/render/include/athenea/render/ReferenceRenderer.h:20-66).
 *   athx_sky_texels / athx_prepare_sky
 *       technique::Environment (modules/technique/include/athenea/technique/
 *       Environment.h:1-155) for the relit path: the sky's prefiltered
 *       octahedral levels and SH as the blend reads them.
 *   athx_sky_sun / athx_sky_residual
 *       Environment::sun (Environment.h:72-84; env_sun.slang) and
 *       env_residual.slang: the sun found in a sky, and the sky without it.
 *
 * Threading: a context is used by one host thread at a time. Every call is
 * asynchronous on the stream it is given, except where a comment says it
 * waits. Callers own Gaussian and target memory; a context owns its scratch.
 */
#ifndef ATHX_H
#define ATHX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#if defined(_WIN32)
#define ATHX_API __declspec(dllexport)
#else
#define ATHX_API __attribute__((visibility("default")))
#endif

/* 0 for success, else athenea's ErrorCode + 1 (Error.h:14-26 numbers
 * Unknown 0, NotFound 1, ... InternalError 10; a Result<T> holds success
 * apart from the code, which a C status cannot). athx_error_code() maps a
 * status back to the ErrorCode value. */
typedef enum athx_status {
    ATHX_OK = 0,
    ATHX_UNKNOWN = 1,
    ATHX_NOT_FOUND = 2,
    ATHX_INVALID_ARGUMENT = 3,
    ATHX_UNSUPPORTED = 4,
    ATHX_IO_FAILURE = 5,
    ATHX_DEVICE_FAILURE = 6,
    ATHX_SHADER_FAILURE = 7,
    ATHX_PLUGIN_FAILURE = 8,
    ATHX_OUT_OF_MEMORY = 9,
    ATHX_CANCELLED = 10,
    ATHX_INTERNAL_ERROR = 11
} athx_status;

/* Why a splat was culled, in frame.slang's numbering
 * (shaders/athenea/splat/frame.slang:42-49). */
enum {
    ATHX_CULL_UNPROJECTED = 0,
    ATHX_CULL_EDIT = 1,
    ATHX_CULL_DEPTH = 2,
    ATHX_CULL_DEGENERATE = 3,
    ATHX_CULL_FAINT = 4,
    ATHX_CULL_OFFSCREEN = 5,
    ATHX_CULL_NO_TILE = 6,
    ATHX_CULL_REASONS = 7
};

/* The athenea ErrorCode value of a failed status (ATHX_UNKNOWN -> 0 =
 * ErrorCode::Unknown, ... ATHX_INTERNAL_ERROR -> 10), or -1 for ATHX_OK and
 * for a value outside the enum. */
ATHX_API int athx_error_code(athx_status status);

typedef struct athx_ctx athx_ctx;

/*
 * One camera for one image size: render::Projection (modules/render/include/
 * athenea/render/Camera.h:50-89) already cast to the floats the kernels read,
 * as setFrame/setObject do (modules/render/src/FrameParams.h:32-95). The view
 * looks down +z with x right and y up; the screen origin is the bottom-left
 * corner and pixel centres sit at +0.5 (frame.slang:22-30).
 */
typedef struct athx_view {
    uint32_t width, height;   /* 1 .. 65535 pixels each */
    float focal_x, focal_y;   /* pixels per unit of x/z (or per unit, orthographic) */
    float centre_x, centre_y; /* the principal point, pixels */
    float near_z, far_z;
    uint32_t orthographic;
    float object_to_view[12]; /* rows of a 3x4 */
    float eye[3];             /* the camera in the cloud's own space */
    float background[4];      /* premultiplied linear */
} athx_view;

/*
 * The Gaussians, structure of arrays, FP32, device memory. This replaces the
 * packed scene::GpuSplats (modules/scene/include/athenea/scene/GpuClouds.h:
 * 44-218) with what its decode produces: log-scales rather than scales,
 * opacity logits rather than opacities, the quaternion as stored (x, y, z, w)
 * and normalised by the projection, and spherical-harmonic coefficients in
 * 3DGS order (coefficient k of channel c at sh[(i * sh_coeffs + k) * 3 + c],
 * k = 0 the constant term, colour = 0.5 + 0.2820948 * sh0 + rest).
 * Exactly one of sh and sh_half is non-null; sh_half holds IEEE binary16.
 */
/*
 * One prepared dome (SplatLights' envTexels / envSh / envSun with envBaseSide,
 * TileRasterizer.h:178-246; layout light/environment.slang:50-110), device
 * memory:
 *   texels  8 octahedral levels of the sky convolved with GGX lobes of
 *           roughness (k/7)^2, level k of side max(max(base_side, 256) >> k,
 *           16), levels back to back, texels row by row; 2 words a texel:
 *           word 0 = red | green << 16, word 1 = blue, IEEE binary16
 *   sh      16 float4: the sky's radiance in real spherical harmonics (the
 *           order of light/environment.slang:202-224), rgb; the first 9 read
 *   sun     2 float4: (direction, disc size), w > 0 when there is a sun;
 *           (irradiance rgb, unused)
 */
typedef struct athx_sky {
    const uint32_t* texels;
    const float* sh;
    const float* sun;
    uint32_t base_side;
} athx_sky;

/*
 * Relighting (SplatInstance::relight, TileRasterizer.h:85-172, with a
 * SplatLights of one prepared dome): the colour becomes relitSplat's dome
 * term for a surface with the plain lobes (splat_relight.slang:1379-1692,
 * 1775-1824) -- albedo (1 - metallic) (irradiance / pi + sun) + the GGX
 * environment BRDF times the prefiltered sky along the mirror direction.
 * Per-Gaussian arrays are FP32 device memory; the struct is read at call
 * time. The eye in the world is object_to_world applied to the view's eye.
 */
typedef struct athx_relight {
    const float* normals;    /* [count][3] stored shading normals (cloud space), or NULL: the disc's short axis */
    const float* albedo;     /* [count][3], or NULL: the evaluated SH colour is the albedo (splat_project.slang:819) */
    const float* metallic;   /* [count] */
    const float* roughness;  /* [count] */
    const uint32_t* normals_packed; /* [count] GpuSplats::normals (octahedral, 2 x 16-bit unorms,
                                       packing.slang:116-135), or NULL: normals is read */
    const uint32_t* pbr;     /* [count] GpuSplats' pbr words (metallic in byte 0, roughness in byte 1,
                                scene/splat_encoding.slang:189-214), or NULL: metallic and roughness */
    float object_to_world[12]; /* rows of a 3x4 */
    athx_sky sky;
} athx_relight;

/*
 * athenea's packed device layout of a cloud (scene::GpuSplats,
 * modules/scene/include/athenea/scene/GpuClouds.h:44-52; the field layout of
 * shaders/athenea/common/packing.slang:1-26), device memory, 124 bytes a
 * Gaussian at degree 3 where the FP32 arrays take 236:
 *   positions  float4 [count]: x, y, z, opacity (linear, 0..1)
 *   shape      uint32 [count][4]:
 *              [0] rotation, smallest three: components in bits 22-31, 12-21
 *                  and 2-11 (unorms of [-sqrt(1/2), sqrt(1/2)]), the index
 *                  of the dropped largest one in bits 0-1
 *              [1] f16 log scale x | f16 log scale y << 16
 *              [2] f16 log scale z | f16 colour red << 16
 *              [3] f16 colour green | f16 colour blue << 16
 *   sh         uint32 [count][(rest_per_colour * 3 + 1) / 2]: the rest
 *              coefficients as IEEE halves, low half first, rgb per basis
 *              function (common/sh.slang:15-22); NULL at rest_per_colour 0
 * The colour is linear or sRGB as athx_settings.linear says. The quaternion
 * is used as decoded (unit to 10 bits), not normalised again.
 */
typedef struct athx_packed_splats {
    const float* positions;
    const uint32_t* shape;
    const uint32_t* sh;
    uint32_t rest_per_colour; /* 0, 3, 8 or 15 */
} athx_packed_splats;

typedef struct athx_gaussians {
    uint32_t count;
    uint32_t sh_coeffs; /* 1, 4, 9 or 16 */
    const float* means;          /* [count][3] */
    const float* log_scales;     /* [count][3] */
    const float* quats;          /* [count][4], (x, y, z, w) */
    const float* opacity_logits; /* [count] */
    const float* sh;             /* [count][sh_coeffs][3] or NULL */
    const uint16_t* sh_half;     /* [count][sh_coeffs][3] or NULL */
    const athx_relight* relight; /* NULL: the SH colour; else the relit colour (forward only) */
    const athx_packed_splats* packed; /* non-NULL: the packed layout is read instead of the arrays
                                         above, which may then be NULL (forward only) */
    const uint32_t* crypto_ids;  /* [count] each Gaussian's Cryptomatte id (GpuSplats::crypto), or NULL:
                                    a cloud without ancestry, whose splats name nothing (id 0) */
} athx_gaussians;

/*
 * RenderSettings (TileRasterizer.h:248-371), the fields a plain frame reads,
 * plus the two launch shapes this rasterizer tunes.
 */
typedef struct athx_settings {
    uint32_t tile;         /* tile side in pixels: 8, 16 or 32 (the reference's is 16) */
    uint32_t batch;        /* Gaussians staged in LDS per blend batch: 32..256, a multiple of 32 */
    uint32_t antialias;    /* 1: the Mip-Splatting compensation (RenderSettings::antialias) */
    uint32_t linear;       /* 1: colours are linear light; 0: sRGB, decoded per splat */
    uint32_t max_sh_degree;
    uint32_t depth_mode;   /* 0 alpha-weighted mean, 1 first crossing of depth_threshold */
    float depth_threshold;
    uint32_t time_stages;  /* 1: wait after each stage and fill athx_frame_stats' times */
    uint32_t count_splats; /* 1: count cull reasons for athx_latest_counters */
    uint64_t counters_tag; /* RenderSettings::countersTag: returned with this frame's counters */
} athx_settings;

/*
 * What a view writes: RenderTargets (TileRasterizer.h:421-450), caller-owned
 * device memory, bottom row first. colour is float4 premultiplied linear,
 * depth float view z. final_t (the transmittance left after the walk) and
 * last (one past the last pair the walk read, relative to the tile's first)
 * are what athx_render_backward replays; either may be NULL when no backward
 * pass follows.
 */
/*
 * The walk's variants (splat_blend.slang:236-578), each on when its pointer
 * is given, for every view of a frame alike:
 *   under_colour / under_depth   an opaque layer under the splats (the
 *       reference's `under` RenderTargets; linear premultiplied float4 and
 *       view z, 0 where it has nothing): a splat at or behind it ends the
 *       pixel's walk, and the layer takes the light that is left
 *       (TileRasterizer.h:465-477, splat_blend.slang:287-300, 512-566)
 *   crypto   the Cryptomatte (RenderTargets::crypto, TileRasterizer.h:427-
 *       431): the six heaviest ids of a pixel (the Gaussians' crypto_ids, the
 *       layer's under_crypto) as three planes of float4 (id bits, coverage,
 *       id bits, coverage), descending
 *   trace    RenderTargets::trace (splat_trace.slang): ATHX_TRACE_WORDS planes
 *       of words a pixel -- the four heaviest slots, their view depths, their
 *       weights as halves, (rest, under) as halves; a pixel nothing reached
 *       holds ATHX_TRACE_EMPTY in its first plane alone
 */
#define ATHX_TRACE_SLOTS 4
#define ATHX_TRACE_WORDS (ATHX_TRACE_SLOTS * 2 + ATHX_TRACE_SLOTS / 2 + 1)
#define ATHX_TRACE_NONE 0xFFFFFFFFu
#define ATHX_TRACE_EMPTY 0xFFFFFFFEu
typedef struct athx_targets {
    float* colour;    /* [height][width][4] */
    float* depth;     /* [height][width] or NULL */
    float* final_t;   /* [height][width] or NULL */
    uint32_t* last;   /* [height][width] or NULL */
    float* crypto;    /* [3][height][width][4] or NULL */
    uint32_t* trace;  /* [ATHX_TRACE_WORDS][height][width] or NULL */
    const float* under_colour;     /* [height][width][4] or NULL */
    const float* under_depth;      /* [height][width], with under_colour */
    const uint32_t* under_crypto;  /* [height][width] or NULL: the layer's ids, for the matte */
} athx_targets;

/* FrameStats (TileRasterizer.h:408-419), summed over the views. The stage
 * times (filled with time_stages) are HIP-event times on the frame's stream:
 * projection (with the no-tile cull settled); compaction + depth sort; the
 * strip entries scanned in depth order; their emission, sorted on the strip
 * (project_ms .. emit_ms); the binning of the entries into tile-sorted pairs
 * in place of the reference's sort of the pair keys (tile_sort_ms); tile
 * ranges; blend. */
typedef struct athx_frame_stats {
    uint32_t splats;
    uint32_t visible;
    uint64_t pairs;
    double project_ms, depth_sort_ms, counts_ms, emit_ms, tile_sort_ms, ranges_ms, blend_ms, total_ms;
    uint64_t frame_serial; /* this frame's number on the context (athx_frame_serial) */
} athx_frame_stats;

/* SplatCounters (TileRasterizer.h:370-401) for the whole frame, summed over
 * its views: slots, the drawn (visible) ones, their (tile, Gaussian) pairs,
 * the most tiles one footprint touched, the culled ones by reason, and for
 * each of the first ATHX_MAX_COUNTED_CLOUDS instances (in the order given)
 * its drawn footprints and their pairs (SplatCounters::clouds). */
#define ATHX_MAX_COUNTED_CLOUDS 64
typedef struct athx_counters {
    uint64_t tag;      /* the frame's athx_settings.counters_tag */
    uint32_t slots;
    uint32_t visible;
    uint64_t pairs;
    uint32_t max_tiles;
    uint32_t culled[ATHX_CULL_REASONS];
    uint32_t clouds;   /* rows below: min(instances, ATHX_MAX_COUNTED_CLOUDS) */
    uint32_t cloud_visible[ATHX_MAX_COUNTED_CLOUDS];
    uint32_t cloud_pairs[ATHX_MAX_COUNTED_CLOUDS];
} athx_counters;

/*
 * Gradients of a scalar loss with respect to every Gaussian parameter, in the
 * layouts of athx_gaussians (caller-owned device memory, FP32). They are
 * ACCUMULATED into (+=), so the caller zeroes them first -- or, with
 * overwrite = 1, WRITTEN (=), every element, so the caller need not clear
 * them. sh is [count][sh_coeffs][3]. Any pointer may be NULL to skip that
 * parameter.
 */
typedef struct athx_grads {
    float* means;
    float* log_scales;
    float* quats;
    float* opacity_logits;
    float* sh;
    uint32_t overwrite;
} athx_grads;

ATHX_API const char* athx_version(void);

/* Opens a context on HIP device `device`. */
ATHX_API athx_status athx_create(int device, athx_ctx** out);
ATHX_API void athx_destroy(athx_ctx* ctx);
/* The message of the last failed call on ctx ("" if none). */
ATHX_API const char* athx_last_error(const athx_ctx* ctx);
/* Frees the scratch a frame grew into; later frames grow it to what they need exactly. */
ATHX_API athx_status athx_release_scratch(athx_ctx* ctx);
/* Bytes of device scratch the context holds now. */
ATHX_API uint64_t athx_scratch_bytes(const athx_ctx* ctx);

/*
 * Renders `nviews` views of one cloud into targets[0..nviews). Waits for the
 * stream once, after the projection and scan, to size the pair buffers (the
 * reference's one readback, TileRasterizer.cpp:1093-1102); everything after is
 * queued without waiting unless settings->time_stages. stats may be NULL.
 * The frame's per-slot records stay in the context for athx_render_backward.
 */
ATHX_API athx_status athx_render(athx_ctx* ctx, const athx_view* views, uint32_t nviews,
                                 const athx_gaussians* gaussians, const athx_settings* settings,
                                 const athx_targets* targets, void* stream, athx_frame_stats* stats);

/*
 * One cloud of a multi-instance frame (render::SplatInstance, TileRasterizer.h:
 * 85-172: a cloud and its objectToWorld). object_to_view and eye are host
 * memory, one entry a view: Projection::worldToView * objectToWorld, and the
 * view's eye taken into the cloud's own space by the inverse of objectToWorld,
 * computed in double and cast to float, as TileRasterizer.cpp:598-600 does for
 * each instance's dispatch. NULL takes the view's own object_to_view / eye
 * (an instance standing where the view's matrix already puts it). A relit
 * instance's relight->object_to_world is its objectToWorld.
 */
typedef struct athx_instance {
    const athx_gaussians* gaussians;
    const float* object_to_view; /* [nviews][12] rows of a 3x4, or NULL */
    const float* eye;            /* [nviews][3], or NULL */
} athx_instance;

/*
 * TileRasterizer::render over a span of SplatInstances (TileRasterizer.h:
 * 465-477): every instance's Gaussians are projected with its own transform
 * and the frame sorts and blends them together. Each view's slots are the
 * instances' Gaussians back to back in the order given: Gaussian i of
 * instance j in view k is slot k * S + slot_bases[j] + i, S the instances'
 * total, slot_bases[j] the counts of the instances before j
 * (FrameStats::slotBases, TileRasterizer.h:408-419). slot_bases (host, may be
 * NULL) receives them. Waits as athx_render does; stats->splats is S. The
 * frame's records serve athx_contribution (gaussians NULL) and the frame
 * readers; athx_render_backward needs a one-instance frame.
 */
ATHX_API athx_status athx_render_instances(athx_ctx* ctx, const athx_view* views, uint32_t nviews,
                                           const athx_instance* instances, uint32_t ninstances,
                                           const athx_settings* settings, const athx_targets* targets, void* stream,
                                           athx_frame_stats* stats, uint32_t* slot_bases);

/*
 * athx_render without its wait: the pair buffers are sized for
 * pair_capacity pairs (0: 1.25x the largest frame this context has seen),
 * the binning's strip entries for as many as the largest frame seen needed
 * (else pair_capacity), and every count is read on the device, so nothing
 * waits for the host. A frame with more pairs (or entries) than that room
 * draws incompletely (the reference's emit drops pairs past its capacity the
 * same way, shaders/athenea/splat/splat_emit.slang:43-46): athx_frame_pairs
 * says so.
 */
ATHX_API athx_status athx_render_async(athx_ctx* ctx, const athx_view* views, uint32_t nviews,
                                       const athx_gaussians* gaussians, const athx_settings* settings,
                                       const athx_targets* targets, uint64_t pair_capacity, void* stream);

/* The last frame's pair total and its pair capacity; waits for an async
 * frame. pairs > capacity: that frame was incomplete, draw it again with
 * a capacity of at least pairs (a frame whose strip entries outgrew their
 * room also reports more pairs than its capacity; the next frame is given
 * room for them). */
ATHX_API athx_status athx_frame_pairs(athx_ctx* ctx, uint64_t* pairs, uint64_t* capacity);

/*
 * A frame captured once into a hipGraph and replayed: the batched, wait-free
 * route. athx_graph_create renders nothing; it sizes a context of its own for
 * these views, Gaussians, targets and pair capacity, uploads the views, and
 * captures the async frame's launches (and, when dl_dcolour and grads are
 * given, the adjoint's: dL/dcolour read from dl_dcolour[k], gradients added
 * into grads, which the caller zeroes). Every replay reads the buffers the
 * pointers name at that time: change the Gaussians in place, replay, and the
 * frame is of the new values. Replays on one graph must not overlap.
 */
typedef struct athx_graph athx_graph;
ATHX_API athx_status athx_graph_create(int device, const athx_view* views, uint32_t nviews,
                                       const athx_gaussians* gaussians, const athx_settings* settings,
                                       const athx_targets* targets, uint64_t pair_capacity,
                                       const float* const* dl_dcolour, const athx_grads* grads, athx_graph** out,
                                       char* error, uint32_t error_bytes);
/* Replays the forward frame on stream. */
ATHX_API athx_status athx_graph_launch(athx_graph* graph, void* stream);
/* Replays the adjoint (the forward must have been replayed before it). */
ATHX_API athx_status athx_graph_launch_backward(athx_graph* graph, void* stream);
/* The last replay's pair total (waits for it) and the graph's capacity. */
ATHX_API athx_status athx_graph_pairs(athx_graph* graph, uint64_t* pairs, uint64_t* capacity);
ATHX_API void athx_graph_destroy(athx_graph* graph);

/*
 * The number of the last frame rendered on ctx (athx_render, athx_render_async;
 * 1 for the first), 0 before any or after athx_release_scratch. A context
 * holds one frame's pairs and records: athx_render_backward and
 * athx_contribution take the number of the frame they follow and refuse any
 * other (INVALID_ARGUMENT), so a frame replaced by a later render is never
 * walked with the later frame's pairs.
 */
ATHX_API uint64_t athx_frame_serial(const athx_ctx* ctx);

/*
 * The adjoint of frame frame_serial, the last athx_render on ctx (same views,
 * Gaussians and settings; targets[k].final_t and .last must have been written
 * by it): dL/dcolour is [height][width][4] per view (premultiplied RGBA,
 * FP32), and gradients are accumulated into grads. Queued, not waited for.
 */
ATHX_API athx_status athx_render_backward(athx_ctx* ctx, const athx_view* views, uint32_t nviews,
                                          const athx_gaussians* gaussians, const athx_settings* settings,
                                          const athx_targets* targets, uint64_t frame_serial,
                                          const float* const* dl_dcolour, const athx_grads* grads, void* stream);

/*
 * What each splat gave the last frame (splat_blend.slang:36-48, 364-398:
 * RenderSettings::contributionMode), taken by walking that frame's pairs
 * again with the same arithmetic and early stop. out (device) is cleared
 * and then holds, at each slot (view * count + gaussian):
 *   mode 1: one word, the sum over its pixels of round(T alpha * 4096);
 *   mode 4: four words (NeuralFit's colour gradient, NeuralFit.cpp:226-228):
 *           for each of r, g, b the sum of floor(T alpha dL/dC * 256 + h),
 *           h = hash01((pixel * 3 + channel) ^ (slot * 2654435761)) of
 *           color.slang:43-47 (signed, two's complement), then the sum of
 *           round(T alpha * 256). dl_dcolour[k]: view k's [height][width][4];
 *   mode 2: two words, low then high, of the 64-bit sum of round(min(|e|^2,
 *           1) * 2^24), e = T alpha (c - (C - P) / max(T', 1e-4)) the change
 *           removing the splat makes (GaussianPOP's score): C the pixel drawn
 *           whole, dl_dcolour[k] [height][width][4]; P the walk once the
 *           splat is in; T' the transmittance after it; |e|^2 summed r, g, b
 *           in that order;
 *   mode 3: two words (splat_lift.slang): the sum of round(T alpha m * 1024)
 *           with m the pixel's mask clamped to [0, 1], then the sum of
 *           round(T alpha * 1024); dl_dcolour[k] the mask, [height][width].
 * The views, gaussians, settings and targets are the last render's, frame
 * frame_serial (athx_frame_serial); gaussians may be NULL (a multi-instance
 * frame's slots are all its instances'), else its count must be the frame's
 * Gaussians a view. Queued.
 */
ATHX_API athx_status athx_contribution(athx_ctx* ctx, const athx_view* views, uint32_t nviews,
                                       const athx_gaussians* gaussians, const athx_settings* settings,
                                       const athx_targets* targets, uint64_t frame_serial, uint32_t mode,
                                       const float* const* dl_dcolour, uint32_t* out, void* stream);

/* TileRasterizer::latestCounters (TileRasterizer.h:479-481): the counters of
 * the newest frame rendered with count_splats whose counts have reached the
 * host -- a ring of three readbacks, each copied on the frame's stream after
 * its counts; never waits. NOT_FOUND when no counted frame has finished
 * (athx_render waits for its frame, so after it returns its counts are the
 * newest). A frame without count_splats leaves the last counts as they were. */
ATHX_API athx_status athx_latest_counters(athx_ctx* ctx, athx_counters* out);

/*
 * The per-splat and per-pair buffers the last athx_render left in the
 * context: the reference's proj / tilesTouched / depthKeys / pairSplats /
 * ranges scratch (TileRasterizer.h:560-620; frame.slang:269-281), which its
 * other readers (trace, counters, NeuralFit) walk. Slot s of view k is
 * k * count + i. Layouts:
 *   RECORDS      float[slots][12]: x, y, view z, cut power, conic a, b, c,
 *                alpha, r, g, b, and the projection's branch bits as a uint
 *                (frustum clamp x lo/hi, y lo/hi = 1/2/4/8, opacity pay-back
 *                16, colour channel > 0 = 32 << c, sRGB power branch = 256 << c),
 *                written for visible slots only
 *   TOUCHED      uint32[slots]: tiles the footprint touches, 0 when culled
 *                (the frame bins its pairs without counting them a slot: the
 *                counts are taken on the stream when this part is asked for)
 *   DEPTH_KEYS   uint32[slots]: asuint(z) >> 7, or 0xFF000000 | cull reason
 *   OFFSETS      uint32[slots]: a visible slot's first pair -- the exclusive
 *                scan of TOUCHED taken in depth order (visible slots stably
 *                sorted by depth key); 0 for culled slots; taken on request
 *   PAIR_KEYS    uint64[pairs]: global tile << 24 | depth key, sorted (the
 *                frame keeps each tile's range and each pair's slot; the key
 *                is rebuilt from them and the slot's depth key, on the stream)
 *   PAIR_VALUES  uint32[pairs]: the slot of each sorted pair
 *   RANGES       uint32[tiles][2]: a tile's [start, end) in the sorted pairs
 */
typedef enum athx_frame_part {
    ATHX_PART_RECORDS = 0,
    ATHX_PART_TOUCHED = 1,
    ATHX_PART_DEPTH_KEYS = 2,
    ATHX_PART_OFFSETS = 3,
    ATHX_PART_PAIR_KEYS = 4,
    ATHX_PART_PAIR_VALUES = 5,
    ATHX_PART_RANGES = 6
} athx_frame_part;

typedef struct athx_frame_info {
    uint32_t slots;
    uint32_t tiles;
    uint32_t key_bits;
    uint64_t pairs;
    uint32_t strips;   /* 16-tile strips of tile rows the pairs are binned through */
    uint64_t entries;  /* (footprint, strip) entries of the binning */
} athx_frame_info;

/* The last frame's sizes; pairs is what its buffers hold (waits for an async frame). */
ATHX_API athx_status athx_frame_info_get(athx_ctx* ctx, athx_frame_info* out);
/* Copies `bytes` of one part (which must be its whole size) to device memory dst. Queued. */
ATHX_API athx_status athx_copy_frame(athx_ctx* ctx, athx_frame_part part, void* dst, uint64_t bytes, void* stream);

/*
 * out[i] = in[0] + ... + in[i-1] over n uint32, and *total (device, one
 * uint64) = the sum of all; single pass, decoupled look-back. in and out may
 * not alias. Queued.
 */
ATHX_API athx_status athx_exclusive_scan(athx_ctx* ctx, const uint32_t* in, uint32_t* out, uint64_t* total,
                                         uint32_t n, void* stream);

/*
 * athx_exclusive_scan that allocates nothing. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). Its look-back scratch is a static device array of the library
 * (32 816 bytes on each GPU it runs on, held until the process ends), so n is at
 * most ATHX_NOALLOC_SCAN_MAX (INVALID_ARGUMENT beyond). Calls that overlap
 * on one GPU (two streams, or two contexts of one device) share that array
 * and race: queue them on one stream. Same results as athx_exclusive_scan.
 */
#define ATHX_NOALLOC_SCAN_MAX (1u << 24)
ATHX_API athx_status athx_exclusive_scan_noalloc(athx_ctx* ctx, const uint32_t* in, uint32_t* out, uint64_t* total,
                                                 uint32_t n, void* stream);

/*
 * Stable LSD radix sort of n (key, value) pairs by the low key_bits bits of
 * the key (1..64), onesweep: one histogram pass, then one pass per 8-bit
 * digit. The result is back in keys/values. Queued.
 */
ATHX_API athx_status athx_sort_pairs(athx_ctx* ctx, uint64_t* keys, uint32_t* values, uint32_t n,
                                     uint32_t key_bits, void* stream);

/*
 * Ingest: a trained splat cloud's PLY into the packed layout above.
 *
 * A binary_little_endian PLY's vertex element as its header lays it out
 * (io::parsePlyHeader, modules/io/src/PlyHeader.cpp:59-128, and the checks of
 * io::readSplatPly, modules/io/src/SplatReaders.cpp:48-195): the vertex
 * block's place in the file, a vertex's bytes, and each property's byte
 * offset and type. The fields the decode reads are property indices.
 */
#define ATHX_PLY_MAX_PROPERTIES 256
typedef enum athx_ply_type {
    ATHX_PLY_F32 = 0,
    ATHX_PLY_F64 = 1,
    ATHX_PLY_U8 = 2,
    ATHX_PLY_I8 = 3,
    ATHX_PLY_U16 = 4,
    ATHX_PLY_I16 = 5,
    ATHX_PLY_U32 = 6,
    ATHX_PLY_I32 = 7
} athx_ply_type;

typedef struct athx_ply_layout {
    uint64_t data_offset;  /* the vertex block's first byte in the file */
    uint64_t count;        /* vertices */
    uint32_t stride;       /* bytes a vertex */
    uint32_t properties;   /* properties a vertex */
    uint32_t offset[ATHX_PLY_MAX_PROPERTIES]; /* byte offset of each in the vertex */
    uint32_t type[ATHX_PLY_MAX_PROPERTIES];   /* athx_ply_type of each */
    /* x y z, opacity (a logit), scale_0..2 (logs), rot_0..3 (w x y z), f_dc_0..2 */
    uint32_t x, y, z, opacity, scale[3], rot[4], dc[3];
    /* f_rest_*: the first one, the file's count a colour (0, 3, 8 or 15;
       red's all, then green's, then blue's) */
    uint32_t rest_base, rest_per_colour;
    /* How the fields are encoded (io::SplatEncoding, modules/io/include/
       athenea/io/RawSplats.h:85-101); all 0 for a 3DGS PLY:
         opacity   0 a logit, 1 linear, 2 a byte (v / 255)
         scale     0 a log (below -30 a flat axis), 1 linear (0 a flat axis,
                   negative a hard surfel's p), 2 SPZ's byte (v / 16 - 10)
         colour    0 the SH DC coefficient, 1 linear, 2 a byte (v / 255),
                   3 SPZ's byte, 4 light
         rotation  0 floats, 1 bytes ((v - 128) / 128), w x y z either way */
    uint32_t opacity_mode, scale_mode, colour_mode, rotation_mode;
} athx_ply_layout;

/*
 * Reads the header at the start of a PLY file (bytes of it, up to the whole
 * file; the reference reads the first 64 KiB). Errors as the reference's:
 * INVALID_ARGUMENT for not a PLY, no end_header, no format, a property before
 * any element, no vertices or a missing required property; UNSUPPORTED for a
 * format other than binary_little_endian, a list in or before the vertex
 * element, a property type outside athx_ply_type, more than 2^32 vertices or
 * ATHX_PLY_MAX_PROPERTIES properties, or harmonics that are not a whole
 * degree or not contiguous. message (may be NULL) receives the reason.
 * Host only: no device is touched.
 */
ATHX_API athx_status athx_ply_parse(const char* header, uint64_t bytes, athx_ply_layout* out, char* message,
                                    uint64_t message_bytes);

/*
 * io::readDotSplat (modules/io/src/SplatReaders.cpp:203-240): the layout of a
 * .splat file of `bytes` bytes, 32-byte records of position f32 x 3, linear
 * scale f32 x 3, rgba bytes (alpha the opacity), rotation bytes w x y z --
 * the vertex block athx_ingest_ply reads, from the file's first byte.
 * INVALID_ARGUMENT when the size is not a whole number of records (or 0).
 * Host only.
 */
ATHX_API athx_status athx_splat_layout(uint64_t bytes, athx_ply_layout* out, char* message, uint64_t message_bytes);

/* Where athx_ingest_ply writes the kept splats (device memory, the caller's);
 * it reads a PLY's vertex block or a .splat file's records alike, as the
 * layout's encodings say. */
typedef struct athx_ingest_out {
    float* positions;  /* [count][4]: x y z, linear opacity clamped to [0, 5] */
    uint32_t* shape;   /* [count][4]: quaternion word; f16 log-scales; f16 base colour */
    uint32_t* sh;      /* [count][sh_words] f16 rest coefficients, rgb a basis; NULL when none kept */
    uint32_t* origin;  /* [count] each kept splat's vertex index, or NULL */
} athx_ingest_out;

/*
 * The vertex block of a PLY (device memory: layout->count * layout->stride
 * bytes, read as bytes at any alignment) validated, decoded and packed on
 * the device (scene::CloudLoader::upload, GpuClouds.h:349-516, over
 * packed_records.slang, splat_validate.slang and splat_decode.slang): each
 * property widened to a float as the file types it; a vertex kept when its
 * position is finite, at most one log-scale is a flat -inf (below -30 in the
 * file) and the others are finite and under 60 in magnitude, and its
 * opacity (sigmoid of the logit) is at least 1/255, or, with keep_faint,
 * whenever its position is finite. Kept vertices keep file order. Rest
 * coefficients kept: min(the file's, max_degree's 0/3/8/15) a colour, whose
 * words a splat are (kept * 3 + 1) / 2. *kept (host) receives the count;
 * the call waits for it.
 */
ATHX_API athx_status athx_ingest_ply(athx_ctx* ctx, const void* vertices, const athx_ply_layout* layout,
                                     uint32_t max_degree, uint32_t keep_faint, const athx_ingest_out* out,
                                     uint32_t* kept);

/*
 * The ground truth (render::ReferenceRenderer::render, ReferenceRenderer.h:
 * 20-37; reference/reference_project.slang, reference_blend.slang): one view
 * of a packed cloud (g->packed; count splats) projected by the reference's
 * own plain projection and blended per pixel over every drawn splat in
 * (depth, index) order -- no tiles, no pairs, no ranges. Writes targets'
 * colour and depth (settings: antialias, linear, max_sh_degree, depth_mode,
 * depth_threshold, background from the view); records (device, may be NULL)
 * receives each splat's projection in the frame's record layout (12 floats:
 * centre xy, view z (-1: culled), cut power, conic a b c, alpha, rgb, 0).
 * *overflow (host) receives the pixels more than 1024 splats reach, which the
 * reference counts because its per-pixel list holds 1024: a comparison
 * means something only at zero. Slow by design: pixels x splats tests.
 * Waits for the count.
 */
ATHX_API athx_status athx_reference_render(athx_ctx* ctx, const athx_view* view, const athx_gaussians* gaussians,
                                           const athx_settings* settings, const athx_targets* targets,
                                           float* records, uint32_t* overflow);
/* The ground truth of several instances (ReferenceRenderer::render over a span
 * of SplatInstances, ReferenceRenderer.h:20-37): each packed cloud projected
 * with its own transform (athx_instance, one view), indices back to back in
 * the order given, every drawn splat of all of them walked per pixel in
 * (depth, index) order. records as above, for all instances in turn. */
ATHX_API athx_status athx_reference_render_instances(athx_ctx* ctx, const athx_view* view,
                                                     const athx_instance* instances, uint32_t ninstances,
                                                     const athx_settings* settings, const athx_targets* targets,
                                                     float* records, uint32_t* overflow);

/*
 * A sky prepared on the device (technique::Environment, modules/technique/
 * include/athenea/technique/Environment.h:1-155), into athx_sky's layout:
 * equirect, [height][width][3] linear radiance (device, FP32); texels,
 * 2 words each of athx_sky_texels(base_side) texels; sh, 16 float4. The sun
 * is the caller's to fill (athx_sky.sun). Waits for the result.
 */
ATHX_API uint32_t athx_sky_texels(uint32_t base_side);
ATHX_API athx_status athx_prepare_sky(athx_ctx* ctx, const float* equirect, uint32_t width, uint32_t height,
                                      uint32_t base_side, uint32_t* texels, float* sh);

/*
 * The sun taken out of a sky (Environment::sun, Environment.h:72-84;
 * technique/env_sun.slang:154-262): equirect as athx_prepare_sky's (device,
 * at most 2^24 texels). The sky is box-filtered 2 x 2 a level at a time until
 * at most 256 texels wide; there a texel is the sun when it is at least 16
 * times the sky's median luminance (the median of a sin-theta-weighted
 * histogram of log2 luminance in half stops from 2^-20) and 0.02 of the
 * brightest texel's, within 0.175 rad of the brightest; a sky whose brightest
 * texel is not 32 times its median has none, and a region whose light spreads
 * past twice a disc's mean square angle is left in the sky. sun (host, 8
 * floats) receives athx_sky.sun's words: (direction, the half angle of the
 * cone of the sun's solid angle, 0 for no sun), (the sun's irradiance rgb,
 * the luminance it was cut at). Waits for the result.
 */
ATHX_API athx_status athx_sky_sun(athx_ctx* ctx, const float* equirect, uint32_t width, uint32_t height,
                                  float* sun);
/*
 * The sky without its sun (technique/env_residual.slang:55-78), what the
 * relit frames' prepared sky is made from when the sun is lit on its own:
 * equirect's texels at least as bright as sun[7] within 0.175 rad of sun's
 * direction set to 0 (none when sun[3] is 0). sun (host) as athx_sky_sun wrote
 * it; residual (device) the equirect's size. Waits for the result.
 */
ATHX_API athx_status athx_sky_residual(athx_ctx* ctx, const float* equirect, uint32_t width, uint32_t height,
                                       const float* sun, float* residual);

/* Two renders compared as 8-bit sRGB code values (ReferenceRenderer.h:52-58):
   each pixel's largest channel difference, alpha included. */
typedef struct athx_image_difference {
    uint64_t pixels;
    uint32_t p99;    /* the smallest difference at least 99% of the pixels are within */
    uint32_t max;
    uint64_t over2;  /* pixels differing by more than 2 code values */
} athx_image_difference;

/* Two linear images compared as radiance (ReferenceRenderer.h:60-66). */
typedef struct athx_hdr_difference {
    uint64_t pixels;
    double rel_mse;       /* mean over pixels of the channel mean of (a - b)^2 / (b^2 + 1e-2) */
    double p99_relative;  /* 99th percentile of max_c |a - b| / max(|b|, 1e-3), to an eighth of a binade */
    double max_relative;
} athx_hdr_difference;

/*
 * a, b: [height][width] float4 linear premultiplied images (device). The
 * histograms are built on the device, a row of pixels a thread as the
 * reference's image_compare.slang and image_compare_hdr.slang do; the
 * percentiles are read off them on the host. Wait for the result.
 */
ATHX_API athx_status athx_compare_images(athx_ctx* ctx, const float* a, const float* b, uint32_t width,
                                         uint32_t height, athx_image_difference* out);
ATHX_API athx_status athx_compare_hdr(athx_ctx* ctx, const float* a, const float* b, uint32_t width, uint32_t height,
                                      athx_hdr_difference* out);

#ifdef __cplusplus
}
#endif

#endif /* ATHX_H */
