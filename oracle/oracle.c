/*
 * oracle.c -- TEST INFRASTRUCTURE. A CPU restatement of athenea's Gaussian
 * splat tile rasterizer, used only as the checker of the HIP path (by tests/,
 * the smoke check and the CPU-baseline leg of bench.py). Nothing in the
 * product links or calls it.
 *
 * What it restates, each function citing the reference it follows:
 *   orc_project      shaders/athenea/splat/splat_project.slang:497-1130
 *                    (projectSplat<0>: no edit, no lens, no shutter, no
 *                    surfels, not relit), helpers frame.slang:283-402,
 *                    sh.slang:25-51, color.slang:7-41
 *   orc_emit_pairs   one (tile<<24 | depth24, slot) pair per touched tile, in
 *                    slot order (splat_emit.slang:17-50's tile test)
 *   orc_sort_pairs   stable LSD radix sort over the key's significant bits
 *                    (RadixSort.cpp:95-261 semantics)
 *   orc_ranges       splat_tiles_clear.slang:8-17 + splat_ranges.slang:11-25
 *   orc_blend        splat_blend.slang:236-578 (splatBlend: composite, matte
 *                    and trace off, contributionMode 0)
 *   orc_blend_f64    the same walk with every FP32 decision (skip, clamp,
 *                    stop) replayed and the arithmetic in FP64
 *   orc_reference    reference/reference_blend.slang:28-123 brute force
 *   orc_render_ref_order  the reference's own two-sort order: depth sort of
 *                    visible slots (24 bits, stable), emit in depth order,
 *                    stable tile sort (TileRasterizer.cpp:1111-1176)
 *   orc_fd_grads     FP64 central finite differences of an L2 image loss
 *                    with every discrete decision frozen
 *   orc_ref_project  reference/reference_project.slang:1-107, the ground
 *                    truth's own projection of the packed layout (its
 *                    blend is orc_reference)
 *   orc_compare_codes / orc_compare_hdr
 *                    reference/image_compare.slang, image_compare_hdr.slang:
 *                    the per-row histograms and relMSE rows, summed
 *   orc_contribution splat_blend.slang:364-408's contribution modes 1 to 4
 *                    over the same per-pixel walk as orc_blend
 *   orc_ingest       a PLY vertex block widened, validated, decoded and
 *                    packed: packed_records.slang:1-110, splat_validate.slang,
 *                    splat_decode.slang:1-173 (PLY's encodings),
 *                    splat_encoding.slang, packing.slang:33-57,319-328
 *
 * Parity of each function is pinned by tests/test_oracle.py (closed forms,
 * invariants, brute force and the reference's own test assertions).
 *
 * Build: oracle/Makefile (gcc, -ffp-contract=off so every a*b+c here is a
 * multiply then an add, as in the kernels; only explicit fmaf() fuses).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/athx_detmath.h"

#define ORC_EXPORT __attribute__((visibility("default")))

/* What a frame's projection is told (frame.slang:51-230, the fields the plain
 * path reads). Layout mirrored by oracle/pyoracle.py. */
typedef struct orc_view {
    uint32_t width, height;
    float focal_x, focal_y, centre_x, centre_y;
    float near_z, far_z;
    uint32_t orthographic;
    uint32_t antialias;
    uint32_t linear;      /* 1: colours are linear light; 0: sRGB, decoded per splat */
    uint32_t sh_limit;    /* highest degree evaluated this frame */
    uint32_t tile;        /* tile side in pixels (16 in the reference) */
    float m[12];          /* object -> view, rows of a 3x4 */
    float eye[3];         /* the camera in the cloud's own space */
    float bg[4];          /* premultiplied linear background */
    uint32_t depth_mode;  /* 0 alpha-weighted mean, 1 first crossing */
    float depth_threshold;
} orc_view;

/* frame.slang:42-48 */
#define CULLED_KEY 0xFF000000u
#define CULL_DEPTH 2u
#define CULL_DEGENERATE 3u
#define CULL_FAINT 4u
#define CULL_OFFSCREEN 5u
#define CULL_NO_TILE 6u

/* Per-slot record layout: screen (x, y, z, cutoff), conic (a, b, c, alpha),
 * colour (r, g, b, branch bits FL_*) -- frame.slang:269-281 without the
 * surfel words. */
#define REC 12

/* Branch flags the FP32 projection took, replayed by the FP64 one. */
#define FL_CLAMP_XLO 1u
#define FL_CLAMP_XHI 2u
#define FL_CLAMP_YLO 4u
#define FL_CLAMP_YHI 8u
#define FL_PAID_BACK 16u /* paidOpacity's spread >= 1/255 branch */
#define FL_COL_R 32u     /* colour channel kept (> 0 for a linear cloud) */
#define FL_COL_G 64u
#define FL_COL_B 128u
#define FL_SRGB_R 256u   /* sRGB decode took the power branch */
#define FL_SRGB_G 512u
#define FL_SRGB_B 1024u
#define FL_RELIT 2048u  /* the colour is the relit one (no SH adjoint through it) */

static const float kSH0 = 0.28209479177387814f;
static const float kC1 = 0.4886025119029199f;
static const float kC2[5] = {1.0925484305920792f, -1.0925484305920792f, 0.31539156525252005f,
                             -1.0925484305920792f, 0.5462742152960396f};
static const float kC3[7] = {-0.5900435899266435f, 2.890611442640554f, -0.4570457994644658f,
                             0.3731763325901154f, -0.4570457994644658f, 1.445305721320277f,
                             -0.5900435899266435f};

static inline float minf_(float a, float b) { return a < b ? a : b; }
static inline float maxf_(float a, float b) { return a > b ? a : b; }
static inline float clampf_(float x, float lo, float hi) { return minf_(maxf_(x, lo), hi); }

/* frame.slang:325-327 */
static float cutoff_power(float alpha) {
    return minf_(2.0f * athx_logf(maxf_(255.0f * alpha, 1.0001f)), 9.0f);
}

/* frame.slang:316-319 */
static float paid_opacity(float alpha, float paid, uint32_t* flags) {
    const float spread = paid * alpha;
    if (spread >= 1.0f / 255.0f) {
        *flags |= FL_PAID_BACK;
        return spread + (1.0f - paid) * (1.0f / 255.0f);
    }
    return spread;
}

/* frame.slang:354-371: the least of a x^2 + 2 b x y + c y^2 over a box, each
 * edge at its clamped 1D minimiser -b * x / c (that is ((-b) x) / c). */
static float min_power_over_box(float a, float b, float c, float lox, float loy, float hix, float hiy) {
    if (lox <= 0.0f && hix >= 0.0f && loy <= 0.0f && hiy >= 0.0f) {
        return 0.0f;
    }
    float best = 3.4e38f;
    for (int e = 0; e < 2; ++e) {
        const float x = e == 0 ? lox : hix;
        const float y = c > 1e-12f ? clampf_(((-b) * x) / c, loy, hiy) : loy;
        best = minf_(best, a * x * x + 2.0f * b * x * y + c * y * y);
    }
    for (int e = 0; e < 2; ++e) {
        const float y = e == 0 ? loy : hiy;
        const float x = a > 1e-12f ? clampf_(((-b) * y) / a, lox, hix) : lox;
        best = minf_(best, a * x * x + 2.0f * b * x * y + c * y * y);
    }
    return best;
}

/* frame.slang:374-379 */
static int tile_touched(const orc_view* v, float cx, float cy, float a, float b, float c, float cutoff,
                        uint32_t tx, uint32_t ty) {
    const uint32_t T = v->tile;
    const float fx = (float)(tx * T) + 0.5f;
    const float fy = (float)(ty * T) + 0.5f;
    const uint32_t ex = (tx + 1) * T < v->width ? (tx + 1) * T : v->width;
    const uint32_t ey = (ty + 1) * T < v->height ? (ty + 1) * T : v->height;
    const float lx = (float)(ex - 1) + 0.5f;
    const float ly = (float)(ey - 1) + 0.5f;
    return min_power_over_box(a, b, c, fx - cx, fy - cy, lx - cx, ly - cy) <= cutoff;
}

/* frame.slang:383-387 then 390-402. Returns 0 when off screen. */
static int tile_rect(const orc_view* v, float cx, float cy, float a, float b, float c, float cutoff,
                     uint32_t rect[4]) {
    const float det = a * c - b * b;
    const float inv = det > 1e-20f ? 1.0f / det : 0.0f;
    const float ex = sqrtf(maxf_(cutoff * c * inv, 0.0f));
    const float ey = sqrtf(maxf_(cutoff * a * inv, 0.0f));
    const float x1 = cx - ex, y1 = cy - ey, x2 = cx + ex, y2 = cy + ey;
    if (x2 < 0.0f || y2 < 0.0f || x1 >= (float)v->width || y1 >= (float)v->height) {
        rect[0] = 1; rect[1] = 1; rect[2] = 0; rect[3] = 0;
        return 0;
    }
    const uint32_t T = v->tile;
    rect[0] = (uint32_t)maxf_(x1, 0.0f) / T;
    rect[1] = (uint32_t)maxf_(y1, 0.0f) / T;
    rect[2] = (uint32_t)minf_(x2, (float)v->width - 1.0f) / T;
    rect[3] = (uint32_t)minf_(y2, (float)v->height - 1.0f) / T;
    return rect[0] <= rect[2] && rect[1] <= rect[3];
}

/* color.slang:7-10 through the deterministic power. */
static float srgb_to_linear(float c, uint32_t* took_pow) {
    c = maxf_(c, 0.0f);
    if (c <= 0.04045f) {
        *took_pow = 0;
        return c / 12.92f;
    }
    *took_pow = 1;
    return athx_powf_pos((c + 0.055f) / 1.055f, 2.4f);
}

/* sh.slang:25-51 for one colour channel; coefficient k of channel ch at sh[k*3+ch]
 * (k = 0 is the DC term, read by the caller). */
static float sh_rest(const float* sh, uint32_t ch, uint32_t degree, float x, float y, float z) {
    if (degree == 0) {
        return 0.0f;
    }
#define B(k) sh[(k) * 3 + ch]
    float r = -kC1 * y * B(1) + kC1 * z * B(2) - kC1 * x * B(3);
    if (degree < 2) {
        return r;
    }
    const float xx = x * x, yy = y * y, zz = z * z;
    const float xy = x * y, yz = y * z, xz = x * z;
    r = r + (kC2[0] * xy * B(4) + kC2[1] * yz * B(5) + kC2[2] * (2.0f * zz - xx - yy) * B(6) +
             kC2[3] * xz * B(7) + kC2[4] * (xx - yy) * B(8));
    if (degree < 3) {
        return r;
    }
    r = r + (kC3[0] * y * (3.0f * xx - yy) * B(9) + kC3[1] * xy * z * B(10) +
             kC3[2] * y * (4.0f * zz - xx - yy) * B(11) + kC3[3] * z * (2.0f * zz - 3.0f * xx - 3.0f * yy) * B(12) +
             kC3[4] * x * (4.0f * zz - xx - yy) * B(13) + kC3[5] * z * (xx - yy) * B(14) +
             kC3[6] * x * (xx - 3.0f * yy) * B(15));
#undef B
    return r;
}

static uint32_t degree_of(uint32_t coeffs) {
    return coeffs >= 16 ? 3u : coeffs >= 9 ? 2u : coeffs >= 4 ? 1u : 0u;
}

/* ---------------------------------------------------------------------------
 * Relit colour, the non-transfer dome branch: relitSplat
 * (shaders/athenea/splat/splat_relight.slang:1775-1824) with one dome whose
 * sky was prepared (relitByDome :1640-1692), a surface with the plain lobes
 * (no coat, sheen, transmission or emission), and the prepared-sky reader of
 * shaders/athenea/light/environment.slang:50-225.
 */
typedef struct orc_relight {
    const float* normals;   /* [n][3] stored shading normals (cloud space), or NULL: the disc's axis */
    const float* albedo;    /* [n][3] or NULL: the evaluated harmonics are the albedo (splat_project.slang:819) */
    const float* metallic;  /* [n] */
    const float* roughness; /* [n] */
    float o2w[12];          /* object -> world rows */
    const uint32_t* texels; /* 2 words a texel: rgb halves */
    const float* sky_sh;    /* 16 float4: radiance coefficients */
    const float* sky_sun;   /* 2 float4: direction + disc size (w > 0: a sun), irradiance */
    uint32_t base_side;
    const uint32_t* normals_packed; /* [n] octahedral words (packing.slang:116-135), or NULL: normals */
    const uint32_t* pbr;            /* [n] metallic byte 0, roughness byte 1 (splat_encoding.slang:212-213), or NULL */
} orc_relight;

#define ENV_LEVELS 8u
#define ENV_COARSE 256u
#define ENV_FLOOR 16u

static inline float lerpf_(float a, float b, float t) { return a + (b - a) * t; }
static inline float satf_(float x) { return clampf_(x, 0.0f, 1.0f); }

/* IEEE binary16 to binary32, exactly. */
static float half_to_float(uint32_t h) {
    const uint32_t sign = (h & 0x8000u) << 16, e = (h >> 10) & 0x1Fu, m = h & 0x3FFu;
    uint32_t bits;
    if (e == 0) {
        if (m == 0) {
            bits = sign;
        } else {
            int ee = -1;
            uint32_t mm = m;
            do { ++ee; mm <<= 1; } while (!(mm & 0x400u));
            bits = sign | ((uint32_t)(127 - 15 - ee) << 23) | ((mm & 0x3FFu) << 13);
        }
    } else if (e == 31) {
        bits = sign | 0x7F800000u | (m << 13);
    } else {
        bits = sign | ((e - 15 + 127) << 23) | (m << 13);
    }
    float f;
    memcpy(&f, &bits, 4);
    return f;
}

/* environment.slang:66-69 */
static uint32_t env_side(uint32_t level, uint32_t base_side) {
    const uint32_t base = base_side > ENV_COARSE ? base_side : ENV_COARSE;
    const uint32_t side = base >> (level < ENV_LEVELS - 1 ? level : ENV_LEVELS - 1);
    return side > ENV_FLOOR ? side : ENV_FLOOR;
}

/* environment.slang:77-92, the walk (the closed form's same sum) */
static uint32_t env_level_offset(uint32_t level, uint32_t base_side) {
    uint32_t at = 0;
    for (uint32_t k = 0; k < (level < ENV_LEVELS ? level : ENV_LEVELS); ++k) {
        const uint32_t side = env_side(k, base_side);
        at += side * side;
    }
    return at;
}

/* common/octahedral.slang:12-23 */
static void oct_encode(float nx, float ny, float nz, float* u, float* v) {
    const float s = maxf_(fabsf(nx) + fabsf(ny) + fabsf(nz), 1.0e-20f);
    nx = nx / s;
    ny = ny / s;
    nz = nz / s;
    float px = nx, py = ny;
    if (!(nz >= 0.0f)) {
        px = (1.0f - fabsf(ny)) * (nx >= 0.0f ? 1.0f : -1.0f);
        py = (1.0f - fabsf(nx)) * (ny >= 0.0f ? 1.0f : -1.0f);
    }
    *u = px * 0.5f + 0.5f;
    *v = py * 0.5f + 0.5f;
}

static void env_words(const uint32_t* texels, uint32_t texel, float out[3]) {
    const uint32_t lo = texels[texel * 2], hi = texels[texel * 2 + 1];
    out[0] = half_to_float(lo & 0xFFFFu);
    out[1] = half_to_float(lo >> 16);
    out[2] = half_to_float(hi & 0xFFFFu);
}

/* environment.slang:133-147: bilinear over one level, clamped at the seam */
static void env_level_at(const orc_relight* rl, uint32_t level, const float dir[3], float out[3]) {
    const uint32_t side_u = env_side(level, rl->base_side);
    const float side = (float)side_u;
    float u, v;
    oct_encode(dir[0], dir[1], dir[2], &u, &v);
    const float atx = clampf_(u * side - 0.5f, 0.0f, side - 1.0f);
    const float aty = clampf_(v * side - 0.5f, 0.0f, side - 1.0f);
    const uint32_t lox = (uint32_t)floorf(atx), loy = (uint32_t)floorf(aty);
    const uint32_t hix = lox + 1 < side_u - 1 ? lox + 1 : side_u - 1;
    const uint32_t hiy = loy + 1 < side_u - 1 ? loy + 1 : side_u - 1;
    const float fx = atx - (float)lox, fy = aty - (float)loy;
    const uint32_t first = env_level_offset(level, rl->base_side);
    float v00[3], v10[3], v01[3], v11[3];
    env_words(rl->texels, first + loy * side_u + lox, v00);
    env_words(rl->texels, first + loy * side_u + hix, v10);
    env_words(rl->texels, first + hiy * side_u + lox, v01);
    env_words(rl->texels, first + hiy * side_u + hix, v11);
    for (int c = 0; c < 3; ++c) {
        out[c] = lerpf_(lerpf_(v00[c], v10[c], fx), lerpf_(v01[c], v11[c], fx), fy);
    }
}

/* environment.slang:162-170: the two levels that bracket sqrt(roughness) */
static void env_radiance(const orc_relight* rl, const float dir[3], float roughness, float out[3]) {
    const float level = sqrtf(satf_(roughness)) * (float)(ENV_LEVELS - 1);
    const uint32_t lo = (uint32_t)floorf(level);
    const uint32_t hi = lo + 1 < ENV_LEVELS - 1 ? lo + 1 : ENV_LEVELS - 1;
    const float f = level - (float)lo;
    float a[3], b[3];
    env_level_at(rl, lo, dir, a);
    env_level_at(rl, hi, dir, b);
    for (int c = 0; c < 3; ++c) out[c] = lerpf_(a[c], b[c], f);
}

/* environment.slang:182-223: irradiance at n from nine radiance coefficients */
static float env_basis(uint32_t k, float x, float y, float z) {
    switch (k) {
    case 0: return 0.28209479177387814f;
    case 1: return -0.4886025119029199f * y;
    case 2: return 0.4886025119029199f * z;
    case 3: return -0.4886025119029199f * x;
    case 4: return 1.0925484305920792f * x * y;
    case 5: return -1.0925484305920792f * y * z;
    case 6: return 0.31539156525252005f * (2.0f * z * z - x * x - y * y);
    case 7: return -1.0925484305920792f * x * z;
    default: return 0.5462742152960396f * (x * x - y * y);
    }
}

static void env_irradiance(const orc_relight* rl, const float n[3], float out[3]) {
    out[0] = out[1] = out[2] = 0.0f;
    for (uint32_t k = 0; k < 9; ++k) {
        const float band = k == 0 ? 3.14159265358979f : (k < 4 ? 2.0943951023932f : 0.78539816339745f);
        const float w = band * env_basis(k, n[0], n[1], n[2]);
        for (int c = 0; c < 3; ++c) out[c] = out[c] + rl->sky_sh[k * 4 + c] * w;
    }
}

/* material/microfacet.slang:66-82 (MaterialX's fit); returns (a, b) */
static void ggx_albedo_ab(float x, float y, float* ra, float* rb) {
    const float x2 = x * x, y2 = y * y;
    static const float c[9][4] = {{0.1003f, 0.9345f, 1.0f, 1.0f},     {-0.6303f, -2.323f, -1.765f, 0.2281f},
                                  {9.748f, 2.229f, 8.263f, 15.94f},    {-2.038f, -3.748f, 11.53f, -55.83f},
                                  {29.34f, 1.424f, 28.96f, 13.08f},    {-8.245f, -0.7684f, -7.507f, 41.26f},
                                  {-26.44f, 1.436f, -36.11f, 54.9f},   {19.99f, 0.2913f, 15.86f, 300.2f},
                                  {-5.448f, 0.6286f, 33.37f, -285.1f}};
    float r[4];
    for (int i = 0; i < 4; ++i) {
        r[i] = c[0][i] + c[1][i] * x + c[2][i] * y + c[3][i] * x * y + c[4][i] * x2 + c[5][i] * y2 +
               c[6][i] * x2 * y + c[7][i] * x * y2 + c[8][i] * x2 * y2;
    }
    *ra = satf_(r[0] / r[2]);
    *rb = satf_(r[1] / r[3]);
}

/* microfacet.slang:85-88 (Turquin 2019) */
static float ggx_compensation(float nv, float alpha, float fss) {
    float a, b;
    ggx_albedo_ab(nv, alpha, &a, &b);
    const float ess = a * 1.0f + 1.0f * b;
    return 1.0f + fss * (1.0f - ess) / ess;
}

/* microfacet.slang:99-107 */
static float fresnel_dielectric(float c, float ior) {
    const float g2 = ior * ior + c * c - 1.0f;
    if (g2 < 0.0f) return 1.0f;
    const float g = sqrtf(g2);
    const float q = (g - c) / (g + c);
    const float t = ((g + c) * c - 1.0f) / ((g - c) * c + 1.0f);
    return 0.5f * (q * q) * (1.0f + t * t);
}

/* splat_relight.slang:455-460 at the plain lobes' index 1.5 (dielectricF0 = 0.04, tint 1) */
static float ggx_env_dielectric(float roughness, float nv) {
    const float alpha = maxf_(roughness * roughness, 1.0e-4f);
    const float F = fresnel_dielectric(nv, 1.5f);
    float a, b;
    ggx_albedo_ab(nv, alpha, &a, &b);
    return (0.04f * a + 1.0f * b) * ggx_compensation(nv, alpha, F);
}

/* splat_relight.slang:472-488 with a white edge: mxArtisticIor (microfacet.slang:117-126),
 * fresnelConductor (:128-143), the fit and its compensation, one channel */
static float ggx_env_conductor(float albedo, float roughness, float nv) {
    const float alpha = maxf_(roughness * roughness, 1.0e-4f);
    const float r = clampf_(satf_(albedo), 0.0f, 0.99f);
    const float root = sqrtf(r);
    const float low = (1.0f - r) / (1.0f + r);
    const float high = (1.0f + root) / (1.0f - root);
    const float n = lerpf_(high, low, 1.0f);
    const float up = n + 1.0f, down = n - 1.0f;
    const float k = sqrtf(maxf_((up * up * r - down * down) / maxf_(1.0f - r, 1.0e-6f), 0.0f));
    const float ct = satf_(nv);
    const float ct2 = ct * ct, st2 = 1.0f - ct2;
    const float n2 = n * n, k2 = k * k;
    const float t0 = n2 - k2 - st2;
    const float a2b2 = sqrtf(t0 * t0 + 4.0f * n2 * k2);
    const float t1 = a2b2 + ct2;
    const float aa = sqrtf(maxf_(0.5f * (a2b2 + t0), 0.0f));
    const float t2 = 2.0f * aa * nv;
    const float rs = (t1 - t2) / (t1 + t2);
    const float t3 = ct2 * a2b2 + st2 * st2;
    const float t4 = t2 * st2;
    const float rp = rs * (t3 - t4) / (t3 + t4);
    const float F = 0.5f * (rp + rs);
    float a, b;
    ggx_albedo_ab(nv, alpha, &a, &b);
    return (F * a + F * b) * ggx_compensation(nv, alpha, F);
}

void orc_unpack_normal(uint32_t word, float out[3]);

/* splat_relight.slang:1860-1880 (splatShadingNormal with splatNormal :1843-1850) */
static void shading_normal(const float R[3][3], float sx, float sy, float sz, const float p[3], const float eye[3],
                           const float* stored, float out[3]) {
    const uint32_t s = sx <= sy ? (sx <= sz ? 0u : 2u) : (sy <= sz ? 1u : 2u);
    float d[3] = {R[0][s], R[1][s], R[2][s]};
    const float dl = 1.0f / sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    d[0] = d[0] * dl; d[1] = d[1] * dl; d[2] = d[2] * dl;
    float e[3] = {eye[0] - p[0], eye[1] - p[1], eye[2] - p[2]};
    const float el = 1.0f / sqrtf(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
    const float toward = d[0] * (e[0] * el) + d[1] * (e[1] * el) + d[2] * (e[2] * el);
    if (toward < 0.0f) {
        d[0] = -d[0]; d[1] = -d[1]; d[2] = -d[2];
    }
    if (!stored) {
        out[0] = d[0]; out[1] = d[1]; out[2] = d[2];
        return;
    }
    const float along = stored[0] * d[0] + stored[1] * d[1] + stored[2] * d[2];
    float flip;
    if (fabsf(along) >= 0.5f) {
        flip = along < 0.0f ? -1.0f : 1.0f;
    } else {
        flip = stored[0] * e[0] + stored[1] * e[1] + stored[2] * e[2] < 0.0f ? -1.0f : 1.0f;
    }
    out[0] = flip * stored[0]; out[1] = flip * stored[1]; out[2] = flip * stored[2];
}

/* The relit colour of slot i: relitSplat's dome term, body * irradiance / pi
 * (+ the sun) + the environment BRDF * the prepared sky along the mirror. */
static void relit_colour(const orc_relight* rl, uint32_t i, const float R[3][3], float sx, float sy, float sz,
                         const float p[3], const float eye_obj[3], const float albedo_sh[3], float out[3]) {
    float albedo[3] = {albedo_sh[0], albedo_sh[1], albedo_sh[2]};
    if (rl->albedo) {
        albedo[0] = rl->albedo[i * 3 + 0]; albedo[1] = rl->albedo[i * 3 + 1]; albedo[2] = rl->albedo[i * 3 + 2];
    }
    const float metallic = rl->pbr ? (float)(rl->pbr[i] & 0xFFu) / 255.0f : rl->metallic[i];
    const float roughness = rl->pbr ? (float)((rl->pbr[i] >> 8) & 0xFFu) / 255.0f : rl->roughness[i];
    float nc[3], unpacked[3];
    const float* stored = NULL;
    if (rl->normals_packed) {
        orc_unpack_normal(rl->normals_packed[i], unpacked);
        stored = unpacked;
    } else if (rl->normals) {
        stored = rl->normals + (size_t)i * 3;
    }
    shading_normal(R, sx, sy, sz, p, eye_obj, stored, nc);
    /* To the world (relightToWorld, relightNormalToWorld :1882-1915). */
    const float* w = rl->o2w;
    const float pw[3] = {w[0] * p[0] + w[1] * p[1] + w[2] * p[2] + w[3] * 1.0f,
                         w[4] * p[0] + w[5] * p[1] + w[6] * p[2] + w[7] * 1.0f,
                         w[8] * p[0] + w[9] * p[1] + w[10] * p[2] + w[11] * 1.0f};
    const float a[3] = {w[0], w[4], w[8]}, b[3] = {w[1], w[5], w[9]}, c[3] = {w[2], w[6], w[10]};
    const float bc[3] = {b[1] * c[2] - b[2] * c[1], b[2] * c[0] - b[0] * c[2], b[0] * c[1] - b[1] * c[0]};
    const float ca[3] = {c[1] * a[2] - c[2] * a[1], c[2] * a[0] - c[0] * a[2], c[0] * a[1] - c[1] * a[0]};
    const float ab[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
    const float side = a[0] * bc[0] + a[1] * bc[1] + a[2] * bc[2] < 0.0f ? -1.0f : 1.0f;
    float n[3];
    for (int k = 0; k < 3; ++k) n[k] = side * (nc[0] * bc[k] + nc[1] * ca[k] + nc[2] * ab[k]);
    const float nn = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
    if (nn > 1.0e-20f) {
        const float inv = 1.0f / sqrtf(nn);
        n[0] = n[0] * inv; n[1] = n[1] * inv; n[2] = n[2] * inv;
    } else {
        n[0] = 0.0f; n[1] = 0.0f; n[2] = 1.0f;
    }
    /* The eye in the world: the object-space eye taken through the same rows. */
    const float* e = eye_obj;
    const float ew[3] = {w[0] * e[0] + w[1] * e[1] + w[2] * e[2] + w[3] * 1.0f,
                         w[4] * e[0] + w[5] * e[1] + w[6] * e[2] + w[7] * 1.0f,
                         w[8] * e[0] + w[9] * e[1] + w[10] * e[2] + w[11] * 1.0f};
    float wo[3] = {ew[0] - pw[0], ew[1] - pw[1], ew[2] - pw[2]};
    const float wl = 1.0f / sqrtf(wo[0] * wo[0] + wo[1] * wo[1] + wo[2] * wo[2]);
    wo[0] = wo[0] * wl; wo[1] = wo[1] * wl; wo[2] = wo[2] * wl;
    /* relitByDome, the prepared dome (:1379-1384, 1640-1692) */
    const float ndv = n[0] * wo[0] + n[1] * wo[1] + n[2] * wo[2];
    const float nv = maxf_(ndv, 1.0e-4f);
    const float i_dot = n[0] * -wo[0] + n[1] * -wo[1] + n[2] * -wo[2];
    const float mirror[3] = {-wo[0] - 2.0f * i_dot * n[0], -wo[1] - 2.0f * i_dot * n[1], -wo[2] - 2.0f * i_dot * n[2]};
    const float over_pi = 1.0f / 3.14159265358979f;
    float above[3];
    env_irradiance(rl, n, above);
    for (int k = 0; k < 3; ++k) above[k] = above[k] * over_pi;
    if (rl->sky_sun[3] > 0.0f) {
        const float cosine = n[0] * rl->sky_sun[0] + n[1] * rl->sky_sun[1] + n[2] * rl->sky_sun[2];
        for (int k = 0; k < 3; ++k) above[k] = above[k] + rl->sky_sun[4 + k] * (maxf_(cosine, 0.0f) * over_pi);
    }
    float reflected[3];
    env_radiance(rl, mirror, roughness, reflected);
    const float diel = ggx_env_dielectric(roughness, nv);
    for (int k = 0; k < 3; ++k) {
        const float body = albedo[k] * (1.0f - metallic);
        const float polish = lerpf_(diel, ggx_env_conductor(albedo[k], roughness, nv), satf_(metallic));
        out[k] = body * above[k] + polish * reflected[k];
    }
}

/* A Gaussian as the projection reads it, from either input layout. */
typedef struct orc_gaussian {
    float p[3], s[3], q[4]; /* position, scales (exp'd), rotation (x, y, z, w) as used */
    float opacity;          /* linear */
    float base[3];          /* the colour before the harmonics' rest */
    float row[48];          /* row[k * 3 + c]: coefficient k >= 1 of channel c */
    uint32_t coeffs;        /* coefficients a channel, constant term included */
} orc_gaussian;

/* The FP32 arrays: exp of the log-scales, sigmoid of the logit, the
 * quaternion normalised, base 0.5 + SH0 dc. */
static void decode_soa(const float* P, const float* LS, const float* Q, const float* LG, const float* SH,
                       uint32_t coeffs, uint32_t i, orc_gaussian* g) {
    for (int k = 0; k < 3; ++k) {
        g->p[k] = P[i * 3 + k];
        g->s[k] = athx_expf(LS[i * 3 + k]);
    }
    g->opacity = athx_sigmoidf(LG[i]);
    float qx = Q[i * 4 + 0], qy = Q[i * 4 + 1], qz = Q[i * 4 + 2], qw = Q[i * 4 + 3];
    const float qn = qx * qx + qy * qy + qz * qz + qw * qw;
    if (qn > 0.0f) {
        const float inv = 1.0f / sqrtf(qn);
        qx = qx * inv; qy = qy * inv; qz = qz * inv; qw = qw * inv;
    } else {
        qx = 0.0f; qy = 0.0f; qz = 0.0f; qw = 1.0f;
    }
    g->q[0] = qx; g->q[1] = qy; g->q[2] = qz; g->q[3] = qw;
    const float* sh = SH + (size_t)i * coeffs * 3;
    for (int c = 0; c < 3; ++c) g->base[c] = 0.5f + kSH0 * sh[c];
    memset(g->row, 0, sizeof g->row);
    for (uint32_t k = 3; k < coeffs * 3; ++k) g->row[k] = sh[k];
    g->coeffs = coeffs;
}

/* packing.slang:59-76 (decodeQuaternion), used as decoded. */
static void decode_quaternion(uint32_t word, float q[4]) {
    const float kSqrtHalf = 0.70710678f;
    const uint32_t largest = word & 3u;
    const float s0 = ((float)((word >> 22) & 1023u) / 1023.0f * 2.0f - 1.0f) * kSqrtHalf;
    const float s1 = ((float)((word >> 12) & 1023u) / 1023.0f * 2.0f - 1.0f) * kSqrtHalf;
    const float s2 = ((float)((word >> 2) & 1023u) / 1023.0f * 2.0f - 1.0f) * kSqrtHalf;
    const float rest = sqrtf(maxf_(0.0f, 1.0f - (s0 * s0 + s1 * s1 + s2 * s2)));
    const float small[3] = {s0, s1, s2};
    uint32_t next = 0;
    for (uint32_t k = 0; k < 4; ++k) {
        if (k == largest) {
            q[k] = rest;
        } else {
            q[k] = small[next++];
        }
    }
}

/* athenea's packed layout (packing.slang:1-26): positions float4 (xyz,
 * linear opacity); shape words (quaternion; f16 log-scales; f16 base
 * colour); the rest coefficients as halves, rgb per basis (sh.slang:15-22). */
static void decode_packed(const float* positions, const uint32_t* shape, const uint32_t* words, uint32_t rest,
                          uint32_t i, orc_gaussian* g) {
    for (int k = 0; k < 3; ++k) g->p[k] = positions[i * 4 + k];
    g->opacity = positions[i * 4 + 3];
    const uint32_t* w = shape + (size_t)i * 4;
    g->s[0] = athx_expf(half_to_float(w[1] & 0xFFFFu));
    g->s[1] = athx_expf(half_to_float(w[1] >> 16));
    g->s[2] = athx_expf(half_to_float(w[2] & 0xFFFFu));
    decode_quaternion(w[0], g->q);
    g->base[0] = half_to_float(w[2] >> 16);
    g->base[1] = half_to_float(w[3] & 0xFFFFu);
    g->base[2] = half_to_float(w[3] >> 16);
    memset(g->row, 0, sizeof g->row);
    const uint32_t nwords = (rest * 3 + 1) / 2;
    for (uint32_t h = 0; h < rest * 3; ++h) {
        const uint32_t word = words[(size_t)i * nwords + h / 2];
        g->row[3 + h] = half_to_float((h & 1u) ? word >> 16 : word & 0xFFFFu);
    }
    g->coeffs = rest + 1;
}

/* One splat projected: splat_project.slang:497-1130 with kTransfer 0. Writes
 * rec, touched and key for slot i; returns the cull reason (0 = visible). */
static uint32_t project_one(const orc_view* v, const orc_gaussian* g, uint32_t i, float* rec, uint32_t* touched,
                            uint32_t* key, uint32_t* flags_out, const orc_relight* rl) {
    uint32_t flags = 0;
    const float px = g->p[0], py = g->p[1], pz = g->p[2];
    const float sx = g->s[0], sy = g->s[1], sz = g->s[2];
    const float opacity = g->opacity;
    const float* m = v->m;
    const float tx = m[0] * px + m[1] * py + m[2] * pz + m[3];
    const float ty = m[4] * px + m[5] * py + m[6] * pz + m[7];
    const float tz = m[8] * px + m[9] * py + m[10] * pz + m[11];
    *touched = 0;
    if (flags_out) *flags_out = 0;
    if (!(tz > v->near_z) || tz > v->far_z) {
        *key = CULLED_KEY | CULL_DEPTH;
        return CULL_DEPTH;
    }
    const float qx = g->q[0], qy = g->q[1], qz = g->q[2], qw = g->q[3];
    const float r00 = 1.0f - 2.0f * (qy * qy + qz * qz), r01 = 2.0f * (qx * qy - qw * qz), r02 = 2.0f * (qx * qz + qw * qy);
    const float r10 = 2.0f * (qx * qy + qw * qz), r11 = 1.0f - 2.0f * (qx * qx + qz * qz), r12 = 2.0f * (qy * qz - qw * qx);
    const float r20 = 2.0f * (qx * qz - qw * qy), r21 = 2.0f * (qy * qz + qw * qx), r22 = 1.0f - 2.0f * (qx * qx + qy * qy);
    const float m00 = r00 * sx, m01 = r01 * sy, m02 = r02 * sz;
    const float m10 = r10 * sx, m11 = r11 * sy, m12 = r12 * sz;
    const float m20 = r20 * sx, m21 = r21 * sy, m22 = r22 * sz;

    float jx0, jx2, jy1, jy2, cx, cy;
    if (v->orthographic) {
        jx0 = v->focal_x; jx2 = 0.0f;
        jy1 = v->focal_y; jy2 = 0.0f;
        cx = v->focal_x * tx + v->centre_x;
        cy = v->focal_y * ty + v->centre_y;
    } else {
        const float invz = 1.0f / tz;
        cx = v->focal_x * tx * invz + v->centre_x;
        cy = v->focal_y * ty * invz + v->centre_y;
        const float W = (float)v->width, H = (float)v->height;
        const float limlox = (-0.15f * W - v->centre_x) / v->focal_x;
        const float limhix = (1.15f * W - v->centre_x) / v->focal_x;
        const float limloy = (-0.15f * H - v->centre_y) / v->focal_y;
        const float limhiy = (1.15f * H - v->centre_y) / v->focal_y;
        const float ux = tx * invz, uy = ty * invz;
        if (ux < limlox) flags |= FL_CLAMP_XLO;
        if (ux > limhix) flags |= FL_CLAMP_XHI;
        if (uy < limloy) flags |= FL_CLAMP_YLO;
        if (uy > limhiy) flags |= FL_CLAMP_YHI;
        const float txc = clampf_(ux, limlox, limhix) * tz;
        const float tyc = clampf_(uy, limloy, limhiy) * tz;
        jx0 = v->focal_x * invz;
        jx2 = -v->focal_x * txc * invz * invz;
        jy1 = v->focal_y * invz;
        jy2 = -v->focal_y * tyc * invz * invz;
    }
    /* T = J W, J's zero entries dropped: row x is (jx0, 0, jx2), row y (0, jy1, jy2). */
    const float t30 = jx0 * m[0] + jx2 * m[8];
    const float t31 = jx0 * m[1] + jx2 * m[9];
    const float t32 = jx0 * m[2] + jx2 * m[10];
    const float u30 = jy1 * m[4] + jy2 * m[8];
    const float u31 = jy1 * m[5] + jy2 * m[9];
    const float u32 = jy1 * m[6] + jy2 * m[10];
    const float ax0 = t30 * m00 + t31 * m10 + t32 * m20;
    const float ax1 = t30 * m01 + t31 * m11 + t32 * m21;
    const float ax2 = t30 * m02 + t31 * m12 + t32 * m22;
    const float ay0 = u30 * m00 + u31 * m10 + u32 * m20;
    const float ay1 = u30 * m01 + u31 * m11 + u32 * m21;
    const float ay2 = u30 * m02 + u31 * m12 + u32 * m22;
    float a = ax0 * ax0 + ax1 * ax1 + ax2 * ax2;
    const float b = ax0 * ay0 + ax1 * ay1 + ax2 * ay2;
    float c = ay0 * ay0 + ay1 * ay1 + ay2 * ay2;
    const float det_before = a * c - b * b;
    a = a + 0.3f;
    c = c + 0.3f;
    const float det = a * c - b * b;
    if (!(det > 1e-12f)) {
        *key = CULLED_KEY | CULL_DEGENERATE;
        if (flags_out) *flags_out = flags;
        return CULL_DEGENERATE;
    }
    const float filtered = sqrtf(maxf_(det_before, 0.0f) / det);
    const float paid = v->antialias ? filtered : 1.0f;
    const float alpha = paid_opacity(opacity, paid, &flags);
    if (!(alpha >= 1.0f / 255.0f)) {
        *key = CULLED_KEY | CULL_FAINT;
        if (flags_out) *flags_out = flags;
        return CULL_FAINT;
    }
    const float bdet = a * c - b * b;
    const float ca = c / bdet, cb = -b / bdet, cc = a / bdet;
    const float cutoff = cutoff_power(alpha);
    uint32_t rect[4];
    if (!tile_rect(v, cx, cy, ca, cb, cc, cutoff, rect)) {
        *key = CULLED_KEY | CULL_OFFSCREEN;
        if (flags_out) *flags_out = flags;
        return CULL_OFFSCREEN;
    }
    uint32_t count = 0;
    for (uint32_t ty_ = rect[1]; ty_ <= rect[3]; ++ty_) {
        for (uint32_t tx_ = rect[0]; tx_ <= rect[2]; ++tx_) {
            count += (uint32_t)tile_touched(v, cx, cy, ca, cb, cc, cutoff, tx_, ty_);
        }
    }
    if (count == 0) {
        *key = CULLED_KEY | CULL_NO_TILE;
        if (flags_out) *flags_out = flags;
        return CULL_NO_TILE;
    }
    /* Colour from this eye, in the cloud's space (splat_project.slang:781-785). */
    const float* sh = g->row;
    const uint32_t degree = degree_of(g->coeffs) < v->sh_limit ? degree_of(g->coeffs) : v->sh_limit;
    const float dx = px - v->eye[0], dy = py - v->eye[1], dz = pz - v->eye[2];
    const float dinv = 1.0f / sqrtf(dx * dx + dy * dy + dz * dz);
    const float ex = dx * dinv, ey = dy * dinv, ez = dz * dinv;
    float col[3];
    for (uint32_t ch = 0; ch < 3; ++ch) {
        const float cval = g->base[ch] + sh_rest(sh, ch, degree, ex, ey, ez);
        if (v->linear) {
            if (cval > 0.0f) flags |= (FL_COL_R << ch);
            col[ch] = maxf_(cval, 0.0f);
        } else {
            uint32_t took = 0;
            col[ch] = srgb_to_linear(cval, &took);
            if (cval > 0.0f) flags |= (FL_COL_R << ch);
            if (took) flags |= (FL_SRGB_R << ch);
        }
    }
    if (rl) {
        /* The relit branch (splat_project.slang:799-976): the evaluated
         * colour is the surface's albedo unless one is given. */
        const float R[3][3] = {{r00, r01, r02}, {r10, r11, r12}, {r20, r21, r22}};
        const float p[3] = {px, py, pz};
        float lit[3];
        relit_colour(rl, i, R, sx, sy, sz, p, v->eye, col, lit);
        col[0] = lit[0]; col[1] = lit[1]; col[2] = lit[2];
        flags |= FL_RELIT;
    }
    float* r = rec + (size_t)i * REC;
    r[0] = cx; r[1] = cy; r[2] = tz; r[3] = cutoff;
    r[4] = ca; r[5] = cb; r[6] = cc; r[7] = alpha;
    r[8] = col[0]; r[9] = col[1]; r[10] = col[2];
    /* The branches taken, as uint bits in the record's last word (the kernels
     * keep them there for the adjoint; same FL_* layout). */
    memcpy(&r[11], &flags, sizeof flags);
    *touched = count;
    uint32_t zb;
    memcpy(&zb, &tz, 4);
    *key = zb >> 7;
    if (flags_out) *flags_out = flags;
    return 0;
}

/* The whole cloud projected into slots [0, n). flags may be NULL; rl NULL:
 * the harmonics' colour, otherwise the relit one. */
ORC_EXPORT void orc_project_lit(const orc_view* v, const float* means, const float* log_scales, const float* quats,
                                const float* logits, const float* sh, uint32_t coeffs, uint32_t n, float* rec,
                                uint32_t* touched, uint32_t* keys, uint32_t* flags, const orc_relight* rl) {
#pragma omp parallel for schedule(dynamic, 4096)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        orc_gaussian g;
        decode_soa(means, log_scales, quats, logits, sh, coeffs, (uint32_t)i, &g);
        project_one(v, &g, (uint32_t)i, rec, touched + i, keys + i, flags ? flags + i : NULL, rl);
    }
}

/* The same from athenea's packed layout. */
ORC_EXPORT void orc_project_packed(const orc_view* v, const float* positions, const uint32_t* shape,
                                   const uint32_t* words, uint32_t rest, uint32_t n, float* rec, uint32_t* touched,
                                   uint32_t* keys, uint32_t* flags, const orc_relight* rl) {
#pragma omp parallel for schedule(dynamic, 4096)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        orc_gaussian g;
        decode_packed(positions, shape, words, rest, (uint32_t)i, &g);
        project_one(v, &g, (uint32_t)i, rec, touched + i, keys + i, flags ? flags + i : NULL, rl);
    }
}

/* packing.slang:129-135 (unpackNormal), for the tests. */
ORC_EXPORT void orc_unpack_normal(uint32_t word, float out[3]) {
    const float fx = (float)(word & 0xFFFFu) / 65535.0f * 2.0f - 1.0f;
    const float fy = (float)(word >> 16) / 65535.0f * 2.0f - 1.0f;
    float nx = fx, ny = fy;
    const float nz = 1.0f - fabsf(fx) - fabsf(fy);
    const float t = satf_(-nz);
    nx = nx + (nx >= 0.0f ? -t : t);
    ny = ny + (ny >= 0.0f ? -t : t);
    const float inv = 1.0f / sqrtf(nx * nx + ny * ny + nz * nz);
    out[0] = nx * inv; out[1] = ny * inv; out[2] = nz * inv;
}

ORC_EXPORT void orc_decode_quaternion(uint32_t word, float q[4]) { decode_quaternion(word, q); }

ORC_EXPORT void orc_project(const orc_view* v, const float* means, const float* log_scales, const float* quats,
                            const float* logits, const float* sh, uint32_t coeffs, uint32_t n, float* rec,
                            uint32_t* touched, uint32_t* keys, uint32_t* flags) {
    orc_project_lit(v, means, log_scales, quats, logits, sh, coeffs, n, rec, touched, keys, flags, NULL);
}

/* The relit colour of one surface point alone, for the tests: n the
 * world-space shading normal, wo the unit direction to the eye. */
ORC_EXPORT void orc_relit_point(const orc_relight* rl, const float n[3], const float wo[3], const float albedo[3],
                                float metallic, float roughness, float out[3]) {
    const float ndv = n[0] * wo[0] + n[1] * wo[1] + n[2] * wo[2];
    const float nv = maxf_(ndv, 1.0e-4f);
    const float i_dot = n[0] * -wo[0] + n[1] * -wo[1] + n[2] * -wo[2];
    const float mirror[3] = {-wo[0] - 2.0f * i_dot * n[0], -wo[1] - 2.0f * i_dot * n[1], -wo[2] - 2.0f * i_dot * n[2]};
    const float over_pi = 1.0f / 3.14159265358979f;
    float above[3];
    env_irradiance(rl, n, above);
    for (int k = 0; k < 3; ++k) above[k] = above[k] * over_pi;
    if (rl->sky_sun[3] > 0.0f) {
        const float cosine = n[0] * rl->sky_sun[0] + n[1] * rl->sky_sun[1] + n[2] * rl->sky_sun[2];
        for (int k = 0; k < 3; ++k) above[k] = above[k] + rl->sky_sun[4 + k] * (maxf_(cosine, 0.0f) * over_pi);
    }
    float reflected[3];
    env_radiance(rl, mirror, roughness, reflected);
    const float diel = ggx_env_dielectric(roughness, nv);
    for (int k = 0; k < 3; ++k) {
        const float body = albedo[k] * (1.0f - metallic);
        const float polish = lerpf_(diel, ggx_env_conductor(albedo[k], roughness, nv), satf_(metallic));
        out[k] = body * above[k] + polish * reflected[k];
    }
}

/* Pieces of the prepared-sky reader and the BRDF fit, for the tests. */
ORC_EXPORT void orc_env_radiance(const orc_relight* rl, const float dir[3], float roughness, float out[3]) {
    env_radiance(rl, dir, roughness, out);
}
ORC_EXPORT void orc_env_irradiance(const orc_relight* rl, const float n[3], float out[3]) {
    env_irradiance(rl, n, out);
}
ORC_EXPORT uint32_t orc_env_level_offset(uint32_t level, uint32_t base_side) {
    return env_level_offset(level, base_side);
}
ORC_EXPORT float orc_half_to_float(uint32_t h) { return half_to_float(h); }
ORC_EXPORT void orc_ggx_albedo(float nv, float alpha, float* a, float* b) { ggx_albedo_ab(nv, alpha, a, b); }
ORC_EXPORT float orc_ggx_dielectric(float roughness, float nv) { return ggx_env_dielectric(roughness, nv); }
ORC_EXPORT float orc_ggx_conductor(float albedo, float roughness, float nv) {
    return ggx_env_conductor(albedo, roughness, nv);
}
ORC_EXPORT float orc_fresnel_dielectric(float c, float ior) { return fresnel_dielectric(c, ior); }
/* The conductor's Fresnel alone: mxArtisticIor with a white edge, then
 * fresnelConductor (microfacet.slang:117-143), as ggx_env_conductor takes it. */
ORC_EXPORT float orc_fresnel_conductor(float albedo, float nv) {
    const float r = clampf_(satf_(albedo), 0.0f, 0.99f);
    const float root = sqrtf(r);
    const float low = (1.0f - r) / (1.0f + r);
    const float high = (1.0f + root) / (1.0f - root);
    const float n = lerpf_(high, low, 1.0f);
    const float up = n + 1.0f, down = n - 1.0f;
    const float k = sqrtf(maxf_((up * up * r - down * down) / maxf_(1.0f - r, 1.0e-6f), 0.0f));
    const float ct = satf_(nv);
    const float ct2 = ct * ct, st2 = 1.0f - ct2;
    const float n2 = n * n, k2 = k * k;
    const float t0 = n2 - k2 - st2;
    const float a2b2 = sqrtf(t0 * t0 + 4.0f * n2 * k2);
    const float t1 = a2b2 + ct2;
    const float aa = sqrtf(maxf_(0.5f * (a2b2 + t0), 0.0f));
    const float t2 = 2.0f * aa * nv;
    const float rs = (t1 - t2) / (t1 + t2);
    const float t3 = ct2 * a2b2 + st2 * st2;
    const float t4 = t2 * st2;
    const float rp = rs * (t3 - t4) / (t3 + t4);
    return 0.5f * (rp + rs);
}

/* Exclusive scan; returns the total (PrefixSum.cpp:58-88 semantics). */
ORC_EXPORT uint64_t orc_exclusive_scan(const uint32_t* in, uint32_t* out, uint64_t n) {
    uint64_t sum = 0;
    for (uint64_t i = 0; i < n; ++i) {
        const uint32_t x = in[i];
        out[i] = (uint32_t)sum;
        sum += x;
    }
    return sum;
}

/* One pair per touched tile, in slot order, each slot's tiles row by row:
 * key = (tile_offset + tile) << 24 | depth24, value = slot. The tile test is
 * splat_emit.slang:37-49's. offsets = exclusive scan of touched. */
ORC_EXPORT void orc_emit_pairs(const orc_view* v, const float* rec, const uint32_t* touched, const uint32_t* keys,
                               const uint32_t* offsets, uint32_t n, uint32_t tile_offset, uint32_t slot_base,
                               uint64_t* pair_keys, uint32_t* pair_vals) {
    const uint32_t tiles_x = (v->width + v->tile - 1) / v->tile;
#pragma omp parallel for schedule(dynamic, 4096)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        if (touched[i] == 0) continue;
        const float* r = rec + (size_t)i * REC;
        uint32_t rect[4];
        tile_rect(v, r[0], r[1], r[4], r[5], r[6], r[3], rect);
        uint64_t at = offsets[i];
        const uint64_t stop = at + touched[i];
        for (uint32_t ty = rect[1]; ty <= rect[3] && at < stop; ++ty) {
            for (uint32_t tx = rect[0]; tx <= rect[2] && at < stop; ++tx) {
                if (tile_touched(v, r[0], r[1], r[4], r[5], r[6], r[3], tx, ty)) {
                    const uint64_t tile = (uint64_t)tile_offset + (uint64_t)ty * tiles_x + tx;
                    pair_keys[at] = (tile << 24) | (uint64_t)(keys[i] & 0xFFFFFFu);
                    pair_vals[at] = slot_base + (uint32_t)i;
                    ++at;
                }
            }
        }
    }
}

/* Stable LSD radix sort of (key, value) over the low `bits` bits, 8 a pass. */
ORC_EXPORT void orc_sort_pairs(uint64_t* keys, uint32_t* vals, uint64_t n, uint32_t bits) {
    if (n < 2) return;
    uint64_t* k2 = (uint64_t*)malloc(n * sizeof(uint64_t));
    uint32_t* v2 = (uint32_t*)malloc(n * sizeof(uint32_t));
    uint64_t* ka = keys;
    uint32_t* va = vals;
    uint64_t* kb = k2;
    uint32_t* vb = v2;
    for (uint32_t shift = 0; shift < bits; shift += 8) {
        uint64_t count[257];
        memset(count, 0, sizeof count);
        for (uint64_t i = 0; i < n; ++i) count[((ka[i] >> shift) & 0xFF) + 1]++;
        for (int d = 0; d < 256; ++d) count[d + 1] += count[d];
        for (uint64_t i = 0; i < n; ++i) {
            const uint64_t at = count[(ka[i] >> shift) & 0xFF]++;
            kb[at] = ka[i];
            vb[at] = va[i];
        }
        uint64_t* tk = ka; ka = kb; kb = tk;
        uint32_t* tv = va; va = vb; vb = tv;
    }
    if (ka != keys) {
        memcpy(keys, ka, n * sizeof(uint64_t));
        memcpy(vals, va, n * sizeof(uint32_t));
    }
    free(k2);
    free(v2);
}

/* ranges[2t] = first pair of tile t, ranges[2t+1] = one past its last; 0, 0 for
 * a tile with none (splat_tiles_clear.slang:8-17, splat_ranges.slang:11-25). */
ORC_EXPORT void orc_ranges(const uint64_t* sorted_keys, uint64_t n, uint32_t tiles, uint32_t* ranges) {
    memset(ranges, 0, (size_t)tiles * 2 * sizeof(uint32_t));
    for (uint64_t i = 0; i < n; ++i) {
        const uint64_t t = sorted_keys[i] >> 24;
        if (t >= tiles) continue;
        if (i == 0 || (sorted_keys[i - 1] >> 24) != t) ranges[t * 2] = (uint32_t)i;
        if (i + 1 == n || (sorted_keys[i + 1] >> 24) != t) ranges[t * 2 + 1] = (uint32_t)(i + 1);
    }
}

/* One pixel's walk, FP32 (splat_blend.slang:287-512 with the plain branches). */
static void blend_pixel_f32(const orc_view* v, const float* rec, const uint32_t* vals, uint32_t start, uint32_t end,
                            uint32_t px, uint32_t py, float out[4], float* depth, float* final_t, uint32_t* last) {
    const float pxc = (float)px + 0.5f, pyc = (float)py + 0.5f;
    float rgb0 = 0.0f, rgb1 = 0.0f, rgb2 = 0.0f;
    float T = 1.0f, depth_sum = 0.0f, depth_at = -1.0f;
    uint32_t stop = 0; /* one past the last pair that contributed, from the tile's first */
    for (uint32_t j = start; j < end; ++j) {
        const float* r = rec + (size_t)vals[j] * REC;
        const float dx = pxc - r[0], dy = pyc - r[1];
        const float power = r[4] * dx * dx + 2.0f * r[5] * dx * dy + r[6] * dy * dy;
        if (power > r[3]) continue;
        const float alpha = minf_(0.99f, r[7] * athx_expf(-0.5f * power));
        if (alpha < 1.0f / 255.0f) continue;
        const float w = T * alpha;
        rgb0 = rgb0 + w * r[8];
        rgb1 = rgb1 + w * r[9];
        rgb2 = rgb2 + w * r[10];
        depth_sum = depth_sum + w * r[2];
        T = T * (1.0f - alpha);
        if (depth_at < 0.0f && 1.0f - T >= v->depth_threshold) depth_at = r[2];
        stop = j - start + 1;
        if (T < 1e-4f) break;
    }
    const float coverage = 1.0f - T;
    out[0] = rgb0 + T * v->bg[0];
    out[1] = rgb1 + T * v->bg[1];
    out[2] = rgb2 + T * v->bg[2];
    out[3] = coverage + T * v->bg[3];
    if (v->depth_mode == 0) {
        *depth = coverage > 1e-6f ? depth_sum / coverage : 0.0f;
    } else {
        *depth = maxf_(depth_at, 0.0f);
    }
    *final_t = T;
    *last = stop;
}

ORC_EXPORT void orc_blend(const orc_view* v, const float* rec, const uint32_t* ranges, const uint32_t* vals,
                          uint32_t tile_offset, float* colour, float* depth, float* final_t, uint32_t* last) {
    const uint32_t W = v->width, H = v->height, T = v->tile;
    const uint32_t tiles_x = (W + T - 1) / T;
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t p = 0; p < (int64_t)W * H; ++p) {
        const uint32_t px = (uint32_t)(p % W), py = (uint32_t)(p / W);
        const uint32_t tile = tile_offset + (py / T) * tiles_x + px / T;
        blend_pixel_f32(v, rec, vals, ranges[tile * 2], ranges[tile * 2 + 1], px, py, colour + p * 4, depth + p,
                        final_t + p, last + p);
    }
}

/* ---------------------------------------------------------------------------
 * The walk's variants (splat_blend.slang:236-578 with composite, matte and
 * traced): an opaque layer under the splats that ends the walk where a splat
 * stands at or behind it and takes what light is left (:287-300, 330-333,
 * 512-566); the Cryptomatte list of the six heaviest ids (:88-143, 421-423);
 * the trace of the four heaviest slots with what the list left out and what
 * closed the walk (splat_trace.slang). Plain splats only: no discs, catchers,
 * media or slopes.
 */
#define ORC_CRYPTO_SLOTS 6
#define ORC_TRACE_SLOTS 4
#define ORC_TRACE_WORDS (ORC_TRACE_SLOTS * 2 + ORC_TRACE_SLOTS / 2 + 1)
#define ORC_TRACE_NONE 0xFFFFFFFFu
#define ORC_TRACE_EMPTY 0xFFFFFFFEu

typedef struct {
    uint32_t ids[ORC_CRYPTO_SLOTS];
    float weights[ORC_CRYPTO_SLOTS];
} crypto_list;

/* splat_blend.slang:93-116 */
static void crypto_add(crypto_list* l, uint32_t id, float weight) {
    if (weight <= 0.0f || id == 0) return;
    uint32_t lightest = 0;
    for (uint32_t k = 0; k < ORC_CRYPTO_SLOTS; ++k) {
        if (l->ids[k] == id && l->weights[k] > 0.0f) {
            l->weights[k] = l->weights[k] + weight;
            return;
        }
        if (l->weights[k] < l->weights[lightest]) lightest = k;
    }
    if (weight > l->weights[lightest]) {
        l->ids[lightest] = id;
        l->weights[lightest] = weight;
    }
}

/* Stable insertion order by descending weight (splat_blend.slang:119-131,
 * splat_trace.slang:102-113). */
static void order_desc(const float* w, uint32_t n, uint32_t* order) {
    for (uint32_t k = 0; k < n; ++k) order[k] = k;
    for (uint32_t i = 1; i < n; ++i) {
        const uint32_t pick = order[i];
        uint32_t j = i;
        while (j > 0 && w[order[j - 1]] < w[pick]) {
            order[j] = order[j - 1];
            --j;
        }
        order[j] = pick;
    }
}

static float bits_float(uint32_t u) {
    float f;
    memcpy(&f, &u, sizeof f);
    return f;
}

static uint32_t float_bits(float f) {
    uint32_t u;
    memcpy(&u, &f, sizeof u);
    return u;
}

static uint32_t ingest_half(float v);

static void blend_pixel_ex(const orc_view* v, const float* rec, const uint32_t* vals, uint32_t start, uint32_t end,
                           uint32_t px, uint32_t py, const float* under_colour, const float* under_depth,
                           const uint32_t* slot_ids, const uint32_t* under_crypto, int matte, int traced,
                           float out[4], float* depth, float* final_t, uint32_t* last, float* crypto,
                           uint32_t* trace) {
    const uint32_t W = v->width, H = v->height, pixels = W * H, at = py * W + px;
    const float pxc = (float)px + 0.5f, pyc = (float)py + 0.5f;
    const int under = under_colour != NULL && under_depth != NULL;
    const float under_z = under ? under_depth[at] : 0.0f;
    const int under_hit = under && under_z > 0.0f;
    float rgb0 = 0.0f, rgb1 = 0.0f, rgb2 = 0.0f;
    float T = 1.0f, depth_sum = 0.0f, depth_at = -1.0f;
    uint32_t stop = 0;
    crypto_list list;
    memset(&list, 0, sizeof list);
    uint32_t tslots[ORC_TRACE_SLOTS];
    float tweights[ORC_TRACE_SLOTS];
    for (uint32_t k = 0; k < ORC_TRACE_SLOTS; ++k) {
        tslots[k] = ORC_TRACE_NONE;
        tweights[k] = 0.0f;
    }
    float ttotal = 0.0f, tfloor = 0.0f;
    uint32_t tlightest = 0;
    for (uint32_t j = start; j < end; ++j) {
        const uint32_t slot = vals[j];
        const float* r = rec + (size_t)slot * REC;
        if (under_hit && r[2] >= under_z) break; /* the layer under the splats ends the walk */
        const float dx = pxc - r[0], dy = pyc - r[1];
        const float power = r[4] * dx * dx + 2.0f * r[5] * dx * dy + r[6] * dy * dy;
        if (power > r[3]) continue;
        const float alpha = minf_(0.99f, r[7] * athx_expf(-0.5f * power));
        if (alpha < 1.0f / 255.0f) continue;
        const float w = T * alpha;
        if (matte) crypto_add(&list, slot_ids ? slot_ids[slot] : 0u, w);
        if (traced) {
            ttotal = ttotal + w;
            if (w > tfloor) {
                /* TraceList::enter (splat_trace.slang:81-90) */
                tslots[tlightest] = slot;
                tweights[tlightest] = w;
                tlightest = 0;
                for (uint32_t k = 1; k < ORC_TRACE_SLOTS; ++k) {
                    if (tweights[k] < tweights[tlightest]) tlightest = k;
                }
                tfloor = tweights[tlightest];
            }
        }
        rgb0 = rgb0 + w * r[8];
        rgb1 = rgb1 + w * r[9];
        rgb2 = rgb2 + w * r[10];
        depth_sum = depth_sum + w * r[2];
        T = T * (1.0f - alpha);
        if (depth_at < 0.0f && 1.0f - T >= v->depth_threshold) depth_at = r[2];
        stop = j - start + 1;
        if (T < 1e-4f) break;
    }
    int opaque = 0;
    float behind[4] = {v->bg[0], v->bg[1], v->bg[2], v->bg[3]};
    float opaque_z = 0.0f;
    uint32_t opaque_id = 0;
    if (under_hit) {
        opaque = 1;
        for (int c = 0; c < 4; ++c) behind[c] = under_colour[(size_t)at * 4 + c];
        opaque_z = under_z;
        if (matte && under_crypto) opaque_id = under_crypto[at];
    }
    if (matte) {
        crypto_add(&list, opaque_id, opaque ? T : 0.0f);
        uint32_t order[ORC_CRYPTO_SLOTS];
        order_desc(list.weights, ORC_CRYPTO_SLOTS, order);
        for (uint32_t layer = 0; layer < ORC_CRYPTO_SLOTS / 2; ++layer) {
            const uint32_t a = order[layer * 2], b = order[layer * 2 + 1];
            float* o = crypto + ((size_t)layer * pixels + at) * 4;
            o[0] = list.weights[a] > 0.0f ? bits_float(list.ids[a]) : 0.0f;
            o[1] = maxf_(list.weights[a], 0.0f);
            o[2] = list.weights[b] > 0.0f ? bits_float(list.ids[b]) : 0.0f;
            o[3] = maxf_(list.weights[b], 0.0f);
        }
    }
    if (traced) {
        const float under_w = opaque ? T : 0.0f;
        if (!(ttotal > 0.0f) && !(under_w > 0.0f)) {
            trace[at] = ORC_TRACE_EMPTY;
        } else {
            uint32_t order[ORC_TRACE_SLOTS];
            order_desc(tweights, ORC_TRACE_SLOTS, order);
            float held = 0.0f;
            for (uint32_t k = 0; k < ORC_TRACE_SLOTS; ++k) {
                const uint32_t from = order[k];
                const int some = tweights[from] > 0.0f;
                trace[(size_t)k * pixels + at] = some ? tslots[from] : ORC_TRACE_NONE;
                trace[(size_t)(ORC_TRACE_SLOTS + k) * pixels + at] =
                    some ? float_bits(rec[(size_t)tslots[from] * REC + 2]) : 0u;
                held = held + tweights[from];
            }
            for (uint32_t pr = 0; pr < ORC_TRACE_SLOTS / 2; ++pr) {
                trace[(size_t)(ORC_TRACE_SLOTS * 2 + pr) * pixels + at] =
                    ingest_half(tweights[order[pr * 2]]) | (ingest_half(tweights[order[pr * 2 + 1]]) << 16);
            }
            trace[(size_t)(ORC_TRACE_WORDS - 1) * pixels + at] =
                ingest_half(maxf_(ttotal - held, 0.0f)) | (ingest_half(maxf_(under_w, 0.0f)) << 16);
        }
    }
    const float coverage = 1.0f - T;
    if (opaque) {
        depth_sum = depth_sum + T * opaque_z;
        if (depth_at < 0.0f) depth_at = opaque_z;
    }
    const float total = opaque ? 1.0f : coverage;
    out[0] = rgb0 + T * behind[0];
    out[1] = rgb1 + T * behind[1];
    out[2] = rgb2 + T * behind[2];
    out[3] = coverage + T * behind[3];
    if (v->depth_mode == 0) {
        *depth = total > 1e-6f ? depth_sum / total : 0.0f;
    } else {
        *depth = maxf_(depth_at, 0.0f);
    }
    *final_t = T;
    *last = stop;
}

/* orc_blend with the under layer (under_colour [H][W][4] and under_depth
 * [H][W], or NULL), the Cryptomatte (slot_ids a slot, under_crypto a pixel
 * or NULL; crypto [3][H][W][4]) and the trace ([11][H][W] words), each off
 * when its output is NULL. */
ORC_EXPORT void orc_blend_ex(const orc_view* v, const float* rec, const uint32_t* ranges, const uint32_t* vals,
                             uint32_t tile_offset, const float* under_colour, const float* under_depth,
                             const uint32_t* slot_ids, const uint32_t* under_crypto, float* colour, float* depth,
                             float* final_t, uint32_t* last, float* crypto, uint32_t* trace) {
    const uint32_t W = v->width, H = v->height, T = v->tile;
    const uint32_t tiles_x = (W + T - 1) / T;
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t p = 0; p < (int64_t)W * H; ++p) {
        const uint32_t px = (uint32_t)(p % W), py = (uint32_t)(p / W);
        const uint32_t tile = tile_offset + (py / T) * tiles_x + px / T;
        blend_pixel_ex(v, rec, vals, ranges[tile * 2], ranges[tile * 2 + 1], px, py, under_colour, under_depth,
                       slot_ids, under_crypto, crypto != NULL, trace != NULL, colour + p * 4, depth + p, final_t + p,
                       last + p, crypto, trace);
    }
}

/* ---------------------------------------------------------------------------
 * FP64: the same projection with the FP32 path's branches replayed.
 * rec64 layout is REC doubles a slot; only slots the FP32 path kept are written.
 */
typedef struct {
    double cx, cy, z, a, b, c, alpha, col[3];
} rec64_t;

static int project_one_f64(const orc_view* v, const double* P, const double* LS, const double* Q, const double* LG,
                           const double* SH, uint32_t coeffs, uint32_t i, uint32_t flags, rec64_t* out) {
    const double px = P[i * 3 + 0], py = P[i * 3 + 1], pz = P[i * 3 + 2];
    const double sx = exp(LS[i * 3 + 0]), sy = exp(LS[i * 3 + 1]), sz = exp(LS[i * 3 + 2]);
    const double opacity = 1.0 / (1.0 + exp(-LG[i]));
    double m[12];
    for (int k = 0; k < 12; ++k) m[k] = (double)v->m[k];
    const double tx = m[0] * px + m[1] * py + m[2] * pz + m[3];
    const double ty = m[4] * px + m[5] * py + m[6] * pz + m[7];
    const double tz = m[8] * px + m[9] * py + m[10] * pz + m[11];
    double qx = Q[i * 4 + 0], qy = Q[i * 4 + 1], qz = Q[i * 4 + 2], qw = Q[i * 4 + 3];
    const double qn = qx * qx + qy * qy + qz * qz + qw * qw;
    if (qn > 0.0) {
        const double inv = 1.0 / sqrt(qn);
        qx *= inv; qy *= inv; qz *= inv; qw *= inv;
    } else {
        qx = 0; qy = 0; qz = 0; qw = 1;
    }
    const double r00 = 1 - 2 * (qy * qy + qz * qz), r01 = 2 * (qx * qy - qw * qz), r02 = 2 * (qx * qz + qw * qy);
    const double r10 = 2 * (qx * qy + qw * qz), r11 = 1 - 2 * (qx * qx + qz * qz), r12 = 2 * (qy * qz - qw * qx);
    const double r20 = 2 * (qx * qz - qw * qy), r21 = 2 * (qy * qz + qw * qx), r22 = 1 - 2 * (qx * qx + qy * qy);
    const double m00 = r00 * sx, m01 = r01 * sy, m02 = r02 * sz;
    const double m10 = r10 * sx, m11 = r11 * sy, m12 = r12 * sz;
    const double m20 = r20 * sx, m21 = r21 * sy, m22 = r22 * sz;
    const double fx = v->focal_x, fy = v->focal_y, ccx = v->centre_x, ccy = v->centre_y;
    double jx0, jx2, jy1, jy2, cx, cy;
    if (v->orthographic) {
        jx0 = fx; jx2 = 0; jy1 = fy; jy2 = 0;
        cx = fx * tx + ccx;
        cy = fy * ty + ccy;
    } else {
        const double invz = 1.0 / tz;
        cx = fx * tx * invz + ccx;
        cy = fy * ty * invz + ccy;
        const double W = v->width, H = v->height;
        /* The limits are the FP32 frame's constants, as the kernel has them. */
        const double limlox = (double)((-0.15f * (float)W - v->centre_x) / v->focal_x);
        const double limhix = (double)((1.15f * (float)W - v->centre_x) / v->focal_x);
        const double limloy = (double)((-0.15f * (float)H - v->centre_y) / v->focal_y);
        const double limhiy = (double)((1.15f * (float)H - v->centre_y) / v->focal_y);
        double ux = tx * invz, uy = ty * invz;
        if (flags & FL_CLAMP_XLO) ux = limlox;
        else if (flags & FL_CLAMP_XHI) ux = limhix;
        if (flags & FL_CLAMP_YLO) uy = limloy;
        else if (flags & FL_CLAMP_YHI) uy = limhiy;
        const double txc = ux * tz, tyc = uy * tz;
        jx0 = fx * invz;
        jx2 = -fx * txc * invz * invz;
        jy1 = fy * invz;
        jy2 = -fy * tyc * invz * invz;
    }
    const double t30 = jx0 * m[0] + jx2 * m[8], t31 = jx0 * m[1] + jx2 * m[9], t32 = jx0 * m[2] + jx2 * m[10];
    const double u30 = jy1 * m[4] + jy2 * m[8], u31 = jy1 * m[5] + jy2 * m[9], u32 = jy1 * m[6] + jy2 * m[10];
    const double ax0 = t30 * m00 + t31 * m10 + t32 * m20;
    const double ax1 = t30 * m01 + t31 * m11 + t32 * m21;
    const double ax2 = t30 * m02 + t31 * m12 + t32 * m22;
    const double ay0 = u30 * m00 + u31 * m10 + u32 * m20;
    const double ay1 = u30 * m01 + u31 * m11 + u32 * m21;
    const double ay2 = u30 * m02 + u31 * m12 + u32 * m22;
    double a = ax0 * ax0 + ax1 * ax1 + ax2 * ax2;
    const double b = ax0 * ay0 + ax1 * ay1 + ax2 * ay2;
    double c = ay0 * ay0 + ay1 * ay1 + ay2 * ay2;
    const double det_before = a * c - b * b;
    a += 0.3;
    c += 0.3;
    const double det = a * c - b * b;
    const double filtered = sqrt((det_before > 0 ? det_before : 0.0) / det);
    const double paid = v->antialias ? filtered : 1.0;
    const double spread = paid * opacity;
    const double alpha = (flags & FL_PAID_BACK) ? spread + (1.0 - paid) / 255.0 : spread;
    out->cx = cx; out->cy = cy; out->z = tz;
    out->a = c / det; out->b = -b / det; out->c = a / det;
    out->alpha = alpha;
    const double* sh = SH + (size_t)i * coeffs * 3;
    const uint32_t dg = degree_of(coeffs) < v->sh_limit ? degree_of(coeffs) : v->sh_limit;
    const double dx = px - v->eye[0], dy = py - v->eye[1], dz = pz - v->eye[2];
    const double dl = sqrt(dx * dx + dy * dy + dz * dz);
    const double x = dx / dl, y = dy / dl, z = dz / dl;
    for (uint32_t ch = 0; ch < 3; ++ch) {
#define B(k) sh[(k) * 3 + ch]
        double cv = 0.5 + 0.28209479177387814 * B(0);
        if (dg >= 1) cv += -0.4886025119029199 * y * B(1) + 0.4886025119029199 * z * B(2) - 0.4886025119029199 * x * B(3);
        if (dg >= 2) {
            const double xx = x * x, yy = y * y, zz = z * z;
            cv += 1.0925484305920792 * x * y * B(4) - 1.0925484305920792 * y * z * B(5) +
                  0.31539156525252005 * (2 * zz - xx - yy) * B(6) - 1.0925484305920792 * x * z * B(7) +
                  0.5462742152960396 * (xx - yy) * B(8);
        }
        if (dg >= 3) {
            const double xx = x * x, yy = y * y, zz = z * z;
            cv += -0.5900435899266435 * y * (3 * xx - yy) * B(9) + 2.890611442640554 * x * y * z * B(10) -
                  0.4570457994644658 * y * (4 * zz - xx - yy) * B(11) +
                  0.3731763325901154 * z * (2 * zz - 3 * xx - 3 * yy) * B(12) -
                  0.4570457994644658 * x * (4 * zz - xx - yy) * B(13) + 1.445305721320277 * z * (xx - yy) * B(14) -
                  0.5900435899266435 * x * (xx - 3 * yy) * B(15);
        }
#undef B
        if (v->linear) {
            out->col[ch] = (flags & (FL_COL_R << ch)) ? cv : 0.0;
        } else {
            const double cc = (flags & (FL_COL_R << ch)) ? cv : 0.0;
            out->col[ch] = (flags & (FL_SRGB_R << ch)) ? pow((cc + 0.055) / 1.055, 2.4) : cc / 12.92;
        }
    }
    return 1;
}

/* The FP64 image: the FP32 walk decides (skip, clamp, stop); the sums are FP64.
 * rec/ranges/vals/flags come from the FP32 path; params are FP64 copies. */
ORC_EXPORT void orc_blend_f64(const orc_view* v, const float* rec, const uint32_t* ranges, const uint32_t* vals,
                              const uint32_t* touched, const uint32_t* flags, uint32_t tile_offset,
                              const double* means, const double* log_scales, const double* quats,
                              const double* logits, const double* sh, uint32_t coeffs, uint32_t n, double* colour,
                              double* depth) {
    rec64_t* r64 = (rec64_t*)calloc(n ? n : 1, sizeof(rec64_t));
#pragma omp parallel for schedule(dynamic, 4096)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        if (touched[i]) {
            project_one_f64(v, means, log_scales, quats, logits, sh, coeffs, (uint32_t)i, flags[i], &r64[i]);
            /* A relit colour is the blend's input as the FP32 projection made
             * it: the FP64 replay covers the geometry and the blend. */
            if (flags[i] & FL_RELIT) {
                for (int ch = 0; ch < 3; ++ch) r64[i].col[ch] = (double)rec[(size_t)i * REC + 8 + ch];
            }
        }
    }
    const uint32_t W = v->width, H = v->height, TS = v->tile;
    const uint32_t tiles_x = (W + TS - 1) / TS;
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t p = 0; p < (int64_t)W * H; ++p) {
        const uint32_t px = (uint32_t)(p % W), py = (uint32_t)(p / W);
        const uint32_t tile = tile_offset + (py / TS) * tiles_x + px / TS;
        const uint32_t start = ranges[tile * 2], end = ranges[tile * 2 + 1];
        const float pxc = (float)px + 0.5f, pyc = (float)py + 0.5f;
        float T32 = 1.0f;
        double T = 1.0, rgb[3] = {0, 0, 0}, dsum = 0.0, dat = -1.0;
        for (uint32_t j = start; j < end; ++j) {
            const uint32_t s = vals[j];
            const float* r = rec + (size_t)s * REC;
            const float dx = pxc - r[0], dy = pyc - r[1];
            const float power = r[4] * dx * dx + 2.0f * r[5] * dx * dy + r[6] * dy * dy;
            if (power > r[3]) continue;
            const float raw = r[7] * athx_expf(-0.5f * power);
            const float alpha32 = minf_(0.99f, raw);
            if (alpha32 < 1.0f / 255.0f) continue;
            const rec64_t* q = &r64[s];
            const double ddx = (px + 0.5) - q->cx, ddy = (py + 0.5) - q->cy;
            const double pw = q->a * ddx * ddx + 2.0 * q->b * ddx * ddy + q->c * ddy * ddy;
            const double alpha = raw > 0.99f ? 0.99 : q->alpha * exp(-0.5 * pw);
            const double w = T * alpha;
            rgb[0] += w * q->col[0];
            rgb[1] += w * q->col[1];
            rgb[2] += w * q->col[2];
            dsum += w * q->z;
            T *= 1.0 - alpha;
            T32 = T32 * (1.0f - alpha32);
            if (dat < 0.0 && 1.0f - T32 >= v->depth_threshold) dat = q->z;
            if (T32 < 1e-4f) break;
        }
        const double cov = 1.0 - T;
        colour[p * 4 + 0] = rgb[0] + T * v->bg[0];
        colour[p * 4 + 1] = rgb[1] + T * v->bg[1];
        colour[p * 4 + 2] = rgb[2] + T * v->bg[2];
        colour[p * 4 + 3] = cov + T * v->bg[3];
        if (v->depth_mode == 0) {
            depth[p] = cov > 1e-6 ? dsum / cov : 0.0;
        } else {
            depth[p] = dat > 0.0 ? dat : 0.0;
        }
    }
    free(r64);
}

/* ---------------------------------------------------------------------------
 * The brute force of reference/reference_blend.slang:28-123: every kept slot
 * for every pixel, insertion-sorted by (view z, slot), same skips and stop. No
 * tiles, no keys. overflow counts pixels past max_hits. */
ORC_EXPORT uint64_t orc_reference(const orc_view* v, const float* rec, const uint32_t* touched, uint32_t n,
                                  uint32_t max_hits, float* colour, float* depth) {
    uint64_t overflow = 0;
    const uint32_t W = v->width, H = v->height;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : overflow)
    for (int64_t p = 0; p < (int64_t)W * H; ++p) {
        const uint32_t px = (uint32_t)(p % W), py = (uint32_t)(p / W);
        const float pxc = (float)px + 0.5f, pyc = (float)py + 0.5f;
        float* hz = (float*)malloc(max_hits * sizeof(float));
        uint32_t* hi = (uint32_t*)malloc(max_hits * sizeof(uint32_t));
        float* ha = (float*)malloc(max_hits * sizeof(float));
        uint32_t hits = 0, over = 0;
        for (uint32_t s = 0; s < n; ++s) {
            if (!touched[s]) continue;
            const float* r = rec + (size_t)s * REC;
            const float dx = pxc - r[0], dy = pyc - r[1];
            const float power = r[4] * dx * dx + 2.0f * r[5] * dx * dy + r[6] * dy * dy;
            if (power > r[3]) continue;
            const float alpha = minf_(0.99f, r[7] * athx_expf(-0.5f * power));
            if (alpha < 1.0f / 255.0f) continue;
            if (hits == max_hits) {
                over = 1;
                continue;
            }
            uint32_t k = hits++;
            while (k > 0 && (hz[k - 1] > r[2] || (hz[k - 1] == r[2] && hi[k - 1] > s))) {
                hz[k] = hz[k - 1];
                hi[k] = hi[k - 1];
                ha[k] = ha[k - 1];
                --k;
            }
            hz[k] = r[2];
            hi[k] = s;
            ha[k] = alpha;
        }
        float rgb0 = 0, rgb1 = 0, rgb2 = 0, T = 1.0f, dsum = 0.0f;
        for (uint32_t k = 0; k < hits; ++k) {
            const float* r = rec + (size_t)hi[k] * REC;
            const float w = T * ha[k];
            rgb0 = rgb0 + w * r[8];
            rgb1 = rgb1 + w * r[9];
            rgb2 = rgb2 + w * r[10];
            dsum = dsum + w * r[2];
            T = T * (1.0f - ha[k]);
            if (T < 1e-4f) break;
        }
        const float cov = 1.0f - T;
        colour[p * 4 + 0] = rgb0 + T * v->bg[0];
        colour[p * 4 + 1] = rgb1 + T * v->bg[1];
        colour[p * 4 + 2] = rgb2 + T * v->bg[2];
        colour[p * 4 + 3] = cov + T * v->bg[3];
        depth[p] = cov > 1e-6f ? dsum / cov : 0.0f;
        overflow += over;
        free(hz);
        free(hi);
        free(ha);
    }
    return overflow;
}

/* ---------------------------------------------------------------------------
 * The reference's own order (TileRasterizer.cpp:1084-1176): compact visible
 * slots with their 24-bit depth keys, stable-sort them by depth, emit each
 * one's tiles in depth order, stable-sort the pairs by tile. Returns P and
 * writes the pair values in sorted order plus the ranges: what the 64-bit
 * key route must reproduce exactly. Timed per stage into ms[6]:
 * project, depth sort, counts (gather + scan), emit, tile sort, ranges. */
static double now_ms(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec * 1e3 + ts.tv_nsec * 1e-6;
}

static void radix_u32_pairs(uint32_t* k, uint32_t* v, uint64_t n, uint32_t bits, uint32_t* tk, uint32_t* tv) {
    uint32_t *ka = k, *va = v, *kb = tk, *vb = tv;
    uint32_t passes = 0;
    for (uint32_t shift = 0; shift < bits; shift += 8, ++passes) {
        uint64_t count[257];
        memset(count, 0, sizeof count);
        for (uint64_t i = 0; i < n; ++i) count[((ka[i] >> shift) & 0xFF) + 1]++;
        for (int d = 0; d < 256; ++d) count[d + 1] += count[d];
        for (uint64_t i = 0; i < n; ++i) {
            const uint64_t at = count[(ka[i] >> shift) & 0xFF]++;
            kb[at] = ka[i];
            vb[at] = va[i];
        }
        uint32_t* t = ka; ka = kb; kb = t;
        t = va; va = vb; vb = t;
    }
    if (ka != k) {
        memcpy(k, ka, n * 4);
        memcpy(v, va, n * 4);
    }
}

static uint32_t bits_for(uint32_t values) {
    uint32_t bits = 1;
    while (bits < 32 && ((uint64_t)1 << bits) < values) ++bits;
    return bits;
}

ORC_EXPORT uint64_t orc_render_ref_order(const orc_view* v, const float* means, const float* log_scales,
                                         const float* quats, const float* logits, const float* sh, uint32_t coeffs,
                                         uint32_t n, float* rec, uint32_t* touched, uint32_t* keys,
                                         uint32_t* pair_vals_out, uint64_t pair_cap, uint32_t* ranges, float* colour,
                                         float* depth, float* final_t, uint32_t* last, double* ms,
                                         const orc_relight* rl) {
    const uint32_t TS = v->tile;
    const uint32_t tiles_x = (v->width + TS - 1) / TS, tiles_y = (v->height + TS - 1) / TS;
    const uint32_t tiles = tiles_x * tiles_y;
    double t0 = now_ms();
    orc_project_lit(v, means, log_scales, quats, logits, sh, coeffs, n, rec, touched, keys, NULL, rl);
    double t1 = now_ms();
    /* compact (splat_compact.slang:16-26) + depth sort, 24 bits */
    uint64_t V = 0;
    for (uint32_t i = 0; i < n; ++i) V += touched[i] != 0;
    uint32_t* dk = (uint32_t*)malloc((V ? V : 1) * 4);
    uint32_t* ds = (uint32_t*)malloc((V ? V : 1) * 4);
    uint32_t* tk = (uint32_t*)malloc((V ? V : 1) * 4);
    uint32_t* tv = (uint32_t*)malloc((V ? V : 1) * 4);
    uint64_t at = 0;
    for (uint32_t i = 0; i < n; ++i) {
        if (touched[i]) {
            dk[at] = keys[i];
            ds[at] = i;
            ++at;
        }
    }
    radix_u32_pairs(dk, ds, V, 24, tk, tv);
    double t2 = now_ms();
    /* gather counts in depth order + scan (splat_gather_counts.slang:11-19) */
    uint32_t* off = (uint32_t*)malloc((V ? V : 1) * 4);
    uint64_t P = 0;
    for (uint64_t k = 0; k < V; ++k) {
        off[k] = (uint32_t)P;
        P += touched[ds[k]];
    }
    double t3 = now_ms();
    if (P > pair_cap) {
        free(dk); free(ds); free(tk); free(tv); free(off);
        return P;
    }
    uint32_t* pk = (uint32_t*)malloc((P ? P : 1) * 4);
    uint32_t* pv = pair_vals_out;
    /* emit in depth order (splat_emit.slang:17-50) */
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t k = 0; k < (int64_t)V; ++k) {
        const uint32_t s = ds[k];
        const float* r = rec + (size_t)s * REC;
        uint32_t rect[4];
        tile_rect(v, r[0], r[1], r[4], r[5], r[6], r[3], rect);
        uint64_t w = off[k];
        const uint64_t stop = w + touched[s];
        for (uint32_t ty = rect[1]; ty <= rect[3] && w < stop; ++ty)
            for (uint32_t tx = rect[0]; tx <= rect[2] && w < stop; ++tx)
                if (tile_touched(v, r[0], r[1], r[4], r[5], r[6], r[3], tx, ty)) {
                    pk[w] = ty * tiles_x + tx;
                    pv[w] = s;
                    ++w;
                }
    }
    double t4 = now_ms();
    uint32_t* tk2 = (uint32_t*)malloc((P ? P : 1) * 4);
    uint32_t* tv2 = (uint32_t*)malloc((P ? P : 1) * 4);
    radix_u32_pairs(pk, pv, P, bits_for(tiles) <= 8 ? 8 : (bits_for(tiles) <= 16 ? 16 : 24), tk2, tv2);
    double t5 = now_ms();
    memset(ranges, 0, (size_t)tiles * 2 * 4);
    for (uint64_t i = 0; i < P; ++i) {
        const uint32_t t = pk[i];
        if (i == 0 || pk[i - 1] != t) ranges[t * 2] = (uint32_t)i;
        if (i + 1 == P || pk[i + 1] != t) ranges[t * 2 + 1] = (uint32_t)(i + 1);
    }
    double t6 = now_ms();
    if (colour) orc_blend(v, rec, ranges, pv, 0, colour, depth, final_t, last);
    double t7 = now_ms();
    if (ms) {
        ms[0] = t1 - t0; ms[1] = t2 - t1; ms[2] = t3 - t2; ms[3] = t4 - t3; ms[4] = t5 - t4; ms[5] = t6 - t5;
        ms[6] = t7 - t6;
    }
    free(dk); free(ds); free(tk); free(tv); free(off); free(pk); free(tk2); free(tv2);
    return P;
}

ORC_EXPORT int orc_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

ORC_EXPORT void orc_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

ORC_EXPORT float orc_expf(float x) { return athx_expf(x); }
ORC_EXPORT float orc_logf(float x) { return athx_logf(x); }

/* ---------------------------------------------------------------------------
 * FP64 central finite differences of the adjoint's product, for its check.
 *
 * L(theta) = sum over pixels and channels of G[px][c] * C[px][c](theta): the
 * image paired with a fixed upstream gradient G (dL/dcolour), which is the
 * vector-Jacobian product an adjoint computes. C is the FP64 image of
 * orc_blend_f64 with every discrete decision (culls, the clamp of the
 * Jacobian, the pay-back branch, max(c, 0), the cut, the per-pair skips, the
 * 0.99 clamp and the early stop) taken from the FP32 frame and frozen, so L is
 * a smooth function of the parameters and its central difference is what the
 * adjoint must reproduce. `magnitudes` gets the sum over pixels of each
 * pixel's share's magnitude, |G . dC/dtheta|: the size of the terms an adjoint
 * sums, which bounds the rounding a float32 sum of them can carry. One view; rec, ranges, vals, touched and flags are
 * that view's FP32 frame (slots 0..n-1, tiles from 0).
 */
static void walk_f64(const orc_view* v, const float* rec, const rec64_t* r64, const uint32_t* vals, uint32_t start,
                     uint32_t end, uint32_t px, uint32_t py, uint32_t sub, const rec64_t* sub_rec, double out[4]) {
    const float pxc = (float)px + 0.5f, pyc = (float)py + 0.5f;
    float T32 = 1.0f;
    double T = 1.0, rgb[3] = {0, 0, 0};
    for (uint32_t j = start; j < end; ++j) {
        const uint32_t s = vals[j];
        const float* r = rec + (size_t)s * REC;
        const float dx = pxc - r[0], dy = pyc - r[1];
        const float power = r[4] * dx * dx + 2.0f * r[5] * dx * dy + r[6] * dy * dy;
        if (power > r[3]) continue;
        const float raw = r[7] * athx_expf(-0.5f * power);
        const float alpha32 = minf_(0.99f, raw);
        if (alpha32 < 1.0f / 255.0f) continue;
        const rec64_t* q = s == sub ? sub_rec : &r64[s];
        const double ddx = (px + 0.5) - q->cx, ddy = (py + 0.5) - q->cy;
        const double pw = q->a * ddx * ddx + 2.0 * q->b * ddx * ddy + q->c * ddy * ddy;
        const double alpha = raw > 0.99f ? 0.99 : q->alpha * exp(-0.5 * pw);
        const double w = T * alpha;
        rgb[0] += w * q->col[0];
        rgb[1] += w * q->col[1];
        rgb[2] += w * q->col[2];
        T *= 1.0 - alpha;
        T32 = T32 * (1.0f - alpha32);
        if (T32 < 1e-4f) break;
    }
    out[0] = rgb[0] + T * v->bg[0];
    out[1] = rgb[1] + T * v->bg[1];
    out[2] = rgb[2] + T * v->bg[2];
    out[3] = (1.0 - T) + T * v->bg[3];
}

/* Over the tiles in `tiles`: sum of G . (C+ - C-) with slot `sub` drawn from
 * rp and then rm, and the sum of its per-pixel magnitudes |G . (C+ - C-)|. */
static void tiles_diff(const orc_view* v, const float* rec, const rec64_t* r64, const uint32_t* ranges,
                       const uint32_t* vals, const uint32_t* tiles, uint32_t ntiles, uint32_t sub, const rec64_t* rp,
                       const rec64_t* rm, const double* upstream, double* sum, double* mag) {
    const uint32_t TS = v->tile, W = v->width, H = v->height;
    const uint32_t tiles_x = (W + TS - 1) / TS;
    double acc = 0.0, acc_mag = 0.0;
    for (uint32_t k = 0; k < ntiles; ++k) {
        const uint32_t t = tiles[k];
        const uint32_t tx = t % tiles_x, ty = t / tiles_x;
        for (uint32_t py = ty * TS; py < (ty + 1) * TS && py < H; ++py) {
            for (uint32_t px = tx * TS; px < (tx + 1) * TS && px < W; ++px) {
                double cp[4], cm[4];
                walk_f64(v, rec, r64, vals, ranges[2 * t], ranges[2 * t + 1], px, py, sub, rp, cp);
                walk_f64(v, rec, r64, vals, ranges[2 * t], ranges[2 * t + 1], px, py, sub, rm, cm);
                const double* g = upstream + ((size_t)py * W + px) * 4;
                const double d = g[0] * (cp[0] - cm[0]) + g[1] * (cp[1] - cm[1]) + g[2] * (cp[2] - cm[2]) +
                                 g[3] * (cp[3] - cm[3]);
                acc += d;
                acc_mag += fabs(d);
            }
        }
    }
    *sum = acc;
    *mag = acc_mag;
}

ORC_EXPORT void orc_fd_grads(const orc_view* v, const float* rec, const uint32_t* ranges, const uint32_t* vals,
                             uint32_t pairs, const uint32_t* touched, const uint32_t* flags, const double* means,
                             const double* log_scales, const double* quats, const double* logits, const double* sh,
                             uint32_t coeffs, uint32_t n, const double* upstream, const uint32_t* ids, uint32_t nids,
                             double h_rel, double* grads, double* magnitudes) {
    const uint32_t TS = v->tile;
    const uint32_t tiles = ((v->width + TS - 1) / TS) * ((v->height + TS - 1) / TS);
    const uint32_t per = 11 + coeffs * 3;
    rec64_t* r64 = (rec64_t*)calloc(n ? n : 1, sizeof(rec64_t));
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        if (touched[i]) project_one_f64(v, means, log_scales, quats, logits, sh, coeffs, (uint32_t)i, flags[i], &r64[i]);
    }
    /* Which tile each pair belongs to: one pass over the ranges. */
    uint32_t* tile_of_pair = (uint32_t*)malloc((pairs ? pairs : 1) * sizeof(uint32_t));
    for (uint32_t t = 0; t < tiles; ++t) {
        for (uint32_t j = ranges[2 * t]; j < ranges[2 * t + 1]; ++j) tile_of_pair[j] = t;
    }
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t q = 0; q < (int64_t)nids; ++q) {
        const uint32_t g = ids[q];
        double* out = grads + (size_t)q * per;
        double* mag = magnitudes + (size_t)q * per;
        for (uint32_t k = 0; k < per; ++k) out[k] = mag[k] = 0.0;
        if (g >= n || !touched[g]) continue;
        uint32_t* mine = (uint32_t*)malloc(touched[g] * sizeof(uint32_t));
        uint32_t nt = 0;
        for (uint32_t j = 0; j < pairs && nt < touched[g]; ++j) {
            if (vals[j] == g) mine[nt++] = tile_of_pair[j];
        }
        double P[3], LS[3], Q[4], LG[1];
        double* SH = (double*)malloc(coeffs * 3 * sizeof(double));
        for (uint32_t k = 0; k < per; ++k) {
            double* slot;
            memcpy(P, means + (size_t)g * 3, sizeof P);
            memcpy(LS, log_scales + (size_t)g * 3, sizeof LS);
            memcpy(Q, quats + (size_t)g * 4, sizeof Q);
            LG[0] = logits[g];
            memcpy(SH, sh + (size_t)g * coeffs * 3, coeffs * 3 * sizeof(double));
            if (k < 3) slot = &P[k];
            else if (k < 6) slot = &LS[k - 3];
            else if (k < 10) slot = &Q[k - 6];
            else if (k < 11) slot = &LG[0];
            else slot = &SH[k - 11];
            const double x0 = *slot;
            const double h = h_rel * (fabs(x0) > 1.0 ? fabs(x0) : 1.0);
            rec64_t rp, rm;
            *slot = x0 + h;
            project_one_f64(v, P, LS, Q, LG, SH, coeffs, 0, flags[g], &rp);
            *slot = x0 - h;
            project_one_f64(v, P, LS, Q, LG, SH, coeffs, 0, flags[g], &rm);
            double diff, diff_mag;
            tiles_diff(v, rec, r64, ranges, vals, mine, nt, g, &rp, &rm, upstream, &diff, &diff_mag);
            out[k] = diff / (2.0 * h);
            mag[k] = diff_mag / (2.0 * h);
        }
        free(SH);
        free(mine);
    }
    free(tile_of_pair);
    free(r64);
}


/* ---- orc_ingest: a PLY vertex block into the packed layout ---- */

/* One property of one vertex as a float: the little-endian bytes of its
 * type, a double rounded to the nearest float (packed_records.slang). */
static float ingest_widen(const uint8_t* rec, uint32_t offset, uint32_t type) {
    const uint8_t* b = rec + offset;
    uint32_t w = 0;
    switch (type) {
    case 2: return (float)b[0];
    case 3: return (float)(int8_t)b[0];
    case 4: return (float)(uint16_t)(b[0] | (b[1] << 8));
    case 5: return (float)(int16_t)(uint16_t)(b[0] | (b[1] << 8));
    case 1: {
        uint64_t u = 0;
        for (int k = 7; k >= 0; --k) {
            u = (u << 8) | b[k];
        }
        double d;
        memcpy(&d, &u, sizeof d);
        return (float)d;
    }
    default:
        w = (uint32_t)b[0] | ((uint32_t)b[1] << 8) | ((uint32_t)b[2] << 16) | ((uint32_t)b[3] << 24);
        if (type == 0) {
            float f;
            memcpy(&f, &w, sizeof f);
            return f;
        }
        return type == 6 ? (float)w : (float)(int32_t)w;
    }
}

/* IEEE binary16 bits of f, round to nearest even (f32tof16), after
 * packing.slang's halfSafe clamp to +-65000 (a NaN clamps to -65000, as
 * max(NaN, lo) is lo). */
static uint32_t ingest_half(float v) {
    float c = v > -65000.0f ? v : -65000.0f;
    c = c < 65000.0f ? c : 65000.0f;
    uint32_t x;
    memcpy(&x, &c, sizeof x);
    const uint32_t sign = (x >> 16) & 0x8000u, ax = x & 0x7FFFFFFFu;
    if (ax < 0x38800000u) {
        /* below 2^-14: a subnormal half, round(|c| / 2^-24) */
        const float scaled = fabsf(c) * 16777216.0f;
        return sign | (uint32_t)rintf(scaled);
    }
    uint32_t h = ((((ax >> 23) - 127u + 15u) & 0x1Fu) << 10) | ((ax & 0x7FFFFFu) >> 13);
    const uint32_t rest = ax & 0x1FFFu;
    if (rest > 0x1000u || (rest == 0x1000u && (h & 1u))) {
        h += 1u;
    }
    return sign | h;
}

static int ingest_finite(float v) { return isfinite(v) ? 1 : 0; }

/* The smallest-three word of a unit-normalised (x, y, z, w): the largest
 * magnitude (first of equals) dropped and made non-negative, the others
 * mapped from [-sqrt(1/2), sqrt(1/2)] to 10 bits, rounded half up. */
static uint32_t ingest_quaternion(float x, float y, float z, float w) {
    const float len = sqrtf(((x * x + y * y) + z * z) + w * w);
    float q[4];
    q[0] = x / len;
    q[1] = y / len;
    q[2] = z / len;
    q[3] = w / len;
    uint32_t largest = 0;
    for (uint32_t k = 1; k < 4; ++k) {
        if (fabsf(q[k]) > fabsf(q[largest])) {
            largest = k;
        }
    }
    if (q[largest] < 0.0f) {
        for (int k = 0; k < 4; ++k) {
            q[k] = -q[k];
        }
    }
    uint32_t word = largest, shift = 22;
    for (uint32_t k = 0; k < 4; ++k) {
        if (k == largest) {
            continue;
        }
        float unit = q[k] / 0.70710678f * 0.5f + 0.5f;
        unit = unit < 0.0f ? 0.0f : (unit > 1.0f ? 1.0f : unit);
        word |= (uint32_t)floorf(unit * 1023.0f + 0.5f) << shift;
        shift -= 10;
    }
    return word;
}

/*
 * count vertices of `stride` bytes; fields[k] = (byte offset, type) of x y z,
 * opacity, scale_0..2, rot_0..3 (w x y z), f_dc_0..2 (k < 14), then `rest`
 * rest coefficients in packing order (rgb a basis). Writes the kept vertices
 * (file order) and returns their count.
 */
/* splat_encoding.slang's decodeOpacity / decodeLogScale / decodeColour /
 * decodeRotation by mode (io::SplatEncoding, RawSplats.h:85-101); a hard
 * surfel's p is kept as packing.slang:393-395's surfelHardnessLog. */
static float decode_opacity(float v, uint32_t mode) {
    return mode == 0 ? athx_sigmoidf(v) : mode == 2 ? v / 255.0f : v;
}

static float decode_log_scale(float v, uint32_t mode) {
    if (mode == 2) return v / 16.0f - 10.0f;
    if (mode == 0) return v < -30.0f ? -INFINITY : v;
    if (v < 0.0f) {
        const float p = -v;
        return p <= 2.001f ? -65000.0f : -(1024.0f + 64.0f * ((p < 32.0f ? p : 32.0f) - 2.0f));
    }
    return v == 0.0f ? -INFINITY : athx_logf(v > 1e-30f ? v : 1e-30f);
}

static float decode_colour(float v, uint32_t mode) {
    const float sh0 = 0.28209479177387814f;
    if (mode == 0) return 0.5f + sh0 * v;
    if (mode == 2) return v / 255.0f;
    if (mode == 3) return 0.5f + sh0 * ((v / 255.0f - 0.5f) / 0.15f);
    return v;
}

static float decode_rotation(float v, uint32_t mode) { return mode == 1 ? (v - 128.0f) / 128.0f : v; }

/* orc_ingest with the layout's encodings: modes = opacity, scale, colour,
 * rotation (athx_ply_layout's). */
ORC_EXPORT uint32_t orc_ingest_modes(const uint8_t* vertices, uint32_t count, uint32_t stride, const uint32_t* offsets,
                                     const uint32_t* types, uint32_t rest, uint32_t keep_faint, const uint32_t* modes,
                                     float* positions, uint32_t* shape, uint32_t* sh, uint32_t* origin) {
    const uint32_t words = (rest + 1) / 2;
    uint32_t kept = 0;
    for (uint32_t i = 0; i < count; ++i) {
        const uint8_t* rec = vertices + (size_t)i * stride;
        float v[14];
        for (int k = 0; k < 14; ++k) {
            v[k] = ingest_widen(rec, offsets[k], types[k]);
        }
        const float opacity = decode_opacity(v[3], modes[0]);
        float ls[3];
        int flats = 0, scales_ok = 1;
        for (int k = 0; k < 3; ++k) {
            ls[k] = decode_log_scale(v[4 + k], modes[1]);
            const int flat = ls[k] <= -1024.0f;
            flats += flat;
            const float kept_ls = flat ? 0.0f : ls[k];
            scales_ok = scales_ok && ingest_finite(kept_ls) && fabsf(kept_ls) < 60.0f;
        }
        const int finite = ingest_finite(v[0]) && ingest_finite(v[1]) && ingest_finite(v[2]);
        const int keep = (finite && flats <= 1 && scales_ok && opacity >= 1.0f / 255.0f) || (keep_faint && finite);
        if (!keep) {
            continue;
        }
        const uint32_t j = kept++;
        if (origin) {
            origin[j] = i;
        }
        float a = opacity < 0.0f ? 0.0f : opacity;
        a = a > 5.0f ? 5.0f : a;
        positions[4 * j + 0] = v[0];
        positions[4 * j + 1] = v[1];
        positions[4 * j + 2] = v[2];
        positions[4 * j + 3] = a;
        /* rot_0 is w; a quaternion of no length is the identity */
        const float qw = decode_rotation(v[7], modes[3]), qx = decode_rotation(v[8], modes[3]),
                    qy = decode_rotation(v[9], modes[3]), qz = decode_rotation(v[10], modes[3]);
        const float qlen = sqrtf(((qx * qx + qy * qy) + qz * qz) + qw * qw);
        const uint32_t rotation = qlen > 1e-8f ? ingest_quaternion(qx, qy, qz, qw)
                                               : ingest_quaternion(0.0f, 0.0f, 0.0f, 1.0f);
        const float base[3] = {decode_colour(v[11], modes[2]), decode_colour(v[12], modes[2]),
                               decode_colour(v[13], modes[2])};
        shape[4 * j + 0] = rotation;
        shape[4 * j + 1] = ingest_half(ls[0]) | (ingest_half(ls[1]) << 16);
        shape[4 * j + 2] = ingest_half(ls[2]) | (ingest_half(base[0]) << 16);
        shape[4 * j + 3] = ingest_half(base[1]) | (ingest_half(base[2]) << 16);
        for (uint32_t k = 0; k < rest; k += 2) {
            const uint32_t lo = ingest_half(ingest_widen(rec, offsets[14 + k], types[14 + k]));
            const uint32_t hi = k + 1 < rest ? ingest_half(ingest_widen(rec, offsets[15 + k], types[15 + k])) : 0u;
            sh[(size_t)j * words + k / 2] = lo | (hi << 16);
        }
    }
    return kept;
}

ORC_EXPORT uint32_t orc_ingest(const uint8_t* vertices, uint32_t count, uint32_t stride, const uint32_t* offsets,
                               const uint32_t* types, uint32_t rest, uint32_t keep_faint, float* positions,
                               uint32_t* shape, uint32_t* sh, uint32_t* origin) {
    static const uint32_t kPly[4] = {0, 0, 0, 0};
    return orc_ingest_modes(vertices, count, stride, offsets, types, rest, keep_faint, kPly, positions, shape, sh,
                            origin);
}

ORC_EXPORT uint32_t orc_half_bits(float v) { return ingest_half(v); }
ORC_EXPORT uint32_t orc_encode_quaternion(float x, float y, float z, float w) { return ingest_quaternion(x, y, z, w); }


/* ---- the ground truth's projection and the image comparisons ---- */

static float ref_half(uint32_t h) { return half_to_float(h & 0xFFFFu); }

/* reference_project.slang over the packed layout, every product written out
 * in the order it is summed: rec[12] a splat (the frame's record layout), z
 * = -1 when culled. rest: coefficients a colour (0, 3, 8, 15). */
ORC_EXPORT void orc_ref_project(const orc_view* v, const float* positions, const uint32_t* shape,
                                const uint32_t* sh, uint32_t n, uint32_t rest, float* records) {
    const uint32_t words = (rest * 3 + 1) / 2;
    for (uint32_t i = 0; i < n; ++i) {
        float* rec = records + (size_t)i * REC;
        memset(rec, 0, REC * sizeof(float));
        rec[2] = -1.0f;
        const float* p = positions + (size_t)i * 4;
        const uint32_t* w4 = shape + (size_t)i * 4;
        const float s[3] = {athx_expf(ref_half(w4[1])), athx_expf(ref_half(w4[1] >> 16)), athx_expf(ref_half(w4[2]))};
        const float* m = v->m;
        const float vx = ((m[0] * p[0] + m[1] * p[1]) + m[2] * p[2]) + m[3];
        const float vy = ((m[4] * p[0] + m[5] * p[1]) + m[6] * p[2]) + m[7];
        const float vz = ((m[8] * p[0] + m[9] * p[1]) + m[10] * p[2]) + m[11];
        if (!(vz > v->near_z) || vz > v->far_z) {
            continue;
        }
        const uint32_t largest = w4[0] & 3u;
        float small[3];
        small[0] = ((float)((w4[0] >> 22) & 1023u) / 1023.0f * 2.0f - 1.0f) * 0.70710678f;
        small[1] = ((float)((w4[0] >> 12) & 1023u) / 1023.0f * 2.0f - 1.0f) * 0.70710678f;
        small[2] = ((float)((w4[0] >> 2) & 1023u) / 1023.0f * 2.0f - 1.0f) * 0.70710678f;
        const float rest_q = sqrtf(fmaxf(0.0f, 1.0f - ((small[0] * small[0] + small[1] * small[1]) + small[2] * small[2])));
        float q[4];
        for (uint32_t k = 0, next = 0; k < 4; ++k) {
            q[k] = k == largest ? rest_q : small[next < 3 ? next++ : 2];
        }
        const float x = q[0], y = q[1], z = q[2], w = q[3];
        const float R[3][3] = {{1.0f - 2.0f * (y * y + z * z), 2.0f * (x * y - w * z), 2.0f * (x * z + w * y)},
                               {2.0f * (x * y + w * z), 1.0f - 2.0f * (x * x + z * z), 2.0f * (y * z - w * x)},
                               {2.0f * (x * z - w * y), 2.0f * (y * z + w * x), 1.0f - 2.0f * (x * x + y * y)}};
        float M[3][3], C3[3][3], Wm[3][3], T1[3][3], CV[3][3];
        for (int r = 0; r < 3; ++r) {
            for (int c = 0; c < 3; ++c) {
                M[r][c] = R[r][c] * s[c];
                Wm[r][c] = m[4 * r + c];
            }
        }
        for (int r = 0; r < 3; ++r) {
            for (int c = 0; c < 3; ++c) {
                C3[r][c] = (M[r][0] * M[c][0] + M[r][1] * M[c][1]) + M[r][2] * M[c][2];
            }
        }
        for (int r = 0; r < 3; ++r) {
            for (int c = 0; c < 3; ++c) {
                T1[r][c] = (Wm[r][0] * C3[0][c] + Wm[r][1] * C3[1][c]) + Wm[r][2] * C3[2][c];
            }
        }
        for (int r = 0; r < 3; ++r) {
            for (int c = 0; c < 3; ++c) {
                CV[r][c] = (T1[r][0] * Wm[c][0] + T1[r][1] * Wm[c][1]) + T1[r][2] * Wm[c][2];
            }
        }
        float cx, cy, J[2][3] = {{0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}};
        if (v->orthographic) {
            cx = v->focal_x * vx + v->centre_x;
            cy = v->focal_y * vy + v->centre_y;
            J[0][0] = v->focal_x;
            J[1][1] = v->focal_y;
        } else {
            cx = v->focal_x * vx / vz + v->centre_x;
            cy = v->focal_y * vy / vz + v->centre_y;
            const float lox = (-0.15f * (float)v->width - v->centre_x) / v->focal_x;
            const float hix = (1.15f * (float)v->width - v->centre_x) / v->focal_x;
            const float loy = (-0.15f * (float)v->height - v->centre_y) / v->focal_y;
            const float hiy = (1.15f * (float)v->height - v->centre_y) / v->focal_y;
            const float tx = fminf(fmaxf(vx / vz, lox), hix) * vz;
            const float ty = fminf(fmaxf(vy / vz, loy), hiy) * vz;
            J[0][0] = v->focal_x / vz;
            J[0][2] = -v->focal_x * tx / (vz * vz);
            J[1][1] = v->focal_y / vz;
            J[1][2] = -v->focal_y * ty / (vz * vz);
        }
        float X[3][2];
        for (int k = 0; k < 3; ++k) {
            for (int c = 0; c < 2; ++c) {
                X[k][c] = (CV[k][0] * J[c][0] + CV[k][1] * J[c][1]) + CV[k][2] * J[c][2];
            }
        }
        const float c00 = (J[0][0] * X[0][0] + J[0][1] * X[1][0]) + J[0][2] * X[2][0];
        const float c01 = (J[0][0] * X[0][1] + J[0][1] * X[1][1]) + J[0][2] * X[2][1];
        const float c11 = (J[1][0] * X[0][1] + J[1][1] * X[1][1]) + J[1][2] * X[2][1];
        const float before = c00 * c11 - c01 * c01;
        const float ca = c00 + 0.3f, cb = c01, cc = c11 + 0.3f;
        const float det = ca * cc - cb * cb;
        if (!(det > 1e-12f)) {
            continue;
        }
        const float filtered = sqrtf(fmaxf(before, 0.0f) / det);
        const float paid = v->antialias ? filtered : 1.0f;
        const float spread = paid * p[3];
        const float alpha = spread >= 1.0f / 255.0f ? spread + (1.0f - paid) * (1.0f / 255.0f) : spread;
        if (!(alpha >= 1.0f / 255.0f)) {
            continue;
        }
        const float cutoff = fminf(2.0f * athx_logf(fmaxf(255.0f * alpha, 1.0001f)), 9.0f);
        const float dx = p[0] - v->eye[0], dy = p[1] - v->eye[1], dz = p[2] - v->eye[2];
        const float len = sqrtf((dx * dx + dy * dy) + dz * dz);
        const float ux = dx / len, uy = dy / len, uz = dz / len;
        const float xx = ux * ux, yy = uy * uy, zz = uz * uz, xy = ux * uy, yz = uy * uz, xz = ux * uz;
        const float c1 = 0.4886025119029199f;
        float Y[16];
        Y[0] = 0.0f;
        Y[1] = -c1 * uy;
        Y[2] = c1 * uz;
        Y[3] = -c1 * ux;
        Y[4] = 1.0925484305920792f * xy;
        Y[5] = -1.0925484305920792f * yz;
        Y[6] = 0.31539156525252005f * (2.0f * zz - xx - yy);
        Y[7] = -1.0925484305920792f * xz;
        Y[8] = 0.5462742152960396f * (xx - yy);
        Y[9] = -0.5900435899266435f * uy * (3.0f * xx - yy);
        Y[10] = 2.890611442640554f * xy * uz;
        Y[11] = -0.4570457994644658f * uy * (4.0f * zz - xx - yy);
        Y[12] = 0.3731763325901154f * uz * (2.0f * zz - 3.0f * xx - 3.0f * yy);
        Y[13] = -0.4570457994644658f * ux * (4.0f * zz - xx - yy);
        Y[14] = 1.445305721320277f * uz * (xx - yy);
        Y[15] = -0.5900435899266435f * ux * (xx - 3.0f * yy);
        const uint32_t rest_degree = rest >= 15 ? 3u : rest >= 8 ? 2u : rest >= 3 ? 1u : 0u;
        const uint32_t degree = rest_degree < v->sh_limit ? rest_degree : v->sh_limit;
        const uint32_t terms = (degree + 1) * (degree + 1);
        const float base[3] = {ref_half(w4[2] >> 16), ref_half(w4[3]), ref_half(w4[3] >> 16)};
        float col[3];
        for (uint32_t ch = 0; ch < 3; ++ch) {
            float val = base[ch];
            for (uint32_t k = 1; k < terms; ++k) {
                const uint32_t h = (k - 1) * 3 + ch;
                const uint32_t word = sh[(size_t)i * words + (h >> 1)];
                val = val + Y[k] * ref_half((h & 1u) ? word >> 16 : word);
            }
            col[ch] = v->linear ? fmaxf(val, 0.0f) : val;
        }
        if (!v->linear) {
            for (uint32_t ch = 0; ch < 3; ++ch) {
                const float c = fmaxf(col[ch], 0.0f);
                col[ch] = c <= 0.04045f ? c / 12.92f : athx_powf_pos((c + 0.055f) / 1.055f, 2.4f);
            }
        }
        rec[0] = cx;
        rec[1] = cy;
        rec[2] = vz;
        rec[3] = cutoff;
        rec[4] = cc / det;
        rec[5] = -cb / det;
        rec[6] = ca / det;
        rec[7] = alpha;
        rec[8] = col[0];
        rec[9] = col[1];
        rec[10] = col[2];
    }
}

static uint32_t ref_code(float linear) {
    const float c = fminf(fmaxf(linear, 0.0f), 1.0f);
    const float e = c <= 0.0031308f ? c * 12.92f : 1.055f * athx_powf_pos(c, 1.0f / 2.4f) - 0.055f;
    return (uint32_t)rintf(e * 255.0f);
}

/* 256 bins of the per-pixel worst code difference (rgb and alpha). */
ORC_EXPORT void orc_compare_codes(const float* a, const float* b, uint32_t width, uint32_t height, uint32_t* bins) {
    memset(bins, 0, 256 * sizeof(uint32_t));
    for (uint64_t p = 0; p < (uint64_t)width * height; ++p) {
        const float* pa = a + p * 4;
        const float* pb = b + p * 4;
        uint32_t worst = 0;
        for (int c = 0; c < 3; ++c) {
            const int d = (int)ref_code(pa[c]) - (int)ref_code(pb[c]);
            const uint32_t ad = (uint32_t)(d < 0 ? -d : d);
            worst = ad > worst ? ad : worst;
        }
        const int da = (int)rintf(fminf(fmaxf(pa[3], 0.0f), 1.0f) * 255.0f) -
                       (int)rintf(fminf(fmaxf(pb[3], 0.0f), 1.0f) * 255.0f);
        const uint32_t ada = (uint32_t)(da < 0 ? -da : da);
        worst = ada > worst ? ada : worst;
        bins[worst] += 1;
    }
}

static uint32_t ref_hdr_bin(float v) {
    if (!(v > 0.0f)) {
        return 0u;
    }
    if (!(v < 3.0e38f)) {
        return 255u;
    }
    int e = 0;
    const float m = frexpf(v, &e) * 2.0f;
    static const float steps[7] = {1.0905077f, 1.1892071f, 1.2968396f, 1.4142135f, 1.5422108f, 1.6817929f,
                                   1.8340081f};
    int frac = 0;
    for (int j = 0; j < 7; ++j) {
        frac += m >= steps[j] ? 1 : 0;
    }
    const int bin = 8 * ((e - 1) + 16) + frac + 1;
    return (uint32_t)(bin < 1 ? 1 : bin > 255 ? 255 : bin);
}

/* 256 bins of the per-pixel worst relative difference and relMSE's sum:
 * Kahan along each row, the rows' sums added with Kahan in row order. */
ORC_EXPORT float orc_compare_hdr(const float* a, const float* b, uint32_t width, uint32_t height, uint32_t* bins) {
    memset(bins, 0, 256 * sizeof(uint32_t));
    float total = 0.0f, total_carry = 0.0f;
    for (uint32_t y = 0; y < height; ++y) {
        float sum = 0.0f, carry = 0.0f;
        for (uint32_t x = 0; x < width; ++x) {
            const float* pa = a + ((uint64_t)y * width + x) * 4;
            const float* pb = b + ((uint64_t)y * width + x) * 4;
            float worst = 0.0f, squared = 0.0f;
            for (int c = 0; c < 4; ++c) {
                const float d = fabsf(pa[c] - pb[c]);
                worst = fmaxf(worst, d / fmaxf(fabsf(pb[c]), 1e-3f));
                squared = squared + d * d / (pb[c] * pb[c] + 1e-2f);
            }
            bins[ref_hdr_bin(worst)] += 1;
            const float term = squared * 0.25f;
            const float t = sum + term;
            carry = carry + (fabsf(sum) >= fabsf(term) ? (sum - t) + term : (term - t) + sum);
            sum = t;
        }
        const float row = sum + carry;
        const float t = total + row;
        total_carry = total_carry + (fabsf(total) >= fabsf(row) ? (total - t) + row : (row - t) + total);
        total = t;
    }
    return total + total_carry;
}


/* ---- contributions: what each slot gave the frame (modes 1 to 4) ---- */

static float contribution_hash01(uint32_t v) {
    const uint32_t state = v * 747796405u + 2891336453u;
    const uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (float)((word >> 22u) ^ word) / 4294967295.0f;
}

/* The walk of blend_pixel_f32 for every pixel of one view, adding at each
 * contributing pair's slot: mode 1 round(w 4096); mode 2 the removal
 * error round(min(|e|^2, 1) 2^24) into a 64-bit (low, high) pair, dl the
 * drawn image [h][w][4]; mode 3 round(w m 1024) and round(w 1024), dl the
 * mask [h][w]; mode 4 floor(w dL/dC 256 + hash) a channel and round(w 256),
 * dl: [h][w][4]. */
ORC_EXPORT void orc_contribution(const orc_view* v, const float* rec, const uint32_t* ranges, const uint32_t* vals,
                                 uint32_t tile_offset, uint32_t mode, const float* dl, uint32_t* out) {
    const uint32_t W = v->width, H = v->height, T = v->tile;
    const uint32_t tiles_x = (W + T - 1) / T;
    for (uint64_t p = 0; p < (uint64_t)W * H; ++p) {
        const uint32_t px = (uint32_t)(p % W), py = (uint32_t)(p / W);
        const uint32_t tile = tile_offset + (py / T) * tiles_x + px / T;
        const float pxc = (float)px + 0.5f, pyc = (float)py + 0.5f;
        float Tr = 1.0f;
        float run[3] = {0.0f, 0.0f, 0.0f};
        for (uint32_t j = ranges[tile * 2]; j < ranges[tile * 2 + 1]; ++j) {
            const uint32_t slot = vals[j];
            const float* r = rec + (size_t)slot * REC;
            const float dx = pxc - r[0], dy = pyc - r[1];
            const float power = r[4] * dx * dx + 2.0f * r[5] * dx * dy + r[6] * dy * dy;
            if (power > r[3]) continue;
            const float alpha = minf_(0.99f, r[7] * athx_expf(-0.5f * power));
            if (alpha < 1.0f / 255.0f) continue;
            const float w = Tr * alpha;
            if (mode == 1) {
                out[slot] += (uint32_t)(w * 4096.0f + 0.5f);
            } else if (mode == 2) {
                /* splat_blend.slang:366-379: the change removing it makes */
                const float keep = maxf_(Tr * (1.0f - alpha), 1.0e-4f);
                float e[3];
                for (uint32_t c = 0; c < 3; ++c) {
                    e[c] = w * (r[8 + c] - (dl[p * 4 + c] - (run[c] + w * r[8 + c])) / keep);
                }
                const float e2 = minf_(e[0] * e[0] + e[1] * e[1] + e[2] * e[2], 1.0f);
                const uint32_t add = (uint32_t)(e2 * 16777216.0f + 0.5f);
                const uint32_t before = out[(size_t)slot * 2];
                out[(size_t)slot * 2] = before + add;
                out[(size_t)slot * 2 + 1] += (before + add < before) ? 1u : 0u;
                for (uint32_t c = 0; c < 3; ++c) {
                    run[c] = run[c] + w * r[8 + c];
                }
            } else if (mode == 3) {
                /* splat_blend.slang:399-408 */
                const uint32_t seen = (uint32_t)(w * 1024.0f + 0.5f);
                if (seen != 0u) {
                    out[(size_t)slot * 2 + 1] += seen;
                    const float m = minf_(maxf_(dl[p], 0.0f), 1.0f);
                    out[(size_t)slot * 2] += (uint32_t)(w * m * 1024.0f + 0.5f);
                }
            } else {
                const uint32_t at = (uint32_t)p;
                for (uint32_t c = 0; c < 3; ++c) {
                    const float g = w * dl[p * 4 + c] * 256.0f;
                    const int step = (int)floorf(g + contribution_hash01((at * 3u + c) ^ (slot * 2654435761u)));
                    out[(size_t)slot * 4 + c] += (uint32_t)step;
                }
                out[(size_t)slot * 4 + 3] += (uint32_t)(w * 256.0f + 0.5f);
            }
            Tr = Tr * (1.0f - alpha);
            if (Tr < 1e-4f) break;
        }
    }
}

/* ------------------------------------------------------------------------
 * THE SUN TAKEN OUT OF A SKY (technique/env_sun.slang:154-262) and the sky
 * without it (technique/env_residual.slang:55-78), for an equirect of linear
 * radiance [h][w][3]: texel (i, j) looks along theta = pi (j + 0.5) / h from
 * +y, phi = 2 pi (i + 0.5) / w from +x towards +z. The search and the sums
 * run on the level halved (2 x 2 means, clamped at an odd edge) until at most
 * 256 wide, as env_sun reads its projection level. The reference's float
 * sums become a histogram of integer weights (sin theta in 2^-24 units: exact
 * in any order) and FP64 sums of the float per-texel terms; its sin, cos and
 * acos are taken in double, the directions in float from those.
 * ------------------------------------------------------------------------ */
static float sky_lum_(const float* rgb) { return rgb[0] * 0.2126f + rgb[1] * 0.7152f + rgb[2] * 0.0722f; }

static void sky_tables_(uint32_t w, uint32_t h, float* rows, float* cols, uint32_t* weights) {
    const double pi = 3.14159265358979323846;
    for (uint32_t j = 0; j < h; ++j) {
        const double theta = pi * ((double)j + 0.5) / (double)h;
        rows[2 * j] = (float)sin(theta);
        rows[2 * j + 1] = (float)cos(theta);
        if (weights) weights[j] = (uint32_t)llround(fmax(sin(theta), 0.0) * 16777216.0);
    }
    for (uint32_t i = 0; i < w; ++i) {
        const double phi = 2.0 * pi * ((double)i + 0.5) / (double)w;
        cols[2 * i] = (float)cos(phi);
        cols[2 * i + 1] = (float)sin(phi);
    }
}

/* env_sun.slang:154-261 into sun[8]: (direction, cone half angle or 0),
 * (irradiance rgb, the luminance cut at). Returns 0, or -1 on no memory. */
int orc_sky_sun(const float* img, uint32_t width, uint32_t height, float* sun) {
    const double pi = 3.14159265358979323846;
    uint32_t w = width, h = height;
    float* level = (float*)malloc((size_t)w * h * 3 * sizeof(float));
    if (!level) return -1;
    memcpy(level, img, (size_t)w * h * 3 * sizeof(float));
    while (w > 256u) {
        const uint32_t dw = w >> 1 > 0 ? w >> 1 : 1u, dh = h >> 1 > 0 ? h >> 1 : 1u;
        float* next = (float*)malloc((size_t)dw * dh * 3 * sizeof(float));
        if (!next) { free(level); return -1; }
        for (uint32_t y = 0; y < dh; ++y) {
            for (uint32_t x = 0; x < dw; ++x) {
                const uint32_t x0 = 2 * x < w - 1 ? 2 * x : w - 1, x1 = 2 * x + 1 < w - 1 ? 2 * x + 1 : w - 1;
                const uint32_t y0 = 2 * y < h - 1 ? 2 * y : h - 1, y1 = 2 * y + 1 < h - 1 ? 2 * y + 1 : h - 1;
                for (int c = 0; c < 3; ++c) {
                    const float a = level[((size_t)y0 * w + x0) * 3 + c], b = level[((size_t)y0 * w + x1) * 3 + c];
                    const float e = level[((size_t)y1 * w + x0) * 3 + c], f = level[((size_t)y1 * w + x1) * 3 + c];
                    next[((size_t)y * dw + x) * 3 + c] = 0.25f * ((a + b) + (e + f));
                }
            }
        }
        free(level);
        level = next;
        w = dw;
        h = dh;
    }
    float* rows = (float*)malloc(2 * (size_t)h * sizeof(float));
    float* cols = (float*)malloc(2 * (size_t)w * sizeof(float));
    uint32_t* weights = (uint32_t*)malloc((size_t)h * sizeof(uint32_t));
    if (!rows || !cols || !weights) { free(level); free(rows); free(cols); free(weights); return -1; }
    sky_tables_(w, h, rows, cols, weights);
    /* sunRegionOf (env_sun.slang:165-204): the histogram, its median, the peak. */
    uint64_t hist[80];
    memset(hist, 0, sizeof(hist));
    uint64_t total = 0;
    float peak = -1.0f;
    uint32_t peak_at = 0;
    for (uint32_t y = 0; y < h; ++y) {
        if (weights[y] == 0u) continue;
        for (uint32_t x = 0; x < w; ++x) {
            const float lum = sky_lum_(level + ((size_t)y * w + x) * 3);
            const float l2 = athx_logf(maxf_(lum, 1.0e-9f)) * 1.44269504088896341f;
            const float fb = minf_(maxf_((l2 + 20.0f) * 2.0f, 0.0f), 79.0f);
            hist[(uint32_t)fb] += weights[y];
            total += weights[y];
            if (maxf_(lum, 0.0f) > peak) {
                peak = maxf_(lum, 0.0f);
                peak_at = y * w + x;
            }
        }
    }
    float median = 0.0f;
    uint64_t below = 0;
    for (uint32_t b = 0; b < 80; ++b) {
        below += hist[b];
        if (2 * below >= total && total > 0) {
            median = (float)exp2((double)b * 0.5 - 20.0 + 0.25);
            break;
        }
    }
    const uint32_t py = peak_at / w, px = peak_at - py * w;
    const float axis[3] = {rows[2 * py] * cols[2 * px], rows[2 * py + 1], rows[2 * py] * cols[2 * px + 1]};
    const float threshold = maxf_(16.0f * median, 0.02f * peak);
    const int found = peak >= 0.0f && peak > 32.0f * maxf_(median, 1.0e-9f);
    const float cos_widest = (float)cos(0.175);
    double sums[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (found) {
        const float area = (float)(2.0 * pi * pi / ((double)w * (double)h));
        for (uint32_t y = 0; y < h; ++y) {
            const float st = rows[2 * y], ct = rows[2 * y + 1];
            if (!(st > 0.0f)) continue;
            for (uint32_t x = 0; x < w; ++x) {
                const float dx = st * cols[2 * x], dy = ct, dz = st * cols[2 * x + 1];
                const float* rgb = level + ((size_t)y * w + x) * 3;
                const float lum = sky_lum_(rgb);
                if (!(lum >= threshold) || !(((dx * axis[0] + dy * axis[1]) + dz * axis[2]) >= cos_widest)) continue;
                const float omega = st * area, lw = lum * omega;
                sums[0] += (double)(rgb[0] * omega);
                sums[1] += (double)(rgb[1] * omega);
                sums[2] += (double)(rgb[2] * omega);
                sums[3] += (double)(dx * lw);
                sums[4] += (double)(dy * lw);
                sums[5] += (double)(dz * lw);
                sums[6] += (double)lw;
                sums[7] += (double)omega;
            }
        }
    }
    /* env_sun.slang:236-261 */
    const double solid = sums[7], weight = sums[6];
    const double clen = sqrt(sums[3] * sums[3] + sums[4] * sums[4] + sums[5] * sums[5]);
    const double disc = solid > 0.0 ? acos(fmin(fmax(1.0 - solid / (2.0 * pi), 0.0), 1.0)) : 0.0;
    const double spread = weight > 0.0 ? 2.0 * fmax(1.0 - clen / weight, 0.0) : 0.0;
    const int compact = spread <= disc * disc + 1.0e-6;
    const int is_sun = found && solid > 0.0 && clen > 0.0 && compact;
    for (int c = 0; c < 3; ++c) {
        sun[c] = is_sun ? (float)(sums[3 + c] / clen) : axis[c];
        sun[4 + c] = (float)sums[c];
    }
    sun[3] = is_sun ? maxf_((float)disc, 1.0e-4f) : 0.0f;
    sun[7] = threshold;
    free(level);
    free(rows);
    free(cols);
    free(weights);
    return 0;
}

/* env_residual.slang:55-78: the sun's texels (at least sun[7] bright, within
 * 0.175 rad of its direction) set to 0; none when sun[3] is 0. */
int orc_sky_residual(const float* img, uint32_t w, uint32_t h, const float* sun, float* out) {
    float* rows = (float*)malloc(2 * (size_t)h * sizeof(float));
    float* cols = (float*)malloc(2 * (size_t)w * sizeof(float));
    if (!rows || !cols) { free(rows); free(cols); return -1; }
    sky_tables_(w, h, rows, cols, NULL);
    const float threshold = sun[3] > 0.0f ? sun[7] : INFINITY;
    const float cos_widest = (float)cos(0.175);
    for (uint32_t y = 0; y < h; ++y) {
        for (uint32_t x = 0; x < w; ++x) {
            const float st = rows[2 * y], ct = rows[2 * y + 1];
            const float dx = st * cols[2 * x], dy = ct, dz = st * cols[2 * x + 1];
            const float* rgb = img + ((size_t)y * w + x) * 3;
            const int sunlit = sky_lum_(rgb) >= threshold && ((dx * sun[0] + dy * sun[1]) + dz * sun[2]) >= cos_widest;
            for (int c = 0; c < 3; ++c) out[((size_t)y * w + x) * 3 + c] = sunlit ? 0.0f : rgb[c];
        }
    }
    free(rows);
    free(cols);
    return 0;
}
