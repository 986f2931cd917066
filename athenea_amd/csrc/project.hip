// project.hip -- the splat projection for gfx950: one thread per Gaussian,
// looping over the frame's views, so each Gaussian's parameters are read
// from HBM once however many views a batch holds.
//
// Restates shaders/athenea/splat/splat_project.slang:497-1130 (projectSplat
// with kTransfer 0: no edit, no lens, no shutter, no surfels, not relit) and
// its helpers in shaders/athenea/splat/frame.slang:283-402. Arithmetic is in
// the order the reference writes it; the file is compiled without
// floating-point contraction, and exp/log come from athx_detmath.h, so the
// depth keys, tile counts and records are the same bits a host restatement
// computes.
#include "athx_internal.h"
#include "relight.h"

namespace athx {
namespace {

constexpr float kSH0 = 0.28209479177387814f;
constexpr float kC1 = 0.4886025119029199f;
constexpr float kC2_0 = 1.0925484305920792f, kC2_1 = -1.0925484305920792f, kC2_2 = 0.31539156525252005f,
                kC2_3 = -1.0925484305920792f, kC2_4 = 0.5462742152960396f;
constexpr float kC3_0 = -0.5900435899266435f, kC3_1 = 2.890611442640554f, kC3_2 = -0.4570457994644658f,
                kC3_3 = 0.3731763325901154f, kC3_4 = -0.4570457994644658f, kC3_5 = 1.445305721320277f,
                kC3_6 = -0.5900435899266435f;

// frame.slang:325-327
__device__ inline float cutoff_power(float alpha) {
    return fmin_(2.0f * athx_logf(fmax_(255.0f * alpha, 1.0001f)), 9.0f);
}

// frame.slang:374-379
__device__ inline bool tile_touched(uint32_t W, uint32_t H, uint32_t T, float cx, float cy, float a, float b,
                                    float c, float cutoff, uint32_t tx, uint32_t ty) {
    const float fx = (float)__umul24(tx, T) + 0.5f;
    const float fy = (float)__umul24(ty, T) + 0.5f;
    const uint32_t ex = __umul24(tx + 1, T) < W ? __umul24(tx + 1, T) : W;
    const uint32_t ey = __umul24(ty + 1, T) < H ? __umul24(ty + 1, T) : H;
    const float lx = (float)(ex - 1) + 0.5f;
    const float ly = (float)(ey - 1) + 0.5f;
    return min_power_over_box(a, b, c, fx - cx, fy - cy, lx - cx, ly - cy) <= cutoff;
}

// frame.slang:383-402: the footprint's tile rectangle, inclusive; false when
// it lies off screen.
__device__ inline bool tile_rect(uint32_t W, uint32_t H, uint32_t T, float cx, float cy, float a, float b, float c,
                                 float cutoff, uint32_t& x0, uint32_t& y0, uint32_t& x1, uint32_t& y1) {
    const float det = a * c - b * b;
    const float inv = det > 1e-20f ? 1.0f / det : 0.0f;
    const float ex = sqrtf(fmax_(cutoff * c * inv, 0.0f));
    const float ey = sqrtf(fmax_(cutoff * a * inv, 0.0f));
    const float lx = cx - ex, ly = cy - ey, hx = cx + ex, hy = cy + ey;
    if (hx < 0.0f || hy < 0.0f || lx >= (float)W || ly >= (float)H) {
        x0 = 1; y0 = 1; x1 = 0; y1 = 0;
        return false;
    }
    // T is 8, 16 or 32: the quotient is a shift.
    const uint32_t shift = (uint32_t)__builtin_ctz(T);
    x0 = (uint32_t)fmax_(lx, 0.0f) >> shift;
    y0 = (uint32_t)fmax_(ly, 0.0f) >> shift;
    x1 = (uint32_t)fmin_(hx, (float)W - 1.0f) >> shift;
    y1 = (uint32_t)fmin_(hy, (float)H - 1.0f) >> shift;
    return x0 <= x1 && y0 <= y1;
}

__device__ inline float half_bits_to_float(uint16_t h) { return (float)__builtin_bit_cast(_Float16, h); }

template <bool kHalf>
__device__ inline float sh_at(const float* sh, const uint16_t* shh, size_t at) {
    if constexpr (kHalf) {
        return half_bits_to_float(shh[at]);
    } else {
        return sh[at];
    }
}

// shaders/athenea/common/sh.slang:25-51, one channel: B(k) is coefficient k
// of the channel (k = 0, the constant term, is read by the caller).
template <typename Coeff>
__device__ inline float sh_rest(Coeff B, uint32_t degree, float x, float y, float z) {
    if (degree == 0) {
        return 0.0f;
    }
    float r = -kC1 * y * B(1) + kC1 * z * B(2) - kC1 * x * B(3);
    if (degree < 2) {
        return r;
    }
    const float xx = x * x, yy = y * y, zz = z * z;
    const float xy = x * y, yz = y * z, xz = x * z;
    r = r + (kC2_0 * xy * B(4) + kC2_1 * yz * B(5) + kC2_2 * (2.0f * zz - xx - yy) * B(6) + kC2_3 * xz * B(7) +
             kC2_4 * (xx - yy) * B(8));
    if (degree < 3) {
        return r;
    }
    r = r + (kC3_0 * y * (3.0f * xx - yy) * B(9) + kC3_1 * xy * z * B(10) + kC3_2 * y * (4.0f * zz - xx - yy) * B(11) +
             kC3_3 * z * (2.0f * zz - 3.0f * xx - 3.0f * yy) * B(12) + kC3_4 * x * (4.0f * zz - xx - yy) * B(13) +
             kC3_5 * z * (xx - yy) * B(14) + kC3_6 * x * (xx - 3.0f * yy) * B(15));
    return r;
}

// All 16 x 3 coefficients of Gaussian i in 16-byte loads (12 of FP32, 6 of
// FP16): a row is 192 (96) bytes, so a lane's loads are whole 16-byte words
// and a wavefront touches 12 (6) lines a lane in all, where one coefficient
// a load touched 48.
template <bool kHalf>
__device__ inline void load_sh16(const float* sh, const uint16_t* shh, uint32_t i, float c[48]) {
    if constexpr (kHalf) {
        const uint4* p = reinterpret_cast<const uint4*>(shh + (size_t)i * 48);
#pragma unroll
        for (uint32_t q = 0; q < 6; ++q) {
            const uint4 v = p[q];
            const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (uint32_t e = 0; e < 4; ++e) {
                c[q * 8 + 2 * e] = half_bits_to_float((uint16_t)(w[e] & 0xFFFFu));
                c[q * 8 + 2 * e + 1] = half_bits_to_float((uint16_t)(w[e] >> 16));
            }
        }
    } else {
        const float4* p = reinterpret_cast<const float4*>(sh + (size_t)i * 48);
#pragma unroll
        for (uint32_t q = 0; q < 12; ++q) {
            const float4 v = p[q];
            c[q * 4 + 0] = v.x;
            c[q * 4 + 1] = v.y;
            c[q * 4 + 2] = v.z;
            c[q * 4 + 3] = v.w;
        }
    }
}

// shaders/athenea/common/color.slang:7-10
__device__ inline float srgb_to_linear(float c) {
    c = fmax_(c, 0.0f);
    return c <= 0.04045f ? c / 12.92f : athx_powf_pos((c + 0.055f) / 1.055f, 2.4f);
}

__device__ inline void cull(uint32_t slot, uint32_t reason, uint32_t* touched, uint32_t* keys) {
    touched[slot] = 0;
    keys[slot] = kCulledKey | reason;
}

// A slot culled before its tiles are counted.
__device__ inline void cull_early(uint32_t slot, uint32_t reason, uint32_t* touched, uint32_t* keys) {
    cull(slot, reason, touched, keys);
}

// packing.slang:59-76 (decodeQuaternion): three 10-bit components of the
// smallest-three encoding and the index of the dropped (largest) one, which
// is rebuilt from the unit length. Not normalised again, as there.
__device__ inline void decode_quaternion(uint32_t word, float& x, float& y, float& z, float& w) {
    constexpr float kSqrtHalf = 0.70710678f;
    const uint32_t largest = word & 3u;
    const float s0 = ((float)((word >> 22) & 1023u) / 1023.0f * 2.0f - 1.0f) * kSqrtHalf;
    const float s1 = ((float)((word >> 12) & 1023u) / 1023.0f * 2.0f - 1.0f) * kSqrtHalf;
    const float s2 = ((float)((word >> 2) & 1023u) / 1023.0f * 2.0f - 1.0f) * kSqrtHalf;
    const float rest = sqrtf(fmax_(0.0f, 1.0f - (s0 * s0 + s1 * s1 + s2 * s2)));
    float q[4];
    const float small[3] = {s0, s1, s2};
    uint32_t next = 0;
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k) {
        if (k == largest) {
            q[k] = rest;
        } else {
            q[k] = small[next < 3 ? next : 2];
            next += 1;
        }
    }
    x = q[0]; y = q[1]; z = q[2]; w = q[3];
}

// The Gaussians' input, a template argument of the projection: the FP32 or
// FP16-SH structure of arrays (Vec: 16 coefficients a channel in aligned
// rows, read whole once for every view), or athenea's packed layout.
constexpr uint32_t kInSoa32 = 0, kInSoa32Vec = 1, kInSoa16 = 2, kInSoa16Vec = 3, kInPacked = 4;

template <uint32_t kIn, bool kRelit>
__global__ void __launch_bounds__(256) project_kernel(ProjectArgs a, const DView* __restrict__ views, uint32_t nviews,
                                                      float* __restrict__ records, uint32_t* __restrict__ touched,
                                                      uint32_t* __restrict__ keys) {
    constexpr bool kHalf = kIn == kInSoa16 || kIn == kInSoa16Vec;
    constexpr bool kVec = kIn == kInSoa32Vec || kIn == kInSoa16Vec;
    constexpr bool kPacked = kIn == kInPacked;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) {
        return;
    }
    float px, py, pz, sx, sy, sz, opacity, qx, qy, qz, qw;
    float stored[3] = {0.0f, 0.0f, 0.0f};  // the packed layout's base colour
    if constexpr (kPacked) {
        // splat_project.slang:507-510, 526, 781-782: position and linear
        // opacity, f16 log-scales, the quaternion and the base colour.
        const float4 po = reinterpret_cast<const float4*>(a.positions)[i];
        const uint4 shape = reinterpret_cast<const uint4*>(a.shape)[i];
        px = po.x; py = po.y; pz = po.z;
        opacity = po.w;
        sx = athx_expf(half_bits_to_float((uint16_t)(shape.y & 0xFFFFu)));
        sy = athx_expf(half_bits_to_float((uint16_t)(shape.y >> 16)));
        sz = athx_expf(half_bits_to_float((uint16_t)(shape.z & 0xFFFFu)));
        decode_quaternion(shape.x, qx, qy, qz, qw);
        stored[0] = half_bits_to_float((uint16_t)(shape.z >> 16));
        stored[1] = half_bits_to_float((uint16_t)(shape.w & 0xFFFFu));
        stored[2] = half_bits_to_float((uint16_t)(shape.w >> 16));
    } else {
        px = a.means[i * 3 + 0]; py = a.means[i * 3 + 1]; pz = a.means[i * 3 + 2];
        sx = athx_expf(a.log_scales[i * 3 + 0]);
        sy = athx_expf(a.log_scales[i * 3 + 1]);
        sz = athx_expf(a.log_scales[i * 3 + 2]);
        opacity = athx_sigmoidf(a.logits[i]);
        // Quaternion (x, y, z, w), normalised (packing.slang:59-76 decodes a unit one).
        qx = a.quats[i * 4 + 0]; qy = a.quats[i * 4 + 1]; qz = a.quats[i * 4 + 2]; qw = a.quats[i * 4 + 3];
        const float qn = qx * qx + qy * qy + qz * qz + qw * qw;
        if (qn > 0.0f) {
            const float inv = 1.0f / sqrtf(qn);
            qx = qx * inv; qy = qy * inv; qz = qz * inv; qw = qw * inv;
        } else {
            qx = 0.0f; qy = 0.0f; qz = 0.0f; qw = 1.0f;
        }
    }

    // The rotation and R S, the same for every view (splat_project.slang:526-532).
    const float r00 = 1.0f - 2.0f * (qy * qy + qz * qz), r01 = 2.0f * (qx * qy - qw * qz),
                r02 = 2.0f * (qx * qz + qw * qy);
    const float r10 = 2.0f * (qx * qy + qw * qz), r11 = 1.0f - 2.0f * (qx * qx + qz * qz),
                r12 = 2.0f * (qy * qz - qw * qx);
    const float r20 = 2.0f * (qx * qz - qw * qy), r21 = 2.0f * (qy * qz + qw * qx),
                r22 = 1.0f - 2.0f * (qx * qx + qy * qy);
    const float m00 = r00 * sx, m01 = r01 * sy, m02 = r02 * sz;
    const float m10 = r10 * sx, m11 = r11 * sy, m12 = r12 * sz;
    const float m20 = r20 * sx, m21 = r21 * sy, m22 = r22 * sz;
    // kVec: 16 coefficients a channel in 16-byte aligned rows, read once for
    // every view.
    float shc[kVec ? 48 : 1];
    if constexpr (kVec) {
        load_sh16<kHalf>(a.sh, a.sh_half, i, shc);
    }

    for (uint32_t v = 0; v < nviews; ++v) {
        const DView& V = views[v];
        const uint32_t slot = v * a.slot_stride + a.slot_base + i;
        const float* m = V.m;
        const float tx = m[0] * px + m[1] * py + m[2] * pz + m[3];
        const float ty = m[4] * px + m[5] * py + m[6] * pz + m[7];
        const float tz = m[8] * px + m[9] * py + m[10] * pz + m[11];
        if (!(tz > V.near_z) || tz > V.far_z) {
            cull_early(slot, ATHX_CULL_DEPTH, touched, keys);
            continue;
        }
        float jx0, jx2, jy1, jy2, cx, cy;
        uint32_t flags = 0;  // the branches taken, kept for the adjoint (kFlag*)
        if (V.orthographic) {
            jx0 = V.focal_x; jx2 = 0.0f;
            jy1 = V.focal_y; jy2 = 0.0f;
            cx = V.focal_x * tx + V.centre_x;
            cy = V.focal_y * ty + V.centre_y;
        } else {
            const float invz = 1.0f / tz;
            cx = V.focal_x * tx * invz + V.centre_x;
            cy = V.focal_y * ty * invz + V.centre_y;
            // The Jacobian at a point clamped to 1.3x the frustum (splat_project.slang:543-557).
            const float limlox = V.lim[0], limhix = V.lim[1], limloy = V.lim[2], limhiy = V.lim[3];
            const float ux = tx * invz, uy = ty * invz;
            flags |= (ux < limlox ? kFlagClampXLo : 0u) | (ux > limhix ? kFlagClampXHi : 0u) |
                     (uy < limloy ? kFlagClampYLo : 0u) | (uy > limhiy ? kFlagClampYHi : 0u);
            const float txc = clamp_(ux, limlox, limhix) * tz;
            const float tyc = clamp_(uy, limloy, limhiy) * tz;
            jx0 = V.focal_x * invz;
            jx2 = -V.focal_x * txc * invz * invz;
            jy1 = V.focal_y * invz;
            jy2 = -V.focal_y * tyc * invz * invz;
        }
        // T = J W, then (T M)(T M)^T (splat_project.slang:558-565).
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
        float ca = ax0 * ax0 + ax1 * ax1 + ax2 * ax2;
        const float cb = ax0 * ay0 + ax1 * ay1 + ax2 * ay2;
        float cc = ay0 * ay0 + ay1 * ay1 + ay2 * ay2;
        // Dilation and the Mip-Splatting compensation (splat_project.slang:651-667, 734-751).
        const float det_before = ca * cc - cb * cb;
        ca = ca + 0.3f;
        cc = cc + 0.3f;
        const float det = ca * cc - cb * cb;
        if (!(det > 1e-12f)) {
            cull_early(slot, ATHX_CULL_DEGENERATE, touched, keys);
            continue;
        }
        const float filtered = sqrtf(fmax_(det_before, 0.0f) / det);
        const float paid = a.antialias ? filtered : 1.0f;
        const float spread = paid * opacity;
        const float alpha = spread >= 1.0f / 255.0f ? spread + (1.0f - paid) * (1.0f / 255.0f) : spread;
        flags |= spread >= 1.0f / 255.0f ? kFlagPaidBack : 0u;
        if (!(alpha >= 1.0f / 255.0f)) {
            cull_early(slot, ATHX_CULL_FAINT, touched, keys);
            continue;
        }
        const float bdet = ca * cc - cb * cb;
        const float ka = cc / bdet, kb = -cb / bdet, kc = ca / bdet;
        const float cutoff = cutoff_power(alpha);
        uint32_t x0, y0, x1, y1;
        if (!tile_rect(V.width, V.height, a.tile, cx, cy, ka, kb, kc, cutoff, x0, y0, x1, y1)) {
            cull_early(slot, ATHX_CULL_OFFSCREEN, touched, keys);
            continue;
        }
        const uint32_t h = y1 - y0 + 1;
        // The reference counts the touched tiles here and culls a zero count
        // (splat_project.slang:764-776); what the frame needs is only whether
        // the count is zero, and the tile under the (clamped) centre nearly
        // always settles that: it is touched whenever the centre lies in its
        // box of pixel centres. The rest are settled by settle_kernel. The
        // slot's word is its strip entries: one a (row, strip) of its
        // rectangle.
        const uint32_t tshift = (uint32_t)__builtin_ctz(a.tile);
        const uint32_t ctx = min(max((uint32_t)clamp_(cx, 0.0f, (float)(V.width - 1u)) >> tshift, x0), x1);
        const uint32_t cty = min(max((uint32_t)clamp_(cy, 0.0f, (float)(V.height - 1u)) >> tshift, y0), y1);
        const bool settled = tile_touched(V.width, V.height, a.tile, cx, cy, ka, kb, kc, cutoff, ctx, cty);
        const uint32_t entries = h * (x1 / kStripTiles - x0 / kStripTiles + 1u);
        // Colour from this eye, in the cloud's space (splat_project.slang:781-785).
        const float dx = px - V.eye[0], dy = py - V.eye[1], dz = pz - V.eye[2];
        const float dinv = 1.0f / sqrtf(dx * dx + dy * dy + dz * dz);
        const float ex = dx * dinv, ey = dy * dinv, ez = dz * dinv;
        const size_t row = (size_t)i * a.sh_coeffs * 3;
        float col[3];
#pragma unroll
        for (uint32_t ch = 0; ch < 3; ++ch) {
            float base, rest;
            if constexpr (kPacked) {
                // sh.slang:15-22: rest coefficient k - 1 of this channel is
                // half (k - 1) * 3 + ch of the Gaussian's words.
                const uint32_t* words = a.packed_sh + (size_t)i * a.packed_words;
                base = stored[ch];
                rest = sh_rest(
                    [&](uint32_t k) {
                        const uint32_t h = (k - 1) * 3 + ch;
                        const uint32_t word = words[h >> 1];
                        return half_bits_to_float((uint16_t)((h & 1u) ? word >> 16 : word & 0xFFFFu));
                    },
                    a.sh_degree, ex, ey, ez);
            } else if constexpr (kVec) {
                base = 0.5f + kSH0 * shc[ch];
                rest = sh_rest([&](uint32_t k) { return shc[k * 3 + ch]; }, a.sh_degree, ex, ey, ez);
            } else {
                base = 0.5f + kSH0 * sh_at<kHalf>(a.sh, a.sh_half, row + ch);
                rest = sh_rest([&](uint32_t k) { return sh_at<kHalf>(a.sh, a.sh_half, row + ch + k * 3); },
                               a.sh_degree, ex, ey, ez);
            }
            const float cv = base + rest;
            col[ch] = a.linear ? fmax_(cv, 0.0f) : srgb_to_linear(cv);
            flags |= (cv > 0.0f ? kFlagColourR : 0u) << ch;
            flags |= (!a.linear && fmax_(cv, 0.0f) > 0.04045f ? kFlagSrgbR : 0u) << ch;
        }
        if constexpr (kRelit) {
            // The relit branch (splat_project.slang:799-976): the evaluated
            // colour is the surface's albedo unless one is given.
            const float R[3][3] = {{r00, r01, r02}, {r10, r11, r12}, {r20, r21, r22}};
            const float3 lit = relit_colour(a.rl, i, R, sx, sy, sz, px, py, pz, V.eye,
                                            make_float3(col[0], col[1], col[2]));
            col[0] = lit.x;
            col[1] = lit.y;
            col[2] = lit.z;
            flags |= kFlagRelit;
        }
        float4* rec = reinterpret_cast<float4*>(records + (size_t)slot * kRecordFloats);
        rec[0] = make_float4(cx, cy, tz, cutoff);
        rec[1] = make_float4(ka, kb, kc, alpha);
        rec[2] = make_float4(col[0], col[1], col[2], __builtin_bit_cast(float, flags));
        keys[slot] = __builtin_bit_cast(uint32_t, tz) >> 7;
        touched[slot] = entries | (settled ? 0u : kUnsettled);
        if (a.slot_ids) {
            // splat_project.slang:1116-1118
            a.slot_ids[slot] = a.crypto_ids ? a.crypto_ids[i] : 0u;
        }
    }
}

__global__ void fill_u32_kernel(uint32_t* p, uint32_t value, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        p[i] = value;
    }
}

// k / w and k % w for k < 64, 1 <= w <= 64: floor(k rcp(w) + 1e-3) is exact
// there (rcp's relative error, under 2^-22, moves k/w by at most 1.6e-5,
// and a k/w that is not whole is at least 1/64 below the next integer).
__device__ inline uint32_t small_row(uint32_t k, uint32_t w) {
    return (uint32_t)((float)k * __builtin_amdgcn_rcpf((float)w) + 1.0e-3f);
}

// Big footprints (rectangles of more than kBigRect tiles) go kBigPerWave to
// a wavefront, in the order the listing scan put them (slot order): lane j
// reads footprint j's record and rectangle, then the wavefront walks them
// one at a time with the footprint's fields read from its lane. The scan
// gave each footprint its first hit word, so nothing is claimed here.
//
// A step covers a run of the rectangle's tiles in row-major order, one tile
// a lane and the ballot of the hits one word: whole rows, 64 / w of them,
// when the rectangle is at most 64 tiles wide (lane = row in the step * w +
// column, so a lane keeps its column for the whole footprint), or 64
// columns of one row when it is wider. The hit words are the emission's.
//
// The tile test is min_power_over_box's, taken apart: what it computes
// from a tile's column alone (the box's x edges, the y that minimises the
// power on each, the x parts of the power there) and from its row alone
// are each computed once for the lane's column or row, and a tile combines
// them in the order the whole function does, so the hits are its bits.
constexpr uint32_t kBigThreads = 256;
constexpr uint32_t kBigBlocks = 4096;
constexpr uint32_t kBigPerWave = 16;

struct BigFootprint {
    float cx, cy, a, b, c, cutoff, ia, ic;
    uint32_t view, x0, y0, w, h;
};

__device__ inline BigFootprint big_footprint(const DView* views, uint32_t n, uint32_t T, const float* records,
                                             uint32_t slot) {
    BigFootprint f;
    f.view = slot / n;
    const DView& V = views[f.view];
    const float4* rec = reinterpret_cast<const float4*>(records + (size_t)slot * kRecordFloats);
    const float4 screen = rec[0], conic = rec[1];
    f.cx = screen.x;
    f.cy = screen.y;
    f.cutoff = screen.w;
    f.a = conic.x;
    f.b = conic.y;
    f.c = conic.z;
    f.ia = recip_of(f.a);
    f.ic = recip_of(f.c);
    uint32_t x1, y1;
    tile_rect(V.width, V.height, T, f.cx, f.cy, f.a, f.b, f.c, f.cutoff, f.x0, f.y0, x1, y1);
    f.w = x1 - f.x0 + 1;
    f.h = y1 - f.y0 + 1;
    return f;
}

__device__ inline uint32_t lane_u(uint32_t v, uint32_t j) { return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)j); }
__device__ inline float lane_f(float v, uint32_t j) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), (int)j));
}
__device__ inline uint64_t lane_u64(uint64_t v, uint32_t j) {
    return (uint64_t)lane_u((uint32_t)v, j) | (uint64_t)lane_u((uint32_t)(v >> 32), j) << 32;
}

// Footprint j of the wavefront's 64, read from lane j (j is wave-uniform, so
// the fields land in scalar registers).
__device__ inline BigFootprint big_of_lane(const BigFootprint& f, uint32_t j) {
    BigFootprint g;
    g.cx = lane_f(f.cx, j);
    g.cy = lane_f(f.cy, j);
    g.a = lane_f(f.a, j);
    g.b = lane_f(f.b, j);
    g.c = lane_f(f.c, j);
    g.cutoff = lane_f(f.cutoff, j);
    g.ia = lane_f(f.ia, j);
    g.ic = lane_f(f.ic, j);
    g.view = lane_u(f.view, j);
    g.x0 = lane_u(f.x0, j);
    g.y0 = lane_u(f.y0, j);
    g.w = lane_u(f.w, j);
    g.h = lane_u(f.h, j);
    return g;
}

// One axis of a tile's box (frame.slang:374-379) and what min_power_over_box
// (:354-371) computes from it alone. Along x: the box's edges lo, hi (pixel
// centres relative to the splat's), the y minimising the power on each edge
// before its clamp, a x^2 and 2b x. Along y: the edges, the x minimising
// the power on each, and c y^2.
struct BoxAxis {
    float lo, hi, q_lo, q_hi, s_lo, s_hi, t_lo, t_hi;
};

__device__ inline BoxAxis box_axis(uint32_t t, uint32_t T, uint32_t extent, float centre, float across, float b,
                                   float along, float b2) {
    BoxAxis r;
    const uint32_t e = __umul24(t + 1, T) < extent ? __umul24(t + 1, T) : extent;
    r.lo = ((float)__umul24(t, T) + 0.5f) - centre;
    r.hi = ((float)(e - 1) + 0.5f) - centre;
    // The other coordinate's minimiser on each edge: ((-b) * edge) / across
    // (read only when across > 1e-12).
    r.q_lo = ((-b) * r.lo) / across;
    r.q_hi = ((-b) * r.hi) / across;
    r.s_lo = along * r.lo * r.lo;
    r.s_hi = along * r.hi * r.hi;
    r.t_lo = b2 * r.lo;
    r.t_hi = b2 * r.hi;
    return r;
}

// min_power_over_box(a, b, c, box) <= cutoff, from the tile's column values
// X = box_axis(tx, ..., cx, c, b, a, 2b) and row values Y = box_axis(ty,
// ..., cy, a, b, c, 2b): the same operations in the same order.
__device__ inline bool box_hit(const BigFootprint& g, float b2, const BoxAxis& X, const BoxAxis& Y) {
    if (X.lo <= 0.0f && X.hi >= 0.0f && Y.lo <= 0.0f && Y.hi >= 0.0f) {
        return true;  // a power of 0; the cut power is above 0
    }
    float best = 3.4e38f;
    {
        const float y = g.c > 1e-12f ? clamp_(X.q_lo, Y.lo, Y.hi) : Y.lo;
        best = fmin_(best, X.s_lo + X.t_lo * y + g.c * y * y);
    }
    {
        const float y = g.c > 1e-12f ? clamp_(X.q_hi, Y.lo, Y.hi) : Y.lo;
        best = fmin_(best, X.s_hi + X.t_hi * y + g.c * y * y);
    }
    {
        const float x = g.a > 1e-12f ? clamp_(Y.q_lo, X.lo, X.hi) : X.lo;
        best = fmin_(best, g.a * x * x + b2 * x * Y.lo + Y.s_lo);
    }
    {
        const float x = g.a > 1e-12f ? clamp_(Y.q_hi, X.lo, X.hi) : X.lo;
        best = fmin_(best, g.a * x * x + b2 * x * Y.hi + Y.s_hi);
    }
    return best <= g.cutoff;
}

__device__ inline BoxAxis column_of(const BigFootprint& g, const DView& V, uint32_t T, float b2, uint32_t tx) {
    return box_axis(tx, T, V.width, g.cx, g.c, g.b, g.a, b2);
}
__device__ inline BoxAxis row_of(const BigFootprint& g, const DView& V, uint32_t T, float b2, uint32_t ty) {
    return box_axis(ty, T, V.height, g.cy, g.a, g.b, g.c, b2);
}

// Walks footprint g's steps, calling step(s, hit, tile) with this lane's tile
// (column, row within the rectangle) and whether the lane has one.
template <typename Step>
__device__ inline void walk_big(const BigFootprint& g, const DView& V, uint32_t T, uint32_t lane, Step&& step) {
    const float b2 = 2.0f * g.b;
    if (g.w <= 64) {
        const uint32_t rows = 64 / g.w;
        const uint32_t dr = small_row(lane, g.w), col = lane - dr * g.w;
        const bool lane_on = dr < rows;
        const BoxAxis X = column_of(g, V, T, b2, g.x0 + col);
        const uint32_t steps = div_up(g.h, rows);
        for (uint32_t s = 0; s < steps; ++s) {
            const uint32_t r = s * rows + dr;
            const bool on = lane_on && r < g.h;
            const bool hit = on && box_hit(g, b2, X, row_of(g, V, T, b2, g.y0 + r));
            step(s, on, hit, col, r);
        }
    } else {
        const uint32_t chunks = div_up(g.w, 64);
        uint32_t s = 0;
        for (uint32_t r = 0; r < g.h; ++r) {
            const BoxAxis Y = row_of(g, V, T, b2, g.y0 + r);
            for (uint32_t ch = 0; ch < chunks; ++ch, ++s) {
                const uint32_t col = ch * 64 + lane;
                const bool on = col < g.w;
                const bool hit = on && box_hit(g, b2, column_of(g, V, T, b2, g.x0 + col), Y);
                step(s, on, hit, col, r);
            }
        }
    }
}

__device__ inline uint32_t wave_inclusive_sum(uint32_t v, uint32_t lane) {
#pragma unroll
    for (uint32_t d = 1; d < 64; d <<= 1) {
        const uint32_t t = __shfl_up(v, d, 64);
        v += lane >= d ? t : 0u;
    }
    return v;
}

// Inclusive max over lanes 0 .. L of a wavefront, by DPP: within each row
// of 16 lanes shifts right by 1, 2, 4 and 8 (lanes shifted in from outside
// the row read 0), then lane 15 of rows 0 and 2 into rows 1 and 3, then
// lane 31 into rows 2 and 3. Every lane must be active.
__device__ inline uint32_t wave_inclusive_max(uint32_t v) {
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xF, 0xF, false));  // row_shr 1
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xF, 0xF, false));  // row_shr 2
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xF, 0xF, false));  // row_shr 4
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xF, 0xF, false));  // row_shr 8
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xA, 0xF, false));  // row_bcast 15
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xC, 0xF, false));  // row_bcast 31
    return v;
}

// Inclusive sum over lanes 0 .. L, by DPP as wave_inclusive_max. Every lane
// must be active.
__device__ inline uint32_t wave_inclusive_sum_dpp(uint32_t v) {
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xF, 0xF, false);  // row_shr 1
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xF, 0xF, false);  // row_shr 2
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xF, 0xF, false);  // row_shr 4
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xF, 0xF, false);  // row_shr 8
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xA, 0xF, false);  // row_bcast 15
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xC, 0xF, false);  // row_bcast 31
    return v;
}

__global__ void __launch_bounds__(kBigThreads) count_big_kernel(const DView* __restrict__ views, uint32_t n,
                                                                uint32_t T, const float* __restrict__ records,
                                                                Footprints fp, uint32_t* __restrict__ touched,
                                                                uint32_t* __restrict__ keys) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t waves = gridDim.x * (kBigThreads / 64);
    const uint32_t total = (uint32_t)(*fp.big_total >> 32);
    for (uint32_t c0 = (blockIdx.x * (kBigThreads / 64) + (threadIdx.x >> 6)) * kBigPerWave; c0 < total;
         c0 += waves * kBigPerWave) {
        const uint32_t b = c0 + lane;
        const bool mine = lane < kBigPerWave && b < total;
        const uint32_t slot = mine ? fp.big_list[b] : 0u;
        BigFootprint f = {};
        uint64_t at = 0;
        if (mine) {
            f = big_footprint(views, n, T, records, slot);
            at = fp.big_mask_base[b];
        }
        // Words past the buffer are not kept: the emission tests those
        // footprints' tiles again.
        const bool keep = mine && at + big_steps(f.w, f.h) <= fp.big_mask_cap;
        const uint32_t here = total - c0 < kBigPerWave ? total - c0 : kBigPerWave;
        uint32_t my_count = 0;
        for (uint32_t j = 0; j < here; ++j) {
            const BigFootprint g = big_of_lane(f, j);
            const DView& V = views[g.view];
            const bool keep_j = lane_u(keep ? 1u : 0u, j) != 0;
            const uint64_t at_j = lane_u64(at, j);
            const uint32_t steps = big_steps(g.w, g.h);
            uint32_t count = 0;
            uint64_t run = 0;  // lane r holds the hit word of step r of the current run of 64 steps
            walk_big(g, V, T, lane, [&](uint32_t s, bool, bool hit, uint32_t, uint32_t) {
                const uint64_t hits = __ballot(hit);
                count += (uint32_t)__popcll(hits);
                const uint32_t r = s & 63u;
                run = lane == r ? hits : run;
                if (keep_j && (r == 63u || s + 1 == steps) && lane <= r) {
                    fp.big_masks[at_j + (s & ~63u) + lane] = run;
                }
            });
            my_count = lane == j ? count : my_count;
        }
        if (mine) {
            touched[slot] = my_count;
            if (my_count == 0) {
                keys[slot] = kCulledKey | ATHX_CULL_NO_TILE;
            }
        }
    }
}

// Small footprints (rectangles of up to kBigRect tiles), 64 slots to a
// wavefront, their tiles spread over the lanes: lane j holds slot c0 + j
// and its rectangle's area A_j (0 for a culled or a big slot), and the
// wavefront walks the A_0 + ... + A_63 tiles 64 at a time, lane L taking
// tile g = g0 + L of the concatenation -- tile k = g - start_j of the
// footprint j whose run holds g (a binary search over the runs' ends in
// LDS). Every lane tests a tile on every step but the last, whatever the
// footprints' sizes, where a thread walking its own rectangle would hold
// its wavefront to the largest of 64.
constexpr uint32_t kSmallThreads = 256;

// The footprint whose run holds g: the first j with end[j] > g.
__device__ inline uint32_t run_of(const uint32_t* ends, uint32_t g) {
    uint32_t lo = 0, hi = 63;
#pragma unroll
    for (uint32_t step = 0; step < 6; ++step) {
        const uint32_t mid = (lo + hi) >> 1;
        if (ends[mid] > g) {
            hi = mid;
        } else {
            lo = mid + 1;
        }
    }
    return lo;
}

// The exact counts' rectangles: each slot's tile rectangle, recomputed
// from its record as the projection computed it (x0 | y0 << 16, w | h <<
// 16), and (0, 0) for a slot culled before its tiles are counted.
__global__ void __launch_bounds__(256) rect_kernel(const DView* __restrict__ views, uint32_t n, uint32_t slots,
                                                   uint32_t T, const float* __restrict__ records,
                                                   const uint32_t* __restrict__ keys, uint2* __restrict__ rects) {
    const uint32_t slot = blockIdx.x * 256u + threadIdx.x;
    if (slot >= slots) {
        return;
    }
    const uint32_t key = keys[slot];
    uint2 r = make_uint2(0u, 0u);
    if (key < kCulledKey || (key & ~kCulledKey) == ATHX_CULL_NO_TILE) {
        const DView& V = views[slot / n];
        const float4* rec = reinterpret_cast<const float4*>(records + (size_t)slot * kRecordFloats);
        const float4 screen = rec[0], conic = rec[1];
        uint32_t x0, y0, x1, y1;
        tile_rect(V.width, V.height, T, screen.x, screen.y, conic.x, conic.y, conic.z, screen.w, x0, y0, x1, y1);
        r = make_uint2(x0 | (y0 << 16), (x1 - x0 + 1u) | ((y1 - y0 + 1u) << 16));
    }
    rects[slot] = r;
}

__global__ void __launch_bounds__(kSmallThreads) count_small_kernel(const DView* __restrict__ views, uint32_t n,
                                                                    uint32_t slots, uint32_t T,
                                                                    const float* __restrict__ records, Footprints fp,
                                                                    uint32_t* __restrict__ touched,
                                                                    uint32_t* __restrict__ keys) {
    __shared__ uint32_t s_end[kSmallThreads];
    __shared__ float4 s_geo[kSmallThreads];     // centre x, y, cut power, conic a
    __shared__ float4 s_shape[kSmallThreads];   // conic b, c, x0 | y0 << 16, w | view << 16 (bits)
    __shared__ uint64_t s_mask[kSmallThreads];
    const uint32_t lane = threadIdx.x & 63u, wbase = threadIdx.x & ~63u;
    const uint32_t slot = blockIdx.x * kSmallThreads + threadIdx.x;
    const uint2 rect = slot < slots ? fp.rects[slot] : make_uint2(0u, 0u);
    const uint32_t w = rect.y & 0xFFFFu, area = w * (rect.y >> 16);
    const uint32_t mine = area <= kBigRect ? area : 0u;
    if (mine != 0) {
        const float4* rec = reinterpret_cast<const float4*>(records + (size_t)slot * kRecordFloats);
        const float4 r0 = rec[0], r1 = rec[1];
        s_geo[threadIdx.x] = make_float4(r0.x, r0.y, r0.w, r1.x);
        s_shape[threadIdx.x] = make_float4(r1.y, r1.z, __builtin_bit_cast(float, rect.x),
                                           __builtin_bit_cast(float, w | (slot / n) << 16));
    }
    const uint32_t end = wave_inclusive_sum(mine, lane);
    s_end[threadIdx.x] = end;
    s_mask[threadIdx.x] = 0;
    const uint32_t total = lane_u(end, 63);
    __syncthreads();
    const uint32_t* ends = s_end + wbase;
    for (uint32_t g0 = 0; g0 < total; g0 += 64) {
        const uint32_t g = g0 + lane;
        const bool on = g < total;
        const uint32_t j = on ? run_of(ends, g) : 0u;
        const uint32_t k = on ? g - (j != 0 ? ends[j - 1] : 0u) : 0u;
        bool hit = false;
        if (on) {
            const float4 geo = s_geo[wbase + j], shape = s_shape[wbase + j];
            const uint32_t corner = __builtin_bit_cast(uint32_t, shape.z);
            const uint32_t wv = __builtin_bit_cast(uint32_t, shape.w);
            const uint32_t wj = wv & 0xFFFFu;
            const DView& V = views[wv >> 16];
            const uint32_t q = small_row(k, wj);
            hit = tile_touched(V.width, V.height, T, geo.x, geo.y, geo.w, shape.x, shape.y, geo.z,
                               (corner & 0xFFFFu) + (k - q * wj), (corner >> 16) + q);
        }
        // The first lane of each footprint's run in this step folds the
        // run's hits into the footprint's mask.
        const uint64_t hits = __ballot(hit);
        const bool leader = on && (k == 0 || lane == 0);
        const uint64_t leaders = __ballot(leader);
        if (leader) {
            const uint64_t after = leaders & ~((2ull << lane) - 1ull);
            const uint32_t len = (after != 0 ? (uint32_t)__builtin_ctzll(after) : 64u) - lane;
            const uint64_t run = (hits >> lane) & (len >= 64 ? ~0ull : (1ull << len) - 1ull);
            s_mask[wbase + j] |= run << k;
        }
    }
    __syncthreads();
    if (mine != 0) {
        const uint64_t mask = s_mask[threadIdx.x];
        const uint32_t count = (uint32_t)__popcll(mask);
        fp.masks[slot] = mask;
        touched[slot] = count;
        if (count == 0) {
            keys[slot] = kCulledKey | ATHX_CULL_NO_TILE;
        }
    }
}

// Where each tile's pairs start and end in the sorted keys, by binary search
// on the tile bits: splat_tiles_clear.slang:8-17 and splat_ranges.slang:11-25
// give a tile with no pairs (0, 0), and so does this.
__device__ inline uint32_t lower_bound_tile(const uint32_t* tiles, uint32_t n, uint64_t tile, uint32_t key_mask) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = lo + ((hi - lo) >> 1);
        if ((uint64_t)(tiles[mid] & key_mask) < tile) {
            lo = mid + 1;
        } else {
            hi = mid;
        }
    }
    return lo;
}

__global__ void __launch_bounds__(256) ranges_kernel(const uint32_t* __restrict__ keys,
                                                     const uint64_t* __restrict__ n_dev, uint32_t cap, uint32_t tiles,
                                                     uint32_t key_mask,
                                                     uint32_t* __restrict__ ranges) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= tiles) {
        return;
    }
    const uint32_t n = *n_dev < cap ? (uint32_t)*n_dev : cap;
    const uint32_t start = lower_bound_tile(keys, n, t, key_mask);
    const uint32_t end = lower_bound_tile(keys, n, (uint64_t)t + 1, key_mask);
    ranges[2 * t] = start < end ? start : 0u;
    ranges[2 * t + 1] = start < end ? end : 0u;
}

// The footprints the projection left unsettled (the tile under the centre
// not touched): a wavefront takes each in turn and tests its rectangle's
// tiles 64 a step, row-major, until one is touched. None touched is the
// reference's zero count: culled, no tile (splat_project.slang:764-776).
// The same pass counts the drawn slots of each block of kCompactTile slots
// into sums[block], the compaction's first step.
// A workgroup a block of slots: 1024 lanes in a small frame, so that it
// still spreads its unsettled footprints over many wavefronts; 256 in a
// large one (kSettleWide slots and up), whose workgroups are then fewer
// wavefronts to launch for the same slots.
constexpr uint32_t kSettleWide = 1u << 24;

template <uint32_t kSettleThreads>
__global__ void __launch_bounds__(kSettleThreads) settle_count_kernel(const DView* __restrict__ views, uint32_t n, uint32_t T,
                                                           uint32_t slots, const float* __restrict__ records,
                                                           uint32_t* __restrict__ touched, uint32_t* __restrict__ keys,
                                                           uint64_t* __restrict__ sums) {
    __shared__ uint32_t s_wave[kSettleThreads / 64];
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t first_block = (uint64_t)blockIdx.x * kCompactTile;
    constexpr uint32_t kPer = kCompactTile / kSettleThreads;
    uint32_t words[kPer];
#pragma unroll
    for (uint32_t item = 0; item < kPer; ++item) {
        const uint64_t i = first_block + (uint64_t)item * kSettleThreads + threadIdx.x;
        words[item] = i < slots ? touched[i] : 0u;
    }
    uint32_t count = 0;
#pragma unroll
    for (uint32_t item = 0; item < kPer; ++item) {
        const uint64_t i = first_block + (uint64_t)item * kSettleThreads + threadIdx.x;
        uint32_t word = words[item];
        uint64_t todo = __ballot((word & kUnsettled) != 0u);
        const uint32_t first = (uint32_t)(i - lane);
        while (todo != 0) {
            const uint32_t j = (uint32_t)__builtin_ctzll(todo);
            todo &= todo - 1;
            const uint32_t slot = first + j;
            const BigFootprint g = big_footprint(views, n, T, records, slot);
            const DView& V = views[g.view];
            const uint32_t area = g.w * g.h;
            bool any = false;
            for (uint32_t k0 = 0; k0 < area && !any; k0 += 64) {
                const uint32_t k = k0 + lane;
                const uint32_t r = k / g.w;
                const bool hit = k < area && tile_touched(V.width, V.height, T, g.cx, g.cy, g.a, g.b, g.c,
                                                          g.cutoff, g.x0 + (k - r * g.w), g.y0 + r);
                any = __ballot(hit) != 0;
            }
            if (lane == j) {
                word = any ? (word & ~kUnsettled) : 0u;
                touched[slot] = word;
                if (!any) {
                    keys[slot] = kCulledKey | ATHX_CULL_NO_TILE;
                }
            }
        }
        count += (uint32_t)__popcll(__ballot(word != 0u));
    }
    if (lane == 0) {
        s_wave[threadIdx.x >> 6] = count;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t t = 0;
        for (uint32_t w = 0; w < kSettleThreads / 64; ++w) {
            t += s_wave[w];
        }
        sums[blockIdx.x] = t;
    }
}

// Deciding most of an entry's tiles without the float tile test. The test
// (box_hit) evaluates a x^2 + 2 b x y + c y^2 at points of the tile's box
// whose coordinates are the box edges rounded relative to the centre, in
// about ten float operations; its result differs from the exact least power
// over the exact box by a few units of 2^-24 M, M = a X^2 + 2 |b| X Y + c Y^2
// with X, Y the box's largest distances from the centre. The exact least
// power over a tile box is at most C exactly when the box's x-range meets
// the x-extent of the ellipse q <= C sliced by the box's row band (the slice
// is convex and spans the band's height). That extent is computed here in
// float, once an entry, at C = cutoff -+ kSlack M: a tile meeting the inner
// one is touched by the float test, a tile missing the outer one is not, and
// only the tiles in between are tested. The extent's own rounding is of the
// same order in power (where its square root is ill-conditioned, at the
// ellipse's top and bottom, the power's x-gradient vanishes), and the
// determinant is taken in FP64, where a thin ellipse's does not cancel;
// kSlack = 256 * 2^-24 covers both with a wide margin.
constexpr float kSlack = 256.0f * 0x1p-24f;

// [xa, xb]: the x-extent of {(x, y) : y0 <= y <= y1, q(x, y) <= C}
// (coordinates relative to the centre; sa = sqrt(a / det), sc = sqrt(c /
// det)); false when it is empty.
__device__ inline bool band_extent(float a, float b, float ia, float ic, float det, float sa, float sc, float C,
                                   float y0, float y1, float& xa, float& xb) {
    if (!(C > 0.0f)) {
        return false;
    }
    // The ellipse's rows: |y| <= sqrt(a C / det); its widest points (x = -+xg)
    // at y = +-b xg / c.
    const float sC = __builtin_amdgcn_sqrtf(C);
    const float ym = sC * sa, xg = sC * sc;
    const float ya = fmax_(y0, -ym), yb = fmin_(y1, ym);
    if (ya > yb) {
        return false;
    }
    // Right end r(y) = (-b y + sqrt(a C - det y^2)) / a is concave, the left
    // end convex: each is extreme at its widest point's row, clamped to the band.
    const float yr = clamp_((-b * xg) * ic, ya, yb);
    const float yl = clamp_((b * xg) * ic, ya, yb);
    const float aC = a * C;
    xb = (-b * yr + __builtin_amdgcn_sqrtf(fmax_(aC - det * yr * yr, 0.0f))) * ia;
    xa = (-b * yl - __builtin_amdgcn_sqrtf(fmax_(aC - det * yl * yl, 0.0f))) * ia;
    return true;
}

// The strip columns c0 .. c1 whose box x-range meets [xa, xb] (relative to
// the centre cx): bit j for column tx0 + j. A column's box is
// [tx T + 0.5, min((tx + 1) T, W) - 0.5] in pixel centres, and both its
// edges, rounded relative to cx, are nondecreasing in the column: the
// columns meeting the extent run from the first whose right edge reaches xa
// to the last whose left edge is at most xb. Each is found from its
// estimate (within a column of it) by stepping while a neighbour still
// qualifies.
__device__ inline uint32_t columns_meeting(float xa, float xb, float cx, uint32_t T, uint32_t W, uint32_t tx0,
                                          uint32_t c0, uint32_t c1) {
    const auto lo_of = [&](int tx) { return ((float)__umul24((uint32_t)tx, T) + 0.5f) - cx; };
    const auto hi_of = [&](int tx) {
        const uint32_t e = __umul24((uint32_t)tx + 1u, T) < W ? __umul24((uint32_t)tx + 1u, T) : W;
        return ((float)(e - 1u) + 0.5f) - cx;
    };
    const int lo = (int)c0, hi = (int)c1;
    const float inv_t = 1.0f / (float)T;
    // First column whose right edge reaches xa (hi + 1 when none).
    int first = (int)ceilf(clamp_((xa + cx + 0.5f) * inv_t - 1.0f, (float)lo, (float)(hi + 1)));
#pragma clang loop unroll(disable) interleave(disable) vectorize(disable)
    while (first > lo && hi_of(first - 1) >= xa) {
        --first;
    }
#pragma clang loop unroll(disable) interleave(disable) vectorize(disable)
    while (first <= hi && hi_of(first) < xa) {
        ++first;
    }
    // Last column whose left edge is at most xb (lo - 1 when none).
    int last = (int)floorf(clamp_((xb + cx - 0.5f) * inv_t, (float)(lo - 1), (float)hi));
#pragma clang loop unroll(disable) interleave(disable) vectorize(disable)
    while (last < hi && lo_of(last + 1) <= xb) {
        ++last;
    }
#pragma clang loop unroll(disable) interleave(disable) vectorize(disable)
    while (last >= lo && lo_of(last) > xb) {
        --last;
    }
    if (first > last) {
        return 0u;
    }
    return ((1u << (uint32_t)(last - first + 1)) - 1u) << ((uint32_t)first - tx0);
}

// An entry's strip columns c0 .. c1 (tile row with band terms Y): the
// columns surely touched, and those the float test must decide.
struct Decided {
    uint32_t sure, test;
};

// A footprint's slice terms: its conic's determinant a c - b^2 (taken in
// FP64, where a thin ellipse's does not cancel, and rounded once),
// sqrt(a / det) and sqrt(c / det); det = 0 marks a conic the extents do
// not handle (not positive definite, or a determinant outside the normal
// float range), whose columns all go to the float test.
struct SliceTerms {
    float det, sa, sc;
};

__device__ inline SliceTerms slice_terms(float a, float b, float c) {
    SliceTerms t{0.0f, 0.0f, 0.0f};
    const double detd = (double)a * (double)c - (double)b * (double)b;
    const float det = (float)detd;
    if (a > 1e-9f && c > 1e-9f && det >= 0x1p-126f && det <= 0x1p126f) {
        const float inv_det = __builtin_amdgcn_rcpf(det);
        t.det = det;
        t.sa = __builtin_amdgcn_sqrtf(a * inv_det);
        t.sc = __builtin_amdgcn_sqrtf(c * inv_det);
    }
    return t;
}

__device__ inline Decided decide_columns(const BigFootprint& g, const SliceTerms& S, const BoxAxis& Y, uint32_t T,
                                         uint32_t W, uint32_t tx0, uint32_t c0, uint32_t c1) {
    Decided d{0u, 0u};
    if (c1 < c0) {
        return d;
    }
    if (S.det == 0.0f) {
        d.test = ((1u << (c1 - c0 + 1u)) - 1u) << (c0 - tx0);
        return d;
    }
    const uint32_t ex = __umul24(c1 + 1u, T) < W ? __umul24(c1 + 1u, T) : W;
    const float X = fmax_(fabsf(((float)__umul24(c0, T) + 0.5f) - g.cx), fabsf(((float)(ex - 1u) + 0.5f) - g.cx));
    const float Yb = fmax_(fabsf(Y.lo), fabsf(Y.hi));
    const float slack = kSlack * (g.a * X * X + 2.0f * fabsf(g.b) * X * Yb + g.c * Yb * Yb);
    float xa = 0.0f, xb = 0.0f;
    d.sure = band_extent(g.a, g.b, g.ia, g.ic, S.det, S.sa, S.sc, g.cutoff - slack, Y.lo, Y.hi, xa, xb)
                 ? columns_meeting(xa, xb, g.cx, T, W, tx0, c0, c1)
                 : 0u;
    const uint32_t maybe = band_extent(g.a, g.b, g.ia, g.ic, S.det, S.sa, S.sc, g.cutoff + slack, Y.lo, Y.hi, xa, xb)
                               ? columns_meeting(xa, xb, g.cx, T, W, tx0, c0, c1)
                               : 0u;
    d.test = maybe & ~d.sure;
    return d;
}

// One strip entry for every (row, strip) a drawn footprint's rectangle
// covers, rank by rank in depth order, a footprint's entries row by row and
// left to right: key = its global strip, value = the rank. The rank's
// footprint is written once beside them, 32 aligned bytes at the rank
// (fp[2k]: centre x, y, conic a, b; fp[2k + 1]: conic c, cut power, first |
// last tile column << 16, slot), so the binning reads one sector an entry
// instead of the slot's record and rectangle. A wavefront's 64
// ranks own one contiguous run of entries (offsets holds every 64th rank's
// first entry); the wavefront walks it 64 entries at a time, lane L taking
// entry g0 + L and finding its rank by a binary search over the ranks' ends
// in LDS, so the stores of a step are contiguous.
constexpr uint32_t kEmitThreads = 256;

__global__ void __launch_bounds__(kEmitThreads) strip_emit_kernel(const DView* __restrict__ views, uint32_t n,
                                                                  uint32_t T,
                                                                  const uint32_t* __restrict__ sorted_slots,
                                                                  const uint32_t* __restrict__ offsets,
                                                                  const uint64_t* __restrict__ visible_dev,
                                                                  uint32_t max_visible, uint32_t cap,
                                                                  const float* __restrict__ records,
                                                                  uint4* __restrict__ fp_out,
                                                                  uint32_t* __restrict__ keys_out,
                                                                  uint32_t* __restrict__ ranks_out) {
    __shared__ uint32_t s_end[kEmitThreads];
    __shared__ uint4 s_rank[kEmitThreads];  // slot, first strip, strips a row, strips between rows
    const uint32_t lane = threadIdx.x & 63u, wbase = threadIdx.x & ~63u;
    const uint32_t visible = *visible_dev < max_visible ? (uint32_t)*visible_dev : max_visible;
    const uint32_t k = blockIdx.x * kEmitThreads + threadIdx.x;
    const uint32_t k0 = k - lane;
    uint32_t count = 0;
    if (k < visible) {
        const uint32_t slot = sorted_slots[k];
        const DView& V = views[slot / n];
        const float4* rec = reinterpret_cast<const float4*>(records + (size_t)slot * kRecordFloats);
        const float4 screen = rec[0], conic = rec[1];
        uint32_t x0, y0, x1, y1;
        tile_rect(V.width, V.height, T, screen.x, screen.y, conic.x, conic.y, conic.z, screen.w, x0, y0, x1, y1);
        const uint32_t w = x1 - x0 + 1u, h = y1 - y0 + 1u;
        const uint32_t sx0 = x0 / kStripTiles;
        const uint32_t cols = x1 / kStripTiles - sx0 + 1u;
        count = h * cols;
        s_rank[threadIdx.x] = make_uint4(k, V.strip_base + y0 * V.strips_x + sx0, cols, V.strips_x);
        fp_out[2u * k] = make_uint4(__builtin_bit_cast(uint32_t, screen.x), __builtin_bit_cast(uint32_t, screen.y),
                                    __builtin_bit_cast(uint32_t, conic.x), __builtin_bit_cast(uint32_t, conic.y));
        fp_out[2u * k + 1u] = make_uint4(__builtin_bit_cast(uint32_t, conic.z), __builtin_bit_cast(uint32_t, screen.w),
                                         x0 | ((x0 + w - 1u) << 16), slot);
    }
    const uint32_t end = wave_inclusive_sum(count, lane);
    s_end[threadIdx.x] = end;
    const uint32_t total = lane_u(end, 63);
    const uint32_t base = k0 < visible ? offsets[k0 / 64u] : 0u;
    __syncthreads();
    const uint32_t* ends = s_end + wbase;
    for (uint32_t g0 = 0; g0 < total; g0 += 64) {
        const uint32_t g = g0 + lane;
        if (g < total) {
            const uint32_t j = run_of(ends, g);
            const uint32_t q = g - (j != 0 ? ends[j - 1] : 0u);
            const uint4 rk = s_rank[wbase + j];
            const uint32_t row = q / rk.z, col = q - row * rk.z;
            const uint64_t at = (uint64_t)base + g;
            if (at < cap) {
                keys_out[at] = rk.y + row * rk.w + col;
                ranks_out[at] = rk.x;
            }
        }
    }
}

// The strip entries with their tiles decided (frames of at most 65 536
// strips): the emission above, but each entry's 16-bit mask of touched strip
// tiles is computed here, where each rank's footprint is read once into LDS
// (decide_columns, then the float test for the few columns it leaves), and
// rides in the key's upper half: key = strip | mask << 16, value = the slot.
// The strip sort orders the strip bits only, and the binning then reads the
// keys and values in order and nothing else.
__global__ void __launch_bounds__(kEmitThreads) strip_mask_emit_kernel(const DView* __restrict__ views, uint32_t n,
                                                                       uint32_t T,
                                                                       const uint32_t* __restrict__ sorted_slots,
                                                                       const uint32_t* __restrict__ offsets,
                                                                       const uint64_t* __restrict__ visible_dev,
                                                                       uint32_t max_visible, uint32_t cap,
                                                                       const float* __restrict__ records,
                                                                       uint32_t* __restrict__ keys_out,
                                                                       uint32_t* __restrict__ slots_out,
                                                                       uint32_t* __restrict__ rank_tiles,
                                                                       uint32_t* __restrict__ view_entry) {
    __shared__ uint32_t s_end[kEmitThreads];
    __shared__ uint32_t s_tiles[kEmitThreads]; // rank: tiles its entries' masks hold (counting frames)
    __shared__ float4 s_geo[kEmitThreads];   // rank: centre x, y, conic a, b
    __shared__ float4 s_geo2[kEmitThreads];  // rank: conic c, cut power, 1 / a, 1 / c
    // (Each rank's strips, first strip column and the view's strips a row
    // are formed again from these where an entry needs them: the smaller a
    // rank's LDS, the more workgroups a CU holds to cover the gathers' wait.)
    __shared__ uint4 s_int[kEmitThreads];    // rank: slot, first | last tile column << 16,
                                             // first row | strips a row << 16, first strip
    __shared__ float4 s_slice[kEmitThreads]; // rank: slice terms det, sa, sc; view width | height << 16
    __shared__ uint32_t s_own[kEmitThreads]; // step position: step << 8 | (rank + 1) of the run starting there
    const uint32_t lane = threadIdx.x & 63u, wbase = threadIdx.x & ~63u;
    const uint32_t visible = *visible_dev < max_visible ? (uint32_t)*visible_dev : max_visible;
    const uint32_t k = blockIdx.x * kEmitThreads + threadIdx.x;
    const uint32_t k0 = k - lane;
    uint32_t count = 0, view = 0;
    bool opens = false;  // the first rank of its view
    if (k < visible) {
        const uint32_t slot = sorted_slots[k];
        // With view_entry given the ranks are grouped by view, so a rank
        // opens its view where the rank before it is another view's: read
        // beside the rank's own slot, not after it.
        const uint32_t before = view_entry && k != 0 ? sorted_slots[k - 1u] : 0u;
        view = slot / n;
        opens = k == 0 || before / n != view;
        const DView& V = views[view];
        const float4* rec = reinterpret_cast<const float4*>(records + (size_t)slot * kRecordFloats);
        const float4 screen = rec[0], conic = rec[1];
        uint32_t x0, y0, x1, y1;
        tile_rect(V.width, V.height, T, screen.x, screen.y, conic.x, conic.y, conic.z, screen.w, x0, y0, x1, y1);
        const uint32_t w = x1 - x0 + 1u, h = y1 - y0 + 1u;
        const uint32_t sx0 = x0 / kStripTiles;
        const uint32_t ns = x1 / kStripTiles - sx0 + 1u;
        count = h * ns;
        s_geo[threadIdx.x] = make_float4(screen.x, screen.y, conic.x, conic.y);
        s_geo2[threadIdx.x] = make_float4(conic.z, screen.w, recip_of(conic.x), recip_of(conic.z));
        s_int[threadIdx.x] = make_uint4(slot, x0 | ((x0 + w - 1u) << 16), y0 | (V.strips_x << 16),
                                        V.strip_base + y0 * V.strips_x + sx0);
        const SliceTerms S = slice_terms(conic.x, conic.y, conic.z);
        s_slice[threadIdx.x] = make_float4(S.det, S.sa, S.sc, __builtin_bit_cast(float, V.width | (V.height << 16)));
    }
    const uint32_t end = wave_inclusive_sum_dpp(count);
    s_end[threadIdx.x] = end;
    s_own[threadIdx.x] = ~0u;
    s_tiles[threadIdx.x] = 0;
    const uint32_t total = lane_u(end, 63);
    const uint32_t base = k0 < visible ? offsets[k0 / 64u] : 0u;
    const uint32_t start = end - count;
    if (view_entry && k < visible) {
        // The first rank of its view records the view's first entry (the
        // segmented strip sort's run of that view).
        if (opens) {
            view_entry[view] = base + start;
        }
    }
    __syncthreads();
    const uint32_t* ends = s_end + wbase;
    // Each step's entries find their ranks without a search: the ranks whose
    // runs start in the step mark their first position (tagged with the
    // step), and a running max over the positions carries each mark on to
    // the run's end; a step that starts inside a run continues the last
    // step's final rank.
    uint32_t carried = 0;
    for (uint32_t g0 = 0, step = 0; g0 < total; g0 += 64, ++step) {
        if (count != 0 && start >= g0 && start < g0 + 64u) {
            s_own[wbase + (start - g0)] = (step << 8) | (lane + 1u);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        const uint32_t mark = s_own[threadIdx.x];
        const uint32_t run = wave_inclusive_max(mark >> 8 == step ? mark & 0xFFu : 0u);
        const uint32_t j = (run > carried + 1u ? run : carried + 1u) - 1u;
        carried = lane_u(j, 63);
        const uint32_t gi = g0 + lane;
        if (gi >= total) {
            continue;
        }
        const uint32_t q = gi - (j != 0 ? ends[j - 1] : 0u);
        const uint4 I = s_int[wbase + j];
        const float4 geo = s_geo[wbase + j], geo2 = s_geo2[wbase + j], sl = s_slice[wbase + j];
        const uint32_t x0 = I.y & 0xFFFFu, x1 = I.y >> 16;
        const uint32_t sx0 = x0 / kStripTiles, ns = x1 / kStripTiles - sx0 + 1u;
        const uint32_t y0 = I.z & 0xFFFFu, strips_x = I.z >> 16;
        // q / ns by a float estimate made exact (q < 2^24, ns <= 4096).
        uint32_t row = (uint32_t)((float)q * __builtin_amdgcn_rcpf((float)ns));
        row = __umul24(row, ns) > q ? row - 1u : row;
        row = __umul24(row + 1u, ns) <= q ? row + 1u : row;
        const uint32_t sxi = q - __umul24(row, ns);
        const uint32_t tx0 = (sx0 + sxi) * kStripTiles;
        const uint32_t c0 = x0 > tx0 ? x0 : tx0;
        const uint32_t c1 = x1 < tx0 + kStripTiles - 1u ? x1 : tx0 + kStripTiles - 1u;
        BigFootprint g{};
        g.cx = geo.x;
        g.cy = geo.y;
        g.a = geo.z;
        g.b = geo.w;
        g.c = geo2.x;
        g.cutoff = geo2.y;
        g.ia = geo2.z;
        g.ic = geo2.w;
        const uint32_t wh = __builtin_bit_cast(uint32_t, sl.w), width = wh & 0xFFFFu, height = wh >> 16;
        const float b2 = 2.0f * g.b;
        const BoxAxis Y = box_axis(y0 + row, T, height, g.cy, g.a, g.b, g.c, b2);
        const Decided d = decide_columns(g, SliceTerms{sl.x, sl.y, sl.z}, Y, T, width, tx0, c0, c1);
        uint32_t mask = d.sure;
        for (uint32_t t = d.test; t != 0; t &= t - 1u) {
            const uint32_t bit = (uint32_t)__builtin_ctz(t);
            if (box_hit(g, b2, box_axis(tx0 + bit, T, width, g.cx, g.c, g.b, g.a, b2), Y)) {
                mask |= 1u << bit;
            }
        }
        const uint64_t at = (uint64_t)base + gi;
        if (at < cap) {
            keys_out[at] = (I.w + __umul24(row, strips_x) + sxi) | (mask << 16);
            slots_out[at] = I.x;
        }
        if (rank_tiles) {
            atomicAdd(&s_tiles[wbase + j], (uint32_t)__builtin_popcount(mask));
        }
    }
    if (rank_tiles && k < visible) {
        // The wavefront's own ranks: its LDS adds are done when it gets here.
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        rank_tiles[k] = s_tiles[threadIdx.x];
    }
}

// A segment's descriptor, two words of uint4: (first entry, one past its
// last, its first table cell = 16 * the strip's first segment + its index
// in the strip, the strip's segments) and (view, tile row, first tile
// column, strip). An empty one (first == last) is a segment past the total.
__global__ void __launch_bounds__(256) seg_desc_kernel(StripBins b, uint4* __restrict__ desc, uint32_t strip_blocks) {
    const uint64_t total = *b.seg_total;
    const uint32_t segs = total < b.seg_cap ? (uint32_t)total : b.seg_cap;
    if (blockIdx.x < strip_blocks) {
        const uint32_t strip = blockIdx.x * 4u + (threadIdx.x >> 6), lane = threadIdx.x & 63u;
        if (strip >= b.strips) {
            return;
        }
        const uint32_t start = b.strip_ranges[2u * strip], end = b.strip_ranges[2u * strip + 1u];
        const uint32_t nseg = div_up(end - start, kStripSeg);
        if (nseg == 0) {
            return;
        }
        const uint32_t base = b.segbase[strip];
        const bool fits = (uint64_t)base + nseg <= segs;
        uint32_t lo = 0, hi = b.nviews;
        while (hi - lo > 1u) {
            const uint32_t mid = (lo + hi) >> 1;
            if (b.strip_bases[mid] <= strip) {
                lo = mid;
            } else {
                hi = mid;
            }
        }
        const DView& V = b.views[lo];
        const uint32_t local = strip - V.strip_base;
        const uint32_t ty = local / V.strips_x;
        const uint32_t tx0 = (local - ty * V.strips_x) * kStripTiles;
        for (uint32_t k = lane; k < nseg; k += 64u) {
            if ((uint64_t)base + k < b.seg_cap) {
                const uint32_t first = start + k * kStripSeg;
                const uint32_t last = end - first < kStripSeg ? end : first + kStripSeg;
                desc[2u * (base + k)] = fits ? make_uint4(first, last, kStripTiles * base + k, nseg) : make_uint4(0, 0, 0, 0);
                desc[2u * (base + k) + 1u] = make_uint4(lo, ty, tx0, strip);
            }
        }
    } else {
        const uint32_t stride = (gridDim.x - strip_blocks) * blockDim.x;
        for (uint32_t i = segs + (blockIdx.x - strip_blocks) * blockDim.x + threadIdx.x; i < b.seg_cap; i += stride) {
            desc[2u * i] = make_uint4(0, 0, 0, 0);
        }
    }
}

// An entry's footprint as the strip test reads it.
struct StripFootprint {
    float4 screen;  // centre x, y, view z, cut power
    float4 conic;   // a, b, c, alpha
    uint32_t x0, x1;
};

__device__ inline StripFootprint strip_footprint(const StripBins& b, uint32_t rank, uint32_t& slot) {
    StripFootprint f;
    const uint4 p = b.fp[2u * rank], q = b.fp[2u * rank + 1u];
    f.screen = make_float4(__builtin_bit_cast(float, p.x), __builtin_bit_cast(float, p.y), 0.0f,
                           __builtin_bit_cast(float, q.y));
    f.conic = make_float4(__builtin_bit_cast(float, p.z), __builtin_bit_cast(float, p.w), __builtin_bit_cast(float, q.x),
                          0.0f);
    f.x0 = q.z & 0xFFFFu;
    f.x1 = q.z >> 16;
    slot = q.w;
    return f;
}

// Which of the strip's tiles the footprint touches (bit j: tile tx0 + j):
// frame.slang:374-379's test of each tile its rectangle holds in the strip,
// the row's half of it taken once.
__device__ inline uint32_t strip_hits(const StripFootprint& sf, const DView& V, uint32_t T, uint32_t ty,
                                      uint32_t tx0) {
    BigFootprint f{};
    f.cx = sf.screen.x;
    f.cy = sf.screen.y;
    f.cutoff = sf.screen.w;
    f.a = sf.conic.x;
    f.b = sf.conic.y;
    f.c = sf.conic.z;
    f.ia = recip_of(f.a);
    f.ic = recip_of(f.c);
    const uint32_t c0 = sf.x0 > tx0 ? sf.x0 : tx0;
    const uint32_t c1 = sf.x1 < tx0 + kStripTiles - 1u ? sf.x1 : tx0 + kStripTiles - 1u;
    const float b2 = 2.0f * f.b;
    const BoxAxis Y = row_of(f, V, T, b2, ty);
    uint32_t mask = 0;
    for (uint32_t tx = c0; tx <= c1; ++tx) {
        if (box_hit(f, b2, column_of(f, V, T, b2, tx), Y)) {
            mask |= 1u << (tx - tx0);
        }
    }
    return mask;
}

// A segment's 1024 entries go 4 to a lane: entry j * 256 + thread of the
// segment, so each step j of the workgroup is 256 consecutive entries.
constexpr uint32_t kSegThreads = 256;
constexpr uint32_t kSegPer = kStripSeg / kSegThreads;

// What a tile test of a strip entry reads: its footprint's column terms
// and its row's terms.
struct StripEntryLds {
    float4 geo;   // centre x, conic a, conic b, conic c
    float4 row;   // the row's box terms: lo, hi, q_lo, q_hi
    float4 more;  // s_lo, s_hi, 1 / c, cut power
};

// Tests the segment's entries and counts each strip tile's hits:
// table[16 base + j nseg + k] for tile j of segment k of a strip whose
// first segment is base, so the table read in order is strip by strip,
// tile by tile, segment by segment -- the order of the tile-sorted pairs.
//
// An entry's footprint covers 1 to 16 of the strip's tiles. A lane testing
// its own entry's tiles would hold its wavefront to the widest of 64, so the
// tests are spread over the lanes: each lane writes its entry's (entry,
// column) items to an LDS list (lane L's items after the ones of lanes
// before it), and the wavefront takes the list 64 items a step, every lane
// one test, folding a hit into its entry's mask with an LDS OR. Each entry's
// footprint and row terms are in LDS beside the list. Every entry's slot,
// then every footprint, is loaded before any is tested.

// With the masks in the keys (strip_mask_emit_kernel), a segment only
// reads them: each entry's mask is kept for the pairs, each tile's hits
// counted into the table.
template <bool kSorted>
__global__ void __launch_bounds__(kSegThreads) strip_mask_count_kernel(StripBins b) {
    __shared__ uint32_t s_hits[kSegThreads / 64][kStripTiles];
    const uint64_t total = *b.seg_total;
    const uint32_t segs = total < b.seg_cap ? (uint32_t)total : b.seg_cap;
    if (blockIdx.x >= segs) {
        return;
    }
    const uint4 d0 = b.seg_desc[2u * blockIdx.x];
    if (d0.x >= d0.y) {
        return;
    }
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    uint32_t count[kStripTiles];
#pragma unroll
    for (uint32_t t = 0; t < kStripTiles; ++t) {
        count[t] = 0;
    }
    // The keys read together (a lane past the segment reads its last key and
    // counts nothing), then counted.
    // (kSorted: the strip sort wrote the masks themselves.)
    uint32_t key[kSegPer];
#pragma unroll
    for (uint32_t j = 0; j < kSegPer; ++j) {
        const uint32_t e = d0.x + j * kSegThreads + threadIdx.x;
        const uint32_t ec = e < d0.y ? e : d0.y - 1u;
        key[j] = kSorted ? (uint32_t)b.masks[ec] << 16 : b.strip_keys[ec];
    }
#pragma unroll
    for (uint32_t j = 0; j < kSegPer; ++j) {
        const uint32_t e = d0.x + j * kSegThreads + threadIdx.x;
        const uint32_t mask = e < d0.y ? key[j] >> 16 : 0u;
        if (!kSorted && e < d0.y) {
            b.masks[e] = (uint16_t)mask;
        }
#pragma unroll
        for (uint32_t t = 0; t < kStripTiles; ++t) {
            count[t] += (uint32_t)__popcll(__ballot((mask >> t) & 1u));
        }
    }
    if (lane == 0) {
#pragma unroll
        for (uint32_t t = 0; t < kStripTiles; ++t) {
            s_hits[wave][t] = count[t];
        }
    }
    __syncthreads();
    if (threadIdx.x < kStripTiles) {
        uint32_t sum = 0;
#pragma unroll
        for (uint32_t w = 0; w < kSegThreads / 64; ++w) {
            sum += s_hits[w][threadIdx.x];
        }
        b.table[d0.z + threadIdx.x * d0.w] = sum;
    }
}

__global__ void __launch_bounds__(kSegThreads) strip_count_kernel(StripBins b) {
    __shared__ StripEntryLds s_entry[kSegThreads];
    __shared__ uint16_t s_items[kSegThreads / 64][64 * kStripTiles];
    __shared__ uint32_t s_mask[kSegThreads];
    __shared__ uint32_t s_hits[kSegThreads / 64][kStripTiles];
    // The live segments, banded over the XCDs (the grid is the capacity's).
    const uint64_t total = *b.seg_total;
    const uint32_t segs = total < b.seg_cap ? (uint32_t)total : b.seg_cap;
    if (blockIdx.x >= segs) {
        return;
    }
    const uint32_t seg = xcd_band(blockIdx.x, segs);
    const uint4 d0 = b.seg_desc[2u * seg];
    if (d0.x >= d0.y) {
        return;  // past the segments: the same for the whole workgroup
    }
    const uint4 d1 = b.seg_desc[2u * seg + 1u];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6, wbase = threadIdx.x & ~63u;
    const DView& V = b.views[d1.x];
    const uint32_t T = b.tile, ty = d1.y, tx0 = d1.z;
    // The entries' ranks, then their footprints; each entry's rank is
    // replaced by its slot, which the pairs carry.
    uint32_t slot[kSegPer];
#pragma unroll
    for (uint32_t j = 0; j < kSegPer; ++j) {
        const uint32_t e = d0.x + j * kSegThreads + threadIdx.x;
        slot[j] = e < d0.y ? b.strip_slots[e] : ~0u;
    }
    StripFootprint f[kSegPer];
#pragma unroll
    for (uint32_t j = 0; j < kSegPer; ++j) {
        uint32_t s = 0;
        f[j] = strip_footprint(b, slot[j] != ~0u ? slot[j] : 0u, s);
        if (slot[j] != ~0u) {
            slot[j] = s;
            b.strip_slots[d0.x + j * kSegThreads + threadIdx.x] = s;
        }
    }
    uint32_t count[kStripTiles];
#pragma unroll
    for (uint32_t t = 0; t < kStripTiles; ++t) {
        count[t] = 0;
    }
    uint16_t* items = s_items[wave];
    for (uint32_t j = 0; j < kSegPer; ++j) {
        // This lane's entry: its columns in the strip, those decided and
        // those to test, and its terms in LDS.
        uint32_t sure = 0, test = 0;
        if (slot[j] != ~0u) {
            BigFootprint g{};
            g.cx = f[j].screen.x;
            g.cy = f[j].screen.y;
            g.cutoff = f[j].screen.w;
            g.a = f[j].conic.x;
            g.b = f[j].conic.y;
            g.c = f[j].conic.z;
            g.ia = recip_of(g.a);
            g.ic = recip_of(g.c);
            const uint32_t c0 = f[j].x0 > tx0 ? f[j].x0 : tx0;
            const uint32_t c1 = f[j].x1 < tx0 + kStripTiles - 1u ? f[j].x1 : tx0 + kStripTiles - 1u;
            const BoxAxis Y = row_of(g, V, T, 2.0f * g.b, ty);
            s_entry[threadIdx.x].geo = make_float4(g.cx, g.a, g.b, g.c);
            s_entry[threadIdx.x].row = make_float4(Y.lo, Y.hi, Y.q_lo, Y.q_hi);
            s_entry[threadIdx.x].more = make_float4(Y.s_lo, Y.s_hi, g.ic, g.cutoff);
            // Sure hits and sure misses from the slice's extent; the rest tested.
            const Decided d = decide_columns(g, slice_terms(g.a, g.b, g.c), Y, T, V.width, tx0, c0, c1);
            sure = d.sure;
            test = d.test;
        }
        s_mask[threadIdx.x] = sure;
        const uint32_t ntest = (uint32_t)__popc(test);
        const uint32_t end = wave_inclusive_sum(ntest, lane);
        const uint32_t total = lane_u(end, 63);
        for (uint32_t k = 0, t = test; t != 0; ++k, t &= t - 1u) {
            items[end - ntest + k] = (uint16_t)(lane | ((uint32_t)__builtin_ctz(t) << 6));
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        for (uint32_t g0 = 0; g0 < total; g0 += 64u) {
            const uint32_t gi = g0 + lane;
            if (gi < total) {
                const uint32_t item = items[gi];
                const uint32_t owner = item & 63u, col = item >> 6;
                const StripEntryLds en = s_entry[wbase + owner];
                BigFootprint g{};
                g.cx = en.geo.x;
                g.a = en.geo.y;
                g.b = en.geo.z;
                g.c = en.geo.w;
                g.ic = en.more.z;
                g.cutoff = en.more.w;
                BoxAxis Y;
                Y.lo = en.row.x;
                Y.hi = en.row.y;
                Y.q_lo = en.row.z;
                Y.q_hi = en.row.w;
                Y.s_lo = en.more.x;
                Y.s_hi = en.more.y;
                const float b2 = 2.0f * g.b;
                if (box_hit(g, b2, column_of(g, V, T, b2, tx0 + col), Y)) {
                    atomicOr(&s_mask[wbase + owner], 1u << col);
                }
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        const uint32_t mask = s_mask[threadIdx.x];
        if (slot[j] != ~0u) {
            b.masks[d0.x + j * kSegThreads + threadIdx.x] = (uint16_t)mask;
        }
#pragma unroll
        for (uint32_t t = 0; t < kStripTiles; ++t) {
            count[t] += (uint32_t)__popcll(__ballot((mask >> t) & 1u));
        }
        __builtin_amdgcn_wave_barrier();
    }
    if (lane == 0) {
#pragma unroll
        for (uint32_t t = 0; t < kStripTiles; ++t) {
            s_hits[wave][t] = count[t];
        }
    }
    __syncthreads();
    if (threadIdx.x < kStripTiles) {
        uint32_t sum = 0;
#pragma unroll
        for (uint32_t w = 0; w < kSegThreads / 64; ++w) {
            sum += s_hits[w][threadIdx.x];
        }
        b.table[d0.z + threadIdx.x * d0.w] = sum;
    }
}

// Writes each hit's slot at its place: the scanned table's start of (tile
// t, this segment), plus the hits of tile t at earlier entries of the
// segment (earlier entries are nearer: the strip's entries are in depth
// order). A tile's hits of one step of a wavefront land side by side.
__global__ void __launch_bounds__(kSegThreads) strip_pairs_kernel(StripBins b) {
    __shared__ uint32_t s_hits[kSegPer][kSegThreads / 64][kStripTiles];
    __shared__ uint32_t s_base[kStripTiles];
    const uint4 d0 = b.seg_desc[2u * blockIdx.x];
    if (d0.x >= d0.y) {
        return;
    }
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint64_t lt_mask = (1ull << lane) - 1ull;
    uint32_t mask[kSegPer], slot[kSegPer];
#pragma unroll
    for (uint32_t j = 0; j < kSegPer; ++j) {
        const uint32_t e = d0.x + j * kSegThreads + threadIdx.x;
        mask[j] = e < d0.y ? b.masks[e] : 0u;
        slot[j] = e < d0.y ? b.strip_slots[e] : 0u;
    }
    if (threadIdx.x < kStripTiles) {
        s_base[threadIdx.x] = b.table[d0.z + threadIdx.x * d0.w];
    }
#pragma unroll
    for (uint32_t j = 0; j < kSegPer; ++j) {
#pragma unroll
        for (uint32_t t = 0; t < kStripTiles; ++t) {
            const uint32_t hits = (uint32_t)__popcll(__ballot((mask[j] >> t) & 1u));
            if (lane == 0) {
                s_hits[j][wave][t] = hits;
            }
        }
    }
    __syncthreads();
    if (threadIdx.x < kStripTiles) {
        // Exclusive prefix over (step, wavefront), the entries' order.
        uint32_t run = s_base[threadIdx.x];
#pragma unroll
        for (uint32_t j = 0; j < kSegPer; ++j) {
#pragma unroll
            for (uint32_t w = 0; w < kSegThreads / 64; ++w) {
                const uint32_t h = s_hits[j][w][threadIdx.x];
                s_hits[j][w][threadIdx.x] = run;
                run += h;
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (uint32_t j = 0; j < kSegPer; ++j) {
#pragma unroll
        for (uint32_t t = 0; t < kStripTiles; ++t) {
            const bool mine = (mask[j] >> t) & 1u;
            const uint64_t bal = __ballot(mine);
            if (mine) {
                const uint32_t at = s_hits[j][wave][t] + (uint32_t)__popcll(bal & lt_mask);
                if (at < b.pair_cap) {
                    b.pair_slots[at] = slot[j];
                }
            }
        }
    }
}

// Each tile's pairs from the scanned table: [table(j, 0), table(j + 1, 0))
// of its strip, the strip's next tile's start being the one past its last
// segment; (0, 0) for a tile with none, as splat_tiles_clear.slang:8-17 and
// splat_ranges.slang:11-25 leave it. Clamped to the pair capacity.
__global__ void __launch_bounds__(256) tile_ranges_kernel(StripBins b, const uint32_t* __restrict__ tile_bases,
                                                          uint32_t tiles, uint32_t* __restrict__ ranges) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= tiles) {
        return;
    }
    uint32_t vlo = 0, vhi = b.nviews;
    while (vhi - vlo > 1u) {
        const uint32_t mid = (vlo + vhi) >> 1;
        if (tile_bases[mid] <= t) {
            vlo = mid;
        } else {
            vhi = mid;
        }
    }
    const DView& V = b.views[vlo];
    const uint32_t local = t - V.tile_base;
    const uint32_t ty = local / V.tiles_x, tx = local - ty * V.tiles_x;
    const uint32_t strip = V.strip_base + ty * V.strips_x + tx / kStripTiles, j = tx % kStripTiles;
    const uint64_t total = *b.seg_total;
    const uint32_t segs = total < b.seg_cap ? (uint32_t)total : b.seg_cap;
    const uint32_t base = b.segbase[strip];
    const uint32_t next = strip + 1u < b.strips ? b.segbase[strip + 1u] : (uint32_t)total;
    const uint32_t nseg = next - base;
    uint32_t start = 0, end = 0;
    if (nseg != 0 && (uint64_t)base + nseg <= segs) {
        const uint64_t pairs = *b.pair_total;
        const uint32_t i0 = kStripTiles * base + j * nseg, i1 = i0 + nseg;
        start = b.table[i0];
        end = i1 < kStripTiles * segs ? b.table[i1] : (uint32_t)pairs;
        start = start < b.pair_cap ? start : b.pair_cap;
        end = end < b.pair_cap ? end : b.pair_cap;
        if (start >= end) {
            start = 0;
            end = 0;
        }
    }
    ranges[2u * t] = start;
    ranges[2u * t + 1u] = end;
}

// The reference's 64-bit keys of the tile-sorted pairs (tile << 24 | depth
// key), a workgroup a tile over its range.
__global__ void __launch_bounds__(256) pair_keys_kernel(const uint32_t* __restrict__ ranges,
                                                        const uint32_t* __restrict__ values,
                                                        const uint32_t* __restrict__ depth_keys,
                                                        uint64_t* __restrict__ keys) {
    const uint32_t t = blockIdx.x;
    const uint32_t start = ranges[2u * t], end = ranges[2u * t + 1u];
    for (uint32_t i = start + threadIdx.x; i < end; i += blockDim.x) {
        keys[i] = (uint64_t)t << kDepthBits | (depth_keys[values[i]] & 0xFFFFFFu);
    }
}

// SplatCounters (splat_frame_counters.slang; words as kCounterWords says):
// each slot's cull reason from its depth key, each drawn footprint's tiles
// from what the frame kept of them, and per cloud its drawn footprints and
// pairs. A workgroup tallies in LDS and adds each nonzero word once.
__global__ void __launch_bounds__(256) frame_counters_kernel(CounterArgs a, uint32_t* __restrict__ words) {
    __shared__ uint32_t local[kCounterWords];
    for (uint32_t i = threadIdx.x; i < kCounterWords; i += 256) {
        local[i] = 0;
    }
    __syncthreads();
    const uint64_t vis_total = *a.visible_dev;
    const uint32_t visible = vis_total < a.slots ? (uint32_t)vis_total : a.slots;
    const uint32_t stride = gridDim.x * 256u;
    for (uint32_t s = blockIdx.x * 256u + threadIdx.x; s < a.slots; s += stride) {
        const uint32_t k = a.depth_keys[s];
        if ((k & 0xFF000000u) == kCulledKey) {
            const uint32_t reason = (k & 0xFFu) < ATHX_CULL_REASONS - 1 ? (k & 0xFFu) : ATHX_CULL_REASONS - 1;
            atomicAdd(&local[8 + reason], 1u);
        }
    }
    uint32_t vis = 0, pairs = 0, most = 0;
    for (uint32_t r = blockIdx.x * 256u + threadIdx.x; r < visible; r += stride) {
        const uint32_t slot = a.sorted_slots[r];
        const uint32_t tiles = a.rank_tiles ? a.rank_tiles[r] : a.slot_tiles[slot];
        vis += 1;
        pairs += tiles;
        most = tiles > most ? tiles : most;
        if (a.clouds != 0) {
            const uint32_t g = slot % a.count;
            if (g < a.base[a.clouds]) {
                uint32_t lo = 0, hi = a.clouds;  // base[lo] <= g < base[hi]
                while (hi - lo > 1u) {
                    const uint32_t mid = (lo + hi) >> 1;
                    if (a.base[mid] <= g) {
                        lo = mid;
                    } else {
                        hi = mid;
                    }
                }
                atomicAdd(&local[16 + 2 * lo], 1u);
                atomicAdd(&local[17 + 2 * lo], tiles);
            }
        }
    }
    // Pairs stay below 2^30 a frame, so one word holds a workgroup's sum.
    atomicAdd(&local[1], vis);
    atomicAdd(&local[2], pairs);
    atomicMax(&local[4], most);
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < kCounterWords; i += 256) {
        if (i == 0 || local[i] == 0) {
            continue;
        }
        if (i == 4) {
            atomicMax(&words[4], local[4]);
        } else {
            atomicAdd(&words[i], local[i]);
        }
    }
}

}  // namespace

hipError_t launch_project(const ProjectArgs& a, const DView* views_dev, uint32_t nviews, float* records,
                          uint32_t* touched, uint32_t* depth_keys, hipStream_t s) {
    if (a.n == 0 || nviews == 0) {
        return hipSuccess;
    }
    const dim3 grid(div_up(a.n, 256));
    // 16-byte loads of whole coefficient rows where the rows are 16 coefficients
    // a channel and the array is 16-byte aligned (torch's allocations are).
    const void* shp = a.sh_half ? static_cast<const void*>(a.sh_half) : static_cast<const void*>(a.sh);
    const bool vec = !a.packed && a.sh_coeffs == 16 && (reinterpret_cast<uintptr_t>(shp) & 15u) == 0;
    using Kernel = void (*)(ProjectArgs, const DView*, uint32_t, float*, uint32_t*, uint32_t*);
    const Kernel kernels[5][2] = {{project_kernel<kInSoa32, false>, project_kernel<kInSoa32, true>},
                                  {project_kernel<kInSoa32Vec, false>, project_kernel<kInSoa32Vec, true>},
                                  {project_kernel<kInSoa16, false>, project_kernel<kInSoa16, true>},
                                  {project_kernel<kInSoa16Vec, false>, project_kernel<kInSoa16Vec, true>},
                                  {project_kernel<kInPacked, false>, project_kernel<kInPacked, true>}};
    const uint32_t in = a.packed ? kInPacked : (a.sh_half ? 2u : 0u) + (vec ? 1u : 0u);
    const Kernel kernel = kernels[in][a.relit ? 1 : 0];
    hipLaunchKernelGGL(kernel, grid, dim3(256), 0, s, a, views_dev, nviews, records, touched, depth_keys);
    return hipGetLastError();
}

hipError_t launch_settle_count(const DView* views_dev, uint32_t n, uint32_t tile, uint32_t slots, const float* records,
                               uint32_t* touched, uint32_t* depth_keys, uint64_t* sums, hipStream_t s) {
    if (slots == 0) {
        return hipSuccess;
    }
    if (slots >= kSettleWide) {
        hipLaunchKernelGGL(settle_count_kernel<256>, dim3(div_up(slots, kCompactTile)), dim3(256), 0, s, views_dev, n,
                           tile, slots, records, touched, depth_keys, sums);
    } else {
        hipLaunchKernelGGL(settle_count_kernel<1024>, dim3(div_up(slots, kCompactTile)), dim3(1024), 0, s, views_dev,
                           n, tile, slots, records, touched, depth_keys, sums);
    }
    return hipGetLastError();
}

hipError_t launch_exact_counts(const DView* views_dev, uint32_t nviews, uint32_t n, uint32_t tile,
                               const float* records, uint32_t* counts, uint32_t* depth_keys, const Footprints& fp,
                               uint32_t* status_words, hipStream_t s) {
    const uint32_t slots = n * nviews;
    if (slots == 0) {
        return launch_big_list(fp, 0, status_words, s);
    }
    if (hipError_t e = launch_fill_u32(counts, 0u, slots, s); e != hipSuccess) {
        return e;
    }
    hipLaunchKernelGGL(rect_kernel, dim3(div_up(slots, 256)), dim3(256), 0, s, views_dev, n, slots, tile, records,
                       depth_keys, fp.rects);
    hipLaunchKernelGGL(count_small_kernel, dim3(div_up(slots, kSmallThreads)), dim3(kSmallThreads), 0, s, views_dev,
                       n, slots, tile, records, fp, counts, depth_keys);
    if (hipError_t e = launch_big_list(fp, slots, status_words, s); e != hipSuccess) {
        return e;
    }
    hipLaunchKernelGGL(count_big_kernel, dim3(kBigBlocks), dim3(kBigThreads), 0, s, views_dev, n, tile, records, fp,
                       counts, depth_keys);
    return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t* p, uint32_t value, uint64_t n, hipStream_t s) {
    if (n == 0) {
        return hipSuccess;
    }
    uint64_t blocks = (n + 255) / 256;
    if (blocks > 4096) {
        blocks = 4096;
    }
    hipLaunchKernelGGL(fill_u32_kernel, dim3((uint32_t)blocks), dim3(256), 0, s, p, value, n);
    return hipGetLastError();
}

hipError_t launch_ranges(const uint32_t* sorted_keys, const uint64_t* n_dev, uint32_t cap, uint32_t tiles,
                         uint32_t* ranges, hipStream_t s, uint32_t key_mask) {
    if (tiles == 0) {
        return hipSuccess;
    }
    hipLaunchKernelGGL(ranges_kernel, dim3(div_up(tiles, 256)), dim3(256), 0, s, sorted_keys, n_dev, cap, tiles,
                       key_mask, ranges);
    return hipGetLastError();
}

hipError_t launch_frame_counters(const CounterArgs& a, uint32_t* words, hipStream_t s) {
    if (a.slots == 0) {
        return hipSuccess;
    }
    // Few workgroups, each over many slots: every one adds its words once.
    const uint32_t blocks = div_up(a.slots, 256u * 32u) < 512u ? div_up(a.slots, 256u * 32u) : 512u;
    hipLaunchKernelGGL(frame_counters_kernel, dim3(blocks), dim3(256), 0, s, a, words);
    return hipGetLastError();
}

hipError_t launch_strip_emit(const DView* views_dev, uint32_t n, uint32_t tile, const uint32_t* sorted_slots,
                             const uint32_t* offsets, const uint64_t* visible_dev, uint32_t max_visible, uint32_t cap,
                             const float* records, uint4* fp, uint32_t* strip_keys, uint32_t* strip_ranks,
                             hipStream_t s) {
    if (max_visible == 0 || cap == 0) {
        return hipSuccess;
    }
    hipLaunchKernelGGL(strip_emit_kernel, dim3(div_up(max_visible, kEmitThreads)), dim3(kEmitThreads), 0, s,
                       views_dev, n, tile, sorted_slots, offsets, visible_dev, max_visible, cap, records, fp,
                       strip_keys, strip_ranks);
    return hipGetLastError();
}

hipError_t launch_seg_desc(const StripBins& b, uint4* desc, hipStream_t s) {
    if (b.seg_cap == 0 || b.strips == 0) {
        return hipSuccess;
    }
    const uint32_t strip_blocks = div_up(b.strips, 4);
    const uint32_t tail_blocks = div_up(b.seg_cap, 256) < 1024 ? div_up(b.seg_cap, 256) : 1024;
    hipLaunchKernelGGL(seg_desc_kernel, dim3(strip_blocks + tail_blocks), dim3(256), 0, s, b, desc, strip_blocks);
    return hipGetLastError();
}

hipError_t launch_strip_count(const StripBins& b, hipStream_t s) {
    if (b.seg_cap == 0) {
        return hipSuccess;
    }
    if (b.masks_sorted) {
        hipLaunchKernelGGL(strip_mask_count_kernel<true>, dim3(b.seg_cap), dim3(kSegThreads), 0, s, b);
    } else if (b.strip_keys) {
        hipLaunchKernelGGL(strip_mask_count_kernel<false>, dim3(b.seg_cap), dim3(kSegThreads), 0, s, b);
    } else {
        hipLaunchKernelGGL(strip_count_kernel, dim3(b.seg_cap), dim3(kSegThreads), 0, s, b);
    }
    return hipGetLastError();
}

hipError_t launch_strip_mask_emit(const DView* views_dev, uint32_t n, uint32_t tile, const uint32_t* sorted_slots,
                                  const uint32_t* offsets, const uint64_t* visible_dev, uint32_t max_visible,
                                  uint32_t cap, const float* records, uint32_t* strip_keys, uint32_t* strip_slots,
                                  uint32_t* rank_tiles, hipStream_t s, uint32_t* view_entry) {
    if (max_visible == 0 || cap == 0) {
        return hipSuccess;
    }
    hipLaunchKernelGGL(strip_mask_emit_kernel, dim3(div_up(max_visible, kEmitThreads)), dim3(kEmitThreads), 0, s,
                       views_dev, n, tile, sorted_slots, offsets, visible_dev, max_visible, cap, records,
                       strip_keys, strip_slots, rank_tiles, view_entry);
    return hipGetLastError();
}

// The segments (one a view) of a multi-view frame's depth sort or strip
// sort: view v's run starts at its given start (the depth sort's: where the
// compaction placed the view's first visible slot) or at the entry its first
// rank recorded in the emission (the strip sort's; a view with no visible
// rank has an empty run where the next starts), cut where the capacity ends
// (the frame is then binned again at its exact size); its partitions follow
// the earlier views', a partial one closing each run. Every workgroup lays the runs out in LDS and maps its 256
// partitions to their views; the first also writes the runs and the totals.
__global__ void __launch_bounds__(256) strip_segments_kernel(SegmentBuild g) {
    __shared__ uint32_t s_start[kMaxSegmentViews + 1];
    __shared__ uint32_t s_part[kMaxSegmentViews + 1];
    __shared__ uint32_t s_first[kMaxSegmentViews], s_entry[kMaxSegmentViews];
    const uint32_t nv = g.nviews;
    // Every view's first rank and first entry read at once (a lane a view),
    // then laid out by one lane from LDS.
    if (threadIdx.x < nv) {
        s_first[threadIdx.x] = g.view_first ? g.view_first[(size_t)threadIdx.x * g.view_stride] : 0u;
        s_entry[threadIdx.x] = g.view_entry[threadIdx.x];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t visible = *g.visible_dev < g.max_visible ? (uint32_t)*g.visible_dev : g.max_visible;
        const uint64_t etotal = *g.entry_total_dev;
        uint32_t next = etotal < g.cap ? (uint32_t)etotal : g.cap;
        s_start[nv] = next;
        uint32_t next_rank = visible;
        for (uint32_t v = nv; v-- > 0;) {
            // A view with ranks starts where its first rank's entries do;
            // one without, where the next view does (view_first null: every
            // view's start is given).
            const uint32_t R = s_first[v];
            if (!g.view_first || R < next_rank) {
                const uint32_t e = s_entry[v];
                next = e < next ? e : next;  // within the capacity, never past the next run
            }
            next_rank = R < next_rank ? R : next_rank;
            s_start[v] = next;
        }
        uint32_t parts = 0;
        for (uint32_t v = 0; v < nv; ++v) {
            s_part[v] = parts;
            parts += div_up(s_start[v + 1] - s_start[v], kSegmentPart);
        }
        s_part[nv] = parts;
        if (blockIdx.x == 0) {
            for (uint32_t v = 0; v < nv; ++v) {
                g.seg_start[v] = s_start[v];
                g.seg_len[v] = s_start[v + 1] - s_start[v];
                g.seg_part[v] = s_part[v];
            }
            *g.parts = parts < g.parts_cap ? parts : g.parts_cap;
            *g.binned_total = s_start[nv] - s_start[0];
        }
    }
    __syncthreads();
    const uint32_t p = blockIdx.x * 256u + threadIdx.x;
    if (p < s_part[nv] && p < g.parts_cap) {
        // The last view whose partitions start at or before p (views with no
        // entries start where the next does).
        uint32_t lo = 0, hi = nv;
        while (hi - lo > 1u) {
            const uint32_t mid = (lo + hi) >> 1;
            if (s_part[mid] <= p) {
                lo = mid;
            } else {
                hi = mid;
            }
        }
        g.part_seg[p] = lo;
    }
}

hipError_t launch_strip_segments(const SegmentBuild& g, hipStream_t s) {
    if (g.nviews == 0 || g.nviews > kMaxSegmentViews || g.parts_cap == 0) {
        return hipErrorInvalidValue;
    }
    hipLaunchKernelGGL(strip_segments_kernel, dim3(div_up(g.parts_cap, 256)), dim3(256), 0, s, g);
    return hipGetLastError();
}


hipError_t launch_strip_pairs(const StripBins& b, hipStream_t s) {
    if (b.seg_cap == 0) {
        return hipSuccess;
    }
    hipLaunchKernelGGL(strip_pairs_kernel, dim3(b.seg_cap), dim3(kSegThreads), 0, s, b);
    return hipGetLastError();
}

hipError_t launch_tile_ranges(const StripBins& b, const uint32_t* tile_bases, uint32_t tiles, uint32_t* ranges,
                              hipStream_t s) {
    if (tiles == 0) {
        return hipSuccess;
    }
    hipLaunchKernelGGL(tile_ranges_kernel, dim3(div_up(tiles, 256)), dim3(256), 0, s, b, tile_bases, tiles, ranges);
    return hipGetLastError();
}

hipError_t launch_pair_keys(const uint32_t* ranges, uint32_t tiles, const uint32_t* values,
                            const uint32_t* depth_keys, uint64_t* keys, hipStream_t s) {
    if (tiles == 0) {
        return hipSuccess;
    }
    hipLaunchKernelGGL(pair_keys_kernel, dim3(tiles), dim3(256), 0, s, ranges, values, depth_keys, keys);
    return hipGetLastError();
}

}  // namespace athx
