// relight.h -- the relit colour of a splat for gfx950: relitSplat's dome
// term (shaders/athenea/splat/splat_relight.slang:1775-1824) for one dome
// whose sky was prepared (relitByDome, :1379-1385 and :1640-1692), a surface
// with the plain lobes (packing.slang:249-261: no coat, sheen, transmission
// or emission; specular weight, colour and index 1, 1, 1.5), and the
// prepared-sky reader of shaders/athenea/light/environment.slang:50-225.
//
// With those lobes the splat's colour is
//   albedo (1 - metallic) (E(n) / pi + sun)  +  polish(n.v) L(mirror, roughness)
// where E is the irradiance of the sky's nine radiance coefficients, L the
// GGX-prefiltered octahedral chain read at sqrt(roughness) of its 8 levels,
// and polish the environment BRDF of splat_relight.slang:455-509 (the
// MaterialX rational fit of GGX's directional albedo with Turquin's energy
// compensation, material/microfacet.slang:66-143), blended between the
// dielectric and the conductor by metallic. Everything is FP32 in the order
// written, so the colour is the same bits a host restatement computes.
#pragma once

#include "athx_internal.h"

namespace athx {

constexpr uint32_t kEnvLevels = 8;
constexpr uint32_t kEnvCoarse = 256;
constexpr uint32_t kEnvFloor = 16;

__device__ inline float rl_min(float a, float b) { return a < b ? a : b; }
__device__ inline float rl_max(float a, float b) { return a > b ? a : b; }
__device__ inline float rl_clamp(float x, float lo, float hi) { return rl_min(rl_max(x, lo), hi); }
__device__ inline float rl_sat(float x) { return rl_clamp(x, 0.0f, 1.0f); }
__device__ inline float rl_lerp(float a, float b, float t) { return a + (b - a) * t; }

// environment.slang:66-69
__device__ inline uint32_t env_side(uint32_t level, uint32_t base_side) {
    const uint32_t base = base_side > kEnvCoarse ? base_side : kEnvCoarse;
    const uint32_t side = base >> (level < kEnvLevels - 1 ? level : kEnvLevels - 1);
    return side > kEnvFloor ? side : kEnvFloor;
}

// environment.slang:77-92: where a level starts, in texels.
__device__ inline uint32_t env_level_offset(uint32_t level, uint32_t base_side) {
    uint32_t at = 0;
    for (uint32_t k = 0; k < level; ++k) {
        const uint32_t side = env_side(k, base_side);
        at += side * side;
    }
    return at;
}

__device__ inline float half_lo(uint32_t w) { return (float)__builtin_bit_cast(_Float16, (uint16_t)(w & 0xFFFFu)); }
__device__ inline float half_hi(uint32_t w) { return (float)__builtin_bit_cast(_Float16, (uint16_t)(w >> 16)); }

// environment.slang:133-147, octEncode from common/octahedral.slang:12-23:
// bilinear over one level, clamped at the seam.
__device__ inline float3 env_level_at(const uint32_t* __restrict__ texels, uint32_t base_side, uint32_t level,
                                      float dx, float dy, float dz) {
    const uint32_t side_u = env_side(level, base_side);
    const float side = (float)side_u;
    const float s = rl_max(fabsf(dx) + fabsf(dy) + fabsf(dz), 1.0e-20f);
    const float nx = dx / s, ny = dy / s, nz = dz / s;
    float px = nx, py = ny;
    if (!(nz >= 0.0f)) {
        px = (1.0f - fabsf(ny)) * (nx >= 0.0f ? 1.0f : -1.0f);
        py = (1.0f - fabsf(nx)) * (ny >= 0.0f ? 1.0f : -1.0f);
    }
    const float u = px * 0.5f + 0.5f, v = py * 0.5f + 0.5f;
    const float atx = rl_clamp(u * side - 0.5f, 0.0f, side - 1.0f);
    const float aty = rl_clamp(v * side - 0.5f, 0.0f, side - 1.0f);
    const uint32_t lox = (uint32_t)floorf(atx), loy = (uint32_t)floorf(aty);
    const uint32_t hix = lox + 1 < side_u - 1 ? lox + 1 : side_u - 1;
    const uint32_t hiy = loy + 1 < side_u - 1 ? loy + 1 : side_u - 1;
    const float fx = atx - (float)lox, fy = aty - (float)loy;
    const uint2* t = reinterpret_cast<const uint2*>(texels) + env_level_offset(level, base_side);
    const uint2 w00 = t[loy * side_u + lox], w10 = t[loy * side_u + hix];
    const uint2 w01 = t[hiy * side_u + lox], w11 = t[hiy * side_u + hix];
    float3 r;
    r.x = rl_lerp(rl_lerp(half_lo(w00.x), half_lo(w10.x), fx), rl_lerp(half_lo(w01.x), half_lo(w11.x), fx), fy);
    r.y = rl_lerp(rl_lerp(half_hi(w00.x), half_hi(w10.x), fx), rl_lerp(half_hi(w01.x), half_hi(w11.x), fx), fy);
    r.z = rl_lerp(rl_lerp(half_lo(w00.y), half_lo(w10.y), fx), rl_lerp(half_lo(w01.y), half_lo(w11.y), fx), fy);
    return r;
}

// environment.slang:162-170
__device__ inline float3 env_radiance(const uint32_t* __restrict__ texels, uint32_t base_side, float dx, float dy,
                                      float dz, float roughness) {
    const float level = sqrtf(rl_sat(roughness)) * (float)(kEnvLevels - 1);
    const uint32_t lo = (uint32_t)floorf(level);
    const uint32_t hi = lo + 1 < kEnvLevels - 1 ? lo + 1 : kEnvLevels - 1;
    const float f = level - (float)lo;
    const float3 a = env_level_at(texels, base_side, lo, dx, dy, dz);
    const float3 b = env_level_at(texels, base_side, hi, dx, dy, dz);
    return make_float3(rl_lerp(a.x, b.x, f), rl_lerp(a.y, b.y, f), rl_lerp(a.z, b.z, f));
}

// environment.slang:182-223: irradiance at n from the first nine radiance
// coefficients (sh: 16 float4, rgb used).
__device__ inline float3 env_irradiance(const float* __restrict__ sh, float x, float y, float z) {
    const float basis[9] = {0.28209479177387814f,
                            -0.4886025119029199f * y,
                            0.4886025119029199f * z,
                            -0.4886025119029199f * x,
                            1.0925484305920792f * x * y,
                            -1.0925484305920792f * y * z,
                            0.31539156525252005f * (2.0f * z * z - x * x - y * y),
                            -1.0925484305920792f * x * z,
                            0.5462742152960396f * (x * x - y * y)};
    float3 e = make_float3(0.0f, 0.0f, 0.0f);
#pragma unroll
    for (uint32_t k = 0; k < 9; ++k) {
        const float band = k == 0 ? 3.14159265358979f : (k < 4 ? 2.0943951023932f : 0.78539816339745f);
        const float w = band * basis[k];
        e.x = e.x + sh[k * 4 + 0] * w;
        e.y = e.y + sh[k * 4 + 1] * w;
        e.z = e.z + sh[k * 4 + 2] * w;
    }
    return e;
}

// material/microfacet.slang:66-82: the fit of GGX's directional albedo,
// returned as its (scale, bias) pair.
__device__ inline float2 ggx_albedo_ab(float x, float y) {
    const float x2 = x * x, y2 = y * y;
    const float r0 = 0.1003f + -0.6303f * x + 9.748f * y + -2.038f * x * y + 29.34f * x2 + -8.245f * y2 +
                     -26.44f * x2 * y + 19.99f * x * y2 + -5.448f * x2 * y2;
    const float r1 = 0.9345f + -2.323f * x + 2.229f * y + -3.748f * x * y + 1.424f * x2 + -0.7684f * y2 +
                     1.436f * x2 * y + 0.2913f * x * y2 + 0.6286f * x2 * y2;
    const float r2 = 1.0f + -1.765f * x + 8.263f * y + 11.53f * x * y + 28.96f * x2 + -7.507f * y2 +
                     -36.11f * x2 * y + 15.86f * x * y2 + 33.37f * x2 * y2;
    const float r3 = 1.0f + 0.2281f * x + 15.94f * y + -55.83f * x * y + 13.08f * x2 + 41.26f * y2 +
                     54.9f * x2 * y + 300.2f * x * y2 + -285.1f * x2 * y2;
    return make_float2(rl_sat(r0 / r2), rl_sat(r1 / r3));
}

// microfacet.slang:85-88 (Turquin 2019)
__device__ inline float ggx_compensation(float nv, float alpha, float fss) {
    const float2 ab = ggx_albedo_ab(nv, alpha);
    const float ess = ab.x * 1.0f + 1.0f * ab.y;
    return 1.0f + fss * (1.0f - ess) / ess;
}

// microfacet.slang:99-107
__device__ inline float fresnel_dielectric(float c, float ior) {
    const float g2 = ior * ior + c * c - 1.0f;
    if (g2 < 0.0f) {
        return 1.0f;
    }
    const float g = sqrtf(g2);
    const float q = (g - c) / (g + c);
    const float t = ((g + c) * c - 1.0f) / ((g - c) * c + 1.0f);
    return 0.5f * (q * q) * (1.0f + t * t);
}

// splat_relight.slang:455-460 at the plain lobes' index 1.5: F0 0.04, tint 1.
__device__ inline float ggx_env_dielectric(float roughness, float nv) {
    const float alpha = rl_max(roughness * roughness, 1.0e-4f);
    const float F = fresnel_dielectric(nv, 1.5f);
    const float2 ab = ggx_albedo_ab(nv, alpha);
    return (0.04f * ab.x + 1.0f * ab.y) * ggx_compensation(nv, alpha, F);
}

// splat_relight.slang:472-488 with a white edge: mxArtisticIor
// (microfacet.slang:117-126) and fresnelConductor (:128-143), one channel.
__device__ inline float ggx_env_conductor(float albedo, float roughness, float nv) {
    const float alpha = rl_max(roughness * roughness, 1.0e-4f);
    const float r = rl_clamp(rl_sat(albedo), 0.0f, 0.99f);
    const float root = sqrtf(r);
    const float low = (1.0f - r) / (1.0f + r);
    const float high = (1.0f + root) / (1.0f - root);
    const float n = rl_lerp(high, low, 1.0f);
    const float up = n + 1.0f, down = n - 1.0f;
    const float k = sqrtf(rl_max((up * up * r - down * down) / rl_max(1.0f - r, 1.0e-6f), 0.0f));
    const float ct = rl_sat(nv);
    const float ct2 = ct * ct, st2 = 1.0f - ct2;
    const float n2 = n * n, k2 = k * k;
    const float t0 = n2 - k2 - st2;
    const float a2b2 = sqrtf(t0 * t0 + 4.0f * n2 * k2);
    const float t1 = a2b2 + ct2;
    const float aa = sqrtf(rl_max(0.5f * (a2b2 + t0), 0.0f));
    const float t2 = 2.0f * aa * nv;
    const float rs = (t1 - t2) / (t1 + t2);
    const float t3 = ct2 * a2b2 + st2 * st2;
    const float t4 = t2 * st2;
    const float rp = rs * (t3 - t4) / (t3 + t4);
    const float F = 0.5f * (rp + rs);
    const float2 ab = ggx_albedo_ab(nv, alpha);
    return (F * ab.x + F * ab.y) * ggx_compensation(nv, alpha, F);
}

// packing.slang:129-135 (unpackNormal): a unit vector from its octahedral
// word, x in the low 16 bits and y in the high, each a unorm.
__device__ inline float3 unpack_normal(uint32_t word) {
    const float fx = (float)(word & 0xFFFFu) / 65535.0f * 2.0f - 1.0f;
    const float fy = (float)(word >> 16) / 65535.0f * 2.0f - 1.0f;
    float nx = fx, ny = fy;
    const float nz = 1.0f - fabsf(fx) - fabsf(fy);
    const float t = rl_sat(-nz);
    nx = nx + (nx >= 0.0f ? -t : t);
    ny = ny + (ny >= 0.0f ? -t : t);
    const float inv = 1.0f / sqrtf(nx * nx + ny * ny + nz * nz);
    return make_float3(nx * inv, ny * inv, nz * inv);
}

// The relit colour of Gaussian i seen from eye (cloud space): the shading
// normal (splatShadingNormal, splat_relight.slang:1843-1880) and the point
// taken to the world (relightToWorld, relightNormalToWorld :1882-1915),
// then the dome term.
__device__ inline float3 relit_colour(const RelightArgs& rl, uint32_t i, const float R[3][3], float sx, float sy,
                                      float sz, float px, float py, float pz, const float* eye, float3 albedo_sh) {
    float3 albedo = albedo_sh;
    if (rl.albedo) {
        albedo = make_float3(rl.albedo[i * 3 + 0], rl.albedo[i * 3 + 1], rl.albedo[i * 3 + 2]);
    }
    // splat_encoding.slang:212-213 (pbrMetallic, pbrRoughness) for a packed word.
    const uint32_t pbr = rl.pbr ? rl.pbr[i] : 0u;
    const float metallic = rl.pbr ? (float)(pbr & 0xFFu) / 255.0f : rl.metallic[i];
    const float roughness = rl.pbr ? (float)((pbr >> 8) & 0xFFu) / 255.0f : rl.roughness[i];
    // The disc's axis: the shortest scale's column, towards the eye.
    const uint32_t s = sx <= sy ? (sx <= sz ? 0u : 2u) : (sy <= sz ? 1u : 2u);
    float d0 = R[0][s], d1 = R[1][s], d2 = R[2][s];
    const float dl = 1.0f / sqrtf(d0 * d0 + d1 * d1 + d2 * d2);
    d0 = d0 * dl; d1 = d1 * dl; d2 = d2 * dl;
    const float e0 = eye[0] - px, e1 = eye[1] - py, e2 = eye[2] - pz;
    const float el = 1.0f / sqrtf(e0 * e0 + e1 * e1 + e2 * e2);
    if (d0 * (e0 * el) + d1 * (e1 * el) + d2 * (e2 * el) < 0.0f) {
        d0 = -d0; d1 = -d1; d2 = -d2;
    }
    float c0 = d0, c1 = d1, c2 = d2;
    if (rl.normals || rl.normals_packed) {
        float s0, s1, s2;
        if (rl.normals_packed) {
            const float3 u = unpack_normal(rl.normals_packed[i]);
            s0 = u.x; s1 = u.y; s2 = u.z;
        } else {
            s0 = rl.normals[i * 3 + 0]; s1 = rl.normals[i * 3 + 1]; s2 = rl.normals[i * 3 + 2];
        }
        const float along = s0 * d0 + s1 * d1 + s2 * d2;
        float flip;
        if (fabsf(along) >= 0.5f) {
            flip = along < 0.0f ? -1.0f : 1.0f;
        } else {
            flip = s0 * e0 + s1 * e1 + s2 * e2 < 0.0f ? -1.0f : 1.0f;
        }
        c0 = flip * s0; c1 = flip * s1; c2 = flip * s2;
    }
    const float* w = rl.o2w;
    const float pw0 = w[0] * px + w[1] * py + w[2] * pz + w[3] * 1.0f;
    const float pw1 = w[4] * px + w[5] * py + w[6] * pz + w[7] * 1.0f;
    const float pw2 = w[8] * px + w[9] * py + w[10] * pz + w[11] * 1.0f;
    // The normal by the cofactor of the linear part, the determinant's sign
    // put back.
    const float a0 = w[0], a1 = w[4], a2 = w[8];
    const float b0 = w[1], b1 = w[5], b2 = w[9];
    const float q0 = w[2], q1 = w[6], q2 = w[10];
    const float bc0 = b1 * q2 - b2 * q1, bc1 = b2 * q0 - b0 * q2, bc2 = b0 * q1 - b1 * q0;
    const float ca0 = q1 * a2 - q2 * a1, ca1 = q2 * a0 - q0 * a2, ca2 = q0 * a1 - q1 * a0;
    const float ab0 = a1 * b2 - a2 * b1, ab1 = a2 * b0 - a0 * b2, ab2 = a0 * b1 - a1 * b0;
    const float side = a0 * bc0 + a1 * bc1 + a2 * bc2 < 0.0f ? -1.0f : 1.0f;
    float n0 = side * (c0 * bc0 + c1 * ca0 + c2 * ab0);
    float n1 = side * (c0 * bc1 + c1 * ca1 + c2 * ab1);
    float n2 = side * (c0 * bc2 + c1 * ca2 + c2 * ab2);
    const float nn = n0 * n0 + n1 * n1 + n2 * n2;
    if (nn > 1.0e-20f) {
        const float inv = 1.0f / sqrtf(nn);
        n0 = n0 * inv; n1 = n1 * inv; n2 = n2 * inv;
    } else {
        n0 = 0.0f; n1 = 0.0f; n2 = 1.0f;
    }
    // The eye in the world: the cloud-space eye through the same rows.
    const float ew0 = w[0] * eye[0] + w[1] * eye[1] + w[2] * eye[2] + w[3] * 1.0f;
    const float ew1 = w[4] * eye[0] + w[5] * eye[1] + w[6] * eye[2] + w[7] * 1.0f;
    const float ew2 = w[8] * eye[0] + w[9] * eye[1] + w[10] * eye[2] + w[11] * 1.0f;
    float wo0 = ew0 - pw0, wo1 = ew1 - pw1, wo2 = ew2 - pw2;
    const float wl = 1.0f / sqrtf(wo0 * wo0 + wo1 * wo1 + wo2 * wo2);
    wo0 = wo0 * wl; wo1 = wo1 * wl; wo2 = wo2 * wl;
    const float ndv = n0 * wo0 + n1 * wo1 + n2 * wo2;
    const float nv = rl_max(ndv, 1.0e-4f);
    const float idot = n0 * -wo0 + n1 * -wo1 + n2 * -wo2;
    const float m0 = -wo0 - 2.0f * idot * n0, m1 = -wo1 - 2.0f * idot * n1, m2 = -wo2 - 2.0f * idot * n2;
    const float over_pi = 1.0f / 3.14159265358979f;
    float3 above = env_irradiance(rl.sky_sh, n0, n1, n2);
    above.x = above.x * over_pi;
    above.y = above.y * over_pi;
    above.z = above.z * over_pi;
    if (rl.sky_sun[3] > 0.0f) {
        const float cosine = n0 * rl.sky_sun[0] + n1 * rl.sky_sun[1] + n2 * rl.sky_sun[2];
        const float g = rl_max(cosine, 0.0f) * over_pi;
        above.x = above.x + rl.sky_sun[4] * g;
        above.y = above.y + rl.sky_sun[5] * g;
        above.z = above.z + rl.sky_sun[6] * g;
    }
    const float3 reflected = env_radiance(rl.texels, rl.base_side, m0, m1, m2, roughness);
    const float diel = ggx_env_dielectric(roughness, nv);
    const float ms = rl_sat(metallic);
    float3 out;
    out.x = albedo.x * (1.0f - metallic) * above.x +
            rl_lerp(diel, ggx_env_conductor(albedo.x, roughness, nv), ms) * reflected.x;
    out.y = albedo.y * (1.0f - metallic) * above.y +
            rl_lerp(diel, ggx_env_conductor(albedo.y, roughness, nv), ms) * reflected.y;
    out.z = albedo.z * (1.0f - metallic) * above.z +
            rl_lerp(diel, ggx_env_conductor(albedo.z, roughness, nv), ms) * reflected.z;
    return out;
}

}  // namespace athx
