// A sky prepared for relighting on the device (technique::Environment,
// modules/technique/include/athenea/technique/Environment.h:1-155, over
// env_project / env_prefilter .slang), into the layout athx_sky reads
// (shaders/athenea/light/environment.slang:50-110): eight octahedral levels
// -- level 0 the sky itself, level k the sky convolved with a GGX lobe of
// roughness (k/7)^2, alpha max(roughness^2, 1e-3), taking n = v = r -- and
// the sky's 16 radiance
// coefficients in real spherical harmonics.
//
// The source is an equirect of linear radiance (texel (i, j) looks along
// theta = pi (j + 0.5) / H from +y, phi = 2 pi (i + 0.5) / W from +x towards
// +z, with solid angle sin(theta) (pi / H) (2 pi / W)); every sum over it runs
// in FP64 and each output texel is rounded once to float, then to f16.
// sky_prefilter_kernel is a texel a lane with the source staged through LDS;
// sky_sh_kernel is one workgroup summing in a fixed order.
#include "athx_internal.h"

namespace athx {
namespace {

constexpr uint32_t kSkyThreads = 256;
constexpr uint32_t kSkyLevels = 8;
constexpr double kPi = 3.14159265358979323846;

__device__ inline void texel_dir(uint32_t i, uint32_t j, uint32_t w, uint32_t h, double d[3], double& omega) {
    const double theta = kPi * ((double)j + 0.5) / (double)h;
    const double phi = 2.0 * kPi * ((double)i + 0.5) / (double)w;
    const double st = sin(theta);
    d[0] = st * cos(phi);
    d[1] = cos(theta);
    d[2] = st * sin(phi);
    omega = st * (kPi / (double)h) * (2.0 * kPi / (double)w);
}

// common/octahedral.slang:25-34: the square [0, 1]^2 back to a unit direction.
__device__ inline void oct_dir(double u, double v, double n[3]) {
    double fx = u * 2.0 - 1.0, fy = v * 2.0 - 1.0, fz = 1.0 - fabs(fx) - fabs(fy);
    const double t = fmin(fmax(-fz, 0.0), 1.0);
    fx += fx >= 0.0 ? -t : t;
    fy += fy >= 0.0 ? -t : t;
    const double len = sqrt(fx * fx + fy * fy + fz * fz);
    n[0] = fx / len;
    n[1] = fy / len;
    n[2] = fz / len;
}

__device__ inline uint32_t half_word(double v) {
    return (uint32_t)__builtin_bit_cast(uint16_t, (_Float16)(float)v);
}

__device__ inline void store_texel(uint32_t* texels, uint32_t at, const double rgb[3]) {
    texels[2u * at] = half_word(rgb[0]) | (half_word(rgb[1]) << 16);
    texels[2u * at + 1u] = half_word(rgb[2]);
}

struct SkyLevels {
    uint32_t side[kSkyLevels];
    uint32_t offset[kSkyLevels];
    uint32_t total;  // texels of the dome
};

__device__ inline uint32_t level_of(const SkyLevels& L, uint32_t t) {
    uint32_t level = 0;
    while (level + 1 < kSkyLevels && t >= L.offset[level + 1]) {
        ++level;
    }
    return level;
}

// Level 0: the sky itself, a bilinear lookup (azimuth wraps, the poles clamp).
__device__ inline void sample_equirect(const float* img, uint32_t w, uint32_t h, const double d[3], double rgb[3]) {
    const double theta = acos(fmin(fmax(d[1], -1.0), 1.0));
    double phi = fmod(atan2(d[2], d[0]), 2.0 * kPi);
    if (phi < 0.0) {
        phi += 2.0 * kPi;
    }
    const double fy = fmin(fmax(theta / kPi * h - 0.5, 0.0), (double)h - 1.0);
    const double fx = phi / (2.0 * kPi) * w - 0.5;
    const int64_t y0 = (int64_t)floor(fy);
    const int64_t y1 = y0 + 1 < (int64_t)h - 1 ? y0 + 1 : (int64_t)h - 1;
    const double ty = fy - (double)y0;
    const double x0f = floor(fx);
    const double tx = fx - x0f;
    const int64_t x0 = (((int64_t)x0f % (int64_t)w) + (int64_t)w) % (int64_t)w;
    const int64_t x1 = (x0 + 1) % (int64_t)w;
    for (int c = 0; c < 3; ++c) {
        const double a = img[((size_t)y0 * w + x0) * 3 + c], b = img[((size_t)y0 * w + x1) * 3 + c];
        const double e = img[((size_t)y1 * w + x0) * 3 + c], f = img[((size_t)y1 * w + x1) * 3 + c];
        const double top = a * (1.0 - tx) + b * tx, bottom = e * (1.0 - tx) + f * tx;
        rgb[c] = top * (1.0 - ty) + bottom * ty;
    }
}

__global__ void __launch_bounds__(kSkyThreads) sky_texels_kernel(const float* __restrict__ img, uint32_t w,
                                                                 uint32_t h, SkyLevels L,
                                                                 uint32_t* __restrict__ texels) {
    __shared__ double s_d[kSkyThreads][3];
    __shared__ double s_l[kSkyThreads][3];
    __shared__ double s_w[kSkyThreads];
    const uint32_t t = blockIdx.x * kSkyThreads + threadIdx.x;
    const bool mine = t < L.total;
    const uint32_t level = mine ? level_of(L, t) : 0u;
    double n[3] = {0.0, 0.0, 1.0};
    if (mine) {
        const uint32_t side = L.side[level], k = t - L.offset[level];
        const uint32_t row = k / side, col = k - row * side;
        oct_dir(((double)col + 0.5) / side, ((double)row + 0.5) / side, n);
    }
    const double rough = (double)level / (double)(kSkyLevels - 1);
    // roughness (k/7)^2 (Environment.cpp:286-287), alpha = max(roughness^2,
    // 1e-3) (env_prefilter.slang:139).
    const double alpha = fmax(rough * rough * rough * rough, 1e-3), a2 = alpha * alpha;
    double acc[3] = {0.0, 0.0, 0.0}, wsum = 0.0;
    // Levels past 0 integrate the whole source; every lane of the workgroup
    // stages it, so the loop is the same for all.
    const bool prefilter = __syncthreads_or(mine && level > 0);
    if (prefilter) {
        const uint32_t texels_src = w * h;
        for (uint32_t s0 = 0; s0 < texels_src; s0 += kSkyThreads) {
            const uint32_t s = s0 + threadIdx.x;
            if (s < texels_src) {
                double d[3], om;
                texel_dir(s % w, s / w, w, h, d, om);
                s_d[threadIdx.x][0] = d[0];
                s_d[threadIdx.x][1] = d[1];
                s_d[threadIdx.x][2] = d[2];
                s_w[threadIdx.x] = om;
                for (int c = 0; c < 3; ++c) {
                    s_l[threadIdx.x][c] = img[(size_t)s * 3 + c];
                }
            }
            __syncthreads();
            const uint32_t count = texels_src - s0 < kSkyThreads ? texels_src - s0 : kSkyThreads;
            if (mine && level > 0) {
                for (uint32_t j = 0; j < count; ++j) {
                    const double* l = s_d[j];
                    const double nl = n[0] * l[0] + n[1] * l[1] + n[2] * l[2];
                    if (!(nl > 0.0)) {
                        continue;
                    }
                    double hv[3] = {n[0] + l[0], n[1] + l[1], n[2] + l[2]};
                    const double hl = fmax(sqrt(hv[0] * hv[0] + hv[1] * hv[1] + hv[2] * hv[2]), 1e-12);
                    const double nh = (n[0] * hv[0] + n[1] * hv[1] + n[2] * hv[2]) / hl;
                    const double q = nh * nh * (a2 - 1.0) + 1.0;
                    const double wt = a2 / (kPi * q * q) * nl * s_w[j];
                    wsum += wt;
                    acc[0] += wt * s_l[j][0];
                    acc[1] += wt * s_l[j][1];
                    acc[2] += wt * s_l[j][2];
                }
            }
            __syncthreads();
        }
    }
    if (!mine) {
        return;
    }
    double rgb[3];
    if (level == 0) {
        sample_equirect(img, w, h, n, rgb);
    } else {
        const double norm = fmax(wsum, 1e-30);
        rgb[0] = acc[0] / norm;
        rgb[1] = acc[1] / norm;
        rgb[2] = acc[2] / norm;
    }
    store_texel(texels, t, rgb);
}

// The 16 radiance coefficients (environment.slang:202-224's basis): one
// workgroup, a lane summing every 256th texel, then the lanes in a fixed
// tree; rgb of coefficient k at sh[4k .. 4k + 2].
__global__ void __launch_bounds__(kSkyThreads) sky_sh_kernel(const float* __restrict__ img, uint32_t w, uint32_t h,
                                                             float* __restrict__ sh) {
    __shared__ double s_sum[kSkyThreads][3];
    for (uint32_t k = 0; k < 16; ++k) {
        double acc[3] = {0.0, 0.0, 0.0};
        for (uint32_t s = threadIdx.x; s < w * h; s += kSkyThreads) {
            double d[3], om;
            texel_dir(s % w, s / w, w, h, d, om);
            const double x = d[0], y = d[1], z = d[2];
            double Y;
            switch (k) {
            case 0: Y = 0.28209479177387814; break;
            case 1: Y = -0.4886025119029199 * y; break;
            case 2: Y = 0.4886025119029199 * z; break;
            case 3: Y = -0.4886025119029199 * x; break;
            case 4: Y = 1.0925484305920792 * x * y; break;
            case 5: Y = -1.0925484305920792 * y * z; break;
            case 6: Y = 0.31539156525252005 * (2 * z * z - x * x - y * y); break;
            case 7: Y = -1.0925484305920792 * x * z; break;
            case 8: Y = 0.5462742152960396 * (x * x - y * y); break;
            case 9: Y = -0.5900435899266435 * y * (3 * x * x - y * y); break;
            case 10: Y = 2.890611442640554 * x * y * z; break;
            case 11: Y = -0.4570457994644658 * y * (4 * z * z - x * x - y * y); break;
            case 12: Y = 0.3731763325901154 * z * (2 * z * z - 3 * x * x - 3 * y * y); break;
            case 13: Y = -0.4570457994644658 * x * (4 * z * z - x * x - y * y); break;
            case 14: Y = 1.445305721320277 * z * (x * x - y * y); break;
            default: Y = -0.5900435899266435 * x * (x * x - 3 * y * y); break;
            }
            for (int c = 0; c < 3; ++c) {
                acc[c] += Y * om * img[(size_t)s * 3 + c];
            }
        }
        for (int c = 0; c < 3; ++c) {
            s_sum[threadIdx.x][c] = acc[c];
        }
        __syncthreads();
        for (uint32_t half = kSkyThreads / 2; half > 0; half >>= 1) {
            if (threadIdx.x < half) {
                for (int c = 0; c < 3; ++c) {
                    s_sum[threadIdx.x][c] += s_sum[threadIdx.x + half][c];
                }
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            sh[4 * k + 0] = (float)s_sum[0][0];
            sh[4 * k + 1] = (float)s_sum[0][1];
            sh[4 * k + 2] = (float)s_sum[0][2];
            sh[4 * k + 3] = 0.0f;
        }
        __syncthreads();
    }
}

// THE SUN, TAKEN OUT OF THE SKY (env_sun.slang:154-262, env_residual.slang:
// 55-78): the brightest compact source of an equirect, found on a box-filtered
// level at most 256 texels wide, and the equirect with its texels set to
// nothing. The texel directions and the rows' solid-angle weights come in as
// tables the host computes once in double (sin and cos nowhere on the device),
// so the host's restatement decides every texel the same way.

// One level down: each texel the mean of its 2 x 2 block, read clamped at an
// odd edge.
__global__ void __launch_bounds__(kSkyThreads) sky_half_kernel(const float* __restrict__ src, uint32_t w, uint32_t h,
                                                               float* __restrict__ dst, uint32_t dw, uint32_t dh) {
    const uint32_t t = blockIdx.x * kSkyThreads + threadIdx.x;
    if (t >= dw * dh) {
        return;
    }
    const uint32_t y = t / dw, x = t - y * dw;
    const uint32_t x0 = min(2u * x, w - 1u), x1 = min(2u * x + 1u, w - 1u);
    const uint32_t y0 = min(2u * y, h - 1u), y1 = min(2u * y + 1u, h - 1u);
    for (int c = 0; c < 3; ++c) {
        const float a = src[((size_t)y0 * w + x0) * 3 + c], b = src[((size_t)y0 * w + x1) * 3 + c];
        const float e = src[((size_t)y1 * w + x0) * 3 + c], f = src[((size_t)y1 * w + x1) * 3 + c];
        dst[(size_t)t * 3 + c] = 0.25f * ((a + b) + (e + f));
    }
}

// env_sun.slang:154-156's grey, as luminance.
__device__ inline float sky_lum(const float* rgb) { return rgb[0] * 0.2126f + rgb[1] * 0.7152f + rgb[2] * 0.0722f; }

// The log-luminance histogram (80 half stops from 2^-20, each texel weighed by
// its row's fixed-point sin theta, so the sums are exact in any order) and the
// brightest texel: peak = max of lum bits << 32 | ~index (the first of equal
// ones, as a scan in raster order keeps).
__global__ void __launch_bounds__(kSkyThreads) sky_sun_scan_kernel(const float* __restrict__ img, uint32_t w,
                                                                   uint32_t h, const uint32_t* __restrict__ row_weight,
                                                                   unsigned long long* __restrict__ hist,
                                                                   unsigned long long* __restrict__ peak) {
    __shared__ unsigned long long s_hist[80];
    if (threadIdx.x < 80) {
        s_hist[threadIdx.x] = 0;
    }
    __syncthreads();
    const uint32_t t = blockIdx.x * kSkyThreads + threadIdx.x;
    unsigned long long best = 0;
    if (t < w * h) {
        const uint32_t y = t / w;
        const float lum = sky_lum(img + (size_t)t * 3);
        const float l2 = athx_logf(fmaxf(lum, 1.0e-9f)) * 1.44269504088896341f;
        const float fb = fminf(fmaxf((l2 + 20.0f) * 2.0f, 0.0f), 79.0f);
        if (row_weight[y] != 0u) {
            atomicAdd(&s_hist[(uint32_t)fb], (unsigned long long)row_weight[y]);
            best = ((unsigned long long)__float_as_uint(fmaxf(lum, 0.0f)) << 32) | (0xFFFFFFFFull - t);
        }
    }
    for (uint32_t d = 32; d >= 1; d >>= 1) {
        const unsigned long long o = __shfl_xor(best, d, 64);
        best = o > best ? o : best;
    }
    if ((threadIdx.x & 63u) == 0 && best != 0) {
        atomicMax(peak, best);
    }
    __syncthreads();
    if (threadIdx.x < 80 && s_hist[threadIdx.x] != 0) {
        atomicAdd(&hist[threadIdx.x], s_hist[threadIdx.x]);
    }
}

// The sun's texels summed (one workgroup, a lane taking every 256th texel,
// then a fixed tree; FP64): its irradiance, its light-weighted direction, the
// light weight and the solid angle. dir = (sin theta cos phi, cos theta,
// sin theta sin phi) from the float tables; omega = sin theta * area.
__global__ void __launch_bounds__(kSkyThreads) sky_sun_sum_kernel(const float* __restrict__ img, uint32_t w,
                                                                  uint32_t h, const float* __restrict__ rows,
                                                                  const float* __restrict__ cols, float area,
                                                                  float threshold, float ax, float ay, float az,
                                                                  float cos_widest, double* __restrict__ sums) {
    __shared__ double s_sum[kSkyThreads][8];
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (uint32_t t = threadIdx.x; t < w * h; t += kSkyThreads) {
        const uint32_t y = t / w, x = t - y * w;
        const float st = rows[2u * y], ct = rows[2u * y + 1u];
        if (!(st > 0.0f)) {
            continue;
        }
        const float dx = st * cols[2u * x], dy = ct, dz = st * cols[2u * x + 1u];
        const float* rgb = img + (size_t)t * 3;
        const float lum = sky_lum(rgb);
        if (!(lum >= threshold) || !(((dx * ax + dy * ay) + dz * az) >= cos_widest)) {
            continue;
        }
        const float omega = st * area;
        const float lw = lum * omega;
        acc[0] += (double)(rgb[0] * omega);
        acc[1] += (double)(rgb[1] * omega);
        acc[2] += (double)(rgb[2] * omega);
        acc[3] += (double)(dx * lw);
        acc[4] += (double)(dy * lw);
        acc[5] += (double)(dz * lw);
        acc[6] += (double)lw;
        acc[7] += (double)omega;
    }
    for (int q = 0; q < 8; ++q) {
        s_sum[threadIdx.x][q] = acc[q];
    }
    __syncthreads();
    for (uint32_t half = kSkyThreads / 2; half > 0; half >>= 1) {
        if (threadIdx.x < half) {
            for (int q = 0; q < 8; ++q) {
                s_sum[threadIdx.x][q] += s_sum[threadIdx.x + half][q];
            }
        }
        __syncthreads();
    }
    if (threadIdx.x < 8) {
        sums[threadIdx.x] = s_sum[0][threadIdx.x];
    }
}

// env_residual.slang:55-78: the equirect with the sun's texels (as bright as
// the threshold, within the widest cone of the sun's axis) set to nothing.
__global__ void __launch_bounds__(kSkyThreads) sky_residual_kernel(const float* __restrict__ img, uint32_t w,
                                                                   uint32_t h, const float* __restrict__ rows,
                                                                   const float* __restrict__ cols, float threshold,
                                                                   float ax, float ay, float az, float cos_widest,
                                                                   float* __restrict__ out) {
    const uint32_t t = blockIdx.x * kSkyThreads + threadIdx.x;
    if (t >= w * h) {
        return;
    }
    const uint32_t y = t / w, x = t - y * w;
    const float st = rows[2u * y], ct = rows[2u * y + 1u];
    const float dx = st * cols[2u * x], dy = ct, dz = st * cols[2u * x + 1u];
    const float* rgb = img + (size_t)t * 3;
    const bool sunlit = sky_lum(rgb) >= threshold && ((dx * ax + dy * ay) + dz * az) >= cos_widest;
    for (int c = 0; c < 3; ++c) {
        out[(size_t)t * 3 + c] = sunlit ? 0.0f : rgb[c];
    }
}

}  // namespace

// environment.slang:66-92: level k's side is max(max(base, 256) >> k, 16),
// the levels back to back.
uint32_t sky_texels(uint32_t base_side, uint32_t* sides, uint32_t* offsets) {
    const uint32_t base = base_side > 256u ? base_side : 256u;
    uint32_t total = 0;
    for (uint32_t k = 0; k < kSkyLevels; ++k) {
        const uint32_t side = (base >> k) > 16u ? (base >> k) : 16u;
        if (sides) {
            sides[k] = side;
            offsets[k] = total;
        }
        total += side * side;
    }
    return total;
}

hipError_t launch_prepare_sky(const float* equirect, uint32_t width, uint32_t height, uint32_t base_side,
                              uint32_t* texels, float* sh, hipStream_t s) {
    SkyLevels L{};
    L.total = sky_texels(base_side, L.side, L.offset);
    hipLaunchKernelGGL(sky_texels_kernel, dim3(div_up(L.total, kSkyThreads)), dim3(kSkyThreads), 0, s, equirect, width,
                       height, L, texels);
    hipLaunchKernelGGL(sky_sh_kernel, dim3(1), dim3(kSkyThreads), 0, s, equirect, width, height, sh);
    return hipGetLastError();
}

hipError_t launch_sky_half(const float* src, uint32_t w, uint32_t h, float* dst, uint32_t dw, uint32_t dh,
                           hipStream_t s) {
    hipLaunchKernelGGL(sky_half_kernel, dim3(div_up(dw * dh, kSkyThreads)), dim3(kSkyThreads), 0, s, src, w, h, dst,
                       dw, dh);
    return hipGetLastError();
}

hipError_t launch_sky_sun_scan(const float* img, uint32_t w, uint32_t h, const uint32_t* row_weight,
                               unsigned long long* hist, unsigned long long* peak, hipStream_t s) {
    hipLaunchKernelGGL(sky_sun_scan_kernel, dim3(div_up(w * h, kSkyThreads)), dim3(kSkyThreads), 0, s, img, w, h,
                       row_weight, hist, peak);
    return hipGetLastError();
}

hipError_t launch_sky_sun_sum(const float* img, uint32_t w, uint32_t h, const float* rows, const float* cols,
                              float area, float threshold, const float axis[3], float cos_widest, double* sums,
                              hipStream_t s) {
    hipLaunchKernelGGL(sky_sun_sum_kernel, dim3(1), dim3(kSkyThreads), 0, s, img, w, h, rows, cols, area, threshold,
                       axis[0], axis[1], axis[2], cos_widest, sums);
    return hipGetLastError();
}

hipError_t launch_sky_residual(const float* img, uint32_t w, uint32_t h, const float* rows, const float* cols,
                               float threshold, const float axis[3], float cos_widest, float* out, hipStream_t s) {
    hipLaunchKernelGGL(sky_residual_kernel, dim3(div_up(w * h, kSkyThreads)), dim3(kSkyThreads), 0, s, img, w, h,
                       rows, cols, threshold, axis[0], axis[1], axis[2], cos_widest, out);
    return hipGetLastError();
}

}  // namespace athx
