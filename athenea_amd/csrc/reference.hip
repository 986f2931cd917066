// The ground truth on the device, and the image comparisons the reference
// measures renders with (render::ReferenceRenderer, modules/render/include/
// athenea/render/ReferenceRenderer.h:20-107).
//
// ref_project_kernel is the reference's own projection
// (reference/reference_project.slang:1-107): written apart from
// project_kernel -- full 3x3 covariance products, no tile rectangle, no tile
// test, no depth key -- so that a mistake in how the rasterizer bins, counts
// or sorts is not also in what it is checked against; the Gaussian itself
// (EWA, dilation, the cut power, the paid opacity) is the same definition. Its
// record has the rasterizer's layout (centre xy, view z, cut power, conic a b
// c, alpha, rgb, 0).
//
// ref_blend_kernel is reference/reference_blend.slang:1-144 re-expressed for
// a GPU: every splat reaching a pixel, blended front to back in (depth,
// index) order. The reference collects each pixel's contributors in a
// per-pixel list and insertion-sorts it; here the drawn splats are sorted
// once by (depth, index) -- a stable sort of the depth's float bits -- and
// every pixel walks that one order, staged through LDS, which is the
// per-pixel order restricted to the pixel. No list, so no list limit: a
// pixel reached by more than kMaxHits splats is counted as the reference
// counts it (a comparison means something only at zero), and blended whole.
//
// compare_codes_kernel, compare_hdr_kernel and the reductions are
// reference/image_compare.slang, image_compare_hdr.slang and
// histogram_reduce.slang: a row of pixels a thread, as there, so the sums
// run in the same order.
#include "athx_internal.h"

namespace athx {
namespace {

constexpr uint32_t kRefThreads = 256;

__device__ inline float half_of(uint32_t h) { return (float)__builtin_bit_cast(_Float16, (uint16_t)h); }

// The real SH basis of sh.slang:8-23 at unit direction (x, y, z), terms
// 1 .. 15 (term 0 is the base colour's).
__device__ inline void sh_basis16(float x, float y, float z, float Y[16]) {
    const float c1 = 0.4886025119029199f;
    const float xx = x * x, yy = y * y, zz = z * z, xy = x * y, yz = y * z, xz = x * z;
    Y[0] = 0.0f;
    Y[1] = -c1 * y;
    Y[2] = c1 * z;
    Y[3] = -c1 * x;
    Y[4] = 1.0925484305920792f * xy;
    Y[5] = -1.0925484305920792f * yz;
    Y[6] = 0.31539156525252005f * (2.0f * zz - xx - yy);
    Y[7] = -1.0925484305920792f * xz;
    Y[8] = 0.5462742152960396f * (xx - yy);
    Y[9] = -0.5900435899266435f * y * (3.0f * xx - yy);
    Y[10] = 2.890611442640554f * xy * z;
    Y[11] = -0.4570457994644658f * y * (4.0f * zz - xx - yy);
    Y[12] = 0.3731763325901154f * z * (2.0f * zz - 3.0f * xx - 3.0f * yy);
    Y[13] = -0.4570457994644658f * x * (4.0f * zz - xx - yy);
    Y[14] = 1.445305721320277f * z * (xx - yy);
    Y[15] = -0.5900435899266435f * x * (xx - 3.0f * yy);
}

struct RefArgs {
    const float* positions;  // [n][4]
    const uint32_t* shape;   // [n][4]
    const uint32_t* sh;      // [n][words] or null
    uint32_t n, rest_per_colour, words, sh_limit, antialias, linear;
    uint32_t base;  // this cloud's first index among the frame's splats (instances back to back)
};

__global__ void __launch_bounds__(kRefThreads) ref_project_kernel(RefArgs a, const DView* __restrict__ view,
                                                                  float* __restrict__ records,
                                                                  uint32_t* __restrict__ keys,
                                                                  uint32_t* __restrict__ values) {
    const uint32_t i = blockIdx.x * kRefThreads + threadIdx.x;
    if (i >= a.n) {
        return;
    }
    const DView& V = *view;
    const uint32_t o = a.base + i;  // the splat's index in the frame
    float* rec = records + (size_t)o * kRecordFloats;
    keys[o] = 0xFFFFFFFFu;  // culled: after every drawn splat
    values[o] = o;
    float4* rec4 = reinterpret_cast<float4*>(rec);
    rec4[0] = make_float4(0.0f, 0.0f, -1.0f, 0.0f);  // a culled record: zeros, z = -1
    rec4[1] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    rec4[2] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const float4 p = reinterpret_cast<const float4*>(a.positions)[i];
    const uint4 sw = reinterpret_cast<const uint4*>(a.shape)[i];
    const float s[3] = {athx_expf(half_of(sw.y & 0xFFFFu)), athx_expf(half_of(sw.y >> 16)),
                        athx_expf(half_of(sw.z & 0xFFFFu))};
    const float* m = V.m;
    const float vx = ((m[0] * p.x + m[1] * p.y) + m[2] * p.z) + m[3];
    const float vy = ((m[4] * p.x + m[5] * p.y) + m[6] * p.z) + m[7];
    const float vz = ((m[8] * p.x + m[9] * p.y) + m[10] * p.z) + m[11];
    if (!(vz > V.near_z) || vz > V.far_z) {
        return;
    }
    // The packed quaternion (x, y, z, w) as the layout decodes it.
    const uint32_t largest = sw.x & 3u;
    const float small[3] = {((float)((sw.x >> 22) & 1023u) / 1023.0f * 2.0f - 1.0f) * 0.70710678f,
                            ((float)((sw.x >> 12) & 1023u) / 1023.0f * 2.0f - 1.0f) * 0.70710678f,
                            ((float)((sw.x >> 2) & 1023u) / 1023.0f * 2.0f - 1.0f) * 0.70710678f};
    const float rest = sqrtf(fmaxf(0.0f, 1.0f - ((small[0] * small[0] + small[1] * small[1]) + small[2] * small[2])));
    float q[4];
    for (uint32_t k = 0, next = 0; k < 4; ++k) {
        q[k] = k == largest ? rest : small[next < 3 ? next++ : 2];
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
    if (V.orthographic) {
        cx = V.focal_x * vx + V.centre_x;
        cy = V.focal_y * vy + V.centre_y;
        J[0][0] = V.focal_x;
        J[1][1] = V.focal_y;
    } else {
        cx = V.focal_x * vx / vz + V.centre_x;
        cy = V.focal_y * vy / vz + V.centre_y;
        const float lox = (-0.15f * (float)V.width - V.centre_x) / V.focal_x;
        const float hix = (1.15f * (float)V.width - V.centre_x) / V.focal_x;
        const float loy = (-0.15f * (float)V.height - V.centre_y) / V.focal_y;
        const float hiy = (1.15f * (float)V.height - V.centre_y) / V.focal_y;
        const float tx = fminf(fmaxf(vx / vz, lox), hix) * vz;
        const float ty = fminf(fmaxf(vy / vz, loy), hiy) * vz;
        J[0][0] = V.focal_x / vz;
        J[0][2] = -V.focal_x * tx / (vz * vz);
        J[1][1] = V.focal_y / vz;
        J[1][2] = -V.focal_y * ty / (vz * vz);
    }
    // cov2 = J (CV J^T): the inner product first, as mul(j, mul(covView, transpose(j))).
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
        return;
    }
    const float filtered = sqrtf(fmaxf(before, 0.0f) / det);
    const float paid = a.antialias ? filtered : 1.0f;
    const float spread = paid * p.w;
    const float alpha = spread >= 1.0f / 255.0f ? spread + (1.0f - paid) * (1.0f / 255.0f) : spread;
    if (!(alpha >= 1.0f / 255.0f)) {
        return;
    }
    const float cutoff = fminf(2.0f * athx_logf(fmaxf(255.0f * alpha, 1.0001f)), 9.0f);
    // The colour from this eye: the base and the rest, degree capped.
    const float dx = p.x - V.eye[0], dy = p.y - V.eye[1], dz = p.z - V.eye[2];
    const float len = sqrtf((dx * dx + dy * dy) + dz * dz);
    float Y[16];
    sh_basis16(dx / len, dy / len, dz / len, Y);
    const uint32_t rest_degree = a.rest_per_colour >= 15 ? 3u : a.rest_per_colour >= 8 ? 2u : a.rest_per_colour >= 3 ? 1u : 0u;
    const uint32_t degree = rest_degree < a.sh_limit ? rest_degree : a.sh_limit;
    const uint32_t terms = (degree + 1) * (degree + 1);
    const float base[3] = {half_of(sw.z >> 16), half_of(sw.w & 0xFFFFu), half_of(sw.w >> 16)};
    float col[3];
    for (uint32_t ch = 0; ch < 3; ++ch) {
        float v = base[ch];
        for (uint32_t k = 1; k < terms; ++k) {
            const uint32_t h = (k - 1) * 3 + ch;
            const uint32_t word = a.sh[(size_t)i * a.words + (h >> 1)];
            v = v + Y[k] * half_of((h & 1u) ? word >> 16 : word & 0xFFFFu);
        }
        col[ch] = a.linear ? fmaxf(v, 0.0f) : v;
    }
    if (!a.linear) {
        // sRGB-stored colours decoded (color.slang:39-41).
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
    rec[11] = 0.0f;
    keys[o] = __builtin_bit_cast(uint32_t, vz);  // a positive float: its bits order as it does
}

struct RefStage {
    float4 screen[kRefThreads];  // centre xy, z, cut power
    float4 conic[kRefThreads];   // a, b, c, alpha
    float4 colour[kRefThreads];
};

__global__ void __launch_bounds__(kRefThreads) ref_blend_kernel(const DView* __restrict__ view,
                                                                const float* __restrict__ records,
                                                                const uint32_t* __restrict__ sorted_keys,
                                                                const uint32_t* __restrict__ sorted,
                                                                uint32_t n, uint32_t max_hits, uint32_t depth_mode,
                                                                float depth_threshold,
                                                                uint32_t* __restrict__ overflow) {
    __shared__ RefStage st;
    const DView& V = *view;
    const uint32_t tiles_x = div_up(V.width, 16u);
    const uint32_t px = (blockIdx.x % tiles_x) * 16u + (threadIdx.x & 15u);
    const uint32_t py = (blockIdx.x / tiles_x) * 16u + (threadIdx.x >> 4);
    const bool inside = px < V.width && py < V.height;
    const float pxc = (float)px + 0.5f, pyc = (float)py + 0.5f;
    float r = 0.0f, g = 0.0f, b = 0.0f, T = 1.0f, dsum = 0.0f, dat = -1.0f;
    uint32_t hits = 0;
    bool done = !inside;
    // Every drawn splat, in (depth, index) order; a pixel keeps counting
    // its contributors past saturation until they exceed max_hits.
    for (uint32_t first = 0; first < n; first += kRefThreads) {
        const bool working = inside && (!done || hits <= max_hits);
        if (!__syncthreads_or(working)) {
            break;
        }
        const uint32_t k = first + threadIdx.x;
        const bool drawn = k < n && sorted_keys[k] != 0xFFFFFFFFu;
        if (drawn) {
            const float4* rec = reinterpret_cast<const float4*>(records + (size_t)sorted[k] * kRecordFloats);
            st.screen[threadIdx.x] = rec[0];
            st.conic[threadIdx.x] = rec[1];
            st.colour[threadIdx.x] = rec[2];
        }
        const uint32_t count = (uint32_t)__popcll(__ballot(drawn));  // drawn splats sort first
        __shared__ uint32_t s_count[kRefThreads / 64];
        if ((threadIdx.x & 63u) == 0) {
            s_count[threadIdx.x >> 6] = count;
        }
        __syncthreads();
        const uint32_t staged = s_count[0] + s_count[1] + s_count[2] + s_count[3];
        if (working) {
            for (uint32_t j = 0; j < staged; ++j) {
                const float4 q = st.screen[j];
                const float4 c = st.conic[j];
                const float dx = pxc - q.x, dy = pyc - q.y;
                const float power = c.x * dx * dx + 2.0f * c.y * dx * dy + c.z * dy * dy;
                if (power > q.w) {
                    continue;
                }
                const float alpha = fminf(0.99f, c.w * athx_expf(-0.5f * power));
                if (alpha < 1.0f / 255.0f) {
                    continue;
                }
                hits += 1;
                if (done) {
                    if (hits > max_hits) {
                        break;
                    }
                    continue;
                }
                const float4 col = st.colour[j];
                const float wgt = T * alpha;
                r = r + wgt * col.x;
                g = g + wgt * col.y;
                b = b + wgt * col.z;
                dsum = dsum + wgt * q.z;
                T = T * (1.0f - alpha);
                if (dat < 0.0f && 1.0f - T >= depth_threshold) {
                    dat = q.z;
                }
                if (T < 1e-4f) {
                    done = true;
                }
            }
        }
        if (staged < kRefThreads) {
            break;  // the rest are culled
        }
        __syncthreads();
    }
    if (!inside) {
        return;
    }
    const size_t at = (size_t)py * V.width + px;
    const float coverage = 1.0f - T;
    reinterpret_cast<float4*>(V.colour)[at] =
        make_float4(r + T * V.bg[0], g + T * V.bg[1], b + T * V.bg[2], coverage + T * V.bg[3]);
    if (V.depth) {
        V.depth[at] = depth_mode == 0 ? (coverage > 1e-6f ? dsum / coverage : 0.0f) : fmaxf(dat, 0.0f);
    }
    if (hits > max_hits) {
        atomicAdd(overflow, 1u);
    }
}

// reference/image_compare.slang's code value of a linear value: the sRGB
// curve (pow through athx_powf_pos), times 255, rounded half to even.
__device__ inline uint32_t code_of(float linear) {
    const float c = fminf(fmaxf(linear, 0.0f), 1.0f);
    const float e = c <= 0.0031308f ? c * 12.92f : 1.055f * athx_powf_pos(c, 1.0f / 2.4f) - 0.055f;
    return (uint32_t)rintf(e * 255.0f);
}

__global__ void compare_codes_kernel(const float4* __restrict__ a, const float4* __restrict__ b, uint32_t width,
                                     uint32_t height, uint32_t* __restrict__ rows) {
    const uint32_t y = blockIdx.x * blockDim.x + threadIdx.x;
    if (y >= height) {
        return;
    }
    uint32_t* row = rows + (size_t)y * 256;
    for (uint32_t k = 0; k < 256; ++k) {
        row[k] = 0;
    }
    for (uint32_t x = 0; x < width; ++x) {
        const float4 pa = a[(size_t)y * width + x], pb = b[(size_t)y * width + x];
        const float ca[3] = {pa.x, pa.y, pa.z}, cb[3] = {pb.x, pb.y, pb.z};
        uint32_t worst = 0;
        for (uint32_t c = 0; c < 3; ++c) {
            const int d = (int)code_of(ca[c]) - (int)code_of(cb[c]);
            worst = max(worst, (uint32_t)abs(d));
        }
        const int da = (int)rintf(fminf(fmaxf(pa.w, 0.0f), 1.0f) * 255.0f) -
                       (int)rintf(fminf(fmaxf(pb.w, 0.0f), 1.0f) * 255.0f);
        worst = max(worst, (uint32_t)abs(da));
        row[worst] += 1;
    }
}

// The HDR bin of a relative difference: 0 for none, else 8 a binade from
// 2^-16, clamped to [1, 255] -- floor((log2 v + 16) * 8) + 1, taken from the
// exponent and comparisons of the mantissa with 2^(j/8), so no log is
// rounded.
__device__ inline uint32_t hdr_bin(float v) {
    if (!(v > 0.0f)) {
        return 0u;
    }
    if (!(v < 3.0e38f)) {
        return 255u;
    }
    int e = 0;
    const float mant = frexpf(v, &e);  // v = mant * 2^e, mant in [0.5, 1)
    const float m = mant * 2.0f;       // [1, 2), exponent e - 1
    const float kSteps[7] = {1.0905077f, 1.1892071f, 1.2968396f, 1.4142135f, 1.5422108f, 1.6817929f, 1.8340081f};
    int frac = 0;
    for (int j = 0; j < 7; ++j) {
        frac += m >= kSteps[j] ? 1 : 0;
    }
    const int bin = 8 * ((e - 1) + 16) + frac + 1;
    return (uint32_t)(bin < 1 ? 1 : bin > 255 ? 255 : bin);
}

__global__ void compare_hdr_kernel(const float4* __restrict__ a, const float4* __restrict__ b, uint32_t width,
                                   uint32_t height, uint32_t* __restrict__ rows, float* __restrict__ row_error) {
    const uint32_t y = blockIdx.x * blockDim.x + threadIdx.x;
    if (y >= height) {
        return;
    }
    uint32_t* row = rows + (size_t)y * 256;
    for (uint32_t k = 0; k < 256; ++k) {
        row[k] = 0;
    }
    float sum = 0.0f, carry = 0.0f;  // Kahan: a row of small terms into a large sum
    for (uint32_t x = 0; x < width; ++x) {
        const float4 pa = a[(size_t)y * width + x], pb = b[(size_t)y * width + x];
        const float va[4] = {pa.x, pa.y, pa.z, pa.w}, vb[4] = {pb.x, pb.y, pb.z, pb.w};
        float worst = 0.0f, squared = 0.0f;
        for (uint32_t c = 0; c < 4; ++c) {
            const float d = fabsf(va[c] - vb[c]);
            worst = fmaxf(worst, d / fmaxf(fabsf(vb[c]), 1e-3f));
            squared = squared + d * d / (vb[c] * vb[c] + 1e-2f);
        }
        row[hdr_bin(worst)] += 1;
        const float term = squared * 0.25f;
        const float t = sum + term;
        carry = carry + (fabsf(sum) >= fabsf(term) ? (sum - t) + term : (term - t) + sum);
        sum = t;
    }
    row_error[y] = sum + carry;
}

__global__ void rows_reduce_kernel(const uint32_t* __restrict__ rows, uint32_t height, uint32_t* __restrict__ bins) {
    const uint32_t bin = threadIdx.x;
    uint32_t sum = 0;
    for (uint32_t r = 0; r < height; ++r) {
        sum += rows[(size_t)r * 256 + bin];
    }
    bins[bin] = sum;
}

__global__ void float_sum_kernel(const float* __restrict__ values, uint32_t count, float* __restrict__ total) {
    float sum = 0.0f, carry = 0.0f;
    for (uint32_t i = 0; i < count; ++i) {
        const float term = values[i];
        const float t = sum + term;
        carry = carry + (fabsf(sum) >= fabsf(term) ? (sum - t) + term : (term - t) + sum);
        sum = t;
    }
    *total = sum + carry;
}

}  // namespace

hipError_t launch_ref_project(const float* positions, const uint32_t* shape, const uint32_t* sh, uint32_t n,
                              uint32_t rest_per_colour, uint32_t sh_limit, uint32_t antialias, uint32_t linear,
                              const DView* view, float* records, uint32_t* keys, uint32_t* values, hipStream_t s,
                              uint32_t base) {
    if (n == 0) {
        return hipSuccess;
    }
    RefArgs a{positions, shape, sh, n, rest_per_colour, (rest_per_colour * 3 + 1) / 2, sh_limit, antialias, linear,
              base};
    hipLaunchKernelGGL(ref_project_kernel, dim3(div_up(n, kRefThreads)), dim3(kRefThreads), 0, s, a, view, records,
                       keys, values);
    return hipGetLastError();
}

hipError_t launch_ref_blend(const DView* view, uint32_t width, uint32_t height, const float* records,
                            const uint32_t* sorted_keys, const uint32_t* sorted, uint32_t n, uint32_t max_hits,
                            uint32_t depth_mode, float depth_threshold, uint32_t* overflow, hipStream_t s) {
    const uint32_t blocks = div_up(width, 16u) * div_up(height, 16u);
    hipLaunchKernelGGL(ref_blend_kernel, dim3(blocks), dim3(kRefThreads), 0, s, view, records, sorted_keys, sorted, n,
                       max_hits, depth_mode, depth_threshold, overflow);
    return hipGetLastError();
}

hipError_t launch_compare_codes(const float* a, const float* b, uint32_t width, uint32_t height, uint32_t* rows,
                                uint32_t* bins, hipStream_t s) {
    hipLaunchKernelGGL(compare_codes_kernel, dim3(div_up(height, 64u)), dim3(64), 0, s,
                       reinterpret_cast<const float4*>(a), reinterpret_cast<const float4*>(b), width, height, rows);
    hipLaunchKernelGGL(rows_reduce_kernel, dim3(1), dim3(256), 0, s, rows, height, bins);
    return hipGetLastError();
}

hipError_t launch_compare_hdr(const float* a, const float* b, uint32_t width, uint32_t height, uint32_t* rows,
                              float* row_error, uint32_t* bins, float* total, hipStream_t s) {
    hipLaunchKernelGGL(compare_hdr_kernel, dim3(div_up(height, 64u)), dim3(64), 0, s,
                       reinterpret_cast<const float4*>(a), reinterpret_cast<const float4*>(b), width, height, rows,
                       row_error);
    hipLaunchKernelGGL(rows_reduce_kernel, dim3(1), dim3(256), 0, s, rows, height, bins);
    hipLaunchKernelGGL(float_sum_kernel, dim3(1), dim3(1), 0, s, row_error, height, total);
    return hipGetLastError();
}

}  // namespace athx
