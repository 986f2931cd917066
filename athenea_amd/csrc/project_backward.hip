// project_backward.hip -- the adjoint of the splat projection: from the
// per-slot gradients the blend's adjoint left (centre, conic, alpha, colour)
// to every Gaussian parameter, one thread per Gaussian summing its views.
//
// No counterpart in the reference, which differentiates nothing but a
// Gaussian's colour (shaders/athenea/splat/splat_blend.slang:381-398). It
// differentiates the forward of project.hip (splat_project.slang:497-1130)
// term by term: the view transform, the frustum-clamped Jacobian (no gradient
// to a clamped component), R(q) S with the quaternion's normalisation, the
// 0.3 px dilation and the Mip-Splatting compensation with its 1/255 pay-back
// (frame.slang:316-319), the conic, the opacity sigmoid, the harmonics with
// their view direction, and max(c, 0) or the sRGB decode (color.slang:7-41).
// The cut radius and the tile set are decisions, not functions: no gradient.
//
// The geometry's arithmetic is FP64: a thin, turned Gaussian's 2D
// determinant is a small difference of large products, and in FP32 the
// chain from it back to the quaternion and scales loses three digits. The
// colour's (a sum of harmonic terms and the view direction's normalisation,
// nothing cancelling) is FP32, in a kernel of its own that runs after the
// geometry's: its 16 harmonics and their three derivatives would otherwise
// hold the geometry's thread to a quarter of its occupancy. Every branch is
// the one the FP32 forward took, read from the record's branch bits, so the
// adjoint differentiates exactly the function the forward computed.
#include "athx_internal.h"

namespace athx {
namespace {

}  // namespace
constexpr double kSH0d = 0.28209479177387814;
constexpr double kC1d = 0.4886025119029199;
namespace {
constexpr double kC2d[5] = {1.0925484305920792, -1.0925484305920792, 0.31539156525252005, -1.0925484305920792,
                            0.5462742152960396};
constexpr double kC3d[7] = {-0.5900435899266435, 2.890611442640554, -0.4570457994644658, 0.3731763325901154,
                            -0.4570457994644658, 1.445305721320277, -0.5900435899266435};

// The 16 real harmonics at dir and their gradients with respect to dir
// (sh.slang:25-51 and shBasisValue :59-80).
template <typename Real>
__device__ inline void sh_basis(uint32_t degree, Real x, Real y, Real z, Real* Y, Real* Dx, Real* Dy, Real* Dz) {
    constexpr Real kSH0 = (Real)athx::kSH0d, kC1 = (Real)athx::kC1d;
    constexpr Real kC2_0 = (Real)kC2d[0], kC2_1 = (Real)kC2d[1], kC2_2 = (Real)kC2d[2], kC2_3 = (Real)kC2d[3],
                   kC2_4 = (Real)kC2d[4];
    constexpr Real kC3_0 = (Real)kC3d[0], kC3_1 = (Real)kC3d[1], kC3_2 = (Real)kC3d[2], kC3_3 = (Real)kC3d[3],
                   kC3_4 = (Real)kC3d[4], kC3_5 = (Real)kC3d[5], kC3_6 = (Real)kC3d[6];
    for (int k = 0; k < 16; ++k) {
        Y[k] = 0; Dx[k] = 0; Dy[k] = 0; Dz[k] = 0;
    }
    Y[0] = kSH0;
    if (degree == 0) {
        return;
    }
    Y[1] = -kC1 * y; Dy[1] = -kC1;
    Y[2] = kC1 * z; Dz[2] = kC1;
    Y[3] = -kC1 * x; Dx[3] = -kC1;
    if (degree < 2) {
        return;
    }
    const double xx = x * x, yy = y * y, zz = z * z;
    Y[4] = kC2_0 * x * y; Dx[4] = kC2_0 * y; Dy[4] = kC2_0 * x;
    Y[5] = kC2_1 * y * z; Dy[5] = kC2_1 * z; Dz[5] = kC2_1 * y;
    Y[6] = kC2_2 * ((Real)2.0 * zz - xx - yy); Dx[6] = -(Real)2.0 * kC2_2 * x; Dy[6] = -(Real)2.0 * kC2_2 * y;
    Dz[6] = (Real)4.0 * kC2_2 * z;
    Y[7] = kC2_3 * x * z; Dx[7] = kC2_3 * z; Dz[7] = kC2_3 * x;
    Y[8] = kC2_4 * (xx - yy); Dx[8] = (Real)2.0 * kC2_4 * x; Dy[8] = -(Real)2.0 * kC2_4 * y;
    if (degree < 3) {
        return;
    }
    Y[9] = kC3_0 * y * ((Real)3.0 * xx - yy); Dx[9] = kC3_0 * (Real)6.0 * x * y;
    Dy[9] = kC3_0 * ((Real)3.0 * xx - (Real)3.0 * yy);
    Y[10] = kC3_1 * x * y * z; Dx[10] = kC3_1 * y * z; Dy[10] = kC3_1 * x * z; Dz[10] = kC3_1 * x * y;
    Y[11] = kC3_2 * y * ((Real)4.0 * zz - xx - yy); Dx[11] = kC3_2 * (-(Real)2.0 * x * y);
    Dy[11] = kC3_2 * ((Real)4.0 * zz - xx - (Real)3.0 * yy); Dz[11] = kC3_2 * (Real)8.0 * y * z;
    Y[12] = kC3_3 * z * ((Real)2.0 * zz - (Real)3.0 * xx - (Real)3.0 * yy); Dx[12] = kC3_3 * (-(Real)6.0 * x * z);
    Dy[12] = kC3_3 * (-(Real)6.0 * y * z); Dz[12] = kC3_3 * ((Real)6.0 * zz - (Real)3.0 * xx - (Real)3.0 * yy);
    Y[13] = kC3_4 * x * ((Real)4.0 * zz - xx - yy); Dx[13] = kC3_4 * ((Real)4.0 * zz - (Real)3.0 * xx - yy);
    Dy[13] = kC3_4 * (-(Real)2.0 * x * y); Dz[13] = kC3_4 * (Real)8.0 * x * z;
    Y[14] = kC3_5 * z * (xx - yy); Dx[14] = kC3_5 * (Real)2.0 * x * z; Dy[14] = kC3_5 * (-(Real)2.0 * y * z);
    Dz[14] = kC3_5 * (xx - yy);
    Y[15] = kC3_6 * x * (xx - (Real)3.0 * yy); Dx[15] = kC3_6 * ((Real)3.0 * xx - (Real)3.0 * yy);
    Dy[15] = kC3_6 * (-(Real)6.0 * x * y);
}

template <bool kHalf>
__device__ inline double sh_load(const float* sh, const uint16_t* shh, size_t at) {
    if constexpr (kHalf) {
        return (double)(float)__builtin_bit_cast(_Float16, shh[at]);
    } else {
        return (double)sh[at];
    }
}

template <bool kHalf>
__global__ void __launch_bounds__(256) project_backward_kernel(ProjectArgs a, const DView* __restrict__ views,
                                                               uint32_t nviews, const uint32_t* __restrict__ touched,
                                                               const float* __restrict__ records,
                                                               const float* __restrict__ grad2d, athx_grads out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) {
        return;
    }
    bool any = false;
    for (uint32_t v = 0; v < nviews; ++v) {
        any = any || touched[(size_t)v * a.slot_stride + a.slot_base + i] != 0;
    }
    if (!any) {
        if (out.overwrite) {
            // Not drawn in any view: its geometry gradients are zero.
            for (int c = 0; c < 3; ++c) {
                if (out.means) out.means[i * 3 + c] = 0.0f;
                if (out.log_scales) out.log_scales[i * 3 + c] = 0.0f;
            }
            for (int c = 0; c < 4; ++c) {
                if (out.quats) out.quats[i * 4 + c] = 0.0f;
            }
            if (out.opacity_logits) out.opacity_logits[i] = 0.0f;
        }
        return;
    }
    const double px = a.means[i * 3 + 0], py = a.means[i * 3 + 1], pz = a.means[i * 3 + 2];
    const double s[3] = {exp((double)a.log_scales[i * 3 + 0]), exp((double)a.log_scales[i * 3 + 1]),
                         exp((double)a.log_scales[i * 3 + 2])};
    const double opacity = 1.0 / (1.0 + exp(-(double)a.logits[i]));
    const double q0x = a.quats[i * 4 + 0], q0y = a.quats[i * 4 + 1], q0z = a.quats[i * 4 + 2],
                 q0w = a.quats[i * 4 + 3];
    const double qn = q0x * q0x + q0y * q0y + q0z * q0z + q0w * q0w;
    double qx = 0.0, qy = 0.0, qz = 0.0, qw = 1.0, qinv = 0.0;
    if (qn > 0.0) {
        qinv = 1.0 / sqrt(qn);
        qx = q0x * qinv; qy = q0y * qinv; qz = q0z * qinv; qw = q0w * qinv;
    }
    double R[3][3];
    R[0][0] = 1.0 - 2.0 * (qy * qy + qz * qz); R[0][1] = 2.0 * (qx * qy - qw * qz); R[0][2] = 2.0 * (qx * qz + qw * qy);
    R[1][0] = 2.0 * (qx * qy + qw * qz); R[1][1] = 1.0 - 2.0 * (qx * qx + qz * qz); R[1][2] = 2.0 * (qy * qz - qw * qx);
    R[2][0] = 2.0 * (qx * qz - qw * qy); R[2][1] = 2.0 * (qy * qz + qw * qx); R[2][2] = 1.0 - 2.0 * (qx * qx + qy * qy);
    double M[3][3];
    for (int r = 0; r < 3; ++r) {
        for (int c = 0; c < 3; ++c) {
            M[r][c] = R[r][c] * s[c];
        }
    }

    double g_p[3] = {0.0, 0.0, 0.0};
    double g_M[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
    double g_opacity = 0.0;
    for (uint32_t v = 0; v < nviews; ++v) {
        const size_t slot = (size_t)v * a.slot_stride + a.slot_base + i;
        if (touched[slot] == 0) {
            continue;
        }
        const uint32_t flags = __builtin_bit_cast(uint32_t, records[slot * kRecordFloats + 11]);
        const float* G = grad2d + slot * kGrad2d;
        const double g_cx = G[0], g_cy = G[1], g_ka = G[2], g_kb = G[3], g_kc = G[4], g_al = G[5];
        const DView& V = views[v];
        double m[12];
        for (int k = 0; k < 12; ++k) {
            m[k] = V.m[k];
        }
        const double fx = V.focal_x, fy = V.focal_y;
        const double tx = m[0] * px + m[1] * py + m[2] * pz + m[3];
        const double ty = m[4] * px + m[5] * py + m[6] * pz + m[7];
        const double tz = m[8] * px + m[9] * py + m[10] * pz + m[11];
        const bool clamp_x = (flags & (kFlagClampXLo | kFlagClampXHi)) != 0;
        const bool clamp_y = (flags & (kFlagClampYLo | kFlagClampYHi)) != 0;
        double jx0, jx2, jy1, jy2, invz = 0.0, lim_x = 0.0, lim_y = 0.0;
        if (V.orthographic) {
            jx0 = fx; jx2 = 0.0; jy1 = fy; jy2 = 0.0;
        } else {
            invz = 1.0 / tz;
            // The limits as the FP32 forward computed them.
            const float W = (float)V.width, H = (float)V.height;
            lim_x = (flags & kFlagClampXLo) ? (double)((-0.15f * W - V.centre_x) / V.focal_x)
                                            : (double)((1.15f * W - V.centre_x) / V.focal_x);
            lim_y = (flags & kFlagClampYLo) ? (double)((-0.15f * H - V.centre_y) / V.focal_y)
                                            : (double)((1.15f * H - V.centre_y) / V.focal_y);
            const double txc = clamp_x ? lim_x * tz : tx;
            const double tyc = clamp_y ? lim_y * tz : ty;
            jx0 = fx * invz;
            jx2 = -fx * txc * invz * invz;
            jy1 = fy * invz;
            jy2 = -fy * tyc * invz * invz;
        }
        const double t3[3] = {jx0 * m[0] + jx2 * m[8], jx0 * m[1] + jx2 * m[9], jx0 * m[2] + jx2 * m[10]};
        const double u3[3] = {jy1 * m[4] + jy2 * m[8], jy1 * m[5] + jy2 * m[9], jy1 * m[6] + jy2 * m[10]};
        double ax[3], ay[3];
        for (int c = 0; c < 3; ++c) {
            ax[c] = t3[0] * M[0][c] + t3[1] * M[1][c] + t3[2] * M[2][c];
            ay[c] = u3[0] * M[0][c] + u3[1] * M[1][c] + u3[2] * M[2][c];
        }
        const double ca = ax[0] * ax[0] + ax[1] * ax[1] + ax[2] * ax[2];
        const double cb = ax[0] * ay[0] + ax[1] * ay[1] + ax[2] * ay[2];
        const double cc = ay[0] * ay[0] + ay[1] * ay[1] + ay[2] * ay[2];
        const double det_before = ca * cc - cb * cb;
        const double A = ca + 0.3, C = cc + 0.3;
        const double D = A * C - cb * cb;
        const double invD = 1.0 / D;

        // alpha = paid * o + (1 - paid) / 255 where the pay-back branch was
        // taken (always, for a projected splat), else paid * o.
        const bool paid_back = (flags & kFlagPaidBack) != 0;
        double g_A = 0.0, g_C = 0.0, g_b = 0.0, g_D = 0.0, g_a = 0.0, g_c = 0.0;
        if (a.antialias) {
            const double paid = sqrt((det_before > 0.0 ? det_before : 0.0) * invD);
            g_opacity += g_al * paid;
            if (paid > 0.0 && det_before > 0.0) {
                const double g_paid = g_al * (paid_back ? opacity - 1.0 / 255.0 : opacity);
                const double g_db = g_paid * 0.5 / (paid * D);
                g_D += -g_paid * 0.5 * paid * invD;
                g_a += g_db * cc;
                g_c += g_db * ca;
                g_b += -2.0 * cb * g_db;
            }
        } else {
            g_opacity += g_al;
        }
        // conic = (C, -b, A) / D.
        g_C += g_ka * invD;
        g_b += -g_kb * invD;
        g_A += g_kc * invD;
        g_D += -(g_ka * C - g_kb * cb + g_kc * A) * invD * invD;
        g_A += g_D * C;
        g_C += g_D * A;
        g_b += -2.0 * cb * g_D;
        g_a += g_A;
        g_c += g_C;
        double g_ax[3], g_ay[3];
        for (int c = 0; c < 3; ++c) {
            g_ax[c] = 2.0 * g_a * ax[c] + g_b * ay[c];
            g_ay[c] = 2.0 * g_c * ay[c] + g_b * ax[c];
        }
        double g_t3[3], g_u3[3];
        for (int r = 0; r < 3; ++r) {
            g_t3[r] = g_ax[0] * M[r][0] + g_ax[1] * M[r][1] + g_ax[2] * M[r][2];
            g_u3[r] = g_ay[0] * M[r][0] + g_ay[1] * M[r][1] + g_ay[2] * M[r][2];
            for (int c = 0; c < 3; ++c) {
                g_M[r][c] += t3[r] * g_ax[c] + u3[r] * g_ay[c];
            }
        }
        double g_t[3] = {0.0, 0.0, 0.0};
        if (!V.orthographic) {
            const double g_jx0 = g_t3[0] * m[0] + g_t3[1] * m[1] + g_t3[2] * m[2];
            const double g_jx2 = g_t3[0] * m[8] + g_t3[1] * m[9] + g_t3[2] * m[10];
            const double g_jy1 = g_u3[0] * m[4] + g_u3[1] * m[5] + g_u3[2] * m[6];
            const double g_jy2 = g_u3[0] * m[8] + g_u3[1] * m[9] + g_u3[2] * m[10];
            const double iz2 = invz * invz;
            g_t[2] += -(g_jx0 * fx + g_jy1 * fy) * iz2;
            if (clamp_x) {
                g_t[2] += g_jx2 * fx * lim_x * iz2;
            } else {
                g_t[0] += -g_jx2 * fx * iz2;
                g_t[2] += g_jx2 * 2.0 * fx * tx * iz2 * invz;
            }
            if (clamp_y) {
                g_t[2] += g_jy2 * fy * lim_y * iz2;
            } else {
                g_t[1] += -g_jy2 * fy * iz2;
                g_t[2] += g_jy2 * 2.0 * fy * ty * iz2 * invz;
            }
            // centre = focal * t.xy / t.z + principal point.
            g_t[0] += g_cx * fx * invz;
            g_t[1] += g_cy * fy * invz;
            g_t[2] += -(g_cx * fx * tx + g_cy * fy * ty) * iz2;
        } else {
            g_t[0] += g_cx * fx;
            g_t[1] += g_cy * fy;
        }
        for (int c = 0; c < 3; ++c) {
            g_p[c] += m[c] * g_t[0] + m[4 + c] * g_t[1] + m[8 + c] * g_t[2];
        }

    }

    // Accumulated (+=), or written (=) when the caller asked for overwrite: the
    // old value is not read then (it may be any bits, NaN included).
    const auto put = [&](float* p, float g) { *p = out.overwrite ? g : *p + g; };
    if (out.means) {
        for (int c = 0; c < 3; ++c) {
            put(&out.means[i * 3 + c], (float)g_p[c]);
        }
    }
    if (out.log_scales) {
        for (int c = 0; c < 3; ++c) {
            const double g_s = g_M[0][c] * R[0][c] + g_M[1][c] * R[1][c] + g_M[2][c] * R[2][c];
            put(&out.log_scales[i * 3 + c], (float)(g_s * s[c]));
        }
    }
    if (out.quats && !(qn > 0.0) && out.overwrite) {
        for (int c = 0; c < 4; ++c) {
            out.quats[i * 4 + c] = 0.0f;
        }
    }
    if (out.quats && qn > 0.0) {
        double gR[3][3];
        for (int r = 0; r < 3; ++r) {
            for (int c = 0; c < 3; ++c) {
                gR[r][c] = g_M[r][c] * s[c];
            }
        }
        const double gx = 2.0 * (qy * (gR[0][1] + gR[1][0]) + qz * (gR[0][2] + gR[2][0]) + qw * (gR[2][1] - gR[1][2])) -
                          4.0 * qx * (gR[1][1] + gR[2][2]);
        const double gy = 2.0 * (qx * (gR[0][1] + gR[1][0]) + qz * (gR[1][2] + gR[2][1]) + qw * (gR[0][2] - gR[2][0])) -
                          4.0 * qy * (gR[0][0] + gR[2][2]);
        const double gz = 2.0 * (qx * (gR[0][2] + gR[2][0]) + qy * (gR[1][2] + gR[2][1]) + qw * (gR[1][0] - gR[0][1])) -
                          4.0 * qz * (gR[0][0] + gR[1][1]);
        const double gw = 2.0 * (qz * (gR[1][0] - gR[0][1]) + qy * (gR[0][2] - gR[2][0]) + qx * (gR[2][1] - gR[1][2]));
        const double d = qx * gx + qy * gy + qz * gz + qw * gw;
        put(&out.quats[i * 4 + 0], (float)((gx - qx * d) * qinv));
        put(&out.quats[i * 4 + 1], (float)((gy - qy * d) * qinv));
        put(&out.quats[i * 4 + 2], (float)((gz - qz * d) * qinv));
        put(&out.quats[i * 4 + 3], (float)((gw - qw * d) * qinv));
    }
    if (out.opacity_logits) {
        put(&out.opacity_logits[i], (float)(g_opacity * opacity * (1.0 - opacity)));
    }
}

// The colour's adjoint in FP32 (after the geometry's, whose mean gradients it
// adds to): 0.5 + SH0 sh0 + rest(dir), then max(c, 0) or the sRGB decode, each
// channel's branch the forward's, to the coefficients and, through the view
// direction, to the mean. A Gaussian's coefficients are read once, in whole
// 16-byte words where its row allows, and their gradients summed over the
// views in registers and added to memory once; every loop over the 16
// harmonics is unrolled, so nothing is indexed at run time.
template <bool kHalf>
__global__ void __launch_bounds__(256) colour_backward_kernel(ProjectArgs a, const DView* __restrict__ views,
                                                              uint32_t nviews, const uint32_t* __restrict__ touched,
                                                              const float* __restrict__ records,
                                                              const float* __restrict__ grad2d, athx_grads out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) {
        return;
    }
    bool any = false;
    for (uint32_t v = 0; v < nviews; ++v) {
        any = any || touched[(size_t)v * a.slot_stride + a.slot_base + i] != 0;
    }
    const uint32_t coeffs = a.sh_coeffs;
    const uint32_t nk = (a.sh_degree + 1) * (a.sh_degree + 1);
    const size_t row = (size_t)i * coeffs * 3;
    if (!any) {
        if (out.overwrite && out.sh) {
            // Not drawn in any view: its coefficients' gradients are zero
            // (its mean's were written by the geometry's adjoint).
            for (uint32_t k = 0; k < coeffs * 3; ++k) {
                out.sh[row + k] = 0.0f;
            }
        }
        return;
    }
    // Coefficient k of channel ch at c[k * 3 + ch]; zero past the row.
    float c[48];
    const bool whole = coeffs == 16 &&
                       ((reinterpret_cast<uintptr_t>(kHalf ? (const void*)a.sh_half : (const void*)a.sh) & 15u) == 0);
    if (whole) {
        if constexpr (kHalf) {
            const uint4* p = reinterpret_cast<const uint4*>(a.sh_half + row);
#pragma unroll
            for (uint32_t q = 0; q < 6; ++q) {
                const uint4 w = p[q];
                const uint32_t u[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
                for (uint32_t e = 0; e < 4; ++e) {
                    c[q * 8 + 2 * e] = (float)__builtin_bit_cast(_Float16, (uint16_t)(u[e] & 0xFFFFu));
                    c[q * 8 + 2 * e + 1] = (float)__builtin_bit_cast(_Float16, (uint16_t)(u[e] >> 16));
                }
            }
        } else {
            const float4* p = reinterpret_cast<const float4*>(a.sh + row);
#pragma unroll
            for (uint32_t q = 0; q < 12; ++q) {
                const float4 w = p[q];
                c[q * 4 + 0] = w.x;
                c[q * 4 + 1] = w.y;
                c[q * 4 + 2] = w.z;
                c[q * 4 + 3] = w.w;
            }
        }
    } else {
#pragma unroll
        for (uint32_t k = 0; k < 48; ++k) {
            c[k] = k < coeffs * 3 ? (float)sh_load<kHalf>(a.sh, a.sh_half, row + k) : 0.0f;
        }
    }
    float g_sh[48];
#pragma unroll
    for (uint32_t k = 0; k < 48; ++k) {
        g_sh[k] = 0.0f;
    }
    const float px = a.means[i * 3 + 0], py = a.means[i * 3 + 1], pz = a.means[i * 3 + 2];
    float g_p[3] = {0.0f, 0.0f, 0.0f};
    for (uint32_t v = 0; v < nviews; ++v) {
        const size_t slot = (size_t)v * a.slot_stride + a.slot_base + i;
        if (touched[slot] == 0) {
            continue;
        }
        const uint32_t flags = __builtin_bit_cast(uint32_t, records[slot * kRecordFloats + 11]);
        const float* G = grad2d + slot * kGrad2d;
        const float g_col[3] = {G[6], G[7], G[8]};
        const DView& V = views[v];
        const float dx = px - V.eye[0], dy = py - V.eye[1], dz = pz - V.eye[2];
        const float dinv = 1.0f / sqrtf(dx * dx + dy * dy + dz * dz);
        const float ex = dx * dinv, ey = dy * dinv, ez = dz * dinv;
        float Y[16], Dx[16], Dy[16], Dz[16];
        sh_basis(a.sh_degree, ex, ey, ez, Y, Dx, Dy, Dz);
        float g_dir[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
        for (uint32_t ch = 0; ch < 3; ++ch) {
            if (!(flags & (kFlagColourR << ch))) {
                continue;  // max(c, 0) took 0
            }
            float g_cv = g_col[ch];
            if (!a.linear) {
                if (flags & (kFlagSrgbR << ch)) {
                    float cv = 0.5f;
#pragma unroll
                    for (uint32_t k = 0; k < 16; ++k) {
                        cv += k < nk ? Y[k] * c[k * 3 + ch] : 0.0f;
                    }
                    g_cv *= (2.4f / 1.055f) * powf((cv + 0.055f) / 1.055f, 1.4f);
                } else {
                    g_cv *= 1.0f / 12.92f;
                }
            }
#pragma unroll
            for (uint32_t k = 0; k < 16; ++k) {
                if (k < nk) {
                    g_sh[k * 3 + ch] += Y[k] * g_cv;
                    if (k != 0) {
                        const float coef = c[k * 3 + ch] * g_cv;
                        g_dir[0] += coef * Dx[k];
                        g_dir[1] += coef * Dy[k];
                        g_dir[2] += coef * Dz[k];
                    }
                }
            }
        }
        const float dot = ex * g_dir[0] + ey * g_dir[1] + ez * g_dir[2];
        g_p[0] += (g_dir[0] - ex * dot) * dinv;
        g_p[1] += (g_dir[1] - ey * dot) * dinv;
        g_p[2] += (g_dir[2] - ez * dot) * dinv;
    }
    if (out.sh) {
        if (coeffs == 16 && (reinterpret_cast<uintptr_t>(out.sh) & 15u) == 0) {
            float4* p = reinterpret_cast<float4*>(out.sh + row);
#pragma unroll
            for (uint32_t q = 0; q < 12; ++q) {
                float4 w = out.overwrite ? make_float4(0.0f, 0.0f, 0.0f, 0.0f) : p[q];
                w.x += g_sh[q * 4 + 0];
                w.y += g_sh[q * 4 + 1];
                w.z += g_sh[q * 4 + 2];
                w.w += g_sh[q * 4 + 3];
                p[q] = w;
            }
        } else {
#pragma unroll
            for (uint32_t k = 0; k < 48; ++k) {
                if (k < coeffs * 3) {
                    out.sh[row + k] = (out.overwrite ? 0.0f : out.sh[row + k]) + g_sh[k];
                }
            }
        }
    }
    if (out.means) {
        for (int cc = 0; cc < 3; ++cc) {
            out.means[i * 3 + cc] += g_p[cc];
        }
    }
}

}  // namespace

hipError_t launch_project_backward(const ProjectArgs& a, const DView* views_dev, uint32_t nviews,
                                   const uint32_t* touched, const float* records, const float* grad2d,
                                   const athx_grads& g, hipStream_t s) {
    if (a.n == 0 || nviews == 0) {
        return hipSuccess;
    }
    const dim3 grid(div_up(a.n, 256));
    if (a.sh_half) {
        hipLaunchKernelGGL(project_backward_kernel<true>, grid, dim3(256), 0, s, a, views_dev, nviews, touched,
                           records, grad2d, g);
        hipLaunchKernelGGL(colour_backward_kernel<true>, grid, dim3(256), 0, s, a, views_dev, nviews, touched,
                           records, grad2d, g);
    } else {
        hipLaunchKernelGGL(project_backward_kernel<false>, grid, dim3(256), 0, s, a, views_dev, nviews, touched,
                           records, grad2d, g);
        hipLaunchKernelGGL(colour_backward_kernel<false>, grid, dim3(256), 0, s, a, views_dev, nviews, touched,
                           records, grad2d, g);
    }
    return hipGetLastError();
}

}  // namespace athx
