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
#include "athx_internal.h"

namespace athx {
namespace {

constexpr float kSH0 = 0.28209479177387814f;
constexpr float kC1 = 0.4886025119029199f;
constexpr float kC2_0 = 1.0925484305920792f, kC2_1 = -1.0925484305920792f, kC2_2 = 0.31539156525252005f,
                kC2_3 = -1.0925484305920792f, kC2_4 = 0.5462742152960396f;
constexpr float kC3_0 = -0.5900435899266435f, kC3_1 = 2.890611442640554f, kC3_2 = -0.4570457994644658f,
                kC3_3 = 0.3731763325901154f, kC3_4 = -0.4570457994644658f, kC3_5 = 1.445305721320277f,
                kC3_6 = -0.5900435899266435f;

// The 16 real harmonics at dir and their gradients with respect to dir
// (sh.slang:25-51 and shBasisValue :59-80, without the constant term's
// derivative, which is zero).
__device__ inline void sh_basis(uint32_t degree, float x, float y, float z, float* Y, float* Dx, float* Dy, float* Dz) {
    for (int k = 0; k < 16; ++k) {
        Y[k] = 0.0f; Dx[k] = 0.0f; Dy[k] = 0.0f; Dz[k] = 0.0f;
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
    const float xx = x * x, yy = y * y, zz = z * z;
    Y[4] = kC2_0 * x * y; Dx[4] = kC2_0 * y; Dy[4] = kC2_0 * x;
    Y[5] = kC2_1 * y * z; Dy[5] = kC2_1 * z; Dz[5] = kC2_1 * y;
    Y[6] = kC2_2 * (2.0f * zz - xx - yy); Dx[6] = -2.0f * kC2_2 * x; Dy[6] = -2.0f * kC2_2 * y; Dz[6] = 4.0f * kC2_2 * z;
    Y[7] = kC2_3 * x * z; Dx[7] = kC2_3 * z; Dz[7] = kC2_3 * x;
    Y[8] = kC2_4 * (xx - yy); Dx[8] = 2.0f * kC2_4 * x; Dy[8] = -2.0f * kC2_4 * y;
    if (degree < 3) {
        return;
    }
    Y[9] = kC3_0 * y * (3.0f * xx - yy); Dx[9] = kC3_0 * 6.0f * x * y; Dy[9] = kC3_0 * (3.0f * xx - 3.0f * yy);
    Y[10] = kC3_1 * x * y * z; Dx[10] = kC3_1 * y * z; Dy[10] = kC3_1 * x * z; Dz[10] = kC3_1 * x * y;
    Y[11] = kC3_2 * y * (4.0f * zz - xx - yy); Dx[11] = kC3_2 * (-2.0f * x * y);
    Dy[11] = kC3_2 * (4.0f * zz - xx - 3.0f * yy); Dz[11] = kC3_2 * 8.0f * y * z;
    Y[12] = kC3_3 * z * (2.0f * zz - 3.0f * xx - 3.0f * yy); Dx[12] = kC3_3 * (-6.0f * x * z);
    Dy[12] = kC3_3 * (-6.0f * y * z); Dz[12] = kC3_3 * (6.0f * zz - 3.0f * xx - 3.0f * yy);
    Y[13] = kC3_4 * x * (4.0f * zz - xx - yy); Dx[13] = kC3_4 * (4.0f * zz - 3.0f * xx - yy);
    Dy[13] = kC3_4 * (-2.0f * x * y); Dz[13] = kC3_4 * 8.0f * x * z;
    Y[14] = kC3_5 * z * (xx - yy); Dx[14] = kC3_5 * 2.0f * x * z; Dy[14] = kC3_5 * (-2.0f * y * z);
    Dz[14] = kC3_5 * (xx - yy);
    Y[15] = kC3_6 * x * (xx - 3.0f * yy); Dx[15] = kC3_6 * (3.0f * xx - 3.0f * yy); Dy[15] = kC3_6 * (-6.0f * x * y);
}

template <bool kHalf>
__device__ inline float sh_load(const float* sh, const uint16_t* shh, size_t at) {
    if constexpr (kHalf) {
        return (float)__builtin_bit_cast(_Float16, shh[at]);
    } else {
        return sh[at];
    }
}

template <bool kHalf>
__global__ void __launch_bounds__(256) project_backward_kernel(ProjectArgs a, const DView* __restrict__ views,
                                                               uint32_t nviews, const uint32_t* __restrict__ touched,
                                                               const float* __restrict__ grad2d, athx_grads out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) {
        return;
    }
    bool any = false;
    for (uint32_t v = 0; v < nviews; ++v) {
        any = any || touched[(size_t)v * a.n + i] != 0;
    }
    if (!any) {
        return;
    }
    const float px = a.means[i * 3 + 0], py = a.means[i * 3 + 1], pz = a.means[i * 3 + 2];
    const float ls[3] = {a.log_scales[i * 3 + 0], a.log_scales[i * 3 + 1], a.log_scales[i * 3 + 2]};
    const float s[3] = {athx_expf(ls[0]), athx_expf(ls[1]), athx_expf(ls[2])};
    const float opacity = athx_sigmoidf(a.logits[i]);
    const float q0x = a.quats[i * 4 + 0], q0y = a.quats[i * 4 + 1], q0z = a.quats[i * 4 + 2], q0w = a.quats[i * 4 + 3];
    const float qn = q0x * q0x + q0y * q0y + q0z * q0z + q0w * q0w;
    float qx = 0.0f, qy = 0.0f, qz = 0.0f, qw = 1.0f, qinv = 0.0f;
    if (qn > 0.0f) {
        qinv = 1.0f / sqrtf(qn);
        qx = q0x * qinv; qy = q0y * qinv; qz = q0z * qinv; qw = q0w * qinv;
    }
    float R[3][3];
    R[0][0] = 1.0f - 2.0f * (qy * qy + qz * qz); R[0][1] = 2.0f * (qx * qy - qw * qz); R[0][2] = 2.0f * (qx * qz + qw * qy);
    R[1][0] = 2.0f * (qx * qy + qw * qz); R[1][1] = 1.0f - 2.0f * (qx * qx + qz * qz); R[1][2] = 2.0f * (qy * qz - qw * qx);
    R[2][0] = 2.0f * (qx * qz - qw * qy); R[2][1] = 2.0f * (qy * qz + qw * qx); R[2][2] = 1.0f - 2.0f * (qx * qx + qy * qy);
    float M[3][3];
    for (int r = 0; r < 3; ++r) {
        for (int c = 0; c < 3; ++c) {
            M[r][c] = R[r][c] * s[c];
        }
    }

    float g_p[3] = {0.f, 0.f, 0.f};
    float g_M[3][3] = {{0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}};
    float g_opacity = 0.0f;
    const size_t row = (size_t)i * a.sh_coeffs * 3;

    for (uint32_t v = 0; v < nviews; ++v) {
        const size_t slot = (size_t)v * a.n + i;
        if (touched[slot] == 0) {
            continue;
        }
        const float* G = grad2d + slot * kGrad2d;
        const float g_cx = G[0], g_cy = G[1], g_ka = G[2], g_kb = G[3], g_kc = G[4], g_al = G[5];
        const float g_col[3] = {G[6], G[7], G[8]};
        const DView& V = views[v];
        const float* m = V.m;
        const float tx = m[0] * px + m[1] * py + m[2] * pz + m[3];
        const float ty = m[4] * px + m[5] * py + m[6] * pz + m[7];
        const float tz = m[8] * px + m[9] * py + m[10] * pz + m[11];
        float jx0, jx2, jy1, jy2;
        bool clamp_x = false, clamp_y = false;
        float lim_x = 0.0f, lim_y = 0.0f, invz = 0.0f;
        if (V.orthographic) {
            jx0 = V.focal_x; jx2 = 0.0f; jy1 = V.focal_y; jy2 = 0.0f;
        } else {
            invz = 1.0f / tz;
            const float W = (float)V.width, H = (float)V.height;
            const float limlox = (-0.15f * W - V.centre_x) / V.focal_x;
            const float limhix = (1.15f * W - V.centre_x) / V.focal_x;
            const float limloy = (-0.15f * H - V.centre_y) / V.focal_y;
            const float limhiy = (1.15f * H - V.centre_y) / V.focal_y;
            const float ux = tx * invz, uy = ty * invz;
            clamp_x = ux < limlox || ux > limhix;
            clamp_y = uy < limloy || uy > limhiy;
            lim_x = ux < limlox ? limlox : limhix;
            lim_y = uy < limloy ? limloy : limhiy;
            const float txc = fminf(fmaxf(ux, limlox), limhix) * tz;
            const float tyc = fminf(fmaxf(uy, limloy), limhiy) * tz;
            jx0 = V.focal_x * invz;
            jx2 = -V.focal_x * txc * invz * invz;
            jy1 = V.focal_y * invz;
            jy2 = -V.focal_y * tyc * invz * invz;
        }
        const float t3[3] = {jx0 * m[0] + jx2 * m[8], jx0 * m[1] + jx2 * m[9], jx0 * m[2] + jx2 * m[10]};
        const float u3[3] = {jy1 * m[4] + jy2 * m[8], jy1 * m[5] + jy2 * m[9], jy1 * m[6] + jy2 * m[10]};
        float ax[3], ay[3];
        for (int c = 0; c < 3; ++c) {
            ax[c] = t3[0] * M[0][c] + t3[1] * M[1][c] + t3[2] * M[2][c];
            ay[c] = u3[0] * M[0][c] + u3[1] * M[1][c] + u3[2] * M[2][c];
        }
        const float ca = ax[0] * ax[0] + ax[1] * ax[1] + ax[2] * ax[2];
        const float cb = ax[0] * ay[0] + ax[1] * ay[1] + ax[2] * ay[2];
        const float cc = ay[0] * ay[0] + ay[1] * ay[1] + ay[2] * ay[2];
        const float det_before = ca * cc - cb * cb;
        const float A = ca + 0.3f, C = cc + 0.3f;
        const float D = A * C - cb * cb;
        const float invD = 1.0f / D;

        // Opacity and the compensation: alpha = paid * o + (1 - paid) / 255
        // (every projected splat took that branch: the other one is culled).
        float g_A = 0.0f, g_C = 0.0f, g_b = 0.0f, g_D = 0.0f, g_a = 0.0f, g_c = 0.0f;
        if (a.antialias) {
            const float paid = sqrtf(fmaxf(det_before, 0.0f) * invD);
            g_opacity += g_al * paid;
            if (paid > 0.0f) {
                const float g_paid = g_al * (opacity - 1.0f / 255.0f);
                const float g_db = g_paid * 0.5f / (paid * D);
                g_D += -g_paid * 0.5f * paid * invD;
                g_a += g_db * cc;
                g_c += g_db * ca;
                g_b += -2.0f * cb * g_db;
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
        g_b += -2.0f * cb * g_D;
        g_a += g_A;
        g_c += g_C;
        float g_ax[3], g_ay[3];
        for (int c = 0; c < 3; ++c) {
            g_ax[c] = 2.0f * g_a * ax[c] + g_b * ay[c];
            g_ay[c] = 2.0f * g_c * ay[c] + g_b * ax[c];
        }
        float g_t3[3], g_u3[3];
        for (int r = 0; r < 3; ++r) {
            g_t3[r] = g_ax[0] * M[r][0] + g_ax[1] * M[r][1] + g_ax[2] * M[r][2];
            g_u3[r] = g_ay[0] * M[r][0] + g_ay[1] * M[r][1] + g_ay[2] * M[r][2];
            for (int c = 0; c < 3; ++c) {
                g_M[r][c] += t3[r] * g_ax[c] + u3[r] * g_ay[c];
            }
        }
        float g_t[3] = {0.f, 0.f, 0.f};
        if (!V.orthographic) {
            const float g_jx0 = g_t3[0] * m[0] + g_t3[1] * m[1] + g_t3[2] * m[2];
            const float g_jx2 = g_t3[0] * m[8] + g_t3[1] * m[9] + g_t3[2] * m[10];
            const float g_jy1 = g_u3[0] * m[4] + g_u3[1] * m[5] + g_u3[2] * m[6];
            const float g_jy2 = g_u3[0] * m[8] + g_u3[1] * m[9] + g_u3[2] * m[10];
            const float iz2 = invz * invz;
            g_t[2] += -(g_jx0 * V.focal_x + g_jy1 * V.focal_y) * iz2;
            if (clamp_x) {
                g_t[2] += g_jx2 * V.focal_x * lim_x * iz2;
            } else {
                g_t[0] += -g_jx2 * V.focal_x * iz2;
                g_t[2] += g_jx2 * 2.0f * V.focal_x * tx * iz2 * invz;
            }
            if (clamp_y) {
                g_t[2] += g_jy2 * V.focal_y * lim_y * iz2;
            } else {
                g_t[1] += -g_jy2 * V.focal_y * iz2;
                g_t[2] += g_jy2 * 2.0f * V.focal_y * ty * iz2 * invz;
            }
            // centre = focal * t.xy / t.z + principal point.
            g_t[0] += g_cx * V.focal_x * invz;
            g_t[1] += g_cy * V.focal_y * invz;
            g_t[2] += -(g_cx * V.focal_x * tx + g_cy * V.focal_y * ty) * iz2;
        } else {
            g_t[0] += g_cx * V.focal_x;
            g_t[1] += g_cy * V.focal_y;
        }
        for (int c = 0; c < 3; ++c) {
            g_p[c] += m[c] * g_t[0] + m[4 + c] * g_t[1] + m[8 + c] * g_t[2];
        }

        // Colour: 0.5 + SH0 * sh0 + rest(dir), then max(c, 0) or sRGB.
        const float dx = px - V.eye[0], dy = py - V.eye[1], dz = pz - V.eye[2];
        const float len = sqrtf(dx * dx + dy * dy + dz * dz);
        const float dinv = 1.0f / len;
        const float ex = dx * dinv, ey = dy * dinv, ez = dz * dinv;
        float Y[16], Dx[16], Dy[16], Dz[16];
        sh_basis(a.sh_degree, ex, ey, ez, Y, Dx, Dy, Dz);
        const uint32_t nk = (a.sh_degree + 1) * (a.sh_degree + 1);
        float g_dir[3] = {0.f, 0.f, 0.f};
        for (uint32_t ch = 0; ch < 3; ++ch) {
            float cv = 0.5f;
            for (uint32_t k = 0; k < nk; ++k) {
                cv += Y[k] * sh_load<kHalf>(a.sh, a.sh_half, row + k * 3 + ch);
            }
            float g_cv;
            if (a.linear) {
                g_cv = cv > 0.0f ? g_col[ch] : 0.0f;
            } else if (cv <= 0.0f) {
                g_cv = 0.0f;
            } else if (cv <= 0.04045f) {
                g_cv = g_col[ch] / 12.92f;
            } else {
                g_cv = g_col[ch] * (2.4f / 1.055f) * athx_powf_pos((cv + 0.055f) / 1.055f, 1.4f);
            }
            if (g_cv == 0.0f) {
                continue;
            }
            if (out.sh) {
                for (uint32_t k = 0; k < nk; ++k) {
                    out.sh[row + k * 3 + ch] += Y[k] * g_cv;
                }
            }
            for (uint32_t k = 1; k < nk; ++k) {
                const float coef = sh_load<kHalf>(a.sh, a.sh_half, row + k * 3 + ch) * g_cv;
                g_dir[0] += coef * Dx[k];
                g_dir[1] += coef * Dy[k];
                g_dir[2] += coef * Dz[k];
            }
        }
        const float dot = ex * g_dir[0] + ey * g_dir[1] + ez * g_dir[2];
        g_p[0] += (g_dir[0] - ex * dot) * dinv;
        g_p[1] += (g_dir[1] - ey * dot) * dinv;
        g_p[2] += (g_dir[2] - ez * dot) * dinv;
    }

    if (out.means) {
        out.means[i * 3 + 0] += g_p[0];
        out.means[i * 3 + 1] += g_p[1];
        out.means[i * 3 + 2] += g_p[2];
    }
    if (out.log_scales) {
        for (int c = 0; c < 3; ++c) {
            const float g_s = g_M[0][c] * R[0][c] + g_M[1][c] * R[1][c] + g_M[2][c] * R[2][c];
            out.log_scales[i * 3 + c] += g_s * s[c];
        }
    }
    if (out.quats && qn > 0.0f) {
        float gR[3][3];
        for (int r = 0; r < 3; ++r) {
            for (int c = 0; c < 3; ++c) {
                gR[r][c] = g_M[r][c] * s[c];
            }
        }
        const float gx = 2.0f * (qy * (gR[0][1] + gR[1][0]) + qz * (gR[0][2] + gR[2][0]) + qw * (gR[2][1] - gR[1][2])) -
                         4.0f * qx * (gR[1][1] + gR[2][2]);
        const float gy = 2.0f * (qx * (gR[0][1] + gR[1][0]) + qz * (gR[1][2] + gR[2][1]) + qw * (gR[0][2] - gR[2][0])) -
                         4.0f * qy * (gR[0][0] + gR[2][2]);
        const float gz = 2.0f * (qx * (gR[0][2] + gR[2][0]) + qy * (gR[1][2] + gR[2][1]) + qw * (gR[1][0] - gR[0][1])) -
                         4.0f * qz * (gR[0][0] + gR[1][1]);
        const float gw = 2.0f * (qz * (gR[1][0] - gR[0][1]) + qy * (gR[0][2] - gR[2][0]) + qx * (gR[2][1] - gR[1][2]));
        const float d = qx * gx + qy * gy + qz * gz + qw * gw;
        out.quats[i * 4 + 0] += (gx - qx * d) * qinv;
        out.quats[i * 4 + 1] += (gy - qy * d) * qinv;
        out.quats[i * 4 + 2] += (gz - qz * d) * qinv;
        out.quats[i * 4 + 3] += (gw - qw * d) * qinv;
    }
    if (out.opacity_logits) {
        out.opacity_logits[i] += g_opacity * opacity * (1.0f - opacity);
    }
}

}  // namespace

hipError_t launch_project_backward(const ProjectArgs& a, const DView* views_dev, uint32_t nviews,
                                   const uint32_t* touched, const float* grad2d, const athx_grads& g, hipStream_t s) {
    if (a.n == 0 || nviews == 0) {
        return hipSuccess;
    }
    const dim3 grid(div_up(a.n, 256));
    if (a.sh_half) {
        hipLaunchKernelGGL(project_backward_kernel<true>, grid, dim3(256), 0, s, a, views_dev, nviews, touched, grad2d,
                           g);
    } else {
        hipLaunchKernelGGL(project_backward_kernel<false>, grid, dim3(256), 0, s, a, views_dev, nviews, touched,
                           grad2d, g);
    }
    return hipGetLastError();
}

}  // namespace athx
