// A trained splat cloud's PLY vertex block into the packed device layout:
// each property widened from the file's bytes (packed_records.slang:1-110),
// each vertex validated (splat_validate.slang:1-31) and the kept ones decoded
// and packed in file order (splat_decode.slang:1-173, splat_encoding.slang,
// common/packing.slang:33-57,319-328) -- by default the 3DGS encodings a PLY
// carries (opacity a logit, log-scales, below -30 a flat axis, float rotation
// w x y z, the SH DC coefficient, rest coefficients red's, green's, blue's),
// or as the layout's modes say (a .splat's bytes and linear scales).
//
// Two passes over the block, a vertex a lane: flags, their exclusive scan
// (each kept vertex's place), then the pack. The block is read as bytes at
// any alignment: a PLY's properties need not be 4-byte aligned.
#include "athx_internal.h"

namespace athx {
namespace {

constexpr uint32_t kIngestThreads = 256;

// Field f of the vertex at rec: the property's bytes (little-endian) as the
// float of their value, a double rounded to the nearest float.
__device__ inline float widen(const uint8_t* __restrict__ rec, uint32_t code) {
    const uint8_t* p = rec + (code & 0xFFFFFFu);
    const uint32_t type = code >> 24;
    if (type == ATHX_PLY_U8) {
        return (float)p[0];
    }
    if (type == ATHX_PLY_I8) {
        return (float)(int8_t)p[0];
    }
    if (type == ATHX_PLY_U16 || type == ATHX_PLY_I16) {
        const uint16_t h = (uint16_t)(p[0] | (p[1] << 8));
        return type == ATHX_PLY_U16 ? (float)h : (float)(int16_t)h;
    }
    const uint32_t lo = (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
    if (type == ATHX_PLY_F64) {
        const uint32_t hi =
            (uint32_t)p[4] | ((uint32_t)p[5] << 8) | ((uint32_t)p[6] << 16) | ((uint32_t)p[7] << 24);
        return (float)__builtin_bit_cast(double, (uint64_t)hi << 32 | lo);
    }
    if (type == ATHX_PLY_F32) {
        return __builtin_bit_cast(float, lo);
    }
    if (type == ATHX_PLY_U32) {
        return (float)lo;
    }
    return (float)(int32_t)lo;
}

__device__ inline bool finite_(float v) { return (__builtin_bit_cast(uint32_t, v) & 0x7F800000u) != 0x7F800000u; }

// splat_encoding.slang's decodeOpacity, decodeLogScale, decodeColour and
// decodeRotation, by the layout's modes (RawSplats.h:85-101).
__device__ inline float opacity_of(float v, uint32_t mode) {
    return mode == 0 ? athx_sigmoidf(v) : mode == 2 ? v / 255.0f : v;
}

// A flat axis: -inf, or a hard surfel's p as packing.slang's surfelHardnessLog keeps it.
__device__ inline float log_scale_of(float v, uint32_t mode) {
    if (mode == 2) {
        return v / 16.0f - 10.0f;
    }
    if (mode == 0) {
        return v < -30.0f ? -__builtin_inff() : v;
    }
    if (v < 0.0f) {
        const float p = -v;
        return p <= 2.001f ? -65000.0f : -(1024.0f + 64.0f * (fminf(p, 32.0f) - 2.0f));
    }
    return v == 0.0f ? -__builtin_inff() : athx_logf(fmaxf(v, 1e-30f));
}

__device__ inline float colour_of(float v, uint32_t mode) {
    constexpr float kSh0 = 0.28209479177387814f;
    if (mode == 0) {
        return 0.5f + kSh0 * v;
    }
    if (mode == 2) {
        return v / 255.0f;
    }
    if (mode == 3) {
        return 0.5f + kSh0 * ((v / 255.0f - 0.5f) / 0.15f);
    }
    return v;
}

__device__ inline float rotation_of(float v, uint32_t mode) { return mode == 1 ? (v - 128.0f) / 128.0f : v; }

// At most one flat axis; the others finite and under 60 in magnitude
// (splat_encoding.slang's scalesValid).
__device__ inline bool scales_valid(float a, float b, float c) {
    const bool fa = a <= -1024.0f, fb = b <= -1024.0f, fc = c <= -1024.0f;
    const uint32_t flats = (fa ? 1u : 0u) + (fb ? 1u : 0u) + (fc ? 1u : 0u);
    const float ka = fa ? 0.0f : a, kb = fb ? 0.0f : b, kc = fc ? 0.0f : c;
    return flats <= 1u && finite_(ka) && finite_(kb) && finite_(kc) && fabsf(ka) < 60.0f && fabsf(kb) < 60.0f &&
           fabsf(kc) < 60.0f;
}

__device__ inline uint32_t half_bits(float v) {
    // Clamped into the f16 range first (packing.slang's halfSafe), then
    // rounded to the nearest half, ties to even.
    const float c = fminf(fmaxf(v, -65000.0f), 65000.0f);
    return (uint32_t)__builtin_bit_cast(uint16_t, (_Float16)c);
}

__device__ inline bool keep_vertex(const uint8_t* rec, const IngestFields& f, float& px, float& py, float& pz,
                                   float& alpha, float ls[3]) {
    px = widen(rec, f.code[0]);
    py = widen(rec, f.code[1]);
    pz = widen(rec, f.code[2]);
    alpha = opacity_of(widen(rec, f.code[3]), f.opacity_mode);
    ls[0] = log_scale_of(widen(rec, f.code[4]), f.scale_mode);
    ls[1] = log_scale_of(widen(rec, f.code[5]), f.scale_mode);
    ls[2] = log_scale_of(widen(rec, f.code[6]), f.scale_mode);
    const bool finite = finite_(px) && finite_(py) && finite_(pz);
    return (finite && scales_valid(ls[0], ls[1], ls[2]) && alpha >= 1.0f / 255.0f) || (f.keep_faint != 0 && finite);
}

__global__ void __launch_bounds__(kIngestThreads) ingest_flags_kernel(const uint8_t* __restrict__ vertices,
                                                                     IngestFields f, uint32_t* __restrict__ keep) {
    const uint32_t i = blockIdx.x * kIngestThreads + threadIdx.x;
    if (i >= f.count) {
        return;
    }
    float px, py, pz, alpha, ls[3];
    keep[i] = keep_vertex(vertices + (size_t)i * f.stride, f, px, py, pz, alpha, ls) ? 1u : 0u;
}

// Smallest-three quaternion word of (x, y, z, w) (packing.slang:33-57):
// normalised, the largest component (the first of equals) dropped and made
// non-negative, the other three as 10-bit codes of [-sqrt(1/2), sqrt(1/2)],
// rounded half up.
__device__ inline uint32_t quaternion_word(float x, float y, float z, float w) {
    const float len = sqrtf(((x * x + y * y) + z * z) + w * w);
    float q[4] = {x / len, y / len, z / len, w / len};
    uint32_t largest = 0;
    float best = fabsf(q[0]);
#pragma unroll
    for (uint32_t k = 1; k < 4; ++k) {
        if (fabsf(q[k]) > best) {
            best = fabsf(q[k]);
            largest = k;
        }
    }
    const float sign = q[largest] < 0.0f ? -1.0f : 1.0f;
    uint32_t word = largest, shift = 22;
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k) {
        if (k == largest) {
            continue;
        }
        const float unit = fminf(fmaxf((sign * q[k]) / 0.70710678f * 0.5f + 0.5f, 0.0f), 1.0f);
        word |= (uint32_t)floorf(unit * 1023.0f + 0.5f) << shift;
        shift -= 10;
    }
    return word;
}

__global__ void __launch_bounds__(kIngestThreads) ingest_pack_kernel(const uint8_t* __restrict__ vertices,
                                                                    IngestFields f,
                                                                    const uint32_t* __restrict__ keep,
                                                                    const uint32_t* __restrict__ dest,
                                                                    athx_ingest_out out) {
    const uint32_t i = blockIdx.x * kIngestThreads + threadIdx.x;
    if (i >= f.count || keep[i] == 0) {
        return;
    }
    const uint8_t* rec = vertices + (size_t)i * f.stride;
    float px, py, pz, alpha, ls[3];
    keep_vertex(rec, f, px, py, pz, alpha, ls);
    const uint32_t j = dest[i];
    reinterpret_cast<float4*>(out.positions)[j] = make_float4(px, py, pz, fminf(fmaxf(alpha, 0.0f), 5.0f));
    if (out.origin) {
        out.origin[j] = i;
    }
    // Rotation: the file's rot_0 is w. A quaternion of no length is the identity.
    const float qw = rotation_of(widen(rec, f.code[7]), f.rotation_mode),
                qx = rotation_of(widen(rec, f.code[8]), f.rotation_mode),
                qy = rotation_of(widen(rec, f.code[9]), f.rotation_mode),
                qz = rotation_of(widen(rec, f.code[10]), f.rotation_mode);
    const bool some = sqrtf(((qx * qx + qy * qy) + qz * qz) + qw * qw) > 1e-8f;
    const uint32_t rotation = some ? quaternion_word(qx, qy, qz, qw) : quaternion_word(0.0f, 0.0f, 0.0f, 1.0f);
    const float r = colour_of(widen(rec, f.code[11]), f.colour_mode);
    const float g = colour_of(widen(rec, f.code[12]), f.colour_mode);
    const float b = colour_of(widen(rec, f.code[13]), f.colour_mode);
    reinterpret_cast<uint4*>(out.shape)[j] =
        make_uint4(rotation, half_bits(ls[0]) | (half_bits(ls[1]) << 16), half_bits(ls[2]) | (half_bits(r) << 16),
                   half_bits(g) | (half_bits(b) << 16));
    if (f.rest == 0 || !out.sh) {
        return;
    }
    // The kept rest coefficients, rgb a basis (their codes in that order),
    // two halves a word, the last word's high half 0 when the count is odd.
    uint32_t* row = out.sh + (size_t)j * f.sh_words;
    for (uint32_t k = 0; k < f.rest; k += 2) {
        const uint32_t lo = half_bits(widen(rec, f.code[kIngestBase + k]));
        const uint32_t hi = k + 1 < f.rest ? half_bits(widen(rec, f.code[kIngestBase + k + 1])) : 0u;
        row[k >> 1] = lo | (hi << 16);
    }
}

}  // namespace

hipError_t launch_ingest_flags(const uint8_t* vertices, const IngestFields& f, uint32_t* keep, hipStream_t s) {
    if (f.count == 0) {
        return hipSuccess;
    }
    hipLaunchKernelGGL(ingest_flags_kernel, dim3(div_up(f.count, kIngestThreads)), dim3(kIngestThreads), 0, s,
                       vertices, f, keep);
    return hipGetLastError();
}

hipError_t launch_ingest_pack(const uint8_t* vertices, const IngestFields& f, const uint32_t* keep,
                              const uint32_t* dest, const athx_ingest_out& out, hipStream_t s) {
    if (f.count == 0) {
        return hipSuccess;
    }
    hipLaunchKernelGGL(ingest_pack_kernel, dim3(div_up(f.count, kIngestThreads)), dim3(kIngestThreads), 0, s,
                       vertices, f, keep, dest, out);
    return hipGetLastError();
}

}  // namespace athx
