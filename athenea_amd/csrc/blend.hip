// blend.hip -- the per-tile front-to-back alpha blend for gfx950 and its
// adjoint.
//
// Forward: shaders/athenea/splat/splat_blend.slang:236-578 (splatBlend,
// blendPixel with composite, matte and trace off, contributionMode 0). One
// workgroup per tile, one thread per pixel; the tile's pairs are staged in
// LDS `batch` at a time (the reference stages 256, splat_blend.slang:300-311)
// and every pixel walks them from there. A wavefront whose 64 pixels are all
// done stops walking at once; the group stops loading when every wavefront
// has stopped. Each pixel also keeps its final transmittance and how many of
// its tile's pairs it read, which is all the adjoint needs to replay the walk.
//
// Adjoint: no counterpart in the reference. It walks the same pairs back to
// front, recovering the transmittance before each Gaussian from the one
// after it (T_k = T_{k+1} / (1 - alpha_k); alpha <= 0.99 keeps it stable),
// and accumulates per-slot gradients of the centre, conic, alpha and colour.
// Lanes of a wavefront walk the same pair at the same step, so a pair's
// gradient is summed across the wavefront first and added to memory once per
// wavefront.
#include "athx_internal.h"

namespace athx {
namespace {

constexpr uint32_t kMaxBatch = 256;

__device__ inline uint32_t view_of_tile(const uint32_t* tile_bases, uint32_t nviews, uint32_t tile) {
    uint32_t lo = 0, hi = nviews;  // tile_bases[nviews] is the total
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (tile_bases[mid] <= tile) {
            lo = mid;
        } else {
            hi = mid;
        }
    }
    return lo;
}

struct LdsBatch {
    float4 xy_cut_op[kMaxBatch];  // centre x, y, cutoff, alpha
    float4 conic[kMaxBatch];      // a, 2b, c, view z
    float4 colour[kMaxBatch];     // r, g, b, conic b
};

__device__ inline void stage(LdsBatch& s, uint32_t k, const float* records, uint32_t slot) {
    const float4* rec = reinterpret_cast<const float4*>(records + (size_t)slot * kRecordFloats);
    const float4 screen = rec[0];
    const float4 conic = rec[1];
    const float4 col = rec[2];
    s.xy_cut_op[k] = make_float4(screen.x, screen.y, screen.w, conic.w);
    s.conic[k] = make_float4(conic.x, 2.0f * conic.y, conic.z, screen.z);
    s.colour[k] = make_float4(col.x, col.y, col.z, conic.y);
}

// A pair's record words as stage() reads them, and their LDS form.
struct Staged {
    float4 screen, conic, col;
};

__device__ inline Staged fetch(const float* records, uint32_t slot) {
    const float4* rec = reinterpret_cast<const float4*>(records + (size_t)slot * kRecordFloats);
    return Staged{rec[0], rec[1], rec[2]};
}

__device__ inline void put(LdsBatch& s, uint32_t k, const Staged& r) {
    s.xy_cut_op[k] = make_float4(r.screen.x, r.screen.y, r.screen.w, r.conic.w);
    s.conic[k] = make_float4(r.conic.x, 2.0f * r.conic.y, r.conic.z, r.screen.z);
    s.colour[k] = make_float4(r.col.x, r.col.y, r.col.z, r.conic.y);
}

// e^(-power / 2) for a drawn pair: power <= the cut power <= 9, and above
// a tiny negative rounding floor for a positive definite conic, so the
// argument lies in [-4.5, 0+] where athx_expf_core equals athx_expf; the
// median only bounds it for inputs no record holds.
__device__ inline float blend_exp(float power) {
    return athx_expf_core(__builtin_amdgcn_fmed3f(-0.5f * power, -80.0f, 80.0f));
}

// A WAVEFRONT'S BAND. A tile's pairs are every Gaussian whose ellipse meets
// the tile; a wavefront's 64 pixels are an 8 x 8 block of it (its band: the
// squarest of 64 pixels, which the fewest ellipses meet -- against a band of
// 64 / TILE whole rows, 10 % fewer walked on the 60-view walk), and a small
// Gaussian often misses the band altogether, so that no
// lane of the wavefront could take it. Each wavefront tests its band against
// 64 staged Gaussians at once, a lane each, and walks only those that meet
// it: the walk and every result are the same, pair for pair, since a
// Gaussian skipped is one whose power exceeds the cut at every pixel centre
// of the band. The test is conservative: the least power over the band's box
// of pixel centres (zero when the box holds the centre, else the least over
// its edges, each at its clamped stationary point) must exceed the cut by
// kBandSlack * M, M = a X^2 + 2 |b| X Y + c Y^2 with X, Y the box's largest
// distances from the centre -- far above the rounding of both this float
// evaluation and the per-pixel one (a few units of 2^-24 M).
constexpr float kBandSlack = 256.0f * 0x1p-24f;

struct Band {
    float x0, x1, y0, y1;  // the band's first and last pixel centres
};

// Wavefront w of a TILE x TILE workgroup takes the 8 x 8 pixel block
// (w % (TILE / 8), w / (TILE / 8)) of the tile, lane l its pixel (l % 8, l / 8).
template <uint32_t TILE>
__device__ inline void block_pixel(uint32_t tid, uint32_t& lx, uint32_t& ly) {
    constexpr uint32_t kBx = TILE / 8;
    const uint32_t w = tid >> 6, l = tid & 63u;
    lx = (w % kBx) * 8u + (l & 7u);
    ly = (w / kBx) * 8u + (l >> 3);
}

template <uint32_t TILE>
__device__ inline Band wave_band(const DView& V, uint32_t tx, uint32_t ty, uint32_t wave) {
    constexpr uint32_t kBx = TILE / 8;
    const uint32_t c0 = tx * TILE + (wave % kBx) * 8u, r0 = ty * TILE + (wave / kBx) * 8u;
    // A block past the view's edge holds no pixel (its lanes never walk).
    const uint32_t c1 = min(c0 + 8u, V.width), r1 = min(r0 + 8u, V.height);
    Band b;
    b.x0 = (float)c0 + 0.5f;
    b.x1 = (float)(c1 > c0 ? c1 - 1u : c0) + 0.5f;
    b.y0 = (float)r0 + 0.5f;
    b.y1 = (float)(r1 > r0 ? r1 - 1u : r0) + 0.5f;
    return b;
}

__device__ inline bool band_meets(const Band& B, float cx, float cy, float a, float b, float c, float cut) {
    const float lox = B.x0 - cx, hix = B.x1 - cx, loy = B.y0 - cy, hiy = B.y1 - cy;
    if (lox <= 0.0f && hix >= 0.0f && loy <= 0.0f && hiy >= 0.0f) {
        return true;
    }
    if (!(a > 1e-12f) || !(c > 1e-12f)) {
        return true;
    }
    const float ia = __builtin_amdgcn_rcpf(a), ic = __builtin_amdgcn_rcpf(c);
    const auto at = [&](float x, float y) { return (a * x * x + 2.0f * b * x * y) + c * y * y; };
    const float ya = clamp_(-b * lox * ic, loy, hiy), yb = clamp_(-b * hix * ic, loy, hiy);
    const float xa = clamp_(-b * loy * ia, lox, hix), xb = clamp_(-b * hiy * ia, lox, hix);
    const float least = fmin_(fmin_(at(lox, ya), at(hix, yb)), fmin_(at(xa, loy), at(xb, hiy)));
    const float X = fmax_(fabsf(lox), fabsf(hix)), Y = fmax_(fabsf(loy), fabsf(hiy));
    const float M = (a * X * X + 2.0f * fabsf(b) * X * Y) + c * Y * Y;
    // Skipped only on a definite answer: a NaN anywhere keeps the Gaussian.
    return !(least > cut + kBandSlack * M);
}

// The staged Gaussians first .. first + 63 of the batch (count staged) that
// meet the band, as a mask of lanes: bit i for Gaussian first + i.
template <uint32_t TILE>
__device__ inline uint64_t band_mask(const LdsBatch& s, const Band& B, uint32_t first, uint32_t count, uint32_t lane) {
    if constexpr (TILE == 8) {
        // A wavefront is the whole tile, which every pair meets.
        const uint32_t n = count - first < 64u ? count - first : 64u;
        return n == 64u ? ~0ull : ((1ull << n) - 1ull);
    } else {
        const uint32_t j = first + lane;
        bool meets = false;
        if (j < count) {
            const float4 q = s.xy_cut_op[j];
            const float4 k = s.conic[j];
            meets = band_meets(B, q.x, q.y, k.x, s.colour[j].w, k.z, q.z);
        }
        return __ballot(meets);
    }
}

template <uint32_t TILE, bool kPiped>
__global__ void __launch_bounds__(TILE * TILE) blend_kernel(const DView* __restrict__ views,
                                                            const uint32_t* __restrict__ tile_bases, uint32_t nviews,
                                                            uint32_t batch, const uint32_t* __restrict__ ranges,
                                                            const uint32_t* __restrict__ pair_slots,
                                                            const float* __restrict__ records, uint32_t depth_mode,
                                                            float depth_threshold) {
    constexpr uint32_t kThreads = TILE * TILE;
    __shared__ LdsBatch s;
    // Tiles in dispatch order, not in bands an XCD: every XCD then walks the
    // same view at once, so the records gathered at any time are one view's
    // (a working set the 256 MB infinity cache holds, where eight views' do
    // not), and the views' unequal loads even out over the XCDs.
    const uint32_t gtile = blockIdx.x;
    const uint32_t v = view_of_tile(tile_bases, nviews, gtile);
    const DView& V = views[v];
    const uint32_t local = gtile - V.tile_base;
    const uint32_t tx = local % V.tiles_x, ty = local / V.tiles_x;
    const uint32_t tid = threadIdx.x;
    uint32_t lx, ly;
    block_pixel<TILE>(tid, lx, ly);
    const uint32_t px = tx * TILE + lx, py = ty * TILE + ly;
    const bool inside = px < V.width && py < V.height;
    const uint32_t start = ranges[2 * gtile], end = ranges[2 * gtile + 1];
    const float pxc = (float)px + 0.5f, pyc = (float)py + 0.5f;

    float r = 0.0f, g = 0.0f, b = 0.0f, T = 1.0f, dsum = 0.0f, dat = -1.0f;
    uint32_t last = 0;
    bool done = !inside;
    const uint32_t lane = tid & 63u;
    const Band band = wave_band<TILE>(V, tx, ty, tid >> 6);
    // With a batch of at most a record a thread (kPiped), the staging is
    // pipelined:
    // while a batch is walked, each thread already reads its record of the
    // next batch (its slot read during the batch before) and its slot of the
    // batch after, so a batch's staging waits on no memory.
    // The reads are unconditional (a thread with no pair reads the tile's
    // last one again, and its record is never staged), so no branch holds
    // them back.
    Staged ahead{};
    uint32_t slot_ahead = 0;
    if (kPiped && start < end) {
        ahead = fetch(records, pair_slots[min(start + tid, end - 1u)]);
        slot_ahead = pair_slots[min(start + batch + tid, end - 1u)];
    }
    for (uint32_t first = start; first < end; first += batch) {
        if (!__syncthreads_or(!done)) {
            break;
        }
        const uint32_t count = end - first < batch ? end - first : batch;
        if constexpr (kPiped) {
            if (tid < count) {
                put(s, tid, ahead);
            }
        } else {
            for (uint32_t k = tid; k < count; k += kThreads) {
                stage(s, k, records, pair_slots[first + k]);
            }
        }
        __syncthreads();
        if constexpr (kPiped) {
            ahead = fetch(records, slot_ahead);
            slot_ahead = pair_slots[min(first + 2u * batch + tid, end - 1u)];
        }
        for (uint32_t c0 = 0; c0 < count && __ballot(!done) != 0; c0 += 64) {
          for (uint64_t mm = band_mask<TILE>(s, band, c0, count, lane); mm != 0 && !done; mm &= mm - 1ull) {
            const uint32_t j = c0 + (uint32_t)__builtin_ctzll(mm);
            const float4 q = s.xy_cut_op[j];
            const float4 k = s.conic[j];
            const float dx = pxc - q.x, dy = pyc - q.y;
            const float power = k.x * dx * dx + k.y * dx * dy + k.z * dy * dy;
            if (power > q.z) {
                continue;
            }
            const float alpha = fminf(0.99f, q.w * blend_exp(power));
            if (alpha < 1.0f / 255.0f) {
                continue;
            }
            const float4 c = s.colour[j];
            const float w = T * alpha;
            r = r + w * c.x;
            g = g + w * c.y;
            b = b + w * c.z;
            dsum = dsum + w * k.w;
            T = T * (1.0f - alpha);
            if (dat < 0.0f && 1.0f - T >= depth_threshold) {
                dat = k.w;
            }
            last = first - start + j + 1;
            if (T < 1e-4f) {
                done = true;
            }
          }
        }
        // Pairs a pixel skipped after its last contribution are read again by
        // nobody: `last` is the walk's own extent.
        __syncthreads();
    }
    if (!inside) {
        return;
    }
    const size_t pix = (size_t)py * V.width + px;
    const float coverage = 1.0f - T;
    float4 out;
    out.x = r + T * V.bg[0];
    out.y = g + T * V.bg[1];
    out.z = b + T * V.bg[2];
    out.w = coverage + T * V.bg[3];
    reinterpret_cast<float4*>(V.colour)[pix] = out;
    if (V.depth) {
        V.depth[pix] = depth_mode == 0 ? (coverage > 1e-6f ? dsum / coverage : 0.0f) : fmaxf(dat, 0.0f);
    }
    if (V.final_t) {
        V.final_t[pix] = T;
    }
    if (V.last) {
        V.last[pix] = last;
    }
}

// The walk's variants (splat_blend.slang:236-578 with composite, matte and
// traced; athx_targets): the same walk as blend_kernel, and
//   under    an opaque layer under the splats (the view's under_colour and
//            under_depth): a splat at or behind its depth ends the pixel's
//            walk, and the layer takes the light left, as the background
//            would, its depth weighing in the mean (:287-300, 330-333,
//            512-566);
//   matte    the six heaviest Cryptomatte ids of the pixel with their
//            coverage, the layer's own id taking what the walk left
//            (:88-143, 421-423, 526-531);
//   traced   the four heaviest slots, their depths and weights, and what the
//            list left out and what closed the walk (splat_trace.slang).
// Both lists live in registers, every access to them unrolled, so they cost
// no memory; the flags are the frame's, uniform across the launch.
constexpr uint32_t kCryptoSlots = 6;

__device__ inline void crypto_add(uint32_t (&ids)[kCryptoSlots], float (&wts)[kCryptoSlots], uint32_t id,
                                  float weight) {
    if (weight <= 0.0f || id == 0u) {
        return;
    }
    uint32_t lightest = 0;
    float least = wts[0];
    bool merged = false;
#pragma unroll
    for (uint32_t k = 0; k < kCryptoSlots; ++k) {
        // The first rank naming this id takes the weight (splat_blend.slang:101-110).
        const bool same = !merged && ids[k] == id && wts[k] > 0.0f;
        if (same) {
            wts[k] = wts[k] + weight;
            merged = true;
        }
        if (!merged && wts[k] < least) {
            least = wts[k];
            lightest = k;
        }
    }
    if (!merged && weight > least) {
#pragma unroll
        for (uint32_t k = 0; k < kCryptoSlots; ++k) {
            if (k == lightest) {
                ids[k] = id;
                wts[k] = weight;
            }
        }
    }
}

// Ranks by descending weight, ties in list order (an insertion sort's).
template <uint32_t N>
__device__ inline void order_desc(const float (&w)[N], uint32_t (&order)[N]) {
#pragma unroll
    for (uint32_t k = 0; k < N; ++k) {
        uint32_t rank = 0;
#pragma unroll
        for (uint32_t m = 0; m < N; ++m) {
            rank += (w[m] > w[k] || (w[m] == w[k] && m < k)) ? 1u : 0u;
        }
#pragma unroll
        for (uint32_t r = 0; r < N; ++r) {
            if (r == rank) {
                order[r] = k;
            }
        }
    }
}

template <uint32_t N, typename T>
__device__ inline T pick(const T (&a)[N], uint32_t k) {
    T v = a[0];
#pragma unroll
    for (uint32_t m = 1; m < N; ++m) {
        v = m == k ? a[m] : v;
    }
    return v;
}

__device__ inline uint32_t half_of(float v) { return (uint32_t)__builtin_bit_cast(uint16_t, (_Float16)v); }

template <uint32_t TILE>
__global__ void __launch_bounds__(TILE * TILE) blend_ex_kernel(const DView* __restrict__ views,
                                                               const uint32_t* __restrict__ tile_bases,
                                                               uint32_t nviews, uint32_t batch,
                                                               const uint32_t* __restrict__ ranges,
                                                               const uint32_t* __restrict__ pair_slots,
                                                               const float* __restrict__ records,
                                                               const uint32_t* __restrict__ slot_ids,
                                                               uint32_t depth_mode, float depth_threshold) {
    constexpr uint32_t kThreads = TILE * TILE;
    __shared__ LdsBatch s;
    __shared__ uint32_t s_slot[kMaxBatch];
    __shared__ uint32_t s_id[kMaxBatch];
    const uint32_t gtile = blockIdx.x;
    const uint32_t v = view_of_tile(tile_bases, nviews, gtile);
    const DView& V = views[v];
    const uint32_t local = gtile - V.tile_base;
    const uint32_t tx = local % V.tiles_x, ty = local / V.tiles_x;
    const uint32_t tid = threadIdx.x;
    uint32_t lx, ly;
    block_pixel<TILE>(tid, lx, ly);
    const uint32_t px = tx * TILE + lx, py = ty * TILE + ly;
    const bool inside = px < V.width && py < V.height;
    const uint32_t start = ranges[2 * gtile], end = ranges[2 * gtile + 1];
    const float pxc = (float)px + 0.5f, pyc = (float)py + 0.5f;
    const uint32_t pixels = V.width * V.height;
    const size_t pix = inside ? (size_t)py * V.width + px : 0;
    const bool matte = V.crypto != nullptr, traced = V.trace != nullptr;
    const bool under = inside && V.under_colour != nullptr && V.under_depth != nullptr;
    const float under_z = under ? V.under_depth[pix] : 0.0f;
    const bool under_hit = under && under_z > 0.0f;

    float r = 0.0f, g = 0.0f, b = 0.0f, T = 1.0f, dsum = 0.0f, dat = -1.0f;
    uint32_t last = 0;
    uint32_t ids[kCryptoSlots];
    float cw[kCryptoSlots];
#pragma unroll
    for (uint32_t k = 0; k < kCryptoSlots; ++k) {
        ids[k] = 0u;
        cw[k] = 0.0f;
    }
    uint32_t tslot[ATHX_TRACE_SLOTS];
    float tw[ATHX_TRACE_SLOTS];
#pragma unroll
    for (uint32_t k = 0; k < ATHX_TRACE_SLOTS; ++k) {
        tslot[k] = ATHX_TRACE_NONE;
        tw[k] = 0.0f;
    }
    float ttotal = 0.0f, tfloor = 0.0f;
    uint32_t tlightest = 0;
    bool done = !inside;
    const uint32_t lane = tid & 63u;
    const Band band = wave_band<TILE>(V, tx, ty, tid >> 6);
    for (uint32_t first = start; first < end; first += batch) {
        if (!__syncthreads_or(!done)) {
            break;
        }
        const uint32_t count = end - first < batch ? end - first : batch;
        for (uint32_t k = tid; k < count; k += kThreads) {
            const uint32_t slot = pair_slots[first + k];
            stage(s, k, records, slot);
            s_slot[k] = slot;
            if (matte) {
                s_id[k] = slot_ids[slot];
            }
        }
        __syncthreads();
        // A pair the band skips holds nothing for these pixels; the under
        // layer's stop comes at the next pair walked, or none is left to stop:
        // pairs run front to back, so any later one lies behind the layer too.
        for (uint32_t c0 = 0; c0 < count && __ballot(!done) != 0; c0 += 64) {
          for (uint64_t mm = band_mask<TILE>(s, band, c0, count, lane); mm != 0 && !done; mm &= mm - 1ull) {
            const uint32_t j = c0 + (uint32_t)__builtin_ctzll(mm);
            const float4 q = s.xy_cut_op[j];
            const float4 k = s.conic[j];
            if (under_hit && k.w >= under_z) {
                done = true;
                break;
            }
            const float dx = pxc - q.x, dy = pyc - q.y;
            const float power = k.x * dx * dx + k.y * dx * dy + k.z * dy * dy;
            if (power > q.z) {
                continue;
            }
            const float alpha = fminf(0.99f, q.w * blend_exp(power));
            if (alpha < 1.0f / 255.0f) {
                continue;
            }
            const float4 c = s.colour[j];
            const float w = T * alpha;
            if (matte) {
                crypto_add(ids, cw, s_id[j], w);
            }
            if (traced) {
                ttotal = ttotal + w;
                if (w > tfloor) {
                    // TraceList::enter (splat_trace.slang:81-90)
                    const uint32_t slot = s_slot[j];
#pragma unroll
                    for (uint32_t m = 0; m < ATHX_TRACE_SLOTS; ++m) {
                        if (m == tlightest) {
                            tslot[m] = slot;
                            tw[m] = w;
                        }
                    }
                    tlightest = 0;
                    float least = tw[0];
#pragma unroll
                    for (uint32_t m = 1; m < ATHX_TRACE_SLOTS; ++m) {
                        if (tw[m] < least) {
                            least = tw[m];
                            tlightest = m;
                        }
                    }
                    tfloor = least;
                }
            }
            r = r + w * c.x;
            g = g + w * c.y;
            b = b + w * c.z;
            dsum = dsum + w * k.w;
            T = T * (1.0f - alpha);
            if (dat < 0.0f && 1.0f - T >= depth_threshold) {
                dat = k.w;
            }
            last = first - start + j + 1;
            if (T < 1e-4f) {
                done = true;
            }
          }
        }
        __syncthreads();
    }
    if (!inside) {
        return;
    }
    float4 behind = make_float4(V.bg[0], V.bg[1], V.bg[2], V.bg[3]);
    float opaque_z = 0.0f;
    uint32_t opaque_id = 0u;
    if (under_hit) {
        behind = reinterpret_cast<const float4*>(V.under_colour)[pix];
        opaque_z = under_z;
        opaque_id = matte && V.under_crypto ? V.under_crypto[pix] : 0u;
    }
    if (matte) {
        crypto_add(ids, cw, opaque_id, under_hit ? T : 0.0f);
        uint32_t order[kCryptoSlots];
        order_desc<kCryptoSlots>(cw, order);
#pragma unroll
        for (uint32_t layer = 0; layer < kCryptoSlots / 2; ++layer) {
            const uint32_t ka = order[layer * 2], kb = order[layer * 2 + 1];
            const float wa = pick(cw, ka), wb = pick(cw, kb);
            const uint32_t ia = pick(ids, ka), ib = pick(ids, kb);
            reinterpret_cast<float4*>(V.crypto)[(size_t)layer * pixels + pix] =
                make_float4(wa > 0.0f ? __builtin_bit_cast(float, ia) : 0.0f, fmaxf(wa, 0.0f),
                            wb > 0.0f ? __builtin_bit_cast(float, ib) : 0.0f, fmaxf(wb, 0.0f));
        }
    }
    if (traced) {
        const float under_w = under_hit ? T : 0.0f;
        if (!(ttotal > 0.0f) && !(under_w > 0.0f)) {
            V.trace[pix] = ATHX_TRACE_EMPTY;
        } else {
            uint32_t order[ATHX_TRACE_SLOTS];
            order_desc<ATHX_TRACE_SLOTS>(tw, order);
            float held = 0.0f;
            float sorted_w[ATHX_TRACE_SLOTS];
#pragma unroll
            for (uint32_t k = 0; k < ATHX_TRACE_SLOTS; ++k) {
                const float wk = pick(tw, order[k]);
                const uint32_t sk = pick(tslot, order[k]);
                const bool some = wk > 0.0f;
                sorted_w[k] = wk;
                V.trace[(size_t)k * pixels + pix] = some ? sk : ATHX_TRACE_NONE;
                V.trace[(size_t)(ATHX_TRACE_SLOTS + k) * pixels + pix] =
                    some ? __builtin_bit_cast(uint32_t, records[(size_t)sk * kRecordFloats + 2]) : 0u;
                held = held + wk;
            }
#pragma unroll
            for (uint32_t pr = 0; pr < ATHX_TRACE_SLOTS / 2; ++pr) {
                V.trace[(size_t)(ATHX_TRACE_SLOTS * 2 + pr) * pixels + pix] =
                    half_of(sorted_w[pr * 2]) | (half_of(sorted_w[pr * 2 + 1]) << 16);
            }
            V.trace[(size_t)(ATHX_TRACE_WORDS - 1) * pixels + pix] =
                half_of(fmaxf(ttotal - held, 0.0f)) | (half_of(fmaxf(under_w, 0.0f)) << 16);
        }
    }
    const float coverage = 1.0f - T;
    if (under_hit) {
        dsum = dsum + T * opaque_z;
        if (dat < 0.0f) {
            dat = opaque_z;
        }
    }
    const float total = under_hit ? 1.0f : coverage;
    float4 out;
    out.x = r + T * behind.x;
    out.y = g + T * behind.y;
    out.z = b + T * behind.z;
    out.w = coverage + T * behind.w;
    reinterpret_cast<float4*>(V.colour)[pix] = out;
    if (V.depth) {
        V.depth[pix] = depth_mode == 0 ? (total > 1e-6f ? dsum / total : 0.0f) : fmaxf(dat, 0.0f);
    }
    if (V.final_t) {
        V.final_t[pix] = T;
    }
    if (V.last) {
        V.last[pix] = last;
    }
}

// What each splat gave the last frame (splat_blend.slang:36-48, 364-398):
// the same walk as blend_kernel -- the same pairs, records, arithmetic and
// early stop -- counting at each pair's slot, with integer atomics (so in
// any order the same sum): mode 1 the blend weight T alpha in 1/4096 of a
// pixel, rounded; mode 4 NeuralFit's colour gradient, the weight times
// dL/dC (the view's dl_dcolour) in 1/256, each channel rounded down or up by
// color.slang's hash01 of (pixel, channel, slot) so that steps below one are
// kept on average, and the weight itself in 1/256 as a fourth word.
__device__ inline float hash01(uint32_t v) {
    const uint32_t state = v * 747796405u + 2891336453u;
    const uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (float)((word >> 22u) ^ word) / 4294967295.0f;
}

template <uint32_t TILE, uint32_t MODE>
__global__ void __launch_bounds__(TILE * TILE) contribution_kernel(const DView* __restrict__ views,
                                                                   const uint32_t* __restrict__ tile_bases,
                                                                   uint32_t nviews, uint32_t batch,
                                                                   const uint32_t* __restrict__ ranges,
                                                                   const uint32_t* __restrict__ pair_slots,
                                                                   const float* __restrict__ records,
                                                                   uint32_t* __restrict__ contribution) {
    constexpr uint32_t kThreads = TILE * TILE;
    __shared__ LdsBatch s;
    __shared__ uint32_t s_slot[kMaxBatch];
    const uint32_t gtile = blockIdx.x;
    const uint32_t v = view_of_tile(tile_bases, nviews, gtile);
    const DView& V = views[v];
    const uint32_t local = gtile - V.tile_base;
    const uint32_t tx = local % V.tiles_x, ty = local / V.tiles_x;
    const uint32_t tid = threadIdx.x;
    uint32_t lx, ly;
    block_pixel<TILE>(tid, lx, ly);
    const uint32_t px = tx * TILE + lx, py = ty * TILE + ly;
    const bool inside = px < V.width && py < V.height;
    const uint32_t start = ranges[2 * gtile], end = ranges[2 * gtile + 1];
    const float pxc = (float)px + 0.5f, pyc = (float)py + 0.5f;
    const uint32_t at = inside ? py * V.width + px : 0u;
    // The per-pixel input: mode 2 the pixel drawn whole, mode 3 the mask,
    // mode 4 dL/dC.
    float4 dl = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if ((MODE == 2 || MODE == 4) && inside) {
        dl = reinterpret_cast<const float4*>(V.dl_dcolour)[at];
    } else if (MODE == 3 && inside) {
        dl.x = fminf(fmaxf(V.dl_dcolour[at], 0.0f), 1.0f);
    }
    float T = 1.0f;
    float r = 0.0f, g = 0.0f, b = 0.0f;
    bool done = !inside;
    const uint32_t lane = tid & 63u;
    const Band band = wave_band<TILE>(V, tx, ty, tid >> 6);
    for (uint32_t first = start; first < end; first += batch) {
        if (!__syncthreads_or(!done)) {
            break;
        }
        const uint32_t count = end - first < batch ? end - first : batch;
        for (uint32_t k = tid; k < count; k += kThreads) {
            const uint32_t slot = pair_slots[first + k];
            stage(s, k, records, slot);
            s_slot[k] = slot;
        }
        __syncthreads();
        for (uint32_t c0 = 0; c0 < count && __ballot(!done) != 0; c0 += 64) {
          for (uint64_t mm = band_mask<TILE>(s, band, c0, count, lane); mm != 0 && !done; mm &= mm - 1ull) {
            const uint32_t j = c0 + (uint32_t)__builtin_ctzll(mm);
            const float4 q = s.xy_cut_op[j];
            const float4 k = s.conic[j];
            const float dx = pxc - q.x, dy = pyc - q.y;
            const float power = k.x * dx * dx + k.y * dx * dy + k.z * dy * dy;
            if (power > q.z) {
                continue;
            }
            const float alpha = fminf(0.99f, q.w * blend_exp(power));
            if (alpha < 1.0f / 255.0f) {
                continue;
            }
            const float w = T * alpha;
            const uint32_t slot = s_slot[j];
            if (MODE == 1) {
                atomicAdd(&contribution[slot], (uint32_t)(w * 4096.0f + 0.5f));
            } else if (MODE == 2) {
                // The pixel without this splat: what is behind it is
                // (C - P) / T' with P the walk once it is in, T' the
                // transmittance after it; the change is w (c - behind).
                const float4 c = s.colour[j];
                const float after = T * (1.0f - alpha);
                const float keep = fmaxf(after, 1.0e-4f);
                const float er = w * (c.x - (dl.x - (r + w * c.x)) / keep);
                const float eg = w * (c.y - (dl.y - (g + w * c.y)) / keep);
                const float eb = w * (c.z - (dl.z - (b + w * c.z)) / keep);
                const float e2 = fminf(er * er + eg * eg + eb * eb, 1.0f);
                const uint32_t add = (uint32_t)(e2 * 16777216.0f + 0.5f);
                const uint32_t before = atomicAdd(&contribution[(size_t)slot * 2], add);
                if (before + add < before) {
                    atomicAdd(&contribution[(size_t)slot * 2 + 1], 1u);
                }
                r = r + w * c.x;
                g = g + w * c.y;
                b = b + w * c.z;
            } else if (MODE == 3) {
                const uint32_t seen = (uint32_t)(w * 1024.0f + 0.5f);
                if (seen != 0u) {
                    atomicAdd(&contribution[(size_t)slot * 2 + 1], seen);
                    const uint32_t within = (uint32_t)(w * dl.x * 1024.0f + 0.5f);
                    if (within != 0u) {
                        atomicAdd(&contribution[(size_t)slot * 2], within);
                    }
                }
            } else {
                const float g[3] = {w * dl.x * 256.0f, w * dl.y * 256.0f, w * dl.z * 256.0f};
                for (uint32_t c = 0; c < 3; ++c) {
                    const int step = (int)floorf(g[c] + hash01((at * 3u + c) ^ (slot * 2654435761u)));
                    if (step != 0) {
                        atomicAdd(&contribution[(size_t)slot * 4 + c], (uint32_t)step);
                    }
                }
                atomicAdd(&contribution[(size_t)slot * 4 + 3], (uint32_t)(w * 256.0f + 0.5f));
            }
            T = T * (1.0f - alpha);
            if (T < 1e-4f) {
                done = true;
            }
          }
        }
        __syncthreads();
    }
}

// The sum over a wavefront's 64 lanes, every lane active: DPP moves within
// each row of 16 lanes (quad swaps, then rotations by 4 and 8) leave each row's
// sum in all of its lanes, and four lane reads add the rows. All VALU, no LDS.
__device__ inline float dpp_add(float x, int ctrl_sel) {
    int moved = 0;
    switch (ctrl_sel) {
    case 0: moved = __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, x), 0xB1, 0xF, 0xF, false); break;   // quad [1,0,3,2]
    case 1: moved = __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, x), 0x4E, 0xF, 0xF, false); break;   // quad [2,3,0,1]
    case 2: moved = __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, x), 0x124, 0xF, 0xF, false); break;  // row_ror 4
    default: moved = __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, x), 0x128, 0xF, 0xF, false); break; // row_ror 8
    }
    return x + __builtin_bit_cast(float, moved);
}

__device__ inline float wave_sum_f(float x) {
    x = dpp_add(x, 0);
    x = dpp_add(x, 1);
    x = dpp_add(x, 2);
    x = dpp_add(x, 3);
    const int bits = __builtin_bit_cast(int, x);
    const float r0 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(bits, 0));
    const float r1 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(bits, 16));
    const float r2 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(bits, 32));
    const float r3 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(bits, 48));
    return (r0 + r1) + (r2 + r3);
}

__device__ inline float as_float(uint32_t u) { return __builtin_bit_cast(float, u); }
__device__ inline uint32_t as_uint(float f) { return __builtin_bit_cast(uint32_t, f); }

// Eight sums over a wavefront's 64 lanes at once, every lane active (a
// reduce-scatter): a swap of the wave's halves (v_permlane32_swap) pairs
// value i with value i + 4 so each lane adds one of them over lanes l and
// l + 32; a swap of rows within each half (v_permlane16_swap) does the same
// for the four values left, over l and l + 16; a rotation by 8 within each
// row of 16, and three DPP steps over each group of 8 lanes, finish. Sum k
// lands in lane 8k. 26 VALU operations, where eight whole-wave sums take 88.
__device__ inline float wave_sum8(const float v[8], uint32_t lane) {
    float r1[4];
#pragma unroll
    for (uint32_t i = 0; i < 4; ++i) {
        const auto p = __builtin_amdgcn_permlane32_swap(as_uint(v[i]), as_uint(v[i + 4]), false, false);
        r1[i] = as_float(p[0]) + as_float(p[1]);  // lanes 0-31: value i; 32-63: value i + 4
    }
    float r2[2];
#pragma unroll
    for (uint32_t i = 0; i < 2; ++i) {
        const auto p = __builtin_amdgcn_permlane16_swap(as_uint(r1[i]), as_uint(r1[i + 2]), false, false);
        r2[i] = as_float(p[0]) + as_float(p[1]);  // row R of 16 lanes: value i + 2R
    }
    const bool upper = (lane & 8u) != 0;
    const float mine = upper ? r2[1] : r2[0];
    const float other = upper ? r2[0] : r2[1];
    float x = mine + as_float((uint32_t)__builtin_amdgcn_mov_dpp((int)as_uint(other), 0x128, 0xF, 0xF, false));
    x = x + as_float((uint32_t)__builtin_amdgcn_mov_dpp((int)as_uint(x), 0x141, 0xF, 0xF, false));  // half mirror
    x = dpp_add(x, 1);
    x = dpp_add(x, 0);
    return x;
}

template <uint32_t TILE, bool kPiped>
__global__ void __launch_bounds__(TILE * TILE) blend_backward_kernel(
    const DView* __restrict__ views, const uint32_t* __restrict__ tile_bases, uint32_t nviews, uint32_t batch,
    const uint32_t* __restrict__ ranges, const uint32_t* __restrict__ pair_slots, const float* __restrict__ records,
    float* __restrict__ grad2d) {
    constexpr uint32_t kThreads = TILE * TILE;
    __shared__ LdsBatch s;
    __shared__ uint32_t s_slot[kMaxBatch];
    // The batch's gradients summed over the workgroup's wavefronts, then added
    // to the slots' words once a batch.
    __shared__ float s_grad[kMaxBatch * kGrad2d];
    __shared__ uint32_t s_max_last;
    __shared__ uint32_t s_wave_last[kThreads / 64];
    const uint32_t gtile = xcd_band(blockIdx.x, gridDim.x);
    const uint32_t v = view_of_tile(tile_bases, nviews, gtile);
    const DView& V = views[v];
    const uint32_t local = gtile - V.tile_base;
    const uint32_t tx = local % V.tiles_x, ty = local / V.tiles_x;
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & 63u;
    uint32_t lx, ly;
    block_pixel<TILE>(tid, lx, ly);
    const uint32_t px = tx * TILE + lx, py = ty * TILE + ly;
    const bool inside = px < V.width && py < V.height;
    const uint32_t start = ranges[2 * gtile];
    const uint32_t tile_pairs = ranges[2 * gtile + 1] - start;
    const float pxc = (float)px + 0.5f, pyc = (float)py + 0.5f;
    const size_t pix = inside ? (size_t)py * V.width + px : 0;

    float T = 1.0f, TN = 1.0f;
    uint32_t last = 0;
    float gr = 0.0f, gg = 0.0f, gb = 0.0f, ga = 0.0f;
    if (inside) {
        T = V.final_t[pix];
        TN = T;
        // A walk never reads past its tile's pairs; the clamp keeps a stale
        // `last` from reading past them either.
        last = min(V.last[pix], tile_pairs);
        const float4 gc = reinterpret_cast<const float4*>(V.dl_dcolour)[pix];
        gr = gc.x; gg = gc.y; gb = gc.z; ga = gc.w;
    }
    // The colour of everything behind the current Gaussian, as seen from just
    // in front of it: the background, then each Gaussian over it.
    float br = V.bg[0], bgc = V.bg[1], bb = V.bg[2];
    // d(alpha out)/d(alpha_k) = T_N (1 - bg.a) / (1 - alpha_k).
    const float ga_scale = ga * TN * (1.0f - V.bg[3]);
    if (tid == 0) {
        s_max_last = 0;
    }
    if (lane == 0) {
        s_wave_last[tid >> 6] = 0;
    }
    __syncthreads();
    if (last != 0) {
        atomicMax(&s_max_last, last);
        atomicMax(&s_wave_last[tid >> 6], last);
    }
    __syncthreads();
    const uint32_t max_last = s_max_last;
    // A wavefront walks only the pairs below its own pixels' furthest.
    const uint32_t wave_last = s_wave_last[tid >> 6];
    const Band band = wave_band<TILE>(V, tx, ty, tid >> 6);
    // kPiped: the staging runs a batch ahead of the walk, as in the forward
    // blend -- while a batch is walked each thread reads its record of the
    // next (lower) batch and its slot of the one after. The reads are
    // unconditional, at in-bounds pairs of the tile (the tile has max_last
    // pairs or more).
    const auto pair_of = [&](uint32_t hi_b) {
        const uint32_t lo_b = hi_b > batch ? hi_b - batch : 0u;
        const uint32_t k = lo_b + tid;
        return start + (k < hi_b ? k : (hi_b != 0u ? hi_b - 1u : 0u));
    };
    uint32_t slot_now = 0, slot_ahead = 0;
    Staged rec_now{};
    if (kPiped && max_last > 0) {
        slot_now = pair_slots[pair_of(max_last)];
        rec_now = fetch(records, slot_now);
        const uint32_t hi1 = max_last > batch ? max_last - batch : 0u;
        slot_ahead = pair_slots[pair_of(hi1)];
    }
    for (uint32_t hi = max_last; hi > 0;) {
        const uint32_t lo = hi > batch ? hi - batch : 0;
        const uint32_t count = hi - lo;
        if constexpr (kPiped) {
            if (tid < count) {
                s_slot[tid] = slot_now;
                put(s, tid, rec_now);
            }
        } else {
            for (uint32_t k = tid; k < count; k += kThreads) {
                const uint32_t slot = pair_slots[start + lo + k];
                s_slot[k] = slot;
                stage(s, k, records, slot);
            }
        }
        for (uint32_t k = tid; k < count * kGrad2d; k += kThreads) {
            s_grad[k] = 0.0f;
        }
        __syncthreads();
        if constexpr (kPiped) {
            // The next batch is [lo - batch, lo); the one after it ends where
            // that one starts.
            slot_now = slot_ahead;
            rec_now = fetch(records, slot_ahead);
            const uint32_t lo1 = lo > batch ? lo - batch : 0u;
            slot_ahead = pair_slots[pair_of(lo1)];
        }
        const uint32_t top = wave_last > lo ? (wave_last - lo < count ? wave_last - lo : count) : 0u;
        // Back to front over the pairs below `top` that meet the band, 64 at a
        // time; the rest hold no pixel of it and add nothing.
        for (uint32_t c1 = top; c1 > 0; c1 = c1 > 64u ? c1 - 64u : 0u) {
          const uint32_t c0 = c1 > 64u ? c1 - 64u : 0u;
          for (uint64_t mm = band_mask<TILE>(s, band, c0, c1, lane); mm != 0;) {
            const uint32_t bit = 63u - (uint32_t)__builtin_clzll(mm);
            mm &= ~(1ull << bit);
            const uint32_t j = c0 + bit;
            const uint32_t idx = lo + j;
            bool contrib = false;
            float g_cx = 0.f, g_cy = 0.f, g_ka = 0.f, g_kb = 0.f, g_kc = 0.f, g_al = 0.f, g_r = 0.f, g_g = 0.f,
                  g_b = 0.f;
            if (idx < last) {
                const float4 q = s.xy_cut_op[j];
                const float4 k = s.conic[j];
                const float dx = pxc - q.x, dy = pyc - q.y;
                const float power = k.x * dx * dx + k.y * dx * dy + k.z * dy * dy;
                if (power <= q.z) {
                    const float e = blend_exp(power);
                    const float raw = q.w * e;
                    const float alpha = fminf(0.99f, raw);
                    if (alpha >= 1.0f / 255.0f) {
                        contrib = true;
                        const float4 c = s.colour[j];
                        const float one_minus = 1.0f - alpha;
                        // 1 / (1 - alpha), alpha <= 0.99: the hardware
                        // reciprocal and one Newton step (within an ulp or so
                        // of the quotient), where an IEEE divide takes a dozen
                        // operations twice a pair. The adjoint answers to
                        // finite differences at 1e-3, not to bits.
                        float inv = __builtin_amdgcn_rcpf(one_minus);
                        inv = inv + inv * (1.0f - one_minus * inv);
                        const float Tb = T * inv;
                        const float w = alpha * Tb;
                        g_r = w * gr;
                        g_g = w * gg;
                        g_b = w * gb;
                        const float dl_dalpha =
                            Tb * ((c.x - br) * gr + (c.y - bgc) * gg + (c.z - bb) * gb) + ga_scale * inv;
                        br = c.x * alpha + br * one_minus;
                        bgc = c.y * alpha + bgc * one_minus;
                        bb = c.z * alpha + bb * one_minus;
                        T = Tb;
                        if (!(raw > 0.99f)) {
                            g_al = dl_dalpha * e;
                            const float dl_dpower = dl_dalpha * (-0.5f * alpha);
                            const float kb = c.w;
                            g_ka = dl_dpower * dx * dx;
                            g_kb = dl_dpower * 2.0f * dx * dy;
                            g_kc = dl_dpower * dy * dy;
                            // power depends on the centre through dx = px - cx.
                            g_cx = -dl_dpower * (2.0f * k.x * dx + 2.0f * kb * dy);
                            g_cy = -dl_dpower * (2.0f * kb * dx + 2.0f * k.z * dy);
                        }
                    }
                }
            }
            if (__any(contrib)) {
                const float eight[8] = {g_cx, g_cy, g_ka, g_kb, g_kc, g_al, g_r, g_g};
                const float sum8 = wave_sum8(eight, lane);
                const float sum_b = wave_sum_f(g_b);
                if ((lane & 7u) == 0) {
                    atomicAdd(&s_grad[j * kGrad2d + (lane >> 3)], sum8);
                } else if (lane == 1) {
                    atomicAdd(&s_grad[j * kGrad2d + 8], sum_b);
                }
            }
          }
        }
        __syncthreads();
        for (uint32_t k = tid; k < count * kGrad2d; k += kThreads) {
            const float g = s_grad[k];
            if (g != 0.0f) {
                atomicAdd(&grad2d[(size_t)s_slot[k / kGrad2d] * kGrad2d + k % kGrad2d], g);
            }
        }
        __syncthreads();
        hi = lo;
    }
}

template <uint32_t TILE>
hipError_t blend_for(const DView* views, const uint32_t* tile_bases, uint32_t nviews, uint32_t tiles, uint32_t batch,
                     const uint32_t* ranges, const uint32_t* slots, const float* records, uint32_t depth_mode,
                     float depth_threshold, hipStream_t s) {
    if (batch <= TILE * TILE) {
        hipLaunchKernelGGL((blend_kernel<TILE, true>), dim3(tiles), dim3(TILE * TILE), 0, s, views, tile_bases, nviews,
                           batch, ranges, slots, records, depth_mode, depth_threshold);
    } else {
        hipLaunchKernelGGL((blend_kernel<TILE, false>), dim3(tiles), dim3(TILE * TILE), 0, s, views, tile_bases,
                           nviews, batch, ranges, slots, records, depth_mode, depth_threshold);
    }
    return hipGetLastError();
}

template <uint32_t TILE>
hipError_t blend_backward_for(const DView* views, const uint32_t* tile_bases, uint32_t nviews, uint32_t tiles,
                              uint32_t batch, const uint32_t* ranges, const uint32_t* slots, const float* records,
                              float* grad2d, hipStream_t s) {
    if (batch <= TILE * TILE) {
        hipLaunchKernelGGL((blend_backward_kernel<TILE, true>), dim3(tiles), dim3(TILE * TILE), 0, s, views,
                           tile_bases, nviews, batch, ranges, slots, records, grad2d);
    } else {
        hipLaunchKernelGGL((blend_backward_kernel<TILE, false>), dim3(tiles), dim3(TILE * TILE), 0, s, views,
                           tile_bases, nviews, batch, ranges, slots, records, grad2d);
    }
    return hipGetLastError();
}

}  // namespace

hipError_t launch_blend(const DView* views_dev, const uint32_t* tile_bases, uint32_t nviews, uint32_t total_tiles,
                        uint32_t tile, uint32_t batch, const uint32_t* ranges, const uint32_t* pair_slots,
                        const float* records, uint32_t depth_mode, float depth_threshold, hipStream_t s,
                        bool variants, const uint32_t* slot_ids) {
    if (total_tiles == 0) {
        return hipSuccess;
    }
    if (batch == 0 || batch > kMaxBatch) {
        return hipErrorInvalidValue;
    }
    if (variants) {
        using Kernel = void (*)(const DView*, const uint32_t*, uint32_t, uint32_t, const uint32_t*, const uint32_t*,
                                const float*, const uint32_t*, uint32_t, float);
        const Kernel k = tile == 8 ? blend_ex_kernel<8> : tile == 16 ? blend_ex_kernel<16> : tile == 32 ? blend_ex_kernel<32>
                                                                                                          : nullptr;
        if (!k) {
            return hipErrorInvalidValue;
        }
        hipLaunchKernelGGL(k, dim3(total_tiles), dim3(tile * tile), 0, s, views_dev, tile_bases, nviews, batch, ranges,
                           pair_slots, records, slot_ids, depth_mode, depth_threshold);
        return hipGetLastError();
    }
    switch (tile) {
    case 8:
        return blend_for<8>(views_dev, tile_bases, nviews, total_tiles, batch, ranges, pair_slots, records, depth_mode,
                            depth_threshold, s);
    case 16:
        return blend_for<16>(views_dev, tile_bases, nviews, total_tiles, batch, ranges, pair_slots, records,
                             depth_mode, depth_threshold, s);
    case 32:
        return blend_for<32>(views_dev, tile_bases, nviews, total_tiles, batch, ranges, pair_slots, records,
                             depth_mode, depth_threshold, s);
    default:
        return hipErrorInvalidValue;
    }
}

template <uint32_t TILE>
static hipError_t contribution_for(const DView* views, const uint32_t* tile_bases, uint32_t nviews, uint32_t tiles,
                            uint32_t batch, const uint32_t* ranges, const uint32_t* slots, const float* records,
                            uint32_t mode, uint32_t* out, hipStream_t s) {
    if (mode == 1) {
        hipLaunchKernelGGL((contribution_kernel<TILE, 1>), dim3(tiles), dim3(TILE * TILE), 0, s, views, tile_bases,
                           nviews, batch, ranges, slots, records, out);
    } else if (mode == 2) {
        hipLaunchKernelGGL((contribution_kernel<TILE, 2>), dim3(tiles), dim3(TILE * TILE), 0, s, views, tile_bases,
                           nviews, batch, ranges, slots, records, out);
    } else if (mode == 3) {
        hipLaunchKernelGGL((contribution_kernel<TILE, 3>), dim3(tiles), dim3(TILE * TILE), 0, s, views, tile_bases,
                           nviews, batch, ranges, slots, records, out);
    } else {
        hipLaunchKernelGGL((contribution_kernel<TILE, 4>), dim3(tiles), dim3(TILE * TILE), 0, s, views, tile_bases,
                           nviews, batch, ranges, slots, records, out);
    }
    return hipGetLastError();
}

hipError_t launch_contribution(const DView* views_dev, const uint32_t* tile_bases, uint32_t nviews,
                               uint32_t total_tiles, uint32_t tile, uint32_t batch, const uint32_t* ranges,
                               const uint32_t* pair_slots, const float* records, uint32_t mode, uint32_t* out,
                               hipStream_t s) {
    if (total_tiles == 0) {
        return hipSuccess;
    }
    if (batch == 0 || batch > kMaxBatch || mode < 1 || mode > 4) {
        return hipErrorInvalidValue;
    }
    switch (tile) {
    case 8:
        return contribution_for<8>(views_dev, tile_bases, nviews, total_tiles, batch, ranges, pair_slots, records,
                                   mode, out, s);
    case 16:
        return contribution_for<16>(views_dev, tile_bases, nviews, total_tiles, batch, ranges, pair_slots, records,
                                    mode, out, s);
    case 32:
        return contribution_for<32>(views_dev, tile_bases, nviews, total_tiles, batch, ranges, pair_slots, records,
                                    mode, out, s);
    default:
        return hipErrorInvalidValue;
    }
}

hipError_t launch_blend_backward(const DView* views_dev, const uint32_t* tile_bases, uint32_t nviews,
                                 uint32_t total_tiles, uint32_t tile, uint32_t batch, const uint32_t* ranges,
                                 const uint32_t* pair_slots, const float* records, float* grad2d, hipStream_t s) {
    if (total_tiles == 0) {
        return hipSuccess;
    }
    if (batch == 0 || batch > kMaxBatch) {
        return hipErrorInvalidValue;
    }
    switch (tile) {
    case 8:
        return blend_backward_for<8>(views_dev, tile_bases, nviews, total_tiles, batch, ranges, pair_slots, records,
                                     grad2d, s);
    case 16:
        return blend_backward_for<16>(views_dev, tile_bases, nviews, total_tiles, batch, ranges, pair_slots, records,
                                      grad2d, s);
    case 32:
        return blend_backward_for<32>(views_dev, tile_bases, nviews, total_tiles, batch, ranges, pair_slots, records,
                                      grad2d, s);
    default:
        return hipErrorInvalidValue;
    }
}

}  // namespace athx
