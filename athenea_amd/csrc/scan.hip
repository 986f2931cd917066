// scan.hip -- exclusive prefix sums in one pass: each workgroup takes a
// ticket (so every partition it waits on has already started and will
// finish), reduces its 4096 elements, publishes the aggregate, and one
// wavefront looks back 64 partitions at a time for the inclusive prefix of
// its predecessors (decoupled look-back, Merrill & Garland 2016).
//
// Three uses share the kernel, each a small policy type:
//   PlainScan    out[i] = in[0] + ... + in[i-1] and the total: gpu::PrefixSum::
//                apply (modules/gpu/include/athenea/gpu/algo/PrefixSum.h:43-44;
//                modules/gpu/src/algo/PrefixSum.cpp:58-88), which the reference
//                runs as three dispatches (prefix_chunk_totals/starts/local).
//   CompactScan  the visible slots, in slot order, with their depth keys:
//                splat_compact.slang:16-26 over the scan of `visible`.
//   GatherScan   the tile counts in depth order, scanned into each slot's first
//                pair: splat_gather_counts.slang:11-19 and the scan after it
//                (TileRasterizer.cpp:1129-1137). Its length is read on the
//                device (the compaction's total), so no host round trip.
//   BigScan      the footprints over kBigRect tiles, listed in slot order with
//                the first of their hit-word steps: one 64-bit sum, the count
//                in the high half and the steps in the low.
//   EntryScan    the strip entries of the depth-sorted slots (the frame's
//                binning, in place of the reference's tile counts): only
//                every 64th rank's offset is kept, one a wavefront of the
//                strip emission.
//   SegScan      each strip's segments (kStripSeg entries each).
//   TableScan    the binning's hit counts in place, their length read on the
//                device: the pair total and every (tile, segment)'s start.
#include "athx_internal.h"

namespace athx {
namespace {

constexpr uint32_t kThreads = 256;
constexpr uint32_t kItems = 16;
constexpr uint32_t kTile = kThreads * kItems;
constexpr uint64_t kFlagAggregate = 1ull << 62;
constexpr uint64_t kFlagPrefix = 2ull << 62;
constexpr uint64_t kValueMask = (1ull << 62) - 1;

__device__ inline uint64_t load_status(const uint64_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void store_status(uint64_t* p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ inline uint64_t wave_inclusive_scan(uint64_t x, uint32_t lane) {
#pragma unroll
    for (uint32_t d = 1; d < 64; d <<= 1) {
        const uint64_t y = __shfl_up(x, d, 64);
        if (lane >= d) {
            x += y;
        }
    }
    return x;
}

__device__ inline uint64_t wave_sum(uint64_t x) {
#pragma unroll
    for (uint32_t d = 32; d >= 1; d >>= 1) {
        x += __shfl_xor(x, d, 64);
    }
    return x;
}

struct PlainScan {
    using Value = uint32_t;
    const uint32_t* in;
    uint32_t* out;
    uint32_t n;
    __device__ uint32_t size() const { return n; }
    __device__ uint32_t load(uint32_t i) const { return in[i]; }
    __device__ void store(uint32_t i, uint32_t, uint64_t excl) const { out[i] = (uint32_t)excl; }
};

struct CompactScan {
    using Value = uint32_t;
    const uint32_t* touched;
    const uint32_t* depth_keys;
    uint32_t* keys_out;
    uint32_t* slots_out;
    uint32_t n;
    __device__ uint32_t size() const { return n; }
    __device__ uint32_t load(uint32_t i) const { return touched[i] != 0 ? 1u : 0u; }
    __device__ void store(uint32_t i, uint32_t v, uint64_t excl) const {
        if (v) {
            keys_out[excl] = depth_keys[i];
            slots_out[excl] = i;
        }
    }
};

struct BigScan {
    using Value = uint64_t;
    const uint2* rects;
    uint32_t* list;
    uint32_t* first_step;
    uint32_t n;
    __device__ uint32_t size() const { return n; }
    __device__ uint64_t load(uint32_t i) const {
        const uint2 r = rects[i];
        const uint32_t w = r.y & 0xFFFFu, h = r.y >> 16;
        return w * h > kBigRect ? (1ull << 32) | big_steps(w, h) : 0ull;
    }
    __device__ void store(uint32_t i, uint64_t v, uint64_t excl) const {
        if (v != 0) {
            list[excl >> 32] = i;
            first_step[excl >> 32] = (uint32_t)excl;
        }
    }
};

struct GatherScan {
    using Value = uint32_t;
    const uint32_t* touched;
    const uint32_t* sorted_slots;
    uint32_t* slot_offsets;
    const uint64_t* n_dev;
    uint32_t cap;
    __device__ uint32_t size() const { return *n_dev < cap ? (uint32_t)*n_dev : cap; }
    __device__ uint32_t load(uint32_t k) const { return touched[sorted_slots[k]]; }
    __device__ void store(uint32_t k, uint32_t, uint64_t excl) const { slot_offsets[sorted_slots[k]] = (uint32_t)excl; }
};

struct EntryScan {
    using Value = uint32_t;
    const uint32_t* touched;
    const uint32_t* sorted_slots;
    uint32_t* offsets;  // [ranks / 64]
    const uint64_t* n_dev;
    uint32_t cap;
    __device__ uint32_t size() const { return *n_dev < cap ? (uint32_t)*n_dev : cap; }
    __device__ uint32_t load(uint32_t k) const { return touched[sorted_slots[k]]; }
    __device__ void store(uint32_t k, uint32_t, uint64_t excl) const {
        if ((k & 63u) == 0) {
            offsets[k >> 6] = (uint32_t)excl;
        }
    }
};

struct SegScan {
    using Value = uint32_t;
    const uint32_t* ranges;
    uint32_t* segbase;
    uint32_t n;
    __device__ uint32_t size() const { return n; }
    __device__ uint32_t load(uint32_t s) const { return div_up(ranges[2u * s + 1u] - ranges[2u * s], kStripSeg); }
    __device__ void store(uint32_t s, uint32_t, uint64_t excl) const { segbase[s] = (uint32_t)excl; }
};

struct TableScan {
    using Value = uint32_t;
    uint32_t* table;
    const uint64_t* segs_dev;
    uint32_t seg_cap;
    __device__ uint32_t size() const { return kStripTiles * (*segs_dev < seg_cap ? (uint32_t)*segs_dev : seg_cap); }
    __device__ uint32_t load(uint32_t i) const { return table[i]; }
    __device__ void store(uint32_t i, uint32_t, uint64_t excl) const { table[i] = (uint32_t)excl; }
};

// status[0] is the ticket counter; status[1 + p] partition p's word. The grid
// may hold more partitions than the scan's length needs (a length read on the
// device): tickets past it leave at once, and nobody waits on them.
template <typename Policy>
__global__ void __launch_bounds__(kThreads) scan_kernel(Policy pol, uint64_t* __restrict__ total,
                                                        uint64_t* __restrict__ status) {
    __shared__ uint32_t s_part;
    __shared__ uint64_t s_wave[kThreads / 64];
    __shared__ uint64_t s_prefix;
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & 63u;
    const uint32_t wave = tid >> 6;
    if (tid == 0) {
        s_part = atomicAdd(reinterpret_cast<uint32_t*>(status), 1u);
    }
    __syncthreads();
    const uint32_t n = pol.size();
    const uint32_t part = s_part;
    const uint32_t parts = div_up(n ? n : 1, kTile);
    if (part >= parts) {
        return;
    }
    const uint64_t first = (uint64_t)part * kTile + (uint64_t)tid * kItems;

    typename Policy::Value v[kItems];
#pragma unroll
    for (uint32_t k = 0; k < kItems; ++k) {
        v[k] = first + k < n ? pol.load((uint32_t)(first + k)) : 0u;
    }
    uint64_t mine = 0;
#pragma unroll
    for (uint32_t k = 0; k < kItems; ++k) {
        mine += v[k];
    }
    const uint64_t incl = wave_inclusive_scan(mine, lane);
    if (lane == 63) {
        s_wave[wave] = incl;
    }
    __syncthreads();
    uint64_t wave_before = 0, aggregate = 0;
#pragma unroll
    for (uint32_t w = 0; w < kThreads / 64; ++w) {
        const uint64_t t = s_wave[w];
        wave_before += w < wave ? t : 0;
        aggregate += t;
    }

    if (wave == 0) {
        uint64_t exclusive = 0;
        if (part == 0) {
            if (lane == 0) {
                store_status(status + 1, kFlagPrefix | (aggregate & kValueMask));
            }
        } else {
            if (lane == 0) {
                store_status(status + 1 + part, kFlagAggregate | (aggregate & kValueMask));
            }
            int64_t j = (int64_t)part - 1 - (int64_t)lane;
            while (true) {
                uint64_t w = j >= 0 ? load_status(status + 1 + j) : kFlagPrefix;
                // Every partition before this one holds a ticket, so each
                // word read here is published in finite time.
                while (!__all((w >> 62) != 0)) {
                    __builtin_amdgcn_s_sleep(1);
                    if ((w >> 62) == 0) {
                        w = load_status(status + 1 + j);
                    }
                }
                const uint64_t prefix_lanes = __ballot((w >> 62) == 2);
                if (prefix_lanes != 0) {
                    const uint32_t stop = __ffsll((unsigned long long)prefix_lanes) - 1;
                    exclusive += wave_sum(lane <= stop ? (w & kValueMask) : 0);
                    break;
                }
                exclusive += wave_sum(w & kValueMask);
                j -= 64;
            }
            if (lane == 0) {
                store_status(status + 1 + part, kFlagPrefix | ((exclusive + aggregate) & kValueMask));
            }
        }
        if (lane == 0) {
            s_prefix = exclusive;
            if (part == parts - 1) {
                *total = exclusive + aggregate;
            }
        }
    }
    __syncthreads();
    uint64_t run = s_prefix + wave_before + (incl - mine);
#pragma unroll
    for (uint32_t k = 0; k < kItems; ++k) {
        if (first + k < n) {
            pol.store((uint32_t)(first + k), v[k], run);
        }
        run += v[k];
    }
}

// Zeroes the tickets and the words of the partitions the scan will use
// (ticket counter included), and the total for an empty scan.
template <typename Policy>
__global__ void clear_status_kernel(Policy pol, uint64_t* __restrict__ status, uint64_t* __restrict__ total) {
    const uint32_t n = pol.size();
    const uint64_t words = 1 + (uint64_t)div_up(n ? n : 1, kTile);
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += (uint64_t)gridDim.x * blockDim.x) {
        status[i] = 0;
    }
    if (n == 0 && blockIdx.x == 0 && threadIdx.x == 0) {
        *total = 0;
    }
}

template <typename Policy>
hipError_t run_scan(const Policy& pol, uint32_t max_n, uint64_t* total, uint32_t* status_words, hipStream_t s) {
    uint64_t* status = reinterpret_cast<uint64_t*>(status_words);
    const uint32_t parts = div_up(max_n ? max_n : 1, kTile);
    const uint32_t clear_blocks = div_up(parts + 1, 256) < 1024 ? div_up(parts + 1, 256) : 1024;
    hipLaunchKernelGGL(clear_status_kernel<Policy>, dim3(clear_blocks), dim3(256), 0, s, pol, status, total);
    hipLaunchKernelGGL(scan_kernel<Policy>, dim3(parts), dim3(kThreads), 0, s, pol, total, status);
    return hipGetLastError();
}

}  // namespace

// u32 words of status a scan of up to n elements needs (u64 words, counted in u32).
size_t scan_status_words(uint32_t n) { return 2 * (1 + (size_t)div_up(n ? n : 1, kTile)); }

hipError_t launch_scan(const uint32_t* in, uint32_t* out, uint64_t* total, uint32_t n, uint32_t* status_words,
                       hipStream_t s) {
    return run_scan(PlainScan{in, out, n}, n, total, status_words, s);
}

hipError_t launch_compact(const uint32_t* touched, const uint32_t* depth_keys, uint32_t slots, uint32_t* keys_out,
                          uint32_t* slots_out, uint64_t* visible_total, uint32_t* status_words, hipStream_t s) {
    return run_scan(CompactScan{touched, depth_keys, keys_out, slots_out, slots}, slots, visible_total, status_words,
                    s);
}

hipError_t launch_big_list(const Footprints& fp, uint32_t slots, uint32_t* status_words, hipStream_t s) {
    return run_scan(BigScan{fp.rects, fp.big_list, fp.big_mask_base, slots}, slots, fp.big_total, status_words, s);
}

hipError_t launch_entry_scan(const uint32_t* touched, const uint32_t* sorted_slots, const uint64_t* visible_dev,
                             uint32_t max_visible, uint32_t* offsets, uint64_t* entry_total, uint32_t* status_words,
                             hipStream_t s) {
    return run_scan(EntryScan{touched, sorted_slots, offsets, visible_dev, max_visible}, max_visible, entry_total,
                    status_words, s);
}

hipError_t launch_seg_scan(const uint32_t* strip_ranges, uint32_t strips, uint32_t* segbase, uint64_t* seg_total,
                           uint32_t* status_words, hipStream_t s) {
    return run_scan(SegScan{strip_ranges, segbase, strips}, strips, seg_total, status_words, s);
}

hipError_t launch_table_scan(uint32_t* table, const uint64_t* seg_total, uint32_t seg_cap, uint64_t* pair_total,
                             uint32_t* status_words, hipStream_t s) {
    return run_scan(TableScan{table, seg_total, seg_cap}, kStripTiles * seg_cap, pair_total, status_words, s);
}

hipError_t launch_gather_scan(const uint32_t* touched, const uint32_t* sorted_slots, const uint64_t* visible_dev,
                              uint32_t max_visible, uint32_t* slot_offsets, uint64_t* pair_total,
                              uint32_t* status_words, hipStream_t s) {
    return run_scan(GatherScan{touched, sorted_slots, slot_offsets, visible_dev, max_visible}, max_visible, pair_total,
                    status_words, s);
}

}  // namespace athx
