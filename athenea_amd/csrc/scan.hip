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

#include <type_traits>

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

// A rank's entry count rides in its sorted depth key's top byte (the
// compaction puts it there, 255 meaning "255 or more: read the slot's").
struct EntryScan {
    using Value = uint32_t;
    const uint32_t* touched;
    const uint32_t* sorted_slots;
    const uint32_t* sorted_keys;
    uint32_t* offsets;  // [ranks / 64]
    const uint64_t* n_dev;
    uint32_t cap;
    // Read in two steps (every key first, then the rare 255s' slots), so a
    // partition's key loads are all in flight at once.
    static constexpr bool kTwoStep = true;
    __device__ uint32_t size() const { return *n_dev < cap ? (uint32_t)*n_dev : cap; }
    __device__ uint32_t raw(uint32_t k) const { return sorted_keys[k]; }
    __device__ uint32_t resolve(uint32_t k, uint32_t key) const {
        const uint32_t c = key >> 24;
        return c != 255u ? c : touched[sorted_slots[k]];
    }
    __device__ uint32_t load(uint32_t k) const { return resolve(k, raw(k)); }
    __device__ void store(uint32_t k, uint32_t, uint64_t excl) const {
        if ((k & 63u) == 0) {
            offsets[k >> 6] = (uint32_t)excl;
        }
    }
};

// Policies whose load splits into a raw read and a resolve (kTwoStep).
template <typename P, typename = void>
struct TwoStep {
    static constexpr bool value = false;
};
template <typename P>
struct TwoStep<P, std::void_t<decltype(P::kTwoStep)>> {
    static constexpr bool value = P::kTwoStep;
};

// Elements at(0) .. at(kItems - 1) (0 past n): a two-step policy's raw reads
// are issued together, at in-bounds indices, before any is resolved.
template <typename Policy, typename At>
__device__ inline void load_items(const Policy& pol, uint32_t n, At at, typename Policy::Value (&v)[kItems]) {
    if constexpr (TwoStep<Policy>::value) {
        if (n == 0) {
#pragma unroll
            for (uint32_t k = 0; k < kItems; ++k) {
                v[k] = 0;
            }
            return;
        }
        uint32_t r[kItems];
#pragma unroll
        for (uint32_t k = 0; k < kItems; ++k) {
            const uint64_t e = at(k);
            r[k] = pol.raw(e < n ? (uint32_t)e : n - 1u);
        }
#pragma unroll
        for (uint32_t k = 0; k < kItems; ++k) {
            const uint64_t e = at(k);
            v[k] = e < n ? pol.resolve((uint32_t)e, r[k]) : 0u;
        }
    } else {
#pragma unroll
        for (uint32_t k = 0; k < kItems; ++k) {
            const uint64_t e = at(k);
            v[k] = e < n ? pol.load((uint32_t)e) : 0u;
        }
    }
}

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

// A partition's elements, kItems consecutive ones a thread (the scan's
// order), read with coalesced loads: row k of kThreads consecutive elements
// a load, through LDS (rows of kItems + 1 words, so neither side's accesses
// collide in a bank).
static_assert(kItems == 16, "the transpose indexes rows of 16");
template <typename Policy>
__device__ inline void load_partition(const Policy& pol, uint64_t part_first, uint32_t n,
                                      typename Policy::Value (&v)[kItems], typename Policy::Value* s_x) {
    const uint32_t tid = threadIdx.x;
    typename Policy::Value x[kItems];
    load_items(pol, n, [&](uint32_t k) { return part_first + k * kThreads + tid; }, x);
#pragma unroll
    for (uint32_t k = 0; k < kItems; ++k) {
        const uint32_t r = k * kThreads + tid;
        s_x[(r >> 4) * (kItems + 1) + (r & 15u)] = x[k];
    }
    __syncthreads();
#pragma unroll
    for (uint32_t k = 0; k < kItems; ++k) {
        v[k] = s_x[tid * (kItems + 1) + k];
    }
}

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

    // Each lane's run read directly: at the sizes this kernel takes, a
    // partition's lines stay in cache between a lane's loads.
    typename Policy::Value v[kItems];
    load_items(pol, n, [&](uint32_t k) { return first + k; }, v);
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

// Long scans (past kLongScan elements, 256 partitions) go in four launches
// instead: each partition's sum, the sums scanned in two levels, then each
// partition scanned from its offset. Past a few hundred partitions in flight
// the look-back above waits on a chain of predecessors (measured on MI355X:
// a 4 M-element scan 53 us that way, about 25 us this way); here nothing
// waits, and the data is read twice.
constexpr uint32_t kLongScan = 1u << 20;
constexpr uint32_t kSumThreads = 1024;

template <typename Policy>
__global__ void __launch_bounds__(kThreads) partition_sum_kernel(Policy pol, uint64_t* __restrict__ sums) {
    __shared__ uint64_t s_wave[kThreads / 64];
    const uint32_t n = pol.size();
    const uint32_t part = blockIdx.x;
    if ((uint64_t)part * kTile >= n && part != 0) {
        return;
    }
    // The sum in any order: rows of kThreads consecutive elements a load.
    const uint64_t first = (uint64_t)part * kTile + threadIdx.x;
    typename Policy::Value x[kItems];
    load_items(pol, n, [&](uint32_t k) { return first + (uint64_t)k * kThreads; }, x);
    uint64_t mine = 0;
#pragma unroll
    for (uint32_t k = 0; k < kItems; ++k) {
        mine += (uint64_t)x[k];
    }
    mine = wave_sum(mine);
    if ((threadIdx.x & 63u) == 0) {
        s_wave[threadIdx.x >> 6] = mine;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t t = 0;
#pragma unroll
        for (uint32_t w = 0; w < kThreads / 64; ++w) {
            t += s_wave[w];
        }
        sums[part] = t;
    }
}

// The partitions' sums scanned in two levels, so that no single workgroup
// walks them all (one did, its loads a lane apart each a line of their own,
// at tens of microseconds past 30 000 partitions): each run of kSumThreads
// sums scanned in place by a workgroup, its total kept in tops[run]; then
// one workgroup scans the tops (at most kSumThreads of them), the grand total
// to *total. A sum's offset is then sums[p] + tops[p / kSumThreads].
template <typename Policy>
__global__ void __launch_bounds__(kSumThreads) run_scan_kernel(Policy pol, uint64_t* __restrict__ sums,
                                                               uint64_t* __restrict__ tops) {
    __shared__ uint64_t s_wave[kSumThreads / 64];
    const uint32_t n = pol.size();
    const uint32_t parts = div_up(n ? n : 1, kTile);
    if (blockIdx.x * kSumThreads >= parts) {
        return;
    }
    const uint32_t i = blockIdx.x * kSumThreads + threadIdx.x;
    const uint64_t x = i < parts ? sums[i] : 0u;
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint64_t incl = wave_inclusive_scan(x, lane);
    if (lane == 63) {
        s_wave[wave] = incl;
    }
    __syncthreads();
    uint64_t before = 0, all = 0;
#pragma unroll
    for (uint32_t w = 0; w < kSumThreads / 64; ++w) {
        const uint64_t t = s_wave[w];
        before += w < wave ? t : 0u;
        all += t;
    }
    if (i < parts) {
        sums[i] = before + incl - x;
    }
    if (threadIdx.x == 0) {
        tops[blockIdx.x] = all;
    }
}

template <typename Policy>
__global__ void __launch_bounds__(kSumThreads) tops_scan_kernel(Policy pol, uint64_t* __restrict__ tops,
                                                                uint64_t* __restrict__ total) {
    __shared__ uint64_t s_wave[kSumThreads / 64];
    const uint32_t n = pol.size();
    const uint32_t runs = div_up(div_up(n ? n : 1, kTile), kSumThreads);
    const uint64_t x = threadIdx.x < runs ? tops[threadIdx.x] : 0u;
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint64_t incl = wave_inclusive_scan(x, lane);
    if (lane == 63) {
        s_wave[wave] = incl;
    }
    __syncthreads();
    uint64_t before = 0, all = 0;
#pragma unroll
    for (uint32_t w = 0; w < kSumThreads / 64; ++w) {
        const uint64_t t = s_wave[w];
        before += w < wave ? t : 0u;
        all += t;
    }
    if (threadIdx.x < runs) {
        tops[threadIdx.x] = before + incl - x;
    }
    if (threadIdx.x == 0) {
        *total = n ? all : 0u;
    }
}

// A count as a policy: what the compaction's two-level scan needs of one.
struct CountOnly {
    uint32_t n;
    __device__ uint32_t size() const { return n; }
};

template <typename Policy>
__global__ void __launch_bounds__(kThreads) partition_scan_kernel(Policy pol, const uint64_t* __restrict__ offsets,
                                                                  const uint64_t* __restrict__ tops) {
    __shared__ uint64_t s_wave[kThreads / 64];
    __shared__ typename Policy::Value s_x[kThreads * (kItems + 1)];
    const uint32_t n = pol.size();
    const uint32_t part = blockIdx.x;
    if ((uint64_t)part * kTile >= n) {
        return;
    }
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint64_t first = (uint64_t)part * kTile + (uint64_t)threadIdx.x * kItems;
    typename Policy::Value v[kItems];
    load_partition(pol, (uint64_t)part * kTile, n, v, s_x);
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
    uint64_t run = offsets[part] + tops[part / kSumThreads] + (incl - mine);
#pragma unroll
    for (uint32_t w = 0; w < kThreads / 64; ++w) {
        run += w < wave ? s_wave[w] : 0u;
    }
#pragma unroll
    for (uint32_t k = 0; k < kItems; ++k) {
        if (first + k < n) {
            pol.store((uint32_t)(first + k), v[k], run);
        }
        run += v[k];
    }
}

template <typename Policy>
hipError_t run_scan(const Policy& pol, uint32_t max_n, uint64_t* total, uint32_t* status_words, hipStream_t s) {
    uint64_t* status = reinterpret_cast<uint64_t*>(status_words);
    const uint32_t parts = div_up(max_n ? max_n : 1, kTile);
    if (max_n > kLongScan) {
        // The status words hold one u64 a partition, then one a run of
        // kSumThreads partitions (scan_status_words).
        uint64_t* tops = status + parts;
        hipLaunchKernelGGL(partition_sum_kernel<Policy>, dim3(parts), dim3(kThreads), 0, s, pol, status);
        hipLaunchKernelGGL(run_scan_kernel<Policy>, dim3(div_up(parts, kSumThreads)), dim3(kSumThreads), 0, s, pol,
                           status, tops);
        hipLaunchKernelGGL(tops_scan_kernel<Policy>, dim3(1), dim3(kSumThreads), 0, s, pol, tops, total);
        hipLaunchKernelGGL(partition_scan_kernel<Policy>, dim3(parts), dim3(kThreads), 0, s, pol, status, tops);
        return hipGetLastError();
    }
    const uint32_t clear_blocks = div_up(parts + 1, 256) < 1024 ? div_up(parts + 1, 256) : 1024;
    hipLaunchKernelGGL(clear_status_kernel<Policy>, dim3(clear_blocks), dim3(256), 0, s, pol, status, total);
    hipLaunchKernelGGL(scan_kernel<Policy>, dim3(parts), dim3(kThreads), 0, s, pol, total, status);
    return hipGetLastError();
}

// The visible slots in slot order (splat_compact.slang:16-26), for any
// number of slots, in three launches that neither wait nor scatter: each
// partition of kTile slots counts its visible ones, one workgroup scans the
// counts, and each partition writes its visible slots from its offset. A
// partition is read in kItems rows of 256 consecutive slots, a slot a lane,
// so every load is coalesced; a row's visible slots are ranked by a ballot,
// so their (key, slot) stores are consecutive too. The key written is the
// slot's depth key with its strip-entry count (capped at 255) in the top
// byte, which the 24-bit depth sort carries along for the entry scan.
__global__ void __launch_bounds__(kThreads) compact_count_kernel(const uint32_t* __restrict__ touched, uint32_t n,
                                                                 uint64_t* __restrict__ sums) {
    __shared__ uint32_t s_wave[kThreads / 64];
    const uint64_t first = (uint64_t)blockIdx.x * kTile;
    uint32_t count = 0;
#pragma unroll
    for (uint32_t k = 0; k < kItems; ++k) {
        const uint64_t i = first + (uint64_t)k * kThreads + threadIdx.x;
        count += (uint32_t)__popcll(__ballot(i < n && touched[i] != 0));
    }
    if ((threadIdx.x & 63u) == 0) {
        s_wave[threadIdx.x >> 6] = count;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t t = 0;
#pragma unroll
        for (uint32_t w = 0; w < kThreads / 64; ++w) {
            t += s_wave[w];
        }
        sums[blockIdx.x] = t;
    }
}

__global__ void __launch_bounds__(kThreads) compact_write_kernel(const uint32_t* __restrict__ touched,
                                                                 const uint32_t* __restrict__ depth_keys, uint32_t n,
                                                                 const uint64_t* __restrict__ offsets,
                                                                 const uint64_t* __restrict__ tops,
                                                                 uint32_t* __restrict__ keys_out,
                                                                 uint32_t* __restrict__ slots_out,
                                                                 uint32_t per_view, uint32_t magic,
                                                                 uint32_t* __restrict__ view_start) {
    __shared__ uint32_t s_count[kItems][kThreads / 64];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint64_t lt_mask = (1ull << lane) - 1ull;
    const uint64_t first = (uint64_t)blockIdx.x * kTile;
    // Every slot's entry count and depth key read at once, before any is
    // waited on (the keys of slots that are not drawn are read and dropped).
    uint32_t entries[kItems], dkey[kItems];
    uint64_t bal[kItems];
#pragma unroll
    for (uint32_t k = 0; k < kItems; ++k) {
        const uint64_t i = first + (uint64_t)k * kThreads + threadIdx.x;
        entries[k] = i < n ? touched[i] : 0u;
        dkey[k] = i < n ? depth_keys[i] : 0u;
    }
#pragma unroll
    for (uint32_t k = 0; k < kItems; ++k) {
        bal[k] = __ballot(entries[k] != 0);
        if (lane == 0) {
            s_count[k][wave] = (uint32_t)__popcll(bal[k]);
        }
    }
    __syncthreads();
    // Exclusive prefix of (row, wavefront) counts, in slot order.
    uint64_t run = offsets[blockIdx.x] + tops[blockIdx.x / kSumThreads];
    // The first view starting at or after the block's first slot.
    const uint32_t boundary = per_view >= kTile ? (uint32_t)((first + per_view - 1u) / per_view) : 0u;
#pragma unroll
    for (uint32_t k = 0; k < kItems; ++k) {
        uint32_t before = 0, row = 0;
#pragma unroll
        for (uint32_t w = 0; w < kThreads / 64; ++w) {
            const uint32_t c = s_count[k][w];
            before += w < wave ? c : 0u;
            row += c;
        }
        const uint64_t i = first + (uint64_t)k * kThreads + threadIdx.x;
        if (view_start && i < n) {
            // A view's first slot (slot = view * per_view + Gaussian) records
            // where the view's visible slots start: the visible ones before it.
            // Views of a block's size or more have at most one first slot in
            // a block, found once; smaller ones divide each slot.
            bool starts;
            uint32_t q;
            if (per_view >= kTile) {
                q = boundary;
                starts = (uint64_t)boundary * per_view == i;
            } else {
                q = (uint32_t)(((uint64_t)(uint32_t)i * magic) >> 32);
                q += (uint64_t)(q + 1u) * per_view <= i ? 1u : 0u;
                starts = (uint64_t)q * per_view == i;
            }
            if (starts) {
                view_start[q] = (uint32_t)(run + before + (uint64_t)__popcll(bal[k] & lt_mask));
            }
        }
        if (entries[k] != 0) {
            const uint64_t at = run + before + (uint64_t)__popcll(bal[k] & lt_mask);
            const uint32_t count = entries[k] < 255u ? entries[k] : 255u;
            keys_out[at] = (dkey[k] & 0xFFFFFFu) | (count << 24);
            slots_out[at] = (uint32_t)i;
        }
        run += row;
    }
}

}  // namespace

// u32 words of status a scan of up to n elements needs: the look-back's u64
// a partition and one more, or the long route's u64 a partition and one a
// run of kSumThreads of them (and one more).
size_t scan_status_words(uint32_t n) {
    const size_t parts = div_up(n ? n : 1, kTile);
    return 2 * (2 + parts + div_up64(parts, kSumThreads));
}

hipError_t launch_scan(const uint32_t* in, uint32_t* out, uint64_t* total, uint32_t n, uint32_t* status_words,
                       hipStream_t s) {
    return run_scan(PlainScan{in, out, n}, n, total, status_words, s);
}

hipError_t launch_compact(const uint32_t* touched, const uint32_t* depth_keys, uint32_t slots, uint32_t* keys_out,
                          uint32_t* slots_out, uint64_t* visible_total, uint32_t* status_words, hipStream_t s,
                          bool counted, uint32_t per_view, uint32_t* view_start) {
    static_assert(kTile == kCompactTile, "the settle pass counts the compaction's blocks");
    uint64_t* sums = reinterpret_cast<uint64_t*>(status_words);
    const uint32_t parts = div_up(slots ? slots : 1, kTile);
    if (!counted || slots == 0) {
        hipLaunchKernelGGL(compact_count_kernel, dim3(parts), dim3(kThreads), 0, s, touched, slots, sums);
    }
    uint64_t* tops = sums + parts;
    const CountOnly count{slots};
    hipLaunchKernelGGL(run_scan_kernel<CountOnly>, dim3(div_up(parts, kSumThreads)), dim3(kSumThreads), 0, s, count,
                       sums, tops);
    hipLaunchKernelGGL(tops_scan_kernel<CountOnly>, dim3(1), dim3(kSumThreads), 0, s, count, tops, visible_total);
    const uint32_t magic = per_view >= 2 ? (uint32_t)((1ull << 32) / per_view) : 0u;
    hipLaunchKernelGGL(compact_write_kernel, dim3(parts), dim3(kThreads), 0, s, touched, depth_keys, slots, sums,
                       tops, keys_out, slots_out, per_view, magic, per_view >= 2 ? view_start : nullptr);
    return hipGetLastError();
}

hipError_t launch_big_list(const Footprints& fp, uint32_t slots, uint32_t* status_words, hipStream_t s) {
    return run_scan(BigScan{fp.rects, fp.big_list, fp.big_mask_base, slots}, slots, fp.big_total, status_words, s);
}

hipError_t launch_entry_scan(const uint32_t* touched, const uint32_t* sorted_slots, const uint32_t* sorted_keys,
                             const uint64_t* visible_dev, uint32_t max_visible, uint32_t* offsets,
                             uint64_t* entry_total, uint32_t* status_words, hipStream_t s) {
    return run_scan(EntryScan{touched, sorted_slots, sorted_keys, offsets, visible_dev, max_visible}, max_visible,
                    entry_total, status_words, s);
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
