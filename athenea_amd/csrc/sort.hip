// sort.hip -- stable LSD radix sort of (uint32 or uint64 key, uint32 value)
// pairs for gfx950, onesweep (Adinets & Merrill 2022): one kernel counts every digit
// pass's histogram in a single read of the keys, then one kernel a pass
// ranks each 4096-key partition in LDS with wavefront ballots (a 64-lane
// match on the 8 digit bits), finds where its digits start among the
// partitions before it by decoupled look-back, and writes each digit's run
// out contiguously from LDS.
//
// Semantics: gpu::RadixSort::sort (modules/gpu/include/athenea/gpu/algo/
// RadixSort.h:63-64; modules/gpu/src/algo/RadixSort.cpp:95-261), a stable
// LSD sort on the low key_bits bits, 8 bits a pass. The reference splits a
// pass into histogram / scan / scatter dispatches over tiles of 1024
// (shaders/athenea/algo/radix_block.slang:79-176).
#include "athx_internal.h"

#ifndef ATHX_REDUCE_SCAN_MIN
#define ATHX_REDUCE_SCAN_MIN (1u << 22)
#endif

namespace athx {
namespace {

constexpr uint32_t kThreads = 256;
constexpr uint32_t kWaves = kThreads / 64;
constexpr uint32_t kItems = 16;
constexpr uint32_t kPart = kThreads * kItems;
constexpr uint32_t kRadix = 256;
constexpr uint32_t kMaxPasses = 8;
constexpr uint32_t kFlagAggregate = 1u << 30;
constexpr uint32_t kFlagPrefix = 2u << 30;
constexpr uint32_t kValueMask = (1u << 30) - 1;

__device__ inline uint32_t load_word(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void store_word(uint32_t* p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Exclusive scan of one uint32 a thread over the kW wavefronts of the block.
template <uint32_t kW = kWaves>
__device__ inline uint32_t block_exclusive_scan(uint32_t x, uint32_t* s_tmp, uint32_t& total) {
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    uint32_t incl = x;
#pragma unroll
    for (uint32_t d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(incl, d, 64);
        if (lane >= d) {
            incl += y;
        }
    }
    if (lane == 63) {
        s_tmp[wave] = incl;
    }
    __syncthreads();
    uint32_t before = 0, all = 0;
#pragma unroll
    for (uint32_t w = 0; w < kW; ++w) {
        const uint32_t t = s_tmp[w];
        before += w < wave ? t : 0u;
        all += t;
    }
    __syncthreads();
    total = all;
    return before + incl - x;
}

// The pair count lives in device memory (the scan's total), bounded by the
// buffers' capacity, so a frame's launches need no host round trip.
__device__ inline uint32_t count_of(const uint64_t* n_dev, uint32_t cap) {
    const uint64_t n = *n_dev;
    return n < cap ? (uint32_t)n : cap;
}

// The digits of a sort over `bits` key bits from lo_bit: ceil(bits / 8)
// passes of as near equal widths as can be (13 bits: 7 then 6), so no pass
// has more digits than it must and each digit's run out of a partition is
// as long as it can be.
__host__ __device__ inline void pass_digit(uint32_t p, uint32_t lo_bit, uint32_t bits, uint32_t passes,
                                           uint32_t& shift, uint32_t& width) {
    const uint32_t base = bits / passes, extra = bits % passes;
    width = base + (p < extra ? 1u : 0u);
    shift = lo_bit + p * base + (p < extra ? p : extra);
}

// The lanes of the wavefront whose digit equals this lane's: a ballot a bit.
__device__ inline uint64_t match_digit(uint32_t d, uint32_t width, bool ok) {
    uint64_t m = __ballot(ok);
#pragma unroll
    for (uint32_t b = 0; b < 8; ++b) {
        if (b < width) {
            const bool bit = (d >> b) & 1u;
            const uint64_t on = __ballot(bit);
            m &= bit ? on : ~on;
        }
    }
    return m;
}

// Every pass's digit counts in one read of the keys, a block's counts in LDS
// and then written out as its own column (partials[digit row][block]); a
// second kernel sums the rows. Adding 256 words a block into one histogram
// with device-scope atomics serialises in the memory system.
constexpr uint32_t kHistBlocks = 1024;
// Sorts of at least this capacity take the reduce-scan-scatter passes;
// smaller ones onesweep's, whose few partitions wait little on each other.
constexpr uint32_t kReduceScanMin = ATHX_REDUCE_SCAN_MIN;

template <typename K>
__global__ void __launch_bounds__(kThreads) histogram_kernel(const K* __restrict__ keys,
                                                             const uint64_t* __restrict__ n_dev, uint32_t cap,
                                                             uint32_t lo_bit, uint32_t bits, uint32_t passes,
                                                             uint32_t* __restrict__ partials) {
    constexpr uint32_t kUnroll = 4;
    __shared__ uint32_t s_hist[kMaxPasses * kRadix];
    const uint32_t n = count_of(n_dev, cap);
    for (uint32_t i = threadIdx.x; i < passes * kRadix; i += kThreads) {
        s_hist[i] = 0;
    }
    __syncthreads();
    for (uint32_t base = blockIdx.x * kThreads * kUnroll; base < n; base += gridDim.x * kThreads * kUnroll) {
        K k[kUnroll];
#pragma unroll
        for (uint32_t u = 0; u < kUnroll; ++u) {
            const uint32_t i = base + u * kThreads + threadIdx.x;
            k[u] = i < n ? keys[i] : K(0);
        }
#pragma unroll
        for (uint32_t u = 0; u < kUnroll; ++u) {
            if (base + u * kThreads + threadIdx.x < n) {
                for (uint32_t p = 0; p < passes; ++p) {
                    uint32_t shift, width;
                    pass_digit(p, lo_bit, bits, passes, shift, width);
                    atomicAdd(&s_hist[p * kRadix + ((uint32_t)(k[u] >> shift) & ((1u << width) - 1u))], 1u);
                }
            }
        }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < passes * kRadix; i += kThreads) {
        partials[(size_t)i * gridDim.x + blockIdx.x] = s_hist[i];
    }
}

// hist[row] = the sum of the blocks' counts in that row, a block a row.
__global__ void __launch_bounds__(kThreads) hist_reduce_kernel(const uint32_t* __restrict__ partials,
                                                               uint32_t blocks, uint32_t* __restrict__ hist) {
    __shared__ uint32_t s_sum[kWaves];
    const uint32_t* row = partials + (size_t)blockIdx.x * blocks;
    uint32_t sum = 0;
    for (uint32_t b = threadIdx.x; b < blocks; b += kThreads) {
        sum += row[b];
    }
#pragma unroll
    for (uint32_t d = 32; d >= 1; d >>= 1) {
        sum += __shfl_xor(sum, d, 64);
    }
    if ((threadIdx.x & 63u) == 0) {
        s_sum[threadIdx.x >> 6] = sum;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t total = 0;
#pragma unroll
        for (uint32_t w = 0; w < kWaves; ++w) {
            total += s_sum[w];
        }
        hist[blockIdx.x] = total;
    }
}

template <typename K>
__global__ void __launch_bounds__(kThreads) onesweep_kernel(const K* __restrict__ kin,
                                                            const uint32_t* __restrict__ vin,
                                                            K* __restrict__ kout, uint32_t* __restrict__ vout,
                                                            const uint64_t* __restrict__ n_dev, uint32_t cap,
                                                            uint32_t shift, uint32_t width,
                                                            const uint32_t* __restrict__ hist,
                                                            uint32_t* __restrict__ lookback,
                                                            uint32_t* __restrict__ counter) {
    __shared__ K s_keys[kPart];
    __shared__ uint32_t s_vals[kPart];
    __shared__ uint32_t s_wave_hist[kWaves * kRadix];
    __shared__ uint32_t s_block_start[kRadix];
    __shared__ uint32_t s_digit_base[kRadix];
    __shared__ uint32_t s_tmp[kWaves];
    __shared__ uint32_t s_part;

    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & 63u;
    const uint32_t wave = tid >> 6;
    if (tid == 0) {
        s_part = atomicAdd(counter, 1u);
    }
    for (uint32_t i = tid; i < kWaves * kRadix; i += kThreads) {
        s_wave_hist[i] = 0;
    }
    __syncthreads();
    const uint32_t n = count_of(n_dev, cap);
    const uint32_t part = s_part;
    // The grid is sized for the capacity: tickets past this frame's
    // partitions leave at once. Nobody waits on them -- a partition waits
    // only on lower tickets.
    if (part >= div_up(n, kPart)) {
        return;
    }
    const uint32_t part_first = part * kPart;
    const uint32_t wave_first = part_first + wave * (64 * kItems);

    // Load: item j of a lane sits at wave_first + j * 64 + lane, so the order
    // of the input is (wave, j, lane) and ranking j by j keeps it stable.
    K key[kItems];
    uint32_t val[kItems];
    uint32_t rank[kItems];
#pragma unroll
    for (uint32_t j = 0; j < kItems; ++j) {
        const uint32_t at = wave_first + j * 64 + lane;
        const bool ok = at < n;
        key[j] = ok ? kin[at] : 0;
        val[j] = ok ? vin[at] : 0;
    }
    const uint64_t lt_mask = (1ull << lane) - 1ull;
    const uint32_t dmask = (1u << width) - 1u;
    uint32_t* my_hist = s_wave_hist + wave * kRadix;
#pragma unroll
    for (uint32_t j = 0; j < kItems; ++j) {
        const uint32_t at = wave_first + j * 64 + lane;
        const bool ok = at < n;
        const uint32_t d = (uint32_t)(key[j] >> shift) & dmask;
        const uint64_t m = match_digit(d, width, ok);
        const uint32_t before = ok ? my_hist[d] : 0u;
        __builtin_amdgcn_wave_barrier();
        rank[j] = before + (uint32_t)__popcll(m & lt_mask);
        const bool leader = ok && (m & lt_mask) == 0;
        if (leader) {
            my_hist[d] = before + (uint32_t)__popcll(m);
        }
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();

    // Thread d owns digit d: offsets of the waves, this partition's count,
    // where the digit starts in the block and in the output.
    const uint32_t d = tid;
    uint32_t count = 0;
#pragma unroll
    for (uint32_t w = 0; w < kWaves; ++w) {
        const uint32_t c = s_wave_hist[w * kRadix + d];
        s_wave_hist[w * kRadix + d] = count;
        count += c;
    }
    // Digits past this pass's radix have no keys and no look-back words.
    const bool digit = d <= dmask;
    uint32_t* my_word = lookback + (size_t)part * kRadix + d;
    if (digit) {
        store_word(my_word, (part == 0 ? kFlagPrefix : kFlagAggregate) | count);
    }
    uint32_t unused;
    s_block_start[d] = block_exclusive_scan(count, s_tmp, unused);
    const uint32_t global_start = block_exclusive_scan(hist[d], s_tmp, unused);
    uint32_t before = 0;
    if (part != 0 && digit) {
        int32_t j = (int32_t)part - 1;
        while (true) {
            uint32_t w = load_word(lookback + (size_t)j * kRadix + d);
            // Partition j took its ticket before this one: it publishes.
            while ((w >> 30) == 0) {
                __builtin_amdgcn_s_sleep(1);
                w = load_word(lookback + (size_t)j * kRadix + d);
            }
            before += w & kValueMask;
            if ((w >> 30) == 2u || j == 0) {
                break;
            }
            --j;
        }
        store_word(my_word, kFlagPrefix | ((before + count) & kValueMask));
    }
    s_digit_base[d] = global_start + before;
    __syncthreads();

    // Into LDS in digit order, then out: each digit's run is contiguous in
    // the output, so neighbouring threads write neighbouring addresses.
#pragma unroll
    for (uint32_t j = 0; j < kItems; ++j) {
        const uint32_t at = wave_first + j * 64 + lane;
        if (at < n) {
            const uint32_t dj = (uint32_t)(key[j] >> shift) & dmask;
            const uint32_t pos = s_block_start[dj] + s_wave_hist[wave * kRadix + dj] + rank[j];
            s_keys[pos] = key[j];
            s_vals[pos] = val[j];
        }
    }
    __syncthreads();
    const uint32_t valid = n - part_first < kPart ? n - part_first : kPart;
    for (uint32_t i = tid; i < valid; i += kThreads) {
        const K k = s_keys[i];
        const uint32_t dk = (uint32_t)(k >> shift) & dmask;
        const uint32_t dst = s_digit_base[dk] + (i - s_block_start[dk]);
        kout[dst] = k;
        vout[dst] = s_vals[i];
    }
}

// Reduce, scan, scatter: the same pass without look-back. upsweep_kernel
// counts each partition's digits into counts[digit][partition], a plain
// scan of that array (digit-major) gives every (digit, partition) its first
// output slot, and downsweep_kernel ranks the partition as onesweep_kernel
// does and scatters from those slots. One more read of the keys a pass, but
// no partition waits on another: on MI355X the look-back's coherent loads
// and the ticket atomics cost more than that read.
// Where a reduce-scan-scatter pass takes its digit from: bits of the key, or
// bits of the key's strip relative to its segment's first (the segmented
// strip sort). `base` is the partition's segment's.
struct KeyDigit {
    uint32_t shift, mask;
    template <typename K>
    __device__ uint32_t operator()(K key, uint32_t, uint32_t) const {
        return (uint32_t)(key >> shift) & mask;
    }
};

struct StripDigit {
    uint32_t shift, mask;
    template <typename K>
    __device__ uint32_t operator()(K key, uint32_t, uint32_t base) const {
        return ((((uint32_t)key & 0xFFFFu) - base) >> shift) & mask;
    }
};

// A partition of a reduce-scan-scatter pass: its elements [first, end) (end
// past first + kP means a whole partition), where its digit counts go
// (cell + digit * stride), and its digit base.
struct Part {
    uint32_t first, end, cell, stride, base;
};

// The first n elements (a count in device memory, bounded by cap), cut into
// parts_cap partitions whose counts lie digit-major: every partition has its
// cells, those past n count nothing.
struct PlainLayout {
    const uint64_t* n_dev;
    uint32_t cap, parts_cap;
    __device__ uint32_t used(uint32_t kP) const { return div_up(count_of(n_dev, cap), kP); }
    __device__ bool part(uint32_t p, uint32_t kP, Part& r) const {
        r.first = p * kP;
        r.end = count_of(n_dev, cap);
        r.cell = p;
        r.stride = parts_cap;
        r.base = 0;
        return true;
    }
};

// Segments (a multi-view frame's views): runs of the array, each cut into
// partitions of its own, so that a partition holds one segment's elements --
// segment s is [seg_start[s], seg_start[s] + seg_len[s]), its partitions
// numbered from seg_part[s]. A segment's counts lie together, digit-major
// within it, so one exclusive scan over every cell gives each (segment,
// digit, partition) its place in the output -- the elements of the earlier
// segments, then of the smaller digits, then of the earlier partitions.
struct SegLayout {
    const uint32_t* part_seg;  // partition -> segment
    const uint32_t* seg_start;
    const uint32_t* seg_len;
    const uint32_t* seg_part;
    const uint32_t* seg_base;  // subtracted from a strip key's low half
    const uint32_t* parts_dev;  // partitions in use
    uint32_t radix;
    __device__ uint32_t used(uint32_t) const { return *parts_dev; }
    __device__ bool part(uint32_t p, uint32_t kP, Part& r) const {
        if (p >= *parts_dev) {
            return false;
        }
        const uint32_t sg = part_seg[p];
        const uint32_t q = p - seg_part[sg];
        r.first = seg_start[sg] + q * kP;
        r.end = seg_start[sg] + seg_len[sg];
        r.cell = radix * seg_part[sg] + q;
        r.stride = div_up(seg_len[sg], kP);
        r.base = seg_base ? seg_base[sg] : 0u;
        return true;
    }
};

// One add a digit a wavefront when every live lane holds the same digit
// (the partitions of a nearly sorted pass), else one a lane.
__device__ inline void hist_add(uint32_t* s_hist, uint32_t d, bool ok) {
    const uint64_t live = __ballot(ok);
    if (live == 0) {
        return;
    }
    const uint32_t first = __builtin_amdgcn_readfirstlane(d);
    if (__ballot(ok && d == first) == live) {
        if ((uint32_t)__builtin_ctzll(live) == (threadIdx.x & 63u)) {
            atomicAdd(&s_hist[first], (uint32_t)__popcll(live));
        }
    } else if (ok) {
        atomicAdd(&s_hist[d], 1u);
    }
}

// The partition a workgroup takes: the partitions in use are dealt to the
// XCDs in contiguous bands (workgroup b runs on XCD b % 8), so neighbouring
// partitions -- whose runs of one digit meet in the output -- share an L2;
// a workgroup past them keeps its own number (and finds no keys).
template <typename L>
__device__ inline uint32_t banded_part(const L& layout, uint32_t kP) {
    const uint32_t used = layout.used(kP);
    return blockIdx.x < used ? xcd_band(blockIdx.x, used) : blockIdx.x;
}

template <typename K, typename D, uint32_t kT, typename L>
__global__ void __launch_bounds__(kT) upsweep_kernel(const K* __restrict__ kin, const uint32_t* __restrict__ vin,
                                                     L layout, D digit, uint32_t width,
                                                     uint32_t* __restrict__ counts) {
    static_assert(kT >= kRadix && kT % 64 == 0, "a thread a digit");
    constexpr uint32_t kP = kT * kItems;
    __shared__ uint32_t s_hist[kRadix];
    const uint32_t tid = threadIdx.x;
    const uint32_t radix = 1u << width;
    Part pt;
    if (!layout.part(blockIdx.x, kP, pt)) {
        return;
    }
    if (tid < kRadix) {
        s_hist[tid] = 0;
    }
    __syncthreads();
    if (pt.first < pt.end) {
        // Every load issued before the first add, so the adds wait on them
        // once: a lane past the end reads the last key (in bounds, unbranched)
        // and adds nothing.
        K key[kItems];
#pragma unroll
        for (uint32_t j = 0; j < kItems; ++j) {
            key[j] = kin[min(pt.first + j * kT + tid, pt.end - 1u)];
        }
#pragma unroll
        for (uint32_t j = 0; j < kItems; ++j) {
            const uint32_t at = pt.first + j * kT + tid;
            hist_add(s_hist, digit(key[j], 0u, pt.base), at < pt.end && at - pt.first < kP);
        }
    }
    __syncthreads();
    if (tid < radix) {
        counts[(size_t)pt.cell + (size_t)tid * pt.stride] = s_hist[tid];
    }
}

template <typename K, typename D, uint32_t kT, typename L>
__global__ void __launch_bounds__(kT) downsweep_kernel(const K* __restrict__ kin, const uint32_t* __restrict__ vin,
                                                       K* __restrict__ kout, uint32_t* __restrict__ vout, L layout,
                                                       D digit, uint32_t width,
                                                       const uint32_t* __restrict__ offsets,
                                                       uint16_t* __restrict__ mout) {
    static_assert(kT >= kRadix && kT % 64 == 0, "a thread a digit");
    constexpr uint32_t kW = kT / 64;
    constexpr uint32_t kP = kT * kItems;
    __shared__ K s_keys[kP];
    __shared__ uint32_t s_vals[kP];
    __shared__ uint32_t s_wave_hist[kW * kRadix];
    __shared__ uint32_t s_block_start[kRadix];
    __shared__ uint32_t s_digit_base[kRadix];
    __shared__ uint32_t s_tmp[kW];

    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & 63u;
    const uint32_t wave = tid >> 6;
    Part pt;
    if (!layout.part(banded_part(layout, kP), kP, pt) || pt.first >= pt.end) {
        return;
    }
    for (uint32_t i = tid; i < kW * kRadix; i += kT) {
        s_wave_hist[i] = 0;
    }
    __syncthreads();
    const uint32_t valid = pt.end - pt.first < kP ? pt.end - pt.first : kP;
    const uint32_t wave_first = wave * (64 * kItems);  // relative to the partition
    K key[kItems];
    uint32_t val[kItems];
    uint32_t rank[kItems];
#pragma unroll
    for (uint32_t j = 0; j < kItems; ++j) {
        const uint32_t at = wave_first + j * 64 + lane;
        const bool ok = at < valid;
        key[j] = ok ? kin[pt.first + at] : 0;
        val[j] = ok ? vin[pt.first + at] : 0;
    }
    const uint64_t lt_mask = (1ull << lane) - 1ull;
    const uint32_t dmask = (1u << width) - 1u;
    uint32_t* my_hist = s_wave_hist + wave * kRadix;
#pragma unroll
    for (uint32_t j = 0; j < kItems; ++j) {
        const bool ok = wave_first + j * 64 + lane < valid;
        const uint32_t d = digit(key[j], val[j], pt.base);
        const uint64_t m = match_digit(d, width, ok);
        const uint32_t before = ok ? my_hist[d] : 0u;
        __builtin_amdgcn_wave_barrier();
        rank[j] = before + (uint32_t)__popcll(m & lt_mask);
        const bool leader = ok && (m & lt_mask) == 0;
        if (leader) {
            my_hist[d] = before + (uint32_t)__popcll(m);
        }
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    // A thread a digit (threads past the radix scan zeros).
    const uint32_t d = tid;
    uint32_t count = 0;
    if (d < kRadix) {
#pragma unroll
        for (uint32_t w = 0; w < kW; ++w) {
            const uint32_t c = s_wave_hist[w * kRadix + d];
            s_wave_hist[w * kRadix + d] = count;
            count += c;
        }
    }
    uint32_t unused;
    const uint32_t start = block_exclusive_scan<kW>(count, s_tmp, unused);
    if (d < kRadix) {
        s_block_start[d] = start;
        s_digit_base[d] = d <= dmask ? offsets[(size_t)pt.cell + (size_t)d * pt.stride] : 0u;
    }
    __syncthreads();
#pragma unroll
    for (uint32_t j = 0; j < kItems; ++j) {
        if (wave_first + j * 64 + lane < valid) {
            const uint32_t dj = digit(key[j], val[j], pt.base);
            const uint32_t pos = s_block_start[dj] + s_wave_hist[wave * kRadix + dj] + rank[j];
            s_keys[pos] = key[j];
            s_vals[pos] = val[j];
        }
    }
    __syncthreads();
    for (uint32_t i = tid; i < valid; i += kT) {
        const K k = s_keys[i];
        const uint32_t v = s_vals[i];
        const uint32_t dk = digit(k, v, pt.base);
        const uint32_t dst = s_digit_base[dk] + (i - s_block_start[dk]);
        // mout: a strip sort's last pass keeps only each key's upper half
        // (its tile mask); the strips' runs come from the counts instead.
        if (mout) {
            mout[dst] = (uint16_t)((uint32_t)k >> 16);
        } else {
            kout[dst] = k;
        }
        vout[dst] = v;
    }
}

// Clears what this frame's passes use: the tickets whole, and each pass's
// look-back words for this frame's partitions only (the histogram is written
// whole by hist_reduce_kernel). Layout, as SortScratch:
// lookback[passes][partitions(cap)][256], counters[passes].
__global__ void clear_kernel(uint32_t* __restrict__ lookback, uint32_t* __restrict__ counters, uint32_t passes,
                             const uint64_t* __restrict__ n_dev, uint32_t cap) {
    const uint32_t n = count_of(n_dev, cap);
    const uint64_t parts_cap = div_up(cap ? cap : 1, kPart);
    const uint64_t per_pass = (uint64_t)div_up(n ? n : 1, kPart) * kRadix;
    const uint64_t look = (uint64_t)passes * per_pass;
    const uint64_t total = look + passes;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (uint64_t)gridDim.x * blockDim.x) {
        if (i < look) {
            const uint64_t p = i / per_pass;
            lookback[p * parts_cap * kRadix + (i - p * per_pass)] = 0;
        } else {
            counters[i - look] = 0;
        }
    }
}

template <typename K>
__global__ void copy_back_kernel(K* __restrict__ keys, uint32_t* __restrict__ values,
                                 const K* __restrict__ kin, const uint32_t* __restrict__ vin,
                                 const uint64_t* __restrict__ n_dev, uint32_t cap) {
    const uint32_t n = count_of(n_dev, cap);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        keys[i] = kin[i];
        values[i] = vin[i];
    }
}

// The reduce-scan-scatter passes rank 8192 32-bit keys a workgroup of 512
// (74 KB of LDS, two workgroups a CU): twice the onesweep partition, so each
// digit's run out of a partition is twice as long and the digit-count scan
// half as long. 64-bit keys keep 4096 (their keys alone take 64 KB).
template <typename K>
constexpr uint32_t rs_threads() {
    return sizeof(K) == 4 ? 512u : 256u;
}

template <typename K>
hipError_t sort_impl(K* keys, uint32_t* values, const uint64_t* n_dev, uint32_t cap, uint32_t lo_bit,
                     uint32_t key_bits, const SortScratch& sc, hipStream_t s, bool* in_alt) {
    if (in_alt) {
        *in_alt = false;
    }
    if (cap < 2 || key_bits <= lo_bit) {
        return hipSuccess;
    }
    if (key_bits > 8 * sizeof(K) || cap > kValueMask) {
        return hipErrorInvalidValue;
    }
    const uint32_t passes = div_up(key_bits - lo_bit, 8);
    const uint32_t parts = sort_partitions(cap);
    const uint32_t spread = parts < 2048 ? parts : 2048;
    if (cap >= kReduceScanMin) {
        constexpr uint32_t kT = rs_threads<K>();
        const uint32_t rparts = div_up(cap, kT * kItems);
        K* kin = keys;
        uint32_t* vin = values;
        K* kout = static_cast<K*>(sc.keys_alt);
        uint32_t* vout = sc.values_alt;
        for (uint32_t p = 0; p < passes; ++p) {
            uint32_t shift, width;
            pass_digit(p, lo_bit, key_bits - lo_bit, passes, shift, width);
            const KeyDigit digit{shift, (1u << width) - 1u};
            const PlainLayout layout{n_dev, cap, rparts};
            hipLaunchKernelGGL((upsweep_kernel<K, KeyDigit, kT, PlainLayout>), dim3(rparts), dim3(kT), 0, s, kin, vin,
                               layout, digit, width, sc.counts);
            if (hipError_t e = launch_scan(sc.counts, sc.offsets, sc.scan_total, (1u << width) * rparts,
                                           sc.scan_status, s);
                e != hipSuccess) {
                return e;
            }
            hipLaunchKernelGGL((downsweep_kernel<K, KeyDigit, kT, PlainLayout>), dim3(rparts), dim3(kT), 0, s, kin,
                               vin, kout, vout, layout, digit, width, sc.offsets, nullptr);
            K* tk = kin; kin = kout; kout = tk;
            uint32_t* tv = vin; vin = vout; vout = tv;
        }
        if (kin != keys) {
            if (in_alt) {
                *in_alt = true;
            } else {
                hipLaunchKernelGGL(copy_back_kernel<K>, dim3(spread), dim3(256), 0, s, keys, values, kin, vin, n_dev,
                                   cap);
            }
        }
        return hipGetLastError();
    }
    const uint32_t hist_blocks = sort_hist_blocks(cap);
    hipLaunchKernelGGL(clear_kernel, dim3(spread), dim3(256), 0, s, sc.lookback, sc.counters, passes, n_dev, cap);
    hipLaunchKernelGGL(histogram_kernel<K>, dim3(hist_blocks), dim3(kThreads), 0, s, keys, n_dev, cap, lo_bit,
                       key_bits - lo_bit, passes, sc.partials);
    hipLaunchKernelGGL(hist_reduce_kernel, dim3(passes * kRadix), dim3(kThreads), 0, s, sc.partials, hist_blocks,
                       sc.hist);
    K* kin = keys;
    uint32_t* vin = values;
    K* kout = static_cast<K*>(sc.keys_alt);
    uint32_t* vout = sc.values_alt;
    for (uint32_t p = 0; p < passes; ++p) {
        uint32_t shift, width;
        pass_digit(p, lo_bit, key_bits - lo_bit, passes, shift, width);
        hipLaunchKernelGGL(onesweep_kernel<K>, dim3(parts), dim3(kThreads), 0, s, kin, vin, kout, vout, n_dev, cap,
                           shift, width, sc.hist + p * kRadix, sc.lookback + (size_t)p * parts * kRadix,
                           sc.counters + p);
        K* tk = kin; kin = kout; kout = tk;
        uint32_t* tv = vin; vin = vout; vout = tv;
    }
    if (kin != keys) {
        if (in_alt) {
            *in_alt = true;
        } else {
            hipLaunchKernelGGL(copy_back_kernel<K>, dim3(spread), dim3(256), 0, s, keys, values, kin, vin, n_dev, cap);
        }
    }
    return hipGetLastError();
}

}  // namespace

uint32_t sort_partitions(uint32_t n) { return div_up(n ? n : 1, kPart); }

uint32_t sort_hist_blocks(uint32_t n) {
    const uint32_t parts = sort_partitions(n);
    return parts < kHistBlocks ? parts : kHistBlocks;
}

// Sorts the first *n_dev (at most cap) of (keys, values) stably by bits
// [lo_bit, key_bits) of the key, the count read on the device. Digits below
// lo_bit are left as the input orders them: a caller whose input is already
// in order of those bits sorts only the rest. The result is back in keys and
// values; sc.keys_alt holds keys of the same width.
hipError_t launch_sort(uint64_t* keys, uint32_t* values, const uint64_t* n_dev, uint32_t cap, uint32_t lo_bit,
                       uint32_t key_bits, const SortScratch& sc, hipStream_t s, bool* in_alt) {
    return sort_impl(keys, values, n_dev, cap, lo_bit, key_bits, sc, s, in_alt);
}

hipError_t launch_sort(uint32_t* keys, uint32_t* values, const uint64_t* n_dev, uint32_t cap, uint32_t lo_bit,
                       uint32_t key_bits, const SortScratch& sc, hipStream_t s, bool* in_alt) {
    return sort_impl(keys, values, n_dev, cap, lo_bit, key_bits, sc, s, in_alt);
}

template <typename D>
hipError_t segment_passes(uint32_t* keys, uint32_t* values, const SegmentArgs& g, uint32_t parts, uint32_t bits,
                          const SortScratch& sc, hipStream_t s, bool& in_alt, uint16_t* mask_out) {
    constexpr uint32_t kT = rs_threads<uint32_t>();
    const uint32_t passes = div_up(bits, 8);
    for (uint32_t p = 0; p < passes; ++p) {
        uint32_t shift, width;
        pass_digit(p, 0, bits, passes, shift, width);
        uint32_t* kin = in_alt ? static_cast<uint32_t*>(sc.keys_alt) : keys;
        uint32_t* vin = in_alt ? sc.values_alt : values;
        uint32_t* kout = in_alt ? keys : static_cast<uint32_t*>(sc.keys_alt);
        uint32_t* vout = in_alt ? values : sc.values_alt;
        const D digit{shift, (1u << width) - 1u};
        const SegLayout layout{g.part_seg, g.seg_start, g.seg_len, g.seg_part, g.seg_base, g.parts, 1u << width};
        hipLaunchKernelGGL((upsweep_kernel<uint32_t, D, kT, SegLayout>), dim3(parts), dim3(kT), 0, s, kin, vin, layout,
                           digit, width, sc.counts);
        if (hipError_t e = launch_scan(sc.counts, sc.offsets, sc.scan_total, (1u << width) * parts, sc.scan_status, s);
            e != hipSuccess) {
            return e;
        }
        hipLaunchKernelGGL((downsweep_kernel<uint32_t, D, kT, SegLayout>), dim3(parts), dim3(kT), 0, s, kin, vin,
                           kout, vout, layout, digit, width, sc.offsets, p + 1 == passes ? mask_out : nullptr);
        in_alt = !in_alt;
    }
    return hipGetLastError();
}

hipError_t launch_segment_sort(uint32_t* keys, uint32_t* values, const SegmentArgs& g, uint32_t parts_cap,
                               uint32_t bits, const SortScratch& sc, hipStream_t s, bool& in_alt, bool strips,
                               uint16_t* mask_out) {
    in_alt = false;
    if (parts_cap == 0 || bits == 0) {
        return hipSuccess;
    }
    if (bits > (strips ? 16u : 32u)) {
        return hipErrorInvalidValue;
    }
    static_assert(rs_threads<uint32_t>() * kItems == kSegmentPart, "a segment's partition is the pass's");
    return strips ? segment_passes<StripDigit>(keys, values, g, parts_cap, bits, sc, s, in_alt, mask_out)
                  : segment_passes<KeyDigit>(keys, values, g, parts_cap, bits, sc, s, in_alt, nullptr);
}

namespace {

// Strip s of view v is digit d = s - (v's first strip) of v's run: its entries
// start at the scanned count of (v, d, v's first partition) and end where the
// next digit's (or the next view's) start, the total past the last cell.
__global__ void __launch_bounds__(256) strip_ranges_seg_kernel(const uint32_t* __restrict__ offsets, SegmentArgs g,
                                                               uint32_t nviews, const uint32_t* __restrict__ strip_bases,
                                                               uint32_t strips, uint32_t radix,
                                                               const uint64_t* __restrict__ total_dev,
                                                               uint32_t* __restrict__ ranges) {
    const uint32_t st = blockIdx.x * blockDim.x + threadIdx.x;
    if (st >= strips) {
        return;
    }
    uint32_t lo = 0, hi = nviews;
    while (hi - lo > 1u) {
        const uint32_t mid = (lo + hi) >> 1;
        if (strip_bases[mid] <= st) {
            lo = mid;
        } else {
            hi = mid;
        }
    }
    const uint32_t v = lo, d = st - strip_bases[v];
    const uint32_t cells = radix * *g.parts;
    const uint32_t stride = div_up(g.seg_len[v], kSegmentPart);
    const uint32_t total = (uint32_t)*total_dev;
    uint32_t start = total, end = total;
    if (d < radix) {
        const uint32_t cell = radix * g.seg_part[v] + d * stride;
        start = cell < cells ? offsets[cell] : total;
        end = cell + stride < cells ? offsets[cell + stride] : total;
    }
    ranges[2u * st] = start < end ? start : 0u;
    ranges[2u * st + 1u] = start < end ? end : 0u;
}

}  // namespace

hipError_t launch_strip_ranges_seg(const uint32_t* offsets, const SegmentArgs& g, uint32_t nviews,
                                   const uint32_t* strip_bases, uint32_t strips, uint32_t radix,
                                   const uint64_t* total_dev, uint32_t* ranges, hipStream_t s) {
    if (strips == 0) {
        return hipSuccess;
    }
    hipLaunchKernelGGL(strip_ranges_seg_kernel, dim3(div_up(strips, 256)), dim3(256), 0, s, offsets, g, nviews,
                       strip_bases, strips, radix, total_dev, ranges);
    return hipGetLastError();
}


}  // namespace athx
