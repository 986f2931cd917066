// scan.hip -- exclusive prefix sum of uint32 in one pass: each workgroup
// takes a ticket (so every partition it waits on has already started and
// will finish), reduces its 4096 elements, publishes the aggregate, and one
// wavefront looks back 64 partitions at a time for the inclusive prefix of
// its predecessors (decoupled look-back, Merrill & Garland 2016).
//
// Semantics: gpu::PrefixSum::apply (modules/gpu/include/athenea/gpu/algo/
// PrefixSum.h:43-44; modules/gpu/src/algo/PrefixSum.cpp:58-88): out[i] =
// in[0] + ... + in[i-1], and the total of all. The reference does it in
// three dispatches (prefix_chunk_totals, prefix_chunk_starts, prefix_local).
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

// status[0] is the ticket counter; status[1 + p] partition p's word.
__global__ void __launch_bounds__(kThreads) scan_kernel(const uint32_t* __restrict__ in, uint32_t* __restrict__ out,
                                                        uint64_t* __restrict__ total, uint32_t n,
                                                        uint64_t* __restrict__ status, bool aligned) {
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
    const uint32_t part = s_part;
    const uint32_t parts = div_up(n, kTile);
    const uint64_t first = (uint64_t)part * kTile + (uint64_t)tid * kItems;

    uint32_t v[kItems];
    if (aligned && first + kItems <= n) {
        const uint4* src = reinterpret_cast<const uint4*>(in + first);
#pragma unroll
        for (uint32_t k = 0; k < kItems / 4; ++k) {
            const uint4 q = src[k];
            v[4 * k + 0] = q.x; v[4 * k + 1] = q.y; v[4 * k + 2] = q.z; v[4 * k + 3] = q.w;
        }
    } else {
#pragma unroll
        for (uint32_t k = 0; k < kItems; ++k) {
            v[k] = first + k < n ? in[first + k] : 0u;
        }
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
    uint32_t o[kItems];
#pragma unroll
    for (uint32_t k = 0; k < kItems; ++k) {
        o[k] = (uint32_t)run;
        run += v[k];
    }
    if (aligned && first + kItems <= n) {
        uint4* dst = reinterpret_cast<uint4*>(out + first);
#pragma unroll
        for (uint32_t k = 0; k < kItems / 4; ++k) {
            dst[k] = make_uint4(o[4 * k + 0], o[4 * k + 1], o[4 * k + 2], o[4 * k + 3]);
        }
    } else {
#pragma unroll
        for (uint32_t k = 0; k < kItems; ++k) {
            if (first + k < n) {
                out[first + k] = o[k];
            }
        }
    }
}

__global__ void zero_u64_kernel(uint64_t* p, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        p[i] = 0;
    }
}

}  // namespace

// u32 words of status a scan of n elements needs (u64 words, counted in u32).
size_t scan_status_words(uint32_t n) { return 2 * (1 + (size_t)div_up(n ? n : 1, kTile)); }

hipError_t launch_scan(const uint32_t* in, uint32_t* out, uint64_t* total, uint32_t n, uint32_t* status_words,
                       hipStream_t s) {
    uint64_t* status = reinterpret_cast<uint64_t*>(status_words);
    if (n == 0) {
        hipLaunchKernelGGL(zero_u64_kernel, dim3(1), dim3(64), 0, s, total, 1);
        return hipGetLastError();
    }
    const uint32_t parts = div_up(n, kTile);
    const uint64_t words = 1 + (uint64_t)parts;
    hipLaunchKernelGGL(zero_u64_kernel, dim3((uint32_t)((words + 255) / 256 < 1024 ? (words + 255) / 256 : 1024)),
                       dim3(256), 0, s, status, words);
    const bool aligned = ((reinterpret_cast<uintptr_t>(in) | reinterpret_cast<uintptr_t>(out)) & 15u) == 0;
    hipLaunchKernelGGL(scan_kernel, dim3(parts), dim3(kThreads), 0, s, in, out, total, n, status, aligned);
    return hipGetLastError();
}

}  // namespace athx
