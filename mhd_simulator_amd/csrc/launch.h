// One kernel body, two launches.  A kernel is a struct K with
//   struct Args { ... };                        (trivially copyable: pointers and scalars)
//   static constexpr int MINB;                  (resident blocks per CU to target; 0 = none)
//   static constexpr int THREADS;               (optional: threads per block, else MHDP_BLOCK)
//   static __device__ void run(const Args&, const VGrid&);
// go<K>(stream, blocks, args) launches it for one problem (`blocks` blocks).  While a host thread
// records (Recorder::active), go<K> appends the launch instead; the batch driver then issues
// the same kernel for many problems as ONE launch (k_many: problem p owns a contiguous run of
// blocks with the numbering it would have had alone, so its results are bit-identical).
#pragma once
#include <cstdint>
#include <cstring>
#include <vector>

#include <type_traits>

#include "mhdp_device.h"

// Threads per block of kernel K: K::THREADS when it declares one, else MHDP_BLOCK.
template <class K, class = void>
struct KThreads : std::integral_constant<int, MHDP_BLOCK> {};
template <class K>
struct KThreads<K, std::void_t<decltype(K::THREADS)>> : std::integral_constant<int, K::THREADS> {};

template <class K>
__global__ void __launch_bounds__(KThreads<K>::value) k_one0(typename K::Args a) {
    K::run(a, vgrid());
}
template <class K>
__global__ void __launch_bounds__(KThreads<K>::value, (K::MINB > 0 ? K::MINB : 1)) k_one(typename K::Args a) {
    K::run(a, vgrid());
}
// starts[0..count]: first block of each problem (a multiple of 8) and the end; counts[p]: its
// block count (the blocks between are idle padding); a: the problems' arguments.  Hardware hands
// workgroups to the 8 XCDs round-robin, so with 8-aligned starts each problem sees the same
// XCD pattern as alone and gets the same logical-block remap (contiguous ranges per XCD).
template <class K>
__device__ __forceinline__ void run_many(const typename K::Args* __restrict__ a, const uint32_t* __restrict__ starts,
                                         uint32_t count) {
    const uint32_t b = blockIdx.x;
    uint32_t lo = 0, hi = count;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (starts[mid] <= b) lo = mid;
        else hi = mid;
    }
    const uint32_t* counts = starts + count + 1;
    const uint32_t j = b - starts[lo], n = counts[lo];
    if (j >= n) return;
    const uint32_t per = n / MHDP_XCDS;
    const uint32_t lb = (per * MHDP_XCDS == n) ? (j % MHDP_XCDS) * per + j / MHDP_XCDS : j;
    const typename K::Args A = a[lo];
    K::run(A, VGrid{lb, n});
}
template <class K>
__global__ void __launch_bounds__(KThreads<K>::value) k_many0(const typename K::Args* a, const uint32_t* starts,
                                                              uint32_t count) {
    run_many<K>(a, starts, count);
}
template <class K>
__global__ void __launch_bounds__(KThreads<K>::value, (K::MINB > 0 ? K::MINB : 1)) k_many(const typename K::Args* a,
                                                                                        const uint32_t* starts,
                                                                                        uint32_t count) {
    run_many<K>(a, starts, count);
}

// One recorded launch of one problem.
struct Recorded {
    const void* kind;  // identifies K (the address of its batched kernel)
    void (*issue)(hipStream_t, const void* args_dev, const uint32_t* starts_dev, uint32_t count, uint32_t blocks);
    std::vector<unsigned char> args;
    uint32_t blocks;
};
struct Recorder {
    std::vector<Recorded> ops;
    // the recorder of the calling host thread, if it is recording
    static Recorder*& active() {
        static thread_local Recorder* r = nullptr;
        return r;
    }
};

// Alignment point of recorded sequences: merging does not move any problem past a mark until
// every problem has reached it (keeps problems of different hierarchy depth in step).
inline void record_mark() {
    if (Recorder* r = Recorder::active()) r->ops.push_back(Recorded{nullptr, nullptr, {}, 0u});
}

template <class K>
void issue_many(hipStream_t s, const void* args_dev, const uint32_t* starts_dev, uint32_t count, uint32_t blocks) {
    const dim3 threads(KThreads<K>::value);
    if constexpr (K::MINB > 0)
        hipLaunchKernelGGL(k_many<K>, dim3(blocks), threads, 0, s, (const typename K::Args*)args_dev, starts_dev,
                           count);
    else
        hipLaunchKernelGGL(k_many0<K>, dim3(blocks), threads, 0, s, (const typename K::Args*)args_dev, starts_dev,
                           count);
}

template <class K>
void go(hipStream_t s, uint32_t blocks, const typename K::Args& a) {
    if (Recorder* r = Recorder::active()) {
        Recorded op;
        if constexpr (K::MINB > 0) op.kind = (const void*)&k_many<K>;
        else op.kind = (const void*)&k_many0<K>;
        op.issue = &issue_many<K>;
        op.args.resize(sizeof a);
        std::memcpy(op.args.data(), &a, sizeof a);
        op.blocks = blocks;
        r->ops.push_back(std::move(op));
        return;
    }
    const dim3 threads(KThreads<K>::value);
    if constexpr (K::MINB > 0) hipLaunchKernelGGL(k_one<K>, dim3(blocks), threads, 0, s, a);
    else hipLaunchKernelGGL(k_one0<K>, dim3(blocks), threads, 0, s, a);
}
