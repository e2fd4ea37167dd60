// anchors() (src/fluid/mod.rs:1447-1493) on the device: connected components of the fluid
// cells over face neighbours (periodic axes wrap), each cell encoded with its component's root,
// the lowest flat index of the component: 2|root<<2 for the root (the anchor), 1|root<<2 for
// the other members, 0 for solid cells.
//
// Lock-free union-find (labels only ever point to smaller indices):
//   init     lab[i] = first cell of i's run of fluid cells along x (the run is one set)
//   hook     fluid cells unite with their fluid lower neighbours along y and z (only where a
//            run starts) and across the periodic x seam; a union links the larger of the two
//            roots to the smaller with a compare-and-swap, so the root that survives in a
//            component is its minimum index - the reference's BFS root
//   compress lab[i] = root(i) (read-only walks: every store is a cell's final root)
//   encode   mask[i] from lab[i]; roots, fluid cells and the first anchor are counted.
// Label reads go through device-coherent atomic loads (not the per-CU L1), and every loop has a
// bound; a bound reached sets a flag and the host reports an error instead of a wrong mask.
#include <algorithm>

#include "mhdp_kernels.h"

namespace {

constexpr uint32_t kNone = 0xffffffffu;
constexpr uint32_t kFindBound = 1u << 24;
constexpr uint32_t kUniteBound = 1u << 16;

__device__ __forceinline__ uint32_t lab_load(const uint32_t* lab, uint32_t i) {
    return __hip_atomic_load(lab + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Root of x.  HALVE: path halving (each store points a cell to an ancestor: still a smaller
// index) - only while hooking: in the compress pass a halving store could overwrite the root
// another thread has just written for that cell with a mere ancestor.
template <bool HALVE>
__device__ __forceinline__ uint32_t find_root(uint32_t* lab, uint32_t x, uint32_t* err) {
    for (uint32_t s = 0; s < kFindBound; s++) {
        const uint32_t p = lab_load(lab, x);
        if (p == x) return x;
        const uint32_t gp = lab_load(lab, p);
        if (HALVE && gp != p) __hip_atomic_store(lab + x, gp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        x = gp;
    }
    atomicOr(err, 1u);
    return x;
}
__device__ __forceinline__ void unite(uint32_t* lab, uint32_t a, uint32_t b, uint32_t* err) {
    for (uint32_t s = 0; s < kUniteBound; s++) {
        uint32_t ra = find_root<true>(lab, a, err), rb = find_root<true>(lab, b, err);
        if (ra == rb) return;
        if (ra < rb) {
            const uint32_t t = ra;
            ra = rb;
            rb = t;
        }
        // link the larger root under the smaller one if it is still a root
        if (atomicCAS(lab + ra, ra, rb) == ra) return;
    }
    atomicOr(err, 2u);
}

}  // namespace

// init: every fluid cell starts labelled with the first cell of its run of fluid cells along x,
// so cells of a run are already united.  One wave per row, 64 cells at a time: the position of
// the last solid cell so far is a running maximum (wave scan plus the carry of earlier chunks).
__global__ void __launch_bounds__(MHDP_BLOCK) k_cc_rows(Grid3 g, const uint8_t* __restrict__ fluid,
                                                        uint32_t* __restrict__ lab) {
    const uint32_t rows = (uint32_t)g.ny * (uint32_t)g.nz;
    const uint32_t nx = (uint32_t)g.nx;
    const int lane = threadIdx.x & 63;
    const uint32_t waves = gridDim.x * (MHDP_BLOCK / 64);
    for (uint32_t row = blockIdx.x * (MHDP_BLOCK / 64) + (threadIdx.x >> 6); row < rows; row += waves) {
        const uint32_t base = row * nx;
        int carry = -1;  // x of the last solid cell before this chunk (-1: none)
        for (uint32_t x0 = 0; x0 < nx; x0 += 64) {
            const uint32_t x = x0 + (uint32_t)lane;
            const bool in = x < nx;
            const bool f = in && fluid[base + x] != 0;
            int last = (in && !f) ? (int)x : -1;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const int v = __shfl_up(last, o, 64);
                if (lane >= o) last = max(last, v);
            }
            last = max(last, carry);
            if (in) lab[base + x] = f ? base + (uint32_t)(last + 1) : kNone;
            carry = __shfl(last, 63, 64);
        }
    }
}

// hook: unions across y and z (and the periodic x wrap).  A y/z pair (i, j) whose x-predecessors
// (i-1, j-1) are both fluid joins the same two runs as that pair, so only pairs at a run start
// (of either run) unite; each pair of touching runs is still united at least once.
__global__ void __launch_bounds__(MHDP_BLOCK) k_cc_hook(Grid3 g, const uint8_t* __restrict__ fluid, uint32_t* lab,
                                                        uint32_t* err) {
    uint32_t lo, hi;
    block_range(g.n, lo, hi);
    const uint32_t sy = (uint32_t)g.nx, sz = (uint32_t)g.nx * (uint32_t)g.ny;
    for (uint32_t i = lo + threadIdx.x; i < hi; i += MHDP_BLOCK) {
        if (!fluid[i]) continue;
        int x, y, z;
        decode(g, i, x, y, z);
        const bool first = x == 0 || !fluid[i - 1];
        if (x == 0 && (g.per & 1u) && g.nx > 1) {
            const uint32_t w = i + (uint32_t)(g.nx - 1);
            if (fluid[w]) unite(lab, i, w, err);
        }
        uint32_t nb[2];
        bool ok[2];
        ok[0] = y > 0 || ((g.per & 2u) && g.ny > 1);
        nb[0] = y > 0 ? i - sy : i + (uint32_t)(g.ny - 1) * sy;
        ok[1] = z > 0 || ((g.per & 4u) && g.nz > 1);
        nb[1] = z > 0 ? i - sz : i + (uint32_t)(g.nz - 1) * sz;
#pragma unroll
        for (int a = 0; a < 2; a++) {
            if (!ok[a] || !fluid[nb[a]]) continue;
            if (!first && fluid[nb[a] - 1]) continue;  // x > 0 here, so nb - 1 is in nb's row
            unite(lab, i, nb[a], err);
        }
    }
}

__global__ void __launch_bounds__(MHDP_BLOCK) k_cc_compress(uint32_t n, uint32_t* lab, uint32_t* err) {
    uint32_t lo, hi;
    block_range(n, lo, hi);
    for (uint32_t i = lo + threadIdx.x; i < hi; i += MHDP_BLOCK) {
        if (lab_load(lab, i) == kNone) continue;
        const uint32_t r = find_root<false>(lab, i, err);
        __hip_atomic_store(lab + i, r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// counts[0] components, counts[1] fluid cells, counts[2] first anchor (lowest root)
__global__ void __launch_bounds__(MHDP_BLOCK) k_cc_encode(uint32_t n, uint32_t* __restrict__ lab,
                                                          uint32_t* __restrict__ counts) {
    uint32_t lo, hi;
    block_range(n, lo, hi);
    uint32_t roots = 0, cells = 0, first = kNone;
    for (uint32_t i = lo + threadIdx.x; i < hi; i += MHDP_BLOCK) {
        const uint32_t r = lab[i];
        if (r == kNone) {
            lab[i] = 0u;
            continue;
        }
        cells++;
        if (r == i) {
            roots++;
            first = min(first, i);
        }
        lab[i] = (r == i ? 2u : 1u) | (r << 2);
    }
    // wave sums, then one atomic per wave
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        roots += __shfl_down(roots, o, 64);
        cells += __shfl_down(cells, o, 64);
        first = min(first, __shfl_down(first, o, 64));
    }
    if ((threadIdx.x & 63) == 0) {
        if (roots) atomicAdd(counts + 0, roots);
        if (cells) atomicAdd(counts + 1, cells);
        if (first != kNone) atomicMin(counts + 2, first);
    }
}

__global__ void k_cc_counts_init(uint32_t* counts) {
    // (0 components, 0 fluid cells, no anchor, no error flags)
    if (threadIdx.x < 4) counts[threadIdx.x] = threadIdx.x == 2 ? kNone : 0u;
}

// fluid (u8, device) -> encoded mask (u32, device); counts (4 x u32, device) as above with
// counts[3] the error flags (0 = labelled).
void launch_components(hipStream_t s, const Grid3& g, const uint8_t* fluid, uint32_t* mask, uint32_t* counts) {
    const dim3 B(MHDP_BLOCK), G(grid_blocks(g.n));
    hipLaunchKernelGGL(k_cc_counts_init, dim3(1), dim3(64), 0, s, counts);
    const uint32_t rows = (uint32_t)g.ny * (uint32_t)g.nz;
    const uint32_t row_blocks = std::min<uint32_t>((rows + MHDP_BLOCK / 64 - 1) / (MHDP_BLOCK / 64), 4096u);
    hipLaunchKernelGGL(k_cc_rows, dim3(row_blocks), B, 0, s, g, fluid, mask);
    hipLaunchKernelGGL(k_cc_hook, G, B, 0, s, g, fluid, mask, counts + 3);
    hipLaunchKernelGGL(k_cc_compress, G, B, 0, s, g.n, mask, counts + 3);
    hipLaunchKernelGGL(k_cc_encode, G, B, 0, s, g.n, mask, counts);
}
