// kernels.hip -- gfx950 kernels of the short-range MD timestep.
//
// Positions, velocities and the integrator are FP64; neighbour-list inclusion and the pair
// cutoff test are exact FP64 comparisons; pair and bond arithmetic is FP32 with FP32
// per-particle sums. Wavefronts are 64 lanes; particle-parallel gathers use an XCD-aware block
// remap so spatially close (Hilbert-sorted) particles share an XCD's L2.
#include <type_traits>

#include "evaluators.h"
#include "kernels.h"

namespace hm
{

// ======================================================================== cell adjacency
// CellList::initializeCellAdj (hoomd/CellList.cc:346-402): the 27 wrapped neighbours of each
// cell, sorted, duplicates removed (fewer than 3 cells along an axis).
__global__ void cell_adj_kernel(unsigned int dx, unsigned int dy, unsigned int dz,
                                unsigned int n_adj, unsigned int* adj)
{
    unsigned int c = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned int ncell = dx * dy * dz;
    if (c >= ncell)
        return;
    int i = c % dx, j = (c / dx) % dy, k = c / (dx * dy);
    unsigned int tmp[27];
    int n = 0;
    for (int nk = k - 1; nk <= k + 1; nk++)
        for (int nj = j - 1; nj <= j + 1; nj++)
            for (int ni = i - 1; ni <= i + 1; ni++)
            {
                int wi = ((ni % (int)dx) + (int)dx) % (int)dx;
                int wj = ((nj % (int)dy) + (int)dy) % (int)dy;
                int wk = ((nk % (int)dz) + (int)dz) % (int)dz;
                tmp[n++] = (unsigned int)((wk * (int)dy + wj) * (int)dx + wi);
            }
    for (int a = 1; a < 27; a++) // insertion sort
    {
        unsigned int v = tmp[a];
        int b = a - 1;
        while (b >= 0 && tmp[b] > v)
        {
            tmp[b + 1] = tmp[b];
            b--;
        }
        tmp[b + 1] = v;
    }
    unsigned int m = 0;
    for (int a = 0; a < 27; a++)
        if (a == 0 || tmp[a] != tmp[a - 1])
        {
            if (m < n_adj)
                adj[(size_t)c * n_adj + m] = tmp[a];
            m++;
        }
}

hipError_t launch_cell_adj(const unsigned int dim[3], unsigned int* adj, hipStream_t s)
{
    unsigned int ncell = dim[0] * dim[1] * dim[2];
    unsigned int n_adj = 1;
    for (int d = 0; d < 3; d++)
        n_adj *= dim[d] > 3 ? 3 : dim[d];
    hipLaunchKernelGGL(cell_adj_kernel, dim3((ncell + 255) / 256), dim3(256), 0, s, dim[0],
                       dim[1], dim[2], n_adj, adj);
    return hipGetLastError();
}

// ======================================================================== persistent grids
// Kernels that only run on a neighbour-list rebuild are launched every step and exit at once
// when the device-side rebuild flag is not set. They walk their work in grid-stride loops over a
// bounded grid, so an idle step costs a few hundred empty workgroups instead of one per cell.
constexpr unsigned int GATED_BLOCKS = 2048;

__device__ __forceinline__ bool gate_closed(const StepState* gate)
{
    return gate && gate->rebuild_at != gate->timestep + 1;
}

// XCD-aware persistent loop: workgroup b runs on XCD (b % 8); it visits the items of that XCD's
// contiguous eighth of [0, n) with stride gridDim.x / 8.
struct XcdRange
{
    unsigned int begin, end, stride;
};

__device__ __forceinline__ XcdRange xcd_range(unsigned int n)
{
    const unsigned int xcd = blockIdx.x & 7u, r = blockIdx.x >> 3, per = gridDim.x >> 3;
    const unsigned int chunk = (n + 7) / 8;
    XcdRange x;
    x.begin = xcd * chunk + r;
    x.end = min(n, (xcd + 1) * chunk);
    x.stride = per;
    return x;
}

// ======================================================================== cell build
// CellList::computeCellList (hoomd/CellList.cc:404-547) / gpu_compute_cell_list_kernel
// (hoomd/CellListGPU.cu:42-171): FP64 makeFraction, truncation to the cell grid, wrap at dim,
// atomic slot, conditions (overflow, NaN, out of box).
// One particle of the binning pass. Called by every lane of a wavefront in step (the slot
// allocation is a wavefront operation); the active lanes must be the low lanes of the wavefront.
__device__ __forceinline__ void bin_particle(unsigned int n, const double4* __restrict__ pos,
                                             const DevBox& box, unsigned int dx, unsigned int dy,
                                             unsigned int dz, unsigned int cmax,
                                             unsigned int* cell_size, hm_cell_entry* entries,
                                             d3* last, unsigned int* cond, unsigned int* keys,
                                             unsigned int* home)
{
#pragma clang fp contract(off)
    {
        double4 p = pos[n];
        if (last)
            last[n] = d3{p.x, p.y, p.z};
        bool ok = true;
        unsigned int bin = 0;
        if (isnan(p.x) || isnan(p.y) || isnan(p.z))
        {
            atomicMax(&cond[1], n + 1);
            ok = false;
        }
        else
        {
            double fx = (p.x - box.lo[0]) * box.Linv[0];
            double fy = (p.y - box.lo[1]) * box.Linv[1];
            double fz = (p.z - box.lo[2]) * box.Linv[2];
            int ib = (int)(fx * (double)dx);
            int jb = (int)(fy * (double)dy);
            int kb = (int)(fz * (double)dz);
            if (ib == (int)dx)
                ib = 0;
            if (jb == (int)dy)
                jb = 0;
            if (kb == (int)dz)
                kb = 0;
            if (fx < -0.00001 || fx >= 1.00001 || fy < -0.00001 || fy >= 1.00001 || fz < -0.00001
                || fz >= 1.00001 || ib < 0 || ib >= (int)dx || jb < 0 || jb >= (int)dy || kb < 0
                || kb >= (int)dz)
            {
                atomicMax(&cond[2], n + 1);
                ok = false;
            }
            else
                bin = (unsigned int)((kb * (int)dy + jb) * (int)dx + ib);
        }
        // slot allocation aggregated over runs of consecutive lanes in one cell (particles in
        // Hilbert order fill a cell from neighbouring indices): the first lane of a run adds the
        // run's length, and every run's atomic is issued by the one instruction. Active lanes are
        // the low lanes of the wave, so a lane's predecessor is active.
        const unsigned int lane = lane_id();
        const unsigned int prev = __shfl_up(bin, 1, 64);
        const bool head = ok && (lane == 0 || prev != bin || !__shfl_up((int)ok, 1, 64));
        const uint64_t heads = wave_ballot(head);
        // lanes that do not continue a run: heads, lanes not binned, inactive lanes
        const uint64_t stop = ~wave_ballot(ok && !head);
        const uint64_t upto = lane == 63 ? ~0ull : ((1ull << (lane + 1)) - 1ull);
        const unsigned int first = 63u - (unsigned int)__clzll((long long)(heads & upto));
        unsigned int base = 0;
        if (head)
        {
            const uint64_t after = lane == 63 ? 0ull : (stop >> (lane + 1));
            const unsigned int len = after ? (unsigned int)__ffsll((long long)after) : 64u - lane;
            base = atomicAdd(&cell_size[bin], len);
        }
        base = __shfl(base, (int)(ok ? first : lane), 64);
        const unsigned int slot = ok ? base + (lane - first) : 0u;
        const bool placed = ok && slot < cmax;
        if (ok && !placed)
            atomicMax(&cond[0], slot + 1);
        if (keys)
        {
            home[n] = placed ? bin * cmax + slot : NO_SLOT;
            if (placed)
                keys[(size_t)bin * cmax + slot] = n;
        }
        else if (placed)
        {
            hm_cell_entry e;
            e.x = p.x;
            e.y = p.y;
            e.z = p.z;
            e.idx = n;
            e.type = (uint32_t)p.w;
            entries[(size_t)bin * cmax + slot] = e;
        }
    }
}

__global__ __launch_bounds__(256) void cell_build_kernel(unsigned int N, const double4* __restrict__ pos,
                                                         DevBox box, unsigned int dx, unsigned int dy,
                                                         unsigned int dz, unsigned int cmax,
                                                         unsigned int* cell_size,
                                                         hm_cell_entry* entries, d3* last,
                                                         unsigned int* cond, const StepState* gate,
                                                         unsigned int* keys, unsigned int* home)
{
    if (gate_closed(gate))
        return;
    for (unsigned int n = blockIdx.x * blockDim.x + threadIdx.x; n < N;
         n += gridDim.x * blockDim.x)
        bin_particle(n, pos, box, dx, dy, dz, cmax, cell_size, entries, last, cond, keys, home);
}

hipError_t launch_cell_build(unsigned int N, const double4* pos, const DevBox& box,
                             const unsigned int dim[3], unsigned int cmax,
                             unsigned int* cell_size, hm_cell_entry* entries, d3* last_pos,
                             unsigned int* cond3, const StepState* gate, hipStream_t s,
                             unsigned int* keys, unsigned int* home)
{
    if (N == 0)
        return hipSuccess;
    if ((keys == nullptr) != (home == nullptr) || (keys == nullptr && entries == nullptr))
        return hipErrorInvalidValue;
    unsigned int blocks = min((N + 255) / 256, GATED_BLOCKS);
    hipLaunchKernelGGL(cell_build_kernel, dim3(blocks), dim3(256), 0, s, N, pos, box, dim[0],
                       dim[1], dim[2], cmax, cell_size, entries, last_pos, cond3, gate, keys,
                       home);
    return hipGetLastError();
}

__global__ void cell_clear_kernel(unsigned int ncell, unsigned int* cell_size,
                                  const StepState* gate)
{
    if (gate_closed(gate))
        return;
    for (unsigned int c = blockIdx.x * blockDim.x + threadIdx.x; c < ncell;
         c += gridDim.x * blockDim.x)
        cell_size[c] = 0;
}

hipError_t launch_cell_clear(unsigned int ncell, unsigned int* cell_size, const StepState* gate,
                             hipStream_t s)
{
    if (ncell == 0)
        return hipSuccess;
    unsigned int blocks = min((ncell + 255) / 256, 256u);
    hipLaunchKernelGGL(cell_clear_kernel, dim3(blocks), dim3(256), 0, s, ncell, cell_size, gate);
    return hipGetLastError();
}

// ======================================================================== neighbour list
// NeighborListBinned::buildNlist (hoomd/md/NeighborListBinned.cc:36-179) in full storage,
// restructured for CDNA:
//  * cell_sort_kernel orders every cell's slots by particle index once per build, so candidate
//    order (adjacency cell order, then index) equals the CPU list's;
//  * one workgroup per home cell stages the adjacent cells' particles in LDS, one wavefront per
//    home particle sweeps the staged candidates 64 at a time, and ballot + mbcnt compaction
//    writes the hits contiguously into the particle's row (and, for the cell-tiled pair kernel,
//    the hit's 16-bit position in the home cell's candidate sequence).
// Inclusion is dr^2 <= (r_cut + r_buff)^2 decided exactly as the FP64 CPU path decides it. The
// fast kernel tests in FP32 on coordinates relative to the home-cell centre and re-decides with
// the CPU's FP64 formula every candidate whose FP32 distance lies within a guard band of the
// threshold, the band being four times the worst-case FP32 error; rows therefore equal the CPU
// rows entry for entry. Grids with fewer than 4 cells along an axis use the FP64 kernel.
#ifndef HM_NLU
#define HM_NLU 4
#endif
constexpr int NLU = HM_NLU;  // candidate chunks of 64 per sweep iteration (FP32 path)
#ifndef HM_NLCAP
#define HM_NLCAP 768
#endif
constexpr int NLCAP = HM_NLCAP;
constexpr int NLHOME = 256;
// Workgroup shapes of the list build. Large cells (tens of particles): four wavefronts share a
// home cell's candidates. Small cells (a few particles, short cutoffs such as WCA): one
// wavefront per workgroup, a smaller stage and two chunks per sweep iteration, so that more
// cells are in flight per CU. WAVES > 1 asks the compiler for that many wavefronts per SIMD
// (it limits the registers; the build waits mostly on memory, so the extra wavefronts in
// flight outweigh the few spills).
struct NlBig
{
    static constexpr int B = 256, CAP = NLCAP, HOME = NLHOME, U = NLU, WAVES = 1;
};
#ifndef HM_MID_WAVES
#define HM_MID_WAVES 8
#endif
struct NlMid
{
    static constexpr int B = 256, CAP = NLCAP, HOME = 64, U = NLU, WAVES = HM_MID_WAVES;
};
struct NlSmall
{
    static constexpr int B = 64, CAP = 256, HOME = 64, U = 2, WAVES = 1;
};
// at most about 4 particles per cell: a stage of 128 candidates, and a workgroup small enough
// in LDS and registers for 8 wavefronts per SIMD
#ifndef HM_TINY_WAVES
#define HM_TINY_WAVES 8
#endif
struct NlTiny
{
    static constexpr int B = 64, CAP = 128, HOME = 64, U = 2, WAVES = HM_TINY_WAVES;
};

// Position relative to the origin of cell c (packed cx | cy << 8 | cz << 16 | type << 24),
// minimum image, rounded to FP32; w keeps the packed word.
__device__ inline float4 rel_of(double px, double py, double pz, uint32_t c, const DevBox& box,
                                const RelFrame& f)
{
#pragma clang fp contract(off)
    double ox = px - (box.lo[0] + (double)(c & 255u) * f.w[0]);
    double oy = py - (box.lo[1] + (double)((c >> 8) & 255u) * f.w[1]);
    double oz = pz - (box.lo[2] + (double)((c >> 16) & 255u) * f.w[2]);
    min_image(box, ox, oy, oz);
    return make_float4((float)ox, (float)oy, (float)oz, __uint_as_float(c));
}

__global__ __launch_bounds__(256) void cell_sort_kernel(unsigned int ncell, unsigned int cmax,
                                                        const unsigned int* __restrict__ cell_size,
                                                        hm_cell_entry* entries,
                                                        unsigned int* cell_idx,
                                                        const StepState* gate, DevBox box,
                                                        RelFrame rf)
{
    // One wavefront per cell, each independent of the others (no workgroup barrier). A cell of
    // at most 64 particles is ranked in registers: lane k holds entry k and counts the smaller
    // indices with lane reads. Larger cells keep their entries in registers (at most
    // NLHOME / 64 per lane) and rank through the wave's own LDS index table.
    constexpr int EPL = NLHOME / 64;
    __shared__ unsigned int sidx[4][NLHOME];
    if (gate_closed(gate))
        return;
    const unsigned int w = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    const unsigned int rounds = (ncell + gridDim.x * 4 - 1) / (gridDim.x * 4);
    auto place = [&](unsigned int c, const hm_cell_entry& e, unsigned int rank) {
        const unsigned int my = e.idx;
        entries[(size_t)c * cmax + rank] = e;
        if (cell_idx)
            cell_idx[(size_t)c * cmax + rank] = my;
        if (rf.rel)
        {
            // cell-slot relative position, packed cell and slot of particle my
            const unsigned int slot = c * cmax + rank;
            const uint32_t pc = pack_cell(c % rf.dim[0], (c / rf.dim[0]) % rf.dim[1],
                                          c / (rf.dim[0] * rf.dim[1]), e.type);
            const float4 rr = rel_of(e.x, e.y, e.z, pc, box, rf);
            rf.rel[slot] = make_float3(rr.x, rr.y, rr.z);
            rf.relw[slot] = pc;
            const_cast<uint32_t*>(rf.pcell)[my] = pc;
            const_cast<uint32_t*>(rf.slot_of)[my] = slot;
        }
    };
    for (unsigned int r = 0; r < rounds; r++)
    {
        const unsigned int c = (r * gridDim.x + blockIdx.x) * 4 + w;
        const unsigned int n = c < ncell ? min(min(cell_size[c], cmax), (unsigned int)NLHOME) : 0u;
        if (n <= 64)
        {
            hm_cell_entry e{};
            unsigned int key = 0xffffffffu;
            if (lane < n)
            {
                e = entries[(size_t)c * cmax + lane];
                key = e.idx;
            }
            // every lane's entry is in registers before any lane stores (the ranks need all keys)
            unsigned int rank = 0;
            for (unsigned int m = 0; m < n; m++)
                rank += __builtin_amdgcn_readlane(key, m) < key ? 1u : 0u;
            if (lane < n)
                place(c, e, rank);
            continue;
        }
        hm_cell_entry e[EPL];
#pragma unroll
        for (int u = 0; u < EPL; u++)
        {
            const unsigned int k = lane + 64u * u;
            if (k < n)
            {
                e[u] = entries[(size_t)c * cmax + k];
                sidx[w][k] = e[u].idx;
            }
        }
        // the table is this wave's own: its LDS accesses are ordered, a wavefront fence suffices
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        __builtin_amdgcn_wave_barrier();
        unsigned int ranks[EPL];
#pragma unroll
        for (int u = 0; u < EPL; u++)
        {
            const unsigned int k = lane + 64u * u;
            ranks[u] = 0;
            if (k < n)
            {
                const unsigned int my = sidx[w][k];
                for (unsigned int m = 0; m < n; m++)
                    ranks[u] += sidx[w][m] < my ? 1u : 0u;
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int u = 0; u < EPL; u++)
            if (lane + 64u * u < n)
                place(c, e[u], ranks[u]);
    }
}

hipError_t launch_cell_sort(unsigned int ncell, unsigned int cmax, const unsigned int* cell_size,
                            hm_cell_entry* entries, unsigned int* cell_idx, const StepState* gate,
                            hipStream_t s, const DevBox* box, const RelFrame* rel)
{
    if (ncell == 0)
        return hipSuccess;
    unsigned int blocks = min((ncell + 3) / 4, GATED_BLOCKS);
    DevBox b{};
    RelFrame rf{};
    if (box && rel)
    {
        b = *box;
        rf = *rel;
    }
    hipLaunchKernelGGL(cell_sort_kernel, dim3(blocks), dim3(256), 0, s, ncell, cmax, cell_size,
                       entries, cell_idx, gate, b, rf);
    return hipGetLastError();
}

// One particle of the per-particle cell ordering.
__device__ __forceinline__ void rank_particle(unsigned int i, unsigned int cmax,
                                              const double4* __restrict__ pos,
                                              const unsigned int* cell_size,
                                              const unsigned int* keys, const unsigned int* home,
                                              hm_cell_entry* entries, unsigned int* cell_idx,
                                              const DevBox& box, const RelFrame& rf)
{
    const unsigned int h = home[i];
    if (h == NO_SLOT)
        return;
    {
        const unsigned int c = h / cmax;
        const unsigned int n = min(cell_size[c], cmax);
        const unsigned int* kc = keys + (size_t)c * cmax;
        unsigned int rank = 0;
        // unrolled so that several key loads are in flight at once (tens of keys per cell at
        // r_cut 2.5)
#pragma unroll 8
        for (unsigned int m = 0; m < n; m++)
            rank += kc[m] < i ? 1u : 0u;
        const double4 p = pos[i];
        const unsigned int slot = c * cmax + rank;
        hm_cell_entry e;
        e.x = p.x;
        e.y = p.y;
        e.z = p.z;
        e.idx = i;
        e.type = (uint32_t)p.w;
        entries[slot] = e;
        cell_idx[slot] = i;
        if (rf.rel)
        {
            const uint32_t pc = pack_cell(c % rf.dim[0], (c / rf.dim[0]) % rf.dim[1],
                                          c / (rf.dim[0] * rf.dim[1]), e.type);
            const float4 rr = rel_of(p.x, p.y, p.z, pc, box, rf);
            rf.rel[slot] = make_float3(rr.x, rr.y, rr.z);
            rf.relw[slot] = pc;
            const_cast<uint32_t*>(rf.pcell)[i] = pc;
            const_cast<uint32_t*>(rf.slot_of)[i] = slot;
        }
    }
}

// Per-particle cell ordering (the engine's form of cell_sort_kernel): particle i's place in its
// cell is the number of smaller particle indices among the cell's keys. Every thread reads only
// the key table and writes only ordered slots, so the cells need no in-place exchange, and the
// work is spread over particles rather than over cells (most cells of a short cutoff hold a few
// particles).
__global__ __launch_bounds__(256) void cell_rank_kernel(unsigned int N, unsigned int cmax,
                                                        const double4* __restrict__ pos,
                                                        const unsigned int* __restrict__ cell_size,
                                                        const unsigned int* __restrict__ keys,
                                                        const unsigned int* __restrict__ home,
                                                        hm_cell_entry* entries,
                                                        unsigned int* cell_idx,
                                                        const StepState* gate, DevBox box,
                                                        RelFrame rf)
{
    if (gate_closed(gate))
        return;
    for (unsigned int i = blockIdx.x * blockDim.x + threadIdx.x; i < N;
         i += gridDim.x * blockDim.x)
        rank_particle(i, cmax, pos, cell_size, keys, home, entries, cell_idx, box, rf);
}

hipError_t launch_cell_rank(unsigned int N, unsigned int cmax, const double4* pos,
                            const unsigned int* cell_size, const unsigned int* keys,
                            const unsigned int* home, hm_cell_entry* entries,
                            unsigned int* cell_idx, const StepState* gate, hipStream_t s,
                            const DevBox& box, const RelFrame& rel)
{
    if (N == 0)
        return hipSuccess;
    if (!pos || !cell_size || !keys || !home || !entries || !cell_idx || cmax == 0)
        return hipErrorInvalidValue;
    const unsigned int blocks = min((N + 255) / 256, GATED_BLOCKS);
    hipLaunchKernelGGL(cell_rank_kernel, dim3(blocks), dim3(256), 0, s, N, cmax, pos, cell_size,
                       keys, home, entries, cell_idx, gate, box, rel);
    return hipGetLastError();
}

// Small systems: binning and ordering in one workgroup, with a workgroup barrier in place of the
// kernel boundary between them. At these sizes a step costs little more than its launches, so
// one launch fewer per step is the saving. The single workgroup's build is slower than the
// two-kernel one, so the limit keeps it to systems that rebuild cheaply (a how-to run at
// N = 16 384 lost 15 % with it, N = 2 048 gained 8 %).
__global__ __launch_bounds__(1024) void cell_bin_rank_kernel(
    unsigned int N, const double4* __restrict__ pos, DevBox box, unsigned int dx, unsigned int dy,
    unsigned int dz, unsigned int cmax, unsigned int* cell_size, d3* last, unsigned int* cond,
    const StepState* gate, unsigned int* keys, unsigned int* home, hm_cell_entry* entries,
    unsigned int* cell_idx, RelFrame rf)
{
    if (gate_closed(gate))
        return;
    for (unsigned int n = threadIdx.x; n < N; n += blockDim.x)
        bin_particle(n, pos, box, dx, dy, dz, cmax, cell_size, nullptr, last, cond, keys, home);
    // the workgroup's keys, slots and counts are visible to all of its wavefronts after this
    __syncthreads();
    for (unsigned int i = threadIdx.x; i < N; i += blockDim.x)
        rank_particle(i, cmax, pos, cell_size, keys, home, entries, cell_idx, box, rf);
}

hipError_t launch_cell_bin_rank(unsigned int N, const unsigned int dim[3], unsigned int cmax,
                                const double4* pos, unsigned int* cell_size, d3* last_pos,
                                unsigned int* cond3, unsigned int* keys, unsigned int* home,
                                hm_cell_entry* entries, unsigned int* cell_idx,
                                const StepState* gate, hipStream_t s, const DevBox& box,
                                const RelFrame& rel)
{
    if (N == 0)
        return hipSuccess;
    if (N > CELL_FUSED_MAX_N || !pos || !cell_size || !keys || !home || !entries || !cell_idx
        || !cond3 || cmax == 0)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(cell_bin_rank_kernel, dim3(1), dim3(1024), 0, s, N, pos, box, dim[0],
                       dim[1], dim[2], cmax, cell_size, last_pos, cond3, gate, keys, home,
                       entries, cell_idx, rel);
    return hipGetLastError();
}

// largest q in [first, last) with aoff[q] <= k (aoff is non-decreasing)
__device__ __forceinline__ unsigned int find_segment(const unsigned int* aoff, unsigned int first,
                                                     unsigned int last, unsigned int k)
{
    unsigned int lo = first, hi = last - 1;
    while (lo < hi)
    {
        unsigned int mid = (lo + hi + 1) >> 1;
        if (aoff[mid] <= k)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// thread 0: offsets of the adjacent cells of one staging pass (at most CAP candidates)
template<int CAP>
__device__ __forceinline__ void pass_bounds(unsigned int n_adj, const unsigned int* asize,
                                            unsigned int* aoff, unsigned int first,
                                            unsigned int* pass_last, unsigned int* pass_total)
{
    // wave 0: prefix sum of the remaining cells' sizes across lanes; when they all fit the
    // stage the pass is every remaining cell, otherwise thread 0 takes cells while they fit
    if (threadIdx.x < 64)
    {
        const unsigned int t = threadIdx.x;
        const unsigned int q = first + t;
        const unsigned int v = q < n_adj ? asize[q] : 0u;
        unsigned int incl = v;
#pragma unroll
        for (unsigned int off = 1; off < 64; off <<= 1)
        {
            const unsigned int u = __shfl_up(incl, off, 64);
            if (t >= off)
                incl += u;
        }
        const unsigned int all = __shfl(incl, 63, 64);
        if (all <= (unsigned int)CAP)
        {
            if (q < n_adj)
                aoff[q] = incl - v;
            if (t == 0)
            {
                *pass_last = n_adj;
                *pass_total = all;
            }
        }
        else if (t == 0)
        {
            unsigned int tot = 0, l = first;
            while (l < n_adj && tot + asize[l] <= CAP)
            {
                aoff[l] = tot;
                tot += asize[l];
                l++;
            }
            *pass_last = l;
            *pass_total = tot;
        }
    }
}

__device__ __forceinline__ void load_exclusions(const NlistKernelArgs& a, unsigned int i,
                                                unsigned int& ex0, unsigned int& ex1,
                                                unsigned int& ex2, unsigned int& ex3)
{
    unsigned int ne = min(a.n_excl[i], a.max_excl);
    const unsigned int* er = a.excl + (size_t)i * a.max_excl;
    if (ne > 0)
        ex0 = er[0];
    if (ne > 1)
        ex1 = er[1];
    if (ne > 2)
        ex2 = er[2];
    if (ne > 3)
        ex3 = er[3];
}

// The k-th smallest (k = 0, 1, 2) of the wrapped coordinates h - 1, h, h + 1 on an axis of
// d >= 3 cells; returns its offset from h and sets n to the coordinate.
__device__ __forceinline__ int sorted_neighbour(int h, int d, int k, int& n)
{
    int off;
    if (h == 0)  // h - 1 wraps to d - 1, the largest
        off = k == 0 ? 0 : (k == 1 ? 1 : -1);
    else if (h == d - 1)  // h + 1 wraps to 0, the smallest
        off = k == 0 ? 1 : (k == 1 ? -1 : 0);
    else
        off = k - 1;
    n = h + off;
    if (n < 0)
        n += d;
    else if (n >= d)
        n -= d;
    return off;
}

// load_exclusions without the count-to-row dependency: the row's first four entries (within the
// row's max_excl) are read together with the count, then masked by it
__device__ __forceinline__ void load_exclusions_flat(const NlistKernelArgs& a, unsigned int i,
                                                     unsigned int& ex0, unsigned int& ex1,
                                                     unsigned int& ex2, unsigned int& ex3)
{
    const unsigned int* er = a.excl + (size_t)i * a.max_excl;
    const unsigned int cnt = a.n_excl[i];
    unsigned int v0 = 0xffffffffu, v1 = 0xffffffffu, v2 = 0xffffffffu, v3 = 0xffffffffu;
    if (a.max_excl > 0)
        v0 = er[0];
    if (a.max_excl > 1)
        v1 = er[1];
    if (a.max_excl > 2)
        v2 = er[2];
    if (a.max_excl > 3)
        v3 = er[3];
    const unsigned int ne = min(cnt, a.max_excl);
    ex0 = ne > 0 ? v0 : 0xffffffffu;
    ex1 = ne > 1 ? v1 : 0xffffffffu;
    ex2 = ne > 2 ? v2 : 0xffffffffu;
    ex3 = ne > 3 ? v3 : 0xffffffffu;
}

// Minimum-image offset of a position from a cell centre, rounded to FP32 (dims >= 4 keeps every
// staged particle within 1.5 cell widths of the centre, well inside half a box).
__device__ __forceinline__ float3 rel_to_centre(const DevBox& b, double x, double y, double z,
                                                double cx, double cy, double cz)
{
#pragma clang fp contract(off)
    double dx = x - cx, dy = y - cy, dz = z - cz;
    min_image(b, dx, dy, dz);
    return make_float3((float)dx, (float)dy, (float)dz);
}

__device__ __forceinline__ void cell_centre(const DevBox& b, const unsigned int* dim,
                                            unsigned int cell, double& cx, double& cy, double& cz)
{
#pragma clang fp contract(off)
    const unsigned int ci = cell % dim[0], cj = (cell / dim[0]) % dim[1],
                       ck = cell / (dim[0] * dim[1]);
    cx = b.lo[0] + ((double)ci + 0.5) * (b.L[0] / (double)dim[0]);
    cy = b.lo[1] + ((double)cj + 0.5) * (b.L[1] / (double)dim[1]);
    cz = b.lo[2] + ((double)ck + 0.5) * (b.L[2] / (double)dim[2]);
}

// One home cell of the neighbour-list build. FAST: FP32 test with FP64 guard-band re-decision
// (needs >= 4 cells per axis); otherwise every candidate is tested in FP64.
template<bool FAST, bool MT, bool EXCL, class C> struct NlistShared
{
    static_assert(C::CAP % 64 == 0, "a pass padded to whole chunks of 64 must fit the stage");
    float4 sc[FAST ? C::CAP : 1];
    double sx[FAST ? 1 : C::CAP], sy[FAST ? 1 : C::CAP], sz[FAST ? 1 : C::CAP];
    unsigned int sj[FAST ? 1 : C::CAP];
    unsigned char stp[MT ? C::CAP : 1];
    float4 sh[C::HOME];
    unsigned char shtp[MT ? C::HOME : 1];
    unsigned int sslot[C::CAP];  // slot mode: candidate slot | offset code
    unsigned int acell[27], asize[27], aoff[27], acode[27];
    float4 ashift[FAST ? 27 : 1];  // FP32 staging: offset of each adjacent cell's origin from the home centre
    unsigned int hcnt[C::HOME];
    unsigned int pass_last, pass_total;
};

template<bool FAST, bool MT, bool EXCL, class C>
__device__ void nlist_cell(const NlistKernelArgs& a, unsigned int cell,
                           NlistShared<FAST, MT, EXCL, C>& S)
{
#pragma clang fp contract(off)
    const unsigned int tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    // one-wavefront workgroups (cmax <= 64): lane s loads slot s of the home cell together with
    // the cell's size and the adjacent cells' sizes, so the three loads overlap (slots at or past
    // the size are read and ignored)
    constexpr bool ONE_WAVE = C::B == 64;
    // FP32 staging from the relative positions (uniform over the workgroup)
    const bool rs = FAST && a.srel != nullptr && a.n_adj == 27;
    const unsigned int nraw = a.cell_size[cell];
    hm_cell_entry h1{};
    float3 hr1 = make_float3(0.f, 0.f, 0.f);
    unsigned int hi1 = 0, hw1 = 0;
    if (ONE_WAVE && tid < a.cmax)
    {
        const size_t sl = (size_t)cell * a.cmax + tid;
        if (rs)
        {
            hr1 = a.srel[sl];
            hi1 = a.sidx[sl];
            if (MT)
                hw1 = a.swrd[sl];
        }
        else
            h1 = a.entries[sl];
    }
    double cx = 0, cy = 0, cz = 0;
    if (FAST)
        cell_centre(a.box, a.dim, cell, cx, cy, cz);
    if (a.n_adj == 27 && a.dim[0] >= 3 && a.dim[1] >= 3 && a.dim[2] >= 3)
    {
        // the 27 adjacent cells computed from the home cell's coordinates, in increasing index
        // order (the order of cell_adj), which saves a dependent load. The index order is the
        // lexicographic order of (z, y, x), so lane t takes the (t % 3)-th smallest x, the
        // ((t / 3) % 3)-th smallest y and the (t / 9)-th smallest z of the wrapped neighbours.
        if (tid < 27)
        {
            const int dx = (int)a.dim[0], dy = (int)a.dim[1], dz = (int)a.dim[2];
            const int hx = (int)(cell % a.dim[0]), hy = (int)((cell / a.dim[0]) % a.dim[1]),
                      hz = (int)(cell / (a.dim[0] * a.dim[1]));
            int nx, ny, nz;
            const int ox = sorted_neighbour(hx, dx, (int)(tid % 3), nx);
            const int oy = sorted_neighbour(hy, dy, (int)((tid / 3) % 3), ny);
            const int oz = sorted_neighbour(hz, dz, (int)(tid / 9), nz);
            const unsigned int c = (unsigned int)(nx + dx * (ny + dy * nz));
            S.acell[tid] = c;
            S.asize[tid] = min(a.cell_size[c], a.cmax);
            S.acode[tid] = a.offset_codes ? offset_code(-ox, -oy, -oz) : 0u;
            if (FAST)
                S.ashift[tid] = make_float4(((float)ox - 0.5f) * a.cwf[0],
                                            ((float)oy - 0.5f) * a.cwf[1],
                                            ((float)oz - 0.5f) * a.cwf[2], 0.f);
        }
    }
    else if (tid < a.n_adj)
    {
        unsigned int c = a.cell_adj[(size_t)cell * a.n_adj + tid];
        S.acell[tid] = c;
        S.asize[tid] = min(a.cell_size[c], a.cmax);
        unsigned int code = 0;
        if (a.offset_codes)
        {
            const int dx = (int)a.dim[0], dy = (int)a.dim[1];
            const int hx = (int)(cell % a.dim[0]), hy = (int)((cell / a.dim[0]) % a.dim[1]),
                      hz = (int)(cell / (a.dim[0] * a.dim[1]));
            const int nx = (int)(c % a.dim[0]), ny = (int)((c / a.dim[0]) % a.dim[1]),
                      nz = (int)(c / (a.dim[0] * a.dim[1]));
            code = offset_code(wrap_offset(hx - nx, dx), wrap_offset(hy - ny, dy),
                               wrap_offset(hz - nz, (int)a.dim[2]));
        }
        S.acode[tid] = code;
    }
    const unsigned int nhome = min(nraw, a.cmax);
    if (nhome == 0)
        return;
    // one-wavefront workgroups: lane s loads home particle s's exclusions here, where their
    // latency overlaps the staging, and the sweep reads them from lane s
    constexpr bool LANE_EXCL = EXCL && ONE_WAVE;
    unsigned int hx0 = 0xffffffffu, hx1 = 0xffffffffu, hx2 = 0xffffffffu, hx3 = 0xffffffffu;
    for (unsigned int s = tid; s < nhome; s += C::B)
    {
        unsigned int hidx, htype;
        float3 r = make_float3(0.f, 0.f, 0.f);
        if (rs)
        {
            // the home cell's own shift (offset 0), the same sum as its candidates get
            float3 rr = hr1;
            unsigned int w = hw1;
            hidx = hi1;
            if (!ONE_WAVE)
            {
                const size_t sl = (size_t)cell * a.cmax + s;
                rr = a.srel[sl];
                hidx = a.sidx[sl];
                w = MT ? a.swrd[sl] : 0u;
            }
            const float sx0 = (0.f - 0.5f) * a.cwf[0], sy0 = (0.f - 0.5f) * a.cwf[1],
                        sz0 = (0.f - 0.5f) * a.cwf[2];
            r = make_float3(rr.x + sx0, rr.y + sy0, rr.z + sz0);
            htype = w >> 24;
        }
        else
        {
            const hm_cell_entry h = ONE_WAVE ? h1 : a.entries[(size_t)cell * a.cmax + s];
            hidx = h.idx;
            htype = h.type;
            if (FAST)
                r = rel_to_centre(a.box, h.x, h.y, h.z, cx, cy, cz);
        }
        if (LANE_EXCL)
            load_exclusions_flat(a, hidx, hx0, hx1, hx2, hx3);
        S.sh[s] = make_float4(r.x, r.y, r.z, __uint_as_float(hidx));
        if (MT)
            S.shtp[s] = (unsigned char)htype;
        S.hcnt[s] = 0;
    }
    __syncthreads();

    unsigned int first = 0, base_k = 0;
    while (first < a.n_adj)
    {
        pass_bounds<C::CAP>(a.n_adj, S.asize, S.aoff, first, &S.pass_last, &S.pass_total);
        __syncthreads();
        const unsigned int last = S.pass_last, total = S.pass_total;
        for (unsigned int k = tid; k < total; k += C::B)
        {
            const unsigned int q = find_segment(S.aoff, first, last, k);
            const unsigned int sl = S.acell[q] * a.cmax + (k - S.aoff[q]);
            if (a.offset_codes)
                S.sslot[k] = sl | S.acode[q];
            if (rs)
            {
                const float3 r = a.srel[sl];
                const float4 sh4 = S.ashift[q];
                S.sc[k] = make_float4(r.x + sh4.x, r.y + sh4.y, r.z + sh4.z,
                                      __uint_as_float(a.sidx[sl]));
                if (MT)
                    S.stp[k] = (unsigned char)(a.swrd[sl] >> 24);
                continue;
            }
            const hm_cell_entry e = a.entries[sl];
            if (FAST)
            {
                const float3 r = rel_to_centre(a.box, e.x, e.y, e.z, cx, cy, cz);
                S.sc[k] = make_float4(r.x, r.y, r.z, __uint_as_float(e.idx));
            }
            else
            {
                S.sx[k] = e.x;
                S.sy[k] = e.y;
                S.sz[k] = e.z;
                S.sj[k] = e.idx;
            }
            if (MT)
                S.stp[k] = (unsigned char)e.type;
        }
        if (FAST)
        {
            // pad the pass to a whole chunk with candidates far outside every cutoff: the sweep
            // then needs no bounds test (their r^2 is far above r_list^2 and the guard band)
            const unsigned int padded = (total + 63u) & ~63u;
            for (unsigned int k = total + tid; k < padded; k += C::B)
            {
                S.sc[k] = make_float4(3.0e18f, 3.0e18f, 3.0e18f, __uint_as_float(0u));
                if (MT)
                    S.stp[k] = 0;
            }
        }
        __syncthreads();
        for (unsigned int s = wave; s < nhome; s += C::B / 64)
        {
            const float4 h4 = S.sh[s];
            const unsigned int i = __float_as_uint(h4.w);
            const unsigned int ti = MT ? S.shtp[s] : 0u;
            unsigned int ex0 = 0xffffffffu, ex1 = 0xffffffffu, ex2 = 0xffffffffu,
                         ex3 = 0xffffffffu;
            if (LANE_EXCL)
            {
                ex0 = __builtin_amdgcn_readlane(hx0, s);
                ex1 = __builtin_amdgcn_readlane(hx1, s);
                ex2 = __builtin_amdgcn_readlane(hx2, s);
                ex3 = __builtin_amdgcn_readlane(hx3, s);
            }
            else if (EXCL)
                load_exclusions(a, i, ex0, ex1, ex2, ex3);
            double hx = 0, hy = 0, hz = 0;
            if (!FAST)
            {
                const hm_cell_entry h = a.entries[(size_t)cell * a.cmax + s];
                hx = h.x;
                hy = h.y;
                hz = h.z;
            }
            unsigned int cnt = S.hcnt[s];
            unsigned int* row = a.nlist + (size_t)i * a.nmax;
            unsigned short* lrow = a.lrow ? a.lrow + (size_t)i * a.nmax : nullptr;
            if (FAST)
            {
                // C::U chunks of 64 candidates per iteration: the LDS reads and distance tests of
                // the chunks are independent, the rare guard-band re-decisions are taken out of
                // the main path, then the hits are compacted chunk by chunk in candidate order
                auto chunks = [&](auto uc, unsigned int b) {
                    constexpr int U = decltype(uc)::value;
                    float4 c4[U];
                    unsigned int tp[U];
                    bool hit[U];
                    unsigned int chk = 0;
                    // positions past the pass's candidates hold the far-away padding
#pragma unroll
                    for (int u = 0; u < U; u++)
                        c4[u] = S.sc[b + 64u * u + lane];
#pragma unroll
                    for (int u = 0; u < U; u++)
                    {
                        const float dx = h4.x - c4[u].x, dy = h4.y - c4[u].y, dz = h4.z - c4[u].z;
                        // fused: at most the rounding error the guard band allows for
                        const float dr2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
                        tp[u] = MT ? ti * a.ntypes + S.stp[b + 64u * u + lane] : 0u;
                        const float t = dr2 - a.rlistsq_f[tp[u]];
                        hit[u] = t <= 0.f;
                        chk |= fabsf(t) <= a.guard[tp[u]] ? (1u << u) : 0u;
                    }
                    if (chk)
                    {
#pragma unroll
                        for (int u = 0; u < U; u++)
                            if (chk & (1u << u))
                            {
                                // exact FP64 decision as NeighborListBinned.cc:151-157 computes it
                                const double4 pi = a.pos[i];
                                const double4 pj = a.pos[__float_as_uint(c4[u].w)];
                                double ddx = pi.x - pj.x, ddy = pi.y - pj.y, ddz = pi.z - pj.z;
                                min_image(a.box, ddx, ddy, ddz);
                                hit[u] = ddx * ddx + ddy * ddy + ddz * ddz <= a.rlistsq[tp[u]];
                            }
                    }
#pragma unroll
                    for (int u = 0; u < U; u++)
                    {
                        const unsigned int j = __float_as_uint(c4[u].w);
                        bool h = hit[u] && (j != i) && a.enabled[tp[u]];
                        if (EXCL)
                            h = h && j != ex0 && j != ex1 && j != ex2 && j != ex3;
                        const uint64_t m = wave_ballot(h);
                        if (h)
                        {
                            const unsigned int k = b + 64u * u + lane;
                            const unsigned int p = cnt + lanes_below(m);
                            if (p < a.nmax)
                            {
                                // plain stores: the hits of one sweep land in a few 128-byte
                                // row blocks, which L2 merges into whole lines before they go
                                // out (streaming stores wrote 3.3x the row bytes to HBM)
                                if (!a.skip_index_rows)
                                {
                                    if (a.offset_codes)
                                        row[row_store_pos(p)] = S.sslot[k];
                                    else
                                        row[p] = j;
                                }
                                if (lrow)
                                    lrow[p] = (unsigned short)(base_k + k);
                            }
                        }
                        cnt += (unsigned int)__popcll(m);
                    }
                };
                // whole groups of C::U chunks, then the remaining chunks one at a time
                unsigned int b = 0;
                for (; b + 64u * C::U <= total; b += 64u * C::U)
                    chunks(std::integral_constant<int, C::U>(), b);
                for (; b < total; b += 64u)
                    chunks(std::integral_constant<int, 1>(), b);
            }
            else
                for (unsigned int b = 0; b < total; b += 64)
                {
                    const unsigned int k = b + lane;
                    bool hit = false;
                    unsigned int j = 0;
                    if (k < total)
                    {
                        j = S.sj[k];
                        double dx = hx - S.sx[k];
                        double dy = hy - S.sy[k];
                        double dz = hz - S.sz[k];
                        min_image(a.box, dx, dy, dz);
                        const double dr2 = dx * dx + dy * dy + dz * dz;
                        const unsigned int tp = MT ? ti * a.ntypes + S.stp[k] : 0u;
                        hit = dr2 <= a.rlistsq[tp];
                        hit = hit && (j != i) && a.enabled[tp];
                        if (EXCL)
                            hit = hit && j != ex0 && j != ex1 && j != ex2 && j != ex3;
                    }
                    const uint64_t m = __ballot(hit);
                    if (hit)
                    {
                        const unsigned int p = cnt + lanes_below(m);
                        if (p < a.nmax)
                        {
                            if (!a.skip_index_rows)
                            {
                                if (a.offset_codes)
                                    row[row_store_pos(p)] = S.sslot[k];
                                else
                                    row[p] = j;
                            }
                            if (lrow)
                                lrow[p] = (unsigned short)(base_k + k);
                        }
                    }
                    cnt += (unsigned int)__popcll(m);
                }
            if (lane == 0)
                S.hcnt[s] = cnt;
        }
        __syncthreads();
        base_k += total;
        first = last;
    }
    for (unsigned int s = tid; s < nhome; s += C::B)
    {
        const unsigned int i = __float_as_uint(S.sh[s].w);
        const unsigned int c = S.hcnt[s];
        a.n_neigh[i] = c;
        if (c > a.nmax)
            atomicMax(a.cond, c);
    }
    __syncthreads();
}

template<bool FAST, bool MT, bool EXCL, class C>
__global__ __launch_bounds__(C::B) __attribute__((amdgpu_waves_per_eu(C::WAVES)))
void nlist_kernel(NlistKernelArgs a)
{
    __shared__ NlistShared<FAST, MT, EXCL, C> S;
    if (gate_closed(a.gate))
        return;
    if (a.counter && blockIdx.x == 0 && threadIdx.x == 0)
    {
        // build statistics of NeighborList::needsUpdating (NeighborList.cc:1349-1415)
        StepState* st = a.counter;
        const unsigned long long t1 = st->timestep + 1;
        if (st->forced_at != t1)
        {
            if (t1 > st->last_update && t1 - st->last_update < st->shortest)
                st->shortest = t1 - st->last_update;
            if (a.dist_check && a.check_delay > 1 && t1 == st->last_update + a.check_delay)
                st->n_dangerous++;
        }
        st->last_update = t1;
        st->n_builds++;
    }
    const XcdRange r = xcd_range(a.ncell);
    for (unsigned int cell = r.begin; cell < r.end; cell += r.stride)
        nlist_cell<FAST, MT, EXCL, C>(a, cell, S);
}

template<bool FAST, class C>
static void nlist_dispatch(const NlistKernelArgs& a, bool mt, bool ex, unsigned int grid,
                           hipStream_t s)
{
    if (mt && ex)
        hipLaunchKernelGGL((nlist_kernel<FAST, true, true, C>), dim3(grid), dim3(C::B), 0, s, a);
    else if (mt)
        hipLaunchKernelGGL((nlist_kernel<FAST, true, false, C>), dim3(grid), dim3(C::B), 0, s, a);
    else if (ex)
        hipLaunchKernelGGL((nlist_kernel<FAST, false, true, C>), dim3(grid), dim3(C::B), 0, s, a);
    else
        hipLaunchKernelGGL((nlist_kernel<FAST, false, false, C>), dim3(grid), dim3(C::B), 0, s,
                           a);
}

hipError_t launch_nlist_build(const NlistKernelArgs& a_in, hipStream_t s)
{
    if (a_in.ncell == 0)
        return hipSuccess;
    NlistKernelArgs a = a_in;
    const unsigned int grid = round_up(a.grid ? a.grid : min(a.ncell, GATED_BLOCKS), 8);
    // small cells: a few particles per cell on average, and every home cell fits the
    // one-wavefront workgroup (cmax <= 64)
    bool small = a.cmax <= (unsigned int)NlSmall::HOME && a.N <= 8ull * a.ncell;
    bool tiny = small && a.N <= 4ull * a.ncell;
    const unsigned int grid_small = round_up(a.grid ? a.grid : min(a.ncell, 4 * GATED_BLOCKS), 8);
    // an explicit shape (the autotuner's) where its home-cell capacity holds cmax; the staging
    // passes of every shape take any number of candidates
    bool mid_ok = a.cmax <= (unsigned int)NlMid::HOME;
    if (a.shape && a.cmax <= (unsigned int)NlSmall::HOME)
    {
        tiny = a.shape == NL_SHAPE_TINY;
        small = tiny || a.shape == NL_SHAPE_SMALL;
        mid_ok = a.shape == NL_SHAPE_MID;
    }
    const bool mt = a.ntypes > 1;
    const bool ex = a.n_excl != nullptr && a.max_excl > 0;
    const bool fast = a.pos != nullptr && a.dim[0] >= 4 && a.dim[1] >= 4 && a.dim[2] >= 4;
    if (fast)
    {
        // worst-case FP32 error of dr^2 near the threshold: coordinates are within
        // dmax = 1.5 cell widths of the centre (rounding u*dmax each), differences within 2*dmax
        const double u = 1.0 / 16777216.0;
        double wmax = 0.0;
        for (int d = 0; d < 3; d++)
            wmax = fmax(wmax, a.box.L[d] / (double)a.dim[d]);
        const double dmax = 1.5 * wmax;
        // entries: each coordinate rounded once (u * dmax), the difference once more (2 u dmax).
        // FP32 staging: the relative position (u w), the FP32 width times (o - 1/2) (3 u w) and
        // the sum (1.5 u w) per coordinate, twice, and the difference (3 u w): 14 u w, taken as 16
        const double e_dx = a.srel ? 16.0 * u * wmax : u * 2.0 * dmax + 2.0 * u * dmax;
        const unsigned int nt = a.ntypes ? a.ntypes : 1;
        for (unsigned int q = 0; q < nt * nt; q++)
        {
            const double rl2 = a.rlistsq[q];
            const double rl = sqrt(rl2);
            a.rlistsq_f[q] = (float)rl2;
            a.guard[q] = (float)(4.0 * (2.0 * sqrt(3.0) * rl * e_dx + 6.0 * u * rl2) + 1e-12);
        }
        if (tiny)
            nlist_dispatch<true, NlTiny>(a, mt, ex, grid_small, s);
        else if (small)
            nlist_dispatch<true, NlSmall>(a, mt, ex, grid_small, s);
        else if (mid_ok)
            nlist_dispatch<true, NlMid>(a, mt, ex, grid, s);
        else
            nlist_dispatch<true, NlBig>(a, mt, ex, grid, s);
    }
    else if (small)
        nlist_dispatch<false, NlSmall>(a, mt, ex, grid_small, s);
    else
        nlist_dispatch<false, NlBig>(a, mt, ex, grid, s);
    return hipGetLastError();
}

// NeighborList::distanceCheck (hoomd/md/NeighborList.cc:1220-1289), standalone form.
__global__ void nlist_check_kernel(unsigned int N, const double4* __restrict__ pos,
                                   const d3* __restrict__ last, DevBox box, double maxsq,
                                   unsigned int* flag)
{
#pragma clang fp contract(off)
    unsigned int n = blockIdx.x * blockDim.x + threadIdx.x;
    bool moved = false;
    if (n < N)
    {
        double4 p = pos[n];
        d3 l = last[n];
        double dx = p.x - 1.0 * l.x, dy = p.y - 1.0 * l.y, dz = p.z - 1.0 * l.z;
        min_image(box, dx, dy, dz);
        moved = dx * dx + dy * dy + dz * dz >= maxsq;
    }
    if (__ballot(moved) != 0 && (threadIdx.x & 63u) == 0)
        atomicOr(flag, 1u);
}

hipError_t launch_nlist_check(unsigned int N, const double4* pos, const d3* last,
                              const DevBox& box, double maxsq, unsigned int* flag, hipStream_t s)
{
    if (N == 0)
        return hipSuccess;
    hipLaunchKernelGGL(nlist_check_kernel, dim3((N + 255) / 256), dim3(256), 0, s, N, pos, last,
                       box, maxsq, flag);
    return hipGetLastError();
}

// ======================================================================== pair forces
// PotentialPair::computeForces (hoomd/md/PotentialPair.h:587-794) with a full list, the
// structure of gpu_compute_pair_forces_shared_kernel (hoomd/md/PotentialPairGPU.cuh:140-404):
// TPP lanes per particle share the particle's neighbour row, TPP * PAIR_BATCH entries per
// batch. All PAIR_BATCH position gathers of a batch are issued before any is used, so each lane
// keeps PAIR_BATCH loads in flight; slots past the row end gather the particle itself and are
// masked out arithmetically. The partial sums are combined by xor-shuffles inside the TPP group.
constexpr int PAIR_BATCH = 8;

// Row loads are unconditional (their index is clamped into the row) so that a batch's loads are
// all in flight together. The empty asm consumes each loaded value at this point, which stops
// the compiler from sinking a load into the branch of its later use, where every load would be
// followed by its own wait.
template<int B> __device__ __forceinline__ void pin_loads(unsigned int (&v)[B])
{
#pragma unroll
    for (int q = 0; q < B; q++)
        asm volatile("" : "+v"(v[q]));
}

template<class Eval> struct ParamTable
{
    typename Eval::param_type p[HM_MAX_TYPES * HM_MAX_TYPES];
};

template<class Eval, int TPP, bool VIRIAL, bool MT, int SHIFT>
__global__ __launch_bounds__(1024) void pair_kernel(PairKernelArgs a, ParamTable<Eval> tab)
{
    const unsigned int lb = xcd_block(blockIdx.x, gridDim.x);
    const unsigned int gt = lb * blockDim.x + threadIdx.x;
    const unsigned int i = gt / TPP;
    const unsigned int sub = gt % TPP;
    const bool active = i < a.N;
    const unsigned int ii = active ? i : a.N - 1;
    const double4 pi = a.pos[ii];
    // a row never holds more than nmax entries (n_neigh reports the full count on overflow)
    const unsigned int n = active ? min(a.n_neigh[ii], a.nmax) : 0u;
    const unsigned int ti = MT ? (unsigned int)pi.w : 0u;
    const unsigned int jmax = a.N - 1;
    const unsigned int elast = a.nmax - 1;

    float fx = 0.f, fy = 0.f, fz = 0.f, pe = 0.f;
    float vxx = 0.f, vxy = 0.f, vxz = 0.f, vyy = 0.f, vyz = 0.f, vzz = 0.f;
    const unsigned int* row = a.nlist + (size_t)ii * a.nmax;

    // Entry k0 + q*TPP + sub belongs to lane sub: in every load instruction the TPP lanes of a
    // particle read TPP consecutive row entries, whose positions (cell order, index-sorted inside
    // a cell, Hilbert-sorted indices) share few cache lines.
    for (unsigned int k0 = 0; k0 < n; k0 += TPP * PAIR_BATCH)
    {
        // entries past n are replaced by the particle itself
        unsigned int js[PAIR_BATCH];
#pragma unroll
        for (int q = 0; q < PAIR_BATCH; q++)
            js[q] = row[min(k0 + q * TPP + sub, elast)];
        pin_loads<PAIR_BATCH>(js);
#pragma unroll
        for (int q = 0; q < PAIR_BATCH; q++)
            js[q] = k0 + q * TPP + sub < n ? js[q] : ii;
        double px[PAIR_BATCH], py[PAIR_BATCH], pz[PAIR_BATCH];
        unsigned int tj[PAIR_BATCH];
#pragma unroll
        for (int q = 0; q < PAIR_BATCH; q++)
        {
            const unsigned int j = min(js[q], jmax);
            const double4 pj = a.pos[j];
            px[q] = pj.x;
            py[q] = pj.y;
            pz[q] = pj.z;
            tj[q] = MT ? (unsigned int)pj.w : 0u;
        }
#pragma unroll
        for (int q = 0; q < PAIR_BATCH; q++)
        {
            double dx, dy, dz, rsq;
            {
#pragma clang fp contract(off)
                dx = pi.x - px[q];
                dy = pi.y - py[q];
                dz = pi.z - pz[q];
                min_image(a.box, dx, dy, dz);
                rsq = dx * dx + dy * dy + dz * dz;
            }
            const unsigned int tp = MT ? ti * a.ntypes + tj[q] : 0u;
            const bool in = (k0 + q * TPP + sub < n) && (rsq < a.rcutsq[tp]);
            const float rsqf = in ? (float)rsq : 1.0f;
            const float rcf = a.rcutsq_f[tp];
            Eval ev(rsqf, rcf, tab.p[tp]);
            float fdr = 0.f, e = 0.f;
            const bool shift = (SHIFT == 1) || (SHIFT == 2 && a.ronsq_f[tp] > rcf);
            ev.evalForceAndEnergy(fdr, e, shift);
            if (SHIFT == 2)
            {
                // XPLOR smoothing (PotentialPair.h:706-731)
                const float ron = a.ronsq_f[tp];
                if (rsqf >= ron && rsqf < rcf)
                {
                    float den = 1.f / ((rcf - ron) * (rcf - ron) * (rcf - ron));
                    float rmc = rsqf - rcf;
                    float sw = rmc * rmc * (rcf + 2.f * rsqf - 3.f * ron) * den;
                    float ds = 12.f * (rsqf - ron) * rmc * den;
                    float e0 = e;
                    e = e0 * sw;
                    fdr = sw * fdr - ds * e0;
                }
            }
            fdr = in ? fdr : 0.f;
            e = in ? e : 0.f;
            const float dxf = (float)dx, dyf = (float)dy, dzf = (float)dz;
            fx += dxf * fdr;
            fy += dyf * fdr;
            fz += dzf * fdr;
            pe += 0.5f * e;
            if (VIRIAL)
            {
                const float f2 = 0.5f * fdr;
                vxx += f2 * dxf * dxf;
                vxy += f2 * dxf * dyf;
                vxz += f2 * dxf * dzf;
                vyy += f2 * dyf * dyf;
                vyz += f2 * dyf * dzf;
                vzz += f2 * dzf * dzf;
            }
        }
    }
    if (TPP > 1)
    {
#pragma unroll
        for (int off = TPP / 2; off > 0; off >>= 1)
        {
            fx += __shfl_xor(fx, off, TPP);
            fy += __shfl_xor(fy, off, TPP);
            fz += __shfl_xor(fz, off, TPP);
            pe += __shfl_xor(pe, off, TPP);
            if (VIRIAL)
            {
                vxx += __shfl_xor(vxx, off, TPP);
                vxy += __shfl_xor(vxy, off, TPP);
                vxz += __shfl_xor(vxz, off, TPP);
                vyy += __shfl_xor(vyy, off, TPP);
                vyz += __shfl_xor(vyz, off, TPP);
                vzz += __shfl_xor(vzz, off, TPP);
            }
        }
    }
    if (active && sub == 0)
    {
        a.force[i] = make_float4(fx, fy, fz, pe);
        if (VIRIAL)
        {
            const size_t P = a.N;
            a.virial[0 * P + i] = vxx;
            a.virial[1 * P + i] = vxy;
            a.virial[2 * P + i] = vxz;
            a.virial[3 * P + i] = vyy;
            a.virial[4 * P + i] = vyz;
            a.virial[5 * P + i] = vzz;
        }
    }
}

template<class Eval, int TPP, bool VIRIAL, bool MT>
static hipError_t pair_dispatch_shift(const PairKernelArgs& a, const ParamTable<Eval>& tab,
                                      dim3 grid, dim3 block, hipStream_t s)
{
    if constexpr (Eval::shift_modes == 1)
    {
        hipLaunchKernelGGL((pair_kernel<Eval, TPP, VIRIAL, MT, 0>), grid, block, 0, s, a, tab);
        return hipGetLastError();
    }
    switch (a.shift_mode)
    {
    case 0:
        hipLaunchKernelGGL((pair_kernel<Eval, TPP, VIRIAL, MT, 0>), grid, block, 0, s, a, tab);
        break;
    case 1:
        hipLaunchKernelGGL((pair_kernel<Eval, TPP, VIRIAL, MT, 1>), grid, block, 0, s, a, tab);
        break;
    default:
        hipLaunchKernelGGL((pair_kernel<Eval, TPP, VIRIAL, MT, 2>), grid, block, 0, s, a, tab);
        break;
    }
    return hipGetLastError();
}

template<class Eval, int TPP>
static hipError_t pair_dispatch_tpp(const PairKernelArgs& a, const ParamTable<Eval>& tab,
                                    unsigned int block, hipStream_t s)
{
    unsigned int threads = a.N * TPP;
    unsigned int nblocks = round_up((threads + block - 1) / block, 8);
    dim3 g(nblocks), b(block);
    bool mt = a.ntypes > 1;
    if (a.compute_virial)
        return mt ? pair_dispatch_shift<Eval, TPP, true, true>(a, tab, g, b, s)
                  : pair_dispatch_shift<Eval, TPP, true, false>(a, tab, g, b, s);
    return mt ? pair_dispatch_shift<Eval, TPP, false, true>(a, tab, g, b, s)
              : pair_dispatch_shift<Eval, TPP, false, false>(a, tab, g, b, s);
}

template<class Eval>
static hipError_t pair_dispatch(const PairKernelArgs& a, const ParamTable<Eval>& tab,
                                PairLaunch cfg, hipStream_t s)
{
    if (a.N == 0)
        return hipSuccess;
    unsigned int block = cfg.block ? cfg.block : 256;
    switch (cfg.tpp ? cfg.tpp : 4)
    {
    case 1:
        return pair_dispatch_tpp<Eval, 1>(a, tab, block, s);
    case 2:
        return pair_dispatch_tpp<Eval, 2>(a, tab, block, s);
    case 8:
        return pair_dispatch_tpp<Eval, 8>(a, tab, block, s);
    case 16:
        return pair_dispatch_tpp<Eval, 16>(a, tab, block, s);
    default:
        return pair_dispatch_tpp<Eval, 4>(a, tab, block, s);
    }
}

hipError_t launch_pair_lj(const PairKernelArgs& a, const hm_lj_params* params, PairLaunch cfg,
                          hipStream_t s)
{
    using E = EvaluatorPairLJ<float>;
    ParamTable<E> tab;
    unsigned int nt = a.ntypes ? a.ntypes : 1;
    for (unsigned int q = 0; q < nt * nt; q++)
        tab.p[q] = E::param_type::make(params[q].sigma, params[q].epsilon);
    return pair_dispatch<E>(a, tab, cfg, s);
}

hipError_t launch_pair_table(const PairKernelArgs& a, const hm_table_params* params,
                             PairLaunch cfg, hipStream_t s)
{
    using E = EvaluatorPairTable<float>;
    ParamTable<E> tab;
    unsigned int nt = a.ntypes ? a.ntypes : 1;
    for (unsigned int q = 0; q < nt * nt; q++)
        tab.p[q] = E::param_type::make(params[q].U, params[q].F, params[q].r_min,
                                       params[q].width);
    return pair_dispatch<E>(a, tab, cfg, s);
}

// ======================================================================== pair forces, FP32 relative
// The pair sums of pair_kernel with one 16-byte gather per neighbour: each particle carries its
// position relative to the origin of the cell it was binned into at the last build (RelFrame),
// and every row entry carries the offset between the two cells of that build, wrapped to
// {-1, 0, 1} per axis (offset codes, written by the neighbour-list build). The pair vector is
// (ri - rj) + offset * w in FP32 -- the continuous displacement, i.e. the minimum image while
// every particle stays within r_buff / 2 of its build position (the distance check's rebuild
// criterion). The cutoff is decided in FP32; pairs whose FP32 r^2 lies within a guard band
// (four times the worst-case FP32 error) of r_cut^2 are re-decided from the FP64 positions with
// the FP64 minimum image, so the set of interacting pairs is the one the FP64 CPU path selects.
struct PairAcc
{
    float fx = 0.f, fy = 0.f, fz = 0.f, pe = 0.f;
    float vxx = 0.f, vxy = 0.f, vxz = 0.f, vyy = 0.f, vyz = 0.f, vzz = 0.f;
};

// Per type pair constants of the relative kernel: the fast evaluator entry (modes none and
// shift) and the standard parameters (XPLOR).
template<class Eval> struct RelTable
{
    typename FastPair<Eval>::entry f[HM_MAX_TYPES * HM_MAX_TYPES];
    typename Eval::param_type p[HM_MAX_TYPES * HM_MAX_TYPES];
};

// FP64 re-decision of r < r_cut for the pair (ii, particle of slot js), as the CPU path decides
__device__ __forceinline__ bool rel_recheck(const PairRelArgs& a, unsigned int ii,
                                            unsigned int js, unsigned int jmax, double rcutsq)
{
#pragma clang fp contract(off)
    const double4 pi = a.pos[ii];
    const double4 pj = a.pos[min(a.part_of[js], jmax)];
    double dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
    min_image(a.box, dx, dy, dz);
    return dx * dx + dy * dy + dz * dz < rcutsq;
}

// Force / r and half energy of one pair that is `in` range.
template<class Eval, int SHIFT>
__device__ __forceinline__ void rel_pair_eval(const float* rcutsq_f, const float* ronsq_f,
                                              const RelTable<Eval>& tab, unsigned int tp,
                                              float rsq, bool in, float& fdr, float& eh)
{
    if (SHIFT != 2)
        FastPair<Eval>::template eval<SHIFT == 1>(rsq, in, tab.f[tp], fdr, eh);
    else
    {
        // XPLOR smoothing (PotentialPair.h:706-731) on the standard evaluator
        const float rcf = rcutsq_f[tp];
        const float rsqf = in ? rsq : 1.0f;
        Eval ev(rsqf, rcf, tab.p[tp]);
        float en = 0.f;
        fdr = 0.f;
        ev.evalForceAndEnergy(fdr, en, ronsq_f[tp] > rcf);
        const float ron = ronsq_f[tp];
        if (rsqf >= ron && rsqf < rcf)
        {
            float den = 1.f / ((rcf - ron) * (rcf - ron) * (rcf - ron));
            float rmc = rsqf - rcf;
            float sw = rmc * rmc * (rcf + 2.f * rsqf - 3.f * ron) * den;
            float ds = 12.f * (rsqf - ron) * rmc * den;
            float e0 = en;
            en = e0 * sw;
            fdr = sw * fdr - ds * e0;
        }
        fdr = in ? fdr : 0.f;
        eh = in ? 0.5f * en : 0.f;
    }
}

template<int B, class Eval, int TPP, bool VIRIAL, bool MT, int SHIFT>
__global__ __launch_bounds__(1024) void pair_rel_kernel(PairRelArgs a, RelTable<Eval> tab)
{
    // shift vectors of the 64 offset codes: code bits 0-1 / 2-3 / 4-5 hold the x / y / z
    // offset in {-1, 0, 1} as 2-bit two's complement
    __shared__ float4 shtab[64];
    if (threadIdx.x < 64)
    {
        const int c = (int)threadIdx.x;
        const int ox = (c << 30) >> 30, oy = (c << 28) >> 30, oz = (c << 26) >> 30;
        shtab[c] = make_float4((float)ox * a.wf[0], (float)oy * a.wf[1], (float)oz * a.wf[2], 0.f);
    }
    __syncthreads();
    const unsigned int lb = xcd_block(blockIdx.x, gridDim.x);
    const unsigned int gt = lb * blockDim.x + threadIdx.x;
    const unsigned int i = gt / TPP;
    const unsigned int sub = gt % TPP;
    const bool active = i < a.N;
    const unsigned int ii = active ? i : a.N - 1;
    const unsigned int smax = a.nslot - 1;
    const unsigned int si = min(a.slot_of[ii], smax);
    const float3 ri = a.rel[si];
    const unsigned int n = active ? min(a.n_neigh[ii], a.nmax) : 0u;
    const unsigned int ti = MT ? (a.relw[si] >> 24) : 0u;
    const unsigned int jmax = a.N - 1;
    constexpr unsigned int S = TPP * B;

    PairAcc acc;
    // Rows are stored in transposed blocks of 32 (row_store_pos): lane sub reads B entries as
    // B / 4 16-byte loads at stored positions p0 + 4 (h TPP + sub); entry q of the lane holds
    // neighbour row_neighbour(p0 + 4 ((q / 4) TPP + sub) + q % 4), so for a fixed q the lanes
    // of a particle hold consecutive neighbours. Whole blocks are read (a block's neighbours
    // are spread over all its positions); rows are followed by NLIST_PAD readable entries, and
    // entries of neighbours >= n are stale but valid slots and are masked.
    static_assert(B % 4 == 0 && (S % ROW_BLOCK == 0 || ROW_BLOCK % S == 0), "row blocks");
    const uint4* row4 = (const uint4*)(a.nlist + (size_t)ii * a.nmax);
    const unsigned int nblk = (n + ROW_BLOCK - 1) & ~(ROW_BLOCK - 1);
    // with whole blocks per iteration (k0 a multiple of 32) entry q holds neighbour
    // k0 + row_neighbour(offset of q), so its validity is k0 < n - row_neighbour(offset of q)
    constexpr bool WHOLE = S % ROW_BLOCK == 0;
    int lim[B];
#pragma unroll
    for (int q = 0; q < B; q++)
        lim[q] = (int)n - (int)row_neighbour(4u * ((unsigned int)(q / 4) * TPP + sub) + (q % 4));
    for (unsigned int k0 = 0; k0 < nblk; k0 += S)
    {
        unsigned int ec[B];
#pragma unroll
        for (int h = 0; h < B / 4; h++)
        {
            const uint4 v = row4[(k0 >> 2) + h * TPP + sub];
            ec[4 * h + 0] = v.x;
            ec[4 * h + 1] = v.y;
            ec[4 * h + 2] = v.z;
            ec[4 * h + 3] = v.w;
        }
        pin_loads<B>(ec);
        unsigned int js[B];
        float3 rj[B];
        unsigned int tj[B];
#pragma unroll
        for (int q = 0; q < B; q++)
        {
            js[q] = min(ec[q] & OFFSET_CODE_MASK, smax);
            rj[q] = a.rel[js[q]];
            tj[q] = MT ? (a.relw[js[q]] >> 24) : 0u;
        }
        // phase 1: pair vectors, FP32 cutoff decisions and the guard-band flags of the batch
        float dx[B], dy[B], dz[B], rsq[B];
        unsigned int tp[B];
        bool in[B];
        unsigned int chk = 0;
#pragma unroll
        for (int q = 0; q < B; q++)
        {
            const float4 sh = shtab[ec[q] >> OFFSET_CODE_SHIFT];
            dx[q] = (ri.x - rj[q].x) + sh.x;
            dy[q] = (ri.y - rj[q].y) + sh.y;
            dz[q] = (ri.z - rj[q].z) + sh.z;
            rsq[q] = dx[q] * dx[q] + dy[q] * dy[q] + dz[q] * dz[q];
            tp[q] = MT ? ti * a.ntypes + tj[q] : 0u;
            const bool valid =
                WHOLE ? (int)k0 < lim[q]
                      : row_neighbour(k0 + 4u * ((unsigned int)(q / 4) * TPP + sub) + (q % 4)) < n;
            const float t = rsq[q] - a.rcutsq_f[tp[q]];
            in[q] = valid && t < 0.f;
            chk |= (valid && fabsf(t) <= a.guard[tp[q]]) ? (1u << q) : 0u;
        }
        // phase 2 (rare): FP64 re-decision of the pairs within the guard band
        if (chk)
        {
#pragma unroll
            for (int q = 0; q < B; q++)
                if (chk & (1u << q))
                    in[q] = rel_recheck(a, ii, js[q], jmax, a.rcutsq[tp[q]]);
        }
        // phase 3: evaluation and accumulation
#pragma unroll
        for (int q = 0; q < B; q++)
        {
            float fdr, eh;
            rel_pair_eval<Eval, SHIFT>(a.rcutsq_f, a.ronsq_f, tab, tp[q], rsq[q], in[q], fdr, eh);
            acc.fx += dx[q] * fdr;
            acc.fy += dy[q] * fdr;
            acc.fz += dz[q] * fdr;
            acc.pe += eh;
            if (VIRIAL)
            {
                const float f2 = 0.5f * fdr;
                acc.vxx += f2 * dx[q] * dx[q];
                acc.vxy += f2 * dx[q] * dy[q];
                acc.vxz += f2 * dx[q] * dz[q];
                acc.vyy += f2 * dy[q] * dy[q];
                acc.vyz += f2 * dy[q] * dz[q];
                acc.vzz += f2 * dz[q] * dz[q];
            }
        }
    }
    if (TPP > 1)
    {
#pragma unroll
        for (int off = TPP / 2; off > 0; off >>= 1)
        {
            acc.fx += __shfl_xor(acc.fx, off, TPP);
            acc.fy += __shfl_xor(acc.fy, off, TPP);
            acc.fz += __shfl_xor(acc.fz, off, TPP);
            acc.pe += __shfl_xor(acc.pe, off, TPP);
            if (VIRIAL)
            {
                acc.vxx += __shfl_xor(acc.vxx, off, TPP);
                acc.vxy += __shfl_xor(acc.vxy, off, TPP);
                acc.vxz += __shfl_xor(acc.vxz, off, TPP);
                acc.vyy += __shfl_xor(acc.vyy, off, TPP);
                acc.vyz += __shfl_xor(acc.vyz, off, TPP);
                acc.vzz += __shfl_xor(acc.vzz, off, TPP);
            }
        }
    }
    if (active && sub == 0)
    {
        a.force[i] = make_float4(acc.fx, acc.fy, acc.fz, acc.pe);
        if (VIRIAL)
        {
            const size_t P = a.N;
            a.virial[0 * P + i] = acc.vxx;
            a.virial[1 * P + i] = acc.vxy;
            a.virial[2 * P + i] = acc.vxz;
            a.virial[3 * P + i] = acc.vyy;
            a.virial[4 * P + i] = acc.vyz;
            a.virial[5 * P + i] = acc.vzz;
        }
    }
}

template<int B, class Eval, int TPP, bool VIRIAL, bool MT>
static void pair_rel_dispatch_shift(const PairRelArgs& a, const RelTable<Eval>& tab, dim3 grid,
                                    dim3 block, hipStream_t s)
{
    if constexpr (Eval::shift_modes == 1)
    {
        hipLaunchKernelGGL((pair_rel_kernel<B, Eval, TPP, VIRIAL, MT, 0>), grid, block, 0, s, a,
                           tab);
        return;
    }
    switch (a.shift_mode)
    {
    case 0:
        hipLaunchKernelGGL((pair_rel_kernel<B, Eval, TPP, VIRIAL, MT, 0>), grid, block, 0, s, a,
                           tab);
        break;
    case 1:
        hipLaunchKernelGGL((pair_rel_kernel<B, Eval, TPP, VIRIAL, MT, 1>), grid, block, 0, s, a,
                           tab);
        break;
    default:
        hipLaunchKernelGGL((pair_rel_kernel<B, Eval, TPP, VIRIAL, MT, 2>), grid, block, 0, s, a,
                           tab);
        break;
    }
}

template<int B, class Eval, int TPP>
static void pair_rel_dispatch_batch(const PairRelArgs& a, const RelTable<Eval>& tab, dim3 g,
                                    dim3 b, hipStream_t s)
{
    const bool mt = a.ntypes > 1;
    if (a.compute_virial)
        mt ? pair_rel_dispatch_shift<B, Eval, TPP, true, true>(a, tab, g, b, s)
           : pair_rel_dispatch_shift<B, Eval, TPP, true, false>(a, tab, g, b, s);
    else
        mt ? pair_rel_dispatch_shift<B, Eval, TPP, false, true>(a, tab, g, b, s)
           : pair_rel_dispatch_shift<B, Eval, TPP, false, false>(a, tab, g, b, s);
}

template<class Eval, int TPP>
static void pair_rel_dispatch_tpp(const PairRelArgs& a, const RelTable<Eval>& tab,
                                  unsigned int block, unsigned int batch, hipStream_t s)
{
    unsigned int threads = a.N * TPP;
    unsigned int nblocks = round_up((threads + block - 1) / block, 8);
    dim3 g(nblocks), b(block);
    if (batch == 4)
        pair_rel_dispatch_batch<4, Eval, TPP>(a, tab, g, b, s);
    else
        pair_rel_dispatch_batch<8, Eval, TPP>(a, tab, g, b, s);
}

// Guard bands of the FP32 cutoff decisions, then the launch with the tuned configuration
template<class E>
static hipError_t pair_rel_launch(PairRelArgs a, const RelTable<E>& tab, PairLaunch cfg,
                                  hipStream_t s)
{
    const unsigned int nt = a.ntypes ? a.ntypes : 1;
    // worst-case FP32 error of r^2 near r_cut^2: |rel| <= R = 1.5 w per axis (cell width plus
    // the r_buff / 2 drift allowance, r_buff <= w), each rel rounded (u R), their difference
    // (2 u R), the cell offset w rounded (u w) and the final sum (u (2 R + w))
    const double u = 1.0 / 16777216.0;
    double wmax = 0.0;
    for (int d = 0; d < 3; d++)
        wmax = fmax(wmax, (double)a.box.L[d] / (double)a.dim[d]);
    const double R = 1.5 * wmax;
    const double e_dx = u * (2.0 * R + 2.0 * R + wmax + 2.0 * R + wmax);
    for (unsigned int q = 0; q < nt * nt; q++)
    {
        const double rc2 = a.rcutsq[q];
        const double rc = sqrt(rc2);
        a.guard[q] = (float)(4.0 * (2.0 * sqrt(3.0) * rc * e_dx + 6.0 * u * rc2) + 1e-12);
    }
    const unsigned int block = cfg.block ? cfg.block : 256;
    const unsigned int batch = cfg.batch == 4 ? 4u : 8u;
    switch (cfg.tpp ? cfg.tpp : 4)
    {
    case 1:
        pair_rel_dispatch_tpp<E, 1>(a, tab, block, batch, s);
        break;
    case 2:
        pair_rel_dispatch_tpp<E, 2>(a, tab, block, batch, s);
        break;
    case 8:
        pair_rel_dispatch_tpp<E, 8>(a, tab, block, batch, s);
        break;
    case 16:
        pair_rel_dispatch_tpp<E, 16>(a, tab, block, batch, s);
        break;
    default:
        pair_rel_dispatch_tpp<E, 4>(a, tab, block, batch, s);
        break;
    }
    return hipGetLastError();
}

hipError_t launch_pair_rel_lj(const PairRelArgs& a_in, const hm_lj_params* params,
                              PairLaunch cfg, hipStream_t s)
{
    if (a_in.N == 0)
        return hipSuccess;
    using E = EvaluatorPairLJ<float>;
    RelTable<E> tab;
    PairRelArgs a = a_in;
    const unsigned int nt = a.ntypes ? a.ntypes : 1;
    for (unsigned int q = 0; q < nt * nt; q++)
    {
        tab.p[q] = E::param_type::make(params[q].sigma, params[q].epsilon);
        tab.f[q] = FastPair<E>::make(tab.p[q], a.rcutsq[q], a.shift_mode == 1);
        // lj1 == 0: evalForceAndEnergy never contributes -- a zero cutoff excludes the pair
        if ((double)tab.p[q].epsilon_x_4 * (double)tab.p[q].sigma_6 == 0.0)
        {
            a.rcutsq[q] = 0.0;
            a.rcutsq_f[q] = 0.f;
        }
    }
    return pair_rel_launch<E>(a, tab, cfg, s);
}

hipError_t launch_pair_rel_table(const PairRelArgs& a_in, const hm_table_params* params,
                                 PairLaunch cfg, hipStream_t s)
{
    if (a_in.N == 0)
        return hipSuccess;
    using E = EvaluatorPairTable<float>;
    RelTable<E> tab;
    PairRelArgs a = a_in;
    a.shift_mode = 0;
    const unsigned int nt = a.ntypes ? a.ntypes : 1;
    for (unsigned int q = 0; q < nt * nt; q++)
    {
        tab.p[q] = E::param_type::make(params[q].U, params[q].F, params[q].r_min,
                                       params[q].width);
        tab.f[q] = FastPair<E>::make(tab.p[q], a.rcutsq[q], 0);
        // an empty table never contributes -- a zero cutoff excludes the pair
        if (params[q].width == 0)
        {
            a.rcutsq[q] = 0.0;
            a.rcutsq_f[q] = 0.f;
        }
    }
    return pair_rel_launch<E>(a, tab, cfg, s);
}

// ======================================================================== pair forces, FP32 cell tiles
// PotentialPair::computeForces (PotentialPair.h:587-794) organised per home cell. The workgroup
// stages the cell-slot relative positions of the 27 adjacent cells (adjacency order, slots in
// order -- the candidate sequence of the neighbour-list build) in LDS, each shifted by
// -offset * w into the home cell's frame, so a pair vector is r_i - r_j with both operands in
// LDS. Rows are the build's 16-bit local rows, read as B contiguous entries per lane. The
// cutoff is decided as in pair_rel_kernel: FP32 with FP64 re-decision inside the guard band.
constexpr unsigned int TCAP = 1024;
constexpr unsigned int TBLOCK = 256;

struct TileShared
{
    float4 sp[TCAP];          // staged home-frame positions; w = packed cell | type
    unsigned int sidx[TCAP];  // particle index of each staged candidate
    float4 ash[27];           // -offset * w per adjacency slot
    unsigned int aoff[28], acell[27];
    unsigned int qhome;
};

// candidate l of the home cell's sequence; staged in LDS when l < TCAP
template<bool ALL>
__device__ __forceinline__ void tile_cand(const PairTileArgs& a, const TileShared& S,
                                          unsigned int l, float4& r, unsigned int& j)
{
    if (ALL || l < TCAP)
    {
        r = S.sp[l];
        j = S.sidx[l];
        return;
    }
    const unsigned int q = find_segment(S.aoff, 0, a.n_adj, l);
    const unsigned int slot = S.acell[q] * a.cmax + (l - S.aoff[q]);
    const float3 rr = a.rel[slot];
    const float4 sh = S.ash[q];
    r = make_float4(rr.x + sh.x, rr.y + sh.y, rr.z + sh.z, __uint_as_float(a.relw[slot]));
    j = a.cell_idx[slot];
}

// CH contiguous 16-bit row entries of one lane, loaded as CH / 8 16-byte vectors
template<int CH> struct RowChunk
{
    uint4 v[CH / 8];
};

template<int CH>
__device__ __forceinline__ void load_chunk(const uint16_t* p, RowChunk<CH>& c)
{
#pragma unroll
    for (int w = 0; w < CH / 8; w++)
        c.v[w] = reinterpret_cast<const uint4*>(p)[w];
}

template<int CH>
__device__ __forceinline__ unsigned int chunk_entry(const RowChunk<CH>& c, int q)
{
    const uint4 v = c.v[q >> 3];
    const int d = (q & 7) >> 1;
    const unsigned int w = d == 0 ? v.x : (d == 1 ? v.y : (d == 2 ? v.z : v.w));
    return (q & 1) ? (w >> 16) : (w & 0xffffu);
}

// the entries [k0 + sub * CH, k0 + sub * CH + CH) of one home particle's row
template<bool ALL, int CH, class Eval, bool VIRIAL, bool MT, int SHIFT>
__device__ __forceinline__ void tile_entries(const PairTileArgs& a, const RelTable<Eval>& tab,
                                             const TileShared& S, const float4& ri,
                                             unsigned int i, unsigned int ti, unsigned int n,
                                             unsigned int kb, const RowChunk<CH>& c,
                                             unsigned int lmax, unsigned int jmax, PairAcc& acc)
{
#pragma unroll
    for (int g = 0; g < CH; g += 4)
    {
        float dx[4], dy[4], dz[4], rsq[4];
        unsigned int tp[4], jj[4];
        bool in[4];
        unsigned int chk = 0;
#pragma unroll
        for (int u = 0; u < 4; u++)
        {
            float4 rj;
            tile_cand<ALL>(a, S, min(chunk_entry<CH>(c, g + u), lmax), rj, jj[u]);
            dx[u] = ri.x - rj.x;
            dy[u] = ri.y - rj.y;
            dz[u] = ri.z - rj.z;
            rsq[u] = dx[u] * dx[u] + dy[u] * dy[u] + dz[u] * dz[u];
            tp[u] = MT ? ti * a.ntypes + (__float_as_uint(rj.w) >> 24) : 0u;
            const bool valid = kb + g + u < n;
            const float t = rsq[u] - a.rcutsq_f[tp[u]];
            in[u] = valid && t < 0.f;
            chk |= (valid && fabsf(t) <= a.guard[tp[u]]) ? (1u << u) : 0u;
        }
        if (chk)
        {
#pragma unroll
            for (int u = 0; u < 4; u++)
                if (chk & (1u << u))
                {
#pragma clang fp contract(off)
                    // re-decide with the FP64 formula of the CPU path
                    const double4 pi = a.pos[i];
                    const double4 pj = a.pos[min(jj[u], jmax)];
                    double ddx = pi.x - pj.x, ddy = pi.y - pj.y, ddz = pi.z - pj.z;
                    min_image(a.box, ddx, ddy, ddz);
                    in[u] = ddx * ddx + ddy * ddy + ddz * ddz < a.rcutsq[tp[u]];
                }
        }
#pragma unroll
        for (int u = 0; u < 4; u++)
        {
            float fdr, eh;
            rel_pair_eval<Eval, SHIFT>(a.rcutsq_f, a.ronsq_f, tab, tp[u], rsq[u], in[u], fdr,
                                       eh);
            acc.fx += dx[u] * fdr;
            acc.fy += dy[u] * fdr;
            acc.fz += dz[u] * fdr;
            acc.pe += eh;
            if (VIRIAL)
            {
                const float f2 = 0.5f * fdr;
                acc.vxx += f2 * dx[u] * dx[u];
                acc.vxy += f2 * dx[u] * dy[u];
                acc.vxz += f2 * dx[u] * dz[u];
                acc.vyy += f2 * dy[u] * dy[u];
                acc.vyz += f2 * dy[u] * dz[u];
                acc.vzz += f2 * dz[u] * dz[u];
            }
        }
    }
}

// One home particle (TPP lanes): chunk 0 of its row has been loaded into c0 already.
template<bool ALL, int CH, class Eval, int TPP, bool VIRIAL, bool MT, int SHIFT>
__device__ __forceinline__ void tile_particle(const PairTileArgs& a, const RelTable<Eval>& tab,
                                              const TileShared& S, bool active, unsigned int hs,
                                              unsigned int i, unsigned int n,
                                              const RowChunk<CH>& c0, unsigned int total)
{
    const unsigned int sub = threadIdx.x % TPP;
    const unsigned int jmax = a.N - 1;
    const unsigned int lmax = total - 1;
    float4 ri;
    unsigned int idum;
    tile_cand<ALL>(a, S, S.aoff[S.qhome] + (active ? hs : 0u), ri, idum);
    const unsigned int ti = MT ? (__float_as_uint(ri.w) >> 24) : 0u;
    PairAcc acc;
    tile_entries<ALL, CH, Eval, VIRIAL, MT, SHIFT>(a, tab, S, ri, i, ti, n, sub * CH, c0, lmax,
                                                   jmax, acc);
    const uint16_t* row = a.lrow + (size_t)i * a.nmax + sub * CH;
    for (unsigned int k0 = TPP * CH; k0 < n; k0 += TPP * CH)
    {
        RowChunk<CH> c;
        load_chunk<CH>(row + k0, c);
        tile_entries<ALL, CH, Eval, VIRIAL, MT, SHIFT>(a, tab, S, ri, i, ti, n, k0 + sub * CH, c,
                                                       lmax, jmax, acc);
    }
#pragma unroll
    for (int off = TPP / 2; off > 0; off >>= 1)
    {
        acc.fx += __shfl_xor(acc.fx, off, TPP);
        acc.fy += __shfl_xor(acc.fy, off, TPP);
        acc.fz += __shfl_xor(acc.fz, off, TPP);
        acc.pe += __shfl_xor(acc.pe, off, TPP);
        if (VIRIAL)
        {
            acc.vxx += __shfl_xor(acc.vxx, off, TPP);
            acc.vxy += __shfl_xor(acc.vxy, off, TPP);
            acc.vxz += __shfl_xor(acc.vxz, off, TPP);
            acc.vyy += __shfl_xor(acc.vyy, off, TPP);
            acc.vyz += __shfl_xor(acc.vyz, off, TPP);
            acc.vzz += __shfl_xor(acc.vzz, off, TPP);
        }
    }
    if (active && sub == 0)
    {
        a.force[i] = make_float4(acc.fx, acc.fy, acc.fz, acc.pe);
        if (VIRIAL)
        {
            const size_t P = a.N;
            a.virial[0 * P + i] = acc.vxx;
            a.virial[1 * P + i] = acc.vxy;
            a.virial[2 * P + i] = acc.vxz;
            a.virial[3 * P + i] = acc.vyy;
            a.virial[4 * P + i] = acc.vyz;
            a.virial[5 * P + i] = acc.vzz;
        }
    }
}

template<int B, class Eval, int TPP, bool VIRIAL, bool MT, int SHIFT>
__global__ __launch_bounds__(TBLOCK) void pair_tile_kernel(PairTileArgs a, RelTable<Eval> tab)
{
    constexpr int CH = 2 * B;  // row entries per lane per chunk
    __shared__ TileShared S;
    const unsigned int tid = threadIdx.x;
    const unsigned int sub = tid % TPP;
    const XcdRange r = xcd_range(a.ncell);
    for (unsigned int cell = r.begin; cell < r.end; cell += r.stride)
    {
        const unsigned int nhome = min(a.cell_size[cell], a.cmax);  // the same for every lane
        if (nhome == 0)
            continue;
        // the first pass's particles and row chunks are requested before the staging, so
        // their latency overlaps it
        const unsigned int hs0 = tid / TPP;
        const bool active0 = hs0 < nhome;
        const unsigned int i0 = active0 ? min(a.cell_idx[(size_t)cell * a.cmax + hs0], a.N - 1)
                                        : 0u;
        const unsigned int n0 = active0 ? min(a.n_neigh[i0], a.nmax) : 0u;
        RowChunk<CH> c0;
        load_chunk<CH>(a.lrow + (size_t)i0 * a.nmax + sub * CH, c0);
        if (tid < 64)
        {
            unsigned int size = 0;
            if (tid < a.n_adj)
            {
                const unsigned int c = a.cell_adj[(size_t)cell * a.n_adj + tid];
                size = min(a.cell_size[c], a.cmax);
                S.acell[tid] = c;
                // -(home - neighbour) * w, the home-cell frame shift of this neighbour cell
                const int hx = (int)(cell % a.dim[0]), hy = (int)((cell / a.dim[0]) % a.dim[1]),
                          hz = (int)(cell / (a.dim[0] * a.dim[1]));
                const int nx = (int)(c % a.dim[0]), ny = (int)((c / a.dim[0]) % a.dim[1]),
                          nz = (int)(c / (a.dim[0] * a.dim[1]));
                const int ox = wrap_offset(hx - nx, (int)a.dim[0]);
                const int oy = wrap_offset(hy - ny, (int)a.dim[1]);
                const int oz = wrap_offset(hz - nz, (int)a.dim[2]);
                S.ash[tid] = make_float4(-(float)ox * a.wf[0], -(float)oy * a.wf[1],
                                         -(float)oz * a.wf[2], 0.f);
                if (c == cell)
                    S.qhome = tid;
            }
            // inclusive prefix sum of the cell sizes over the wave
            unsigned int incl = size;
#pragma unroll
            for (int off = 1; off < 64; off <<= 1)
            {
                const unsigned int v = __shfl_up(incl, off, 64);
                if (tid >= (unsigned int)off)
                    incl += v;
            }
            if (tid < a.n_adj)
                S.aoff[tid + 1] = incl;
            if (tid == 0)
                S.aoff[0] = 0;
        }
        __syncthreads();
        const unsigned int total = S.aoff[a.n_adj];
        const unsigned int nst = min(total, TCAP);
        for (unsigned int t = tid; t < nst; t += TBLOCK)
        {
            const unsigned int q = find_segment(S.aoff, 0, a.n_adj, t);
            const unsigned int slot = S.acell[q] * a.cmax + (t - S.aoff[q]);
            const float3 rr = a.rel[slot];
            const float4 sh = S.ash[q];
            S.sp[t] = make_float4(rr.x + sh.x, rr.y + sh.y, rr.z + sh.z,
                                  __uint_as_float(a.relw[slot]));
            S.sidx[t] = a.cell_idx[slot];
        }
        __syncthreads();
        const bool all = total <= TCAP;
        if (all)
            tile_particle<true, CH, Eval, TPP, VIRIAL, MT, SHIFT>(a, tab, S, active0, hs0, i0, n0,
                                                                  c0, total);
        else
            tile_particle<false, CH, Eval, TPP, VIRIAL, MT, SHIFT>(a, tab, S, active0, hs0, i0,
                                                                   n0, c0, total);
        // cells holding more than TBLOCK / TPP particles: further passes, rows loaded inline
        for (unsigned int s0 = TBLOCK / TPP; s0 < nhome; s0 += TBLOCK / TPP)
        {
            const unsigned int hs = s0 + tid / TPP;
            const bool active = hs < nhome;
            const unsigned int i = active ? min(a.cell_idx[(size_t)cell * a.cmax + hs], a.N - 1)
                                          : 0u;
            const unsigned int n = active ? min(a.n_neigh[i], a.nmax) : 0u;
            RowChunk<CH> c;
            load_chunk<CH>(a.lrow + (size_t)i * a.nmax + sub * CH, c);
            if (all)
                tile_particle<true, CH, Eval, TPP, VIRIAL, MT, SHIFT>(a, tab, S, active, hs, i, n,
                                                                      c, total);
            else
                tile_particle<false, CH, Eval, TPP, VIRIAL, MT, SHIFT>(a, tab, S, active, hs, i,
                                                                       n, c, total);
        }
        __syncthreads();
    }
}

template<int B, class Eval, int TPP, bool VIRIAL, bool MT>
static void pair_tile_dispatch_shift(const PairTileArgs& a, const RelTable<Eval>& tab,
                                     unsigned int grid, hipStream_t s)
{
    switch (a.shift_mode)
    {
    case 0:
        hipLaunchKernelGGL((pair_tile_kernel<B, Eval, TPP, VIRIAL, MT, 0>), dim3(grid),
                           dim3(TBLOCK), 0, s, a, tab);
        break;
    case 1:
        hipLaunchKernelGGL((pair_tile_kernel<B, Eval, TPP, VIRIAL, MT, 1>), dim3(grid),
                           dim3(TBLOCK), 0, s, a, tab);
        break;
    default:
        hipLaunchKernelGGL((pair_tile_kernel<B, Eval, TPP, VIRIAL, MT, 2>), dim3(grid),
                           dim3(TBLOCK), 0, s, a, tab);
        break;
    }
}

template<int B, class Eval, int TPP>
static void pair_tile_dispatch_batch(const PairTileArgs& a, const RelTable<Eval>& tab,
                                     unsigned int grid, hipStream_t s)
{
    const bool mt = a.ntypes > 1;
    if (a.compute_virial)
        mt ? pair_tile_dispatch_shift<B, Eval, TPP, true, true>(a, tab, grid, s)
           : pair_tile_dispatch_shift<B, Eval, TPP, true, false>(a, tab, grid, s);
    else
        mt ? pair_tile_dispatch_shift<B, Eval, TPP, false, true>(a, tab, grid, s)
           : pair_tile_dispatch_shift<B, Eval, TPP, false, false>(a, tab, grid, s);
}

template<class Eval, int TPP>
static void pair_tile_dispatch_tpp(const PairTileArgs& a, const RelTable<Eval>& tab,
                                   unsigned int batch, unsigned int grid, hipStream_t s)
{
    if (batch == 4)
        pair_tile_dispatch_batch<4, Eval, TPP>(a, tab, grid, s);
    else
        pair_tile_dispatch_batch<8, Eval, TPP>(a, tab, grid, s);
}

hipError_t launch_pair_tile_lj(const PairTileArgs& a_in, const hm_lj_params* params,
                               PairLaunch cfg, hipStream_t s)
{
    if (a_in.N == 0 || a_in.ncell == 0)
        return hipSuccess;
    using E = EvaluatorPairLJ<float>;
    RelTable<E> tab;
    PairTileArgs a = a_in;
    const unsigned int nt = a.ntypes ? a.ntypes : 1;
    for (unsigned int q = 0; q < nt * nt; q++)
    {
        tab.p[q] = E::param_type::make(params[q].sigma, params[q].epsilon);
        tab.f[q] = FastPair<E>::make(tab.p[q], a.rcutsq[q], a.shift_mode == 1);
        if ((double)tab.p[q].epsilon_x_4 * (double)tab.p[q].sigma_6 == 0.0)
        {
            a.rcutsq[q] = 0.0;
            a.rcutsq_f[q] = 0.f;
        }
    }
    // worst-case FP32 error of r^2 near r_cut^2 (rel rounding, the home-frame shift and the
    // difference, |rel| <= 1.5 w): below that of the gather form, whose bound is used
    const double u = 1.0 / 16777216.0;
    double wmax = 0.0;
    for (int d = 0; d < 3; d++)
        wmax = fmax(wmax, (double)a.box.L[d] / (double)a.dim[d]);
    const double R = 1.5 * wmax;
    const double e_dx = u * (2.0 * R + 2.0 * R + wmax + 2.0 * R + wmax);
    for (unsigned int q = 0; q < nt * nt; q++)
    {
        const double rc2 = a.rcutsq[q];
        const double rc = sqrt(rc2);
        a.guard[q] = (float)(4.0 * (2.0 * sqrt(3.0) * rc * e_dx + 6.0 * u * rc2) + 1e-12);
    }
    const unsigned int grid = round_up(min(cfg.block ? cfg.block : 2048u, a.ncell), 8);
    const unsigned int batch = cfg.batch == 4 ? 4u : 8u;
    switch (cfg.tpp ? cfg.tpp : 8)
    {
    case 4:
        pair_tile_dispatch_tpp<E, 4>(a, tab, batch, grid, s);
        break;
    case 16:
        pair_tile_dispatch_tpp<E, 16>(a, tab, batch, grid, s);
        break;
    default:
        pair_tile_dispatch_tpp<E, 8>(a, tab, batch, grid, s);
        break;
    }
    return hipGetLastError();
}

// ======================================================================== cell-tiled pair forces
// The same sums as pair_kernel, organised per home cell: the workgroup stages the current FP64
// positions of the 27 adjacent cells' particles in LDS (shifted into the home cell's periodic
// frame, so no per-pair minimum image is needed), then TPP lanes per home particle walk the
// particle's 16-bit local row, which indexes that staged sequence. Each position is fetched from
// memory once per adjacent home cell instead of once per neighbour pair, and the row stream is
// half the size of the 32-bit list. Cells whose candidates exceed PC_CAP fall back to global
// gathers with the full minimum image.
constexpr int PC_CAP = 1024;
constexpr int PC_BLOCK = 256;

template<class Eval, int TPP, bool VIRIAL, bool MT, int SHIFT>
__global__ __launch_bounds__(PC_BLOCK) void pair_cell_kernel(PairCellArgs a, ParamTable<Eval> tab)
{
    __shared__ double sx[PC_CAP], sy[PC_CAP], sz[PC_CAP];
    __shared__ unsigned char stp[MT ? PC_CAP : 1];
    __shared__ unsigned int acell[27], asize[27], aoff[27];
    __shared__ unsigned int s_home_off, s_total;
    const unsigned int tid = threadIdx.x;
    const XcdRange range = xcd_range(a.ncell);
    for (unsigned int cell = range.begin; cell < range.end; cell += range.stride)
    {
        const unsigned int nhome = min(a.cell_size[cell], a.cmax);
        if (nhome == 0)
            continue;
        if (tid < a.n_adj)
        {
            const unsigned int c = a.cell_adj[(size_t)cell * a.n_adj + tid];
            acell[tid] = c;
            asize[tid] = min(a.cell_size[c], a.cmax);
        }
        __syncthreads();
        if (tid == 0)
        {
            unsigned int tot = 0, home = 0;
            for (unsigned int q = 0; q < a.n_adj; q++)
            {
                aoff[q] = tot;
                if (acell[q] == cell)
                    home = tot;
                tot += asize[q];
            }
            s_total = tot;
            s_home_off = home;
        }
        __syncthreads();
        const unsigned int total = s_total, home_off = s_home_off;
        const bool staged = total <= (unsigned int)PC_CAP;
        double cx, cy, cz;
        cell_centre(a.box, a.dim, cell, cx, cy, cz);
        if (staged)
        {
            for (unsigned int k = tid; k < total; k += PC_BLOCK)
            {
#pragma clang fp contract(off)
                const unsigned int q = find_segment(aoff, 0, a.n_adj, k);
                const unsigned int idx = a.cell_idx[(size_t)acell[q] * a.cmax + (k - aoff[q])];
                const double4 p = a.pos[idx];
                sx[k] = p.x - a.box.L[0] * rint((p.x - cx) * a.box.Linv[0]);
                sy[k] = p.y - a.box.L[1] * rint((p.y - cy) * a.box.Linv[1]);
                sz[k] = p.z - a.box.L[2] * rint((p.z - cz) * a.box.Linv[2]);
                if (MT)
                    stp[k] = (unsigned char)p.w;
            }
        }
        __syncthreads();
        for (unsigned int hb = 0; hb < nhome; hb += PC_BLOCK / TPP)
        {
            const unsigned int hp = hb + tid / TPP, sub = tid % TPP;
            const bool active = hp < nhome;
            const unsigned int i = a.cell_idx[(size_t)cell * a.cmax + (active ? hp : 0u)];
            double xi, yi, zi;
            unsigned int ti = 0;
            if (staged)
            {
                const unsigned int kh = home_off + (active ? hp : 0u);
                xi = sx[kh];
                yi = sy[kh];
                zi = sz[kh];
                if (MT)
                    ti = stp[kh];
            }
            else
            {
                const double4 p = a.pos[i];
                xi = p.x;
                yi = p.y;
                zi = p.z;
                ti = MT ? (unsigned int)p.w : 0u;
            }
            const unsigned int n = active ? min(a.n_neigh[i], a.nmax) : 0u;
            const unsigned short* row = a.lrow + (size_t)i * a.nmax;
            float fx = 0.f, fy = 0.f, fz = 0.f, pe = 0.f;
            float vxx = 0.f, vxy = 0.f, vxz = 0.f, vyy = 0.f, vyz = 0.f, vzz = 0.f;
            for (unsigned int k0 = 0; k0 < n; k0 += TPP * PAIR_BATCH)
            {
                double px[PAIR_BATCH], py[PAIR_BATCH], pz[PAIR_BATCH];
                unsigned int tj[PAIR_BATCH];
                bool val[PAIR_BATCH];
                unsigned int lr[PAIR_BATCH];
#pragma unroll
                for (int q = 0; q < PAIR_BATCH; q++)
                    lr[q] = row[min(k0 + q * TPP + sub, a.nmax - 1)];
                pin_loads<PAIR_BATCH>(lr);
#pragma unroll
                for (int q = 0; q < PAIR_BATCH; q++)
                {
                    const unsigned int e = k0 + q * TPP + sub;
                    val[q] = e < n;
                    // local rows index [0, total); clamp so a stale row cannot leave the tables
                    const unsigned int l = val[q] ? min(lr[q], total - 1) : 0u;
                    if (staged)
                    {
                        const unsigned int ls = val[q] ? l : home_off + (active ? hp : 0u);
                        px[q] = sx[ls];
                        py[q] = sy[ls];
                        pz[q] = sz[ls];
                        tj[q] = MT ? stp[ls] : 0u;
                    }
                    else
                    {
#pragma clang fp contract(off)
                        const unsigned int qq = find_segment(aoff, 0, a.n_adj, l);
                        const unsigned int j =
                            val[q] ? a.cell_idx[(size_t)acell[qq] * a.cmax + (l - aoff[qq])] : i;
                        const double4 pj = a.pos[j];
                        double ddx = xi - pj.x, ddy = yi - pj.y, ddz = zi - pj.z;
                        min_image(a.box, ddx, ddy, ddz);
                        px[q] = xi - ddx;
                        py[q] = yi - ddy;
                        pz[q] = zi - ddz;
                        tj[q] = MT ? (unsigned int)pj.w : 0u;
                    }
                }
#pragma unroll
                for (int q = 0; q < PAIR_BATCH; q++)
                {
                    double dx, dy, dz, rsq;
                    {
#pragma clang fp contract(off)
                        dx = xi - px[q];
                        dy = yi - py[q];
                        dz = zi - pz[q];
                        rsq = dx * dx + dy * dy + dz * dz;
                    }
                    const unsigned int tp = MT ? ti * a.ntypes + tj[q] : 0u;
                    const bool in = val[q] && (rsq < a.rcutsq[tp]);
                    const float rsqf = in ? (float)rsq : 1.0f;
                    const float rcf = a.rcutsq_f[tp];
                    Eval ev(rsqf, rcf, tab.p[tp]);
                    float fdr = 0.f, e = 0.f;
                    const bool shift = (SHIFT == 1) || (SHIFT == 2 && a.ronsq_f[tp] > rcf);
                    ev.evalForceAndEnergy(fdr, e, shift);
                    if (SHIFT == 2)
                    {
                        const float ron = a.ronsq_f[tp];
                        if (rsqf >= ron && rsqf < rcf)
                        {
                            float den = 1.f / ((rcf - ron) * (rcf - ron) * (rcf - ron));
                            float rmc = rsqf - rcf;
                            float sw = rmc * rmc * (rcf + 2.f * rsqf - 3.f * ron) * den;
                            float ds = 12.f * (rsqf - ron) * rmc * den;
                            float e0 = e;
                            e = e0 * sw;
                            fdr = sw * fdr - ds * e0;
                        }
                    }
                    fdr = in ? fdr : 0.f;
                    e = in ? e : 0.f;
                    const float dxf = (float)dx, dyf = (float)dy, dzf = (float)dz;
                    fx += dxf * fdr;
                    fy += dyf * fdr;
                    fz += dzf * fdr;
                    pe += 0.5f * e;
                    if (VIRIAL)
                    {
                        const float f2 = 0.5f * fdr;
                        vxx += f2 * dxf * dxf;
                        vxy += f2 * dxf * dyf;
                        vxz += f2 * dxf * dzf;
                        vyy += f2 * dyf * dyf;
                        vyz += f2 * dyf * dzf;
                        vzz += f2 * dzf * dzf;
                    }
                }
            }
            if (TPP > 1)
            {
#pragma unroll
                for (int off = TPP / 2; off > 0; off >>= 1)
                {
                    fx += __shfl_xor(fx, off, TPP);
                    fy += __shfl_xor(fy, off, TPP);
                    fz += __shfl_xor(fz, off, TPP);
                    pe += __shfl_xor(pe, off, TPP);
                    if (VIRIAL)
                    {
                        vxx += __shfl_xor(vxx, off, TPP);
                        vxy += __shfl_xor(vxy, off, TPP);
                        vxz += __shfl_xor(vxz, off, TPP);
                        vyy += __shfl_xor(vyy, off, TPP);
                        vyz += __shfl_xor(vyz, off, TPP);
                        vzz += __shfl_xor(vzz, off, TPP);
                    }
                }
            }
            if (active && sub == 0)
            {
                a.force[i] = make_float4(fx, fy, fz, pe);
                if (VIRIAL)
                {
                    const size_t P = a.N;
                    a.virial[0 * P + i] = vxx;
                    a.virial[1 * P + i] = vxy;
                    a.virial[2 * P + i] = vxz;
                    a.virial[3 * P + i] = vyy;
                    a.virial[4 * P + i] = vyz;
                    a.virial[5 * P + i] = vzz;
                }
            }
        }
        __syncthreads();
    }
}

template<class Eval, int TPP, bool VIRIAL, bool MT>
static void pair_cell_shift(const PairCellArgs& a, const ParamTable<Eval>& tab, dim3 g,
                            hipStream_t s)
{
    const dim3 b(PC_BLOCK);
    switch (a.shift_mode)
    {
    case 0:
        hipLaunchKernelGGL((pair_cell_kernel<Eval, TPP, VIRIAL, MT, 0>), g, b, 0, s, a, tab);
        break;
    case 1:
        hipLaunchKernelGGL((pair_cell_kernel<Eval, TPP, VIRIAL, MT, 1>), g, b, 0, s, a, tab);
        break;
    default:
        hipLaunchKernelGGL((pair_cell_kernel<Eval, TPP, VIRIAL, MT, 2>), g, b, 0, s, a, tab);
        break;
    }
}

template<class Eval, int TPP>
static void pair_cell_tpp(const PairCellArgs& a, const ParamTable<Eval>& tab, dim3 g,
                          hipStream_t s)
{
    const bool mt = a.ntypes > 1;
    if (a.compute_virial)
        mt ? pair_cell_shift<Eval, TPP, true, true>(a, tab, g, s)
           : pair_cell_shift<Eval, TPP, true, false>(a, tab, g, s);
    else
        mt ? pair_cell_shift<Eval, TPP, false, true>(a, tab, g, s)
           : pair_cell_shift<Eval, TPP, false, false>(a, tab, g, s);
}

hipError_t launch_pair_cell_lj(const PairCellArgs& a, const hm_lj_params* params, PairLaunch cfg,
                               hipStream_t s)
{
    if (a.N == 0 || a.ncell == 0)
        return hipSuccess;
    using E = EvaluatorPairLJ<float>;
    ParamTable<E> tab;
    const unsigned int nt = a.ntypes ? a.ntypes : 1;
    for (unsigned int q = 0; q < nt * nt; q++)
        tab.p[q] = E::param_type::make(params[q].sigma, params[q].epsilon);
    // cfg.block is the number of workgroups for the persistent cell loop (0: one per cell)
    const unsigned int grid = round_up(cfg.block ? min(a.ncell, cfg.block) : a.ncell, 8);
    const dim3 g(grid);
    switch (cfg.tpp ? cfg.tpp : 8)
    {
    case 4:
        pair_cell_tpp<E, 4>(a, tab, g, s);
        break;
    case 16:
        pair_cell_tpp<E, 16>(a, tab, g, s);
        break;
    default:
        pair_cell_tpp<E, 8>(a, tab, g, s);
        break;
    }
    return hipGetLastError();
}

// ======================================================================== bonds
// PotentialBond::computeForces (hoomd/md/PotentialBond.h:167-315) in the per-particle form of
// gpu_compute_bond_forces_kernel (hoomd/md/PotentialBondGPU.cuh:98-239). Bonds are held by tag
// (CSR), so a Hilbert sort only changes the tag->index map.
template<class Eval> struct BondTable
{
    typename Eval::param_type p[HM_MAX_TYPES];
};

template<class Eval>
__global__ __launch_bounds__(256) void bond_kernel(BondKernelArgs a, BondTable<Eval> tab)
{
    unsigned int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.N)
        return;
    // partners in bond order: from the index-ordered table (all of its loads issued together),
    // else through tag -> CSR row -> partner tag -> index
    constexpr unsigned int MAXB = 4;
    unsigned int part[MAXB], ptype[MAXB];
    unsigned int nb = 0;
    if (a.bpart)
    {
        nb = a.max_bonds;
#pragma unroll
        for (unsigned int k = 0; k < MAXB; k++)
        {
            part[k] = k < nb ? a.bpart[(size_t)k * a.N + i] : NO_SLOT;
            ptype[k] = (k < nb && a.bptype) ? a.bptype[(size_t)k * a.N + i] : 0u;
        }
    }
    const unsigned int tg = a.bpart ? 0u : a.tag[i];
    const unsigned int b0 = a.bpart ? 0u : a.row[tg];
    const unsigned int b1 = a.bpart ? nb : a.row[tg + 1];
    const double4 pi = a.pos[i];
    float fx = 0.f, fy = 0.f, fz = 0.f, pe = 0.f;
    float v[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    bool bad = false;
    for (unsigned int b = b0; b < b1; b++)
    {
        unsigned int j, bt;
        if (a.bpart)
        {
            // constant indices into part / ptype keep them in registers
            j = part[0];
            bt = ptype[0];
#pragma unroll
            for (unsigned int k = 1; k < MAXB; k++)
                if (b == k)
                {
                    j = part[k];
                    bt = ptype[k];
                }
            if (j == NO_SLOT)
                break;
        }
        else
        {
            j = a.rtag[a.other_tag[b]];
            bt = a.btype ? a.btype[b] : 0u;
        }
        const double4 pj = a.pos[j];
        double dx, dy, dz, rsq;
        {
#pragma clang fp contract(off)
            dx = pi.x - pj.x;
            dy = pi.y - pj.y;
            dz = pi.z - pj.z;
            min_image(a.box, dx, dy, dz);
            rsq = dx * dx + dy * dy + dz * dz;
        }
        Eval ev((float)rsq, tab.p[bt]);
        float fdr = 0.f, e = 0.f;
        if (!ev.evalForceAndEnergy(fdr, e))
        {
            bad = true;
            continue;
        }
        const float dxf = (float)dx, dyf = (float)dy, dzf = (float)dz;
        fx += fdr * dxf;
        fy += fdr * dyf;
        fz += fdr * dzf;
        pe += 0.5f * e;
        if (a.compute_virial)
        {
            float f2 = 0.5f * fdr;
            v[0] += dxf * dxf * f2;
            v[1] += dxf * dyf * f2;
            v[2] += dxf * dzf * f2;
            v[3] += dyf * dyf * f2;
            v[4] += dyf * dzf * f2;
            v[5] += dzf * dzf * f2;
        }
    }
    if (bad)
    {
        if (a.flag)
            atomicOr(a.flag, 1u);
        if (a.st)
            atomicOr(&a.st->error, (unsigned int)ERR_BOND);
    }
    a.force[i] = make_float4(fx, fy, fz, pe);
    if (a.compute_virial)
        for (int m = 0; m < 6; m++)
            a.virial[(size_t)m * a.N + i] = v[m];
}

hipError_t launch_bond_fene(const BondKernelArgs& a, const hm_fene_params* params, hipStream_t s)
{
    if (a.N == 0)
        return hipSuccess;
    using E = EvaluatorBondFENE<float>;
    BondTable<E> tab;
    unsigned int nt = a.nbond_types ? a.nbond_types : 1;
    for (unsigned int q = 0; q < nt; q++)
        tab.p[q] = E::param_type::make(params[q].k, params[q].r0, params[q].epsilon,
                                       params[q].sigma, params[q].delta);
    if (a.bpart && a.max_bonds > 4)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(bond_kernel<E>, dim3((a.N + 255) / 256), dim3(256), 0, s, a, tab);
    return hipGetLastError();
}

hipError_t launch_bond_harmonic(const BondKernelArgs& a, const hm_harmonic_params* params,
                                hipStream_t s)
{
    if (a.N == 0)
        return hipSuccess;
    using E = EvaluatorBondHarmonic<float>;
    BondTable<E> tab;
    unsigned int nt = a.nbond_types ? a.nbond_types : 1;
    for (unsigned int q = 0; q < nt; q++)
        tab.p[q] = E::param_type::make(params[q].k, params[q].r0);
    if (a.bpart && a.max_bonds > 4)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(bond_kernel<E>, dim3((a.N + 255) / 256), dim3(256), 0, s, a, tab);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void bond_index_kernel(unsigned int N,
                                                         const unsigned int* __restrict__ tag,
                                                         const unsigned int* __restrict__ rtag,
                                                         const unsigned int* __restrict__ row,
                                                         const unsigned int* __restrict__ other,
                                                         const unsigned int* __restrict__ btype,
                                                         unsigned int max_bonds,
                                                         unsigned int* bpart, unsigned int* bptype)
{
    const unsigned int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N)
        return;
    const unsigned int tg = tag[i];
    const unsigned int b0 = row[tg], deg = row[tg + 1] - b0;
    for (unsigned int k = 0; k < max_bonds; k++)
    {
        const bool has = k < deg;
        bpart[(size_t)k * N + i] = has ? rtag[other[b0 + k]] : NO_SLOT;
        if (bptype)
            bptype[(size_t)k * N + i] = has && btype ? btype[b0 + k] : 0u;
    }
}

hipError_t launch_bond_index(unsigned int N, const unsigned int* tag, const unsigned int* rtag,
                             const unsigned int* row, const unsigned int* other_tag,
                             const unsigned int* btype, unsigned int max_bonds,
                             unsigned int* bpart, unsigned int* bptype, hipStream_t s)
{
    if (N == 0 || max_bonds == 0)
        return hipSuccess;
    if (max_bonds > 4 || !tag || !rtag || !row || !other_tag || !bpart)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(bond_index_kernel, dim3((N + 255) / 256), dim3(256), 0, s, N, tag, rtag,
                       row, other_tag, btype, max_bonds, bpart, bptype);
    return hipGetLastError();
}

// ======================================================================== integration
// TwoStepConstantVolume::integrateStepOne (hoomd/md/TwoStepConstantVolume.cc:7-80), fused with
// NeighborList::distanceCheck for step t + 1 (NeighborList.cc:1220-1289) and with clearing the
// cell occupancy counters for a possible rebuild.
__global__ __launch_bounds__(256) void step_one_kernel(StepOneArgs a)
{
#pragma clang fp contract(off)
    const unsigned int gtid = blockIdx.x * blockDim.x + threadIdx.x;
    if (a.cell_size)
        for (unsigned int c = gtid; c < a.ncell; c += gridDim.x * blockDim.x)
            a.cell_size[c] = 0;
    const unsigned int n = gtid;
    bool moved = false;
    if (n < a.N)
    {
        const double alpha = (a.st && a.use_state_alpha) ? a.st->alpha : a.alpha;
        const double dt = a.dt;
        double4 v = a.vel[n];
        double4 p = a.pos[n];
        const d3 acc = a.accel[n];
        i3 img = a.image[n];
        double vx = v.x + (1.0 / 2.0) * acc.x * dt;
        double vy = v.y + (1.0 / 2.0) * acc.y * dt;
        double vz = v.z + (1.0 / 2.0) * acc.z * dt;
        vx *= alpha;
        vy *= alpha;
        vz *= alpha;
        double px = p.x + dt * vx;
        double py = p.y + dt * vy;
        double pz = p.z + dt * vz;
        wrap_axis(px, img.x, a.box.lo[0], a.box.hi[0], a.box.L[0]);
        wrap_axis(py, img.y, a.box.lo[1], a.box.hi[1], a.box.L[1]);
        wrap_axis(pz, img.z, a.box.lo[2], a.box.hi[2], a.box.L[2]);
        v.x = vx;
        v.y = vy;
        v.z = vz;
        p.x = px;
        p.y = py;
        p.z = pz;
        a.vel[n] = v;
        a.pos[n] = p;
        a.image[n] = img;
        if (a.pos32)
            a.pos32[n] = make_float4((float)px, (float)py, (float)pz, 0.f);
        if (a.rel.rel)
        {
            const float4 r = rel_of(px, py, pz, a.rel.pcell[n], a.box, a.rel);
            a.rel.rel[a.rel.slot_of[n]] = make_float3(r.x, r.y, r.z);
        }
        if (a.last)
        {
            const d3 l = a.last[n];
            double dx = px - 1.0 * l.x, dy = py - 1.0 * l.y, dz = pz - 1.0 * l.z;
            min_image(a.box, dx, dy, dz);
            moved = dx * dx + dy * dy + dz * dz >= a.maxsq;
        }
    }
    if (a.begin && gtid == 0)
        a.begin->timestep = a.begin->next;
    if (a.flag_state)
    {
        // NeighborList::needsUpdating (NeighborList.cc:1325-1418) for step t1: no check before
        // rebuild_check_delay steps have passed since the last build; an unconditional build when
        // the buffer is zero or distance checks are off and the period is hit; else the distance
        // check. The conditions are the same on every lane, so the ballot stays uniform.
        const StepState* st = a.flag_state;
        const unsigned long long t1 = (a.begin ? a.begin->next : st->timestep) + 1;
        const bool should = !(t1 < st->last_update + a.check_delay);
        const bool always = a.maxsq < 0.0
                            || (!a.dist_check
                                && (a.check_delay == 0
                                    || (a.check_delay > 1
                                        && t1 == st->last_update + a.check_delay)));
        if (should && (always || __ballot(moved) != 0) && (threadIdx.x & 63u) == 0)
            atomicMax(&a.flag_state->rebuild_at, t1);
    }
}

hipError_t launch_step_one(const StepOneArgs& a, hipStream_t s)
{
    unsigned int work = a.N > a.ncell ? a.N : a.ncell;
    if (work == 0)
        return hipSuccess;
    unsigned int blocks = (a.N + 255) / 256;
    if (blocks == 0)
        blocks = 1;
    hipLaunchKernelGGL(step_one_kernel, dim3(blocks), dim3(256), 0, s, a);
    return hipGetLastError();
}

// FP64 block sum over 256 threads in a fixed order (deterministic)
__device__ __forceinline__ double block_sum_256(double v, double* sh)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        v += __shfl_xor(v, off, 64);
    const unsigned int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63u) == 0)
        sh[w] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        t = ((sh[0] + sh[1]) + sh[2]) + sh[3];
    return t;
}

// TwoStepConstantVolume::integrateStepTwo (TwoStepConstantVolume.cc:213-260) or
// TwoStepLangevin::integrateStepTwo (hoomd/md/TwoStepLangevin.cc:206-275), with the net force
// summed from up to four force arrays and the FP64 kinetic-energy partial of each block.
template<bool LANGEVIN>
__global__ __launch_bounds__(256) void step_two_kernel(StepTwoArgs a)
{
#pragma clang fp contract(off)
    __shared__ double sh[4];
    const unsigned int n = blockIdx.x * blockDim.x + threadIdx.x;
    double ke = 0.0;
    if (n < a.N)
    {
        double4 v = a.vel[n];
        const double m = v.w;
        const double minv = 1.0 / m;
        double Fx = 0.0, Fy = 0.0, Fz = 0.0;
        for (unsigned int q = 0; q < a.nforces; q++)
        {
            const float4 f = a.f[q][n];
            Fx += (double)f.x;
            Fy += (double)f.y;
            Fz += (double)f.z;
        }
        double ax, ay, az;
        if (LANGEVIN)
        {
            const uint64_t t = a.st ? a.st->timestep : a.timestep;
            Philox g(25, t, a.seed, a.tag[n]);
            const double rx = uniform(g, -1.0, 1.0);
            const double ry = uniform(g, -1.0, 1.0);
            const double rz = uniform(g, -1.0, 1.0);
            const double gm = a.gamma_by_type[(unsigned int)a.pos[n].w];
            const double coeff = sqrt(6.0 * gm * a.kT / a.dt);
            const double bx = rx * coeff - gm * v.x;
            const double by = ry * coeff - gm * v.y;
            const double bz = rz * coeff - gm * v.z;
            ax = (Fx + bx) * minv;
            ay = (Fy + by) * minv;
            az = (Fz + bz) * minv;
            v.x += (1.0 / 2.0) * ax * a.dt;
            v.y += (1.0 / 2.0) * ay * a.dt;
            v.z += (1.0 / 2.0) * az * a.dt;
        }
        else
        {
            ax = Fx * minv;
            ay = Fy * minv;
            az = Fz * minv;
            v.x *= a.alpha;
            v.y *= a.alpha;
            v.z *= a.alpha;
            v.x += (1.0 / 2.0) * a.dt * ax;
            v.y += (1.0 / 2.0) * a.dt * ay;
            v.z += (1.0 / 2.0) * a.dt * az;
        }
        a.vel[n] = v;
        a.accel[n] = d3{ax, ay, az};
        ke = m * (v.x * v.x + v.y * v.y + v.z * v.z);
    }
    if (a.ke_partial)
    {
        double t = block_sum_256(ke, sh);
        if (threadIdx.x == 0)
            a.ke_partial[blockIdx.x] = t;
    }
    if (a.finish && n == 0)
        a.finish->next = a.finish->timestep + 1;
}

hipError_t launch_step_two(const StepTwoArgs& a, hipStream_t s, unsigned int* nblocks_out)
{
    unsigned int blocks = (a.N + 255) / 256;
    if (nblocks_out)
        *nblocks_out = blocks;
    if (blocks == 0)
        return hipSuccess;
    if (a.gamma_by_type)
        hipLaunchKernelGGL(step_two_kernel<true>, dim3(blocks), dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL(step_two_kernel<false>, dim3(blocks), dim3(256), 0, s, a);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void ke_partial_kernel(unsigned int N, const double4* vel,
                                                         double* partial)
{
#pragma clang fp contract(off)
    __shared__ double sh[4];
    const unsigned int n = blockIdx.x * blockDim.x + threadIdx.x;
    double ke = 0.0;
    if (n < N)
    {
        double4 v = vel[n];
        ke = v.w * (v.x * v.x + v.y * v.y + v.z * v.z);
    }
    double t = block_sum_256(ke, sh);
    if (threadIdx.x == 0)
        partial[blockIdx.x] = t;
}

hipError_t launch_ke_partial(unsigned int N, const double4* vel, double* partial,
                             unsigned int* nblocks_out, hipStream_t s)
{
    unsigned int blocks = (N + 255) / 256;
    *nblocks_out = blocks;
    if (blocks == 0)
        return hipSuccess;
    hipLaunchKernelGGL(ke_partial_kernel, dim3(blocks), dim3(256), 0, s, N, vel, partial);
    return hipGetLastError();
}

// Step-begin: advance the device timestep and draw the Bussi factor from the kinetic energy
// of the current velocities (BussiThermostat::getRescalingFactorsOne, Thermostat.h:357-391).
__global__ __launch_bounds__(256) void step_begin_kernel(StepState* st, const double* ke_partial,
                                                         unsigned int npartial, int bussi,
                                                         double ndof, double dt, double tau,
                                                         double kT, uint16_t seed)
{
#pragma clang fp contract(off)
    __shared__ double sh[4];
    double acc = 0.0;
    if (bussi)
        for (unsigned int q = threadIdx.x; q < npartial; q += 256)
            acc += ke_partial[q];
    double tot = block_sum_256(acc, sh);
    if (threadIdx.x == 0)
    {
        const unsigned long long t = st->next;
        st->timestep = t;
        st->next = t + 1;
        double alpha = 1.0;
        if (bussi)
        {
            const double K = tot * 0.5;
            if (K == 0.0)
                st->error |= ERR_ZERO_KE;
            else
                alpha = bussi_alpha(K, ndof, dt, tau, kT, t, seed, 0u);
        }
        st->alpha = alpha;
    }
}

hipError_t launch_step_begin(StepState* st, const double* ke_partial, unsigned int npartial,
                             int bussi, double ndof, double dt, double tau, double kT,
                             uint16_t seed, hipStream_t s)
{
    hipLaunchKernelGGL(step_begin_kernel, dim3(1), dim3(256), 0, s, st, ke_partial, npartial,
                       bussi, ndof, dt, tau, kT, seed);
    return hipGetLastError();
}

__global__ void set_rebuild_kernel(StepState* st)
{
    st->rebuild_at = st->next + 1;
    st->forced_at = st->next + 1;
}

hipError_t launch_set_rebuild(StepState* st, hipStream_t s)
{
    hipLaunchKernelGGL(set_rebuild_kernel, dim3(1), dim3(1), 0, s, st);
    return hipGetLastError();
}

// Integrator::computeAccelerations: a = F_net / m
__global__ void accel_init_kernel(unsigned int N, const double4* vel, d3* accel,
                                  const float4* f0, const float4* f1, const float4* f2,
                                  const float4* f3, unsigned int nforces)
{
#pragma clang fp contract(off)
    unsigned int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N)
        return;
    const float4* fs[4] = {f0, f1, f2, f3};
    double Fx = 0.0, Fy = 0.0, Fz = 0.0;
    for (unsigned int q = 0; q < nforces; q++)
    {
        float4 f = fs[q][n];
        Fx += (double)f.x;
        Fy += (double)f.y;
        Fz += (double)f.z;
    }
    const double minv = 1.0 / vel[n].w;
    accel[n] = d3{Fx * minv, Fy * minv, Fz * minv};
}

hipError_t launch_accel_init(unsigned int N, const double4* vel, d3* accel, const float4* const* f,
                             unsigned int nforces, hipStream_t s)
{
    if (N == 0)
        return hipSuccess;
    const float4* fs[4] = {nullptr, nullptr, nullptr, nullptr};
    for (unsigned int q = 0; q < nforces && q < 4; q++)
        fs[q] = f[q];
    hipLaunchKernelGGL(accel_init_kernel, dim3((N + 255) / 256), dim3(256), 0, s, N, vel, accel,
                       fs[0], fs[1], fs[2], fs[3], nforces);
    return hipGetLastError();
}

// ======================================================================== thermo
// ComputeThermo::computeProperties (hoomd/md/ComputeThermo.cc:76-292) sums in FP64; two passes
// with a fixed block count so the summation order does not depend on timing.
struct ForcePtrs
{
    const float4* f[4];
    const float* v[4];
    unsigned int n;
};

__global__ __launch_bounds__(256) void thermo_partial_kernel(unsigned int N, const double4* vel,
                                                             ForcePtrs fp, double* partial)
{
#pragma clang fp contract(off)
    __shared__ double sh[4];
    double acc[THERMO_SUMS] = {};
    for (unsigned int n = blockIdx.x * blockDim.x + threadIdx.x; n < N;
         n += gridDim.x * blockDim.x)
    {
        double4 v = vel[n];
        acc[0] += v.w * (v.x * v.x + v.y * v.y + v.z * v.z);
        // kinetic part of the pressure tensor, m v_a v_b (ComputeThermo.cc:236-262)
        acc[9] += v.w * v.x * v.x;
        acc[10] += v.w * v.x * v.y;
        acc[11] += v.w * v.x * v.z;
        acc[12] += v.w * v.y * v.y;
        acc[13] += v.w * v.y * v.z;
        acc[14] += v.w * v.z * v.z;
        for (unsigned int q = 0; q < fp.n; q++)
        {
            acc[1] += (double)fp.f[q][n].w;
            if (fp.v[q])
                for (int m = 0; m < 6; m++)
                    acc[3 + m] += (double)fp.v[q][(size_t)m * N + n];
        }
    }
    for (unsigned int k = 0; k < THERMO_SUMS; k++)
    {
        double t = block_sum_256(acc[k], sh);
        if (threadIdx.x == 0)
            partial[blockIdx.x * THERMO_SUMS + k] = t;
    }
}

__global__ __launch_bounds__(256) void thermo_final_kernel(const double* partial,
                                                           unsigned int nblocks, double* out)
{
#pragma clang fp contract(off)
    __shared__ double sh[4];
    for (unsigned int k = 0; k < THERMO_SUMS; k++)
    {
        double acc = 0.0;
        for (unsigned int b = threadIdx.x; b < nblocks; b += 256)
            acc += partial[b * THERMO_SUMS + k];
        double t = block_sum_256(acc, sh);
        if (threadIdx.x == 0)
        {
            if (k == 0)
                out[0] = 0.5 * t;
            else if (k == 1)
                out[1] = t;
            else if (k >= 3)
                out[k] = t;
        }
    }
    if (threadIdx.x == 0)
        out[2] = (1.0 / 3.0) * (out[3] + out[6] + out[8]);
}

hipError_t launch_thermo(unsigned int N, const double4* vel, const float4* const* f,
                         const float* const* vir, unsigned int nforces, double* partial,
                         double* out, hipStream_t s)
{
    ForcePtrs fp;
    fp.n = nforces > 4 ? 4 : nforces;
    for (int q = 0; q < 4; q++)
    {
        fp.f[q] = q < (int)fp.n ? f[q] : nullptr;
        fp.v[q] = (q < (int)fp.n && vir) ? vir[q] : nullptr;
    }
    hipLaunchKernelGGL(thermo_partial_kernel, dim3(THERMO_BLOCKS), dim3(256), 0, s, N, vel, fp,
                       partial);
    hipLaunchKernelGGL(thermo_final_kernel, dim3(1), dim3(256), 0, s, partial, THERMO_BLOCKS, out);
    return hipGetLastError();
}

// ======================================================================== thermalize
// ParticleGroup::thermalizeParticleMomenta (hoomd/ParticleGroup.cc:690-806): Philox
// Seed(35, timestep, seed), Counter(tag); three normals of width sqrt(kT/m); COM momentum
// removed afterwards.
__global__ __launch_bounds__(256) void thermalize_kernel(unsigned int N, double4* vel,
                                                         const unsigned int* tag, double kT,
                                                         uint64_t timestep, uint16_t seed,
                                                         double* partial)
{
#pragma clang fp contract(off)
    __shared__ double sh[4];
    double p[3] = {0, 0, 0};
    for (unsigned int n = blockIdx.x * blockDim.x + threadIdx.x; n < N;
         n += gridDim.x * blockDim.x)
    {
        double4 v = vel[n];
        Philox g(35, timestep, seed, tag[n]);
        double sigma = sqrt(kT / v.w);
        v.x = normal(g, sigma, 0.0);
        v.y = normal(g, sigma, 0.0);
        v.z = normal(g, sigma, 0.0);
        vel[n] = v;
        p[0] += v.w * v.x;
        p[1] += v.w * v.y;
        p[2] += v.w * v.z;
    }
    for (int k = 0; k < 3; k++)
    {
        double t = block_sum_256(p[k], sh);
        if (threadIdx.x == 0)
            partial[blockIdx.x * 3 + k] = t;
    }
}

__global__ __launch_bounds__(256) void remove_com_kernel(unsigned int N, double4* vel,
                                                         const double* partial,
                                                         unsigned int nblocks)
{
#pragma clang fp contract(off)
    double tot[3] = {0, 0, 0};
    for (unsigned int b = 0; b < nblocks; b++)
        for (int k = 0; k < 3; k++)
            tot[k] += partial[b * 3 + k];
    const double cx = tot[0] / (double)N, cy = tot[1] / (double)N, cz = tot[2] / (double)N;
    for (unsigned int n = blockIdx.x * blockDim.x + threadIdx.x; n < N;
         n += gridDim.x * blockDim.x)
    {
        double4 v = vel[n];
        v.x = v.x - cx / v.w;
        v.y = v.y - cy / v.w;
        v.z = v.z - cz / v.w;
        vel[n] = v;
    }
}

hipError_t launch_thermalize(unsigned int N, double4* vel, const unsigned int* tag, double kT,
                             uint64_t timestep, uint16_t seed, double* scratch3, hipStream_t s)
{
    if (N == 0)
        return hipSuccess;
    const unsigned int nb = THERMO_BLOCKS;
    hipLaunchKernelGGL(thermalize_kernel, dim3(nb), dim3(256), 0, s, N, vel, tag, kT, timestep,
                       seed, scratch3);
    hipLaunchKernelGGL(remove_com_kernel, dim3(nb), dim3(256), 0, s, N, vel, scratch3, nb);
    return hipGetLastError();
}

// ======================================================================== Hilbert keys
// SFCPackTuner::generateTraversalOrder (hoomd/SFCPackTuner.cc:290-487) visits the eight
// children of each sub-box in an order that is permuted per child by eight fixed rules. The
// rank of a bin is therefore a base-8 number whose digit at each level is the position of the
// bin's child in the current order; it is computed directly instead of storing the grid^3
// table. Walking tables and rules are the reference's (they define the curve).
__constant__ unsigned char c_child_of[8] = {0, 1, 3, 2, 7, 6, 4, 5}; // (di*4+dj*2+dk) -> child
__constant__ unsigned char c_istep[8] = {0, 0, 0, 0, 1, 1, 1, 1};
__constant__ unsigned char c_jstep[8] = {0, 0, 1, 1, 1, 1, 0, 0};
__constant__ unsigned char c_kstep[8] = {0, 1, 1, 0, 0, 1, 1, 0};
__constant__ unsigned char c_rule[8][8] = {
    {0, 3, 4, 7, 6, 5, 2, 1}, {0, 7, 6, 1, 2, 5, 4, 3}, {0, 7, 6, 1, 2, 5, 4, 3},
    {2, 3, 0, 1, 6, 7, 4, 5}, {2, 3, 0, 1, 6, 7, 4, 5}, {4, 3, 2, 5, 6, 1, 0, 7},
    {4, 3, 2, 5, 6, 1, 0, 7}, {6, 5, 2, 1, 0, 3, 4, 7}};

__device__ unsigned int hilbert_rank(unsigned int ib, unsigned int jb, unsigned int kb,
                                     unsigned int grid)
{
    unsigned char order[8] = {0, 1, 2, 3, 4, 5, 6, 7};
    unsigned int rank = 0;
    unsigned int i0 = 0, j0 = 0, k0 = 0;
    for (unsigned int w = grid >> 1; w >= 1; w >>= 1)
    {
        unsigned int di = (ib - i0) >= w, dj = (jb - j0) >= w, dk = (kb - k0) >= w;
        unsigned char child = c_child_of[di * 4 + dj * 2 + dk];
        unsigned int m = 0;
        while (order[m] != child)
            m++;
        rank = rank * 8 + m;
        i0 += w * c_istep[child];
        j0 += w * c_jstep[child];
        k0 += w * c_kstep[child];
        unsigned char nxt[8];
        for (int q = 0; q < 8; q++)
            nxt[q] = order[c_rule[m][q]];
        for (int q = 0; q < 8; q++)
            order[q] = nxt[q];
        if (w == 1)
            break;
    }
    return rank;
}

// gpu_sfc_bin_particles_kernel (hoomd/SFCPackTunerGPU.cu:26-79): bin = (uint)(f * grid) % grid,
// clamped; key = (rank << 32) | index so a key sort is the stable (rank, index) sort.
__global__ void sfc_keys_kernel(unsigned int N, const double4* pos, DevBox box, unsigned int grid,
                                unsigned long long* keys)
{
#pragma clang fp contract(off)
    unsigned int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N)
        return;
    double4 p = pos[n];
    double f[3] = {(p.x - box.lo[0]) * box.Linv[0], (p.y - box.lo[1]) * box.Linv[1],
                   (p.z - box.lo[2]) * box.Linv[2]};
    int b[3];
    for (int d = 0; d < 3; d++)
    {
        int v = (int)((unsigned int)(f[d] * (double)grid) % grid);
        if (v < 0)
            v = 0;
        if (v >= (int)grid)
            v = (int)grid - 1;
        b[d] = v;
    }
    unsigned int rank = hilbert_rank((unsigned)b[0], (unsigned)b[1], (unsigned)b[2], grid);
    keys[n] = ((unsigned long long)rank << 32) | n;
}

hipError_t launch_sfc_keys(unsigned int N, const double4* pos, const DevBox& box,
                           unsigned int grid, unsigned long long* keys, hipStream_t s)
{
    if (N == 0)
        return hipSuccess;
    hipLaunchKernelGGL(sfc_keys_kernel, dim3((N + 255) / 256), dim3(256), 0, s, N, pos, box, grid,
                       keys);
    return hipGetLastError();
}

__global__ void keys_to_order_kernel(unsigned int N, const unsigned long long* keys,
                                     unsigned int* order)
{
    unsigned int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n < N)
        order[n] = (unsigned int)(keys[n] & 0xffffffffull);
}

hipError_t launch_keys_to_order(unsigned int N, const unsigned long long* keys,
                                unsigned int* order, hipStream_t s)
{
    if (N == 0)
        return hipSuccess;
    hipLaunchKernelGGL(keys_to_order_kernel, dim3((N + 255) / 256), dim3(256), 0, s, N, keys,
                       order);
    return hipGetLastError();
}

__global__ void traversal_kernel(unsigned int grid, unsigned int* traversal)
{
    unsigned int bin = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned int nb = grid * grid * grid;
    if (bin >= nb)
        return;
    unsigned int i = bin / (grid * grid), j = (bin / grid) % grid, k = bin % grid;
    traversal[bin] = hilbert_rank(i, j, k, grid);
}

hipError_t launch_traversal(unsigned int grid, unsigned int* traversal, hipStream_t s)
{
    unsigned int nb = grid * grid * grid;
    hipLaunchKernelGGL(traversal_kernel, dim3((nb + 255) / 256), dim3(256), 0, s, grid, traversal);
    return hipGetLastError();
}

// ======================================================================== permutation helpers
// SFCPackTuner::applySortOrder (hoomd/SFCPackTuner.cc:119-285): dst[n] = src[order[n]]
__global__ void gather_words_kernel(const uint32_t* __restrict__ src, uint32_t* __restrict__ dst,
                                    const unsigned int* __restrict__ order, unsigned int N,
                                    unsigned int words)
{
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t total = (size_t)N * words;
    if (t >= total)
        return;
    unsigned int n = (unsigned int)(t / words), w = (unsigned int)(t % words);
    dst[t] = src[(size_t)order[n] * words + w];
}

hipError_t gather_bytes(const void* src, void* dst, const unsigned int* order, unsigned int N,
                        unsigned int elem_bytes, hipStream_t s)
{
    if (N == 0)
        return hipSuccess;
    unsigned int words = elem_bytes / 4;
    size_t total = (size_t)N * words;
    hipLaunchKernelGGL(gather_words_kernel, dim3((unsigned int)((total + 255) / 256)), dim3(256), 0,
                       s, (const uint32_t*)src, (uint32_t*)dst, order, N, words);
    return hipGetLastError();
}

__global__ void rtag_kernel(unsigned int N, const unsigned int* tag, unsigned int* rtag)
{
    unsigned int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n < N)
        rtag[tag[n]] = n;
}

hipError_t launch_rtag(unsigned int N, const unsigned int* tag, unsigned int* rtag, hipStream_t s)
{
    if (N == 0)
        return hipSuccess;
    hipLaunchKernelGGL(rtag_kernel, dim3((N + 255) / 256), dim3(256), 0, s, N, tag, rtag);
    return hipGetLastError();
}

// NeighborList::updateExListIdx (hoomd/md/NeighborList.cc): exclusions held by tag, mapped to
// indices after every sort.
__global__ void excl_index_kernel(unsigned int N, const unsigned int* tag, const unsigned int* rtag,
                                  const unsigned int* n_excl_tag, const unsigned int* excl_tag,
                                  unsigned int max_excl, unsigned int* n_excl, unsigned int* excl)
{
    unsigned int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N)
        return;
    unsigned int tg = tag[n];
    unsigned int ne = n_excl_tag[tg];
    for (unsigned int q = 0; q < max_excl; q++)
        excl[(size_t)n * max_excl + q] =
            q < ne ? rtag[excl_tag[(size_t)tg * max_excl + q]] : 0xffffffffu;
    n_excl[n] = ne;
}

hipError_t launch_excl_index(unsigned int N, const unsigned int* tag, const unsigned int* rtag,
                             const unsigned int* n_excl_tag, const unsigned int* excl_tag,
                             unsigned int max_excl, unsigned int* n_excl, unsigned int* excl,
                             hipStream_t s)
{
    if (N == 0)
        return hipSuccess;
    hipLaunchKernelGGL(excl_index_kernel, dim3((N + 255) / 256), dim3(256), 0, s, N, tag, rtag,
                       n_excl_tag, excl_tag, max_excl, n_excl, excl);
    return hipGetLastError();
}

} // namespace hm
