// batch.hip -- many independent LJ systems ("replicas") of uneven size advanced together.
//
// The reference advances one system per Simulation; a study of many small systems (state
// points, seeds) runs them one after another, each step a launch sequence sized for one small
// system that cannot fill the chip. Here every replica lives in one set of concatenated arrays
// (replica r owns particles [off_r, off_r + n_r) and cells [cell_off_r, cell_off_r + dim_r^3))
// and every phase of the timestep is ONE launch over all replicas:
//
//   step one   (tiles)      Bussi factor of the tile's replica, velocity-Verlet first half,
//                           wrap, distance check -> the replica's rebuild flag
//   cell clear (cells)      only cells of replicas that rebuild
//   cell bin   (tiles)      only particles of replicas that rebuild
//   cell rank  (slots)      each rebuilt cell's slots ordered by particle index
//   list build (tiles)      full neighbour rows, only replicas that rebuild
//   pair       (particles)  LJ force/energy, FP32 arithmetic on FP64 minimum-image separations
//   step two   (tiles)      second half kick, kinetic-energy partial per tile
//
// A tile is up to BTILE consecutive particles of one replica, so per-tile reductions never mix
// replicas and every per-replica sum (kinetic energy) is a fixed-order sum of its tiles'
// partials: results do not depend on launch shape or on the other replicas. Every decision is
// taken on the device (per-replica flags), so chunks of steps are captured as hipGraphs.
//
// Per-replica algorithms are the reference's: TwoStepConstantVolume step one/two
// (hoomd/md/TwoStepConstantVolume.cc:7-80,213-260) with the Bussi factor of
// hoomd/md/Thermostat.h:357-446 (RNG instance = tag 0 of the replica), NeighborList's distance
// check (hoomd/md/NeighborList.cc:1220-1289, rebuild_check_delay 1), the binned full list
// (hoomd/md/NeighborListGPUBinned.cu:52-298: j != i, r^2 <= r_list^2) over the unique adjacent
// cells of CellList::initializeCellAdj (hoomd/CellList.cc:346-402), and PotentialPair with
// EvaluatorPairLJ (hoomd/md/PotentialPair.h:587-794, EvaluatorPairLJ.h:190-211).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <new>
#include <string>
#include <vector>

#include "capi_util.h"
#include "evaluators.h"
#include "kernels.h"

using namespace hm;

namespace
{

constexpr unsigned int BTILE = 256;  // particles per tile (= block of the tile kernels)

struct BError
{
    int code;
    std::string what;
};

inline void bchk(hipError_t e, const char* what)
{
    if (e != hipSuccess)
        throw BError{HM_ERR_LAUNCH, std::string(what) + ": " + hipGetErrorString(e)};
}

template<class T> T* balloc(size_t n)
{
    void* p = nullptr;
    if (hipMalloc(&p, sizeof(T) * (n ? n : 1)) != hipSuccess)
        throw BError{HM_ERR_ALLOC, "device allocation failed"};
    return (T*)p;
}

template<class T> void bfree(T*& p)
{
    if (p)
        (void)hipFree((void*)p);
    p = nullptr;
}

// per-replica constants
struct Rep
{
    double L, Linv, hi;         // cubic box edge, 1/L, L/2
    double kT, ndof;            // Bussi target and translational degrees of freedom
    double cell_inv;            // dim / L
    unsigned int off, n;        // particle range
    unsigned int cell_off, dim; // cell range (dim^3 cells)
    unsigned int tile_off, ntile;
    unsigned int seed;          // 16-bit user seed
    unsigned int pad;
    double w;                   // cell width L / dim
};

// Replicas with at least REL_MIN_DIM cells per axis keep FP32 positions relative to the centre
// of each particle's cell at the last build (rel, refreshed every step) and store row entries
// as the neighbour index plus the code (ox+1) + 3 (oy+1) + 9 (oz+1) of the neighbour cell's
// offset from the particle's cell, so the pair vector is rel_i - rel_j - o w in FP32.
constexpr unsigned int REL_MIN_DIM = 4;
constexpr unsigned int BCODE_SHIFT = 26;
constexpr unsigned int BJ_MASK = (1u << BCODE_SHIFT) - 1u;
constexpr unsigned int BCODE_SELF = 13u; // offset (0, 0, 0)

// per-replica device state
struct RepState
{
    unsigned int rebuild;      // rebuild before the forces of the current step
    unsigned int n_builds;
    unsigned int error;        // ERR_ZERO_KE
    unsigned int pad;
};

struct BatchStep
{
    unsigned long long timestep; // step that the next step-one kernel starts
};

struct BCond
{
    unsigned int cell_need;
    unsigned int nlist_need;
    unsigned int nan_idx;
    unsigned int pad;
};

struct PairConst
{
    float lj1, lj2;
    double rcutsq;
    float rcutsq_f;
    float guard_w, guard_0; // FP32 cutoff guard band: guard_w * w + guard_0
};

// FP32 offset of x from the centre of the cell with packed coordinates cc (x | y << 10 | z << 20),
// wrapped to the nearest image in FP64 first
__device__ __forceinline__ float4 centre_rel(const double4& x, unsigned int cc, const Rep& R)
{
#pragma clang fp contract(off)
    const double c[3] = {-R.hi + ((double)(cc & 1023u) + 0.5) * R.w,
                         -R.hi + ((double)((cc >> 10) & 1023u) + 0.5) * R.w,
                         -R.hi + ((double)(cc >> 20) + 0.5) * R.w};
    double d[3] = {x.x - c[0], x.y - c[1], x.z - c[2]};
    for (int k = 0; k < 3; k++)
        d[k] -= R.L * rint(d[k] * R.Linv);
    return make_float4((float)d[0], (float)d[1], (float)d[2], 0.f);
}

__device__ __forceinline__ double block_reduce(double v, double* sh)
{
    // BTILE = 256 threads = 4 wavefronts; fixed-order sums
    for (int o = 32; o > 0; o >>= 1)
        v += __shfl_xor(v, o, 64);
    const unsigned int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63u) == 0)
        sh[w] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        t = (sh[0] + sh[1]) + (sh[2] + sh[3]);
    __syncthreads();
    return t;
}

__device__ __forceinline__ void tile_range(const Rep& R, unsigned int t, unsigned int& first,
                                           unsigned int& count)
{
    const unsigned int k = t - R.tile_off;
    first = R.off + k * BTILE;
    count = min(BTILE, R.n - k * BTILE);
}

// -------------------------------------------------------------------------- step one
// TwoStepConstantVolume::integrateStepOne with the Bussi factor of the replica's kinetic energy
// (sum of its tiles' partials, fixed order, computed redundantly by every tile of the replica),
// then the distance check against the positions of the last build.
__global__ __launch_bounds__(BTILE) void bstep_one_kernel(
    const Rep* __restrict__ reps, const unsigned int* __restrict__ tile_rep,
    RepState* __restrict__ rst, const BatchStep* __restrict__ bst,
    const double* __restrict__ ke_partial, double4* __restrict__ pos, double4* __restrict__ vel,
    const d3* __restrict__ accel, i3* __restrict__ image, const d3* __restrict__ last,
    BCond* __restrict__ cond, double dt, double tau, int bussi, double maxsq,
    const unsigned int* __restrict__ cellc, float4* __restrict__ rel)
{
#pragma clang fp contract(off)
    __shared__ double sh[4];
    __shared__ double s_alpha;
    __shared__ unsigned int s_move;
    const unsigned int t = blockIdx.x;
    const unsigned int r = tile_rep[t];
    const Rep R = reps[r];
    unsigned int first, count;
    tile_range(R, t, first, count);
    if (threadIdx.x == 0)
        s_move = 0;
    double alpha = 1.0;
    if (bussi)
    {
        double acc = 0.0;
        if (threadIdx.x < R.ntile)
            acc = ke_partial[R.tile_off + threadIdx.x];
        // replicas of more than BTILE tiles: remaining partials in a fixed order
        for (unsigned int q = threadIdx.x + BTILE; q < R.ntile; q += BTILE)
            acc += ke_partial[R.tile_off + q];
        const double tot = block_reduce(acc, sh);
        if (threadIdx.x == 0)
        {
            const double K = tot * 0.5;
            double a = 1.0;
            if (K == 0.0)
            {
                if (t == R.tile_off)
                    atomicOr(&rst[r].error, ERR_ZERO_KE);
            }
            else
                a = bussi_alpha(K, R.ndof, dt, tau, R.kT, bst->timestep, (uint16_t)R.seed, 0u);
            s_alpha = a;
        }
        __syncthreads();
        alpha = s_alpha;
    }
    unsigned int moved = 0;
    if (threadIdx.x < count)
    {
        const unsigned int p = first + threadIdx.x;
        double4 x = pos[p];
        double4 v = vel[p];
        const d3 a = accel[p];
        v.x = (v.x + (1.0 / 2.0) * a.x * dt) * alpha;
        v.y = (v.y + (1.0 / 2.0) * a.y * dt) * alpha;
        v.z = (v.z + (1.0 / 2.0) * a.z * dt) * alpha;
        x.x += dt * v.x;
        x.y += dt * v.y;
        x.z += dt * v.z;
        i3 im = image[p];
        wrap_axis(x.x, im.x, -R.hi, R.hi, R.L);
        wrap_axis(x.y, im.y, -R.hi, R.hi, R.L);
        wrap_axis(x.z, im.z, -R.hi, R.hi, R.L);
        pos[p] = x;
        vel[p] = v;
        image[p] = im;
        if (R.dim >= REL_MIN_DIM)
            rel[p] = centre_rel(x, cellc[p], R);
        if (x.x != x.x || x.y != x.y || x.z != x.z)
            atomicMax(&cond->nan_idx, p + 1);
        const d3 l = last[p];
        double dx = x.x - l.x, dy = x.y - l.y, dz = x.z - l.z;
        dz -= R.L * rint(dz * R.Linv);
        dy -= R.L * rint(dy * R.Linv);
        dx -= R.L * rint(dx * R.Linv);
        moved = (dx * dx + dy * dy + dz * dz >= maxsq) ? 1u : 0u;
    }
    if (wave_ballot(moved != 0) != 0 && lane_id() == 0)
        s_move = 1;  // benign race: every writer stores 1
    __syncthreads();
    if (threadIdx.x == 0 && s_move)
        atomicOr(&rst[r].rebuild, 1u);
}

// -------------------------------------------------------------------------- cells
__global__ __launch_bounds__(256) void bcell_clear_kernel(const Rep* __restrict__ reps,
                                                          const unsigned int* __restrict__ cell_rep,
                                                          RepState* __restrict__ rst,
                                                          unsigned int ncell,
                                                          unsigned int* __restrict__ cell_size)
{
    const unsigned int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell)
        return;
    const unsigned int r = cell_rep[c];
    if (!rst[r].rebuild)
        return;
    cell_size[c] = 0;
    if (c == reps[r].cell_off)
        rst[r].n_builds++;
}

__device__ __forceinline__ unsigned int cell_coord(double x, const Rep& R)
{
    // CellList::computeCellList bin (hoomd/CellList.cc:466-490) for a cubic box at the origin:
    // floor(fractional * dim), the upper edge folded back
    int c = (int)((x + R.hi) * R.cell_inv);
    if (c >= (int)R.dim)
        c = 0;
    if (c < 0)
        c = 0;
    return (unsigned int)c;
}

// Also zeroes the particle's neighbour count: the list build writes the counts of the particles
// it finds in the cells, so a particle left out of an overflowed cell keeps an empty row (the
// overflow then replays the chunk) instead of a count from an earlier build.
__global__ __launch_bounds__(BTILE) void bcell_bin_kernel(
    const Rep* __restrict__ reps, const unsigned int* __restrict__ tile_rep,
    const RepState* __restrict__ rst, const double4* __restrict__ pos, d3* __restrict__ last,
    unsigned int cmax, unsigned int* __restrict__ cell_size, unsigned int* __restrict__ cell_raw,
    unsigned int* __restrict__ n_neigh, BCond* __restrict__ cond, unsigned int* __restrict__ cellc,
    float4* __restrict__ rel)
{
    const unsigned int t = blockIdx.x;
    const unsigned int r = tile_rep[t];
    if (!rst[r].rebuild)
        return;
    const Rep R = reps[r];
    unsigned int first, count;
    tile_range(R, t, first, count);
    if (threadIdx.x >= count)
        return;
    const unsigned int p = first + threadIdx.x;
    const double4 x = pos[p];
    last[p] = d3{x.x, x.y, x.z};
    n_neigh[p] = 0;
    const unsigned int ccx = cell_coord(x.x, R), ccy = cell_coord(x.y, R), ccz = cell_coord(x.z, R);
    const unsigned int c = R.cell_off + (ccz * R.dim + ccy) * R.dim + ccx;
    const unsigned int cc = ccx | (ccy << 10) | (ccz << 20);
    cellc[p] = cc;
    if (R.dim >= REL_MIN_DIM)
        rel[p] = centre_rel(x, cc, R);
    const unsigned int slot = atomicAdd(&cell_size[c], 1u);
    if (slot < cmax)
        cell_raw[(size_t)c * cmax + slot] = p;
    else
        atomicMax(&cond->cell_need, slot + 1);
}

// Order every rebuilt cell's slots by particle index (slots were claimed by atomics in arrival
// order): slot k's particle goes to its rank among the cell's particles. Rows, and so the FP32
// force sums, then do not depend on scheduling: a run is reproducible and a replica follows the
// same trajectory alone or in any batch.
__global__ __launch_bounds__(256) void bcell_rank_kernel(const unsigned int* __restrict__ cell_rep,
                                                         const RepState* __restrict__ rst,
                                                         unsigned int ncell, unsigned int cmax,
                                                         const unsigned int* __restrict__ cell_size,
                                                         const unsigned int* __restrict__ cell_raw,
                                                         unsigned int* __restrict__ cell_idx)
{
    const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (size_t)ncell * cmax)
        return;
    const unsigned int c = (unsigned int)(g / cmax), k = (unsigned int)(g % cmax);
    if (!rst[cell_rep[c]].rebuild)
        return;
    const unsigned int n = min(cell_size[c], cmax);
    if (k >= n)
        return;
    const unsigned int* raw = cell_raw + (size_t)c * cmax;
    const unsigned int my = raw[k];
    unsigned int rank = 0;
    for (unsigned int m = 0; m < n; m++)
        rank += raw[m] < my ? 1u : 0u;
    cell_idx[(size_t)c * cmax + rank] = my;
}

// -------------------------------------------------------------------------- list build
// One workgroup of NL_WAVES wavefronts per home cell of a rebuilding replica. The candidates
// (the particles of the unique adjacent cells, in cell order: offsets {-1, 0, 1} per axis where
// the axis has at least 3 cells, {0, 1} for 2, {0} for 1, as CellList::initializeCellAdj keeps
// each adjacent cell once) are staged in LDS, NL_CAP at a time; then each wavefront takes every
// NL_WAVES-th home particle, its 64 lanes test 64 candidates at once, and a ballot places the hits
// contiguously in the particle's row, in candidate order (NeighborListGPUBinned.cu:52-298 with
// the lanes over candidates instead of over particles).
//
// With at least 4 cells per axis, candidates and home particles are staged as FP32 offsets from
// the home-cell centre (each wrapped to its nearest image once, in FP64): every offset is at most
// 1.5 w per axis (w = L / dim >= r_list), strictly below L / 2 = 2 w or more, and a pair within
// r_list has one image shorter than L / 2, so the FP32 difference of offsets is the minimum
// image. Its r^2 is off the FP64 value by at
// most 8 sqrt(3) u w r_list + 3 u r_list^2 (u = 2^-24: offsets rounded to FP32, one subtraction
// per axis, three squares and two sums near r^2 = r_list^2); pairs within 32 u (w r_list +
// r_list^2) of the cut are re-decided from the FP64 positions with the FP64 minimum image, so the
// rows equal those of the FP64 test. Boxes with fewer than 4 cells per axis take the FP64 test
// for every pair (with 3, an offset of exactly 1.5 w = L / 2 has two nearest images).
constexpr unsigned int NL_CAP = 1024;
constexpr unsigned int NL_HOME_MAX = 256;
constexpr unsigned int NL_WAVES = 4;

__device__ __forceinline__ bool within_fp64(const double4& xi, const double4& xj, const Rep& R,
                                            double rlistsq)
{
#pragma clang fp contract(off)
    double dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
    dz -= R.L * rint(dz * R.Linv);
    dy -= R.L * rint(dy * R.Linv);
    dx -= R.L * rint(dx * R.Linv);
    return dx * dx + dy * dy + dz * dz <= rlistsq;
}

__device__ __forceinline__ float centre_offset(double x, double centre, const Rep& R)
{
#pragma clang fp contract(off)
    double d = x - centre;
    d -= R.L * rint(d * R.Linv);
    return (float)d;
}

__global__ __launch_bounds__(64 * NL_WAVES) void bnlist_kernel(
    const Rep* __restrict__ reps, const unsigned int* __restrict__ cell_rep,
    const RepState* __restrict__ rst, const double4* __restrict__ pos, unsigned int cmax,
    const unsigned int* __restrict__ cell_size, const unsigned int* __restrict__ cell_idx,
    double rlistsq, unsigned int nmax, unsigned int* __restrict__ nlist,
    unsigned int* __restrict__ n_neigh, BCond* __restrict__ cond)
{
#pragma clang fp contract(off)
    __shared__ float4 sc[NL_CAP];  // offsets from the home-cell centre, candidate index in w
    __shared__ unsigned int acell[27], aoff[28];
    __shared__ unsigned int hcnt[NL_HOME_MAX];
    const unsigned int c = blockIdx.x;
    const unsigned int r = cell_rep[c];
    if (!rst[r].rebuild)
        return;
    const Rep R = reps[r];
    const unsigned int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const unsigned int lc = c - R.cell_off;
    const int d = (int)R.dim;
    const int cx = (int)(lc % R.dim), cy = (int)((lc / R.dim) % R.dim), cz = (int)(lc / (R.dim * R.dim));
    const int lo = d >= 3 ? -1 : 0;
    const int span = d >= 3 ? 3 : (d == 2 ? 2 : 1);
    const unsigned int nadj = (unsigned int)(span * span * span);
    if (wave == 0)
    {
        // adjacent cells in (z, y, x) offset order and their sizes, prefix sums by a wave scan
        unsigned int asz = 0;
        if (lane < nadj)
        {
            const int ox = lo + (int)(lane % span), oy = lo + (int)((lane / span) % span),
                      oz = lo + (int)(lane / (span * span));
            const int x = (cx + ox + d) % d, y = (cy + oy + d) % d, z = (cz + oz + d) % d;
            const unsigned int ac = R.cell_off + ((unsigned int)z * R.dim + y) * R.dim + x;
            acell[lane] = ac;
            asz = min(cell_size[ac], cmax);
        }
        unsigned int incl = asz;
        for (unsigned int o = 1; o < 32; o <<= 1)
        {
            const unsigned int t = __shfl_up(incl, o, 64);
            if (lane >= o)
                incl += t;
        }
        if (lane < nadj)
            aoff[lane + 1] = incl;
        if (lane == 0)
            aoff[0] = 0;
    }
    const unsigned int nhome = min(min(cell_size[c], cmax), NL_HOME_MAX);
    const unsigned int* home = cell_idx + (size_t)c * cmax;
    for (unsigned int h = tid; h < nhome; h += blockDim.x)
        hcnt[h] = 0;
    __syncthreads();
    const unsigned int total = aoff[nadj];
    const bool rel = d >= (int)REL_MIN_DIM;
    const double w = R.L / (double)d;
    const double ccx = -R.hi + ((double)cx + 0.5) * w, ccy = -R.hi + ((double)cy + 0.5) * w,
                 ccz = -R.hi + ((double)cz + 0.5) * w;
    const double rl = sqrt(rlistsq);
    const float rlsq_f = (float)rlistsq;
    const float guard = (float)(32.0 * 0x1p-24 * (w * rl + rlistsq));
    for (unsigned int base = 0; base < total; base += NL_CAP)
    {
        const unsigned int npass = min(NL_CAP, total - base);
        for (unsigned int k = tid; k < npass; k += blockDim.x)
        {
            const unsigned int g = base + k;
            unsigned int seg = 0;
            while (aoff[seg + 1] <= g)
                seg++;
            const unsigned int j = cell_idx[(size_t)acell[seg] * cmax + (g - aoff[seg])];
            // with offsets, the stored entry carries the cell-offset code: seg enumerates the
            // 27 offsets as (ox + 1) + 3 (oy + 1) + 9 (oz + 1)
            float4 q = make_float4(0.f, 0.f, 0.f, __uint_as_float(rel ? j | (seg << BCODE_SHIFT) : j));
            if (rel)
            {
                const double4 xj = pos[j];
                q.x = centre_offset(xj.x, ccx, R);
                q.y = centre_offset(xj.y, ccy, R);
                q.z = centre_offset(xj.z, ccz, R);
            }
            sc[k] = q;
        }
        __syncthreads();
        for (unsigned int h = wave; h < nhome; h += NL_WAVES)
        {
            const unsigned int i = home[h];
            const double4 xi = pos[i];
            const float hx = rel ? centre_offset(xi.x, ccx, R) : 0.f;
            const float hy = rel ? centre_offset(xi.y, ccy, R) : 0.f;
            const float hz = rel ? centre_offset(xi.z, ccz, R) : 0.f;
            unsigned int cnt = hcnt[h];
            unsigned int* row = nlist + (size_t)i * nmax;
            for (unsigned int k0 = 0; k0 < npass; k0 += 64)
            {
                const unsigned int k = k0 + lane;
                bool hit = false;
                unsigned int e = 0;
                if (k < npass)
                {
                    const float4 q = sc[k];
                    e = __float_as_uint(q.w);
                    const unsigned int j = e & BJ_MASK;
                    if (rel)
                    {
                        const float dx = hx - q.x, dy = hy - q.y, dz = hz - q.z;
                        const float d2 = dx * dx + dy * dy + dz * dz;
                        hit = d2 <= rlsq_f;
                        if (fabsf(d2 - rlsq_f) <= guard)
                            hit = within_fp64(xi, pos[j], R, rlistsq);
                    }
                    else
                        hit = within_fp64(xi, pos[j], R, rlistsq);
                    hit = hit && j != i;
                }
                const uint64_t m = wave_ballot(hit);
                const unsigned int at = cnt + lanes_below(m);
                if (hit && at < nmax)
                    row[at] = e;
                cnt += (unsigned int)__popcll(m);
            }
            if (lane == 0)
                hcnt[h] = cnt;
        }
        __syncthreads();
    }
    for (unsigned int h = tid; h < nhome; h += blockDim.x)
    {
        const unsigned int cnt = hcnt[h];
        n_neigh[home[h]] = min(cnt, nmax);
        if (cnt > nmax)
            atomicMax(&cond->nlist_need, cnt);
    }
}

// -------------------------------------------------------------------------- pair forces
// tpp lanes per particle split its row; FP32 evaluator arithmetic and FP32 per-particle sums
// (force, pair energy halved as PotentialPair.h:739-772). Replicas with offsets: the FP32 pair
// vector rel_i - rel_j - o w (16-byte gathers) and an FP32 cutoff test whose guard band is
// re-decided from the FP64 positions; others: FP64 separation and exact cutoff test.
__device__ __forceinline__ void lj_accumulate(const PairConst& pc, float rsq, float dx, float dy,
                                              float dz, float& fx, float& fy, float& fz,
                                              float& e)
{
#pragma clang fp contract(off)
    const float r2inv = recip(rsq);
    const float r6inv = r2inv * r2inv * r2inv;
    const float fdr = r2inv * r6inv * (12.f * pc.lj1 * r6inv - 6.f * pc.lj2);
    e += r6inv * (pc.lj1 * r6inv - pc.lj2);
    fx += dx * fdr;
    fy += dy * fdr;
    fz += dz * fdr;
}

template<unsigned int TPP>
__global__ __launch_bounds__(256) void bpair_kernel(
    const Rep* __restrict__ reps, const unsigned int* __restrict__ part_rep, unsigned int N,
    const double4* __restrict__ pos, const unsigned int* __restrict__ nlist, unsigned int nmax,
    const unsigned int* __restrict__ n_neigh, PairConst pc, float4* __restrict__ force,
    const float4* __restrict__ rel)
{
#pragma clang fp contract(off)
    const unsigned int g = blockIdx.x * blockDim.x + threadIdx.x;
    const unsigned int p = g / TPP;
    const unsigned int sub = g % TPP;
    float fx = 0.f, fy = 0.f, fz = 0.f, e = 0.f;
    constexpr unsigned int UNR = 4;
    if (p < N && reps[part_rep[p]].dim >= REL_MIN_DIM)
    {
        const Rep R = reps[part_rep[p]];
        const float4 ri = rel[p];
        const float w = (float)R.w;
        const float guard = pc.guard_w * w + pc.guard_0;
        const unsigned int nn = min(n_neigh[p], nmax);
        const unsigned int* row = nlist + (size_t)p * nmax;
        for (unsigned int k0 = sub; k0 < nn; k0 += TPP * UNR)
        {
            unsigned int en[UNR];
#pragma unroll
            for (unsigned int u = 0; u < UNR; u++)
            {
                const unsigned int k = k0 + u * TPP;
                // the particle itself, at zero offset, stands in for missing entries
                en[u] = k < nn ? row[k] : (p | (BCODE_SELF << BCODE_SHIFT));
            }
            float4 rj[UNR];
#pragma unroll
            for (unsigned int u = 0; u < UNR; u++)
                rj[u] = rel[en[u] & BJ_MASK];
#pragma unroll
            for (unsigned int u = 0; u < UNR; u++)
            {
                const unsigned int j = en[u] & BJ_MASK, code = en[u] >> BCODE_SHIFT;
                const float ox = (float)((int)(code % 3u) - 1);
                const float oy = (float)((int)((code / 3u) % 3u) - 1);
                const float oz = (float)((int)(code / 9u) - 1);
                const float dx = (ri.x - rj[u].x) - ox * w;
                const float dy = (ri.y - rj[u].y) - oy * w;
                const float dz = (ri.z - rj[u].z) - oz * w;
                const float rsq = dx * dx + dy * dy + dz * dz;
                const float t = rsq - pc.rcutsq_f;
                bool in = t < 0.f;
                if (fabsf(t) <= guard)
                {
                    // exact FP64 decision (the FP64 form of this kernel)
                    const double4 xi = pos[p], xj = pos[j];
                    double ddx = xi.x - xj.x, ddy = xi.y - xj.y, ddz = xi.z - xj.z;
                    ddz -= R.L * rint(ddz * R.Linv);
                    ddy -= R.L * rint(ddy * R.Linv);
                    ddx -= R.L * rint(ddx * R.Linv);
                    in = ddx * ddx + ddy * ddy + ddz * ddz < pc.rcutsq;
                }
                if (in && j != p)
                    lj_accumulate(pc, rsq, dx, dy, dz, fx, fy, fz, e);
            }
        }
    }
    else if (p < N)
    {
        const Rep R = reps[part_rep[p]];
        const double4 xi = pos[p];
        const unsigned int nn = min(n_neigh[p], nmax);
        const unsigned int* row = nlist + (size_t)p * nmax;
        // UNR neighbours per lane per pass: every index, then every position load is issued
        // before the first is used, so several gathers are in flight per lane
        for (unsigned int k0 = sub; k0 < nn; k0 += TPP * UNR)
        {
            unsigned int j[UNR];
#pragma unroll
            for (unsigned int u = 0; u < UNR; u++)
            {
                const unsigned int k = k0 + u * TPP;
                j[u] = k < nn ? row[k] : p;
            }
            double4 xj[UNR];
#pragma unroll
            for (unsigned int u = 0; u < UNR; u++)
                xj[u] = pos[j[u]];
#pragma unroll
            for (unsigned int u = 0; u < UNR; u++)
            {
                double dx = xi.x - xj[u].x, dy = xi.y - xj[u].y, dz = xi.z - xj[u].z;
                dz -= R.L * rint(dz * R.Linv);
                dy -= R.L * rint(dy * R.Linv);
                dx -= R.L * rint(dx * R.Linv);
                const double rsq = dx * dx + dy * dy + dz * dz;
                // the particle itself stands in for missing entries: rsq = 0 is excluded
                if (rsq < pc.rcutsq && j[u] != p)
                    lj_accumulate(pc, (float)rsq, (float)dx, (float)dy, (float)dz, fx, fy, fz, e);
            }
        }
    }
    for (unsigned int o = TPP / 2; o > 0; o >>= 1)
    {
        fx += __shfl_xor(fx, o, 64);
        fy += __shfl_xor(fy, o, 64);
        fz += __shfl_xor(fz, o, 64);
        e += __shfl_xor(e, o, 64);
    }
    if (p < N && sub == 0)
        force[p] = make_float4(fx, fy, fz, 0.5f * e);
}

// -------------------------------------------------------------------------- step two
__global__ __launch_bounds__(BTILE) void bstep_two_kernel(
    const Rep* __restrict__ reps, const unsigned int* __restrict__ tile_rep,
    RepState* __restrict__ rst, BatchStep* __restrict__ bst, const float4* __restrict__ force,
    double4* __restrict__ vel, d3* __restrict__ accel, double* __restrict__ ke_partial, double dt,
    int advance)
{
#pragma clang fp contract(off)
    __shared__ double sh[4];
    const unsigned int t = blockIdx.x;
    const unsigned int r = tile_rep[t];
    const Rep R = reps[r];
    unsigned int first, count;
    tile_range(R, t, first, count);
    double ke = 0.0;
    if (threadIdx.x < count)
    {
        const unsigned int p = first + threadIdx.x;
        double4 v = vel[p];
        const float4 f = force[p];
        const double minv = 1.0 / v.w;
        d3 a{(double)f.x * minv, (double)f.y * minv, (double)f.z * minv};
        if (advance)
        {
            v.x += (1.0 / 2.0) * dt * a.x;
            v.y += (1.0 / 2.0) * dt * a.y;
            v.z += (1.0 / 2.0) * dt * a.z;
            vel[p] = v;
        }
        accel[p] = a;
        ke = v.w * (v.x * v.x + v.y * v.y + v.z * v.z);
    }
    const double tot = block_reduce(ke, sh);
    if (threadIdx.x == 0)
    {
        ke_partial[t] = tot;
        if (t == R.tile_off)
            rst[r].rebuild = 0;
        if (t == 0 && advance)
            bst->timestep++;
    }
}

// -------------------------------------------------------------------------- cell order
// The particles of each replica re-laid in cell order (cells in index order, each cell's
// particles by index), so that the particles a wavefront gathers in the list build and the pair
// kernel share cache lines; the reference's SFCPackTuner does this with a Hilbert order
// (hoomd/SFCPackTuner.cc:119-285) for the same reason. One wavefront per replica turns its cell
// sizes into first slots (exclusive prefix sums, offset by the replica's first particle).
__global__ __launch_bounds__(64) void bcell_start_kernel(const Rep* __restrict__ reps,
                                                         const unsigned int* __restrict__ cell_size,
                                                         unsigned int cmax,
                                                         unsigned int* __restrict__ cell_start)
{
    const Rep R = reps[blockIdx.x];
    const unsigned int ncell = R.dim * R.dim * R.dim;
    const unsigned int lane = threadIdx.x;
    unsigned int base = R.off;
    for (unsigned int c0 = 0; c0 < ncell; c0 += 64)
    {
        const unsigned int c = c0 + lane;
        const unsigned int n = c < ncell ? min(cell_size[R.cell_off + c], cmax) : 0u;
        unsigned int incl = n;
        for (unsigned int o = 1; o < 64; o <<= 1)
        {
            const unsigned int t = __shfl_up(incl, o, 64);
            if (lane >= o)
                incl += t;
        }
        if (c < ncell)
            cell_start[R.cell_off + c] = base + incl - n;
        base += __shfl(incl, 63, 64);
    }
}

// order[new index] = old index
__global__ __launch_bounds__(256) void bcell_order_kernel(unsigned int ncell, unsigned int cmax,
                                                          const unsigned int* __restrict__ cell_size,
                                                          const unsigned int* __restrict__ cell_start,
                                                          const unsigned int* __restrict__ cell_idx,
                                                          unsigned int* __restrict__ order)
{
    const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (size_t)ncell * cmax)
        return;
    const unsigned int c = (unsigned int)(g / cmax), k = (unsigned int)(g % cmax);
    if (k < min(cell_size[c], cmax))
        order[cell_start[c] + k] = cell_idx[g];
}

__global__ void bset_rebuild_kernel(RepState* rst, unsigned int nrep)
{
    const unsigned int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < nrep)
        rst[r].rebuild = 1;
}

} // namespace

// ============================================================================ host object
struct hm_batch
{
    unsigned int nrep = 0, N = 0, ntile = 0, ncell = 0;
    std::vector<Rep> hrep;
    double dt = 0.005, tau = 0.0, r_cut = 2.5, r_buff = 0.4, rlistsq = 0.0, maxsq = 0.0;
    PairConst pc{};
    unsigned int cmax = 0, nmax = 0;
    unsigned int pair_tpp = 4, pair_block = 256;
    bool tuned = false;
    unsigned int chunk = 100;
    uint64_t timestep = 0, steps_run = 0, rollbacks = 0;
    bool prepared = false;
    hipStream_t stream = nullptr;

    Rep* reps = nullptr;
    RepState* rst = nullptr;
    BatchStep* bst = nullptr;
    BCond* cond = nullptr;
    BCond* h_cond = nullptr;
    unsigned int* tile_rep = nullptr;
    unsigned int* cell_rep = nullptr;
    unsigned int* part_rep = nullptr;
    double4* pos = nullptr;
    double4* vel = nullptr;
    d3* accel = nullptr;
    i3* image = nullptr;
    d3* last = nullptr;
    float4* force = nullptr;
    double* ke_partial = nullptr;
    unsigned int* cell_size = nullptr;
    unsigned int* cell_raw = nullptr;
    unsigned int* cell_idx = nullptr;
    unsigned int* nlist = nullptr;
    unsigned int* n_neigh = nullptr;
    // cell-order sort: replica-local tag of each particle (its index at creation, minus the
    // replica's first index), the new order and a staging buffer for the permutation
    unsigned int* tag = nullptr;
    // offsets from the particle's cell centre at the last build, and that cell (packed)
    float4* rel = nullptr;
    unsigned int* cellc = nullptr;
    unsigned int* cell_start = nullptr;
    unsigned int* order = nullptr;
    void* stage = nullptr;
    unsigned int sort_period = 100;
    uint64_t last_sort = ~0ull, n_sorts = 0;
    // chunk-start snapshot for overflow replay
    double4 *s_pos = nullptr, *s_vel = nullptr;
    d3 *s_accel = nullptr, *s_last = nullptr;
    i3* s_image = nullptr;
    double* s_ke = nullptr;
    unsigned int* s_tag = nullptr;
    RepState* s_rst = nullptr;
    BatchStep* s_bst = nullptr;

    std::map<unsigned int, hipGraphExec_t> graphs;
    std::vector<float> tune_log; // rows of (tpp, block, ms)

    ~hm_batch()
    {
        clear_graphs();
        for (void** p : {(void**)&reps, (void**)&rst, (void**)&bst, (void**)&cond,
                         (void**)&tile_rep, (void**)&cell_rep, (void**)&part_rep, (void**)&pos,
                         (void**)&vel, (void**)&accel, (void**)&image, (void**)&last,
                         (void**)&force, (void**)&ke_partial, (void**)&cell_size,
                         (void**)&cell_raw, (void**)&cell_idx, (void**)&nlist, (void**)&n_neigh, (void**)&tag, (void**)&rel, (void**)&cellc, (void**)&cell_start, (void**)&order, (void**)&stage, (void**)&s_pos,
                         (void**)&s_vel, (void**)&s_accel, (void**)&s_last, (void**)&s_image,
                         (void**)&s_ke, (void**)&s_tag, (void**)&s_rst, (void**)&s_bst})
            if (*p)
            {
                (void)hipFree(*p);
                *p = nullptr;
            }
        if (h_cond)
            (void)hipHostFree(h_cond);
        if (stream)
            (void)hipStreamDestroy(stream);
    }

    void clear_graphs()
    {
        for (auto& kv : graphs)
            (void)hipGraphExecDestroy(kv.second);
        graphs.clear();
    }

    void alloc_lists()
    {
        if (cmax > NL_HOME_MAX)
            throw BError{HM_ERR_LAUNCH, "batch: more than 256 particles in one cell"};
        bfree(cell_raw);
        bfree(cell_idx);
        bfree(nlist);
        cell_raw = balloc<unsigned int>((size_t)ncell * cmax);
        cell_idx = balloc<unsigned int>((size_t)ncell * cmax);
        nlist = balloc<unsigned int>((size_t)N * nmax);
        // zero-filled: every row entry and count is a valid index / length before any build
        bchk(hipMemset(nlist, 0, sizeof(unsigned int) * (size_t)N * nmax), "clear rows");
        bchk(hipMemset(cell_idx, 0, sizeof(unsigned int) * (size_t)ncell * cmax), "clear cells");
        if (n_neigh)
            bchk(hipMemset(n_neigh, 0, sizeof(unsigned int) * N), "clear counts");
        clear_graphs();
    }

    // ---------------------------------------------------------------- launches
    void launch_pair()
    {
        const unsigned int nthreads = N * pair_tpp;
        const dim3 grid((nthreads + pair_block - 1) / pair_block), block(pair_block);
        switch (pair_tpp)
        {
        case 1:
            hipLaunchKernelGGL(bpair_kernel<1>, grid, block, 0, stream, reps, part_rep, N, pos,
                               nlist, nmax, n_neigh, pc, force, rel);
            break;
        case 2:
            hipLaunchKernelGGL(bpair_kernel<2>, grid, block, 0, stream, reps, part_rep, N, pos,
                               nlist, nmax, n_neigh, pc, force, rel);
            break;
        case 4:
            hipLaunchKernelGGL(bpair_kernel<4>, grid, block, 0, stream, reps, part_rep, N, pos,
                               nlist, nmax, n_neigh, pc, force, rel);
            break;
        case 8:
            hipLaunchKernelGGL(bpair_kernel<8>, grid, block, 0, stream, reps, part_rep, N, pos,
                               nlist, nmax, n_neigh, pc, force, rel);
            break;
        default:
            hipLaunchKernelGGL(bpair_kernel<16>, grid, block, 0, stream, reps, part_rep, N, pos,
                               nlist, nmax, n_neigh, pc, force, rel);
            break;
        }
        bchk(hipGetLastError(), "batch pair");
    }

    void launch_rebuild()
    {
        hipLaunchKernelGGL(bcell_clear_kernel, dim3((ncell + 255) / 256), dim3(256), 0, stream,
                           reps, cell_rep, rst, ncell, cell_size);
        hipLaunchKernelGGL(bcell_bin_kernel, dim3(ntile), dim3(BTILE), 0, stream, reps, tile_rep,
                           rst, pos, last, cmax, cell_size, cell_raw, n_neigh, cond, cellc, rel);
        const size_t nslot = (size_t)ncell * cmax;
        hipLaunchKernelGGL(bcell_rank_kernel, dim3((unsigned int)((nslot + 255) / 256)), dim3(256),
                           0, stream, cell_rep, rst, ncell, cmax, cell_size, cell_raw, cell_idx);
        hipLaunchKernelGGL(bnlist_kernel, dim3(ncell), dim3(64 * NL_WAVES), 0, stream, reps,
                           cell_rep, rst,
                           pos, cmax, cell_size, cell_idx, rlistsq, nmax, nlist, n_neigh, cond);
        bchk(hipGetLastError(), "batch rebuild");
    }

    void launch_step()
    {
        hipLaunchKernelGGL(bstep_one_kernel, dim3(ntile), dim3(BTILE), 0, stream, reps, tile_rep,
                           rst, bst, ke_partial, pos, vel, accel, image, last, cond, dt, tau, 1,
                           maxsq, cellc, rel);
        bchk(hipGetLastError(), "batch step one");
        launch_rebuild();
        launch_pair();
        hipLaunchKernelGGL(bstep_two_kernel, dim3(ntile), dim3(BTILE), 0, stream, reps, tile_rep,
                           rst, bst, force, vel, accel, ke_partial, dt, 1);
        bchk(hipGetLastError(), "batch step two");
    }

    // forces, accelerations and kinetic-energy partials of the current state (all lists rebuilt)
    void prep()
    {
        hipLaunchKernelGGL(bset_rebuild_kernel, dim3((nrep + 255) / 256), dim3(256), 0, stream,
                           rst, nrep);
        bchk(hipMemsetAsync(cond, 0, sizeof(BCond), stream), "clear conditions");
        launch_rebuild();
        launch_pair();
        hipLaunchKernelGGL(bstep_two_kernel, dim3(ntile), dim3(BTILE), 0, stream, reps, tile_rep,
                           rst, bst, force, vel, accel, ke_partial, dt, 0);
        bchk(hipGetLastError(), "batch prep");
        read_cond();
        if (h_cond->cell_need > cmax || h_cond->nlist_need > nmax)
        {
            cmax = std::max(cmax, round_up(h_cond->cell_need + 8, 8));
            nmax = std::max(nmax, round_up(h_cond->nlist_need + 16, 16));
            alloc_lists();
            return prep();
        }
        prepared = true;
    }

    void read_cond()
    {
        bchk(hipMemcpyAsync(h_cond, cond, sizeof(BCond), hipMemcpyDeviceToHost, stream), "cond");
        bchk(hipStreamSynchronize(stream), "sync");
    }

    hipGraphExec_t capture_steps(unsigned int K)
    {
        hipGraph_t g;
        bchk(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "begin capture");
        try
        {
            for (unsigned int k = 0; k < K; k++)
                launch_step();
        }
        catch (...)
        {
            (void)hipStreamEndCapture(stream, &g);
            throw;
        }
        bchk(hipStreamEndCapture(stream, &g), "end capture");
        hipGraphExec_t ex;
        hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        bchk(e, "graph instantiate");
        return ex;
    }

    void snapshot()
    {
        if (!s_pos)
        {
            s_pos = balloc<double4>(N);
            s_vel = balloc<double4>(N);
            s_accel = balloc<d3>(N);
            s_last = balloc<d3>(N);
            s_image = balloc<i3>(N);
            s_ke = balloc<double>(ntile);
            s_tag = balloc<unsigned int>(N);
            s_rst = balloc<RepState>(nrep);
            s_bst = balloc<BatchStep>(1);
        }
        auto cp = [&](void* d, const void* s, size_t b) {
            bchk(hipMemcpyAsync(d, s, b, hipMemcpyDeviceToDevice, stream), "snapshot");
        };
        cp(s_pos, pos, (size_t)N * sizeof(double4));
        cp(s_vel, vel, (size_t)N * sizeof(double4));
        cp(s_accel, accel, (size_t)N * sizeof(d3));
        cp(s_last, last, (size_t)N * sizeof(d3));
        cp(s_image, image, (size_t)N * sizeof(i3));
        cp(s_ke, ke_partial, (size_t)ntile * sizeof(double));
        cp(s_tag, tag, (size_t)N * sizeof(unsigned int));
        cp(s_rst, rst, (size_t)nrep * sizeof(RepState));
        cp(s_bst, bst, sizeof(BatchStep));
    }

    void restore()
    {
        auto cp = [&](void* d, const void* s, size_t b) {
            bchk(hipMemcpyAsync(d, s, b, hipMemcpyDeviceToDevice, stream), "restore");
        };
        cp(pos, s_pos, (size_t)N * sizeof(double4));
        cp(vel, s_vel, (size_t)N * sizeof(double4));
        cp(accel, s_accel, (size_t)N * sizeof(d3));
        cp(last, s_last, (size_t)N * sizeof(d3));
        cp(image, s_image, (size_t)N * sizeof(i3));
        cp(ke_partial, s_ke, (size_t)ntile * sizeof(double));
        cp(tag, s_tag, (size_t)N * sizeof(unsigned int));
        cp(rst, s_rst, (size_t)nrep * sizeof(RepState));
        cp(bst, s_bst, sizeof(BatchStep));
    }

    uint64_t chunk_len(uint64_t t, uint64_t end) const
    {
        uint64_t K = std::min<uint64_t>(chunk ? chunk : 1, end - t);
        if (sort_period)
            K = std::min<uint64_t>(K, (t / sort_period + 1) * sort_period - t);
        return K;
    }

    // Re-lay every replica's particles in cell order (between chunks, outside any graph): a
    // fresh binning of all replicas, the per-replica cell starts, the order, the permutation of
    // every per-particle array, then the kinetic-energy partials of the new tiles and a forced
    // list build before the next forces.
    void sort_cells()
    {
        if (!cell_start)
        {
            cell_start = balloc<unsigned int>(ncell);
            order = balloc<unsigned int>(N);
            stage = balloc<double4>(N);
        }
        for (int attempt = 0;; attempt++)
        {
            bchk(hipMemsetAsync(cond, 0, sizeof(BCond), stream), "clear conditions");
            hipLaunchKernelGGL(bset_rebuild_kernel, dim3((nrep + 255) / 256), dim3(256), 0,
                               stream, rst, nrep);
            hipLaunchKernelGGL(bcell_clear_kernel, dim3((ncell + 255) / 256), dim3(256), 0,
                               stream, reps, cell_rep, rst, ncell, cell_size);
            hipLaunchKernelGGL(bcell_bin_kernel, dim3(ntile), dim3(BTILE), 0, stream, reps,
                               tile_rep, rst, pos, last, cmax, cell_size, cell_raw, n_neigh,
                               cond, cellc, rel);
            const size_t nslot = (size_t)ncell * cmax;
            hipLaunchKernelGGL(bcell_rank_kernel, dim3((unsigned int)((nslot + 255) / 256)),
                               dim3(256), 0, stream, cell_rep, rst, ncell, cmax, cell_size,
                               cell_raw, cell_idx);
            bchk(hipGetLastError(), "sort binning");
            read_cond();
            if (h_cond->cell_need <= cmax)
                break;
            if (attempt > 4)
                throw BError{HM_ERR_LAUNCH, "batch sort: cell list kept overflowing"};
            cmax = round_up(h_cond->cell_need + 8, 8);
            alloc_lists();
        }
        const size_t nslot = (size_t)ncell * cmax;
        hipLaunchKernelGGL(bcell_start_kernel, dim3(nrep), dim3(64), 0, stream, reps, cell_size,
                           cmax, cell_start);
        hipLaunchKernelGGL(bcell_order_kernel, dim3((unsigned int)((nslot + 255) / 256)),
                           dim3(256), 0, stream, ncell, cmax, cell_size, cell_start, cell_idx,
                           order);
        bchk(hipGetLastError(), "sort order");
        auto permute = [&](void* arr, unsigned int bytes) {
            bchk(gather_bytes(arr, stage, order, N, bytes, stream), "permute");
            bchk(hipMemcpyAsync(arr, stage, (size_t)N * bytes, hipMemcpyDeviceToDevice, stream),
                 "permute copy");
        };
        permute(pos, sizeof(double4));
        permute(vel, sizeof(double4));
        permute(accel, sizeof(d3));
        permute(image, sizeof(i3));
        permute(last, sizeof(d3));
        permute(force, sizeof(float4));
        permute(tag, sizeof(unsigned int));
        // kinetic-energy partials of the new tiles (accelerations rewritten unchanged), then
        // every replica rebuilds its lists (rows hold old indices) before the next forces
        hipLaunchKernelGGL(bstep_two_kernel, dim3(ntile), dim3(BTILE), 0, stream, reps, tile_rep,
                           rst, bst, force, vel, accel, ke_partial, dt, 0);
        hipLaunchKernelGGL(bset_rebuild_kernel, dim3((nrep + 255) / 256), dim3(256), 0, stream,
                           rst, nrep);
        bchk(hipGetLastError(), "sort finish");
        last_sort = timestep;
        n_sorts++;
    }

    int prime(uint64_t nsteps)
    {
        if (!prepared)
            prep();
        int added = 0;
        for (uint64_t t = timestep, end = timestep + nsteps; t < end;)
        {
            const unsigned int K = (unsigned int)chunk_len(t, end);
            if (graphs.find(K) == graphs.end())
            {
                graphs.emplace(K, capture_steps(K));
                added++;
            }
            t += K;
        }
        return added;
    }

    int run(uint64_t nsteps)
    {
        if (!prepared)
            prep();
        const uint64_t end = timestep + nsteps;
        int retries = 0;
        while (timestep < end)
        {
            if (sort_period && timestep % sort_period == 0 && last_sort != timestep)
                sort_cells();
            const unsigned int K = (unsigned int)chunk_len(timestep, end);
            bchk(hipMemsetAsync(cond, 0, sizeof(BCond), stream), "clear conditions");
            snapshot();
            auto it = graphs.find(K);
            if (it == graphs.end())
                it = graphs.emplace(K, capture_steps(K)).first;
            bchk(hipGraphLaunch(it->second, stream), "graph launch");
            read_cond();
            const BCond c = *h_cond;
            if (c.cell_need > cmax || c.nlist_need > nmax)
            {
                // regrow and replay the chunk from its start (the reference's regrow-and-repeat,
                // NeighborList.cc:411-427, CellList.cc:183-196)
                restore();
                if (c.cell_need > cmax)
                    cmax = round_up(c.cell_need + 8, 8);
                if (c.nlist_need > nmax)
                    nmax = round_up(c.nlist_need + 16, 16);
                alloc_lists();
                hipLaunchKernelGGL(bset_rebuild_kernel, dim3((nrep + 255) / 256), dim3(256), 0,
                                   stream, rst, nrep);
                bchk(hipGetLastError(), "force rebuild");
                rollbacks++;
                if (++retries > 16)
                    throw BError{HM_ERR_LAUNCH, "batch: neighbour list kept overflowing"};
                continue;
            }
            if (c.nan_idx)
            {
                restore();
                bchk(hipStreamSynchronize(stream), "sync");
                throw BError{HM_ERR_NAN, "batch: particle position is NaN"};
            }
            retries = 0;
            timestep += K;
            steps_run += K;
        }
        std::vector<RepState> h(nrep);
        bchk(hipMemcpy(h.data(), rst, sizeof(RepState) * nrep, hipMemcpyDeviceToHost), "state");
        for (unsigned int r = 0; r < nrep; r++)
            if (h[r].error & ERR_ZERO_KE)
                throw BError{HM_ERR_THERMO,
                             "batch: Bussi thermostat requires non-zero initial momenta"};
        return HM_OK;
    }

    // ---------------------------------------------------------------- autotuner
    // Times every (tpp, block) of the pair kernel on the current lists (median of 5 launches)
    // and keeps the fastest. The choice is cached per (total particles, mean density) bucket for
    // the life of the process, so batches of the same shape skip the sweep.
    float time_pair(unsigned int iters)
    {
        hipEvent_t a, b;
        bchk(hipEventCreate(&a), "event");
        bchk(hipEventCreate(&b), "event");
        launch_pair();
        bchk(hipEventRecord(a, stream), "event");
        for (unsigned int k = 0; k < iters; k++)
            launch_pair();
        bchk(hipEventRecord(b, stream), "event");
        bchk(hipEventSynchronize(b), "event sync");
        float ms = 0.f;
        bchk(hipEventElapsedTime(&ms, a, b), "elapsed");
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        return ms / (float)iters;
    }

    void tune();

    // index at creation of the particle now at each position (replica offset + its tag)
    std::vector<unsigned int> creation_index()
    {
        std::vector<unsigned int> t(N), pr(N);
        bchk(hipMemcpy(t.data(), tag, sizeof(unsigned int) * N, hipMemcpyDeviceToHost), "tag");
        for (const Rep& R : hrep)
            for (unsigned int k = 0; k < R.n; k++)
                t[R.off + k] += R.off;
        return t;
    }
};

namespace
{
struct TuneKey
{
    unsigned int n_bucket, rho_bucket;
    bool operator<(const TuneKey& o) const
    {
        return n_bucket != o.n_bucket ? n_bucket < o.n_bucket : rho_bucket < o.rho_bucket;
    }
};
std::map<TuneKey, std::pair<unsigned int, unsigned int>> g_batch_tune_cache;

TuneKey tune_key(const hm_batch& b)
{
    double vol = 0.0;
    for (const Rep& R : b.hrep)
        vol += R.L * R.L * R.L;
    const double rho = b.N / vol;
    return TuneKey{(unsigned int)std::lround(std::log2((double)b.N) * 2.0),
                   (unsigned int)std::lround(rho * 20.0)};
}
} // namespace

void hm_batch::tune()
{
    if (!prepared)
        prep();
    const TuneKey key = tune_key(*this);
    auto it = g_batch_tune_cache.find(key);
    tune_log.clear();
    if (it != g_batch_tune_cache.end())
    {
        pair_tpp = it->second.first;
        pair_block = it->second.second;
        tuned = true;
        return;
    }
    float best = 1e30f;
    unsigned int bt = pair_tpp, bb = pair_block;
    for (unsigned int tpp : {1u, 2u, 4u, 8u, 16u})
        for (unsigned int blk : {64u, 128u, 256u})
        {
            pair_tpp = tpp;
            pair_block = blk;
            float s[5];
            for (float& v : s)
                v = time_pair(2);
            std::sort(s, s + 5);
            tune_log.insert(tune_log.end(), {(float)tpp, (float)blk, s[2]});
            if (s[2] < best)
            {
                best = s[2];
                bt = tpp;
                bb = blk;
            }
        }
    pair_tpp = bt;
    pair_block = bb;
    tuned = true;
    g_batch_tune_cache[key] = {bt, bb};
    clear_graphs();
}

// ============================================================================ C ABI
namespace
{
template<class F> int bguard(F&& f)
{
    try
    {
        return f();
    }
    catch (const BError& e)
    {
        return fail(e.code, e.what.c_str());
    }
    catch (const std::bad_alloc&)
    {
        return fail(HM_ERR_ALLOC, "host allocation failed");
    }
}
} // namespace

extern "C"
{

hm_batch* hm_batch_create(unsigned int nrep, const unsigned int* n, const double* L,
                          const double* kT, const uint16_t* seed, const double* pos3,
                          const double* vel3, const double* mass, double epsilon, double sigma,
                          double r_cut, double r_buff, double dt, double tau)
{
    if (nrep == 0 || !n || !L || !kT || !seed || !pos3 || !(r_cut > 0.0) || !(r_buff >= 0.0)
        || !(dt > 0.0) || !(tau >= 0.0))
    {
        fail(HM_ERR_ARGS, "hm_batch_create: invalid arguments");
        return nullptr;
    }
    hm_batch* b = new (std::nothrow) hm_batch();
    if (!b)
        return nullptr;
    int r = bguard([&]() {
        const double rlist = r_cut + r_buff;
        b->nrep = nrep;
        b->dt = dt;
        b->tau = tau;
        b->r_cut = r_cut;
        b->r_buff = r_buff;
        b->rlistsq = rlist * rlist;
        b->maxsq = distance_check_maxsq(r_cut, r_buff);
        // EvaluatorPairLJ parameters (lj1 = 4 eps sigma^12, lj2 = 4 eps sigma^6)
        const double s6 = sigma * sigma * sigma * sigma * sigma * sigma;
        b->pc.lj1 = (float)(4.0 * epsilon * s6 * s6);
        b->pc.lj2 = (float)(4.0 * epsilon * s6);
        b->pc.rcutsq = r_cut * r_cut;
        b->pc.rcutsq_f = (float)(r_cut * r_cut);
        {
            // FP32 pair-vector error per component, u (3 w + 2 r_buff + r_cut): the two offsets
            // (each within w / 2 + r_buff / 2 of its cell centre), their difference, o w and
            // the final subtraction; then r^2 within 2 sqrt(3) r_cut e + 3 u r_cut^2, taken 4x
            const double u = 1.0 / 16777216.0, rc = r_cut, s3 = std::sqrt(3.0);
            b->pc.guard_w = (float)(4.0 * u * 2.0 * s3 * rc * 3.0);
            b->pc.guard_0 = (float)(4.0 * u * (2.0 * s3 * rc * (2.0 * r_buff + rc) + 3.0 * rc * rc));
        }
        b->hrep.resize(nrep);
        unsigned long long off = 0, coff = 0, toff = 0;
        double rho_max = 0.0, cell_fill_max = 0.0;
        for (unsigned int q = 0; q < nrep; q++)
        {
            Rep& R = b->hrep[q];
            if (n[q] == 0 || !(L[q] > 0.0) || !(kT[q] >= 0.0))
                throw BError{HM_ERR_ARGS, "hm_batch_create: replica with no particles or box"};
            // NeighborList::checkBoxSize (hoomd/md/NeighborList.cc:538-561): the list range
            // must be below half the box
            if (L[q] <= 2.0 * rlist)
                throw BError{HM_ERR_BOXSMALL,
                             "hm_batch_create: simulation box too small for the neighbour list"};
            R.L = L[q];
            R.Linv = 1.0 / L[q];
            R.hi = L[q] / 2.0;
            R.kT = kT[q];
            R.ndof = 3.0 * n[q] - 3.0;
            // CellList::computeDimensions (hoomd/CellList.cc:83-131): floor(L / r_list) cells
            R.dim = std::max(1u, (unsigned int)std::floor(L[q] / rlist));
            R.cell_inv = R.dim / L[q];
            R.w = L[q] / R.dim;
            R.off = (unsigned int)off;
            R.n = n[q];
            R.cell_off = (unsigned int)coff;
            R.tile_off = (unsigned int)toff;
            R.ntile = (n[q] + BTILE - 1) / BTILE;
            R.seed = seed[q];
            off += n[q];
            coff += (unsigned long long)R.dim * R.dim * R.dim;
            toff += R.ntile;
            // densest cell: the replica's density times its real cell volume (w = L / dim can
            // be up to twice r_list)
            const double w = L[q] / R.dim;
            cell_fill_max = std::max(cell_fill_max, n[q] / (L[q] * L[q] * L[q]) * w * w * w);
            rho_max = std::max(rho_max, n[q] / (L[q] * L[q] * L[q]));
            // row entries hold a particle index in BCODE_SHIFT bits; cell coordinates in 10
            if (off >= (1ull << BCODE_SHIFT) || coff >= (1ull << 31) || R.dim >= 1024u)
                throw BError{HM_ERR_ARGS, "hm_batch_create: too many particles or cells"};
        }
        b->N = (unsigned int)off;
        b->ncell = (unsigned int)coff;
        b->ntile = (unsigned int)toff;
        // first list sizes from the densest replica; overflow regrows them
        b->cmax = std::min(NL_HOME_MAX, round_up((unsigned int)(cell_fill_max * 1.5) + 16, 8));
        b->nmax = round_up((unsigned int)(4.18879 * rlist * rlist * rlist * rho_max * 1.4) + 32,
                           16);
        bchk(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking), "stream");
        const unsigned int N = b->N;
        b->reps = balloc<Rep>(nrep);
        b->rst = balloc<RepState>(nrep);
        b->bst = balloc<BatchStep>(1);
        b->cond = balloc<BCond>(1);
        if (hipHostMalloc((void**)&b->h_cond, sizeof(BCond), hipHostMallocDefault) != hipSuccess)
            throw BError{HM_ERR_ALLOC, "pinned allocation failed"};
        b->tile_rep = balloc<unsigned int>(b->ntile);
        b->cell_rep = balloc<unsigned int>(b->ncell);
        b->part_rep = balloc<unsigned int>(N);
        b->pos = balloc<double4>(N);
        b->vel = balloc<double4>(N);
        b->accel = balloc<d3>(N);
        b->image = balloc<i3>(N);
        b->last = balloc<d3>(N);
        b->force = balloc<float4>(N);
        b->ke_partial = balloc<double>(b->ntile);
        b->cell_size = balloc<unsigned int>(b->ncell);
        b->n_neigh = balloc<unsigned int>(N);
        b->tag = balloc<unsigned int>(N);
        b->rel = balloc<float4>(N);
        b->cellc = balloc<unsigned int>(N);
        bchk(hipMemset(b->rel, 0, sizeof(float4) * N), "clear offsets");
        bchk(hipMemset(b->cellc, 0, sizeof(unsigned int) * N), "clear cells");
        b->alloc_lists();
        std::vector<unsigned int> trep(b->ntile), crep(b->ncell), prep_(N), htag(N);
        std::vector<double4> hp(N), hv(N);
        for (unsigned int q = 0; q < nrep; q++)
        {
            const Rep& R = b->hrep[q];
            for (unsigned int t = 0; t < R.ntile; t++)
                trep[R.tile_off + t] = q;
            for (unsigned int c = 0; c < R.dim * R.dim * R.dim; c++)
                crep[R.cell_off + c] = q;
            for (unsigned int i = 0; i < R.n; i++)
            {
                const size_t p = R.off + i;
                double x[3];
                for (int d = 0; d < 3; d++)
                {
                    // positions are wrapped into [-L/2, L/2) as the snapshot reader does
                    double w = pos3[3 * p + d];
                    w -= R.L * std::floor((w + R.hi) * R.Linv);
                    if (w >= R.hi)
                        w -= R.L;
                    x[d] = w;
                }
                hp[p] = make_double4(x[0], x[1], x[2], 0.0);
                hv[p] = make_double4(vel3 ? vel3[3 * p] : 0.0, vel3 ? vel3[3 * p + 1] : 0.0,
                                     vel3 ? vel3[3 * p + 2] : 0.0, mass ? mass[p] : 1.0);
                prep_[p] = q;
                htag[p] = i;
            }
        }
        auto up = [&](void* d, const void* s, size_t bytes) {
            bchk(hipMemcpy(d, s, bytes, hipMemcpyHostToDevice), "upload");
        };
        up(b->reps, b->hrep.data(), sizeof(Rep) * nrep);
        up(b->tile_rep, trep.data(), sizeof(unsigned int) * b->ntile);
        up(b->cell_rep, crep.data(), sizeof(unsigned int) * b->ncell);
        up(b->part_rep, prep_.data(), sizeof(unsigned int) * N);
        up(b->tag, htag.data(), sizeof(unsigned int) * N);
        up(b->pos, hp.data(), sizeof(double4) * N);
        up(b->vel, hv.data(), sizeof(double4) * N);
        bchk(hipMemset(b->image, 0, sizeof(i3) * N), "clear");
        bchk(hipMemset(b->accel, 0, sizeof(d3) * N), "clear");
        bchk(hipMemset(b->rst, 0, sizeof(RepState) * nrep), "clear");
        bchk(hipMemset(b->bst, 0, sizeof(BatchStep)), "clear");
        bchk(hipMemset(b->cell_size, 0, sizeof(unsigned int) * b->ncell), "clear");
        return HM_OK;
    });
    if (r != HM_OK)
    {
        delete b;
        return nullptr;
    }
    return b;
}

void hm_batch_destroy(hm_batch* b)
{
    delete b;
}

int hm_batch_thermalize(hm_batch* b, uint64_t timestep)
{
    if (!b)
        return fail(HM_ERR_ARGS, "hm_batch_thermalize: NULL batch");
    return bguard([&]() {
        // ParticleGroup::thermalizeParticleMomenta per replica: tags 0..n-1 of the replica, its
        // seed and kT, its own COM momentum removed
        double* scratch = balloc<double>(THERMO_BLOCKS * 3);
        hipError_t e = hipSuccess;
        for (const Rep& R : b->hrep)
        {
            if (e != hipSuccess)
                break;
            e = launch_thermalize(R.n, b->vel + R.off, b->tag + R.off, R.kT, timestep,
                                  (uint16_t)R.seed, scratch, b->stream);
        }
        if (e == hipSuccess)
            e = hipStreamSynchronize(b->stream);
        bfree(scratch);
        bchk(e, "thermalize");
        b->prepared = false;
        return HM_OK;
    });
}

int hm_batch_set_timestep(hm_batch* b, uint64_t timestep)
{
    if (!b)
        return fail(HM_ERR_ARGS, "hm_batch_set_timestep: NULL batch");
    return bguard([&]() {
        BatchStep s{timestep};
        bchk(hipMemcpy(b->bst, &s, sizeof(s), hipMemcpyHostToDevice), "timestep");
        b->timestep = timestep;
        return HM_OK;
    });
}

int hm_batch_run(hm_batch* b, uint64_t nsteps)
{
    if (!b)
        return fail(HM_ERR_ARGS, "hm_batch_run: NULL batch");
    return bguard([&]() { return b->run(nsteps); });
}

int hm_batch_prime(hm_batch* b, uint64_t nsteps)
{
    if (!b)
        return fail(HM_ERR_ARGS, "hm_batch_prime: NULL batch");
    return bguard([&]() { return b->prime(nsteps); });
}

int hm_batch_set_sort_period(hm_batch* b, unsigned int period)
{
    if (!b)
        return fail(HM_ERR_ARGS, "hm_batch_set_sort_period: NULL batch");
    b->sort_period = period;
    b->clear_graphs();
    return HM_OK;
}

int hm_batch_set_chunk(hm_batch* b, unsigned int steps)
{
    if (!b || steps == 0)
        return fail(HM_ERR_ARGS, "hm_batch_set_chunk: invalid arguments");
    b->chunk = steps;
    return HM_OK;
}

int hm_batch_tune(hm_batch* b, float* log3, unsigned int max_rows)
{
    if (!b)
        return fail(HM_ERR_ARGS, "hm_batch_tune: NULL batch");
    return bguard([&]() {
        b->tune();
        const unsigned int rows = (unsigned int)(b->tune_log.size() / 3);
        if (log3)
            std::memcpy(log3, b->tune_log.data(), sizeof(float) * 3 * std::min(rows, max_rows));
        return (int)rows;
    });
}

int hm_batch_set_pair_launch(hm_batch* b, unsigned int tpp, unsigned int block)
{
    if (!b || (tpp != 1 && tpp != 2 && tpp != 4 && tpp != 8 && tpp != 16)
        || (block != 64 && block != 128 && block != 256))
        return fail(HM_ERR_ARGS, "hm_batch_set_pair_launch: tpp in {1,2,4,8,16}, block in "
                                 "{64,128,256}");
    b->pair_tpp = tpp;
    b->pair_block = block;
    b->clear_graphs();
    return HM_OK;
}

int hm_batch_config(const hm_batch* b, unsigned int out[6])
{
    if (!b || !out)
        return fail(HM_ERR_ARGS, "hm_batch_config: invalid arguments");
    out[0] = b->pair_tpp;
    out[1] = b->pair_block;
    out[2] = b->tuned ? 1u : 0u;
    out[3] = b->cmax;
    out[4] = b->nmax;
    out[5] = b->ncell;
    return HM_OK;
}

int hm_batch_get_state(hm_batch* b, double* pos3, double* vel3, int* image3)
{
    if (!b)
        return fail(HM_ERR_ARGS, "hm_batch_get_state: NULL batch");
    return bguard([&]() {
        const unsigned int N = b->N;
        std::vector<double4> p(N), v(N);
        std::vector<i3> im(N);
        bchk(hipStreamSynchronize(b->stream), "sync");
        bchk(hipMemcpy(p.data(), b->pos, sizeof(double4) * N, hipMemcpyDeviceToHost), "pos");
        bchk(hipMemcpy(v.data(), b->vel, sizeof(double4) * N, hipMemcpyDeviceToHost), "vel");
        bchk(hipMemcpy(im.data(), b->image, sizeof(i3) * N, hipMemcpyDeviceToHost), "image");
        const std::vector<unsigned int> where = b->creation_index();
        for (unsigned int n = 0; n < N; n++)
        {
            const unsigned int i = where[n];
            if (pos3)
            {
                pos3[3 * i] = p[n].x;
                pos3[3 * i + 1] = p[n].y;
                pos3[3 * i + 2] = p[n].z;
            }
            if (vel3)
            {
                vel3[3 * i] = v[n].x;
                vel3[3 * i + 1] = v[n].y;
                vel3[3 * i + 2] = v[n].z;
            }
            if (image3)
            {
                image3[3 * i] = im[n].x;
                image3[3 * i + 1] = im[n].y;
                image3[3 * i + 2] = im[n].z;
            }
        }
        return HM_OK;
    });
}

int hm_batch_get_forces(hm_batch* b, double* force4)
{
    if (!b || !force4)
        return fail(HM_ERR_ARGS, "hm_batch_get_forces: invalid arguments");
    return bguard([&]() {
        if (!b->prepared)
            b->prep();
        std::vector<float4> f(b->N);
        bchk(hipStreamSynchronize(b->stream), "sync");
        bchk(hipMemcpy(f.data(), b->force, sizeof(float4) * b->N, hipMemcpyDeviceToHost),
             "force");
        const std::vector<unsigned int> where = b->creation_index();
        for (unsigned int n = 0; n < b->N; n++)
        {
            const unsigned int i = where[n];
            force4[4 * i] = f[n].x;
            force4[4 * i + 1] = f[n].y;
            force4[4 * i + 2] = f[n].z;
            force4[4 * i + 3] = f[n].w;
        }
        return HM_OK;
    });
}

// per replica: out[r * 4 + 0] = kinetic energy, +1 = potential energy, +2 = list builds since
// creation, +3 = neighbour-list entries
int hm_batch_replica_stats(hm_batch* b, double* out4)
{
    if (!b || !out4)
        return fail(HM_ERR_ARGS, "hm_batch_replica_stats: invalid arguments");
    return bguard([&]() {
        if (!b->prepared)
            b->prep();
        const unsigned int N = b->N;
        std::vector<double4> v(N);
        std::vector<float4> f(N);
        std::vector<unsigned int> nn(N);
        std::vector<RepState> rs(b->nrep);
        bchk(hipStreamSynchronize(b->stream), "sync");
        bchk(hipMemcpy(v.data(), b->vel, sizeof(double4) * N, hipMemcpyDeviceToHost), "vel");
        bchk(hipMemcpy(f.data(), b->force, sizeof(float4) * N, hipMemcpyDeviceToHost), "force");
        bchk(hipMemcpy(nn.data(), b->n_neigh, sizeof(unsigned int) * N, hipMemcpyDeviceToHost),
             "n_neigh");
        bchk(hipMemcpy(rs.data(), b->rst, sizeof(RepState) * b->nrep, hipMemcpyDeviceToHost),
             "state");
        for (unsigned int q = 0; q < b->nrep; q++)
        {
            const Rep& R = b->hrep[q];
            double ke = 0.0, pe = 0.0, ent = 0.0;
            for (unsigned int i = R.off; i < R.off + R.n; i++)
            {
                ke += 0.5 * v[i].w * (v[i].x * v[i].x + v[i].y * v[i].y + v[i].z * v[i].z);
                pe += f[i].w;
                ent += nn[i];
            }
            out4[4 * q] = ke;
            out4[4 * q + 1] = pe;
            out4[4 * q + 2] = rs[q].n_builds;
            out4[4 * q + 3] = ent;
        }
        return HM_OK;
    });
}

// totals: out[0] particles, out[1] steps run, out[2] overflow replays, out[3] list builds (all
// replicas)
int hm_batch_stats(hm_batch* b, uint64_t out4[4])
{
    if (!b || !out4)
        return fail(HM_ERR_ARGS, "hm_batch_stats: invalid arguments");
    return bguard([&]() {
        std::vector<RepState> rs(b->nrep);
        bchk(hipStreamSynchronize(b->stream), "sync");
        bchk(hipMemcpy(rs.data(), b->rst, sizeof(RepState) * b->nrep, hipMemcpyDeviceToHost),
             "state");
        uint64_t nb = 0;
        for (const RepState& s : rs)
            nb += s.n_builds;
        out4[0] = b->N;
        out4[1] = b->steps_run;
        out4[2] = b->rollbacks;
        out4[3] = nb;
        return HM_OK;
    });
}

// Replace the cell and row capacities (test hook: capacities below the need make the next
// rebuild overflow, which the run loop answers by replaying the chunk with grown lists). Every
// replica rebuilds its lists before the forces of the next step.
int hm_batch_set_list_sizes(hm_batch* b, unsigned int cmax, unsigned int nmax)
{
    if (!b || cmax == 0 || nmax == 0)
        return fail(HM_ERR_ARGS, "hm_batch_set_list_sizes: invalid arguments");
    return bguard([&]() {
        if (!b->prepared)
            b->prep();
        b->cmax = cmax;
        b->nmax = nmax;
        b->alloc_lists();
        hipLaunchKernelGGL(bset_rebuild_kernel, dim3((b->nrep + 255) / 256), dim3(256), 0,
                           b->stream, b->rst, b->nrep);
        bchk(hipGetLastError(), "force rebuild");
        bchk(hipStreamSynchronize(b->stream), "sync");
        return HM_OK;
    });
}

// Average duration of the pair launch over iters back-to-back launches on the batch stream.
int hm_batch_time_pair(hm_batch* b, unsigned int iters, double* avg_ms)
{
    if (!b || !avg_ms || iters == 0)
        return fail(HM_ERR_ARGS, "hm_batch_time_pair: invalid arguments");
    return bguard([&]() {
        if (!b->prepared)
            b->prep();
        *avg_ms = b->time_pair(iters);
        return HM_OK;
    });
}

} // extern "C"
