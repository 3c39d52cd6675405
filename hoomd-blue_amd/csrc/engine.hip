// engine.hip -- device-resident particle system that runs whole timesteps natively.
//
// Host structure (what the reference spreads over System, IntegratorTwoStep, NeighborList,
// CellList, PotentialPair, PotentialBond, TwoStepConstantVolume / TwoStepLangevin, the Bussi
// thermostat and SFCPackTuner) collapses into one object that owns every per-particle array in
// HBM. A timestep is a fixed launch sequence
//     step-begin (timestep, Bussi factor) -> step one (+ distance check, clear cell counts)
//     -> cell build* -> neighbour list build* -> pair/bond forces -> step two (+ KE partials)
// where * runs only when the device-side distance check (or a forced update) requested a
// rebuild for this step. Nothing in the sequence reads back to the host, so chunks of steps are
// captured once as a hipGraph and replayed. Conditions (overflow, NaN, out-of-box, bond stretch)
// are read once per chunk; an overflow restores the chunk's starting state, grows the buffer
// and replays the chunk, which gives the reference's regrow-and-repeat semantics
// (NeighborList.cc:411-427, CellList.cc:183-196) without a per-step host sync.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "capi_util.h"
#include "mlp.h"
#include "kernels.h"

using namespace hm;

namespace
{
struct HmError
{
    int code;
    std::string what;
};

inline void hchk(hipError_t e, const char* what)
{
    if (e != hipSuccess)
        throw HmError{HM_ERR_LAUNCH, std::string(what) + ": " + hipGetErrorString(e)};
}

template<class T> T* dalloc(size_t n)
{
    void* p = nullptr;
    if (hipMalloc(&p, sizeof(T) * (n ? n : 1)) != hipSuccess)
        throw HmError{HM_ERR_ALLOC, "device allocation failed"};
    return (T*)p;
}

template<class T> void dfree(T*& p)
{
    if (p)
        (void)hipFree((void*)p);
    p = nullptr;
}

template<class T> T* halloc(size_t n)
{
    void* p = nullptr;
    if (hipHostMalloc(&p, sizeof(T) * (n ? n : 1), hipHostMallocDefault) != hipSuccess)
        throw HmError{HM_ERR_ALLOC, "pinned host allocation failed"};
    memset(p, 0, sizeof(T) * (n ? n : 1));
    return (T*)p;
}

struct ForceSlot
{
    int kind = 0; // 0 LJ pair, 1 FENE bond, 2 Table pair, 3 learned CG site energy, 4 harmonic bond
    float4* f = nullptr;
    float* v = nullptr;
    // pair forces
    double rcut[HM_MAX_TYPES * HM_MAX_TYPES] = {};
    double ron[HM_MAX_TYPES * HM_MAX_TYPES] = {};
    hm_lj_params p[HM_MAX_TYPES * HM_MAX_TYPES] = {};
    int mode = 0;
    // Table: U then F of every type pair in one device array
    float* tab = nullptr;
    hm_table_params tp[HM_MAX_TYPES * HM_MAX_TYPES] = {};
    // FENE, CSR by tag
    unsigned int* row = nullptr;
    unsigned int* other = nullptr;
    unsigned int* btype = nullptr;
    unsigned int nbt = 1;
    hm_fene_params fp[HM_MAX_TYPES] = {};
    unsigned int* flag = nullptr;
    // partners (and types, with several bond types) in the current index order, [maxb * N]
    unsigned int* bpart = nullptr;
    unsigned int* bptype = nullptr;
    unsigned int maxb = 0;
    hm_harmonic_params hp[HM_MAX_TYPES] = {};
    // learned CG site energy: FP32 parameters, bf16 operand copies, per-site densities,
    // dE/dG and site energies; training buffers are allocated by the first training call
    hm_mlp_cg_model mm{};
    float* mparams = nullptr;
    uint16_t* mbf16 = nullptr;
    float4 *mG = nullptr, *mA = nullptr;
    uint2* mAb = nullptr;  // bf16 copy of mA for the force pass
    float* mE = nullptr;
    float4 *mT = nullptr, *mErr = nullptr, *mC = nullptr;
    float *mLoss = nullptr, *mGpart = nullptr, *mM1 = nullptr, *mM2 = nullptr, *mGrad = nullptr;
    float* mGacc = nullptr;  // minibatch gradient sum
    double* mLossOut = nullptr;
    // optional tabulated pair term of the force pass: U then F, mtab_w points on
    // [mtab_rmin, r_cut)
    float* mtab = nullptr;
    unsigned int mtab_w = 0;
    double mtab_rmin = 0.0;
    unsigned int adam_t = 0;
};
bool is_pair(const ForceSlot& f)
{
    return f.kind == 0 || f.kind == 2 || f.kind == 3;
}

// out[3 tag + d] -= f[i].d for every particle i (frames are stored by tag)
__global__ __launch_bounds__(256) void subtract_by_tag_kernel(unsigned int N,
                                                              const unsigned int* __restrict__ tag,
                                                              const float4* __restrict__ f,
                                                              float* __restrict__ out)
{
    const unsigned int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N)
        return;
    const float4 v = f[i];
    const size_t t = tag[i];
    out[3 * t] -= v.x;
    out[3 * t + 1] -= v.y;
    out[3 * t + 2] -= v.z;
}
bool is_bond(const ForceSlot& f)
{
    return f.kind == 1 || f.kind == 4;
}
} // namespace

namespace
{
struct TuneChoice
{
    PairLaunch pair;
    int nl_shape;
    unsigned int nl_grid;
};
// launch shapes chosen by the autotuner, per kind of system, for the life of the process
std::map<std::string, TuneChoice>& g_tune_cache()
{
    static std::map<std::string, TuneChoice> cache;
    return cache;
}
} // namespace

struct hm_system
{
    unsigned int N = 0, ntypes = 1;
    hm_box hbox{};
    DevBox box{};
    uint16_t seed = 0;
    hipStream_t stream = nullptr;

    double4 *pos = nullptr, *vel = nullptr;
    d3* accel = nullptr;
    i3* image = nullptr;
    unsigned int *tag = nullptr, *rtag = nullptr;
    d3* last = nullptr;

    std::vector<ForceSlot> forces;

    // neighbour list
    double r_buff = 0.4;
    int exclude_bonds = 1;
    unsigned int check_delay = 1;  // NeighborList rebuild_check_delay
    int dist_check = 1;            // NeighborList check_dist
    unsigned long long run_builds_base = 0;
    unsigned int dim[3] = {0, 0, 0};
    unsigned int ncell = 0, n_adj = 0, cmax = 0, nmax = 0;
    unsigned int *cell_size = nullptr, *cell_adj = nullptr, *nlist = nullptr, *n_neigh = nullptr;
    hm_cell_entry* entries = nullptr;
    unsigned int* cell_idx = nullptr;
    // binning pass output: unordered particle index of each slot [ncell * cmax], unordered slot
    // of each particle [N]
    unsigned int *ckey = nullptr, *chome = nullptr;
    unsigned short* lrow = nullptr;
    // pair kernel kind 2: cell-slot relative FP32 positions [ncell * cmax], packed cell and
    // slot of every particle [N] (RelFrame)
    float3* rel = nullptr;
    uint32_t* relw = nullptr;
    uint32_t *pcell = nullptr, *slot_of = nullptr;
    double rcut_mat[HM_MAX_TYPES * HM_MAX_TYPES] = {};
    double rcut_max = 0.0;
    unsigned int max_excl = 0;
    unsigned int *n_excl_tag = nullptr, *excl_tag = nullptr, *n_excl = nullptr, *excl = nullptr;

    // device state
    StepState* st = nullptr;
    Conditions* cond = nullptr;
    StepState* h_st = nullptr;
    Conditions* h_cond = nullptr;
    double* ke_partial = nullptr;
    unsigned int n_partial = 0;
    double* thermo_partial = nullptr;
    double* thermo_out = nullptr;
    double* h_thermo = nullptr;

    // method
    int method = -1;
    double dt = 0.005, kT = 1.0, tau = 0.0;
    double gamma[HM_MAX_TYPES] = {1.0, 1.0, 1.0, 1.0};
    double* d_gamma = nullptr;

    // sorter
    unsigned int sort_period = 0, sort_grid = 256;
    unsigned long long *keys_a = nullptr, *keys_b = nullptr;
    void* sort_tmp = nullptr;
    size_t sort_tmp_bytes = 0;
    unsigned int* order = nullptr;
    void* scratch = nullptr; // N * 32 bytes
    uint64_t last_sort_step = ~0ull;

    // run control
    uint64_t timestep = 0;
    bool prepared = false;
    int compute_virial = 0;
    // FP32 copy of the positions for the learned-CG pair passes (allocated with the first such
    // force); written by step one, so inside a step the forces read it as it is
    float4* pos32 = nullptr;
    bool in_step = false;
    PairLaunch pair_cfg{4, 256};      // gather kernels: lanes per particle, block size
    PairLaunch pair_cell_cfg{8, 0};   // cell-tiled kernel: lanes per particle, workgroups
    PairLaunch pair_tile_cfg{8, 2048, 8};  // FP32 cell tiles: lanes, workgroups, row batch
    // 0 cell-tiled (FP64 staging), 1 neighbour-index gather of FP64 positions, 2 gather of FP32
    // cell-slot relative positions, 3 cell-tiled FP32 relative positions (kinds 2 and 3 need
    // 3..255 cells per axis, else kind 1 runs)
    int pair_kernel_kind = 2;
    bool autotune = true;             // sweep the pair launch configuration at the next prep
    bool pair_tuned = false;
    // list-build launch shape (0: from the occupancy) and persistent grid, from the autotuner
    int nl_shape = 0;
    unsigned int nl_grid = 0;
    // second sweep once the system has run this many steps since its first prep (the lattice
    // start is more ordered than the liquid it melts into)
    uint64_t retune_after = 500;
    bool retuned = false;
    bool tune_cached = false;         // the last sweep was answered by the shape cache
    unsigned int chunk = 100;
    bool use_graph = true;
    std::map<unsigned int, hipGraphExec_t> graphs;

    // rollback snapshot
    double4 *s_pos = nullptr, *s_vel = nullptr;
    d3 *s_accel = nullptr, *s_last = nullptr;
    i3* s_image = nullptr;
    unsigned int* s_tag = nullptr;
    double* s_ke = nullptr;
    StepState s_state{};

    uint64_t n_sorts = 0, steps_run = 0, rollbacks = 0;

    // coarse-grained site map (hm_sys_set_site_map): CSR of particle tags per site
    unsigned int nsites = 0;
    unsigned int *site_row = nullptr, *site_tags = nullptr;

    ~hm_system() { release(); }

    void release()
    {
        clear_graphs();
        dfree(pos);
        dfree(pos32);
        pos32 = nullptr;
        dfree(vel);
        dfree(accel);
        dfree(image);
        dfree(tag);
        dfree(rtag);
        dfree(last);
        for (auto& f : forces)
        {
            dfree(f.f);
            dfree(f.v);
            dfree(f.row);
            dfree(f.other);
            dfree(f.btype);
            dfree(f.bpart);
            dfree(f.bptype);
            dfree(f.flag);
            dfree(f.tab);
            dfree(f.mparams);
            dfree(f.mbf16);
            dfree(f.mG);
            dfree(f.mA);
            dfree(f.mAb);
            dfree(f.mE);
            dfree(f.mT);
            dfree(f.mErr);
            dfree(f.mC);
            dfree(f.mLoss);
            dfree(f.mGpart);
            dfree(f.mGacc);
            dfree(f.mtab);
            dfree(f.mM1);
            dfree(f.mM2);
            dfree(f.mGrad);
            dfree(f.mLossOut);
        }
        dfree(site_row);
        dfree(site_tags);
        forces.clear();
        dfree(cell_size);
        dfree(cell_adj);
        dfree(nlist);
        dfree(n_neigh);
        dfree(entries);
        dfree(cell_idx);
        dfree(ckey);
        dfree(chome);
        dfree(lrow);
        dfree(rel);
        dfree(relw);
        dfree(pcell);
        dfree(slot_of);
        dfree(n_excl_tag);
        dfree(excl_tag);
        dfree(n_excl);
        dfree(excl);
        dfree(st);
        dfree(cond);
        if (h_st)
            (void)hipHostFree(h_st);
        if (h_cond)
            (void)hipHostFree(h_cond);
        if (h_thermo)
            (void)hipHostFree(h_thermo);
        h_st = nullptr;
        h_cond = nullptr;
        h_thermo = nullptr;
        dfree(ke_partial);
        dfree(thermo_partial);
        dfree(thermo_out);
        dfree(d_gamma);
        dfree(keys_a);
        dfree(keys_b);
        if (sort_tmp)
            (void)hipFree(sort_tmp);
        sort_tmp = nullptr;
        dfree(order);
        if (scratch)
            (void)hipFree(scratch);
        scratch = nullptr;
        dfree(s_pos);
        dfree(s_vel);
        dfree(s_accel);
        dfree(s_last);
        dfree(s_image);
        dfree(s_tag);
        dfree(s_ke);
        if (stream)
            (void)hipStreamDestroy(stream);
        stream = nullptr;
    }

    void clear_graphs()
    {
        for (auto& kv : graphs)
            (void)hipGraphExecDestroy(kv.second);
        graphs.clear();
    }

    // ------------------------------------------------------------------ configuration
    void configure_nlist()
    {
        // NeighborList::updateRList: r_cut matrix = max over the pair forces (NeighborList.cc:452-532)
        for (auto& v : rcut_mat)
            v = 0.0;
        rcut_max = 0.0;
        for (auto& f : forces)
            if (is_pair(f))
                for (unsigned int q = 0; q < ntypes * ntypes; q++)
                    rcut_mat[q] = std::max(rcut_mat[q], f.rcut[q]);
        for (unsigned int q = 0; q < ntypes * ntypes; q++)
            rcut_max = std::max(rcut_max, rcut_mat[q]);
        if (rcut_max <= 0.0)
            return;
        // NeighborList::checkBoxSize (NeighborList.cc:538-561)
        double rmax = rcut_max + r_buff;
        for (int d = 0; d < 3; d++)
            if (hbox.L[d] <= rmax * 2.0)
                throw HmError{HM_ERR_BOXSMALL, "nlist: simulation box is too small for r_cut + buffer"};
        unsigned int nd[3];
        hm_cell_dims(&hbox, rmax, nd);
        unsigned int nc = nd[0] * nd[1] * nd[2];
        if (nc != ncell || !cell_adj)
        {
            dim[0] = nd[0];
            dim[1] = nd[1];
            dim[2] = nd[2];
            ncell = nc;
            n_adj = hm_cell_n_adj(dim);
            dfree(cell_size);
            dfree(cell_adj);
            dfree(entries);
            cell_size = dalloc<unsigned int>(ncell);
            cell_adj = dalloc<unsigned int>((size_t)ncell * n_adj);
            hchk(launch_cell_adj(dim, cell_adj, stream), "cell adjacency");
            cmax = 0;
        }
        double vol = hbox.L[0] * hbox.L[1] * hbox.L[2];
        double rho = N / vol;
        double w3 = (hbox.L[0] / dim[0]) * (hbox.L[1] / dim[1]) * (hbox.L[2] / dim[2]);
        if (cmax == 0)
        {
            cmax = (unsigned int)std::min(256.0, std::ceil(rho * w3 * 2.0 + 16.0));
            alloc_cells();
        }
        if (nmax == 0)
        {
            double expect = rho * 4.0 / 3.0 * M_PI * rmax * rmax * rmax;
            nmax = (unsigned int)std::ceil(expect * 1.3 + 16.0);
            nmax = (nmax + ROW_BLOCK - 1) & ~(ROW_BLOCK - 1);
            alloc_rows();
        }
        if (!pcell)
        {
            pcell = dalloc<uint32_t>(N);
            slot_of = dalloc<uint32_t>(N);
            chome = dalloc<unsigned int>(N);
            hchk(hipMemsetAsync(pcell, 0, sizeof(uint32_t) * N, stream), "clear");
            hchk(hipMemsetAsync(slot_of, 0, sizeof(uint32_t) * N, stream), "clear");
        }
        if (!n_neigh)
        {
            n_neigh = dalloc<unsigned int>(N);
            hchk(hipMemsetAsync(n_neigh, 0, sizeof(unsigned int) * N, stream), "clear nlist");
        }
        clear_graphs();
    }

    void grow_cmax(unsigned int need)
    {
        unsigned int c = std::min(256u, (unsigned int)std::ceil(need * 1.3 + 4.0));
        if (need > 256)
            throw HmError{HM_ERR_ARGS, "cell list: more than 256 particles in one cell"};
        if (c <= cmax)
            c = std::min(256u, cmax + 8);
        cmax = c;
        alloc_cells();
        clear_graphs();
    }

    void alloc_cells()
    {
        dfree(entries);
        dfree(cell_idx);
        dfree(ckey);
        dfree(rel);
        dfree(relw);
        entries = dalloc<hm_cell_entry>((size_t)ncell * cmax);
        cell_idx = dalloc<unsigned int>((size_t)ncell * cmax);
        ckey = dalloc<unsigned int>((size_t)ncell * cmax);
        rel = dalloc<float3>((size_t)ncell * cmax);
        relw = dalloc<uint32_t>((size_t)ncell * cmax);
        hchk(hipMemsetAsync(cell_idx, 0, sizeof(unsigned int) * ncell * cmax, stream), "clear");
        hchk(hipMemsetAsync(rel, 0, sizeof(float3) * ncell * cmax, stream), "clear");
        hchk(hipMemsetAsync(relw, 0, sizeof(uint32_t) * ncell * cmax, stream), "clear");
        if (slot_of)  // slots of the old table may exceed the new one until the next build
            hchk(hipMemsetAsync(slot_of, 0, sizeof(uint32_t) * N, stream), "clear");
    }

    // neighbour rows (32-bit particle indices) and local rows (16-bit staged positions);
    // zero-filled so that stale rows never hold an out-of-range index
    void alloc_rows()
    {
        dfree(nlist);
        dfree(lrow);
        nlist = dalloc<unsigned int>((size_t)N * nmax + NLIST_PAD);
        lrow = dalloc<unsigned short>((size_t)N * nmax + NLIST_PAD);
        hchk(hipMemsetAsync(nlist, 0, sizeof(unsigned int) * ((size_t)N * nmax + NLIST_PAD),
                            stream),
             "clear nlist");
        hchk(hipMemsetAsync(lrow, 0, sizeof(unsigned short) * ((size_t)N * nmax + NLIST_PAD),
                            stream),
             "clear nlist");
    }

    void grow_nmax(unsigned int need)
    {
        // Nmax rounds up to a multiple of ROW_BLOCK (the reference uses 4,
        // NeighborList.cc:1601-1617; rows with offset codes are stored in transposed blocks of
        // ROW_BLOCK entries); extra headroom avoids repeated replays when the maximum drifts
        // upward by one or two.
        unsigned int m = (unsigned int)std::ceil(need * 1.15 + 8.0);
        m = (m + ROW_BLOCK - 1) & ~(ROW_BLOCK - 1);
        if (m <= nmax)
            m = nmax + ROW_BLOCK;
        nmax = m;
        alloc_rows();
        clear_graphs();
    }

    void update_bond_index()
    {
        for (auto& f : forces)
            if (is_bond(f) && f.bpart)
                hchk(launch_bond_index(N, tag, rtag, f.row, f.other, f.btype, f.maxb, f.bpart,
                                       f.bptype, stream),
                     "bond index");
    }

    void update_exclusion_index()
    {
        if (max_excl && exclude_bonds)
            hchk(launch_excl_index(N, tag, rtag, n_excl_tag, excl_tag, max_excl, n_excl, excl,
                                   stream),
                 "exclusion index");
    }

    void fill_nlist_args(NlistKernelArgs& k, bool gated)
    {
        memset(&k, 0, sizeof(k));
        k.N = N;
        k.box = box;
        k.dim[0] = dim[0];
        k.dim[1] = dim[1];
        k.dim[2] = dim[2];
        k.pos = pos;
        k.ncell = ncell;
        k.cmax = cmax;
        k.cell_size = cell_size;
        k.entries = entries;
        k.cell_adj = cell_adj;
        k.n_adj = n_adj;
        k.ntypes = ntypes;
        for (unsigned int q = 0; q < ntypes * ntypes; q++)
        {
            double rc = rcut_mat[q];
            double rl = rc > 0.0 ? rc + r_buff : 0.0;
            k.rlistsq[q] = rl * rl;
            k.enabled[q] = rc > 0.0 ? 1u : 0u;
        }
        bool ex = max_excl && exclude_bonds;
        k.n_excl = ex ? n_excl : nullptr;
        k.excl = ex ? excl : nullptr;
        k.max_excl = ex ? max_excl : 0;
        k.nmax = nmax;
        k.nlist = nlist;
        k.n_neigh = n_neigh;
        // local rows feed the cell-tiled kernels, which need no index rows
        k.lrow = (pair_kind() == 0 || pair_kind() == 3) ? lrow : nullptr;
        k.skip_index_rows = pair_kind() == 3 ? 1 : 0;
        k.offset_codes = pair_kind() == 2 ? 1 : 0;
        k.cond = &cond->nlist_need;
        k.shape = nl_shape;
        k.grid = nl_grid;
        k.gate = gated ? st : nullptr;
        k.counter = st;
        k.check_delay = check_delay;
        k.dist_check = dist_check;
        if (uses_rel())
        {
            // the cell ordering of this build wrote the relative positions the list stages
            k.srel = rel;
            k.sidx = cell_idx;
            k.swrd = relw;
            for (int d = 0; d < 3; d++)
                k.cwf[d] = (float)(box.L[d] / (double)dim[d]);
        }
    }

    bool rel_ok() const
    {
        for (int d = 0; d < 3; d++)
            if (dim[d] < 3 || dim[d] > 255)
                return false;
        return (size_t)ncell * cmax <= OFFSET_CODE_MAX_N;
    }
    // the requested kernel form, where it applies: the cell-tiled forms (0, 3) are built for LJ
    // only, so a Table force runs the relative gather (2); the relative forms need rel_ok()
    int pair_kind() const
    {
        // the learned site energy reads plain particle-index rows
        for (const auto& f : forces)
            if (f.kind == 3)
                return 1;
        int k = pair_kernel_kind;
        if (k == 0 || k == 3)
            for (const auto& f : forces)
                if (f.kind == 2)
                    k = 2;
        return (k == 2 || k == 3) && !rel_ok() ? 1 : k;
    }
    bool uses_rel() const { return pair_kind() == 2 || pair_kind() == 3; }
    // Pair list (tags, a < b) from the 16-bit local rows: entry l of a particle in home cell c
    // is candidate l of c's adjacency-ordered sequence (cells in cell_adj order, slots in order)
    long long pairs_from_local_rows(unsigned int* pairs, long long max_pairs)
    {
        std::vector<unsigned int> csize(ncell), adj((size_t)ncell * n_adj), cidx((size_t)ncell * cmax),
            nn(N), tg(N), pc(N);
        std::vector<unsigned short> lr((size_t)N * nmax);
        hchk(hipStreamSynchronize(stream), "sync");
        auto dl = [&](void* d, const void* sp, size_t b) {
            hchk(hipMemcpy(d, sp, b, hipMemcpyDeviceToHost), "download");
        };
        dl(csize.data(), cell_size, sizeof(unsigned int) * csize.size());
        dl(adj.data(), cell_adj, sizeof(unsigned int) * adj.size());
        dl(cidx.data(), cell_idx, sizeof(unsigned int) * cidx.size());
        dl(nn.data(), n_neigh, sizeof(unsigned int) * N);
        dl(tg.data(), tag, sizeof(unsigned int) * N);
        dl(pc.data(), pcell, sizeof(uint32_t) * N);
        dl(lr.data(), lrow, sizeof(unsigned short) * lr.size());
        long long out = 0;
        std::vector<unsigned int> seq;
        for (unsigned int i = 0; i < N; i++)
        {
            const uint32_t p = pc[i];
            const unsigned int c = (p & 255u) + dim[0] * (((p >> 8) & 255u) + dim[1] * ((p >> 16) & 255u));
            seq.clear();
            for (unsigned int q = 0; q < n_adj; q++)
            {
                const unsigned int nc = adj[(size_t)c * n_adj + q];
                const unsigned int sz = std::min(csize[nc], cmax);
                for (unsigned int k = 0; k < sz; k++)
                    seq.push_back(cidx[(size_t)nc * cmax + k]);
            }
            for (unsigned int k = 0; k < nn[i] && k < nmax; k++)
            {
                const unsigned int l = lr[(size_t)i * nmax + k];
                if (l >= seq.size())
                    throw HmError{HM_ERR_STATE, "local row entry outside the candidate sequence"};
                const unsigned int a = tg[i], b = tg[seq[l]];
                if (a < b)
                {
                    if (pairs && out < max_pairs)
                    {
                        pairs[2 * out] = a;
                        pairs[2 * out + 1] = b;
                    }
                    out++;
                }
            }
        }
        return out;
    }

    RelFrame rel_frame() const
    {
        return make_rel_frame(uses_rel() ? rel : nullptr, relw, pcell, slot_of, box, dim);
    }

    // ------------------------------------------------------------------ kernels of one step
    void launch_pair(ForceSlot& f, int with_virial)
    {
        const int kind = pair_kind();
        if (kind == 3)
        {
            PairTileArgs k;
            memset(&k, 0, sizeof(k));
            k.N = N;
            k.pos = pos;
            k.rel = rel;
            k.relw = relw;
            k.cell_idx = cell_idx;
            k.cell_size = cell_size;
            k.cell_adj = cell_adj;
            k.n_adj = n_adj;
            k.ncell = ncell;
            k.cmax = cmax;
            const RelFrame fr = rel_frame();
            for (int d = 0; d < 3; d++)
            {
                k.dim[d] = dim[d];
                k.wf[d] = fr.wf[d];
            }
            k.box = box;
            k.lrow = lrow;
            k.nmax = nmax;
            k.n_neigh = n_neigh;
            k.ntypes = ntypes;
            for (unsigned int q = 0; q < ntypes * ntypes; q++)
            {
                const double rc = f.rcut[q] > 0.0 ? f.rcut[q] : 0.0;
                k.rcutsq[q] = rc * rc;
                k.rcutsq_f[q] = (float)(rc * rc);
                k.ronsq_f[q] = (float)(f.ron[q] * f.ron[q]);
            }
            k.shift_mode = f.mode;
            k.compute_virial = with_virial;
            k.force = f.f;
            k.virial = f.v;
            hchk(launch_pair_tile_lj(k, f.p, pair_tile_cfg, stream), "pair force");
            return;
        }
        if (kind == 2)
        {
            PairRelArgs k;
            memset(&k, 0, sizeof(k));
            k.N = N;
            k.pos = pos;
            k.rel = rel;
            k.relw = relw;
            k.slot_of = slot_of;
            k.part_of = cell_idx;
            k.nslot = ncell * cmax;
            k.box = box;
            const RelFrame fr = rel_frame();
            for (int d = 0; d < 3; d++)
            {
                k.dim[d] = dim[d];
                k.wf[d] = fr.wf[d];
            }
            k.nlist = nlist;
            k.nmax = nmax;
            k.n_neigh = n_neigh;
            k.ntypes = ntypes;
            for (unsigned int q = 0; q < ntypes * ntypes; q++)
            {
                const double rc = f.rcut[q] > 0.0 ? f.rcut[q] : 0.0;
                k.rcutsq[q] = rc * rc;
                k.rcutsq_f[q] = (float)(rc * rc);
                k.ronsq_f[q] = (float)(f.ron[q] * f.ron[q]);
            }
            k.shift_mode = f.mode;
            k.compute_virial = with_virial;
            k.force = f.f;
            k.virial = f.v;
            if (f.kind == 2)
                hchk(launch_pair_rel_table(k, f.tp, pair_cfg, stream), "pair force");
            else
                hchk(launch_pair_rel_lj(k, f.p, pair_cfg, stream), "pair force");
            return;
        }
        if (kind == 1)
        {
            PairKernelArgs k;
            fill_pair_args(k, N, pos, hbox, nlist, nmax, n_neigh, ntypes, f.rcut, f.ron, f.mode,
                           with_virial, f.f, f.v);
            if (f.kind == 2)
                hchk(launch_pair_table(k, f.tp, pair_cfg, stream), "pair force");
            else
                hchk(launch_pair_lj(k, f.p, pair_cfg, stream), "pair force");
            return;
        }
        PairCellArgs c;
        memset(&c, 0, sizeof(c));
        c.N = N;
        c.pos = pos;
        c.box = box;
        c.dim[0] = dim[0];
        c.dim[1] = dim[1];
        c.dim[2] = dim[2];
        c.ncell = ncell;
        c.cmax = cmax;
        c.cell_size = cell_size;
        c.cell_idx = cell_idx;
        c.cell_adj = cell_adj;
        c.n_adj = n_adj;
        c.lrow = lrow;
        c.nmax = nmax;
        c.n_neigh = n_neigh;
        c.ntypes = ntypes;
        for (unsigned int q = 0; q < ntypes * ntypes; q++)
        {
            const double rc = f.rcut[q] > 0.0 ? f.rcut[q] : 0.0;
            c.rcutsq[q] = rc * rc;
            c.rcutsq_f[q] = (float)(rc * rc);
            c.ronsq_f[q] = (float)(f.ron[q] * f.ron[q]);
        }
        c.shift_mode = f.mode;
        c.compute_virial = with_virial;
        c.force = f.f;
        c.virial = f.v;
        hchk(launch_pair_cell_lj(c, f.p, pair_cell_cfg, stream), "pair force");
    }

    MlpArgs mlp_args(ForceSlot& f, int with_virial)
    {
        MlpArgs a;
        memset(&a, 0, sizeof(a));
        a.N = N;
        a.pos = pos;
        a.box = box;
        a.nlist = nlist;
        a.nmax = nmax;
        a.n_neigh = n_neigh;
        a.d = make_mlp_descr(f.mm);
        a.master = f.mparams;
        a.wb = f.mbf16;
        a.G = f.mG;
        a.A = f.mA;
        a.Ab = f.mAb;
        a.Esite = f.mE;
        a.force = f.f;
        a.virial = with_virial ? f.v : nullptr;
        a.target = f.mT;
        a.err = f.mErr;
        a.C = f.mC;
        a.loss_partial = f.mLoss;
        a.gpart = f.mGpart;
        a.loss_scale = (float)(2.0 / (3.0 * (double)N));
        a.ptab = f.mtab;
        a.ptab_rmin = (float)f.mtab_rmin;
        a.ptab_width = f.mtab ? f.mtab_w : 0;
        return a;
    }

    void launch_mlp(ForceSlot& f, int with_virial)
    {
        MlpArgs a = mlp_args(f, with_virial);
        a.A = nullptr;  // the force pass reads the bf16 copy; the FP32 dE/dG is for training
        if (pos32)
        {
            // step one of this step wrote the copy; elsewhere (set-up, thermo, after a sort)
            // it is refreshed first
            if (!in_step)
                hchk(launch_pos_f32(N, pos, pos32, stream), "fp32 positions");
            a.pos32 = pos32;
        }
        hchk(launch_mlp_descriptor(a, stream), "mlp descriptor");
        hchk(launch_mlp_site(a, stream), "mlp site energies");
        hchk(launch_mlp_force(a, false, stream), "mlp forces");
    }

    void launch_forces(int with_virial)
    {
        for (auto& f : forces)
        {
            if (f.kind == 3)
                launch_mlp(f, with_virial);
            else if (is_pair(f))
                launch_pair(f, with_virial);
            else
            {
                BondKernelArgs k;
                memset(&k, 0, sizeof(k));
                k.N = N;
                k.pos = pos;
                k.box = box;
                k.tag = tag;
                k.rtag = rtag;
                k.row = f.row;
                k.other_tag = f.other;
                k.btype = f.btype;
                k.nbond_types = f.nbt;
                k.compute_virial = with_virial;
                k.force = f.f;
                k.virial = f.v;
                k.flag = f.flag;
                k.st = st;
                k.bpart = f.bpart;
                k.bptype = f.bptype;
                k.max_bonds = f.maxb;
                if (f.kind == 4)
                    hchk(launch_bond_harmonic(k, f.hp, stream), "bond force");
                else
                    hchk(launch_bond_fene(k, f.fp, stream), "bond force");
            }
        }
    }

    unsigned int force_ptrs(const float4** fp) const
    {
        unsigned int n = 0;
        for (auto& f : forces)
            if (n < 4)
                fp[n++] = f.f;
        return n;
    }

    double ndof() const
    {
        // IntegratorTwoStep::getTranslationalDOF: one momentum-conserving method over all
        // particles removes D degrees of freedom (IntegratorTwoStep.cc:154-188)
        return method == 2 ? 3.0 * N : 3.0 * N - 3.0;
    }

    void launch_step()
    {
        // the Bussi factor needs the kinetic energy of every particle before step one: a
        // step-begin kernel of its own. Without it, step one begins the step and step two ends it
        const bool bussi = method == 1;
        if (bussi)
            hchk(launch_step_begin(st, ke_partial, n_partial, 1, ndof(), dt, tau, kT, seed,
                                   stream),
                 "step begin");
        StepOneArgs a1;
        memset(&a1, 0, sizeof(a1));
        a1.N = N;
        a1.pos = pos;
        a1.vel = vel;
        a1.accel = accel;
        a1.image = image;
        a1.box = box;
        a1.dt = dt;
        a1.alpha = 1.0;
        a1.st = st;
        a1.pos32 = pos32;
        a1.use_state_alpha = bussi ? 1 : 0;
        a1.begin = bussi ? nullptr : st;
        bool have_nl = rcut_max > 0.0;
        if (have_nl)
        {
            a1.last = last;
            a1.maxsq = r_buff < 1e-6 ? -1.0 : distance_check_maxsq(rcut_max, r_buff);
            a1.flag_state = st;
            a1.check_delay = check_delay;
            a1.dist_check = dist_check;
            a1.rel = rel_frame();
        }
        // the gather pair kernels read no cell list between builds: step one clears the
        // occupancy counters every step, which saves a launch; the cell-tiled kernels keep them
        const int kind = pair_kind();
        const bool fused_clear = have_nl && (kind == 1 || kind == 2);
        if (fused_clear)
        {
            a1.cell_size = cell_size;
            a1.ncell = ncell;
        }
        hchk(launch_step_one(a1, stream), "step one");
        if (have_nl)
        {
            if (!fused_clear)
                hchk(launch_cell_clear(ncell, cell_size, st, stream), "cell clear");
            if (N <= CELL_FUSED_MAX_N)
                hchk(launch_cell_bin_rank(N, dim, cmax, pos, cell_size, last, &cond->cell_need,
                                          ckey, chome, entries, cell_idx, st, stream, box,
                                          rel_frame()),
                     "cell build and order");
            else
            {
                hchk(launch_cell_build(N, pos, box, dim, cmax, cell_size, nullptr, last,
                                       &cond->cell_need, st, stream, ckey, chome),
                     "cell build");
                hchk(launch_cell_rank(N, cmax, pos, cell_size, ckey, chome, entries, cell_idx,
                                      st, stream, box, rel_frame()),
                     "cell order");
            }
            NlistKernelArgs k;
            fill_nlist_args(k, true);
            hchk(launch_nlist_build(k, stream), "nlist build");
        }
        {
            struct InStep
            {
                bool& f;
                explicit InStep(bool& x) : f(x) { f = true; }
                ~InStep() { f = false; }
            } guard(in_step);
            launch_forces(compute_virial);
        }
        StepTwoArgs a2;
        memset(&a2, 0, sizeof(a2));
        a2.N = N;
        a2.vel = vel;
        a2.accel = accel;
        a2.nforces = force_ptrs(a2.f);
        a2.dt = dt;
        a2.alpha = 1.0;
        if (method == 2)
        {
            a2.gamma_by_type = d_gamma;
            a2.pos = pos;
            a2.tag = tag;
            a2.kT = kT;
            a2.seed = seed;
        }
        a2.st = st;
        a2.ke_partial = ke_partial;
        a2.finish = bussi ? nullptr : st;
        hchk(launch_step_two(a2, stream, nullptr), "step two");
    }

    // ------------------------------------------------------------------ prep / forces
    void build_nlist_now()
    {
        // CellList::compute + NeighborList::compute loops with host-side regrow (the reference's
        // behaviour on the first build, NeighborList.cc:409-427)
        for (int attempt = 0; attempt < 64; attempt++)
        {
            hchk(hipMemsetAsync(cond, 0, sizeof(Conditions), stream), "clear conditions");
            hchk(hipMemsetAsync(cell_size, 0, sizeof(unsigned int) * ncell, stream), "clear cells");
            hchk(launch_cell_build(N, pos, box, dim, cmax, cell_size, nullptr, last,
                                   &cond->cell_need, nullptr, stream, ckey, chome),
                 "cell build");
            hchk(hipMemcpyAsync(h_cond, cond, sizeof(Conditions), hipMemcpyDeviceToHost, stream),
                 "read conditions");
            hchk(hipStreamSynchronize(stream), "sync");
            if (h_cond->nan_idx)
                throw HmError{HM_ERR_NAN, "particle position is NaN"};
            if (h_cond->out_idx)
                throw HmError{HM_ERR_OUTSIDE, "particle is no longer in the simulation box"};
            if (h_cond->cell_need > cmax)
            {
                grow_cmax(h_cond->cell_need);
                continue;
            }
            hchk(launch_cell_rank(N, cmax, pos, cell_size, ckey, chome, entries, cell_idx, nullptr,
                                  stream, box, rel_frame()),
                 "cell order");
            NlistKernelArgs k;
            fill_nlist_args(k, false);
            k.counter = nullptr;
            hchk(launch_nlist_build(k, stream), "nlist build");
            hchk(hipMemcpyAsync(h_cond, cond, sizeof(Conditions), hipMemcpyDeviceToHost, stream),
                 "read conditions");
            hchk(hipStreamSynchronize(stream), "sync");
            if (h_cond->nlist_need > nmax)
            {
                grow_nmax(h_cond->nlist_need);
                continue;
            }
            return;
        }
        throw HmError{HM_ERR_LAUNCH, "neighbour list did not converge"};
    }

    // Autotuner (the role of Autotuner.h): every candidate launch configuration of the pair
    // kernel in use is timed over 5 launches with HIP events on the engine stream, and the one
    // with the smallest median launch time is kept -- the reference's protocol of sampling each
    // parameter tuple 5 times and taking the median (Autotuner.h:503-664). Candidates: lanes per
    // particle {1, 2, 4, 8, 16} x block size {64, 128, 256, 512, 1024} for the gather kernels,
    // lanes per particle {4, 8, 16} for the cell-tiled kernel. The pair launches only rewrite
    // the force arrays with the same values, so tuning does not change the state.
    // Sweep of the list build's workgroup shape and persistent grid on the lists of the current
    // state (cells already built; every candidate writes the same rows), median of 3 launches.
    struct NlSample
    {
        int shape;
        unsigned int grid;
        float ms;
    };
    std::vector<NlSample> nl_log;
    void tune_nlist()
    {
        nl_log.clear();
        if (rcut_max <= 0.0 || !nlist)
            return;
        std::vector<int> shapes{0};
        if (cmax <= 64)
            shapes = {NL_SHAPE_TINY, NL_SHAPE_SMALL, NL_SHAPE_MID, NL_SHAPE_BIG};
        hipEvent_t ev[4];
        for (auto& e : ev)
            hchk(hipEventCreate(&e), "event");
        float best = 1e30f;
        int bs = nl_shape;
        unsigned int bg = nl_grid;
        for (int sh : shapes)
        {
            const bool one_wave = sh == NL_SHAPE_TINY || sh == NL_SHAPE_SMALL;
            std::vector<unsigned int> grids =
                one_wave ? std::vector<unsigned int>{2048u, 4096u, 8192u, 16384u}
                         : std::vector<unsigned int>{1024u, 2048u, 4096u, 8192u};
            // one-wavefront shapes also try one workgroup per cell (no persistent loop)
            if (one_wave && round_up(ncell, 8) > 16384u)
                grids.push_back(round_up(ncell, 8));
            for (unsigned int g : grids)
            {
                if (g > round_up(ncell, 8) && g != 1024u && g != 2048u)
                    continue;
                nl_shape = sh;
                nl_grid = g;
                NlistKernelArgs k;
                fill_nlist_args(k, false);
                k.counter = nullptr;
                hchk(launch_nlist_build(k, stream), "nlist tune");
                for (int r = 0; r < 3; r++)
                {
                    hchk(hipEventRecord(ev[r], stream), "event");
                    hchk(launch_nlist_build(k, stream), "nlist tune");
                }
                hchk(hipEventRecord(ev[3], stream), "event");
                hchk(hipEventSynchronize(ev[3]), "event");
                float t[3];
                for (int r = 0; r < 3; r++)
                    hchk(hipEventElapsedTime(&t[r], ev[r], ev[r + 1]), "event");
                std::sort(t, t + 3);
                nl_log.push_back({sh, g, t[1]});
                if (t[1] < best)
                {
                    best = t[1];
                    bs = sh;
                    bg = g;
                }
            }
        }
        for (auto& e : ev)
            (void)hipEventDestroy(e);
        nl_shape = bs;
        nl_grid = bg;
        clear_graphs();
    }

    // the whole sweep: the shape cache (same kind of system seen before in this process), else
    // the list build then the pair kernel, and the result stored in the cache
    void tune_all(bool use_cache)
    {
        const std::string key = tune_key();
        auto it = g_tune_cache().find(key);
        if (use_cache && it != g_tune_cache().end())
        {
            const TuneChoice& c = it->second;
            pair_launch_cfg() = c.pair;
            nl_shape = c.nl_shape;
            nl_grid = c.nl_grid;
            pair_tuned = true;
            tune_cached = true;
            tune_log.clear();
            nl_log.clear();
            clear_graphs();
            return;
        }
        tune_cached = false;
        tune_nlist();
        tune_pair();
        g_tune_cache()[key] = TuneChoice{pair_launch_cfg(), nl_shape, nl_grid};
    }

    PairLaunch& pair_launch_cfg()
    {
        const int kind = pair_kind();
        return kind == 0 ? pair_cell_cfg : (kind == 3 ? pair_tile_cfg : pair_cfg);
    }

    // bucket of systems that share launch shapes: pair kind, force kinds, types, exclusions,
    // log2 N in half steps, density, list range and cells per axis (coarse)
    std::string tune_key() const
    {
        const double vol = hbox.L[0] * hbox.L[1] * hbox.L[2];
        char buf[256];
        std::string forces_kinds;
        for (auto& f : forces)
            forces_kinds += char('0' + f.kind);
        std::snprintf(buf, sizeof(buf), "k%d f%s t%u x%d n%ld r%ld l%ld c%u", pair_kind(),
                      forces_kinds.c_str(), ntypes, (max_excl && exclude_bonds) ? 1 : 0,
                      std::lround(2.0 * std::log2((double)N)), std::lround(N / vol * 40.0),
                      std::lround((rcut_max + r_buff) * 10.0), std::min(dim[0], 64u) / 8);
        return buf;
    }

    void tune_pair()
    {
        ForceSlot* f = nullptr;
        for (auto& x : forces)
            if (x.kind == 0 || x.kind == 2)
            {
                f = &x;
                break;
            }
        pair_tuned = true;
        if (!f || N == 0)
            return;
        const int kind = pair_kind();
        std::vector<PairLaunch> cands;
        if (kind == 0)
            for (unsigned int t : {4u, 8u, 16u})
                cands.push_back(PairLaunch{t, 0});
        else if (kind == 3)
            for (unsigned int bt : {4u, 8u})
                for (unsigned int t : {4u, 8u, 16u})
                    for (unsigned int g : {1024u, 2048u, 4096u})
                        cands.push_back(PairLaunch{t, g, bt});
        else
            for (unsigned int bt : kind == 2 ? std::vector<unsigned int>{4u, 8u}
                                             : std::vector<unsigned int>{8u})
                for (unsigned int t : {1u, 2u, 4u, 8u, 16u})
                    for (unsigned int b : {64u, 128u, 256u, 512u, 1024u})
                        cands.push_back(PairLaunch{t, b, bt});
        PairLaunch& cfg = kind == 0 ? pair_cell_cfg : (kind == 3 ? pair_tile_cfg : pair_cfg);
        hipEvent_t ev[6];
        for (auto& e : ev)
            hchk(hipEventCreate(&e), "event");
        float best = 1e30f;
        PairLaunch best_cfg = cfg;
        tune_log.clear();
        for (const PairLaunch& c : cands)
        {
            cfg = c;
            launch_pair(*f, compute_virial);  // warm-up launch
            for (int r = 0; r < 5; r++)
            {
                hchk(hipEventRecord(ev[r], stream), "event");
                launch_pair(*f, compute_virial);
            }
            hchk(hipEventRecord(ev[5], stream), "event");
            hchk(hipEventSynchronize(ev[5]), "event");
            float t[5];
            for (int r = 0; r < 5; r++)
                hchk(hipEventElapsedTime(&t[r], ev[r], r < 4 ? ev[r + 1] : ev[5]), "event");
            std::sort(t, t + 5);
            tune_log.push_back({c.tpp, c.block, c.batch, t[2]});
            if (t[2] < best)
            {
                best = t[2];
                best_cfg = c;
            }
        }
        for (auto& e : ev)
            (void)hipEventDestroy(e);
        cfg = best_cfg;
        clear_graphs();
    }
    struct TuneSample
    {
        unsigned int tpp, size, batch;
        float ms;
    };
    std::vector<TuneSample> tune_log;

    void prep()
    {
        if (method < 0)
            throw HmError{HM_ERR_STATE, "no integration method set"};
        configure_nlist();
        if (!st)
        {
            st = dalloc<StepState>(1);
            cond = dalloc<Conditions>(1);
            h_st = halloc<StepState>(1);
            h_cond = halloc<Conditions>(1);
        }
        StepState s0{};
        s0.next = timestep;
        s0.timestep = timestep;
        s0.rebuild_at = 0;
        s0.alpha = 1.0;
        // the build below is the forced build of the first step (NeighborList.cc:1358-1368)
        s0.last_update = timestep;
        s0.forced_at = timestep;
        s0.shortest = ~0ull;
        s0.n_builds = rcut_max > 0.0 ? 1 : 0;
        hchk(hipMemcpyAsync(st, &s0, sizeof(StepState), hipMemcpyHostToDevice, stream), "state");
        update_exclusion_index();
        update_bond_index();
        if (rcut_max > 0.0)
            build_nlist_now();
        if (autotune && !pair_tuned)
            tune_all(true);
        launch_forces(compute_virial);
        const float4* fp[4];
        unsigned int nf = force_ptrs(fp);
        hchk(launch_accel_init(N, vel, accel, fp, nf, stream), "accelerations");
        hchk(launch_ke_partial(N, vel, ke_partial, &n_partial, stream), "kinetic energy");
        hchk(hipStreamSynchronize(stream), "prep sync");
        check_errors_after(false);
        prepared = true;
    }

    void check_errors_after(bool chunk)
    {
        hchk(hipMemcpyAsync(h_st, st, sizeof(StepState), hipMemcpyDeviceToHost, stream), "state");
        hchk(hipMemcpyAsync(h_cond, cond, sizeof(Conditions), hipMemcpyDeviceToHost, stream),
             "conditions");
        hchk(hipStreamSynchronize(stream), "sync");
        (void)chunk;
    }

    // ------------------------------------------------------------------ sort
    void do_sort()
    {
        // SFCPackTuner::update (hoomd/SFCPackTuner.cc:81-285): order, permute, rtag, force update
        int end_bit = 32 + 3 * ilog2(sort_grid);
        if (!keys_a)
        {
            keys_a = dalloc<unsigned long long>(N);
            keys_b = dalloc<unsigned long long>(N);
            order = dalloc<unsigned int>(N);
            hchk(sort_keys64(nullptr, nullptr, N, end_bit, nullptr, &sort_tmp_bytes, stream),
                 "sort size");
            hchk(hipMalloc(&sort_tmp, sort_tmp_bytes ? sort_tmp_bytes : 16), "sort scratch");
        }
        if (!scratch)
            hchk(hipMalloc(&scratch, (size_t)N * 32), "sort scratch");
        hchk(launch_sfc_keys(N, pos, box, sort_grid, keys_a, stream), "sfc keys");
        size_t tb = sort_tmp_bytes;
        hchk(sort_keys64(keys_a, keys_b, N, end_bit, sort_tmp, &tb, stream), "sfc sort");
        hchk(launch_keys_to_order(N, keys_b, order, stream), "sfc order");
        auto permute = [&](void* arr, unsigned int bytes) {
            hchk(gather_bytes(arr, scratch, order, N, bytes, stream), "permute");
            hchk(hipMemcpyAsync(arr, scratch, (size_t)N * bytes, hipMemcpyDeviceToDevice, stream),
                 "permute copy");
        };
        permute(pos, 32);
        permute(vel, 32);
        permute(accel, 24);
        permute(image, 12);
        permute(tag, 4);
        permute(last, 24);
        hchk(launch_rtag(N, tag, rtag, stream), "rtag");
        update_exclusion_index();
        update_bond_index();
        hchk(launch_set_rebuild(st, stream), "force rebuild");
        n_sorts++;
        last_sort_step = timestep;
    }

    // ------------------------------------------------------------------ snapshot / rollback
    void snapshot()
    {
        if (!s_pos)
        {
            s_pos = dalloc<double4>(N);
            s_vel = dalloc<double4>(N);
            s_accel = dalloc<d3>(N);
            s_last = dalloc<d3>(N);
            s_image = dalloc<i3>(N);
            s_tag = dalloc<unsigned int>(N);
            s_ke = dalloc<double>(n_partial);
        }
        auto cp = [&](void* d, const void* s, size_t b) {
            hchk(hipMemcpyAsync(d, s, b, hipMemcpyDeviceToDevice, stream), "snapshot");
        };
        cp(s_pos, pos, (size_t)N * 32);
        cp(s_vel, vel, (size_t)N * 32);
        cp(s_accel, accel, (size_t)N * 24);
        cp(s_last, last, (size_t)N * 24);
        cp(s_image, image, (size_t)N * 12);
        cp(s_tag, tag, (size_t)N * 4);
        cp(s_ke, ke_partial, (size_t)n_partial * 8);
        hchk(hipMemcpyAsync(&s_state, st, sizeof(StepState), hipMemcpyDeviceToHost, stream),
             "snapshot state");
    }

    void restore()
    {
        auto cp = [&](void* d, const void* s, size_t b) {
            hchk(hipMemcpyAsync(d, s, b, hipMemcpyDeviceToDevice, stream), "restore");
        };
        cp(pos, s_pos, (size_t)N * 32);
        cp(vel, s_vel, (size_t)N * 32);
        cp(accel, s_accel, (size_t)N * 24);
        cp(last, s_last, (size_t)N * 24);
        cp(image, s_image, (size_t)N * 12);
        cp(tag, s_tag, (size_t)N * 4);
        cp(ke_partial, s_ke, (size_t)n_partial * 8);
        StepState s = s_state;
        s.error = 0;
        hchk(hipMemcpyAsync(st, &s, sizeof(StepState), hipMemcpyHostToDevice, stream), "restore");
        hchk(launch_rtag(N, tag, rtag, stream), "rtag");
        update_exclusion_index();
        update_bond_index();
        hchk(hipStreamSynchronize(stream), "restore sync");
    }

    // ------------------------------------------------------------------ run
    void launch_chunk(unsigned int K)
    {
        if (!use_graph)
        {
            for (unsigned int k = 0; k < K; k++)
                launch_step();
            return;
        }
        auto it = graphs.find(K);
        if (it == graphs.end())
            it = graphs.emplace(K, capture_steps(K)).first;
        hchk(hipGraphLaunch(it->second, stream), "graph launch");
    }

    // K steps captured from the stream into an instantiated graph (nothing is launched)
    hipGraphExec_t capture_steps(unsigned int K)
    {
        hipGraph_t g;
        hchk(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "begin capture");
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
        hchk(hipStreamEndCapture(stream, &g), "end capture");
        hipGraphExec_t ex;
        hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        hchk(e, "graph instantiate");
        return ex;
    }

    // Capture and instantiate the graphs of every chunk length that run(nsteps) from the
    // current step would launch, without launching anything, so that a timed run that follows
    // spends no time in capture or instantiation. Returns the number of graphs added.
    int prime(uint64_t nsteps)
    {
        if (!prepared)
            prep();
        if (!use_graph)
            return 0;
        int added = 0;
        uint64_t t = timestep, end = timestep + nsteps;
        while (t < end)
        {
            uint64_t K = std::min<uint64_t>(chunk ? chunk : 1, end - t);
            if (sort_period)
            {
                uint64_t next_sort = (t / sort_period + 1) * sort_period;
                K = std::min<uint64_t>(K, next_sort - t);
            }
            const uint64_t ran = steps_run + (t - timestep);
            if (autotune && !retuned && !tune_cached && ran < retune_after)
                K = std::min<uint64_t>(K, retune_after - ran);
            if (graphs.find((unsigned int)K) == graphs.end())
            {
                graphs.emplace((unsigned int)K, capture_steps((unsigned int)K));
                added++;
            }
            t += K;
        }
        return added;
    }

    int run(uint64_t nsteps)
    {
        // build statistics count from the start of each run (System::run -> resetStats,
        // System.cc:90; NeighborList::resetStats, NeighborList.cc:1420-1426)
        run_builds_base = 0;
        if (prepared && st)
        {
            StepState h;
            hchk(hipStreamSynchronize(stream), "sync");
            hchk(hipMemcpy(&h, st, sizeof(StepState), hipMemcpyDeviceToHost), "download");
            run_builds_base = h.n_builds;
            const unsigned long long reset[2] = {~0ull, 0ull};
            hchk(hipMemcpy(&st->shortest, reset, sizeof(reset), hipMemcpyHostToDevice), "reset");
        }
        if (!prepared)
            prep();
        uint64_t end = timestep + nsteps;
        int retries = 0;
        while (timestep < end)
        {
            if (autotune && !retuned && steps_run >= retune_after && !tune_cached)
            {
                // second sweep on the state the run has reached (outside any captured chunk)
                retuned = true;
                tune_all(false);
                // the sweep's list builds reused the last cell binning: rebuild both before the
                // forces of the next step
                hchk(launch_set_rebuild(st, stream), "force rebuild");
            }
            if (sort_period && timestep % sort_period == 0 && last_sort_step != timestep)
                do_sort();
            uint64_t K = std::min<uint64_t>(chunk ? chunk : 1, end - timestep);
            if (sort_period)
            {
                uint64_t next_sort = (timestep / sort_period + 1) * sort_period;
                K = std::min<uint64_t>(K, next_sort - timestep);
            }
            if (autotune && !retuned && !tune_cached && steps_run < retune_after)
                K = std::min<uint64_t>(K, retune_after - steps_run);
            hchk(hipMemsetAsync(cond, 0, sizeof(Conditions), stream), "clear conditions");
            snapshot();
            launch_chunk((unsigned int)K);
            check_errors_after(true);
            const Conditions c = *h_cond;
            const StepState s = *h_st;
            if (c.cell_need > cmax || c.nlist_need > nmax)
            {
                restore();
                if (c.cell_need > cmax)
                    grow_cmax(c.cell_need);
                if (c.nlist_need > nmax)
                    grow_nmax(c.nlist_need);
                hchk(launch_set_rebuild(st, stream), "force rebuild");
                rollbacks++;
                if (++retries > 32)
                    throw HmError{HM_ERR_LAUNCH, "neighbour list kept overflowing"};
                continue;
            }
            if (c.nan_idx || c.out_idx || (s.error & (ERR_BOND | ERR_ZERO_KE)))
            {
                restore();
                if (c.nan_idx)
                    throw HmError{HM_ERR_NAN, "particle position is NaN"};
                if (c.out_idx)
                    throw HmError{HM_ERR_OUTSIDE, "particle is no longer in the simulation box"};
                if (s.error & ERR_BOND)
                    throw HmError{HM_ERR_BOND, "bond.fene: bond out of bounds"};
                throw HmError{HM_ERR_THERMO, "Bussi thermostat requires non-zero initial momenta"};
            }
            retries = 0;
            timestep += K;
            steps_run += K;
        }
        return HM_OK;
    }
};

// ====================================================================== C ABI
namespace
{
int guard_fail(const HmError& e)
{
    return fail(e.code, e.what.c_str());
}

template<class F> int guarded(F&& f)
{
    try
    {
        return f();
    }
    catch (const HmError& e)
    {
        return guard_fail(e);
    }
    catch (const std::bad_alloc&)
    {
        return fail(HM_ERR_ALLOC, "host allocation failed");
    }
    catch (const std::exception& e)
    {
        return fail(HM_ERR_LAUNCH, e.what());
    }
}
} // namespace

extern "C"
{

hm_system* hm_sys_create(unsigned int N, hm_box box, unsigned int ntypes, const double* pos3,
                         const double* vel3, const unsigned int* type_id, const double* mass,
                         const int* image3, uint16_t seed)
{
    if (N == 0 || !pos3 || ntypes == 0 || ntypes > HM_MAX_TYPES)
    {
        fail(HM_ERR_ARGS, "hm_sys_create: invalid arguments");
        return nullptr;
    }
    hm_system* s = new (std::nothrow) hm_system();
    if (!s)
        return nullptr;
    int r = guarded([&]() {
        s->N = N;
        s->ntypes = ntypes;
        s->hbox = box;
        s->box = make_devbox(box);
        s->seed = seed;
        hchk(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking), "stream");
        std::vector<double4> p(N), v(N);
        std::vector<i3> im(N);
        std::vector<unsigned int> tg(N);
        for (unsigned int i = 0; i < N; i++)
        {
            unsigned int t = type_id ? type_id[i] : 0u;
            if (t >= ntypes)
                throw HmError{HM_ERR_ARGS, "type id out of range"};
            p[i] = make_double4(pos3[3 * i], pos3[3 * i + 1], pos3[3 * i + 2], (double)t);
            v[i] = make_double4(vel3 ? vel3[3 * i] : 0.0, vel3 ? vel3[3 * i + 1] : 0.0,
                                vel3 ? vel3[3 * i + 2] : 0.0, mass ? mass[i] : 1.0);
            im[i] = image3 ? i3{image3[3 * i], image3[3 * i + 1], image3[3 * i + 2]} : i3{0, 0, 0};
            tg[i] = i;
        }
        s->pos = dalloc<double4>(N);
        s->vel = dalloc<double4>(N);
        s->accel = dalloc<d3>(N);
        s->image = dalloc<i3>(N);
        s->tag = dalloc<unsigned int>(N);
        s->rtag = dalloc<unsigned int>(N);
        s->last = dalloc<d3>(N);
        s->ke_partial = dalloc<double>((N + 255) / 256);
        s->thermo_partial = dalloc<double>(THERMO_SUMS * THERMO_BLOCKS);
        s->thermo_out = dalloc<double>(THERMO_SUMS);
        s->h_thermo = halloc<double>(THERMO_SUMS);
        s->d_gamma = dalloc<double>(HM_MAX_TYPES);
        hchk(hipMemcpy(s->pos, p.data(), sizeof(double4) * N, hipMemcpyHostToDevice), "upload");
        hchk(hipMemcpy(s->vel, v.data(), sizeof(double4) * N, hipMemcpyHostToDevice), "upload");
        hchk(hipMemcpy(s->image, im.data(), sizeof(i3) * N, hipMemcpyHostToDevice), "upload");
        hchk(hipMemcpy(s->tag, tg.data(), sizeof(unsigned int) * N, hipMemcpyHostToDevice),
             "upload");
        hchk(hipMemcpy(s->rtag, tg.data(), sizeof(unsigned int) * N, hipMemcpyHostToDevice),
             "upload");
        hchk(hipMemset(s->accel, 0, sizeof(d3) * N), "clear");
        hchk(hipMemset(s->last, 0, sizeof(d3) * N), "clear");
        return HM_OK;
    });
    if (r != HM_OK)
    {
        delete s;
        return nullptr;
    }
    return s;
}

void hm_sys_destroy(hm_system* s)
{
    if (s)
    {
        (void)hipStreamSynchronize(s->stream);
        delete s;
    }
}

int hm_sys_set_nlist(hm_system* s, double r_buff, int exclude_bonds)
{
    if (!s || r_buff < 0.0)
        return fail(HM_ERR_ARGS, "hm_sys_set_nlist: invalid arguments");
    s->r_buff = r_buff;
    s->exclude_bonds = exclude_bonds;
    s->prepared = false;
    s->nmax = 0;
    s->clear_graphs();
    return HM_OK;
}

int hm_sys_add_pair_lj(hm_system* s, const double* r_cut, const double* r_on,
                       const hm_lj_params* params, int mode)
{
    if (!s || !r_cut || !params || mode < 0 || mode > 2 || s->forces.size() >= 4)
        return fail(HM_ERR_ARGS, "hm_sys_add_pair_lj: invalid arguments");
    return guarded([&]() {
        ForceSlot f;
        f.kind = 0;
        unsigned int nt = s->ntypes;
        for (unsigned int q = 0; q < nt * nt; q++)
        {
            f.rcut[q] = r_cut[q];
            f.ron[q] = r_on ? r_on[q] : 0.0;
            f.p[q] = params[q];
        }
        f.mode = mode;
        f.f = dalloc<float4>(s->N);
        f.v = dalloc<float>((size_t)6 * s->N);
        hchk(hipMemset(f.f, 0, sizeof(float4) * s->N), "clear");
        hchk(hipMemset(f.v, 0, sizeof(float) * 6 * s->N), "clear");
        s->forces.push_back(f);
        s->prepared = false;
        s->nmax = 0;
        s->clear_graphs();
        return (int)s->forces.size() - 1;
    });
}

int hm_sys_add_pair_table(hm_system* s, const double* r_cut, const double* r_min,
                          const unsigned int* width, const double* U, const double* F)
{
    if (!s || !r_cut || !r_min || !width || s->forces.size() >= 4)
        return fail(HM_ERR_ARGS, "hm_sys_add_pair_table: invalid arguments");
    const unsigned int nt = s->ntypes;
    size_t total = 0;
    for (unsigned int q = 0; q < nt * nt; q++)
    {
        if (width[q] && r_cut[q] > 0.0 && !(r_min[q] < r_cut[q]))
            return fail(HM_ERR_ARGS, "hm_sys_add_pair_table: r_min must be below r_cut");
        total += width[q];
    }
    if (total && (!U || !F))
        return fail(HM_ERR_ARGS, "hm_sys_add_pair_table: U and F are required");
    return guarded([&]() {
        ForceSlot f;
        f.kind = 2;
        std::vector<float> h(2 * total + 1, 0.f);
        for (size_t k = 0; k < total; k++)
        {
            h[k] = (float)U[k];
            h[total + k] = (float)F[k];
        }
        f.tab = dalloc<float>(h.size());
        hchk(hipMemcpy(f.tab, h.data(), sizeof(float) * h.size(), hipMemcpyHostToDevice),
             "upload");
        size_t off = 0;
        for (unsigned int q = 0; q < nt * nt; q++)
        {
            f.rcut[q] = r_cut[q];
            f.ron[q] = 0.0;
            f.tp[q].r_min = r_min[q];
            f.tp[q].width = width[q];
            f.tp[q].U = f.tab + off;
            f.tp[q].F = f.tab + total + off;
            off += width[q];
        }
        f.mode = 0;
        f.f = dalloc<float4>(s->N);
        f.v = dalloc<float>((size_t)6 * s->N);
        hchk(hipMemset(f.f, 0, sizeof(float4) * s->N), "clear");
        hchk(hipMemset(f.v, 0, sizeof(float) * 6 * s->N), "clear");
        s->forces.push_back(f);
        s->prepared = false;
        s->nmax = 0;
        s->clear_graphs();
        return (int)s->forces.size() - 1;
    });
}

namespace
{
// the CSR-by-tag bond table, the index-order partner table and the bond exclusions of a bond
// force (FENE: kind 1, harmonic: kind 4)
int add_bond_force(hm_system* s, unsigned int nb, const unsigned int* bond_tags,
                   const unsigned int* bond_typeid, unsigned int nbond_types, int kind,
                   const hm_fene_params* fene, const hm_harmonic_params* harm)
{
    return guarded([&]() {
        unsigned int N = s->N;
        std::vector<unsigned int> deg(N + 1, 0);
        for (unsigned int b = 0; b < nb; b++)
        {
            unsigned int a = bond_tags[2 * b], c = bond_tags[2 * b + 1];
            if (a >= N || c >= N || a == c)
                throw HmError{HM_ERR_ARGS, "bond tag out of range"};
            if (bond_typeid && bond_typeid[b] >= nbond_types)
                throw HmError{HM_ERR_ARGS, "bond type out of range"};
            deg[a]++;
            deg[c]++;
        }
        std::vector<unsigned int> row(N + 1, 0);
        for (unsigned int i = 0; i < N; i++)
            row[i + 1] = row[i] + deg[i];
        std::vector<unsigned int> other(row[N] ? row[N] : 1), btype(row[N] ? row[N] : 1);
        std::vector<unsigned int> fill(row.begin(), row.end() - 1);
        for (unsigned int b = 0; b < nb; b++)
        {
            unsigned int a = bond_tags[2 * b], c = bond_tags[2 * b + 1];
            unsigned int t = bond_typeid ? bond_typeid[b] : 0u;
            other[fill[a]] = c;
            btype[fill[a]++] = t;
            other[fill[c]] = a;
            btype[fill[c]++] = t;
        }
        ForceSlot f;
        f.kind = kind;
        f.nbt = nbond_types;
        for (unsigned int q = 0; q < nbond_types; q++)
        {
            if (fene)
                f.fp[q] = fene[q];
            if (harm)
                f.hp[q] = harm[q];
        }
        f.row = dalloc<unsigned int>(N + 1);
        f.other = dalloc<unsigned int>(other.size());
        f.btype = dalloc<unsigned int>(btype.size());
        f.flag = dalloc<unsigned int>(1);
        f.f = dalloc<float4>(N);
        f.v = dalloc<float>((size_t)6 * N);
        hchk(hipMemcpy(f.row, row.data(), sizeof(unsigned int) * (N + 1), hipMemcpyHostToDevice),
             "upload");
        hchk(hipMemcpy(f.other, other.data(), sizeof(unsigned int) * other.size(),
                       hipMemcpyHostToDevice),
             "upload");
        hchk(hipMemcpy(f.btype, btype.data(), sizeof(unsigned int) * btype.size(),
                       hipMemcpyHostToDevice),
             "upload");
        hchk(hipMemset(f.flag, 0, sizeof(unsigned int)), "clear");
        hchk(hipMemset(f.f, 0, sizeof(float4) * N), "clear");
        hchk(hipMemset(f.v, 0, sizeof(float) * 6 * N), "clear");
        // bond exclusions by tag (NeighborList::addExclusionsFromBonds, NeighborList.cc:813-846)
        unsigned int mx = 0;
        for (unsigned int i = 0; i < N; i++)
            mx = std::max(mx, deg[i]);
        if (mx > 4)
            throw HmError{HM_ERR_ARGS, "at most 4 bonds per particle are supported"};
        if (mx > 0)
        {
            f.maxb = mx;
            f.bpart = dalloc<unsigned int>((size_t)N * mx);
            if (nbond_types > 1)
                f.bptype = dalloc<unsigned int>((size_t)N * mx);
        }
        if (mx > 0)
        {
            std::vector<unsigned int> ne(N, 0), ex((size_t)N * mx, 0xffffffffu);
            for (unsigned int i = 0; i < N; i++)
                for (unsigned int b = row[i]; b < row[i + 1]; b++)
                {
                    unsigned int o = other[b];
                    bool dup = false;
                    for (unsigned int q = 0; q < ne[i]; q++)
                        dup = dup || ex[(size_t)i * mx + q] == o;
                    if (!dup)
                        ex[(size_t)i * mx + ne[i]++] = o;
                }
            dfree(s->n_excl_tag);
            dfree(s->excl_tag);
            dfree(s->n_excl);
            dfree(s->excl);
            s->max_excl = mx;
            s->n_excl_tag = dalloc<unsigned int>(N);
            s->excl_tag = dalloc<unsigned int>((size_t)N * mx);
            s->n_excl = dalloc<unsigned int>(N);
            s->excl = dalloc<unsigned int>((size_t)N * mx);
            hchk(hipMemcpy(s->n_excl_tag, ne.data(), sizeof(unsigned int) * N,
                           hipMemcpyHostToDevice),
                 "upload");
            hchk(hipMemcpy(s->excl_tag, ex.data(), sizeof(unsigned int) * ex.size(),
                           hipMemcpyHostToDevice),
                 "upload");
        }
        s->forces.push_back(f);
        s->prepared = false;
        s->clear_graphs();
        return (int)s->forces.size() - 1;
    });
}
} // namespace

int hm_sys_update_pair_table(hm_system* s, int force_id, const double* U, const double* F)
{
    if (!s || force_id < 0 || force_id >= (int)s->forces.size() || s->forces[force_id].kind != 2
        || !U || !F)
        return fail(HM_ERR_ARGS, "hm_sys_update_pair_table: invalid arguments");
    return guarded([&]() {
        ForceSlot& f = s->forces[force_id];
        size_t total = 0;
        for (unsigned int q = 0; q < s->ntypes * s->ntypes; q++)
            total += f.tp[q].width;
        std::vector<float> h(2 * total + 1, 0.f);
        for (size_t k = 0; k < total; k++)
        {
            h[k] = (float)U[k];
            h[total + k] = (float)F[k];
        }
        // the captured graphs read the table through the same pointers: new values take
        // effect at the next step, no re-capture needed
        hchk(hipStreamSynchronize(s->stream), "sync");
        hchk(hipMemcpyAsync(f.tab, h.data(), sizeof(float) * h.size(), hipMemcpyHostToDevice,
                            s->stream),
             "upload");
        hchk(hipStreamSynchronize(s->stream), "sync");
        return HM_OK;
    });
}

int hm_sys_add_bond_fene(hm_system* s, unsigned int nb, const unsigned int* bond_tags,
                         const unsigned int* bond_typeid, unsigned int nbond_types,
                         const hm_fene_params* params)
{
    if (!s || (nb && !bond_tags) || !params || nbond_types == 0 || nbond_types > HM_MAX_TYPES
        || s->forces.size() >= 4)
        return fail(HM_ERR_ARGS, "hm_sys_add_bond_fene: invalid arguments");
    return add_bond_force(s, nb, bond_tags, bond_typeid, nbond_types, 1, params, nullptr);
}

int hm_sys_add_bond_harmonic(hm_system* s, unsigned int nb, const unsigned int* bond_tags,
                             const unsigned int* bond_typeid, unsigned int nbond_types,
                             const hm_harmonic_params* params)
{
    if (!s || (nb && !bond_tags) || !params || nbond_types == 0 || nbond_types > HM_MAX_TYPES
        || s->forces.size() >= 4)
        return fail(HM_ERR_ARGS, "hm_sys_add_bond_harmonic: invalid arguments");
    return add_bond_force(s, nb, bond_tags, bond_typeid, nbond_types, 4, nullptr, params);
}

int hm_sys_set_method(hm_system* s, int method, double dt, double kT, const double* param)
{
    if (!s || method < 0 || method > 2 || !(dt >= 0.0))
        return fail(HM_ERR_ARGS, "hm_sys_set_method: invalid arguments");
    return guarded([&]() {
        s->method = method;
        s->dt = dt;
        s->kT = kT;
        s->tau = 0.0;
        if (method == 1 && param)
            s->tau = param[0];
        if (method == 2)
        {
            for (unsigned int q = 0; q < s->ntypes; q++)
                s->gamma[q] = param ? param[q] : 1.0;
            hchk(hipMemcpy(s->d_gamma, s->gamma, sizeof(double) * HM_MAX_TYPES,
                           hipMemcpyHostToDevice),
                 "upload");
        }
        s->clear_graphs();
        return HM_OK;
    });
}

int hm_sys_set_sorter(hm_system* s, unsigned int period, unsigned int grid)
{
    if (!s || (period && (grid == 0 || grid > 1024 || (grid & (grid - 1)))))
        return fail(HM_ERR_ARGS, "hm_sys_set_sorter: invalid arguments");
    s->sort_period = period;
    s->sort_grid = grid;
    return HM_OK;
}

int hm_sys_set_compute_virial(hm_system* s, int on)
{
    if (!s)
        return fail(HM_ERR_ARGS, "hm_sys_set_compute_virial: NULL system");
    s->compute_virial = on ? 1 : 0;
    s->clear_graphs();
    return HM_OK;
}

int hm_sys_set_pair_launch(hm_system* s, unsigned int tpp, unsigned int block_size)
{
    if (!s || (tpp && tpp != 1 && tpp != 2 && tpp != 4 && tpp != 8 && tpp != 16)
        || (block_size && (block_size % 64 || block_size > 1024)))
        return fail(HM_ERR_ARGS, "hm_sys_set_pair_launch: invalid arguments");
    if (tpp)
        s->pair_cfg.tpp = tpp;
    if (block_size)
        s->pair_cfg.block = block_size;
    if (tpp || block_size)
        s->autotune = false;
    s->clear_graphs();
    return HM_OK;
}

int hm_sys_set_pair_kernel(hm_system* s, int kind, unsigned int tpp, unsigned int size)
{
    if (!s || kind < 0 || kind > 3)
        return fail(HM_ERR_ARGS, "hm_sys_set_pair_kernel: invalid arguments");
    if ((kind == 0 || kind == 3) && tpp && tpp != 4 && tpp != 8 && tpp != 16)
        return fail(HM_ERR_ARGS, "hm_sys_set_pair_kernel: cell kernel tpp must be 4, 8 or 16");
    s->pair_kernel_kind = kind;
    // the relative positions of kind 2 are written by the cell-list build: build again
    s->prepared = false;
    // tpp = size = 0: the autotuner picks the launch configuration at the next run
    s->autotune = tpp == 0 && size == 0;
    s->pair_tuned = false;
    if (kind == 0)
    {
        if (tpp)
            s->pair_cell_cfg.tpp = tpp;
        s->pair_cell_cfg.block = size;
        s->clear_graphs();
        return HM_OK;
    }
    if (kind == 3)
    {
        if (tpp)
            s->pair_tile_cfg.tpp = tpp;
        if (size)
            s->pair_tile_cfg.block = size;
        s->clear_graphs();
        return HM_OK;
    }
    return hm_sys_set_pair_launch(s, tpp, size);
}

int hm_sys_set_autotune(hm_system* s, int on)
{
    if (!s)
        return fail(HM_ERR_ARGS, "hm_sys_set_autotune: NULL system");
    s->autotune = on != 0;
    s->pair_tuned = false;
    if (s->autotune)
        s->prepared = false;
    return HM_OK;
}

int hm_sys_pair_config(const hm_system* s, unsigned int out[5])
{
    if (!s || !out)
        return fail(HM_ERR_ARGS, "hm_sys_pair_config: invalid arguments");
    const int kind = s->pair_kind();
    const PairLaunch& c = kind == 0 ? s->pair_cell_cfg : (kind == 3 ? s->pair_tile_cfg : s->pair_cfg);
    out[0] = (unsigned int)kind;
    out[1] = c.tpp;
    out[2] = c.block;
    out[3] = s->pair_tuned ? 1u : 0u;
    out[4] = c.batch;
    return HM_OK;
}

int hm_sys_tune_log(const hm_system* s, float* out, unsigned int max_rows)
{
    if (!s || (!out && max_rows))
        return fail(HM_ERR_ARGS, "hm_sys_tune_log: invalid arguments");
    unsigned int n = (unsigned int)std::min<size_t>(max_rows, s->tune_log.size());
    for (unsigned int q = 0; q < n; q++)
    {
        out[4 * q + 0] = (float)s->tune_log[q].tpp;
        out[4 * q + 1] = (float)s->tune_log[q].size;
        out[4 * q + 2] = (float)s->tune_log[q].batch;
        out[4 * q + 3] = s->tune_log[q].ms;
    }
    return (int)s->tune_log.size();
}

void hm_tune_cache_clear(void)
{
    g_tune_cache().clear();
}

int hm_sys_nlist_launch(const hm_system* s, int out[3], float* log3, unsigned int max_rows)
{
    if (!s || !out || (!log3 && max_rows))
        return fail(HM_ERR_ARGS, "hm_sys_nlist_launch: invalid arguments");
    out[0] = s->nl_shape;
    out[1] = (int)s->nl_grid;
    out[2] = s->tune_cached ? 1 : 0;
    const unsigned int n = (unsigned int)std::min<size_t>(max_rows, s->nl_log.size());
    for (unsigned int q = 0; q < n; q++)
    {
        log3[3 * q + 0] = (float)s->nl_log[q].shape;
        log3[3 * q + 1] = (float)s->nl_log[q].grid;
        log3[3 * q + 2] = s->nl_log[q].ms;
    }
    return (int)s->nl_log.size();
}

int hm_sys_set_chunk(hm_system* s, unsigned int steps, int use_graph)
{
    if (!s)
        return fail(HM_ERR_ARGS, "hm_sys_set_chunk: NULL system");
    s->chunk = steps ? steps : 1;
    s->use_graph = use_graph != 0;
    s->clear_graphs();
    return HM_OK;
}

int hm_sys_run(hm_system* s, uint64_t nsteps)
{
    if (!s)
        return fail(HM_ERR_ARGS, "hm_sys_run: NULL system");
    return guarded([&]() { return s->run(nsteps); });
}

int hm_sys_prime(hm_system* s, uint64_t nsteps)
{
    if (!s)
        return fail(HM_ERR_ARGS, "hm_sys_prime: NULL system");
    return guarded([&]() { return s->prime(nsteps); });
}

uint64_t hm_sys_timestep(const hm_system* s)
{
    return s ? s->timestep : 0;
}

int hm_sys_set_timestep(hm_system* s, uint64_t timestep)
{
    if (!s)
        return fail(HM_ERR_ARGS, "hm_sys_set_timestep: NULL system");
    s->timestep = timestep;
    s->prepared = false;
    return HM_OK;
}

int hm_sys_get_state(hm_system* s, double* pos3, double* vel3, int* image3, double* accel3)
{
    if (!s)
        return fail(HM_ERR_ARGS, "hm_sys_get_state: NULL system");
    return guarded([&]() {
        unsigned int N = s->N;
        std::vector<double4> p(N), v(N);
        std::vector<i3> im(N);
        std::vector<d3> ac(N);
        std::vector<unsigned int> tg(N);
        hchk(hipStreamSynchronize(s->stream), "sync");
        hchk(hipMemcpy(p.data(), s->pos, sizeof(double4) * N, hipMemcpyDeviceToHost), "download");
        hchk(hipMemcpy(v.data(), s->vel, sizeof(double4) * N, hipMemcpyDeviceToHost), "download");
        hchk(hipMemcpy(im.data(), s->image, sizeof(i3) * N, hipMemcpyDeviceToHost), "download");
        hchk(hipMemcpy(ac.data(), s->accel, sizeof(d3) * N, hipMemcpyDeviceToHost), "download");
        hchk(hipMemcpy(tg.data(), s->tag, sizeof(unsigned int) * N, hipMemcpyDeviceToHost),
             "download");
        for (unsigned int n = 0; n < N; n++)
        {
            size_t t = tg[n];
            if (pos3)
            {
                pos3[3 * t] = p[n].x;
                pos3[3 * t + 1] = p[n].y;
                pos3[3 * t + 2] = p[n].z;
            }
            if (vel3)
            {
                vel3[3 * t] = v[n].x;
                vel3[3 * t + 1] = v[n].y;
                vel3[3 * t + 2] = v[n].z;
            }
            if (image3)
            {
                image3[3 * t] = im[n].x;
                image3[3 * t + 1] = im[n].y;
                image3[3 * t + 2] = im[n].z;
            }
            if (accel3)
            {
                accel3[3 * t] = ac[n].x;
                accel3[3 * t + 1] = ac[n].y;
                accel3[3 * t + 2] = ac[n].z;
            }
        }
        return HM_OK;
    });
}

int hm_sys_set_state(hm_system* s, const double* pos3, const double* vel3, const int* image3)
{
    if (!s)
        return fail(HM_ERR_ARGS, "hm_sys_set_state: NULL system");
    return guarded([&]() {
        unsigned int N = s->N;
        std::vector<double4> p(N), v(N);
        std::vector<i3> im(N);
        std::vector<unsigned int> tg(N);
        hchk(hipStreamSynchronize(s->stream), "sync");
        hchk(hipMemcpy(p.data(), s->pos, sizeof(double4) * N, hipMemcpyDeviceToHost), "download");
        hchk(hipMemcpy(v.data(), s->vel, sizeof(double4) * N, hipMemcpyDeviceToHost), "download");
        hchk(hipMemcpy(im.data(), s->image, sizeof(i3) * N, hipMemcpyDeviceToHost), "download");
        hchk(hipMemcpy(tg.data(), s->tag, sizeof(unsigned int) * N, hipMemcpyDeviceToHost),
             "download");
        for (unsigned int n = 0; n < N; n++)
        {
            size_t t = tg[n];
            if (pos3)
            {
                p[n].x = pos3[3 * t];
                p[n].y = pos3[3 * t + 1];
                p[n].z = pos3[3 * t + 2];
            }
            if (vel3)
            {
                v[n].x = vel3[3 * t];
                v[n].y = vel3[3 * t + 1];
                v[n].z = vel3[3 * t + 2];
            }
            if (image3)
                im[n] = i3{image3[3 * t], image3[3 * t + 1], image3[3 * t + 2]};
        }
        hchk(hipMemcpy(s->pos, p.data(), sizeof(double4) * N, hipMemcpyHostToDevice), "upload");
        hchk(hipMemcpy(s->vel, v.data(), sizeof(double4) * N, hipMemcpyHostToDevice), "upload");
        hchk(hipMemcpy(s->image, im.data(), sizeof(i3) * N, hipMemcpyHostToDevice), "upload");
        s->prepared = false;
        return HM_OK;
    });
}

int hm_sys_get_forces(hm_system* s, int force_id, double* force4, double* virial6)
{
    if (!s || force_id < 0 || force_id >= (int)s->forces.size())
        return fail(HM_ERR_ARGS, "hm_sys_get_forces: invalid force id");
    return guarded([&]() {
        if (!s->prepared)
            s->prep();
        unsigned int N = s->N;
        auto& f = s->forces[force_id];
        std::vector<float4> fh(N);
        std::vector<float> vh((size_t)6 * N);
        std::vector<unsigned int> tg(N);
        hchk(hipStreamSynchronize(s->stream), "sync");
        if (virial6 && !s->compute_virial)
        {
            s->launch_forces(1);
            hchk(hipStreamSynchronize(s->stream), "sync");
        }
        hchk(hipMemcpy(fh.data(), f.f, sizeof(float4) * N, hipMemcpyDeviceToHost), "download");
        hchk(hipMemcpy(vh.data(), f.v, sizeof(float) * 6 * N, hipMemcpyDeviceToHost), "download");
        hchk(hipMemcpy(tg.data(), s->tag, sizeof(unsigned int) * N, hipMemcpyDeviceToHost),
             "download");
        for (unsigned int n = 0; n < N; n++)
        {
            size_t t = tg[n];
            if (force4)
            {
                force4[4 * t] = fh[n].x;
                force4[4 * t + 1] = fh[n].y;
                force4[4 * t + 2] = fh[n].z;
                force4[4 * t + 3] = fh[n].w;
            }
            if (virial6)
                for (int m = 0; m < 6; m++)
                    virial6[(size_t)m * N + t] = vh[(size_t)m * N + n];
        }
        return HM_OK;
    });
}

long long hm_sys_get_pairs(hm_system* s, unsigned int* pairs, long long max_pairs)
{
    if (!s)
        return fail(HM_ERR_ARGS, "hm_sys_get_pairs: NULL system");
    long long out = 0;
    int r = guarded([&]() {
        if (!s->prepared)
            s->prep();
        unsigned int N = s->N;
        if (!s->nlist)
            return HM_OK;
        std::vector<unsigned int> nl((size_t)N * s->nmax), nn(N), tg(N);
        hchk(hipStreamSynchronize(s->stream), "sync");
        hchk(hipMemcpy(nl.data(), s->nlist, sizeof(unsigned int) * nl.size(),
                       hipMemcpyDeviceToHost),
             "download");
        hchk(hipMemcpy(nn.data(), s->n_neigh, sizeof(unsigned int) * N, hipMemcpyDeviceToHost),
             "download");
        hchk(hipMemcpy(tg.data(), s->tag, sizeof(unsigned int) * N, hipMemcpyDeviceToHost),
             "download");
        if (s->pair_kind() == 3)
        {
            out = s->pairs_from_local_rows(pairs, max_pairs);
            return HM_OK;
        }
        // kind 2 rows hold cell slots with offset codes: map slots back to particles
        const bool slots = s->pair_kind() == 2;
        const unsigned int jmask = slots ? OFFSET_CODE_MASK : 0xffffffffu;
        std::vector<unsigned int> part;
        if (slots)
        {
            part.resize((size_t)s->ncell * s->cmax);
            hchk(hipMemcpy(part.data(), s->cell_idx, sizeof(unsigned int) * part.size(),
                           hipMemcpyDeviceToHost),
                 "download");
        }
        for (unsigned int i = 0; i < N; i++)
            for (unsigned int k = 0; k < nn[i] && k < s->nmax; k++)
            {
                // slot rows are stored in transposed blocks (row_store_pos)
                unsigned int j = nl[(size_t)i * s->nmax + (slots ? row_store_pos(k) : k)] & jmask;
                if (slots)
                    j = part[j];
                unsigned int a = tg[i], b = tg[j];
                if (a < b)
                {
                    if (pairs && out < max_pairs)
                    {
                        pairs[2 * out] = a;
                        pairs[2 * out + 1] = b;
                    }
                    out++;
                }
            }
        return HM_OK;
    });
    return r == HM_OK ? out : r;
}

int hm_sys_thermo(hm_system* s, double* out15)
{
    if (!s || !out15)
        return fail(HM_ERR_ARGS, "hm_sys_thermo: invalid arguments");
    return guarded([&]() {
        if (!s->prepared)
            s->prep();
        if (!s->compute_virial)
            s->launch_forces(1);
        const float4* fp[4];
        const float* vp[4] = {nullptr, nullptr, nullptr, nullptr};
        unsigned int nf = s->force_ptrs(fp);
        for (unsigned int q = 0; q < nf; q++)
            vp[q] = s->forces[q].v;
        hchk(launch_thermo(s->N, s->vel, fp, vp, nf, s->thermo_partial, s->thermo_out, s->stream),
             "thermo");
        hchk(hipMemcpyAsync(s->h_thermo, s->thermo_out, sizeof(double) * THERMO_SUMS, hipMemcpyDeviceToHost,
                            s->stream),
             "thermo");
        hchk(hipStreamSynchronize(s->stream), "sync");
        for (unsigned int k = 0; k < THERMO_SUMS; k++)
            out15[k] = s->h_thermo[k];
        return HM_OK;
    });
}

int hm_sys_stats(hm_system* s, uint64_t* out9)
{
    if (!s || !out9)
        return fail(HM_ERR_ARGS, "hm_sys_stats: invalid arguments");
    return guarded([&]() {
        StepState h{};
        h.shortest = ~0ull;
        if (s->st)
        {
            hchk(hipStreamSynchronize(s->stream), "sync");
            hchk(hipMemcpy(&h, s->st, sizeof(StepState), hipMemcpyDeviceToHost), "download");
        }
        out9[0] = h.n_builds;
        out9[1] = s->nmax;
        out9[2] = s->cmax;
        out9[3] = s->n_sorts;
        out9[4] = s->steps_run;
        out9[5] = s->rollbacks;
        out9[6] = h.n_builds - std::min<uint64_t>(h.n_builds, s->run_builds_base);
        // NeighborList::getSmallestRebuild: periods are tracked below 100 (NeighborList.cc:259)
        out9[7] = std::min<uint64_t>(h.shortest, 100);
        out9[8] = h.n_dangerous;
        return HM_OK;
    });
}

int hm_sys_set_list_capacity(hm_system* s, unsigned int cmax, unsigned int nmax)
{
    if (!s || cmax == 0 || cmax > 256 || nmax == 0)
        return fail(HM_ERR_ARGS, "hm_sys_set_list_capacity: 0 < cmax <= 256, nmax > 0");
    return guarded([&]() {
        if (!s->prepared)
            s->prep();
        // fresh zero-filled tables (every slot and row index stays in range until the forced
        // build below refills them), then a forced build before the forces of the next step
        s->cmax = cmax;
        s->alloc_cells();
        s->nmax = (nmax + ROW_BLOCK - 1) & ~(ROW_BLOCK - 1);
        s->alloc_rows();
        s->clear_graphs();
        hchk(launch_set_rebuild(s->st, s->stream), "force rebuild");
        hchk(hipStreamSynchronize(s->stream), "sync");
        return HM_OK;
    });
}

int hm_sys_set_nlist_policy(hm_system* s, unsigned int rebuild_check_delay, int check_dist)
{
    if (!s)
        return fail(HM_ERR_ARGS, "hm_sys_set_nlist_policy: NULL system");
    s->check_delay = rebuild_check_delay;
    s->dist_check = check_dist != 0;
    s->clear_graphs();
    return HM_OK;
}

int hm_sys_thermalize(hm_system* s, double kT)
{
    if (!s)
        return fail(HM_ERR_ARGS, "hm_sys_thermalize: NULL system");
    return guarded([&]() {
        hchk(launch_thermalize(s->N, s->vel, s->tag, kT, s->timestep, s->seed, s->thermo_partial,
                               s->stream),
             "thermalize");
        unsigned int np = 0;
        hchk(launch_ke_partial(s->N, s->vel, s->ke_partial, &np, s->stream), "kinetic energy");
        s->n_partial = np;
        hchk(hipStreamSynchronize(s->stream), "sync");
        return HM_OK;
    });
}

hm_stream_t hm_sys_stream(hm_system* s)
{
    return s ? (hm_stream_t)s->stream : nullptr;
}

// ------------------------------------------------------------------ learned CG site energy
namespace
{
ForceSlot* mlp_slot(hm_system* s, int force_id)
{
    if (!s || force_id < 0 || force_id >= (int)s->forces.size()
        || s->forces[force_id].kind != 3)
        return nullptr;
    return &s->forces[force_id];
}

void mlp_alloc_training(hm_system* s, ForceSlot& f)
{
    if (f.mT)
        return;
    const unsigned int N = s->N;
    f.mT = dalloc<float4>(N);
    f.mErr = dalloc<float4>(N);
    f.mC = dalloc<float4>((size_t)4 * N);
    f.mLoss = dalloc<float>(mlp_pair_blocks(N));
    f.mGpart = dalloc<float>((size_t)MLP_GRAD_BLOCKS * MLP_NPARAM);
    f.mM1 = dalloc<float>(MLP_NPARAM);
    f.mM2 = dalloc<float>(MLP_NPARAM);
    f.mGrad = dalloc<float>(MLP_NPARAM);
    hchk(hipMemset(f.mM1, 0, sizeof(float) * MLP_NPARAM), "clear");
    hchk(hipMemset(f.mM2, 0, sizeof(float) * MLP_NPARAM), "clear");
}

// the neighbour list of the current positions, built now (host-side regrow on overflow)
void mlp_build_now(hm_system* s)
{
    if (!s->prepared)
        s->prep();  // builds the list
    else if (s->rcut_max > 0.0)
        s->build_nlist_now();
}

// one training frame: positions and targets in, list built, forces, loss and gradient rows
void mlp_frame_pass(hm_system* s, ForceSlot& f, const float* fpos, const float* fforce,
                    double* loss_out)
{
    if (!s->prepared)
        s->prep();
    hchk(launch_mlp_load_frame(s->N, s->rtag, fpos, fforce, s->box, s->pos, f.mT, s->stream),
         "mlp load frame");
    if (s->rcut_max > 0.0)
        s->build_nlist_now();
    MlpArgs a = s->mlp_args(f, 0);
    // training reads the FP32 dE/dG: the loss and its gradient are those of the FP32 forces
    // (the bf16 copy of the production force pass moves forces by about 1e-3 of their size,
    // far below the force-matching residual)
    a.Ab = nullptr;
    hchk(launch_mlp_descriptor(a, s->stream), "mlp descriptor");
    hchk(launch_mlp_site(a, s->stream), "mlp site energies");
    hchk(launch_mlp_force(a, true, s->stream), "mlp forces");
    hchk(launch_mlp_backward(a, s->stream), "mlp backward");
    hchk(launch_mlp_loss(f.mLoss, mlp_pair_blocks(s->N), loss_out, s->stream), "mlp loss");
}
} // namespace

int hm_sys_add_mlp_cg(hm_system* s, const hm_mlp_cg_model* model, const float* params)
{
    if (!s || !model || !params || !(model->r_cut > 0.0) || !(model->width > 0.0)
        || s->forces.size() >= 4)
        return fail(HM_ERR_ARGS, "hm_sys_add_mlp_cg: invalid arguments");
    return guarded([&]() {
        ForceSlot f;
        f.kind = 3;
        f.mm = *model;
        for (unsigned int q = 0; q < s->ntypes * s->ntypes; q++)
            f.rcut[q] = model->r_cut;
        const unsigned int N = s->N;
        f.f = dalloc<float4>(N);
        f.v = dalloc<float>((size_t)6 * N);
        f.mparams = dalloc<float>(MLP_NPARAM);
        f.mbf16 = dalloc<uint16_t>(MLP_NBF16);
        f.mG = dalloc<float4>((size_t)4 * N);
        f.mA = dalloc<float4>((size_t)4 * N);
        f.mAb = dalloc<uint2>((size_t)4 * N);
        if (!s->pos32)
            s->pos32 = dalloc<float4>(N);
        f.mE = dalloc<float>(N);
        hchk(hipMemset(f.f, 0, sizeof(float4) * N), "clear");
        hchk(hipMemset(f.v, 0, sizeof(float) * 6 * N), "clear");
        hchk(hipMemcpy(f.mparams, params, sizeof(float) * MLP_NPARAM, hipMemcpyHostToDevice),
             "upload");
        hchk(launch_mlp_pack(f.mparams, f.mbf16, s->stream), "mlp pack");
        hchk(hipStreamSynchronize(s->stream), "sync");
        s->forces.push_back(f);
        s->prepared = false;
        s->nmax = 0;
        s->clear_graphs();
        return (int)s->forces.size() - 1;
    });
}

int hm_sys_mlp_set_params(hm_system* s, int force_id, const float* params)
{
    ForceSlot* f = mlp_slot(s, force_id);
    if (!f || !params)
        return fail(HM_ERR_ARGS, "hm_sys_mlp_set_params: invalid arguments");
    return guarded([&]() {
        hchk(hipMemcpyAsync(f->mparams, params, sizeof(float) * MLP_NPARAM,
                            hipMemcpyHostToDevice, s->stream),
             "upload");
        hchk(launch_mlp_pack(f->mparams, f->mbf16, s->stream), "mlp pack");
        if (f->mM1)
        {
            hchk(hipMemsetAsync(f->mM1, 0, sizeof(float) * MLP_NPARAM, s->stream), "clear");
            hchk(hipMemsetAsync(f->mM2, 0, sizeof(float) * MLP_NPARAM, s->stream), "clear");
        }
        f->adam_t = 0;
        hchk(hipStreamSynchronize(s->stream), "sync");
        s->prepared = false;
        return HM_OK;
    });
}

int hm_sys_mlp_set_pair_table(hm_system* s, int force_id, unsigned int width, double r_min,
                              const double* U, const double* F)
{
    ForceSlot* f = mlp_slot(s, force_id);
    if (!f || (width && (!U || !F || !(r_min >= 0.0) || !(r_min < f->mm.r_cut))))
        return fail(HM_ERR_ARGS, "hm_sys_mlp_set_pair_table: invalid arguments (0 <= r_min < "
                                 "r_cut, U and F of width points)");
    return guarded([&]() {
        hchk(hipStreamSynchronize(s->stream), "sync");
        if (width == 0)
        {
            dfree(f->mtab);
            f->mtab = nullptr;
            f->mtab_w = 0;
            s->clear_graphs();
            s->prepared = false;
            return HM_OK;
        }
        std::vector<float> h(2 * (size_t)width);
        for (unsigned int k = 0; k < width; k++)
        {
            h[k] = (float)U[k];
            h[width + k] = (float)F[k];
        }
        // same shape: overwrite in place, so captured step graphs stay valid
        if (!f->mtab || f->mtab_w != width || f->mtab_rmin != r_min)
        {
            dfree(f->mtab);
            f->mtab = nullptr;
            f->mtab = dalloc<float>(2 * (size_t)width);
            f->mtab_w = width;
            f->mtab_rmin = r_min;
            s->clear_graphs();
        }
        hchk(hipMemcpy(f->mtab, h.data(), sizeof(float) * h.size(), hipMemcpyHostToDevice),
             "upload");
        s->prepared = false;
        return HM_OK;
    });
}

int hm_sys_mlp_get_params(hm_system* s, int force_id, float* params)
{
    ForceSlot* f = mlp_slot(s, force_id);
    if (!f || !params)
        return fail(HM_ERR_ARGS, "hm_sys_mlp_get_params: invalid arguments");
    return guarded([&]() {
        hchk(hipStreamSynchronize(s->stream), "sync");
        hchk(hipMemcpy(params, f->mparams, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToHost),
             "download");
        return HM_OK;
    });
}

int hm_sys_mlp_set_model(hm_system* s, int force_id, const hm_mlp_cg_model* model)
{
    ForceSlot* f = mlp_slot(s, force_id);
    if (!f || !model || model->r_cut != f->mm.r_cut || !(model->width > 0.0))
        return fail(HM_ERR_ARGS, "hm_sys_mlp_set_model: invalid arguments (r_cut is fixed)");
    f->mm = *model;
    s->prepared = false;
    s->clear_graphs();
    return HM_OK;
}

int hm_sys_mlp_descriptors(hm_system* s, int force_id, const float* frame_pos, float* G)
{
    ForceSlot* f = mlp_slot(s, force_id);
    if (!f || !G)
        return fail(HM_ERR_ARGS, "hm_sys_mlp_descriptors: invalid arguments");
    return guarded([&]() {
        if (frame_pos)
        {
            if (!s->prepared)
                s->prep();
            hchk(launch_mlp_load_frame(s->N, s->rtag, frame_pos, nullptr, s->box, s->pos,
                                       nullptr, s->stream),
                 "mlp load frame");
            s->prepared = false;  // the next run starts from these positions: prepare again
            if (s->rcut_max > 0.0)
                s->build_nlist_now();
        }
        else
            mlp_build_now(s);
        MlpArgs a = s->mlp_args(*f, 0);
        hchk(launch_mlp_descriptor(a, s->stream), "mlp descriptor");
        std::vector<float> g((size_t)16 * s->N);
        std::vector<unsigned int> tg(s->N);
        hchk(hipStreamSynchronize(s->stream), "sync");
        hchk(hipMemcpy(g.data(), f->mG, sizeof(float) * g.size(), hipMemcpyDeviceToHost),
             "download");
        hchk(hipMemcpy(tg.data(), s->tag, sizeof(unsigned int) * s->N, hipMemcpyDeviceToHost),
             "download");
        for (unsigned int i = 0; i < s->N; i++)
            memcpy(G + (size_t)16 * tg[i], &g[(size_t)16 * i], sizeof(float) * 16);
        return HM_OK;
    });
}

int hm_sys_frame_subtract_force(hm_system* s, int force_id, const float* frame_pos,
                                float* frame_force)
{
    if (!s || force_id < 0 || force_id >= (int)s->forces.size() || !frame_pos || !frame_force
        || !(is_pair(s->forces[force_id]) && s->forces[force_id].kind != 3))
        return fail(HM_ERR_ARGS, "hm_sys_frame_subtract_force: a pair force and device frames");
    return guarded([&]() {
        if (!s->prepared)
            s->prep();
        hchk(launch_mlp_load_frame(s->N, s->rtag, frame_pos, nullptr, s->box, s->pos, nullptr,
                                   s->stream),
             "load frame");
        s->prepared = false;  // the next run starts from these positions: prepare again
        s->build_nlist_now();
        ForceSlot& f = s->forces[force_id];
        s->launch_pair(f, 0);
        hipLaunchKernelGGL(subtract_by_tag_kernel, dim3((s->N + 255) / 256), dim3(256), 0,
                           s->stream, s->N, s->tag, f.f, frame_force);
        hchk(hipGetLastError(), "subtract force");
        hchk(hipStreamSynchronize(s->stream), "sync");
        return HM_OK;
    });
}

int hm_sys_mlp_loss_grad(hm_system* s, int force_id, const float* frame_pos,
                         const float* frame_force, double* loss, float* grad)
{
    ForceSlot* f = mlp_slot(s, force_id);
    if (!f || !frame_pos || !frame_force || !loss)
        return fail(HM_ERR_ARGS, "hm_sys_mlp_loss_grad: invalid arguments");
    return guarded([&]() {
        mlp_alloc_training(s, *f);
        if (!f->mLossOut)
            f->mLossOut = dalloc<double>(1);
        mlp_frame_pass(s, *f, frame_pos, frame_force, f->mLossOut);
        hchk(launch_mlp_grad_sum(f->mGpart, f->mGrad, s->stream), "mlp gradient");
        hchk(hipStreamSynchronize(s->stream), "sync");
        double sq = 0.0;
        hchk(hipMemcpy(&sq, f->mLossOut, sizeof(double), hipMemcpyDeviceToHost), "download");
        *loss = sq / (3.0 * (double)s->N);
        if (grad)
            hchk(hipMemcpy(grad, f->mGrad, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToHost),
                 "download");
        s->prepared = false;
        return HM_OK;
    });
}

int hm_sys_mlp_train(hm_system* s, int force_id, const hm_mlp_train_args* t)
{
    ForceSlot* f = mlp_slot(s, force_id);
    if (!f || !t || !t->frame_pos || !t->frame_force || t->nframes == 0
        || !(t->lr > 0.0) || !(t->beta1 >= 0.0 && t->beta1 < 1.0)
        || !(t->beta2 >= 0.0 && t->beta2 < 1.0) || !(t->eps > 0.0))
        return fail(HM_ERR_ARGS, "hm_sys_mlp_train: invalid arguments");
    const unsigned int B = t->frames_per_step ? t->frames_per_step : 1;
    if (t->order)
        for (size_t k = 0; k < (size_t)t->nsteps * B; k++)
            if (t->order[k] >= t->nframes)
                return fail(HM_ERR_ARGS, "hm_sys_mlp_train: frame order out of range");
    return guarded([&]() {
        mlp_alloc_training(s, *f);
        dfree(f->mLossOut);
        f->mLossOut = dalloc<double>((size_t)t->nsteps * B + 1);
        if (B > 1 && !f->mGacc)
            f->mGacc = dalloc<float>(MLP_NPARAM);
        const size_t stride = (size_t)3 * s->N;
        for (unsigned int k = 0; k < t->nsteps; k++)
        {
            for (unsigned int b = 0; b < B; b++)
            {
                const size_t q = (size_t)k * B + b;
                const unsigned int fr = t->order ? t->order[q] : (unsigned int)(q % t->nframes);
                mlp_frame_pass(s, *f, t->frame_pos + stride * fr, t->frame_force + stride * fr,
                               f->mLossOut + q);
                if (B > 1)
                    hchk(launch_mlp_grad_accum(f->mGpart, f->mGacc, b == 0, s->stream),
                         "mlp gradient sum");
            }
            if (B > 1)
                hchk(launch_mlp_adam_acc(f->mparams, f->mM1, f->mM2, f->mGacc, 1.0f / (float)B,
                                         f->mbf16, (float)t->lr, (float)t->beta1,
                                         (float)t->beta2, (float)t->eps, ++f->adam_t, s->stream),
                     "mlp adam");
            else
                hchk(launch_mlp_adam(f->mparams, f->mM1, f->mM2, f->mGpart, f->mbf16,
                                     (float)t->lr, (float)t->beta1, (float)t->beta2,
                                     (float)t->eps, ++f->adam_t, s->stream),
                     "mlp adam");
        }
        hchk(hipStreamSynchronize(s->stream), "sync");
        if (t->loss && t->nsteps)
        {
            std::vector<double> h((size_t)t->nsteps * B);
            hchk(hipMemcpy(h.data(), f->mLossOut, sizeof(double) * h.size(),
                           hipMemcpyDeviceToHost),
                 "download");
            for (unsigned int k = 0; k < t->nsteps; k++)
            {
                double sum = 0.0;
                for (unsigned int b = 0; b < B; b++)
                    sum += h[(size_t)k * B + b];
                t->loss[k] = sum / B / (3.0 * (double)s->N);
            }
        }
        s->prepared = false;
        return HM_OK;
    });
}

int hm_sys_set_site_map(hm_system* s, unsigned int nsites, const unsigned int* row,
                        const unsigned int* tags)
{
    if (!s || (nsites && (!row || !tags)))
        return fail(HM_ERR_ARGS, "hm_sys_set_site_map: invalid arguments");
    for (unsigned int I = 0; I < nsites; I++)
        if (row[I + 1] <= row[I])
            return fail(HM_ERR_ARGS, "hm_sys_set_site_map: every site needs a particle");
    for (unsigned int b = 0; b < (nsites ? row[nsites] : 0u); b++)
        if (tags[b] >= s->N)
            return fail(HM_ERR_ARGS, "hm_sys_set_site_map: tag out of range");
    return guarded([&]() {
        dfree(s->site_row);
        dfree(s->site_tags);
        s->nsites = nsites;
        if (nsites)
        {
            s->site_row = dalloc<unsigned int>(nsites + 1);
            s->site_tags = dalloc<unsigned int>(row[nsites]);
            hchk(hipMemcpy(s->site_row, row, sizeof(unsigned int) * (nsites + 1),
                           hipMemcpyHostToDevice),
                 "upload");
            hchk(hipMemcpy(s->site_tags, tags, sizeof(unsigned int) * row[nsites],
                           hipMemcpyHostToDevice),
                 "upload");
        }
        return HM_OK;
    });
}

int hm_sys_map_frame(hm_system* s, float* d_pos, float* d_force)
{
    if (!s || !d_pos || !s->nsites)
        return fail(HM_ERR_ARGS, "hm_sys_map_frame: invalid arguments (no site map)");
    return guarded([&]() {
        if (!s->prepared)
            s->prep();
        SiteMapArgs a;
        memset(&a, 0, sizeof(a));
        a.nsites = s->nsites;
        a.row = s->site_row;
        a.tags = s->site_tags;
        a.rtag = s->rtag;
        a.pos = s->pos;
        a.vel = s->vel;
        a.nforces = s->force_ptrs(a.f);
        a.box = s->box;
        a.out_pos = d_pos;
        a.out_force = d_force;
        hchk(launch_map_sites(a, s->stream), "map sites");
        hchk(hipStreamSynchronize(s->stream), "sync");
        return HM_OK;
    });
}

int hm_sys_time_pair(hm_system* s, int force_id, unsigned int iters, double* avg_ms,
                     uint64_t* nlist_entries)
{
    if (!s || force_id < 0 || force_id >= (int)s->forces.size() || !is_pair(s->forces[force_id])
        || iters == 0 || !avg_ms)
        return fail(HM_ERR_ARGS, "hm_sys_time_pair: invalid arguments");
    return guarded([&]() {
        if (!s->prepared)
            s->prep();
        auto& f = s->forces[force_id];
        hipEvent_t e0, e1;
        hchk(hipEventCreate(&e0), "event");
        hchk(hipEventCreate(&e1), "event");
        s->launch_pair(f, s->compute_virial);
        hchk(hipEventRecord(e0, s->stream), "event");
        for (unsigned int q = 0; q < iters; q++)
            s->launch_pair(f, s->compute_virial);
        hchk(hipEventRecord(e1, s->stream), "event");
        hchk(hipEventSynchronize(e1), "event");
        float ms = 0.f;
        hchk(hipEventElapsedTime(&ms, e0, e1), "event");
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        *avg_ms = (double)ms / iters;
        if (nlist_entries)
        {
            std::vector<unsigned int> nn(s->N);
            hchk(hipMemcpy(nn.data(), s->n_neigh, sizeof(unsigned int) * s->N,
                           hipMemcpyDeviceToHost),
                 "download");
            uint64_t tot = 0;
            for (unsigned int v : nn)
                tot += std::min(v, s->nmax);
            *nlist_entries = tot;
        }
        return HM_OK;
    });
}

int hm_sys_time_build(hm_system* s, unsigned int iters, double* avg_ms2)
{
    if (!s || iters == 0 || !avg_ms2)
        return fail(HM_ERR_ARGS, "hm_sys_time_build: invalid arguments");
    return guarded([&]() {
        if (!s->prepared)
            s->prep();
        if (s->rcut_max <= 0.0)
            throw HmError{HM_ERR_STATE, "hm_sys_time_build: no neighbour list"};
        hipEvent_t e[3];
        for (auto& x : e)
            hchk(hipEventCreate(&x), "event");
        float tc = 0.f, tn = 0.f;
        for (unsigned int q = 0; q < iters; q++)
        {
            // a complete rebuild at the current positions, launched as a step launches it
            hchk(hipEventRecord(e[0], s->stream), "event");
            hchk(launch_cell_clear(s->ncell, s->cell_size, nullptr, s->stream), "cell clear");
            if (s->N <= CELL_FUSED_MAX_N)
                hchk(launch_cell_bin_rank(s->N, s->dim, s->cmax, s->pos, s->cell_size, s->last,
                                          &s->cond->cell_need, s->ckey, s->chome, s->entries,
                                          s->cell_idx, nullptr, s->stream, s->box,
                                          s->rel_frame()),
                     "cell build and order");
            else
            {
                hchk(launch_cell_build(s->N, s->pos, s->box, s->dim, s->cmax, s->cell_size,
                                       nullptr, s->last, &s->cond->cell_need, nullptr, s->stream,
                                       s->ckey, s->chome),
                     "cell build");
                hchk(launch_cell_rank(s->N, s->cmax, s->pos, s->cell_size, s->ckey, s->chome,
                                      s->entries, s->cell_idx, nullptr, s->stream, s->box,
                                      s->rel_frame()),
                     "cell order");
            }
            hchk(hipEventRecord(e[1], s->stream), "event");
            NlistKernelArgs k;
            s->fill_nlist_args(k, false);
            k.counter = nullptr;
            hchk(launch_nlist_build(k, s->stream), "nlist build");
            hchk(hipEventRecord(e[2], s->stream), "event");
            hchk(hipEventSynchronize(e[2]), "event");
            float a = 0.f, b = 0.f;
            hchk(hipEventElapsedTime(&a, e[0], e[1]), "event");
            hchk(hipEventElapsedTime(&b, e[1], e[2]), "event");
            tc += a;
            tn += b;
        }
        for (auto& x : e)
            (void)hipEventDestroy(x);
        avg_ms2[0] = (double)tc / iters;
        avg_ms2[1] = (double)tn / iters;
        // the run continues with the regular build machinery
        hchk(launch_set_rebuild(s->st, s->stream), "force rebuild");
        hchk(hipStreamSynchronize(s->stream), "sync");
        return HM_OK;
    });
}

int hm_sys_time_mlp(hm_system* s, int force_id, unsigned int iters, double* avg_ms3)
{
    if (!s || force_id < 0 || force_id >= (int)s->forces.size()
        || s->forces[force_id].kind != 3 || iters == 0 || !avg_ms3)
        return fail(HM_ERR_ARGS, "hm_sys_time_mlp: invalid arguments");
    return guarded([&]() {
        if (!s->prepared)
            s->prep();
        MlpArgs a = s->mlp_args(s->forces[force_id], s->compute_virial);
        a.A = nullptr;  // as in a step: the bf16 copy only
        if (s->pos32)
        {
            hchk(launch_pos_f32(s->N, s->pos, s->pos32, s->stream), "fp32 positions");
            a.pos32 = s->pos32;
        }
        hipEvent_t e[4];
        for (auto& x : e)
            hchk(hipEventCreate(&x), "event");
        // one untimed pass, then iters launches of each kernel in turn (each pass reads the
        // previous one's output, as in a step)
        hchk(launch_mlp_descriptor(a, s->stream), "mlp descriptor");
        hchk(launch_mlp_site(a, s->stream), "mlp site energies");
        hchk(launch_mlp_force(a, false, s->stream), "mlp forces");
        hchk(hipEventRecord(e[0], s->stream), "event");
        for (unsigned int q = 0; q < iters; q++)
            hchk(launch_mlp_descriptor(a, s->stream), "mlp descriptor");
        hchk(hipEventRecord(e[1], s->stream), "event");
        for (unsigned int q = 0; q < iters; q++)
            hchk(launch_mlp_site(a, s->stream), "mlp site energies");
        hchk(hipEventRecord(e[2], s->stream), "event");
        for (unsigned int q = 0; q < iters; q++)
            hchk(launch_mlp_force(a, false, s->stream), "mlp forces");
        hchk(hipEventRecord(e[3], s->stream), "event");
        hchk(hipEventSynchronize(e[3]), "event");
        for (int k = 0; k < 3; k++)
        {
            float ms = 0.f;
            hchk(hipEventElapsedTime(&ms, e[k], e[k + 1]), "event");
            avg_ms3[k] = (double)ms / iters;
        }
        for (auto& x : e)
            (void)hipEventDestroy(x);
        return HM_OK;
    });
}

} // extern "C"
