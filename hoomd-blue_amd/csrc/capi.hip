// capi.hip -- extern "C" kernel drivers declared in include/hoomd_amd.h.
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>

#include "capi_util.h"
#include "kernels.h"

using namespace hm;

namespace hm
{
thread_local std::string g_last_error;

int fail(int code, const char* what)
{
    g_last_error = what;
    return code;
}

int check(hipError_t e, const char* what)
{
    if (e == hipSuccess)
        return HM_OK;
    g_last_error = std::string(what) + ": " + hipGetErrorString(e);
    return HM_ERR_LAUNCH;
}
} // namespace hm

namespace
{
// Partial sums of hm_thermo_noalloc: one row of THERMO_SUMS doubles per partial-sum workgroup
// (THERMO_BLOCKS rows, 60 KiB), a copy per device, held until the process ends. Host code uses
// only the per-device addresses taken with hipGetSymbolAddress below, never this name.
constexpr size_t THERMO_SCRATCH = (size_t)THERMO_BLOCKS * THERMO_SUMS;
static __device__ __align__(256) double scratch[THERMO_SCRATCH];

constexpr int SCRATCH_MAX_DEVICES = 64;
std::mutex g_scratch_mutex;
double* g_scratch_ptr[SCRATCH_MAX_DEVICES] = {};

// Address of the current device's scratch array, looked up once per device (with that device
// current, as every call here runs).
int scratch_for_current_device(double** out)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= SCRATCH_MAX_DEVICES)
        return fail(HM_ERR_LAUNCH, "hm_thermo_noalloc: no usable current device");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev])
    {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return check(e != hipSuccess ? e : hipErrorInvalidSymbol,
                         "hm_thermo_noalloc: scratch address");
        g_scratch_ptr[dev] = (double*)p;
    }
    *out = g_scratch_ptr[dev];
    return HM_OK;
}
} // namespace

extern "C"
{

const char* hm_version(void)
{
    return "hoomd_amd 0.1 (gfx950)";
}

int hm_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess)
        return 0;
    return n;
}

const char* hm_last_error(void)
{
    return g_last_error.c_str();
}

int hm_malloc(void** ptr, size_t bytes)
{
    if (!ptr)
        return fail(HM_ERR_ARGS, "hm_malloc: NULL ptr");
    if (hipMalloc(ptr, bytes ? bytes : 16) != hipSuccess)
        return fail(HM_ERR_ALLOC, "hm_malloc: hipMalloc failed");
    return HM_OK;
}

int hm_free(void* ptr)
{
    return check(hipFree(ptr), "hm_free");
}

int hm_memcpy_htod(void* dst, const void* src, size_t bytes)
{
    return check(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice), "hm_memcpy_htod");
}

int hm_memcpy_dtoh(void* dst, const void* src, size_t bytes)
{
    return check(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost), "hm_memcpy_dtoh");
}

int hm_memcpy_dtod(void* dst, const void* src, size_t bytes)
{
    return check(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice), "hm_memcpy_dtod");
}

int hm_memset(void* dst, int value, size_t bytes)
{
    return check(hipMemset(dst, value, bytes), "hm_memset");
}

int hm_synchronize(void)
{
    return check(hipDeviceSynchronize(), "hm_synchronize");
}

int hm_cell_dims(const hm_box* box, double w, unsigned int dim[3])
{
    if (!box || !dim || !(w > 0.0))
        return fail(HM_ERR_ARGS, "hm_cell_dims: invalid arguments");
    for (int d = 0; d < 3; d++)
    {
        dim[d] = (unsigned int)((box->L[d] + 2.0 * 0.0) / w);
        if (dim[d] == 0)
            dim[d] = 1;
    }
    return HM_OK;
}

unsigned int hm_cell_n_adj(const unsigned int dim[3])
{
    unsigned int n = 1;
    for (int d = 0; d < 3; d++)
        n *= dim[d] > 3 ? 3 : dim[d];
    return n;
}

int hm_cell_adj(const unsigned int dim[3], unsigned int* d_adj, hm_stream_t stream)
{
    if (!dim || !d_adj || dim[0] == 0 || dim[1] == 0 || dim[2] == 0)
        return fail(HM_ERR_ARGS, "hm_cell_adj: invalid arguments");
    return check(launch_cell_adj(dim, d_adj, (hipStream_t)stream), "hm_cell_adj");
}

int hm_cell_list_build(const hm_cell_args* a, hm_stream_t stream)
{
    if (!a || !a->pos || !a->cell_size || !a->entries || !a->conditions || a->cmax == 0)
        return fail(HM_ERR_ARGS, "hm_cell_list_build: invalid arguments");
    if (a->cmax > 256)
        return fail(HM_ERR_ARGS, "hm_cell_list_build: cmax above 256");
    hipStream_t s = (hipStream_t)stream;
    size_t ncell = (size_t)a->dim[0] * a->dim[1] * a->dim[2];
    int r = check(hipMemsetAsync(a->cell_size, 0, ncell * sizeof(unsigned int), s),
                  "hm_cell_list_build: clear");
    if (r)
        return r;
    DevBox box = make_devbox(a->box);
    r = check(launch_cell_build(a->N, (const double4*)a->pos, box, a->dim, a->cmax, a->cell_size,
                                a->entries, (d3*)a->last_pos, a->conditions, nullptr, s),
              "hm_cell_list_build");
    if (r)
        return r;
    return check(launch_cell_sort((unsigned int)ncell, a->cmax, a->cell_size, a->entries,
                                  a->cell_idx, nullptr, s),
                 "hm_cell_list_build: order");
}

int hm_nlist_build(const hm_nlist_args* a, hm_stream_t stream)
{
    if (!a || !a->cell_size || !a->entries || !a->cell_adj || !a->nlist || !a->n_neigh
        || !a->conditions || a->ntypes == 0 || a->ntypes > HM_MAX_TYPES || (a->nmax & 7u)
        || a->nmax == 0)
        return fail(HM_ERR_ARGS, "hm_nlist_build: invalid arguments");
    if (a->cmax > 256)
        return fail(HM_ERR_ARGS, "hm_nlist_build: cmax above 256");
    if (a->n_excl && a->max_excl > 4)
        return fail(HM_ERR_ARGS, "hm_nlist_build: at most 4 exclusions per particle");
    NlistKernelArgs k;
    memset(&k, 0, sizeof(k));
    k.N = a->N;
    k.box = make_devbox(a->box);
    k.dim[0] = a->dim[0];
    k.dim[1] = a->dim[1];
    k.dim[2] = a->dim[2];
    k.pos = (const double4*)a->pos;
    k.ncell = a->dim[0] * a->dim[1] * a->dim[2];
    k.cmax = a->cmax;
    k.cell_size = a->cell_size;
    k.entries = a->entries;
    k.cell_adj = a->cell_adj;
    k.n_adj = a->n_adj;
    k.ntypes = a->ntypes;
    for (unsigned int q = 0; q < a->ntypes * a->ntypes; q++)
    {
        // NeighborList::updateRList (NeighborList.cc:515-520)
        double rc = a->r_cut[q];
        double rl = rc > 0.0 ? rc + a->r_buff : 0.0;
        k.rlistsq[q] = rl * rl;
        k.enabled[q] = rc > 0.0 ? 1u : 0u;
    }
    k.n_excl = a->n_excl;
    k.excl = a->excl;
    k.max_excl = a->n_excl ? a->max_excl : 0;
    k.nmax = a->nmax;
    k.nlist = a->nlist;
    k.n_neigh = a->n_neigh;
    k.lrow = a->local_rows;
    k.cond = a->conditions;
    return check(launch_nlist_build(k, (hipStream_t)stream), "hm_nlist_build");
}

int hm_pair_lj_cells(const hm_pair_cell_args* a, const hm_lj_params* params, hm_stream_t stream)
{
    if (!a || !params || !a->pos || !a->cell_size || !a->cell_idx || !a->cell_adj
        || !a->local_rows || !a->n_neigh || !a->force || (a->compute_virial && !a->virial)
        || a->ntypes == 0 || a->ntypes > HM_MAX_TYPES || a->nmax == 0 || a->cmax == 0
        || a->cmax > 256 || a->dim[0] == 0 || a->dim[1] == 0 || a->dim[2] == 0)
        return fail(HM_ERR_ARGS, "hm_pair_lj_cells: invalid arguments");
    if (a->tpp && a->tpp != 4 && a->tpp != 8 && a->tpp != 16)
        return fail(HM_ERR_ARGS, "hm_pair_lj_cells: tpp must be 4, 8 or 16");
    if (a->shift_mode < 0 || a->shift_mode > 2)
        return fail(HM_ERR_ARGS, "hm_pair_lj_cells: shift_mode must be 0, 1 or 2");
    PairCellArgs c;
    memset(&c, 0, sizeof(c));
    c.N = a->N;
    c.pos = (const double4*)a->pos;
    c.box = make_devbox(a->box);
    for (int d = 0; d < 3; d++)
        c.dim[d] = a->dim[d];
    c.ncell = a->dim[0] * a->dim[1] * a->dim[2];
    c.cmax = a->cmax;
    c.cell_size = a->cell_size;
    c.cell_idx = a->cell_idx;
    c.cell_adj = a->cell_adj;
    c.n_adj = a->n_adj;
    c.lrow = a->local_rows;
    c.nmax = a->nmax;
    c.n_neigh = a->n_neigh;
    c.ntypes = a->ntypes;
    for (unsigned int q = 0; q < a->ntypes * a->ntypes; q++)
    {
        const double rc = a->r_cut[q] > 0.0 ? a->r_cut[q] : 0.0;
        c.rcutsq[q] = rc * rc;
        c.rcutsq_f[q] = (float)(rc * rc);
        c.ronsq_f[q] = (float)(a->r_on[q] * a->r_on[q]);
    }
    c.shift_mode = a->shift_mode;
    c.compute_virial = a->compute_virial;
    c.force = (float4*)a->force;
    c.virial = a->virial;
    PairLaunch cfg{a->tpp, a->workgroups};
    return check(launch_pair_cell_lj(c, params, cfg, (hipStream_t)stream), "hm_pair_lj_cells");
}

int hm_nlist_check(unsigned int N, const double* pos, const double* last_pos, hm_box box,
                   double r_cut_max, double r_buff, unsigned int* flag, hm_stream_t stream)
{
    if (!pos || !last_pos || !flag)
        return fail(HM_ERR_ARGS, "hm_nlist_check: invalid arguments");
    double maxsq = distance_check_maxsq(r_cut_max, r_buff);
    return check(launch_nlist_check(N, (const double4*)pos, (const d3*)last_pos, make_devbox(box),
                                    maxsq, flag, (hipStream_t)stream),
                 "hm_nlist_check");
}

int hm_pair_lj(const hm_pair_args* a, const hm_lj_params* params, hm_stream_t stream)
{
    if (!a || !params || !a->pos || !a->nlist || !a->n_neigh || !a->force
        || (a->compute_virial && !a->virial) || a->ntypes == 0 || a->ntypes > HM_MAX_TYPES
        || (a->nmax & 7u) || a->nmax == 0)
        return fail(HM_ERR_ARGS, "hm_pair_lj: invalid arguments");
    if (a->tpp && a->tpp != 1 && a->tpp != 2 && a->tpp != 4 && a->tpp != 8 && a->tpp != 16)
        return fail(HM_ERR_ARGS, "hm_pair_lj: tpp must be 1, 2, 4, 8 or 16");
    if (a->block_size && (a->block_size % 64 || a->block_size > 1024))
        return fail(HM_ERR_ARGS, "hm_pair_lj: block_size must be a multiple of 64 <= 1024");
    if (a->shift_mode < 0 || a->shift_mode > 2)
        return fail(HM_ERR_ARGS, "hm_pair_lj: shift_mode must be 0, 1 or 2");
    PairKernelArgs k;
    fill_pair_args(k, a->N, (const double4*)a->pos, a->box, a->nlist, a->nmax, a->n_neigh,
                   a->ntypes, a->r_cut, a->r_on, a->shift_mode, a->compute_virial,
                   (float4*)a->force, a->virial);
    PairLaunch cfg{a->tpp, a->block_size};
    return check(launch_pair_lj(k, params, cfg, (hipStream_t)stream), "hm_pair_lj");
}

int hm_pair_table(const hm_pair_args* a, const hm_table_params* params, hm_stream_t stream)
{
    if (!a || !params || !a->pos || !a->nlist || !a->n_neigh || !a->force
        || (a->compute_virial && !a->virial) || a->ntypes == 0 || a->ntypes > HM_MAX_TYPES
        || (a->nmax & 7u) || a->nmax == 0)
        return fail(HM_ERR_ARGS, "hm_pair_table: invalid arguments");
    if (a->tpp && a->tpp != 1 && a->tpp != 2 && a->tpp != 4 && a->tpp != 8 && a->tpp != 16)
        return fail(HM_ERR_ARGS, "hm_pair_table: tpp must be 1, 2, 4, 8 or 16");
    if (a->block_size && (a->block_size % 64 || a->block_size > 1024))
        return fail(HM_ERR_ARGS, "hm_pair_table: block_size must be a multiple of 64 <= 1024");
    if (a->shift_mode != 0)
        return fail(HM_ERR_ARGS, "hm_pair_table: tables support shift_mode 0 (none) only");
    const unsigned int nt = a->ntypes;
    for (unsigned int q = 0; q < nt * nt; q++)
        if (params[q].width && a->r_cut[q] > 0.0
            && (!params[q].U || !params[q].F || !(params[q].r_min < a->r_cut[q])))
            return fail(HM_ERR_ARGS, "hm_pair_table: a table needs U, F and r_min < r_cut");
    PairKernelArgs k;
    fill_pair_args(k, a->N, (const double4*)a->pos, a->box, a->nlist, a->nmax, a->n_neigh,
                   a->ntypes, a->r_cut, a->r_on, 0, a->compute_virial, (float4*)a->force,
                   a->virial);
    PairLaunch cfg{a->tpp, a->block_size};
    return check(launch_pair_table(k, params, cfg, (hipStream_t)stream), "hm_pair_table");
}

namespace
{
bool bond_args_ok(const hm_bond_args* a)
{
    return a && a->pos && a->tag && a->rtag && a->bond_row && a->bond_other && a->force
           && (!a->compute_virial || a->virial) && a->nbond_types <= HM_MAX_TYPES;
}

BondKernelArgs bond_kernel_args(const hm_bond_args* a)
{
    BondKernelArgs k;
    memset(&k, 0, sizeof(k));
    k.N = a->N;
    k.pos = (const double4*)a->pos;
    k.box = make_devbox(a->box);
    k.tag = a->tag;
    k.rtag = a->rtag;
    k.row = a->bond_row;
    k.other_tag = a->bond_other;
    k.btype = a->bond_type;
    k.nbond_types = a->nbond_types ? a->nbond_types : 1;
    k.compute_virial = a->compute_virial;
    k.force = (float4*)a->force;
    k.virial = a->virial;
    k.flag = a->flag;
    return k;
}
} // namespace

int hm_bond_harmonic(const hm_bond_args* a, const hm_harmonic_params* params, hm_stream_t stream)
{
    if (!bond_args_ok(a) || !params)
        return fail(HM_ERR_ARGS, "hm_bond_harmonic: invalid arguments");
    return check(launch_bond_harmonic(bond_kernel_args(a), params, (hipStream_t)stream),
                 "hm_bond_harmonic");
}

int hm_bond_fene(const hm_bond_args* a, const hm_fene_params* params, hm_stream_t stream)
{
    if (!a || !params || !a->pos || !a->tag || !a->rtag || !a->bond_row || !a->bond_other
        || !a->force || (a->compute_virial && !a->virial) || a->nbond_types > HM_MAX_TYPES)
        return fail(HM_ERR_ARGS, "hm_bond_fene: invalid arguments");
    BondKernelArgs k;
    memset(&k, 0, sizeof(k));
    k.N = a->N;
    k.pos = (const double4*)a->pos;
    k.box = make_devbox(a->box);
    k.tag = a->tag;
    k.rtag = a->rtag;
    k.row = a->bond_row;
    k.other_tag = a->bond_other;
    k.btype = a->bond_type;
    k.nbond_types = a->nbond_types ? a->nbond_types : 1;
    k.compute_virial = a->compute_virial;
    k.force = (float4*)a->force;
    k.virial = a->virial;
    k.flag = a->flag;
    return check(launch_bond_fene(k, params, (hipStream_t)stream), "hm_bond_fene");
}

int hm_nvt_step_one(unsigned int N, double* pos, double* vel, const double* accel, int* image,
                    hm_box box, double dt, double alpha, hm_stream_t stream)
{
    if (!pos || !vel || !accel || !image)
        return fail(HM_ERR_ARGS, "hm_nvt_step_one: invalid arguments");
    StepOneArgs a;
    memset(&a, 0, sizeof(a));
    a.N = N;
    a.pos = (double4*)pos;
    a.vel = (double4*)vel;
    a.accel = (const d3*)accel;
    a.image = (i3*)image;
    a.box = make_devbox(box);
    a.dt = dt;
    a.alpha = alpha;
    return check(launch_step_one(a, (hipStream_t)stream), "hm_nvt_step_one");
}

int hm_nvt_step_two(unsigned int N, double* vel, double* accel, const float* const* forces,
                    unsigned int nforces, double dt, double alpha, hm_stream_t stream)
{
    if (!vel || !accel || (nforces && !forces) || nforces > 4)
        return fail(HM_ERR_ARGS, "hm_nvt_step_two: invalid arguments");
    StepTwoArgs a;
    memset(&a, 0, sizeof(a));
    a.N = N;
    a.vel = (double4*)vel;
    a.accel = (d3*)accel;
    for (unsigned int q = 0; q < nforces; q++)
        a.f[q] = (const float4*)forces[q];
    a.nforces = nforces;
    a.dt = dt;
    a.alpha = alpha;
    return check(launch_step_two(a, (hipStream_t)stream, nullptr), "hm_nvt_step_two");
}

int hm_langevin_step_two(unsigned int N, double* vel, double* accel, const double* pos,
                         const unsigned int* tag, const float* const* forces,
                         unsigned int nforces, const double* gamma_by_type, unsigned int ntypes,
                         double kT, double dt, uint64_t timestep, uint16_t seed,
                         hm_stream_t stream)
{
    if (!vel || !accel || !pos || !tag || !gamma_by_type || (nforces && !forces) || nforces > 4
        || ntypes == 0)
        return fail(HM_ERR_ARGS, "hm_langevin_step_two: invalid arguments");
    StepTwoArgs a;
    memset(&a, 0, sizeof(a));
    a.N = N;
    a.vel = (double4*)vel;
    a.accel = (d3*)accel;
    for (unsigned int q = 0; q < nforces; q++)
        a.f[q] = (const float4*)forces[q];
    a.nforces = nforces;
    a.dt = dt;
    a.gamma_by_type = gamma_by_type;
    a.pos = (const double4*)pos;
    a.tag = tag;
    a.kT = kT;
    a.seed = seed;
    a.timestep = timestep;
    return check(launch_step_two(a, (hipStream_t)stream, nullptr), "hm_langevin_step_two");
}

int hm_thermo(unsigned int N, const double* vel, const float* const* forces,
              const float* const* virials, unsigned int nforces, double* out, hm_stream_t stream)
{
    if (!vel || !out || nforces > 4 || (nforces && !forces))
        return fail(HM_ERR_ARGS, "hm_thermo: invalid arguments");
    hipStream_t s = (hipStream_t)stream;
    double* partial = nullptr;
    if (hipMallocAsync((void**)&partial, sizeof(double) * THERMO_SUMS * THERMO_BLOCKS, s)
        != hipSuccess)
        return fail(HM_ERR_ALLOC, "hm_thermo: scratch allocation failed");
    int r = check(launch_thermo(N, (const double4*)vel, (const float4* const*)forces, virials,
                                nforces, partial, out, s),
                  "hm_thermo");
    (void)hipFreeAsync(partial, s);
    return r;
}

int hm_thermo_noalloc(unsigned int N, const double* vel, const float* const* forces,
                      const float* const* virials, unsigned int nforces, double* out,
                      hm_stream_t stream)
{
    if (!vel || !out || nforces > 4 || (nforces && !forces))
        return fail(HM_ERR_ARGS, "hm_thermo_noalloc: invalid arguments");
    // the partial-sum grid is THERMO_BLOCKS workgroups whatever N is, and every workgroup writes
    // its whole row, so the scratch needs no clearing and no N outgrows it
    double* partial = nullptr;
    int r = scratch_for_current_device(&partial);
    if (r != HM_OK)
        return r;
    return check(launch_thermo(N, (const double4*)vel, (const float4* const*)forces, virials,
                               nforces, partial, out, (hipStream_t)stream),
                 "hm_thermo_noalloc");
}

double hm_bussi_alpha_host(double K, double ndof, double dt, double tau, double kT,
                           uint64_t timestep, uint16_t seed, uint32_t instance_tag)
{
    return bussi_alpha(K, ndof, dt, tau, kT, timestep, seed, instance_tag);
}

int hm_thermalize(unsigned int N, double* vel, const unsigned int* tag, double kT,
                  uint64_t timestep, uint16_t seed, hm_stream_t stream)
{
    if (!vel || !tag)
        return fail(HM_ERR_ARGS, "hm_thermalize: invalid arguments");
    hipStream_t s = (hipStream_t)stream;
    double* partial = nullptr;
    if (hipMallocAsync((void**)&partial, sizeof(double) * 3 * THERMO_BLOCKS, s) != hipSuccess)
        return fail(HM_ERR_ALLOC, "hm_thermalize: scratch allocation failed");
    int r = check(launch_thermalize(N, (double4*)vel, tag, kT, timestep, seed, partial, s),
                  "hm_thermalize");
    (void)hipFreeAsync(partial, s);
    return r;
}

int hm_sfc_order(unsigned int N, const double* pos, hm_box box, unsigned int grid,
                 unsigned int* order, hm_stream_t stream)
{
    if (!pos || !order || grid == 0 || grid > 1024 || (grid & (grid - 1)))
        return fail(HM_ERR_ARGS, "hm_sfc_order: invalid arguments");
    hipStream_t s = (hipStream_t)stream;
    unsigned long long *ka = nullptr, *kb = nullptr;
    void* tmp = nullptr;
    size_t tmp_bytes = 0;
    int end_bit = 32 + 3 * ilog2(grid);
    (void)sort_keys64(nullptr, nullptr, N, end_bit, nullptr, &tmp_bytes, s);
    if (hipMallocAsync((void**)&ka, sizeof(unsigned long long) * (N ? N : 1), s) != hipSuccess
        || hipMallocAsync((void**)&kb, sizeof(unsigned long long) * (N ? N : 1), s) != hipSuccess
        || hipMallocAsync(&tmp, tmp_bytes ? tmp_bytes : 16, s) != hipSuccess)
        return fail(HM_ERR_ALLOC, "hm_sfc_order: scratch allocation failed");
    int r = check(launch_sfc_keys(N, (const double4*)pos, make_devbox(box), grid, ka, s),
                  "hm_sfc_order: keys");
    if (!r)
        r = check(sort_keys64(ka, kb, N, end_bit, tmp, &tmp_bytes, s), "hm_sfc_order: sort");
    if (!r)
        r = check(launch_keys_to_order(N, kb, order, s), "hm_sfc_order: order");
    (void)hipFreeAsync(ka, s);
    (void)hipFreeAsync(kb, s);
    (void)hipFreeAsync(tmp, s);
    return r;
}

int hm_sfc_traversal(unsigned int grid, unsigned int* traversal, hm_stream_t stream)
{
    if (!traversal || grid == 0 || grid > 512 || (grid & (grid - 1)))
        return fail(HM_ERR_ARGS, "hm_sfc_traversal: invalid arguments");
    return check(launch_traversal(grid, traversal, (hipStream_t)stream), "hm_sfc_traversal");
}

} // extern "C"
