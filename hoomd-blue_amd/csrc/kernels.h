// kernels.h -- launch interfaces of the gfx950 kernels (used by the C-ABI drivers and the engine).
#pragma once

#include "hm_device.h"

namespace hm
{

struct d3
{
    double x, y, z;
};
struct i3
{
    int x, y, z;
};

// Device-resident step state: every per-step decision lives here, so a whole chunk of steps can
// be captured once in a hipGraph and replayed.
struct StepState
{
    unsigned long long next;        // step that the next step-begin kernel starts
    unsigned long long timestep;    // current step t (set by the step-begin kernel)
    unsigned long long rebuild_at;  // neighbour list must be rebuilt before forces of this step
    unsigned long long n_builds;    // rebuild counter
    double alpha;                   // Bussi translational factor of the current step
    unsigned int error;             // sticky error bits (see ERR_*)
    unsigned int pad;
    // rebuild policy and statistics (NeighborList::needsUpdating, NeighborList.cc:1325-1418)
    unsigned long long last_update; // step of the last build
    unsigned long long forced_at;   // step of a forced build (sort, overflow recovery)
    unsigned long long shortest;    // shortest period between distance-check builds this run
    unsigned long long n_dangerous; // builds on the first check after a build with delay > 1
};

enum : unsigned int
{
    ERR_CELL_OVERFLOW = 1u,
    ERR_NLIST_OVERFLOW = 2u,
    ERR_NAN = 4u,
    ERR_OUTSIDE = 8u,
    ERR_BOND = 16u,
    ERR_ZERO_KE = 32u,
};

// device-side conditions words (atomicMax'd)
struct Conditions
{
    unsigned int cell_need;   // required cmax
    unsigned int nan_idx;     // particle + 1
    unsigned int out_idx;     // particle + 1
    unsigned int nlist_need;  // required nmax
    unsigned int bond_flag;
    unsigned int pad[3];
};

struct PairLaunch
{
    unsigned int tpp;
    unsigned int block;
    unsigned int batch = 8;   // row entries per lane per gather batch (relative kernel: 4 or 8)
};

// Cell-slot relative positions. At every cell-list build the cell sort gives particle p the
// slot c * cmax + k (cell c, k-th particle of the cell in index order) and writes
//   rel[slot]   = position - origin of cell c (minimum image, FP32; packed 12-byte records, so a
//                 cell's particles share as few cache lines as possible),
//   relw[slot]  = pcell[p] = cx | cy << 8 | cz << 16 | type << 24,
//   slot_of[p]  = slot.
// Step one refreshes rel[slot_of[p]] from the new position every step; the cell (and slot) of a
// particle stay those of the last build. A neighbour cell's particles are therefore contiguous
// in rel. Needs 3 <= dim <= 255 per axis and ncell * cmax <= OFFSET_CODE_MAX_N.
struct RelFrame
{
    float3* rel;              // [ncell * cmax], slot order
    uint32_t* relw;           // [ncell * cmax], slot order
    const uint32_t* pcell;    // [N]
    const uint32_t* slot_of;  // [N]
    unsigned int dim[3];
    double w[3];              // cell widths (FP64)
    float wf[3];              // cell widths rounded to FP32
};

// Neighbour-row entries with offset codes: bits 0-25 particle index, bits 26-27 / 28-29 / 30-31
// the x / y / z offset (home cell - neighbour cell) in {-1, 0, 1} as 2-bit two's complement.
constexpr unsigned int OFFSET_CODE_SHIFT = 26;
constexpr unsigned int OFFSET_CODE_MASK = (1u << OFFSET_CODE_SHIFT) - 1u;
constexpr unsigned int OFFSET_CODE_MAX_N = 1u << OFFSET_CODE_SHIFT;
// entries allocated after the last neighbour row, so batched row reads need no clamp
constexpr unsigned int NLIST_PAD = 256;

// Rows with offset codes are stored in blocks of ROW_BLOCK = 32 entries, transposed so that the
// four entries of one 16-byte load are neighbours r, r + 8, r + 16, r + 24 of the block: a lane
// group of 8 reads 32 consecutive entries with one load each, and its q-th entries are 8
// consecutive neighbours. Neighbour p (candidate order) is stored at position row_store_pos(p);
// stored position P holds neighbour row_neighbour(P). Row strides are multiples of ROW_BLOCK.
constexpr unsigned int ROW_BLOCK = 32;
__host__ __device__ inline unsigned int row_store_pos(unsigned int p)
{
    return (p & ~31u) | ((p & 7u) << 2) | ((p >> 3) & 3u);
}
__host__ __device__ inline unsigned int row_neighbour(unsigned int P)
{
    return (P & ~31u) | ((P >> 2) & 7u) | ((P & 3u) << 3);
}

__host__ __device__ inline unsigned int offset_code(int ox, int oy, int oz)
{
    return (((unsigned int)ox & 3u) | (((unsigned int)oy & 3u) << 2)
            | (((unsigned int)oz & 3u) << 4))
           << OFFSET_CODE_SHIFT;
}

__host__ __device__ inline int wrap_offset(int d, int dim)
{
    return d > 1 ? d - dim : (d < -1 ? d + dim : d);
}

__host__ __device__ inline uint32_t pack_cell(unsigned int cx, unsigned int cy, unsigned int cz,
                                              unsigned int type)
{
    return cx | (cy << 8) | (cz << 16) | (type << 24);
}

inline RelFrame make_rel_frame(float3* rel, uint32_t* relw, const uint32_t* pcell,
                               const uint32_t* slot_of, const DevBox& box,
                               const unsigned int dim[3])
{
    RelFrame f;
    f.rel = rel;
    f.relw = relw;
    f.pcell = pcell;
    f.slot_of = slot_of;
    for (int d = 0; d < 3; d++)
    {
        f.dim[d] = dim[d];
        f.w[d] = dim[d] ? box.L[d] / (double)dim[d] : 0.0;
        f.wf[d] = (float)f.w[d];
    }
    return f;
}

// ---------------------------------------------------------------- cell list / nlist
hipError_t launch_cell_adj(const unsigned int dim[3], unsigned int* adj, hipStream_t s);

// Bin every particle. With keys == NULL the particle's entry goes to its (unordered) slot of
// entries; otherwise keys[slot] receives the particle index and home[i] the slot (NO_SLOT when
// the particle was not binned), and launch_cell_rank writes the ordered entries.
constexpr unsigned int NO_SLOT = 0xffffffffu;
hipError_t launch_cell_build(unsigned int N, const double4* pos, const DevBox& box,
                             const unsigned int dim[3], unsigned int cmax,
                             unsigned int* cell_size, hm_cell_entry* entries, d3* last_pos,
                             unsigned int* cond3, const StepState* gate, hipStream_t s,
                             unsigned int* keys = nullptr, unsigned int* home = nullptr);

struct NlistKernelArgs
{
    unsigned int N;
    DevBox box;
    unsigned int dim[3];
    const double4* pos;       // current positions (FP64 re-decision in the guard band)
    float rlistsq_f[HM_MAX_TYPES * HM_MAX_TYPES];  // set by launch_nlist_build
    float guard[HM_MAX_TYPES * HM_MAX_TYPES];      // set by launch_nlist_build
    unsigned int ncell;
    unsigned int cmax;
    const unsigned int* cell_size;
    const hm_cell_entry* entries;
    const unsigned int* cell_adj;
    unsigned int n_adj;
    unsigned int ntypes;
    double rlistsq[HM_MAX_TYPES * HM_MAX_TYPES];  // 0 for disabled pairs -> never included
    unsigned int enabled[HM_MAX_TYPES * HM_MAX_TYPES];
    const unsigned int* n_excl;
    const unsigned int* excl;
    unsigned int max_excl;
    unsigned int nmax;
    unsigned int* nlist;
    unsigned int* n_neigh;
    unsigned short* lrow;     // optional: local rows (position in the home cell's candidates)
    // slot mode: entries are cell slots (c * cmax + k, see RelFrame) carrying the candidate
    // cell's offset from the home cell, wrapped to {-1, 0, 1} per axis, as 2-bit two's
    // complement fields above the slot (OFFSET_CODE_SHIFT); needs 3 <= dim per axis and
    // ncell * cmax <= OFFSET_CODE_MAX_N
    int offset_codes;
    int skip_index_rows;      // write only the local rows (nlist untouched)
    unsigned int* cond;       // needed nmax
    const StepState* gate;
    StepState* counter;       // n_builds and the rebuild statistics updated by block 0
    unsigned int check_delay;
    int dist_check;
    // FP32 staging (the engine with the relative-position pair kernel): the build's cell-slot
    // relative positions, slot -> particle index and packed cell + type. With 27 adjacent cells
    // and the FP32 test, candidates are staged from these (16 bytes, three FP32 adds) instead of
    // from entries (32 bytes, an FP64 minimum image); the guard band widens to match.
    const float3* srel;
    const unsigned int* sidx;
    const uint32_t* swrd;
    float cwf[3];             // cell widths rounded to FP32
    // launch shape, 0 = chosen from the cell occupancy. shape: workgroup form of the FP32 path
    // (NL_SHAPE_*; a form whose home-cell capacity is below cmax falls back to the default);
    // grid: persistent workgroups (rounded up to a multiple of 8)
    int shape;
    unsigned int grid;
};
enum : int
{
    NL_SHAPE_TINY = 1,
    NL_SHAPE_SMALL = 2,
    NL_SHAPE_MID = 3,
    NL_SHAPE_BIG = 4,
};

// Cell-tiled pair kernel: neighbours read from the home cell's LDS-staged candidates through
// the local rows written by the neighbour-list build.
struct PairCellArgs
{
    unsigned int N;
    const double4* pos;
    DevBox box;
    unsigned int dim[3];
    unsigned int ncell;
    unsigned int cmax;
    const unsigned int* cell_size;
    const unsigned int* cell_idx;   // [ncell*cmax] particle index of each slot (index order)
    const unsigned int* cell_adj;
    unsigned int n_adj;
    const unsigned short* lrow;     // [N*nmax]
    unsigned int nmax;
    const unsigned int* n_neigh;
    unsigned int ntypes;
    double rcutsq[HM_MAX_TYPES * HM_MAX_TYPES];
    float rcutsq_f[HM_MAX_TYPES * HM_MAX_TYPES];
    float ronsq_f[HM_MAX_TYPES * HM_MAX_TYPES];
    int shift_mode;
    int compute_virial;
    float4* force;
    float* virial;
};
hipError_t launch_pair_cell_lj(const PairCellArgs& a, const hm_lj_params* params, PairLaunch cfg,
                               hipStream_t s);
// clear the cell occupancy counters when a rebuild is due (gate) or unconditionally (NULL)
hipError_t launch_cell_clear(unsigned int ncell, unsigned int* cell_size, const StepState* gate,
                             hipStream_t s);
// order each cell's slots by particle index (deterministic candidate order)
hipError_t launch_cell_sort(unsigned int ncell, unsigned int cmax, const unsigned int* cell_size,
                            hm_cell_entry* entries, unsigned int* cell_idx, const StepState* gate,
                            hipStream_t s, const DevBox* box = nullptr,
                            const RelFrame* rel = nullptr);
// order each cell's slots by particle index from the key table of launch_cell_build: one thread
// per particle, writes the ordered entries, index table and relative positions
hipError_t launch_cell_rank(unsigned int N, unsigned int cmax, const double4* pos,
                            const unsigned int* cell_size, const unsigned int* keys,
                            const unsigned int* home, hm_cell_entry* entries,
                            unsigned int* cell_idx, const StepState* gate, hipStream_t s,
                            const DevBox& box, const RelFrame& rel);
// launch_cell_build (key table form) and launch_cell_rank in one single-workgroup launch, for
// systems of at most CELL_FUSED_MAX_N particles
constexpr unsigned int CELL_FUSED_MAX_N = 4096;
hipError_t launch_cell_bin_rank(unsigned int N, const unsigned int dim[3], unsigned int cmax,
                                const double4* pos, unsigned int* cell_size, d3* last_pos,
                                unsigned int* cond3, unsigned int* keys, unsigned int* home,
                                hm_cell_entry* entries, unsigned int* cell_idx,
                                const StepState* gate, hipStream_t s, const DevBox& box,
                                const RelFrame& rel);
hipError_t launch_nlist_build(const NlistKernelArgs& a, hipStream_t s);

hipError_t launch_nlist_check(unsigned int N, const double4* pos, const d3* last,
                              const DevBox& box, double maxsq, unsigned int* flag, hipStream_t s);

// ---------------------------------------------------------------- pair
struct PairKernelArgs
{
    unsigned int N;
    const double4* pos;
    DevBox box;
    const unsigned int* nlist;
    unsigned int nmax;
    const unsigned int* n_neigh;
    unsigned int ntypes;
    double rcutsq[HM_MAX_TYPES * HM_MAX_TYPES];
    double ronsq[HM_MAX_TYPES * HM_MAX_TYPES];
    float rcutsq_f[HM_MAX_TYPES * HM_MAX_TYPES];
    float ronsq_f[HM_MAX_TYPES * HM_MAX_TYPES];
    int shift_mode;
    int compute_virial;
    float4* force;
    float* virial;
};
hipError_t launch_pair_lj(const PairKernelArgs& a, const hm_lj_params* params,
                          PairLaunch cfg, hipStream_t s);
// The same kernel with the Table evaluator (mode none)
hipError_t launch_pair_table(const PairKernelArgs& a, const hm_table_params* params,
                             PairLaunch cfg, hipStream_t s);

// Pair kernel reading one 16-byte relative position per neighbour (see RelFrame).
struct PairRelArgs
{
    unsigned int N;
    const double4* pos;      // FP64 positions for the guard-band re-decision of the cutoff
    const float3* rel;       // slot order (RelFrame)
    const uint32_t* relw;    // slot order (RelFrame): packed cell and type
    const uint32_t* slot_of; // [N]
    const uint32_t* part_of; // [ncell * cmax]: particle of each slot (the cell slot table)
    unsigned int nslot;      // ncell * cmax
    // nlist holds N * nmax entries followed by at least NLIST_PAD readable (zeroed) entries
    DevBox box;
    unsigned int dim[3];
    float wf[3];
    const unsigned int* nlist;
    unsigned int nmax;
    const unsigned int* n_neigh;
    unsigned int ntypes;
    double rcutsq[HM_MAX_TYPES * HM_MAX_TYPES];
    float rcutsq_f[HM_MAX_TYPES * HM_MAX_TYPES];
    float ronsq_f[HM_MAX_TYPES * HM_MAX_TYPES];
    float guard[HM_MAX_TYPES * HM_MAX_TYPES];
    int shift_mode;
    int compute_virial;
    float4* force;
    float* virial;
};
hipError_t launch_pair_rel_lj(const PairRelArgs& a, const hm_lj_params* params, PairLaunch cfg,
                              hipStream_t s);
hipError_t launch_pair_rel_table(const PairRelArgs& a, const hm_table_params* params,
                                 PairLaunch cfg, hipStream_t s);

// Cell-tiled form of the relative kernel: one workgroup per home cell stages the adjacent
// cells' cell-slot relative positions in LDS, shifted into the home cell's frame, and reads
// each particle's neighbours through its 16-bit local row (the position of the neighbour in
// the home cell's adjacency-ordered candidate sequence, as the neighbour-list build writes it).
struct PairTileArgs
{
    unsigned int N;
    const double4* pos;       // FP64 positions for the guard-band re-decision of the cutoff
    const float3* rel;        // slot order (RelFrame)
    const uint32_t* relw;     // slot order (RelFrame): packed cell and type
    const uint32_t* cell_idx; // [ncell * cmax] particle of each slot
    const unsigned int* cell_size;
    const unsigned int* cell_adj;
    unsigned int n_adj;       // 27 (3 <= dim per axis)
    unsigned int ncell;
    unsigned int cmax;
    unsigned int dim[3];
    float wf[3];
    DevBox box;
    const uint16_t* lrow;     // [N * nmax + NLIST_PAD]
    unsigned int nmax;
    const unsigned int* n_neigh;
    unsigned int ntypes;
    double rcutsq[HM_MAX_TYPES * HM_MAX_TYPES];
    float rcutsq_f[HM_MAX_TYPES * HM_MAX_TYPES];
    float ronsq_f[HM_MAX_TYPES * HM_MAX_TYPES];
    float guard[HM_MAX_TYPES * HM_MAX_TYPES];
    int shift_mode;
    int compute_virial;
    float4* force;
    float* virial;
};
// cfg.tpp: lanes per particle (4, 8, 16); cfg.batch: entries per lane per row load (4, 8);
// cfg.block: workgroups of the persistent cell loop (0 = 2048)
hipError_t launch_pair_tile_lj(const PairTileArgs& a, const hm_lj_params* params,
                               PairLaunch cfg, hipStream_t s);

// ---------------------------------------------------------------- bond
struct BondKernelArgs
{
    unsigned int N;
    const double4* pos;
    DevBox box;
    const unsigned int* tag;        // index -> tag
    const unsigned int* rtag;       // tag -> index
    const unsigned int* row;        // [N+1] CSR by tag
    const unsigned int* other_tag;  // partner tag
    const unsigned int* btype;      // bond type (may be NULL)
    unsigned int nbond_types;
    int compute_virial;
    float4* force;
    float* virial;
    unsigned int* flag;
    StepState* st;                  // error bit on stretched bonds (may be NULL)
    // optional partner table in the current index order (launch_bond_index): bond k of particle
    // i is bpart[k * N + i] (NO_SLOT past its bonds), of type bptype[k * N + i]; when set, tag,
    // rtag, row, other_tag and btype are not read
    const unsigned int* bpart;
    const unsigned int* bptype;
    unsigned int max_bonds;         // bonds per particle in bpart (at most 4)
};
hipError_t launch_bond_fene(const BondKernelArgs& a, const hm_fene_params* params, hipStream_t s);
hipError_t launch_bond_harmonic(const BondKernelArgs& a, const hm_harmonic_params* params,
                                hipStream_t s);
// the partner table of BondKernelArgs from the CSR-by-tag bonds; rebuilt whenever the particle
// order changes (sort, restore)
hipError_t launch_bond_index(unsigned int N, const unsigned int* tag, const unsigned int* rtag,
                             const unsigned int* row, const unsigned int* other_tag,
                             const unsigned int* btype, unsigned int max_bonds,
                             unsigned int* bpart, unsigned int* bptype, hipStream_t s);

// ---------------------------------------------------------------- integration
struct StepOneArgs
{
    unsigned int N;
    double4* pos;
    double4* vel;
    const d3* accel;
    i3* image;
    DevBox box;
    double dt;
    double alpha;              // used when st == NULL or !use_state_alpha
    const StepState* st;
    int use_state_alpha;
    // fused distance check (NeighborList::distanceCheck for step t + 1)
    const d3* last;
    double maxsq;
    StepState* flag_state;     // rebuild_at = t + 1 when the rebuild policy asks for a build
    unsigned int check_delay;  // rebuild_check_delay (>= 0)
    int dist_check;            // check_dist
    // fused refresh of the reference-cell relative positions (rel.rel may be NULL)
    RelFrame rel;
    // fused clearing of the cell occupancy counters
    unsigned int* cell_size;
    unsigned int ncell;
    // FP32 copy of the new positions (x, y, z, 0) for the learned-CG pair passes, or NULL
    float4* pos32;
    // when set, this launch begins the step without a step-begin kernel: every workgroup takes
    // the current step from begin->next (which no kernel of the step writes before step two)
    // and one thread stores it in begin->timestep for the kernels that follow
    StepState* begin;
};
hipError_t launch_step_one(const StepOneArgs& a, hipStream_t s);

struct StepTwoArgs
{
    unsigned int N;
    double4* vel;
    d3* accel;
    const float4* f[4];
    unsigned int nforces;
    double dt;
    double alpha;
    // Langevin (gamma_by_type != NULL)
    const double* gamma_by_type;
    const double4* pos;
    const unsigned int* tag;
    double kT;
    uint16_t seed;
    uint64_t timestep;       // used when st == NULL
    const StepState* st;
    // FP64 kinetic energy partial sums per block (may be NULL)
    double* ke_partial;
    // when set (with StepOneArgs::begin), one thread ends the step: finish->next =
    // finish->timestep + 1; no kernel of the step reads next after step one
    StepState* finish;
};
hipError_t launch_step_two(const StepTwoArgs& a, hipStream_t s, unsigned int* nblocks_out);

// step-begin kernel: advances the device timestep; Bussi factor from KE partial sums
hipError_t launch_step_begin(StepState* st, const double* ke_partial, unsigned int npartial,
                             int bussi, double ndof, double dt, double tau, double kT,
                             uint16_t seed, hipStream_t s);
hipError_t launch_set_rebuild(StepState* st, hipStream_t s);  // rebuild before next forces

// initial accelerations a = F/m (IntegratorTwoStep::prepRun -> computeAccelerations)
hipError_t launch_accel_init(unsigned int N, const double4* vel, d3* accel, const float4* const* f,
                             unsigned int nforces, hipStream_t s);

// thermo sums: out[0] KE, out[1] PE, out[2] W, out[3..8] virial tensor; partial scratch of
// THERMO_BLOCKS * THERMO_SUMS doubles of partial sums; THERMO_SUMS doubles of output
constexpr unsigned int THERMO_BLOCKS = 512;
constexpr unsigned int THERMO_SUMS = 15;
hipError_t launch_thermo(unsigned int N, const double4* vel, const float4* const* f,
                         const float* const* vir, unsigned int nforces, double* partial,
                         double* out, hipStream_t s);
hipError_t launch_ke_partial(unsigned int N, const double4* vel, double* partial,
                             unsigned int* nblocks_out, hipStream_t s);

hipError_t launch_thermalize(unsigned int N, double4* vel, const unsigned int* tag, double kT,
                             uint64_t timestep, uint16_t seed, double* scratch3, hipStream_t s);

// ---------------------------------------------------------------- Hilbert sort
hipError_t launch_sfc_keys(unsigned int N, const double4* pos, const DevBox& box,
                           unsigned int grid, unsigned long long* keys, hipStream_t s);
hipError_t sort_keys64(unsigned long long* keys_in, unsigned long long* keys_out, unsigned int N,
                       int end_bit, void* tmp, size_t* tmp_bytes, hipStream_t s);
hipError_t launch_keys_to_order(unsigned int N, const unsigned long long* keys,
                                unsigned int* order, hipStream_t s);
hipError_t launch_traversal(unsigned int grid, unsigned int* traversal, hipStream_t s);

// gather helpers for applying a sort order
hipError_t gather_bytes(const void* src, void* dst, const unsigned int* order, unsigned int N,
                        unsigned int elem_bytes, hipStream_t s);
hipError_t launch_rtag(unsigned int N, const unsigned int* tag, unsigned int* rtag, hipStream_t s);
hipError_t launch_excl_index(unsigned int N, const unsigned int* tag, const unsigned int* rtag,
                             const unsigned int* n_excl_tag, const unsigned int* excl_tag,
                             unsigned int max_excl, unsigned int* n_excl, unsigned int* excl,
                             hipStream_t s);

} // namespace hm
