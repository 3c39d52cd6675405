/*
 * hoomd_amd.h -- C ABI of the MI355X (gfx950) short-range MD timestep library.
 *
 * Two layers share this header:
 *
 *  1. Kernel drivers (hm_*): one call per GPU pass of the reference's hot path. Each replaces a
 *     HOOMD-blue "driver function" (the C-style layer between the *GPU C++ classes and the
 *     kernels, ARCHITECTURE.md:198-208) and takes plain device pointers, sizes and a stream.
 *     They are asynchronous; data errors (overflow, NaN, out-of-box, stretched bond) come back
 *     through device condition words, as in the reference (CellList.cc:549-607,
 *     NeighborList.cc:1601-1617).
 *
 *  2. The system engine (hm_sys_*): a device-resident particle system that runs whole timesteps
 *     natively (cell list, neighbour list, pair and bond forces, integration, thermostats and
 *     the Hilbert sort) with every rebuild and thermostat decision taken on the device. It is
 *     what the hoomd_amd Python package (the hoomd.md surface) drives.
 *
 * Memory layout in HBM (structure of arrays, index order = current sorted order):
 *   pos      double[4N]  x, y, z, type (type stored as an exact double)
 *   vel      double[4N]  vx, vy, vz, mass
 *   accel    double[3N]
 *   image    int[3N]
 *   tag      uint32[N], rtag uint32[N]
 *   force    float[4N]   fx, fy, fz, energy (per force compute, FP32 accumulation)
 *   virial   float[6N]   xx, xy, xz, yy, yz, zz rows of pitch N
 *   nlist    uint32[N*nmax] row-major (row i holds the neighbours of particle i), n_neigh uint32[N]
 *   cell     cell_size uint32[ncell]; cell entries hm_cell_entry[ncell*cmax]
 *
 * Return codes: HM_OK (0) or a negative HM_ERR_* value. No call throws.
 * Boxes are orthorhombic and periodic in x, y and z (every workload here is cubic).
 */
#ifndef HOOMD_AMD_H
#define HOOMD_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HM_OK 0
#define HM_ERR_LAUNCH (-1)   /* a kernel launch or HIP runtime call failed */
#define HM_ERR_ARGS (-2)     /* invalid arguments (sizes, NULL pointers, limits) */
#define HM_ERR_ALLOC (-3)    /* device allocation failed */
#define HM_ERR_NAN (-4)      /* a particle position is NaN */
#define HM_ERR_OUTSIDE (-5)  /* a particle left the box */
#define HM_ERR_BOND (-6)     /* a FENE bond is stretched beyond r0 */
#define HM_ERR_THERMO (-7)   /* Bussi thermostat with zero kinetic energy */
#define HM_ERR_STATE (-8)    /* engine used in the wrong order */
#define HM_ERR_BOXSMALL (-9) /* r_cut + buffer exceeds half the box */

#define HM_MAX_TYPES 4

typedef void* hm_stream_t; /* a hipStream_t; NULL is the null stream */

/* Orthorhombic periodic box; lo = -L/2, Linv = 1/L are derived as BoxDim::setL does. */
typedef struct hm_box
{
    double L[3];
} hm_box;

/* One cell-list slot (32 bytes). */
typedef struct hm_cell_entry
{
    double x, y, z;
    uint32_t idx;
    uint32_t type;
} hm_cell_entry;

/* ------------------------------------------------------------------------------------------ */
/* Library                                                                                     */
/* ------------------------------------------------------------------------------------------ */

/* Version string of this library. */
const char* hm_version(void);
/* Number of visible GPUs (0 when none); does not initialise a context on failure. */
int hm_device_count(void);
/* Text of the last error recorded on this thread. */
const char* hm_last_error(void);
/* Device memory helpers so callers without a HIP runtime binding can stage buffers. */
int hm_malloc(void** ptr, size_t bytes);
int hm_free(void* ptr);
int hm_memcpy_htod(void* dst, const void* src, size_t bytes);
int hm_memcpy_dtoh(void* dst, const void* src, size_t bytes);
int hm_memset(void* dst, int value, size_t bytes);
int hm_memcpy_dtod(void* dst, const void* src, size_t bytes);
int hm_synchronize(void);

/* ------------------------------------------------------------------------------------------ */
/* Kernel drivers                                                                              */
/* ------------------------------------------------------------------------------------------ */

/* Cell grid for a nominal width: dim = max(1, floor(L / width)) per axis.
 * Replaces CellList::computeDimensions (hoomd/CellList.cc:83-131). Host-only. */
int hm_cell_dims(const hm_box* box, double nominal_width, unsigned int dim[3]);
/* Number of unique adjacent cells: prod(min(dim, 3)) (hoomd/CellList.cc:258-275). Host-only. */
unsigned int hm_cell_n_adj(const unsigned int dim[3]);
/* Adjacency table adj[cell*n_adj + k], sorted and unique per cell, built on the device.
 * Replaces CellList::initializeCellAdj (hoomd/CellList.cc:346-402). */
int hm_cell_adj(const unsigned int dim[3], unsigned int* d_adj, hm_stream_t stream);

typedef struct hm_cell_args
{
    unsigned int N;
    const double* pos;          /* [4N] */
    hm_box box;
    unsigned int dim[3];
    unsigned int cmax;          /* slots per cell */
    unsigned int* cell_size;    /* [ncell]; cleared by this call */
    hm_cell_entry* entries;     /* [ncell*cmax] */
    double* last_pos;           /* [3N] or NULL: receives the binned positions */
    unsigned int* conditions;   /* [3] atomicMax'd: needed cmax, NaN particle + 1, outside + 1 */
    unsigned int* cell_idx;     /* [ncell*cmax] or NULL: particle index of each (ordered) slot */
} hm_cell_args;
/* Bin every particle, then order each cell's slots by particle index (cmax <= 256). Replaces
 * gpu_compute_cell_list
 * (hoomd/CellListGPU.cuh, kernel CellListGPU.cu:42-171) driven by CellListGPU::computeCellList
 * (hoomd/CellListGPU.cc:39-232). */
int hm_cell_list_build(const hm_cell_args* args, hm_stream_t stream);

typedef struct hm_nlist_args
{
    unsigned int N;
    const double* pos;          /* [4N] */
    hm_box box;
    unsigned int dim[3];
    unsigned int cmax;
    const unsigned int* cell_size;
    const hm_cell_entry* entries;
    const unsigned int* cell_adj; /* [ncell*n_adj] */
    unsigned int n_adj;
    unsigned int ntypes;        /* <= HM_MAX_TYPES */
    double r_cut[HM_MAX_TYPES * HM_MAX_TYPES]; /* r_cut[ti*ntypes+tj]; <= 0 disables the pair */
    double r_buff;
    const unsigned int* n_excl; /* [N] or NULL: excluded neighbours (indices) per particle */
    const unsigned int* excl;   /* [N*max_excl] */
    unsigned int max_excl;
    unsigned int nmax;          /* row length of nlist (multiple of 8) */
    unsigned int* nlist;        /* [N*nmax] */
    unsigned int* n_neigh;      /* [N] */
    unsigned int* conditions;   /* [1] atomicMax'd: largest n_neigh when it exceeds nmax */
    uint16_t* local_rows;       /* [N*nmax] or NULL: each neighbour's position in its home
                                   cell's candidate sequence (adjacent cells in order, slots in
                                   order), the row format of hm_pair_lj_cells */
} hm_nlist_args;
/* Full neighbour list: for each i every j != i with |r_ij|^2 <= (r_cut + r_buff)^2 in FP64,
 * candidates in adjacency-cell order and particle-index order within a cell.
 * Replaces gpu_compute_nlist_binned (hoomd/md/NeighborListGPUBinned.cuh, kernel
 * NeighborListGPUBinned.cu:52-298) together with gpu_nlist_filter (NeighborListGPU.cu:158-223). */
int hm_nlist_build(const hm_nlist_args* args, hm_stream_t stream);

/* Distance check: flag[0] = 1 when any particle moved by >= ((r_cut_max + r_buff) - r_cut_max)/2
 * since last_pos. Replaces gpu_nlist_needs_update_check_new (hoomd/md/NeighborListGPU.cuh,
 * kernel NeighborListGPU.cu:42-84). flag must be cleared by the caller. */
int hm_nlist_check(unsigned int N, const double* pos, const double* last_pos, hm_box box,
                   double r_cut_max, double r_buff, unsigned int* flag, hm_stream_t stream);

/* LJ parameters for one type pair, built as EvaluatorPairLJ::param_type(sigma, epsilon). */
typedef struct hm_lj_params
{
    double sigma;
    double epsilon;
} hm_lj_params;

typedef struct hm_pair_args
{
    unsigned int N;
    const double* pos;
    hm_box box;
    const unsigned int* nlist;
    unsigned int nmax;
    const unsigned int* n_neigh;
    unsigned int ntypes;
    double r_cut[HM_MAX_TYPES * HM_MAX_TYPES];
    double r_on[HM_MAX_TYPES * HM_MAX_TYPES];
    int shift_mode;             /* 0 none, 1 shift, 2 xplor */
    int compute_virial;
    float* force;               /* [4N] */
    float* virial;              /* [6N] pitch N (unused when compute_virial == 0) */
    unsigned int tpp;           /* lanes per particle: 1, 2, 4, 8 or 16; 0 = tuned default */
    unsigned int block_size;    /* 64..1024 step 64; 0 = tuned default */
} hm_pair_args;
/* Pair force, energy and virial with a full neighbour list: FP64 positions and minimum image,
 * FP64 cutoff test (rsq < r_cut^2), FP32 evaluator arithmetic and FP32 per-particle sums;
 * energy and virial carry the 1/2 of each pair. Replaces
 * gpu_compute_pair_forces<EvaluatorPairLJ> (hoomd/md/PotentialPairGPU.cuh:579-656, kernel
 * :140-404) as instantiated by PotentialPairGPUKernel.cu.inc:19-21. */
int hm_pair_lj(const hm_pair_args* args, const hm_lj_params* params, hm_stream_t stream);

/* Tabulated pair parameters for one type pair, as EvaluatorPairTable::param_type
 * (hoomd/md/EvaluatorPairTable.h:38-118): width points U[k], F[k] (F = -dU/dr) at
 * r_min + k (r_cut - r_min) / width. U and F are DEVICE arrays of width floats. */
typedef struct hm_table_params
{
    double r_min;
    unsigned int width;
    const float* U;
    const float* F;
} hm_table_params;

/* Pair force with the Table evaluator: the same sums and outputs as hm_pair_lj, mode none
 * (shift_mode must be 0; r_on is ignored). Replaces
 * gpu_compute_pair_forces<EvaluatorPairTable> (hoomd/md/PotentialPairGPU.cuh:579-656) as
 * instantiated by PotentialPairGPUKernel.cu.inc:19-21 for PotentialPairTable. */
int hm_pair_table(const hm_pair_args* args, const hm_table_params* params, hm_stream_t stream);

typedef struct hm_pair_cell_args
{
    unsigned int N;
    const double* pos;
    hm_box box;
    unsigned int dim[3];
    unsigned int cmax;
    const unsigned int* cell_size;
    const unsigned int* cell_idx;   /* from hm_cell_list_build */
    const unsigned int* cell_adj;
    unsigned int n_adj;
    const uint16_t* local_rows;     /* from hm_nlist_build */
    unsigned int nmax;
    const unsigned int* n_neigh;
    unsigned int ntypes;
    double r_cut[HM_MAX_TYPES * HM_MAX_TYPES];
    double r_on[HM_MAX_TYPES * HM_MAX_TYPES];
    int shift_mode;
    int compute_virial;
    float* force;
    float* virial;
    unsigned int tpp;               /* 4, 8 or 16 lanes per particle; 0 = 8 */
    unsigned int workgroups;        /* persistent cell loop size; 0 = one workgroup per cell */
} hm_pair_cell_args;
/* The same pair sums as hm_pair_lj, tiled by cell: one workgroup per home cell stages the
 * adjacent cells' current positions in LDS and reads each particle's neighbours through its
 * 16-bit local row. The cell list and local rows must come from the last build; positions may
 * have moved since (the list buffer). */
int hm_pair_lj_cells(const hm_pair_cell_args* args, const hm_lj_params* params,
                     hm_stream_t stream);

typedef struct hm_fene_params
{
    double k, r0, epsilon, sigma, delta;
} hm_fene_params;

/* Harmonic bond parameters, EvaluatorBondHarmonic::param_type (hoomd/md/EvaluatorBondHarmonic.h):
 * U = k (r - r0)^2 / 2. */
typedef struct hm_harmonic_params
{
    double k, r0;
} hm_harmonic_params;

typedef struct hm_bond_args
{
    unsigned int N;
    const double* pos;
    hm_box box;
    const unsigned int* tag;        /* [N] index -> tag */
    const unsigned int* rtag;       /* [N] tag -> index */
    const unsigned int* bond_row;   /* [N+1] CSR offsets by TAG: bonds of the particle with tag t */
    const unsigned int* bond_other; /* partner TAG of each row entry */
    const unsigned int* bond_type;  /* type of each row entry (may be NULL: type 0) */
    unsigned int nbond_types;       /* <= HM_MAX_TYPES */
    int compute_virial;
    float* force;                   /* [4N] */
    float* virial;                  /* [6N] pitch N */
    unsigned int* flag;             /* [1] set to 1 when a bond is stretched beyond r0 */
} hm_bond_args;
/* FENE+WCA bond force, one thread per particle over its bond row (both endpoints hold the
 * bond; energy and virial carry 1/2 per endpoint). Bonds are stored by tag so a particle sort
 * needs no table rebuild. Replaces gpu_compute_bond_forces<EvaluatorBondFENE>
 * (hoomd/md/PotentialBondGPU.cuh:98-239). */
int hm_bond_fene(const hm_bond_args* args, const hm_fene_params* params, hm_stream_t stream);

/* Harmonic bond force with the same per-particle form and outputs as hm_bond_fene (flag is
 * never set). Replaces gpu_compute_bond_forces<EvaluatorBondHarmonic>
 * (hoomd/md/PotentialBondGPU.cuh:98-239, instantiated for md.bond.Harmonic). */
int hm_bond_harmonic(const hm_bond_args* args, const hm_harmonic_params* params,
                     hm_stream_t stream);
/* Velocity-Verlet step one with velocity rescale alpha (1 for NVE): v = (v + a dt/2) * alpha;
 * r += v dt; wrap. Replaces gpu_nvt_rescale_step_one (hoomd/md/TwoStepConstantVolumeGPU.cu:
 * 35-100) and gpu_nve_step_one (hoomd/md/TwoStepNVEGPU.cu:46-126). */
int hm_nvt_step_one(unsigned int N, double* pos, double* vel, const double* accel, int* image,
                    hm_box box, double dt, double alpha, hm_stream_t stream);
/* Step two: a = (sum of forces)/m; v = v*alpha + a dt/2. Up to 4 force arrays are summed
 * (the gpu_integrator_sum_net_force pass, hoomd/Integrator.cu:68, fused). Replaces
 * gpu_nvt_rescale_step_two (TwoStepConstantVolumeGPU.cu:166-213). */
int hm_nvt_step_two(unsigned int N, double* vel, double* accel, const float* const* forces,
                    unsigned int nforces, double dt, double alpha, hm_stream_t stream);
/* Langevin step two with per-particle Philox noise Seed(25, timestep, seed), Counter(tag).
 * Replaces gpu_langevin_step_two (hoomd/md/TwoStepLangevinGPU.cu:51-182). gamma[ntypes]. */
int hm_langevin_step_two(unsigned int N, double* vel, double* accel, const double* pos,
                         const unsigned int* tag, const float* const* forces,
                         unsigned int nforces, const double* gamma_by_type, unsigned int ntypes,
                         double kT, double dt, uint64_t timestep, uint16_t seed,
                         hm_stream_t stream);
/* Thermodynamic sums in FP64 (deterministic order), out[15]: out[0] = KE, out[1] = PE,
 * out[2] = trace of the virial / 3 (W), out[3..8] = virial tensor sums (xx, xy, xz, yy, yz, zz),
 * out[9..14] = kinetic tensor sums m v_a v_b in the same order. Replaces
 * gpu_compute_thermo_partial_sums + final_sums (hoomd/md/ComputeThermoGPU.cu:46,141). forces may
 * hold 0..4 arrays. Allocates its partial-sum scratch on the stream (hipMallocAsync). */
int hm_thermo(unsigned int N, const double* vel, const float* const* forces,
              const float* const* virials, unsigned int nforces, double* out, hm_stream_t stream);
/* hm_thermo without any allocation (no counterpart in the reference; for callers that cannot
 * allocate: code under hipGraph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, results and errors as hm_thermo. The partial sums live in a library array of
 * THERMO_BLOCKS x 15 doubles (60 KiB) per device, held until the process ends; the partial-sum
 * grid is that fixed size for every N, so no N exceeds it. Every call on one device shares that
 * device's array: calls that overlap on one device (different streams or threads) race. */
int hm_thermo_noalloc(unsigned int N, const double* vel, const float* const* forces,
                      const float* const* virials, unsigned int nforces, double* out,
                      hm_stream_t stream);
/* Bussi rescale factor for the translational DOF from the kinetic energy K, drawn on the host
 * thread exactly as BussiThermostat::compute_rescale_factor (hoomd/md/Thermostat.h:357-446). */
double hm_bussi_alpha_host(double K, double ndof, double dt, double tau, double kT,
                           uint64_t timestep, uint16_t seed, uint32_t instance_tag);
/* Maxwell-Boltzmann velocities from Philox Seed(35, timestep, seed), Counter(tag), then the COM
 * momentum removed. Replaces ParticleGroup::thermalizeParticleMomenta (hoomd/ParticleGroup.cc:
 * 690-806). */
int hm_thermalize(unsigned int N, double* vel, const unsigned int* tag, double kT,
                  uint64_t timestep, uint16_t seed, hm_stream_t stream);
/* Hilbert-curve order: order[new] = old index, sorted by (traversal rank of the particle's
 * bin on a grid^3 lattice, index). Replaces gpu_sfc_bin_particles + thrust sort
 * (hoomd/SFCPackTunerGPU.cu:26-146, SFCPackTunerGPU.cc:65-148). grid is a power of two <= 1024. */
int hm_sfc_order(unsigned int N, const double* pos, hm_box box, unsigned int grid,
                 unsigned int* order, hm_stream_t stream);
/* Hilbert traversal rank of every bin of a grid^3 lattice (diagnostic; SFCPackTuner.cc:456-487). */
int hm_sfc_traversal(unsigned int grid, unsigned int* traversal, hm_stream_t stream);

/* ------------------------------------------------------------------------------------------ */
/* Learned coarse-grained site energy (no counterpart in the reference; the two-pass structure  */
/* is the reference's EAM: hoomd/metal/EAMForceGPU.cu:20 density pass, :147 force pass)          */
/* ------------------------------------------------------------------------------------------ */
/* Site i of a coarse-grained system carries the 16 radial-basis densities
 *   G_k(i) = sum_{j != i, r_ij < r_cut} exp(-(r_ij - mu_k)^2 / (2 width^2)) f_c(r_ij),
 *   f_c(r) = (cos(pi r / r_cut) + 1) / 2,  mu_k = mu0 + k dmu,
 * and the site energy E_i = w3 . silu(W2 silu(W1 x_i + b1) + b2) + b3 of x_i = (G(i) - gmean)
 * gscale (16 -> 64 -> 64 -> 1). Forces come from the analytic backward:
 *   F_i = -sum_j sum_k (dE_i/dG_k(i) + dE_j/dG_k(j)) d phi_k(r_ij)/dr (r_i - r_j) / r_ij.
 * The matrix products run on bf16 MFMA with FP32 accumulation; weights are kept in FP32
 * (params) with bf16 operand copies (params_bf16) refreshed by hm_mlp_cg_pack / _adam. */
#define HM_MLP_NK 16
#define HM_MLP_HIDDEN 64
/* flat FP32 parameter vector: W1[64][16], b1[64], W2[64][64] (row = output unit), b2[64],
 * w3[64], b3 */
#define HM_MLP_NPARAM 5313
#define HM_MLP_NBF16 10240     /* bf16 copies: W1, W1^T, W2, W2^T */
#define HM_MLP_GRAD_BLOCKS 256 /* rows of partial gradients written by hm_mlp_cg_backward */

typedef struct hm_mlp_cg_model
{
    double r_cut;                  /* cutoff of the basis (3.0) */
    double width;                  /* Gaussian width (0.2) */
    double mu0, dmu;               /* centres mu0 + k dmu, k < 16 (0, 0.2) */
    double gmean[HM_MLP_NK];       /* input normalisation x = (G - gmean) * gscale */
    double gscale[HM_MLP_NK];
} hm_mlp_cg_model;

typedef struct hm_mlp_cg_args
{
    unsigned int N;
    const double* pos;             /* [4N] */
    hm_box box;
    const unsigned int* nlist;     /* [N*nmax] full list, row i = neighbour indices of i */
    unsigned int nmax;
    const unsigned int* n_neigh;   /* [N] */
    const float* params;           /* [HM_MLP_NPARAM] */
    const uint16_t* params_bf16;   /* [HM_MLP_NBF16] */
    float* G;                      /* [16N] densities */
    float* A;                      /* [16N] dE_i/dG(i) */
    float* Esite;                  /* [N] site energies */
    float* force;                  /* [4N] fx, fy, fz, site energy */
    float* virial;                 /* [6N] rows of pitch N, or NULL */
    /* training (force matching, loss = sum_i |F_i - T_i|^2 / (3N)) */
    const float* target;           /* [4N] target forces (xyz used) */
    float* err;                    /* [4N] dLoss/dF */
    float* C;                      /* [16N] adjoint of the densities (times gscale) */
    float* loss_partial;           /* [hm_mlp_cg_loss_blocks(N)] */
    float* grad_partial;           /* [HM_MLP_GRAD_BLOCKS * HM_MLP_NPARAM] */
    /* optional bf16 copy of A (round to nearest even; 16-byte aligned), or NULL. When set,
     * hm_mlp_cg_forward writes it beside A (A may then be NULL: only the copy is written) and
     * hm_mlp_cg_forces reads it for both partners of every pair (32 instead of 64 gathered
     * bytes per pair). */
    uint16_t* A_bf16;              /* [16N] */
    /* optional FP32 positions (x, y, z, unused; 16-byte aligned), or NULL. When set, the
     * density and force passes gather these (16 instead of 24 bytes per pair) and take the
     * minimum image in FP32 with the box length split into two FP32 parts. */
    const float* pos32;            /* [4N] */
} hm_mlp_cg_args;

/* Entries of loss_partial for N sites. */
unsigned int hm_mlp_cg_loss_blocks(unsigned int N);
/* bf16 operand copies of the weights from the FP32 parameters. */
int hm_mlp_cg_pack(const float* params, uint16_t* params_bf16, hm_stream_t stream);
/* Pass 1: the densities G (EAMForceGPU.cu:20, the density sum of gpu_kernel_1). */
int hm_mlp_cg_descriptor(const hm_mlp_cg_args* args, const hm_mlp_cg_model* model,
                         hm_stream_t stream);
/* Site MLP on bf16 MFMA: E_i into Esite and dE_i/dG(i) into A (the embedding derivative of
 * gpu_kernel_1, EAMForceGPU.cu:20-145). */
int hm_mlp_cg_forward(const hm_mlp_cg_args* args, const hm_mlp_cg_model* model,
                      hm_stream_t stream);
/* Pass 2: forces, site energies and (when virial != NULL) the per-site virial, half of each
 * pair (EAMForceGPU.cu:147, gpu_kernel_2). With train != 0 also err = dLoss/dF against
 * target and the block sums of |F - T|^2 into loss_partial. */
int hm_mlp_cg_forces(const hm_mlp_cg_args* args, const hm_mlp_cg_model* model, int train,
                     hm_stream_t stream);
/* Training backward: the adjoint pass into C, then the parameter gradient of the loss as
 * HM_MLP_GRAD_BLOCKS rows of partial sums in grad_partial (matrix products on bf16 MFMA). */
int hm_mlp_cg_backward(const hm_mlp_cg_args* args, const hm_mlp_cg_model* model,
                       hm_stream_t stream);
/* out[0] = sum_i |F_i - T_i|^2 (FP64) from loss_partial of N sites. */
int hm_mlp_cg_loss(const float* loss_partial, unsigned int N, double* out, hm_stream_t stream);
/* grad = sum of the grad_partial rows (fixed order). */
int hm_mlp_cg_grad(const float* grad_partial, float* grad, hm_stream_t stream);
/* One Adam step t (t >= 1) on params with moments m1, m2 (each [HM_MLP_NPARAM]) from the
 * grad_partial rows; refreshes params_bf16. */
int hm_mlp_cg_adam(float* params, float* m1, float* m2, const float* grad_partial,
                   uint16_t* params_bf16, double lr, double beta1, double beta2, double eps,
                   unsigned int t, hm_stream_t stream);

/* ------------------------------------------------------------------------------------------ */
/* System engine                                                                               */
/* ------------------------------------------------------------------------------------------ */

typedef struct hm_system hm_system;

/* Create a device-resident system from host arrays given in TAG order. type_id may be NULL
 * (all type 0); mass may be NULL (all 1). Positions must lie inside the box. */
hm_system* hm_sys_create(unsigned int N, hm_box box, unsigned int ntypes, const double* pos3,
                         const double* vel3, const unsigned int* type_id, const double* mass,
                         const int* image3, uint16_t seed);
void hm_sys_destroy(hm_system* s);

/* Neighbour list (md.nlist.Cell): buffer, bond exclusions on/off. */
int hm_sys_set_nlist(hm_system* s, double r_buff, int exclude_bonds);
/* Rebuild policy of NeighborList::needsUpdating (hoomd/md/NeighborList.cc:1325-1418): no check
 * until rebuild_check_delay steps after a build; with check_dist the list is rebuilt when a
 * particle moved by r_buff / 2, without it every rebuild_check_delay steps. Defaults 1, 1. */
int hm_sys_set_nlist_policy(hm_system* s, unsigned int rebuild_check_delay, int check_dist);
/* LJ pair force (md.pair.LJ): r_cut/params per type pair (ntypes^2, row-major, symmetric);
 * r_cut <= 0 disables a pair. mode: 0 none, 1 shift, 2 xplor. Returns the force id (>= 0). */
int hm_sys_add_pair_lj(hm_system* s, const double* r_cut, const double* r_on,
                       const hm_lj_params* params, int mode);
/* Tabulated pair force (md.pair.Table, mode none): r_cut per type pair; per type pair q
 * (row-major, symmetric) r_min[q] and width[q] points of U and F, stored one pair after another
 * in the host arrays U and F (pair q starts at the sum of the earlier widths). width 0 or
 * r_cut <= 0 disables a pair. Returns the force id. */
int hm_sys_add_pair_table(hm_system* s, const double* r_cut, const double* r_min,
                          const unsigned int* width, const double* U, const double* F);
/* New U and F values of a Table force, same layout and widths as at hm_sys_add_pair_table
 * (a ramped potential changes its values between runs without rebuilding the system). */
int hm_sys_update_pair_table(hm_system* s, int force_id, const double* U, const double* F);
/* FENE+WCA bonds (md.bond.FENEWCA): bonds by tag [2*nb], bond types [nb] (may be NULL),
 * params per bond type. Returns the force id. */
int hm_sys_add_bond_fene(hm_system* s, unsigned int nb, const unsigned int* bond_tags,
                         const unsigned int* bond_typeid, unsigned int nbond_types,
                         const hm_fene_params* params);
/* Integration method: 0 ConstantVolume (NVE), 1 ConstantVolume + Bussi(kT, tau),
 * 2 Langevin(kT, gamma per type in param[0..ntypes-1]). */
int hm_sys_set_method(hm_system* s, int method, double dt, double kT, const double* param);
/* Hilbert sorter (tune.ParticleSorter): period in steps (0 disables), grid (power of 2). */
int hm_sys_set_sorter(hm_system* s, unsigned int period, unsigned int grid);
/* Request the pressure tensor (virial) on every step (default off; on demand for logging). */
int hm_sys_set_compute_virial(hm_system* s, int on);
/* Launch configuration of the neighbour-index gather pair kernel (0 = keep/tuned). */
int hm_sys_set_pair_launch(hm_system* s, unsigned int tpp, unsigned int block_size);
/* Pair kernel used by the engine: kind 0 = cell-tiled (tpp 4, 8 or 16 lanes per particle,
 * size = workgroups of the persistent cell loop, 0 = one per cell), kind 1 = neighbour-index
 * gather of FP64 positions (tpp, size = block size), kind 2 (default) = neighbour-index gather
 * of FP32 positions relative to the particle's cell at the last build, with the cutoff
 * re-decided in FP64 near r_cut (tpp, size = block size; kind 1 runs when an axis has fewer
 * than 3 or more than 255 cells). */
int hm_sys_set_pair_kernel(hm_system* s, int kind, unsigned int tpp, unsigned int size);
/* Autotuner (replaces Autotuner.h, hoomd/Autotuner.h:503-664): when on (the default, and after
 * hm_sys_set_pair_kernel with tpp = size = 0) the next run first times every launch
 * configuration of the pair kernel in use (5 launches each, median) and keeps the fastest.
 * An explicit tpp / block size turns it off. */
int hm_sys_set_autotune(hm_system* s, int on);
/* Pair launch in use: out[0] = kind actually run, out[1] = tpp, out[2] = block size (kind 1, 2)
 * or workgroups (kind 0), out[3] = 1 when the autotuner has chosen it, out[4] = row entries per
 * lane per gather batch (kind 2: 4 or 8). */
int hm_sys_pair_config(const hm_system* s, unsigned int out[5]);
/* Autotuner samples of the last sweep: rows of (tpp, size, batch, median ms), at most max_rows
 * written; returns the number of rows of the sweep. */
int hm_sys_tune_log(const hm_system* s, float* out, unsigned int max_rows);
/* List-build launch chosen by the autotuner: out[0] = workgroup shape (0 default, 1 tiny,
 * 2 small, 3 mid, 4 big), out[1] = persistent grid (0 default), out[2] = 1 when the choice came
 * from the process's cache of earlier sweeps (same pair kind, forces, types, exclusions, size,
 * density, list range and cell grid). log3 receives rows (shape, grid, median ms) of the last
 * sweep; returns its number of rows. The autotuner sweeps at the first run and once more after
 * 500 steps (on the state the run has reached), unless the cache answered. */
int hm_sys_nlist_launch(const hm_system* s, int out[3], float* log3, unsigned int max_rows);
/* Forget every launch shape the autotuner has cached in this process. */
void hm_tune_cache_clear(void);
/* Steps per captured hipGraph chunk (0 disables graphs; default 100). */
int hm_sys_set_chunk(hm_system* s, unsigned int steps, int use_graph);
/* Advance nsteps timesteps. Returns HM_OK or a negative error (state is left at the last good
 * chunk boundary on data errors). */
int hm_sys_run(hm_system* s, uint64_t nsteps);
/* Capture and instantiate the hipGraphs of every chunk length that hm_sys_run(s, nsteps) from
 * the current step would launch, launching nothing (prepares the system first if needed).
 * Returns the number of graphs added (>= 0) or a negative error. */
int hm_sys_prime(hm_system* s, uint64_t nsteps);
uint64_t hm_sys_timestep(const hm_system* s);
/* Set the current timestep (before the first run, or to continue a trajectory). */
int hm_sys_set_timestep(hm_system* s, uint64_t timestep);
/* Copy out state by tag (each pointer may be NULL): pos3, vel3, image3, accel3. */
int hm_sys_get_state(hm_system* s, double* pos3, double* vel3, int* image3, double* accel3);
/* Replace positions/velocities/images by tag (forces a neighbour-list rebuild). */
int hm_sys_set_state(hm_system* s, const double* pos3, const double* vel3, const int* image3);
/* Per-particle force (fx, fy, fz, energy) [4N] and virial [6N] of one force id, by tag. */
int hm_sys_get_forces(hm_system* s, int force_id, double* force4, double* virial6);
/* Neighbour list of the last build as (tag_i, tag_j) pairs with tag_i < tag_j; returns the
 * number of pairs (<= max_pairs) or a negative error. */
long long hm_sys_get_pairs(hm_system* s, unsigned int* pairs, long long max_pairs);
/* Thermo sums of the current state into out[15]: see hm_thermo. Evaluates the virials first
 * when they are not computed every step. */
int hm_sys_thermo(hm_system* s, double* out15);
/* Statistics, out[9]: out[0] = nlist builds since creation, out[1] = nmax, out[2] = cmax,
 * out[3] = sorts, out[4] = steps run, out[5] = rollbacks after overflow, out[6] = builds in the
 * last run (NeighborList num_builds), out[7] = shortest period between distance-check builds in
 * the last run, 100 when none (getSmallestRebuild), out[8] = dangerous builds in the last run. */
int hm_sys_stats(hm_system* s, uint64_t* out9);
/* Replace the cell capacity (particles per cell, <= 256) and the row capacity (neighbours per
 * particle, rounded up to 32) and rebuild the lists before the forces of the next step. A
 * capacity below the need makes that build overflow: the run restores the chunk's first state,
 * grows the capacity and replays the chunk (the reference regrows and repeats the build,
 * NeighborList.cc:411-427, CellList.cc:183-196). */
int hm_sys_set_list_capacity(hm_system* s, unsigned int cmax, unsigned int nmax);
/* Thermalize velocities of all particles (ParticleGroup::thermalizeParticleMomenta). */
int hm_sys_thermalize(hm_system* s, double kT);
/* Launch stream used by the engine (a hipStream_t). */
hm_stream_t hm_sys_stream(hm_system* s);
/* Harmonic bonds (md.bond.Harmonic): bonds by tag [2*nb], bond types [nb] (may be NULL),
 * params per bond type. Returns the force id. */
int hm_sys_add_bond_harmonic(hm_system* s, unsigned int nb, const unsigned int* bond_tags,
                             const unsigned int* bond_typeid, unsigned int nbond_types,
                             const hm_harmonic_params* params);
/* Learned CG site energy as a force of the system (all type pairs use model->r_cut), from
 * HM_MLP_NPARAM FP32 host parameters. It reads the system's neighbour list, so bonded pairs
 * are left out when the list excludes bonds (hm_sys_set_nlist). Returns the force id. */
int hm_sys_add_mlp_cg(hm_system* s, const hm_mlp_cg_model* model, const float* params);
/* Replace the parameters (host, HM_MLP_NPARAM floats); resets the Adam moments. */
int hm_sys_mlp_set_params(hm_system* s, int force_id, const float* params);
int hm_sys_mlp_get_params(hm_system* s, int force_id, float* params);
/* Replace the normalisation and basis constants (r_cut cannot change). */
int hm_sys_mlp_set_model(hm_system* s, int force_id, const hm_mlp_cg_model* model);
/* A tabulated pair term evaluated in the site energy's force pass over the same pairs, as
 * md.pair.Table would (EvaluatorPairTable.h:152-196: U and F = -dU/dr on width points r_min +
 * k (r_cut - r_min) / width, linear interpolation), so a structural prior costs no pass of its
 * own. width 0 removes it. Same width and r_min: overwritten in place (captured step graphs
 * stay valid). Forces, energies, virials and the training loss include it. */
int hm_sys_mlp_set_pair_table(hm_system* s, int force_id, unsigned int width, double r_min,
                              const double* U, const double* F);
/* Densities G by tag [16N] (host) of the current state, or, when frame_pos is not NULL, of
 * the positions frame_pos (device [N][3] FP32 by tag), which then replace the state's. */
int hm_sys_mlp_descriptors(hm_system* s, int force_id, const float* frame_pos, float* G);

/* Force-matching training on device-resident frames: step k loads frame order[k] (or
 * k mod nframes), builds the list, evaluates the forces, the loss sum |F - T|^2 / (3N) and
 * its gradient, and takes one Adam step (on the mean gradient of frames_per_step frames). */
typedef struct hm_mlp_train_args
{
    unsigned int nframes;
    const float* frame_pos;        /* device [nframes][N][3] FP32 positions by tag */
    const float* frame_force;      /* device [nframes][N][3] FP32 target forces by tag */
    unsigned int nsteps;
    const unsigned int* order;     /* host [nsteps] frame of each step, or NULL */
    double lr, beta1, beta2, eps;
    double* loss;                  /* host [nsteps] loss of each step before its update, or NULL */
    unsigned int frames_per_step;  /* frames whose gradients one Adam step averages (0 or 1:
                                    * one); order then holds nsteps * frames_per_step frames */
} hm_mlp_train_args;
int hm_sys_mlp_train(hm_system* s, int force_id, const hm_mlp_train_args* args);
/* Loss and gradient (host [HM_MLP_NPARAM], may be NULL) of one frame (device pointers as in
 * hm_mlp_train_args), with no update. */
int hm_sys_mlp_loss_grad(hm_system* s, int force_id, const float* frame_pos,
                         const float* frame_force, double* loss, float* grad);

/* Force-matching targets minus a known pair force (a prior): the system's positions become
 * frame_pos (device [N][3] FP32 by tag), the lists are rebuilt, pair force force_id (LJ or
 * Table) is evaluated and subtracted from frame_force (device [N][3] FP32 by tag). */
int hm_sys_frame_subtract_force(hm_system* s, int force_id, const float* frame_pos,
                                float* frame_force);

/* Coarse-grained site map: site I holds the particles with tags tags[row[I]] ..
 * tags[row[I+1]-1] (host arrays, row[nsites+1]). */
int hm_sys_set_site_map(hm_system* s, unsigned int nsites, const unsigned int* row,
                        const unsigned int* tags);
/* Map the current state: d_pos[3*nsites] = mass-weighted site centres wrapped into the box,
 * d_force[3*nsites] (may be NULL) = summed forces of all force computes (device FP32). */
int hm_sys_map_frame(hm_system* s, float* d_pos, float* d_force);

/* Benchmark helper (no counterpart in the reference): average durations in ms of a complete
 * list rebuild at the current positions, iters times, HIP events on the engine stream:
 * avg_ms2[0] = cell binning and ordering, avg_ms2[1] = neighbour-list build. The next step
 * rebuilds again through the regular machinery. */
int hm_sys_time_build(hm_system* s, unsigned int iters, double* avg_ms2);
/* Benchmark helper (no counterpart in the reference): average durations in ms of the learned
 * site energy's three kernels (densities, site MLP on bf16 MFMA, forces) of force force_id over
 * iters launches each, HIP events on the engine stream. */
int hm_sys_time_mlp(hm_system* s, int force_id, unsigned int iters, double* avg_ms3);
/* Force-matching targets (no counterpart in the reference; device arrays): out_force = frame
 * forces minus the harmonic bond forces k (r - r0)^2 / 2 of each site's bonded partners
 * (partner[row[i] .. row[i+1]-1] for site i), frames [nframes][nsites][3] FP32. out_force may
 * be frame_force. */
int hm_frames_sub_harmonic(unsigned int nframes, unsigned int nsites, const float* frame_pos,
                           const float* frame_force, float* out_force, const unsigned int* row,
                           const unsigned int* partner, double k, double r0, hm_box box,
                           hm_stream_t stream);

/* Benchmark helper (no counterpart in the reference): average duration in ms of the pair-force
 * kernel of force force_id over iters back-to-back launches on the engine stream, timed with
 * HIP events on that stream, and the number of neighbour-list entries it reads. */
int hm_sys_time_pair(hm_system* s, int force_id, unsigned int iters, double* avg_ms,
                     uint64_t* nlist_entries);


/* ---------------------------------------------------------------------------------------------
 * Batched replicas: many independent LJ systems of uneven size in one set of device arrays, each
 * phase of the timestep one launch over all of them (no counterpart in the reference, which runs
 * one System per Simulation; each replica follows ConstantVolume + Bussi, the binned full list
 * with distance-check rebuilds and PotentialPair<EvaluatorPairLJ> with mode "none", as the
 * single-system path does: hoomd/md/TwoStepConstantVolume.cc:7-260, Thermostat.h:357-446,
 * NeighborList.cc:1220-1289, NeighborListGPUBinned.cu:52-298, PotentialPair.h:587-794).
 * ------------------------------------------------------------------------------------------- */
typedef struct hm_batch hm_batch;
/* nrep replicas; replica r holds n[r] particles in a cubic periodic box of edge L[r] and is
 * thermostatted at kT[r] with RNG seed seed[r] (16 bits, as Simulation.seed). pos3/vel3/mass
 * hold all replicas' particles concatenated in replica order (vel3, mass may be NULL: zero, 1).
 * LJ epsilon/sigma/r_cut and buffer r_buff are shared; dt and Bussi tau (0: tau = 0) too.
 * Returns NULL on error (hm_last_error; HM_ERR_BOXSMALL when L[r] <= 2 (r_cut + r_buff)). */
hm_batch* hm_batch_create(unsigned int nrep, const unsigned int* n, const double* L,
                          const double* kT, const uint16_t* seed, const double* pos3,
                          const double* vel3, const double* mass, double epsilon, double sigma,
                          double r_cut, double r_buff, double dt, double tau);
void hm_batch_destroy(hm_batch* b);
/* Maxwell-Boltzmann velocities per replica (ParticleGroup::thermalizeParticleMomenta with the
 * replica's tags 0..n-1, seed and kT; COM momentum of each replica removed). */
int hm_batch_thermalize(hm_batch* b, uint64_t timestep);
int hm_batch_set_timestep(hm_batch* b, uint64_t timestep);
/* Advance every replica nsteps steps (chunks captured as hipGraphs). */
int hm_batch_run(hm_batch* b, uint64_t nsteps);
/* Capture the graphs hm_batch_run(b, nsteps) would launch, launching nothing. */
int hm_batch_prime(hm_batch* b, uint64_t nsteps);
/* Steps per captured chunk (default 100). */
int hm_batch_set_chunk(hm_batch* b, unsigned int steps);
/* Re-lay each replica's particles in cell order every period steps (default 100; 0 never),
 * as the reference's particle sorter does with a Hilbert order (SFCPackTuner.cc:119-285). */
int hm_batch_set_sort_period(hm_batch* b, unsigned int period);
/* Autotune the pair launch (tpp x block, median of 5) on the current lists; cached per
 * (particles, density) bucket for the process. log3 receives rows (tpp, block, ms); returns
 * the number of rows of the sweep (0 when the cache answered). */
int hm_batch_tune(hm_batch* b, float* log3, unsigned int max_rows);
/* Fixed pair launch: tpp in {1,2,4,8,16}, block in {64,128,256}. */
int hm_batch_set_pair_launch(hm_batch* b, unsigned int tpp, unsigned int block);
/* out = {pair tpp, pair block, autotuned, cmax, nmax, total cells}. */
int hm_batch_config(const hm_batch* b, unsigned int out[6]);
/* State of all particles in replica order (any pointer may be NULL). */
int hm_batch_get_state(hm_batch* b, double* pos3, double* vel3, int* image3);
/* Pair force (fx, fy, fz, energy) of every particle [4N] for the current state. */
int hm_batch_get_forces(hm_batch* b, double* force4);
/* Per replica r: out4[4r] kinetic energy, [4r+1] pair energy, [4r+2] list builds, [4r+3]
 * neighbour-list entries. */
int hm_batch_replica_stats(hm_batch* b, double* out4);
/* out4 = {particles, steps run, overflow replays, list builds over all replicas}. */
int hm_batch_stats(hm_batch* b, uint64_t out4[4]);
/* Replace the cell and row capacities and rebuild every list at the next step (capacities below
 * the need exercise the overflow replay: the chunk restarts from its first step with grown
 * lists, as the reference regrows and repeats, NeighborList.cc:411-427). */
int hm_batch_set_list_sizes(hm_batch* b, unsigned int cmax, unsigned int nmax);
/* Benchmark helper: average pair-kernel launch duration (ms) over iters launches, HIP events on
 * the batch stream. */
int hm_batch_time_pair(hm_batch* b, unsigned int iters, double* avg_ms);

#ifdef __cplusplus
}
#endif
#endif
