/*
 * hoomd_oracle.h -- CPU oracle for the short-range MD timestep (TEST INFRASTRUCTURE ONLY).
 *
 * This is a plain-C, FP64 restatement of the HOOMD-blue v4.8.2 CPU algorithms on the
 * short-range MD hot path. It exists to CHECK the HIP implementation: only the tests,
 * the smoke entry and the benchmark's CPU-baseline leg may load it. The product library
 * never links, includes or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative to the
 * HOOMD-blue source root). Compile with -ffp-contract=off so that the FP64 distance
 * tests round exactly as written (the reference is compiled without FMA contraction on
 * x86-64 baseline targets).
 *
 * Pinning status (what checks each function; the checks are in tests/test_oracle.py):
 *   ob_* box functions ......... pinned: reference BoxDim.h compiled in oracle/_ref + closed forms
 *   oe_lj / oe_fene ............ pinned: forces_and_energies.json, test_bond.py vectors,
 *                                 reference evaluator headers compiled in oracle/_ref
 *   oc_* cell list ............. pinned: test_cell_list.cc dimension known answers, invariants
 *   on_* neighbour list ........ pinned: true_pair_list.json, O(N^2) brute force
 *   op_pair_* .................. pinned: 3rd-law vs full-list equivalence, finite differences
 *   op_bond_* .................. pinned: test_bond.py vectors, finite differences
 *   oi_* integrators ........... pinned: closed-form single steps, NVE energy conservation,
 *                                 time reversibility
 *   orng_philox ................ pinned: Random123 published known-answer vectors
 *   orng_seed / counter / u01 .. pinned: random_numbers_test.cc known answers
 *   orng_normal / gamma ........ PARITY UNPINNED as streams: only their distributions are
 *                                 checked (moments, and Kolmogorov-Smirnov tests of the
 *                                 thermalize normals, Bussi's K' against Gamma(g/2, kT) and the
 *                                 Langevin uniforms); the reference's RandomNumbers.h needs the
 *                                 absent Random123 headers and its tests hold no normal/gamma
 *                                 vectors; oi_bussi_alpha and oi_thermalize inherit this
 *   os_hilbert_* ............... pinned: permutation/adjacency invariants of the Hilbert curve
 *   nvp_run (nve_precision.c) .. pinned: its FP64 variant reproduces os_run's NVE trajectory
 *                                 (tests/test_oracle_nve_precision.py); the FP32 variants
 *                                 differ from it in the declared types only
 */
#ifndef HOOMD_ORACLE_H
#define HOOMD_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ----------------------------------------------------------------------------- box */
typedef struct
{
    double lo[3];
    double hi[3];
    double L[3];
    double Linv[3];
    int periodic[3];
} ob_box;

void ob_box_init(ob_box* b, double Lx, double Ly, double Lz);
void ob_make_fraction(const ob_box* b, const double v[3], double f[3]);
void ob_min_image(const ob_box* b, double w[3]);
void ob_wrap(const ob_box* b, double w[3], int img[3]);

/* ----------------------------------------------------------------------------- RNG */
void orng_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

typedef struct
{
    uint32_t key[2];
    uint32_t ctr[4];
} orng_gen;

void orng_seed(uint8_t id, uint64_t timestep, uint16_t seed, uint32_t key[2]);
void orng_counter(uint32_t a, uint32_t b, uint32_t c, uint16_t d, uint32_t ctr[4]);
void orng_init(orng_gen* g, uint8_t id, uint64_t timestep, uint16_t seed, uint32_t a,
               uint32_t b, uint32_t c, uint16_t d);
void orng_next(orng_gen* g, uint32_t out[4]);
double orng_u01_u64(uint64_t in);
double orng_uneg11_u64(uint64_t in);
double orng_uniform(orng_gen* g, double a, double b);
double orng_normal(orng_gen* g, double sigma, double mu);
double orng_gamma(orng_gen* g, double alpha, double b);

/* ----------------------------------------------------------------------------- evaluators */
/* returns 1 when evaluated; shift = 1 subtracts V(r_cut) */
int oe_lj(double rsq, double rcutsq, double sigma, double epsilon, int shift, double* force_divr,
          double* energy);
int oe_fene(double rsq, double k, double r0, double sigma, double epsilon, double delta,
            double* force_divr, double* energy);
/* tabulated pair: width points V, F from r_min in steps of (r_cut - r_min) / width */
int oe_table(double rsq, double rcutsq, double rmin, unsigned int width, const double* V,
             const double* F, double* force_divr, double* energy);

/* ----------------------------------------------------------------------------- cell list */
void oc_compute_dims(const ob_box* b, double nominal_width, unsigned int dim[3]);
/* returns number of adjacent cells per cell (n_adj); adj has ncell*n_adj entries */
unsigned int oc_n_adj(const unsigned int dim[3]);
void oc_cell_adj(const unsigned int dim[3], unsigned int* adj);
/* bins particles in index order; returns 0, or -(n+1) for NaN / out-of-box particle n.
 * cell_size[ncell]; cell_idx[ncell*cmax] (slot order = increasing particle index).
 * *overflow receives the largest required cmax when it exceeds cmax (else 0). */
int oc_compute_cell_list(const ob_box* b, const unsigned int dim[3], size_t N, const double* pos3,
                         unsigned int cmax, unsigned int* cell_size, unsigned int* cell_idx,
                         unsigned int* cell_of, unsigned int* overflow);

/* ----------------------------------------------------------------------------- neighbour list */
/* Binned build (NeighborListBinned::buildNlist). r_list = r_cut + r_buff for every type pair
 * with r_cut > 0. full = 1 stores every j != i, full = 0 stores only j > i.
 * types[N] (may be NULL for one type), rcut[ntypes*ntypes].
 * excl: optional per-particle exclusion list (indices), n_excl[N], excl[N*max_excl].
 * nlist[N*nmax], n_neigh[N]. Returns the largest n_neigh seen (caller regrows when > nmax). */
unsigned int on_build_binned(const ob_box* b, size_t N, const double* pos3, const unsigned int* types,
                             unsigned int ntypes, const double* rcut, double r_buff, int full,
                             const unsigned int* n_excl, const unsigned int* excl,
                             unsigned int max_excl, unsigned int nmax, unsigned int* nlist,
                             unsigned int* n_neigh);
/* O(N^2) brute force with the same inclusion rule (for pinning the binned build) */
unsigned int on_build_brute(const ob_box* b, size_t N, const double* pos3, double r_list, int full,
                            unsigned int nmax, unsigned int* nlist, unsigned int* n_neigh);
/* NeighborList::distanceCheck for a fixed box: 1 when any particle moved >= r_buff/2 */
int on_distance_check(const ob_box* b, size_t N, const double* pos3, const double* last3,
                      double rcut_max, double r_buff);

/* ----------------------------------------------------------------------------- forces */
/* PotentialPair::computeForces with an LJ evaluator for one particle type.
 * third_law = 1 expects a half list. force4[N*4] = fx,fy,fz,energy; virial6[6*N] (pitch N).
 * shift_mode: 0 none, 1 shift, 2 xplor (ron). */
void op_pair_lj(const ob_box* b, size_t N, const double* pos3, const unsigned int* nlist,
                const unsigned int* n_neigh, unsigned int nmax, int third_law, double sigma,
                double epsilon, double rcut, double ron, int shift_mode, int compute_virial,
                double* force4, double* virial6);

/* PotentialPair::computeForces with the Table evaluator for one particle type: V, F hold width
 * points from r_min in steps of (r_cut - r_min) / width. Same outputs as op_pair_lj. */
void op_pair_table(const ob_box* b, size_t N, const double* pos3, const unsigned int* nlist,
                   const unsigned int* n_neigh, unsigned int nmax, int third_law, double rcut,
                   double rmin, unsigned int width, const double* V, const double* F,
                   int compute_virial, double* force4, double* virial6);

/* PotentialBond::computeForces with the FENE+WCA evaluator. bonds[2*nb] are indices.
 * Returns 0, or -(b+1) if bond b is stretched beyond r0. */
int op_bond_fene(const ob_box* b, size_t N, const double* pos3, size_t nb, const unsigned int* bonds,
                 double k, double r0, double sigma, double epsilon, double delta,
                 int compute_virial, double* force4, double* virial6);

/* ----------------------------------------------------------------------------- integrators */
/* TwoStepConstantVolume::integrateStepOne with a rescale factor alpha (1 for NVE) */
void oi_nvt_step_one(const ob_box* b, size_t N, double* pos3, double* vel3, const double* accel3,
                     int* image3, double dt, double alpha);
/* TwoStepConstantVolume::integrateStepTwo: a = F/m, v = v*alpha + a dt/2 */
void oi_nvt_step_two(size_t N, double* vel3, double* accel3, const double* mass,
                     const double* net_force4, double dt, double alpha);
/* TwoStepLangevin::integrateStepTwo */
void oi_langevin_step_two(size_t N, double* vel3, double* accel3, const double* mass,
                          const unsigned int* tag, const double* net_force4, const double* gamma,
                          const unsigned int* types, double kT, double dt, uint64_t timestep,
                          uint16_t seed);
/* ComputeThermo translational KE (double sum, index order) */
double oi_kinetic_energy(size_t N, const double* vel3, const double* mass);
/* BussiThermostat::compute_rescale_factor for the translational DOF */
double oi_bussi_alpha(double K, double ndof, double dt, double tau, double kT, uint64_t timestep,
                      uint16_t seed, uint32_t instance_tag);
/* ParticleGroup::thermalizeParticleMomenta for the all group (3D, no rigid bodies) */
void oi_thermalize(size_t N, double* vel3, const double* mass, const unsigned int* tag, double kT,
                   uint64_t timestep, uint16_t seed);

/* ----------------------------------------------------------------------------- Hilbert sort */
/* SFCPackTuner::generateTraversalOrder: traversal[bin] = rank, grid^3 entries */
void os_hilbert_traversal(unsigned int grid, unsigned int* traversal);
/* SFCPackTuner::getSortedOrder3D: order[new] = old index */
void os_sort_order(const ob_box* b, size_t N, const double* pos3, unsigned int grid,
                   const unsigned int* traversal, unsigned int* order);

/* ----------------------------------------------------------------------------- full system */
/* A complete single-type LJ system stepped by the restated CPU path (half nlist, 3rd law),
 * used as the trajectory oracle and as the timed CPU baseline. */
typedef struct os_system os_system;

os_system* os_create(size_t N, double Lx, double Ly, double Lz, const double* pos3,
                     const double* vel3, const double* mass, double sigma, double epsilon,
                     double rcut, int shift_mode, double r_buff, double dt, uint16_t seed);
void os_destroy(os_system* s);
/* thermostat: 0 none (NVE), 1 Bussi (kT, tau), 2 Langevin (kT, gamma) */
void os_set_thermostat(os_system* s, int kind, double kT, double tau_or_gamma);
void os_set_sorter(os_system* s, unsigned int period, unsigned int grid);
void os_set_bonds(os_system* s, size_t nb, const unsigned int* bond_tags, double k, double r0,
                  double sigma, double epsilon, double delta, int exclude_bonds);
/* returns 0 or a negative error code */
int os_run(os_system* s, uint64_t nsteps);
/* state by tag: pos3, vel3, image3 (each may be NULL) */
void os_get_state(const os_system* s, double* pos3, double* vel3, int* image3);
/* per-tag forces of the last evaluation: pair (force4, virial6) and bond (may be NULL) */
void os_get_forces(const os_system* s, double* pair_force4, double* pair_virial6,
                   double* bond_force4);
uint64_t os_timestep(const os_system* s);
uint64_t os_n_builds(const os_system* s);
/* NeighborList rebuild_check_delay / check_dist (NeighborList.cc:1313-1418); defaults 1, 1 */
/* Host threads of the step loop (default 1; the tests keep the serial summation order). With
 * n > 1 the list build, pair forces (per-thread third-law buffers) and integrator loops run on
 * n OpenMP threads: the benchmark's CPU baseline at P = host cores. */
void oracle_set_threads(int n);
int oracle_threads(void);
void os_set_nlist_policy(os_system* s, unsigned int rebuild_check_delay, int check_dist);
uint64_t os_n_dangerous(const os_system* s);
unsigned int os_shortest_rebuild(const os_system* s);
double os_kinetic_energy(const os_system* s);
double os_potential_energy(const os_system* s);

#ifdef __cplusplus
}
#endif
#endif
