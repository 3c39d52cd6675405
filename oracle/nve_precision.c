/*
 * nve_precision.c -- NVE energy-drift precision study (TEST INFRASTRUCTURE ONLY).
 *
 * The plain LJ NVE timestep (velocity Verlet, TwoStepConstantVolume without a thermostat,
 * hoomd/md/TwoStepConstantVolume.cc:7-80,213-260; LJ pair forces, hoomd/md/EvaluatorPairLJ.h;
 * binned list with buffer and distance check, hoomd/md/NeighborList.cc:1238-1418) written once
 * over two floating-point types and instantiated three times:
 *
 *   variant 0  FP64 state, FP64 pairs   the reference's default build (HOOMD_LONGREAL_SIZE 64,
 *                                       hoomd/HOOMDMath.h:72-86; its MD code uses Scalar only)
 *   variant 1  FP64 state, FP32 pairs   the HIP engine's split: positions, velocities, wrapping
 *                                       and the update in FP64, pair arithmetic and the
 *                                       per-particle force sum in FP32
 *   variant 2  FP32 state, FP32 pairs   the reference's all-FP32 build (HOOMD_LONGREAL_SIZE 32,
 *                                       hoomd/HOOMDMath.h:57-71)
 *
 * Cubic periodic box, unit masses, eps = sigma = 1, unshifted forces. Energies are measured in
 * FP64 from each variant's state, with the pair energy shifted to zero at r_cut: the truncated
 * force conserves exactly that energy (shifting changes no force), so its drift is the
 * integration error alone.
 *
 * Pinning: variant 0 reproduces the oracle's System (hoomd_oracle.c os_run, NVE) energies step
 * by step (tests/test_oracle_nve_precision.py); variants 1 and 2 differ from it only in the
 * declared types.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "nve_precision.h"

typedef struct
{
    size_t N;
    double L, rcut, rbuff, rlist, dt;
    unsigned int dim, nmax;
    int threads;
    unsigned int *cell_of, *cell_start, *cell_fill, *cell_members;
    unsigned int *nlist, *n_neigh;
    long long n_builds;
} nvp_state;

#define PX double
#define PF double
#define VN(x) x##_dd
#include "nve_precision_body.h"
#undef PX
#undef PF
#undef VN

#define PX double
#define PF float
#define VN(x) x##_df
#include "nve_precision_body.h"
#undef PX
#undef PF
#undef VN

#define PX float
#define PF float
#define VN(x) x##_ff
#include "nve_precision_body.h"
#undef PX
#undef PF
#undef VN

long long nvp_run(int variant, size_t N, double L, const double* pos3, const double* vel3,
                  double r_cut, double r_buff, double dt, uint64_t nsteps, uint64_t every,
                  int threads, double* ke_out, double* pe_out, double* pos_out, double* vel_out)
{
    if (variant < 0 || variant > 2 || N < 2 || !(L > 0) || every == 0 || nsteps % every
        || threads < 1 || !pos3 || !vel3 || !ke_out || !pe_out)
        return -1;
    nvp_state s;
    memset(&s, 0, sizeof(s));
    s.N = N;
    s.L = L;
    s.rcut = r_cut;
    s.rbuff = r_buff;
    s.rlist = r_cut + r_buff;
    s.dt = dt;
    s.threads = threads;
    s.dim = (unsigned int)floor(L / s.rlist);
    if (s.dim < 3)
        return -1;
    size_t nc = (size_t)s.dim * s.dim * s.dim;
    s.nmax = 128;
    s.cell_of = (unsigned int*)malloc(sizeof(unsigned int) * N);
    s.cell_start = (unsigned int*)malloc(sizeof(unsigned int) * (nc + 1));
    s.cell_fill = (unsigned int*)malloc(sizeof(unsigned int) * nc);
    s.cell_members = (unsigned int*)malloc(sizeof(unsigned int) * N);
    s.n_neigh = (unsigned int*)malloc(sizeof(unsigned int) * N);
    s.nlist = (unsigned int*)malloc(sizeof(unsigned int) * N * s.nmax);
    long long rc = -1;
    if (s.cell_of && s.cell_start && s.cell_fill && s.cell_members && s.n_neigh && s.nlist)
    {
        int r = variant == 0   ? run_dd(&s, pos3, vel3, nsteps, every, ke_out, pe_out, pos_out,
                                        vel_out)
                : variant == 1 ? run_df(&s, pos3, vel3, nsteps, every, ke_out, pe_out, pos_out,
                                        vel_out)
                               : run_ff(&s, pos3, vel3, nsteps, every, ke_out, pe_out, pos_out,
                                        vel_out);
        if (r == 0)
            rc = s.n_builds;
    }
    free(s.cell_of);
    free(s.cell_start);
    free(s.cell_fill);
    free(s.cell_members);
    free(s.n_neigh);
    free(s.nlist);
    return rc;
}
