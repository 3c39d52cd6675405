/*
 * nve_precision.h -- NVE energy-drift precision study (TEST INFRASTRUCTURE ONLY; see
 * nve_precision.c for the three variants and what pins them).
 */
#ifndef NVE_PRECISION_H
#define NVE_PRECISION_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Runs nsteps NVE steps of the LJ system (cubic box L, unit masses, r_cut, list buffer r_buff)
 * in precision variant 0 (FP64), 1 (FP64 state, FP32 pairs) or 2 (FP32). ke_out / pe_out
 * receive nsteps / every + 1 samples (steps 0, every, 2 every, ...): kinetic and r_cut-shifted
 * pair energy in FP64. pos_out / vel_out (or NULL) receive the final state. Returns the number
 * of list builds, or -1 on bad arguments (nsteps not a multiple of every, fewer than 3 cells per
 * side) or failed allocation. */
long long nvp_run(int variant, size_t N, double L, const double* pos3, const double* vel3,
                  double r_cut, double r_buff, double dt, uint64_t nsteps, uint64_t every,
                  int threads, double* ke_out, double* pe_out, double* pos_out, double* vel_out);

#ifdef __cplusplus
}
#endif
#endif
