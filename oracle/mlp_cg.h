/*
 * mlp_cg.h -- CPU oracle of the learned coarse-grained site-energy potential (TEST
 * INFRASTRUCTURE ONLY). The model and what pins each function are described in mlp_cg.c.
 */
#ifndef MLP_CG_ORACLE_H
#define MLP_CG_ORACLE_H

#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

#ifdef __cplusplus
extern "C" {
#endif

#define OCG_NK 16 /* radial-basis features */
#define OCG_H 64  /* hidden width */
/* flat parameter vector: W1[64][16], b1[64], W2[64][64], b2[64], w3[64], b3 */
#define OCG_NPARAM (OCG_H * OCG_NK + OCG_H + OCG_H * OCG_H + OCG_H + OCG_H + 1)

typedef struct
{
    double rcut, width, mu0, dmu;
    double gmean[OCG_NK], gscale[OCG_NK];
    int bf16; /* 1: round matrix-product operands to bfloat16 as the GPU kernels do; 2: also
                 round dE/dG where the force pass reads it (the GPU's bf16 copy) */
} ocg_model;

void ocg_basis(const ocg_model* m, double r, double* phi, double* dphi);
void ocg_site(const ocg_model* m, const double* params, const double* x, double* E, double* dEdx);
void ocg_descriptor(const ocg_model* m, const double* L, const double* pos, size_t N,
                    const unsigned int* nl, const unsigned int* nn, unsigned int nmax, double* G);
/* forces F[3N], site energies Esite[N] (or NULL), per-site virial [6N] rows of pitch N (or
 * NULL) and dE/dG A[16N] (or NULL); returns the total energy */
double ocg_forces(const ocg_model* m, const double* params, const double* L, const double* pos,
                  size_t N, const unsigned int* nl, const unsigned int* nn, unsigned int nmax,
                  double* F, double* Esite, double* virial, double* A);
/* force-matching loss against T[3N] and its gradient grad[OCG_NPARAM]; F_out[3N] or NULL */
double ocg_loss_grad(const ocg_model* m, const double* params, const double* L, const double* pos,
                     size_t N, const unsigned int* nl, const unsigned int* nn, unsigned int nmax,
                     const double* T, double* grad, double* F_out);
void ocg_adam(double* params, double* m1, double* m2, const double* grad, size_t n, double lr,
              double beta1, double beta2, double eps, unsigned int t);
double ocg_bf16_round(double v);

#ifdef __cplusplus
}
#endif
#endif
