/*
 * mlp_cg.c -- CPU oracle of the learned coarse-grained site-energy potential (TEST
 * INFRASTRUCTURE ONLY: loaded by tests/, smoke() and bench.py's checks, never by the product).
 *
 * The model has no counterpart in the reference. Its structure is the reference's two-pass
 * many-body form of EAM (hoomd/metal/EAMForceCompute.cc:359-424 sums a per-particle density
 * over neighbours, :428-444 takes the embedding energy and its derivative, :447-560 is the pair
 * pass whose force carries the embedding derivatives of both partners; on the GPU
 * EAMForceGPU.cu:20 and :147), with a 16-component density and an MLP as the embedding
 * function:
 *
 *   phi_k(r)  = exp(-(r - mu_k)^2 / (2 w^2)) f_c(r),  f_c(r) = (cos(pi r / r_c) + 1) / 2, r < r_c
 *   G_k(i)    = sum_{j != i, r_ij < r_c} phi_k(r_ij)                      k = 0..15
 *   x(i)      = (G(i) - gmean) * gscale
 *   E_i       = w3 . silu(W2 silu(W1 x + b1) + b2) + b3                   (16 -> 64 -> 64 -> 1)
 *   a(i)      = dE_i/dG(i) = gscale * W1^T [(W2^T [w3 * silu'(h2)]) * silu'(h1)]
 *   F_i       = -sum_j sum_k (a_k(i) + a_k(j)) phi_k'(r_ij) (r_i - r_j) / r_ij
 *
 * Force matching: L = sum_i |F_i - T_i|^2 / (3 N). With e_i = dL/dF_i and
 *   c_k(i) = -sum_j phi_k'(r_ij) (e_i - e_j) . (r_i - r_j) / r_ij,
 * dL/dtheta = sum_i d/dtheta [ c(i) . a(i) ], evaluated per site as the parameter gradient of
 * the directional derivative of E along gscale * c (the forward tangent through the MLP,
 * then its reverse pass).
 *
 * Everything is FP64. With bf16 = 1 the operands of every matrix product are rounded to
 * bfloat16 (round to nearest even, after rounding to FP32) at the same points as the GPU
 * kernels, which multiply bf16 operands and accumulate in FP32; FP32 rounding of the other
 * intermediate values is not emulated. With bf16 = 2 the force pass (and so the loss) also
 * reads a(i) rounded to bfloat16, as the GPU force pass does when it gathers the bf16 copy
 * of dE/dG; the a(i) returned stay unrounded.
 *
 * Pinning (tests/test_oracle_mlp.py): forces against the finite-difference gradient of the
 * total energy and against torch autograd of the same energy; parameter gradients against
 * finite differences of the loss and against torch double-backward autograd; the descriptor
 * and forces of a two-site system against closed forms; one Adam step against its closed
 * form (the first step moves every parameter by -lr g / (|g| + eps)).
 */
#include "mlp_cg.h"

#include <math.h>
#include <string.h>

static double bf16r(double v)
{
    float f = (float)v;
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u)
        return f;
    u += 0x7fffu + ((u >> 16) & 1u);
    u &= 0xffff0000u;
    memcpy(&f, &u, 4);
    return (double)f;
}

static double rnd(int bf16, double v)
{
    return bf16 ? bf16r(v) : v;
}

static double silu(double h)
{
    return h / (1.0 + exp(-h));
}
static double silu_d(double h)
{
    double s = 1.0 / (1.0 + exp(-h));
    return s * (1.0 + h * (1.0 - s));
}
static double silu_dd(double h)
{
    double s = 1.0 / (1.0 + exp(-h));
    return s * (1.0 - s) * (2.0 + h * (1.0 - 2.0 * s));
}

void ocg_basis(const ocg_model* m, double r, double* phi, double* dphi)
{
    const double w2 = m->width * m->width;
    double fc = 0.0, dfc = 0.0;
    if (r < m->rcut)
    {
        fc = 0.5 * (cos(M_PI * r / m->rcut) + 1.0);
        dfc = -0.5 * M_PI / m->rcut * sin(M_PI * r / m->rcut);
    }
    for (int k = 0; k < OCG_NK; k++)
    {
        const double d = r - (m->mu0 + k * m->dmu);
        const double g = exp(-d * d / (2.0 * w2));
        phi[k] = g * fc;
        dphi[k] = g * (dfc - fc * d / w2);
    }
}

/* parameter offsets in the flat vector */
#define P_W1 0
#define P_B1 (P_W1 + OCG_H * OCG_NK)
#define P_W2 (P_B1 + OCG_H)
#define P_B2 (P_W2 + OCG_H * OCG_H)
#define P_W3 (P_B2 + OCG_H)
#define P_B3 (P_W3 + OCG_H)

typedef struct
{
    double x[OCG_NK];
    double h1[OCG_H], a1[OCG_H], h2[OCG_H], a2[OCG_H];
    double d2[OCG_H], u[OCG_H], d1[OCG_H];
    double E;
    double dEdx[OCG_NK];
} site_work;

/* forward and input gradient of one site; x already normalised */
static void site_eval(const ocg_model* m, const double* p, const double* x_in, site_work* w)
{
    const int b = m->bf16;
    for (int k = 0; k < OCG_NK; k++)
        w->x[k] = x_in[k];
    for (int j = 0; j < OCG_H; j++)
    {
        double s = 0.0;
        for (int k = 0; k < OCG_NK; k++)
            s += rnd(b, p[P_W1 + j * OCG_NK + k]) * rnd(b, w->x[k]);
        w->h1[j] = s + p[P_B1 + j];
        w->a1[j] = silu(w->h1[j]);
    }
    for (int j = 0; j < OCG_H; j++)
    {
        double s = 0.0;
        for (int k = 0; k < OCG_H; k++)
            s += rnd(b, p[P_W2 + j * OCG_H + k]) * rnd(b, w->a1[k]);
        w->h2[j] = s + p[P_B2 + j];
        w->a2[j] = silu(w->h2[j]);
    }
    double E = p[P_B3];
    for (int j = 0; j < OCG_H; j++)
        E += p[P_W3 + j] * w->a2[j];
    w->E = E;
    for (int j = 0; j < OCG_H; j++)
        w->d2[j] = p[P_W3 + j] * silu_d(w->h2[j]);
    for (int k = 0; k < OCG_H; k++)
    {
        double s = 0.0;
        for (int j = 0; j < OCG_H; j++)
            s += rnd(b, p[P_W2 + j * OCG_H + k]) * rnd(b, w->d2[j]);
        w->u[k] = s;
        w->d1[k] = s * silu_d(w->h1[k]);
    }
    for (int k = 0; k < OCG_NK; k++)
    {
        double s = 0.0;
        for (int j = 0; j < OCG_H; j++)
            s += rnd(b, p[P_W1 + j * OCG_NK + k]) * rnd(b, w->d1[j]);
        w->dEdx[k] = s;
    }
}

void ocg_site(const ocg_model* m, const double* params, const double* x, double* E, double* dEdx)
{
    site_work w;
    site_eval(m, params, x, &w);
    *E = w.E;
    for (int k = 0; k < OCG_NK; k++)
        dEdx[k] = w.dEdx[k];
}

static void min_image(const double* L, double* d)
{
    for (int a = 0; a < 3; a++)
        d[a] -= L[a] * rint(d[a] / L[a]);
}

/* pair loop over a full neighbour list (rows of nmax entries); pairs at r >= r_cut skipped */
#define FOR_PAIRS(BODY)                                                                          \
    for (size_t i = 0; i < N; i++)                                                               \
        for (unsigned int q = 0; q < nn[i]; q++)                                                 \
        {                                                                                        \
            const size_t j = nl[i * nmax + q];                                                   \
            double d[3] = {pos[3 * i] - pos[3 * j], pos[3 * i + 1] - pos[3 * j + 1],             \
                           pos[3 * i + 2] - pos[3 * j + 2]};                                     \
            min_image(L, d);                                                                     \
            const double rsq = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];                          \
            if (!(rsq < m->rcut * m->rcut))                                                      \
                continue;                                                                        \
            const double r = sqrt(rsq);                                                          \
            double phi[OCG_NK], dphi[OCG_NK];                                                    \
            ocg_basis(m, r, phi, dphi);                                                          \
            BODY                                                                                 \
        }

void ocg_descriptor(const ocg_model* m, const double* L, const double* pos, size_t N,
                    const unsigned int* nl, const unsigned int* nn, unsigned int nmax, double* G)
{
    memset(G, 0, sizeof(double) * N * OCG_NK);
    FOR_PAIRS({
        (void)dphi;
        for (int k = 0; k < OCG_NK; k++)
            G[i * OCG_NK + k] += phi[k];
    })
}

static void normalise(const ocg_model* m, const double* G, double* x)
{
    for (int k = 0; k < OCG_NK; k++)
        x[k] = (G[k] - m->gmean[k]) * m->gscale[k];
}

double ocg_forces(const ocg_model* m, const double* params, const double* L, const double* pos,
                  size_t N, const unsigned int* nl, const unsigned int* nn, unsigned int nmax,
                  double* F, double* Esite, double* virial, double* A)
{
    double* G = (double*)malloc(sizeof(double) * N * OCG_NK);
    double* a = A ? A : (double*)malloc(sizeof(double) * N * OCG_NK);
    ocg_descriptor(m, L, pos, N, nl, nn, nmax, G);
    double Etot = 0.0;
    for (size_t i = 0; i < N; i++)
    {
        double x[OCG_NK];
        site_work w;
        normalise(m, G + i * OCG_NK, x);
        site_eval(m, params, x, &w);
        if (Esite)
            Esite[i] = w.E;
        Etot += w.E;
        for (int k = 0; k < OCG_NK; k++)
            a[i * OCG_NK + k] = m->gscale[k] * w.dEdx[k];
    }
    memset(F, 0, sizeof(double) * 3 * N);
    if (virial)
        memset(virial, 0, sizeof(double) * 6 * N);
    const int ra = m->bf16 >= 2;
    FOR_PAIRS({
        (void)phi;
        double s = 0.0;
        for (int k = 0; k < OCG_NK; k++)
            s += (rnd(ra, a[i * OCG_NK + k]) + rnd(ra, a[j * OCG_NK + k])) * dphi[k];
        /* f_ij = -s r_hat; the virial of i takes half of each pair, d_a f_b */
        const double fdr = -s / r;
        for (int c = 0; c < 3; c++)
            F[3 * i + c] += fdr * d[c];
        if (virial)
        {
            virial[0 * N + i] += 0.5 * fdr * d[0] * d[0];
            virial[1 * N + i] += 0.5 * fdr * d[0] * d[1];
            virial[2 * N + i] += 0.5 * fdr * d[0] * d[2];
            virial[3 * N + i] += 0.5 * fdr * d[1] * d[1];
            virial[4 * N + i] += 0.5 * fdr * d[1] * d[2];
            virial[5 * N + i] += 0.5 * fdr * d[2] * d[2];
        }
    })
    free(G);
    if (!A)
        free(a);
    return Etot;
}

double ocg_loss_grad(const ocg_model* m, const double* params, const double* L, const double* pos,
                     size_t N, const unsigned int* nl, const unsigned int* nn, unsigned int nmax,
                     const double* T, double* grad, double* F_out)
{
    const int b = m->bf16;
    double* G = (double*)malloc(sizeof(double) * N * OCG_NK);
    double* a = (double*)malloc(sizeof(double) * N * OCG_NK);
    double* F = (double*)malloc(sizeof(double) * N * 3);
    double* e = (double*)malloc(sizeof(double) * N * 3);
    double* c = (double*)calloc(N * OCG_NK, sizeof(double));
    ocg_forces(m, params, L, pos, N, nl, nn, nmax, F, NULL, NULL, a);
    ocg_descriptor(m, L, pos, N, nl, nn, nmax, G);
    double loss = 0.0;
    for (size_t i = 0; i < 3 * N; i++)
    {
        const double r = F[i] - T[i];
        loss += r * r;
        e[i] = 2.0 * r / (3.0 * (double)N);
    }
    loss /= 3.0 * (double)N;
    FOR_PAIRS({
        (void)phi;
        const double de = ((e[3 * i] - e[3 * j]) * d[0] + (e[3 * i + 1] - e[3 * j + 1]) * d[1]
                           + (e[3 * i + 2] - e[3 * j + 2]) * d[2])
                          / r;
        for (int k = 0; k < OCG_NK; k++)
            c[i * OCG_NK + k] -= dphi[k] * de;
    })
    memset(grad, 0, sizeof(double) * OCG_NPARAM);
    for (size_t i = 0; i < N; i++)
    {
        double x[OCG_NK], ct[OCG_NK];
        site_work w;
        normalise(m, G + i * OCG_NK, x);
        site_eval(m, params, x, &w);
        for (int k = 0; k < OCG_NK; k++)
            ct[k] = m->gscale[k] * c[i * OCG_NK + k];
        /* tangent of the forward pass along ct */
        double t1[OCG_H], ta1[OCG_H], t2[OCG_H], ta2[OCG_H];
        for (int j = 0; j < OCG_H; j++)
        {
            double s = 0.0;
            for (int k = 0; k < OCG_NK; k++)
                s += rnd(b, params[P_W1 + j * OCG_NK + k]) * rnd(b, ct[k]);
            t1[j] = s;
            ta1[j] = silu_d(w.h1[j]) * s;
        }
        for (int j = 0; j < OCG_H; j++)
        {
            double s = 0.0;
            for (int k = 0; k < OCG_H; k++)
                s += rnd(b, params[P_W2 + j * OCG_H + k]) * rnd(b, ta1[k]);
            t2[j] = s;
            ta2[j] = silu_d(w.h2[j]) * s;
        }
        /* reverse pass of s = w3 . ta2 through the tangent and the primal */
        double Tt2[OCG_H], Th2[OCG_H], v[OCG_H], Tt1[OCG_H], Th1[OCG_H];
        for (int j = 0; j < OCG_H; j++)
        {
            grad[P_W3 + j] += ta2[j];
            Tt2[j] = w.d2[j];
            Th2[j] = params[P_W3 + j] * silu_dd(w.h2[j]) * t2[j];
            grad[P_B2 + j] += Th2[j];
        }
        for (int j = 0; j < OCG_H; j++)
            for (int k = 0; k < OCG_H; k++)
                grad[P_W2 + j * OCG_H + k] += rnd(b, Tt2[j]) * rnd(b, ta1[k])
                                              + rnd(b, Th2[j]) * rnd(b, w.a1[k]);
        for (int k = 0; k < OCG_H; k++)
        {
            double s = 0.0;
            for (int j = 0; j < OCG_H; j++)
                s += rnd(b, params[P_W2 + j * OCG_H + k]) * rnd(b, Th2[j]);
            v[k] = s;
            Tt1[k] = w.d1[k];
            Th1[k] = w.u[k] * silu_dd(w.h1[k]) * t1[k] + v[k] * silu_d(w.h1[k]);
            grad[P_B1 + k] += Th1[k];
        }
        for (int j = 0; j < OCG_H; j++)
            for (int k = 0; k < OCG_NK; k++)
                grad[P_W1 + j * OCG_NK + k] += rnd(b, Tt1[j]) * rnd(b, ct[k])
                                               + rnd(b, Th1[j]) * rnd(b, x[k]);
    }
    if (F_out)
        memcpy(F_out, F, sizeof(double) * 3 * N);
    free(G);
    free(a);
    free(F);
    free(e);
    free(c);
    return loss;
}

void ocg_adam(double* params, double* m1, double* m2, const double* grad, size_t n, double lr,
              double beta1, double beta2, double eps, unsigned int t)
{
    const double c1 = 1.0 - pow(beta1, (double)t), c2 = 1.0 - pow(beta2, (double)t);
    for (size_t k = 0; k < n; k++)
    {
        m1[k] = beta1 * m1[k] + (1.0 - beta1) * grad[k];
        m2[k] = beta2 * m2[k] + (1.0 - beta2) * grad[k] * grad[k];
        params[k] -= lr * (m1[k] / c1) / (sqrt(m2[k] / c2) + eps);
    }
}

double ocg_bf16_round(double v)
{
    return bf16r(v);
}
