/*
 * hoomd_oracle.c -- CPU oracle for the short-range MD timestep (TEST INFRASTRUCTURE ONLY).
 *
 * FP64 restatement of the HOOMD-blue v4.8.2 CPU path. Only tests/, the smoke entry and the
 * benchmark's cpu_baseline leg load this file's library; the product never does.
 * Build with -ffp-contract=off (see oracle/Makefile).
 */
#include "hoomd_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Host threads of the step loop's list build, pair forces and integrator (1 by default: the
 * tests compare against the serial order of every sum). The reference parallelises its CPU
 * path over MPI ranks with spatial domains (hoomd/Communicator.h:90-125); the benchmark's CPU
 * baseline runs this restatement on P threads instead. With P > 1 the step loop keeps full
 * lists (each pair evaluated from both ends, as the reference's GPU path does) so that rows are
 * independent and the threads share no writes; each thread takes contiguous ranges of the
 * Hilbert-sorted particles. Results differ from the serial run only in summation order. */
static int g_threads = 1;

void oracle_set_threads(int n)
{
    g_threads = n > 1 ? n : 1;
}

int oracle_threads(void)
{
    return g_threads;
}

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* =========================================================================== box
 * hoomd/BoxDim.h:276-283 (makeFraction), :317-389 (minImage, host branch form),
 * :407-475 (wrap). Orthorhombic boxes only (all tilt factors zero), fully periodic. */

void ob_box_init(ob_box* b, double Lx, double Ly, double Lz)
{
    /* BoxDim::setL: hi = L/2, lo = -hi, Linv = 1/L */
    double L[3] = {Lx, Ly, Lz};
    for (int d = 0; d < 3; d++)
    {
        b->L[d] = L[d];
        b->hi[d] = L[d] / 2.0;
        b->lo[d] = -b->hi[d];
        b->Linv[d] = 1.0 / L[d];
        b->periodic[d] = 1;
    }
}

void ob_make_fraction(const ob_box* b, const double v[3], double f[3])
{
    /* BoxDim.h:276-283 with ghost_width = 0 and zero tilt: (v - lo) * Linv */
    for (int d = 0; d < 3; d++)
        f[d] = (v[d] - b->lo[d]) * b->Linv[d];
}

void ob_min_image(const ob_box* b, double w[3])
{
    /* BoxDim.h:343-386 (host branches). z first, then y, then x; tilts are zero. */
    if (b->periodic[2])
    {
        if (w[2] >= b->hi[2])
            w[2] -= b->L[2];
        else if (w[2] < b->lo[2])
            w[2] += b->L[2];
    }
    if (b->periodic[1])
    {
        if (w[1] >= b->hi[1])
        {
            int i = (int)(w[1] * b->Linv[1] + 0.5);
            w[1] -= (double)i * b->L[1];
        }
        else if (w[1] < b->lo[1])
        {
            int i = (int)(-w[1] * b->Linv[1] + 0.5);
            w[1] += (double)i * b->L[1];
        }
    }
    if (b->periodic[0])
    {
        if (w[0] >= b->hi[0])
        {
            int i = (int)(w[0] * b->Linv[0] + 0.5);
            w[0] -= (double)i * b->L[0];
        }
        else if (w[0] < b->lo[0])
        {
            int i = (int)(-w[0] * b->Linv[0] + 0.5);
            w[0] += (double)i * b->L[0];
        }
    }
}

void ob_wrap(const ob_box* b, double w[3], int img[3])
{
    /* BoxDim.h:407-475 with zero tilt: one image shift per axis */
    for (int d = 0; d < 3; d++)
    {
        if (!b->periodic[d])
            continue;
        if (w[d] >= b->hi[d])
        {
            w[d] -= b->L[d];
            img[d]++;
        }
        else if (w[d] < b->lo[d])
        {
            w[d] += b->L[d];
            img[d]--;
        }
    }
}

/* =========================================================================== RNG
 * Philox4x32-10: Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3",
 * SC'11, restated from the paper (round function, multipliers 0xD2511F53 / 0xCD9E8D57,
 * Weyl key increments 0x9E3779B9 / 0xBB67AE85). Random123 itself is an un-vendored submodule
 * of the reference (.gitmodules:10-12); HOOMD's wrapper is hoomd/RandomNumbers.h:136-529. */

static void philox_round(uint32_t c[4], const uint32_t k[2])
{
    uint64_t p0 = (uint64_t)0xD2511F53u * (uint64_t)c[0];
    uint64_t p1 = (uint64_t)0xCD9E8D57u * (uint64_t)c[2];
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c[1] ^ k[0];
    uint32_t n2 = hi0 ^ c[3] ^ k[1];
    c[0] = n0;
    c[1] = lo1;
    c[2] = n2;
    c[3] = lo0;
}

void orng_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
    uint32_t k[2] = {key[0], key[1]};
    for (int r = 0; r < 10; r++)
    {
        if (r > 0)
        {
            k[0] += 0x9E3779B9u;
            k[1] += 0xBB67AE85u;
        }
        philox_round(c, k);
    }
    memcpy(out, c, sizeof(c));
}

void orng_seed(uint8_t id, uint64_t timestep, uint16_t seed, uint32_t key[2])
{
    /* RandomNumbers.h:150-156 */
    key[0] = ((uint32_t)id << 24) | ((uint32_t)seed << 8)
             | (uint32_t)((timestep & 0x000000ff00000000ull) >> 32);
    key[1] = (uint32_t)(timestep & 0x00000000ffffffffull);
}

void orng_counter(uint32_t a, uint32_t b, uint32_t c, uint16_t d, uint32_t ctr[4])
{
    /* RandomNumbers.h:185-188: {d << 16, c, b, a} */
    ctr[0] = (uint32_t)d << 16;
    ctr[1] = c;
    ctr[2] = b;
    ctr[3] = a;
}

void orng_init(orng_gen* g, uint8_t id, uint64_t timestep, uint16_t seed, uint32_t a, uint32_t b,
               uint32_t c, uint16_t d)
{
    orng_seed(id, timestep, seed, g->key);
    orng_counter(a, b, c, d, g->ctr);
}

void orng_next(orng_gen* g, uint32_t out[4])
{
    /* RandomNumbers.h:264-270: evaluate then increment ctr[0] */
    orng_philox4x32_10(g->ctr, g->key, out);
    g->ctr[0] += 1;
}

double orng_u01_u64(uint64_t in)
{
    /* Random123 u01<double,uint64_t>: in * 2^-64 + 2^-65 (RandomNumbers.h:72-84) */
    const double factor = 1.0 / 18446744073709551616.0;
    const double half = 0.5 * factor;
    return (double)in * factor + half;
}

double orng_uneg11_u64(uint64_t in)
{
    /* Random123 uneg11<double,uint64_t>: (int64)in * 2^-63 + 2^-64 (RandomNumbers.h:103-114) */
    const double factor = 1.0 / 9223372036854775808.0;
    const double half = 0.5 * factor;
    return (double)(int64_t)in * factor + half;
}

static uint64_t gen_u64(orng_gen* g)
{
    uint32_t u[4];
    orng_next(g, u);
    return ((uint64_t)u[0] << 32) | u[1];
}

double orng_uniform(orng_gen* g, double a, double b)
{
    /* UniformDistribution (RandomNumbers.h:302-330) */
    double width = b - a;
    return a + width * orng_u01_u64(gen_u64(g));
}

double orng_normal(orng_gen* g, double sigma, double mu)
{
    /* NormalDistribution single draw (RandomNumbers.h:353-365); the host build evaluates
     * sincospi(x) as sincos(M_PI * x) (HOOMDMath.h:330-338). */
    uint32_t u[4];
    orng_next(g, u);
    uint64_t u0 = ((uint64_t)u[0] << 32) | u[1];
    uint64_t u1 = ((uint64_t)u[2] << 32) | u[3];
    double x = sin(M_PI * orng_uneg11_u64(u0));
    double r = sqrt(-2.0 * log(orng_u01_u64(u1)));
    x *= r;
    return x * sigma + mu;
}

double orng_gamma(orng_gen* g, double m_alpha, double m_b)
{
    /* GammaDistribution, Marsaglia-Tsang with squeeze (RandomNumbers.h:452-515) */
    if (m_alpha <= 0)
        return 0.0;
    double alpha = m_alpha;
    if (m_alpha < 1.0)
        alpha += 1.0;
    double d = alpha - 1.0 / 3.0;
    double c = (1.0 / sqrt(d)) / 3.0;
    double v;
    for (;;)
    {
        double x;
        do
        {
            x = orng_normal(g, 1.0, 0.0);
            v = 1.0 + c * x;
        } while (v <= 0.0);
        v = v * v * v;
        const double x2 = x * x;
        double u = orng_u01_u64(gen_u64(g));
        if (u < 1.0 - 0.0331 * x2 * x2)
            break;
        if (log(u) < 0.5 * x2 + d * (1.0 - v + log(v)))
            break;
    }
    v *= d * m_b;
    if (m_alpha < 1.0)
        v *= pow(orng_u01_u64(gen_u64(g)), 1.0 / m_alpha);
    return v;
}

/* =========================================================================== evaluators */

int oe_lj(double rsq, double rcutsq, double sigma, double epsilon, int shift, double* force_divr,
          double* energy)
{
    /* param_type(sigma, epsilon): sigma_6, epsilon_x_4 (hoomd/md/EvaluatorPairLJ.h:124-134);
     * lj1 = 4 eps sigma^12, lj2 = 4 eps sigma^6 (:162-166); evalForceAndEnergy (:190-211) */
    double sigma_6 = sigma * sigma * sigma * sigma * sigma * sigma;
    double epsilon_x_4 = 4.0 * epsilon;
    double lj1 = epsilon_x_4 * sigma_6 * sigma_6;
    double lj2 = epsilon_x_4 * sigma_6;
    if (rsq < rcutsq && lj1 != 0)
    {
        double r2inv = 1.0 / rsq;
        double r6inv = r2inv * r2inv * r2inv;
        *force_divr = r2inv * r6inv * (12.0 * lj1 * r6inv - 6.0 * lj2);
        *energy = r6inv * (lj1 * r6inv - lj2);
        if (shift)
        {
            double rcut2inv = 1.0 / rcutsq;
            double rcut6inv = rcut2inv * rcut2inv * rcut2inv;
            *energy -= rcut6inv * (lj1 * rcut6inv - lj2);
        }
        return 1;
    }
    return 0;
}

int oe_table(double rsq, double rcutsq, double rmin, unsigned int width, const double* V,
             const double* F, double* force_divr, double* energy)
{
    /* EvaluatorPairTable::evalForceAndEnergy (hoomd/md/EvaluatorPairTable.h:152-196): linear
     * interpolation on width points r_min + k dr, dr = (r_cut - r_min) / width, the point past
     * the last one taken as zero. The index is kept below width (the reference reads past the
     * table when rounding puts r - r_min at width dr). */
    double r = sqrt(rsq);
    if (rsq >= rcutsq || r < rmin || width == 0)
        return 0;
    double rcut = sqrt(rcutsq);
    double delta_r = (rcut - rmin) / (double)width;
    double value_f = (r - rmin) / delta_r;
    unsigned int value_i = (unsigned int)floor(value_f);
    if (value_i >= width)
        value_i = width - 1;
    double V0 = V[value_i], F0 = F[value_i], V1 = 0.0, F1 = 0.0;
    if (value_i + 1 < width)
    {
        V1 = V[value_i + 1];
        F1 = F[value_i + 1];
    }
    double f = value_f - (double)value_i;
    double Vr = V0 + f * (V1 - V0);
    double Fr = F0 + f * (F1 - F0);
    if (rsq > 0.0)
        *force_divr = Fr / r;
    *energy = Vr;
    return 1;
}

int oe_fene(double rsq, double K, double r_0, double sigma, double epsilon, double delta,
            double* force_divr, double* bond_eng)
{
    /* fene_params(dict) (hoomd/md/EvaluatorBondFENE.h:45-55), evaluator ctor (:84-90),
     * evalForceAndEnergy (:116-156) */
    double sigma_6 = sigma * sigma * sigma * sigma * sigma * sigma;
    double epsilon_x_4 = 4.0 * epsilon;
    double lj1 = epsilon_x_4 * sigma_6 * sigma_6;
    double lj2 = epsilon_x_4 * sigma_6;

    double rtemp = sqrt(rsq) - delta;
    double rmdoverr = rtemp / sqrt(rsq);
    rsq = rtemp * rtemp;

    double r2inv = 1.0 / rsq;
    double r6inv = r2inv * r2inv * r2inv;
    double wca_divr = 0.0;
    double pair_eng = 0.0;
    double sigma6inv = lj2 / lj1;
    double eps = lj2 * lj2 / 4.0 / lj1;
    if (lj1 != 0 && r6inv > sigma6inv / 2.0)
    {
        wca_divr = r2inv * r6inv * (12.0 * lj1 * r6inv - 6.0 * lj2);
        pair_eng = (r6inv * (lj1 * r6inv - lj2) + eps);
    }
    *force_divr = wca_divr * rmdoverr;
    *bond_eng = pair_eng;
    if (K != 0.0)
    {
        if (rsq >= r_0 * r_0)
            return 0;
        *force_divr += -K / (1.0 - rsq / (r_0 * r_0)) * rmdoverr;
        *bond_eng += -0.5 * K * (r_0 * r_0) * log(1.0 - rsq / (r_0 * r_0));
    }
    return 1;
}

/* =========================================================================== cell list */

void oc_compute_dims(const ob_box* b, double w, unsigned int dim[3])
{
    /* CellList::computeDimensions (hoomd/CellList.cc:83-131), no ghost layer, multiple 1 */
    for (int d = 0; d < 3; d++)
    {
        dim[d] = (unsigned int)((b->L[d] + 2.0 * 0.0) / w);
        if (dim[d] == 0)
            dim[d] = 1;
    }
}

unsigned int oc_n_adj(const unsigned int dim[3])
{
    /* CellList.cc:258-275: unique neighbours when fewer than 3 cells along an axis */
    unsigned int n = 1;
    for (int d = 0; d < 3; d++)
        n *= dim[d] > 3 ? 3 : dim[d];
    return n;
}

static int cmp_uint(const void* a, const void* b)
{
    unsigned int x = *(const unsigned int*)a, y = *(const unsigned int*)b;
    return (x > y) - (x < y);
}

void oc_cell_adj(const unsigned int dim[3], unsigned int* adj)
{
    /* CellList::initializeCellAdj (CellList.cc:346-402): 27 wrapped neighbours, sorted, unique */
    unsigned int n_adj = oc_n_adj(dim);
    int mx = (int)dim[0], my = (int)dim[1], mz = (int)dim[2];
    unsigned int tmp[27];
    for (int k = 0; k < mz; k++)
        for (int j = 0; j < my; j++)
            for (int i = 0; i < mx; i++)
            {
                unsigned int cur = (unsigned int)((k * my + j) * mx + i);
                int n = 0;
                for (int nk = k - 1; nk <= k + 1; nk++)
                    for (int nj = j - 1; nj <= j + 1; nj++)
                        for (int ni = i - 1; ni <= i + 1; ni++)
                        {
                            int wi = ni % mx;
                            if (wi < 0)
                                wi += mx;
                            int wj = nj % my;
                            if (wj < 0)
                                wj += my;
                            int wk = nk % mz;
                            if (wk < 0)
                                wk += mz;
                            tmp[n++] = (unsigned int)((wk * my + wj) * mx + wi);
                        }
                qsort(tmp, 27, sizeof(unsigned int), cmp_uint);
                unsigned int m = 0;
                for (int q = 0; q < 27; q++)
                    if (m == 0 || tmp[q] != adj[(size_t)cur * n_adj + m - 1])
                        adj[(size_t)cur * n_adj + m++] = tmp[q];
            }
}

static int bin_particle(const ob_box* b, const unsigned int dim[3], const double p[3],
                        unsigned int* bin)
{
    /* CellList.cc:440-490 */
    if (isnan(p[0]) || isnan(p[1]) || isnan(p[2]))
        return -1;
    double f[3];
    ob_make_fraction(b, p, f);
    int ib = (int)(f[0] * dim[0]);
    int jb = (int)(f[1] * dim[1]);
    int kb = (int)(f[2] * dim[2]);
    for (int d = 0; d < 3; d++)
        if (f[d] < -0.00001 || f[d] >= 1.00001)
            return -2;
    if (ib == (int)dim[0] && b->periodic[0])
        ib = 0;
    if (jb == (int)dim[1] && b->periodic[1])
        jb = 0;
    if (kb == (int)dim[2] && b->periodic[2])
        kb = 0;
    if (ib < 0 || ib >= (int)dim[0] || jb < 0 || jb >= (int)dim[1] || kb < 0 || kb >= (int)dim[2])
        return -2;
    *bin = (unsigned int)((kb * (int)dim[1] + jb) * (int)dim[0] + ib);
    return 0;
}

int oc_compute_cell_list(const ob_box* b, const unsigned int dim[3], size_t N, const double* pos3,
                         unsigned int cmax, unsigned int* cell_size, unsigned int* cell_idx,
                         unsigned int* cell_of, unsigned int* overflow)
{
    /* CellList::computeCellList (CellList.cc:404-547) storing the particle index (flag index) */
    size_t ncell = (size_t)dim[0] * dim[1] * dim[2];
    memset(cell_size, 0, ncell * sizeof(unsigned int));
    unsigned int ovf = 0;
    int err = 0;
    for (size_t n = 0; n < N; n++)
    {
        unsigned int bin = 0;
        int r = bin_particle(b, dim, &pos3[3 * n], &bin);
        if (r != 0)
        {
            err = -(int)(n + 1);
            if (cell_of)
                cell_of[n] = 0xffffffffu;
            continue;
        }
        if (cell_of)
            cell_of[n] = bin;
        unsigned int off = cell_size[bin];
        if (off < cmax)
            cell_idx[(size_t)bin * cmax + off] = (unsigned int)n;
        else if (off + 1 > ovf)
            ovf = off + 1;
        cell_size[bin]++;
    }
    if (overflow)
        *overflow = ovf;
    return err;
}

/* =========================================================================== neighbour list */

static double dist2_min_image(const ob_box* b, const double* pi, const double* pj)
{
    double dx[3] = {pi[0] - pj[0], pi[1] - pj[1], pi[2] - pj[2]};
    ob_min_image(b, dx);
    return dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2];
}

static int is_excluded(unsigned int i, unsigned int j, const unsigned int* n_excl,
                       const unsigned int* excl, unsigned int max_excl)
{
    if (!n_excl)
        return 0;
    for (unsigned int q = 0; q < n_excl[i]; q++)
        if (excl[(size_t)i * max_excl + q] == j)
            return 1;
    return 0;
}

unsigned int on_build_binned(const ob_box* b, size_t N, const double* pos3, const unsigned int* types,
                             unsigned int ntypes, const double* rcut, double r_buff, int full,
                             const unsigned int* n_excl, const unsigned int* excl,
                             unsigned int max_excl, unsigned int nmax, unsigned int* nlist,
                             unsigned int* n_neigh)
{
    /* NeighborList::updateRList (hoomd/md/NeighborList.cc:452-532): r_list = r_cut + r_buff,
     * cell width = max r_cut + r_buff; NeighborListBinned::buildNlist
     * (hoomd/md/NeighborListBinned.cc:36-179). Exclusions: NeighborList::filterNlist removes
     * excluded pairs after the build (NeighborList.cc:1448-1536); equivalent to skipping them. */
    if (ntypes == 0)
        ntypes = 1;
    double rmax = 0.0;
    for (unsigned int q = 0; q < ntypes * ntypes; q++)
        if (rcut[q] > rmax)
            rmax = rcut[q];
    double* rlistsq = (double*)malloc(sizeof(double) * ntypes * ntypes);
    for (unsigned int q = 0; q < ntypes * ntypes; q++)
    {
        double rl = rcut[q] > 0.0 ? rcut[q] + r_buff : 0.0;
        rlistsq[q] = rl * rl;
    }
    unsigned int dim[3];
    oc_compute_dims(b, rmax + r_buff, dim);
    size_t ncell = (size_t)dim[0] * dim[1] * dim[2];
    unsigned int n_adj = oc_n_adj(dim);
    unsigned int* adj = (unsigned int*)malloc(sizeof(unsigned int) * ncell * n_adj);
    oc_cell_adj(dim, adj);
    /* cell capacity: regrow until no overflow (CellList::compute loop, CellList.cc:183-196) */
    unsigned int cmax = (unsigned int)ceil((double)N / (double)ncell);
    if (cmax == 0)
        cmax = 1;
    unsigned int* cell_size = (unsigned int*)malloc(sizeof(unsigned int) * ncell);
    unsigned int* cell_idx = NULL;
    for (;;)
    {
        free(cell_idx);
        cell_idx = (unsigned int*)malloc(sizeof(unsigned int) * ncell * cmax);
        unsigned int ovf = 0;
        oc_compute_cell_list(b, dim, N, pos3, cmax, cell_size, cell_idx, NULL, &ovf);
        if (ovf <= cmax)
            break;
        cmax = ovf;
    }
    unsigned int max_seen = 0;
#pragma omp parallel for schedule(dynamic, 512) reduction(max : max_seen) if (g_threads > 1) num_threads(g_threads)
    for (size_t i = 0; i < N; i++)
    {
        unsigned int cnt = 0;
        const double* pi = &pos3[3 * i];
        unsigned int ti = types ? types[i] : 0;
        unsigned int my_cell;
        if (bin_particle(b, dim, pi, &my_cell) != 0)
        {
            n_neigh[i] = 0;
            continue;
        }
        for (unsigned int a = 0; a < n_adj; a++)
        {
            unsigned int nc = adj[(size_t)my_cell * n_adj + a];
            unsigned int size = cell_size[nc];
            for (unsigned int o = 0; o < size; o++)
            {
                unsigned int j = cell_idx[(size_t)nc * cmax + o];
                unsigned int tj = types ? types[j] : 0;
                double rc = rcut[ti * ntypes + tj];
                if (j == i || rc <= 0.0)
                    continue;
                if (is_excluded((unsigned int)i, j, n_excl, excl, max_excl))
                    continue;
                double dr2 = dist2_min_image(b, pi, &pos3[3 * j]);
                if (dr2 <= rlistsq[ti * ntypes + tj])
                {
                    if (full || i < j)
                    {
                        if (cnt < nmax)
                            nlist[i * nmax + cnt] = j;
                        cnt++;
                    }
                }
            }
        }
        n_neigh[i] = cnt;
        if (cnt > max_seen)
            max_seen = cnt;
    }
    free(rlistsq);
    free(adj);
    free(cell_size);
    free(cell_idx);
    return max_seen;
}

unsigned int on_build_brute(const ob_box* b, size_t N, const double* pos3, double r_list, int full,
                            unsigned int nmax, unsigned int* nlist, unsigned int* n_neigh)
{
    double rl2 = r_list * r_list;
    unsigned int max_seen = 0;
    for (size_t i = 0; i < N; i++)
    {
        unsigned int cnt = 0;
        for (size_t j = 0; j < N; j++)
        {
            if (j == i || (!full && j < i))
                continue;
            if (dist2_min_image(b, &pos3[3 * i], &pos3[3 * j]) <= rl2)
            {
                if (cnt < nmax)
                    nlist[i * nmax + cnt] = (unsigned int)j;
                cnt++;
            }
        }
        n_neigh[i] = cnt;
        if (cnt > max_seen)
            max_seen = cnt;
    }
    return max_seen;
}

int on_distance_check(const ob_box* b, size_t N, const double* pos3, const double* last3,
                      double rcut_max, double r_buff)
{
    /* NeighborList::distanceCheck (hoomd/md/NeighborList.cc:1220-1289) for an unchanged box:
     * lambda = 1, delta_max = ((rcut_max + r_buff) * 1 - rcut_max) / 2 */
    double old_rmin = rcut_max;
    double rmax = old_rmin + r_buff;
    double delta_max = (rmax * 1.0 - old_rmin) / 2.0;
    double maxsq = delta_max > 0 ? delta_max * delta_max : 0;
    int moved = 0;
#pragma omp parallel for reduction(| : moved) if (g_threads > 1) num_threads(g_threads)
    for (size_t i = 0; i < N; i++)
    {
        double dx[3];
        for (int d = 0; d < 3; d++)
            dx[d] = pos3[3 * i + d] - 1.0 * last3[3 * i + d];
        ob_min_image(b, dx);
        if (dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2] >= maxsq)
            moved = 1;
    }
    return moved;
}

/* =========================================================================== forces */

/* One pair evaluation: force / r and pair energy for rsq; 0 when the pair does not interact */
typedef int (*pair_eval_fn)(double rsq, const void* ctx, double* force_divr, double* energy);

/* PotentialPair<Evaluator>::computeForces (hoomd/md/PotentialPair.h:587-794) with the evaluator
 * as a callback: per particle, force sum, half pair energy and half virial (PotentialPair.h:
 * 739-772); with a half list (third_law) the partner gets the opposite force and the other
 * halves. */
/* Particle i's row: its force, half energies and half virials, and with a half list
 * (third_law) the partners' opposite forces and other halves, added into force4 / virial6. */
static inline void pair_row(const ob_box* b, size_t N, size_t i, const double* pos3,
                            const unsigned int* nlist, const unsigned int* n_neigh,
                            unsigned int nmax, int third_law, int compute_virial,
                            pair_eval_fn eval, const void* ctx, double* force4, double* virial6)
{
    const double* pi = &pos3[3 * i];
    double fi[3] = {0, 0, 0}, pei = 0.0;
    double vxx = 0, vxy = 0, vxz = 0, vyy = 0, vyz = 0, vzz = 0;
    for (unsigned int k = 0; k < n_neigh[i]; k++)
    {
        unsigned int j = nlist[i * nmax + k];
        double dx[3] = {pi[0] - pos3[3 * j], pi[1] - pos3[3 * j + 1], pi[2] - pos3[3 * j + 2]};
        ob_min_image(b, dx);
        double rsq = dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2];
        double force_divr = 0.0, pair_eng = 0.0;
        if (!eval(rsq, ctx, &force_divr, &pair_eng))
            continue;
        double f2 = force_divr * 0.5;
        fi[0] += dx[0] * force_divr;
        fi[1] += dx[1] * force_divr;
        fi[2] += dx[2] * force_divr;
        pei += pair_eng * 0.5;
        if (compute_virial)
        {
            vxx += f2 * dx[0] * dx[0];
            vxy += f2 * dx[0] * dx[1];
            vxz += f2 * dx[0] * dx[2];
            vyy += f2 * dx[1] * dx[1];
            vyz += f2 * dx[1] * dx[2];
            vzz += f2 * dx[2] * dx[2];
        }
        if (third_law)
        {
            force4[4 * j + 0] -= dx[0] * force_divr;
            force4[4 * j + 1] -= dx[1] * force_divr;
            force4[4 * j + 2] -= dx[2] * force_divr;
            force4[4 * j + 3] += pair_eng * 0.5;
            if (compute_virial)
            {
                virial6[0 * N + j] += f2 * dx[0] * dx[0];
                virial6[1 * N + j] += f2 * dx[0] * dx[1];
                virial6[2 * N + j] += f2 * dx[0] * dx[2];
                virial6[3 * N + j] += f2 * dx[1] * dx[1];
                virial6[4 * N + j] += f2 * dx[1] * dx[2];
                virial6[5 * N + j] += f2 * dx[2] * dx[2];
            }
        }
    }
    force4[4 * i + 0] += fi[0];
    force4[4 * i + 1] += fi[1];
    force4[4 * i + 2] += fi[2];
    force4[4 * i + 3] += pei;
    if (compute_virial)
    {
        virial6[0 * N + i] += vxx;
        virial6[1 * N + i] += vxy;
        virial6[2 * N + i] += vxz;
        virial6[3 * N + i] += vyy;
        virial6[4 * N + i] += vyz;
        virial6[5 * N + i] += vzz;
    }
}

static void pair_accumulate(const ob_box* b, size_t N, const double* pos3,
                            const unsigned int* nlist, const unsigned int* n_neigh,
                            unsigned int nmax, int third_law, int compute_virial,
                            pair_eval_fn eval, const void* ctx, double* force4, double* virial6)
{
    memset(force4, 0, sizeof(double) * 4 * N);
    if (virial6)
        memset(virial6, 0, sizeof(double) * 6 * N);
    /* rows of a full list are independent: threads take contiguous ranges of them */
#pragma omp parallel for schedule(dynamic, 512) if (g_threads > 1 && !third_law) num_threads(g_threads)
    for (size_t i = 0; i < N; i++)
        pair_row(b, N, i, pos3, nlist, n_neigh, nmax, third_law, compute_virial, eval, ctx,
                 force4, virial6);
}

typedef struct
{
    double sigma, epsilon, rcutsq, ronsq;
    int shift_mode;
} lj_ctx;

static int lj_eval(double rsq, const void* c, double* force_divr, double* pair_eng)
{
    const lj_ctx* p = (const lj_ctx*)c;
    int energy_shift = 0;
    if (p->shift_mode == 1)
        energy_shift = 1;
    else if (p->shift_mode == 2 && p->ronsq > p->rcutsq)
        energy_shift = 1;
    if (!oe_lj(rsq, p->rcutsq, p->sigma, p->epsilon, energy_shift, force_divr, pair_eng))
        return 0;
    if (p->shift_mode == 2 && rsq >= p->ronsq && rsq < p->rcutsq)
    {
        /* XPLOR smoothing (PotentialPair.h:706-731) */
        double old_e = *pair_eng, old_f = *force_divr;
        double den = 1.0 / ((p->rcutsq - p->ronsq) * (p->rcutsq - p->ronsq) * (p->rcutsq - p->ronsq));
        double rmc = rsq - p->rcutsq;
        double s = rmc * rmc * (p->rcutsq + 2.0 * rsq - 3.0 * p->ronsq) * den;
        double ds = 12.0 * (rsq - p->ronsq) * rmc * den;
        *pair_eng = old_e * s;
        *force_divr = s * old_f - ds * old_e;
    }
    return 1;
}

void op_pair_lj(const ob_box* b, size_t N, const double* pos3, const unsigned int* nlist,
                const unsigned int* n_neigh, unsigned int nmax, int third_law, double sigma,
                double epsilon, double rcut, double ron, int shift_mode, int compute_virial,
                double* force4, double* virial6)
{
    /* PotentialPair<EvaluatorPairLJ>::computeForces (hoomd/md/PotentialPair.h:587-794) */
    lj_ctx c = {sigma, epsilon, rcut * rcut, ron * ron, shift_mode};
    pair_accumulate(b, N, pos3, nlist, n_neigh, nmax, third_law, compute_virial, lj_eval, &c,
                    force4, virial6);
}

typedef struct
{
    double rcutsq, rmin;
    unsigned int width;
    const double *V, *F;
} table_ctx;

static int table_eval(double rsq, const void* c, double* force_divr, double* pair_eng)
{
    const table_ctx* p = (const table_ctx*)c;
    return oe_table(rsq, p->rcutsq, p->rmin, p->width, p->V, p->F, force_divr, pair_eng);
}

void op_pair_table(const ob_box* b, size_t N, const double* pos3, const unsigned int* nlist,
                   const unsigned int* n_neigh, unsigned int nmax, int third_law, double rcut,
                   double rmin, unsigned int width, const double* V, const double* F,
                   int compute_virial, double* force4, double* virial6)
{
    /* PotentialPair<EvaluatorPairTable>::computeForces (hoomd/md/PotentialPair.h:587-794) */
    table_ctx c = {rcut * rcut, rmin, width, V, F};
    pair_accumulate(b, N, pos3, nlist, n_neigh, nmax, third_law, compute_virial, table_eval, &c,
                    force4, virial6);
}

int op_bond_fene(const ob_box* b, size_t N, const double* pos3, size_t nb, const unsigned int* bonds,
                 double k, double r0, double sigma, double epsilon, double delta,
                 int compute_virial, double* force4, double* virial6)
{
    /* PotentialBond<EvaluatorBondFENE>::computeForces (hoomd/md/PotentialBond.h:167-315) */
    memset(force4, 0, sizeof(double) * 4 * N);
    if (virial6)
        memset(virial6, 0, sizeof(double) * 6 * N);
    for (size_t q = 0; q < nb; q++)
    {
        unsigned int a = bonds[2 * q], c = bonds[2 * q + 1];
        double dx[3] = {pos3[3 * c] - pos3[3 * a], pos3[3 * c + 1] - pos3[3 * a + 1],
                        pos3[3 * c + 2] - pos3[3 * a + 2]};
        ob_min_image(b, dx);
        double rsq = dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2];
        double fdr = 0.0, e = 0.0;
        int ev = oe_fene(rsq, k, r0, sigma, epsilon, delta, &fdr, &e);
        e *= 0.5;
        if (!ev)
            return -(int)(q + 1);
        double bv[6] = {0, 0, 0, 0, 0, 0};
        if (compute_virial)
        {
            double f2 = 0.5 * fdr;
            bv[0] = dx[0] * dx[0] * f2;
            bv[1] = dx[0] * dx[1] * f2;
            bv[2] = dx[0] * dx[2] * f2;
            bv[3] = dx[1] * dx[1] * f2;
            bv[4] = dx[1] * dx[2] * f2;
            bv[5] = dx[2] * dx[2] * f2;
        }
        for (int d = 0; d < 3; d++)
        {
            force4[4 * c + d] += fdr * dx[d];
            force4[4 * a + d] -= fdr * dx[d];
        }
        force4[4 * c + 3] += e;
        force4[4 * a + 3] += e;
        if (compute_virial)
            for (int m = 0; m < 6; m++)
            {
                virial6[m * N + c] += bv[m];
                virial6[m * N + a] += bv[m];
            }
    }
    return 0;
}

/* =========================================================================== integrators */

void oi_nvt_step_one(const ob_box* b, size_t N, double* pos3, double* vel3, const double* accel3,
                     int* image3, double dt, double alpha)
{
    /* TwoStepConstantVolume::integrateStepOne (hoomd/md/TwoStepConstantVolume.cc:7-80) */
#pragma omp parallel for if (g_threads > 1) num_threads(g_threads)
    for (size_t j = 0; j < N; j++)
    {
        for (int d = 0; d < 3; d++)
        {
            double v = vel3[3 * j + d];
            v = v + (1.0 / 2.0) * accel3[3 * j + d] * dt;
            v *= alpha;
            pos3[3 * j + d] += dt * v;
            vel3[3 * j + d] = v;
        }
    }
#pragma omp parallel for if (g_threads > 1) num_threads(g_threads)
    for (size_t j = 0; j < N; j++)
        ob_wrap(b, &pos3[3 * j], &image3[3 * j]);
}

void oi_nvt_step_two(size_t N, double* vel3, double* accel3, const double* mass,
                     const double* net_force4, double dt, double alpha)
{
    /* TwoStepConstantVolume::integrateStepTwo (TwoStepConstantVolume.cc:213-260) */
#pragma omp parallel for if (g_threads > 1) num_threads(g_threads)
    for (size_t j = 0; j < N; j++)
    {
        double minv = 1.0 / mass[j];
        for (int d = 0; d < 3; d++)
        {
            double a = net_force4[4 * j + d] * minv;
            double v = vel3[3 * j + d];
            v *= alpha;
            v += (1.0 / 2.0) * dt * a;
            vel3[3 * j + d] = v;
            accel3[3 * j + d] = a;
        }
    }
}

void oi_langevin_step_two(size_t N, double* vel3, double* accel3, const double* mass,
                          const unsigned int* tag, const double* net_force4, const double* gamma,
                          const unsigned int* types, double kT, double dt, uint64_t timestep,
                          uint16_t seed)
{
    /* TwoStepLangevin::integrateStepTwo (hoomd/md/TwoStepLangevin.cc:206-275), no tally,
     * translational part only; RNG id TwoStepLangevin = 25 (hoomd/RNGIdentifiers.h:51) */
    for (size_t j = 0; j < N; j++)
    {
        orng_gen g;
        orng_init(&g, 25, timestep, seed, tag[j], 0, 0, 0);
        double rx = orng_uniform(&g, -1.0, 1.0);
        double ry = orng_uniform(&g, -1.0, 1.0);
        double rz = orng_uniform(&g, -1.0, 1.0);
        double gm = gamma[types ? types[j] : 0];
        double coeff = sqrt(6.0 * gm * kT / dt);
        double bd[3] = {rx * coeff - gm * vel3[3 * j], ry * coeff - gm * vel3[3 * j + 1],
                        rz * coeff - gm * vel3[3 * j + 2]};
        double minv = 1.0 / mass[j];
        for (int d = 0; d < 3; d++)
        {
            accel3[3 * j + d] = (net_force4[4 * j + d] + bd[d]) * minv;
            vel3[3 * j + d] += (1.0 / 2.0) * accel3[3 * j + d] * dt;
        }
    }
}

double oi_kinetic_energy(size_t N, const double* vel3, const double* mass)
{
    /* ComputeThermo::computeProperties, no pressure tensor (hoomd/md/ComputeThermo.cc:133-149) */
    double ke = 0.0;
#pragma omp parallel for reduction(+ : ke) if (g_threads > 1) num_threads(g_threads)
    for (size_t j = 0; j < N; j++)
        ke += mass[j]
              * (vel3[3 * j] * vel3[3 * j] + vel3[3 * j + 1] * vel3[3 * j + 1]
                 + vel3[3 * j + 2] * vel3[3 * j + 2]);
    return ke * 0.5;
}

double oi_bussi_alpha(double K, double ndof, double dt, double tau, double kT, uint64_t timestep,
                      uint16_t seed, uint32_t instance_tag)
{
    /* BussiThermostat::getRescalingFactorsOne + compute_rescale_factor
     * (hoomd/md/Thermostat.h:357-446); RNG id BussiThermostat = 45 (RNGIdentifiers.h:70) */
    if (dt == 0.0 || ndof == 0)
        return 1.0;
    orng_gen g;
    orng_init(&g, 45, timestep, seed, instance_tag, 0, 0, 0);
    double c = 0.0;
    if (tau != 0.0)
        c = exp(-dt / tau);
    double r1 = orng_normal(&g, 1.0, 0.0);
    double rg = 0.0;
    if (ndof > 1.0)
        rg = 2.0 * orng_gamma(&g, (ndof - 1.0) / 2.0, 1.0);
    double v = kT / 2.0 / K;
    double term1 = v * (1.0 - c) * (rg + r1 * r1);
    double term2 = 2.0 * r1 * sqrt(v * (1.0 - c) * c);
    return sqrt(c + term1 + term2);
}

void oi_thermalize(size_t N, double* vel3, const double* mass, const unsigned int* tag, double kT,
                   uint64_t timestep, uint16_t seed)
{
    /* ParticleGroup::thermalizeParticleMomenta (hoomd/ParticleGroup.cc:690-806), 3D, no
     * rigid bodies, zero moment of inertia; RNG id ParticleGroupThermalize = 35 */
    double tot[3] = {0, 0, 0};
    for (size_t j = 0; j < N; j++)
    {
        orng_gen g;
        orng_init(&g, 35, timestep, seed, tag[j], 0, 0, 0);
        double sigma = sqrt(kT / mass[j]);
        vel3[3 * j] = orng_normal(&g, sigma, 0.0);
        vel3[3 * j + 1] = orng_normal(&g, sigma, 0.0);
        vel3[3 * j + 2] = orng_normal(&g, sigma, 0.0);
        for (int d = 0; d < 3; d++)
            tot[d] += mass[j] * vel3[3 * j + d];
    }
    double com[3] = {tot[0] / (double)N, tot[1] / (double)N, tot[2] / (double)N};
    for (size_t j = 0; j < N; j++)
        for (int d = 0; d < 3; d++)
            vel3[3 * j + d] = vel3[3 * j + d] - com[d] / mass[j];
}

/* =========================================================================== Hilbert sort
 * hoomd/SFCPackTuner.cc:290-487 (walking tables, the eight child-order permutation rules and the
 * recursion) and :541-650 (binning and stable sort). */

static const int h_istep[8] = {0, 0, 0, 0, 1, 1, 1, 1};
static const int h_jstep[8] = {0, 0, 1, 1, 1, 1, 0, 0};
static const int h_kstep[8] = {0, 1, 1, 0, 0, 1, 1, 0};
/* child order rule m: result[q] = in[h_perm[m][q]] */
static const unsigned char h_perm[8][8] = {
    {0, 3, 4, 7, 6, 5, 2, 1}, {0, 7, 6, 1, 2, 5, 4, 3}, {0, 7, 6, 1, 2, 5, 4, 3},
    {2, 3, 0, 1, 6, 7, 4, 5}, {2, 3, 0, 1, 6, 7, 4, 5}, {4, 3, 2, 5, 6, 1, 0, 7},
    {4, 3, 2, 5, 6, 1, 0, 7}, {6, 5, 2, 1, 0, 3, 4, 7}};

static void hilbert_rec(int i, int j, int k, int w, int Mx, const unsigned int order[8],
                        unsigned int* traversal, unsigned int* rank)
{
    if (w == 1)
    {
        traversal[(size_t)i * Mx * Mx + (size_t)j * Mx + k] = (*rank)++;
        return;
    }
    w /= 2;
    for (int m = 0; m < 8; m++)
    {
        unsigned int cur = order[m];
        unsigned int child[8];
        for (int q = 0; q < 8; q++)
            child[q] = order[h_perm[m][q]];
        hilbert_rec(i + w * h_istep[cur], j + w * h_jstep[cur], k + w * h_kstep[cur], w, Mx, child,
                    traversal, rank);
    }
}

void os_hilbert_traversal(unsigned int grid, unsigned int* traversal)
{
    unsigned int order[8] = {0, 1, 2, 3, 4, 5, 6, 7};
    unsigned int rank = 0;
    hilbert_rec(0, 0, 0, (int)grid, (int)grid, order, traversal, &rank);
}

typedef struct
{
    unsigned int key;
    unsigned int idx;
} sort_pair;

static int cmp_pair(const void* a, const void* b)
{
    const sort_pair* x = (const sort_pair*)a;
    const sort_pair* y = (const sort_pair*)b;
    if (x->key != y->key)
        return (x->key > y->key) - (x->key < y->key);
    return (x->idx > y->idx) - (x->idx < y->idx);
}

void os_sort_order(const ob_box* b, size_t N, const double* pos3, unsigned int grid,
                   const unsigned int* traversal, unsigned int* order)
{
    sort_pair* bins = (sort_pair*)malloc(sizeof(sort_pair) * N);
    for (size_t n = 0; n < N; n++)
    {
        double f[3];
        ob_make_fraction(b, &pos3[3 * n], f);
        int ib = (int)((unsigned int)(f[0] * grid) % grid);
        int jb = (int)((unsigned int)(f[1] * grid) % grid);
        int kb = (int)((unsigned int)(f[2] * grid) % grid);
        if (ib < 0)
            ib = 0;
        if (ib >= (int)grid)
            ib = (int)grid - 1;
        if (jb < 0)
            jb = 0;
        if (jb >= (int)grid)
            jb = (int)grid - 1;
        if (kb < 0)
            kb = 0;
        if (kb >= (int)grid)
            kb = (int)grid - 1;
        size_t bin = (size_t)ib * grid * grid + (size_t)jb * grid + (size_t)kb;
        bins[n].key = traversal[bin];
        bins[n].idx = (unsigned int)n;
    }
    qsort(bins, N, sizeof(sort_pair), cmp_pair);
    for (size_t n = 0; n < N; n++)
        order[n] = bins[n].idx;
    free(bins);
}

/* =========================================================================== full system
 * The step loop of System::run (hoomd/System.cc:80-181) with tuners first, then
 * IntegratorTwoStep::update (hoomd/md/IntegratorTwoStep.cc:51-123); prepRun computes the first
 * net force and accelerations (IntegratorTwoStep.cc:226-282). The neighbour list rebuild policy
 * follows NeighborList::compute / needsUpdating (NeighborList.cc:363-435,1325-1418) with
 * rebuild_check_delay = 1 and distance checks on. CPU storage mode is half (pair.py:135-138). */

struct os_system
{
    size_t N;
    ob_box box;
    double *pos, *vel, *accel, *mass, *last;
    int* image;
    unsigned int *tag, *rtag;
    double sigma, epsilon, rcut;
    int shift_mode;
    double r_buff, dt;
    uint16_t seed;
    int thermostat;
    double kT, tau_gamma;
    unsigned int sort_period, sort_grid;
    unsigned int* traversal;
    /* neighbour list (half) */
    unsigned int nmax;
    unsigned int *nlist, *n_neigh;
    int force_update;
    uint64_t n_builds;
    /* rebuild policy and statistics (NeighborList.cc:1313-1426) */
    unsigned int check_delay;
    int dist_check;
    uint64_t last_update, n_dangerous;
    uint64_t update_periods[100];
    /* bonds (by tag) */
    size_t nb;
    unsigned int* bond_tags;
    double bk, br0, bsigma, beps, bdelta;
    int exclude_bonds;
    unsigned int max_excl;
    unsigned int *n_excl, *excl;
    /* outputs (index order) */
    double *pair_f, *pair_v, *bond_f, *net_f;
    uint64_t timestep;
    int prepared;
};

os_system* os_create(size_t N, double Lx, double Ly, double Lz, const double* pos3,
                     const double* vel3, const double* mass, double sigma, double epsilon,
                     double rcut, int shift_mode, double r_buff, double dt, uint16_t seed)
{
    os_system* s = (os_system*)calloc(1, sizeof(os_system));
    s->N = N;
    s->check_delay = 1;
    s->dist_check = 1;
    ob_box_init(&s->box, Lx, Ly, Lz);
    s->pos = (double*)malloc(sizeof(double) * 3 * N);
    s->vel = (double*)malloc(sizeof(double) * 3 * N);
    s->accel = (double*)calloc(3 * N, sizeof(double));
    s->mass = (double*)malloc(sizeof(double) * N);
    s->last = (double*)malloc(sizeof(double) * 3 * N);
    s->image = (int*)calloc(3 * N, sizeof(int));
    s->tag = (unsigned int*)malloc(sizeof(unsigned int) * N);
    s->rtag = (unsigned int*)malloc(sizeof(unsigned int) * N);
    memcpy(s->pos, pos3, sizeof(double) * 3 * N);
    memcpy(s->vel, vel3, sizeof(double) * 3 * N);
    for (size_t i = 0; i < N; i++)
    {
        s->mass[i] = mass ? mass[i] : 1.0;
        s->tag[i] = s->rtag[i] = (unsigned int)i;
    }
    s->sigma = sigma;
    s->epsilon = epsilon;
    s->rcut = rcut;
    s->shift_mode = shift_mode;
    s->r_buff = r_buff;
    s->dt = dt;
    s->seed = seed;
    s->nmax = 4; /* NeighborList.cc:145-152: Nmax starts at 4 */
    s->nlist = (unsigned int*)malloc(sizeof(unsigned int) * N * s->nmax);
    s->n_neigh = (unsigned int*)calloc(N, sizeof(unsigned int));
    s->force_update = 1;
    s->pair_f = (double*)calloc(4 * N, sizeof(double));
    s->pair_v = (double*)calloc(6 * N, sizeof(double));
    s->bond_f = (double*)calloc(4 * N, sizeof(double));
    s->net_f = (double*)calloc(4 * N, sizeof(double));
    return s;
}

void os_destroy(os_system* s)
{
    if (!s)
        return;
    free(s->pos);
    free(s->vel);
    free(s->accel);
    free(s->mass);
    free(s->last);
    free(s->image);
    free(s->tag);
    free(s->rtag);
    free(s->traversal);
    free(s->nlist);
    free(s->n_neigh);
    free(s->bond_tags);
    free(s->n_excl);
    free(s->excl);
    free(s->pair_f);
    free(s->pair_v);
    free(s->bond_f);
    free(s->net_f);
    free(s);
}

void os_set_thermostat(os_system* s, int kind, double kT, double tau_or_gamma)
{
    s->thermostat = kind;
    s->kT = kT;
    s->tau_gamma = tau_or_gamma;
}

void os_set_sorter(os_system* s, unsigned int period, unsigned int grid)
{
    s->sort_period = period;
    s->sort_grid = grid;
    free(s->traversal);
    s->traversal = NULL;
    if (period)
    {
        s->traversal = (unsigned int*)malloc(sizeof(unsigned int) * grid * grid * grid);
        os_hilbert_traversal(grid, s->traversal);
    }
}

void os_set_bonds(os_system* s, size_t nb, const unsigned int* bond_tags, double k, double r0,
                  double sigma, double epsilon, double delta, int exclude_bonds)
{
    s->nb = nb;
    free(s->bond_tags);
    s->bond_tags = (unsigned int*)malloc(sizeof(unsigned int) * 2 * nb);
    memcpy(s->bond_tags, bond_tags, sizeof(unsigned int) * 2 * nb);
    s->bk = k;
    s->br0 = r0;
    s->bsigma = sigma;
    s->beps = epsilon;
    s->bdelta = delta;
    s->exclude_bonds = exclude_bonds;
    s->force_update = 1;
}

static void os_build_exclusions(os_system* s)
{
    /* NeighborList::addExclusionsFromBonds (hoomd/md/NeighborList.cc:813-846), index space */
    size_t N = s->N;
    free(s->n_excl);
    free(s->excl);
    s->n_excl = NULL;
    s->excl = NULL;
    if (!s->exclude_bonds || s->nb == 0)
        return;
    unsigned int* cnt = (unsigned int*)calloc(N, sizeof(unsigned int));
    for (size_t q = 0; q < s->nb; q++)
    {
        cnt[s->bond_tags[2 * q]]++;
        cnt[s->bond_tags[2 * q + 1]]++;
    }
    unsigned int mx = 1;
    for (size_t i = 0; i < N; i++)
        if (cnt[i] > mx)
            mx = cnt[i];
    s->max_excl = mx;
    s->n_excl = (unsigned int*)calloc(N, sizeof(unsigned int));
    s->excl = (unsigned int*)malloc(sizeof(unsigned int) * N * mx);
    for (size_t q = 0; q < s->nb; q++)
    {
        unsigned int a = s->rtag[s->bond_tags[2 * q]], c = s->rtag[s->bond_tags[2 * q + 1]];
        s->excl[(size_t)a * mx + s->n_excl[a]++] = c;
        s->excl[(size_t)c * mx + s->n_excl[c]++] = a;
    }
    free(cnt);
}

static void os_apply_sort(os_system* s)
{
    /* SFCPackTuner::update -> getSortedOrder3D + applySortOrder (SFCPackTuner.cc:81-285) */
    size_t N = s->N;
    unsigned int* order = (unsigned int*)malloc(sizeof(unsigned int) * N);
    os_sort_order(&s->box, N, s->pos, s->sort_grid, s->traversal, order);
    double* t3 = (double*)malloc(sizeof(double) * 3 * N);
    int* i3 = (int*)malloc(sizeof(int) * 3 * N);
    double* t1 = (double*)malloc(sizeof(double) * N);
    unsigned int* u1 = (unsigned int*)malloc(sizeof(unsigned int) * N);
    double* t4 = (double*)malloc(sizeof(double) * 4 * N);
#define PERMUTE3(arr, tmp)                                                                          \
    for (size_t n = 0; n < N; n++)                                                                 \
        for (int d = 0; d < 3; d++)                                                                \
            tmp[3 * n + d] = arr[3 * (size_t)order[n] + d];                                        \
    memcpy(arr, tmp, sizeof(tmp[0]) * 3 * N);
    PERMUTE3(s->pos, t3)
    PERMUTE3(s->vel, t3)
    PERMUTE3(s->accel, t3)
    PERMUTE3(s->last, t3)
    PERMUTE3(s->image, i3)
#undef PERMUTE3
    for (size_t n = 0; n < N; n++)
        t1[n] = s->mass[order[n]];
    memcpy(s->mass, t1, sizeof(double) * N);
    for (size_t n = 0; n < N; n++)
        u1[n] = s->tag[order[n]];
    memcpy(s->tag, u1, sizeof(unsigned int) * N);
    for (size_t n = 0; n < N; n++)
        s->rtag[s->tag[n]] = (unsigned int)n;
    for (size_t n = 0; n < N; n++)
        for (int d = 0; d < 4; d++)
            t4[4 * n + d] = s->net_f[4 * (size_t)order[n] + d];
    memcpy(s->net_f, t4, sizeof(double) * 4 * N);
    free(order);
    free(t3);
    free(i3);
    free(t1);
    free(u1);
    free(t4);
    /* the sort signal forces a neighbour-list rebuild (NeighborList.cc:222) */
    s->force_update = 1;
}

static void os_nlist_compute(os_system* s, uint64_t timestep)
{
    /* NeighborList::needsUpdating (hoomd/md/NeighborList.cc:1325-1418), check once per step */
    int rebuild = 0, dangerous = 0;
    if (s->force_update)
    {
        /* a forced build: no statistics, never dangerous (:1358-1368) */
        rebuild = 1;
    }
    else if (!(timestep < s->last_update + s->check_delay)) /* shouldCheckDistance (:1313-1316) */
    {
        dangerous = s->dist_check && s->check_delay > 1
                    && timestep == s->last_update + s->check_delay;
        if (s->r_buff < 1e-6
            || (!s->dist_check
                && (s->check_delay == 0
                    || (s->check_delay > 1 && timestep == s->last_update + s->check_delay))))
            rebuild = 1;
        else
            rebuild = on_distance_check(&s->box, s->N, s->pos, s->last, s->rcut, s->r_buff);
        if (rebuild && timestep > s->last_update)
        {
            uint64_t period = timestep - s->last_update;
            if (period >= 100)
                period = 99;
            s->update_periods[period]++;
        }
    }
    if (!rebuild)
        return;
    if (dangerous)
        s->n_dangerous++;
    s->last_update = timestep;
    s->force_update = 0;
    os_build_exclusions(s);
    for (;;)
    {
        unsigned int need = on_build_binned(&s->box, s->N, s->pos, NULL, 1, &s->rcut, s->r_buff,
                                            g_threads > 1,
                                            s->n_excl, s->excl, s->max_excl, s->nmax, s->nlist,
                                            s->n_neigh);
        if (need <= s->nmax)
            break;
        /* NeighborList::checkConditions: Nmax = round up to multiple of 4 (NeighborList.cc:1601) */
        s->nmax = need > 4 ? (need + 3) & ~3u : 4;
        free(s->nlist);
        s->nlist = (unsigned int*)malloc(sizeof(unsigned int) * s->N * s->nmax);
    }
    memcpy(s->last, s->pos, sizeof(double) * 3 * s->N);
    s->n_builds++;
}

static int os_compute_net_force(os_system* s, uint64_t timestep)
{
    size_t N = s->N;
    os_nlist_compute(s, timestep);
    op_pair_lj(&s->box, N, s->pos, s->nlist, s->n_neigh, s->nmax, g_threads > 1 ? 0 : 1, s->sigma,
               s->epsilon, s->rcut, 0.0, s->shift_mode, 1, s->pair_f, s->pair_v);
    memcpy(s->net_f, s->pair_f, sizeof(double) * 4 * N);
    if (s->nb)
    {
        unsigned int* bidx = (unsigned int*)malloc(sizeof(unsigned int) * 2 * s->nb);
        for (size_t q = 0; q < 2 * s->nb; q++)
            bidx[q] = s->rtag[s->bond_tags[q]];
        int r = op_bond_fene(&s->box, N, s->pos, s->nb, bidx, s->bk, s->br0, s->bsigma, s->beps,
                             s->bdelta, 0, s->bond_f, NULL);
        free(bidx);
        if (r != 0)
            return -3;
        for (size_t q = 0; q < 4 * N; q++)
            s->net_f[q] += s->bond_f[q];
    }
    return 0;
}

int os_run(os_system* s, uint64_t nsteps)
{
    size_t N = s->N;
    double ndof = 3.0 * (double)N - (s->thermostat == 2 ? 0.0 : 3.0);
    if (!s->prepared)
    {
        int r = os_compute_net_force(s, s->timestep);
        if (r)
            return r;
        for (size_t j = 0; j < N; j++)
            for (int d = 0; d < 3; d++)
                s->accel[3 * j + d] = s->net_f[4 * j + d] / s->mass[j];
        s->prepared = 1;
    }
    for (uint64_t c = 0; c < nsteps; c++)
    {
        uint64_t t = s->timestep;
        if (s->sort_period && t % s->sort_period == 0)
            os_apply_sort(s);
        double alpha = 1.0;
        if (s->thermostat == 1)
        {
            double K = oi_kinetic_energy(N, s->vel, s->mass);
            /* instance id = tag of the first group member; the all group's first tag is 0 */
            alpha = oi_bussi_alpha(K, ndof, s->dt, s->tau_gamma, s->kT, t, s->seed, 0);
        }
        oi_nvt_step_one(&s->box, N, s->pos, s->vel, s->accel, s->image, s->dt, alpha);
        int r = os_compute_net_force(s, t + 1);
        if (r)
            return r;
        if (s->thermostat == 2)
        {
            double gm = s->tau_gamma;
            oi_langevin_step_two(N, s->vel, s->accel, s->mass, s->tag, s->net_f, &gm, NULL, s->kT,
                                 s->dt, t, s->seed);
        }
        else
            oi_nvt_step_two(N, s->vel, s->accel, s->mass, s->net_f, s->dt, 1.0);
        s->timestep++;
    }
    return 0;
}

void os_get_state(const os_system* s, double* pos3, double* vel3, int* image3)
{
    for (size_t n = 0; n < s->N; n++)
    {
        unsigned int t = s->tag[n];
        for (int d = 0; d < 3; d++)
        {
            if (pos3)
                pos3[3 * t + d] = s->pos[3 * n + d];
            if (vel3)
                vel3[3 * t + d] = s->vel[3 * n + d];
            if (image3)
                image3[3 * t + d] = s->image[3 * n + d];
        }
    }
}

void os_get_forces(const os_system* s, double* pf, double* pv, double* bf)
{
    size_t N = s->N;
    for (size_t n = 0; n < N; n++)
    {
        unsigned int t = s->tag[n];
        for (int d = 0; d < 4; d++)
        {
            if (pf)
                pf[4 * t + d] = s->pair_f[4 * n + d];
            if (bf)
                bf[4 * t + d] = s->bond_f[4 * n + d];
        }
        if (pv)
            for (int m = 0; m < 6; m++)
                pv[m * N + t] = s->pair_v[m * N + n];
    }
}

uint64_t os_timestep(const os_system* s) { return s->timestep; }
uint64_t os_n_builds(const os_system* s) { return s->n_builds; }

void os_set_nlist_policy(os_system* s, unsigned int rebuild_check_delay, int check_dist)
{
    s->check_delay = rebuild_check_delay;
    s->dist_check = check_dist != 0;
}

uint64_t os_n_dangerous(const os_system* s) { return s->n_dangerous; }

unsigned int os_shortest_rebuild(const os_system* s)
{
    /* NeighborList::getSmallestRebuild (hoomd/md/NeighborList.cc:1428-1436) */
    for (unsigned int i = 0; i < 100; i++)
        if (s->update_periods[i] != 0)
            return i;
    return 100;
}
double os_kinetic_energy(const os_system* s) { return oi_kinetic_energy(s->N, s->vel, s->mass); }

double os_potential_energy(const os_system* s)
{
    double pe = 0.0;
    for (size_t n = 0; n < s->N; n++)
        pe += s->net_f[4 * n + 3];
    return pe;
}
