/*
 * nve_precision_body.h -- one precision variant of the NVE precision study (TEST INFRASTRUCTURE
 * ONLY; included by nve_precision.c once per variant).
 *
 * Before inclusion define:
 *   PX       the state type: positions, velocities, accelerations, wrapping, the list distance
 *            tests and the velocity-Verlet update (the reference's Scalar / LongReal)
 *   PF       the pair type: the pair separation after the minimum image, r^2, the LJ
 *            evaluation and the per-particle force sum
 *   VN(x)    pastes the variant's suffix onto x
 */

/* ob_wrap for a cubic box (hoomd/BoxDim.h:330-380 restated for one length) */
static void VN(wrap)(PX* x, PX L)
{
    const PX h = L / (PX)2;
    if (*x >= h)
        *x -= L;
    else if (*x < -h)
        *x += L;
}

/* nearest periodic image of a separation, cubic box */
static PX VN(min_image)(PX d, PX L, PX Linv)
{
    return d - L * (PX)rint((double)(d * Linv));
}

/* binned full list over the 27 neighbour cells; rows in cell order, particles of a cell in
 * index order, so the list does not depend on the thread count. Returns the longest row. */
static unsigned int VN(build)(nvp_state* s, const PX* pos)
{
    const size_t N = s->N;
    const unsigned int D = s->dim;
    const PX L = (PX)s->L, Linv = (PX)1 / (PX)s->L;
    const PX rl2 = (PX)(s->rlist * s->rlist);
    const double scale = (double)D / s->L;
    memset(s->cell_start, 0, sizeof(unsigned int) * ((size_t)D * D * D + 1));
    for (size_t i = 0; i < N; i++)
    {
        int c[3];
        for (int d = 0; d < 3; d++)
        {
            int k = (int)floor(((double)pos[3 * i + d] + 0.5 * s->L) * scale);
            c[d] = k < 0 ? 0 : (k >= (int)D ? (int)D - 1 : k);
        }
        s->cell_of[i] = ((unsigned int)c[0] * D + (unsigned int)c[1]) * D + (unsigned int)c[2];
        s->cell_start[s->cell_of[i] + 1]++;
    }
    for (size_t c = 0; c < (size_t)D * D * D; c++)
        s->cell_start[c + 1] += s->cell_start[c];
    memcpy(s->cell_fill, s->cell_start, sizeof(unsigned int) * (size_t)D * D * D);
    for (size_t i = 0; i < N; i++)
        s->cell_members[s->cell_fill[s->cell_of[i]]++] = (unsigned int)i;

    unsigned int longest = 0;
#pragma omp parallel for num_threads(s->threads) schedule(static) reduction(max : longest)
    for (size_t i = 0; i < N; i++)
    {
        unsigned int c = s->cell_of[i];
        int cx = (int)(c / (D * D)), cy = (int)((c / D) % D), cz = (int)(c % D);
        unsigned int n = 0;
        for (int ox = -1; ox <= 1; ox++)
            for (int oy = -1; oy <= 1; oy++)
                for (int oz = -1; oz <= 1; oz++)
                {
                    unsigned int nc = ((unsigned int)((cx + ox + (int)D) % (int)D) * D
                                       + (unsigned int)((cy + oy + (int)D) % (int)D))
                                          * D
                                      + (unsigned int)((cz + oz + (int)D) % (int)D);
                    for (unsigned int q = s->cell_start[nc]; q < s->cell_start[nc + 1]; q++)
                    {
                        unsigned int j = s->cell_members[q];
                        if (j == i)
                            continue;
                        PX dx = VN(min_image)(pos[3 * i] - pos[3 * j], L, Linv);
                        PX dy = VN(min_image)(pos[3 * i + 1] - pos[3 * j + 1], L, Linv);
                        PX dz = VN(min_image)(pos[3 * i + 2] - pos[3 * j + 2], L, Linv);
                        if (dx * dx + dy * dy + dz * dz < rl2)
                        {
                            if (n < s->nmax)
                                s->nlist[i * s->nmax + n] = j;
                            n++;
                        }
                    }
                }
        s->n_neigh[i] = n;
        if (n > longest)
            longest = n;
    }
    return longest;
}

/* LJ forces over the full list (EvaluatorPairLJ.h:101-140 with mode "none"), no third law */
static void VN(forces)(nvp_state* s, const PX* pos, PX* accel)
{
    const PX L = (PX)s->L, Linv = (PX)1 / (PX)s->L;
    const PF rcsq = (PF)(s->rcut * s->rcut);
    const PF lj1 = (PF)4, lj2 = (PF)4; /* 4 eps sigma^12, 4 eps sigma^6 at eps = sigma = 1 */
#pragma omp parallel for num_threads(s->threads) schedule(static)
    for (size_t i = 0; i < s->N; i++)
    {
        PF fx = 0, fy = 0, fz = 0;
        const unsigned int* row = s->nlist + i * s->nmax;
        for (unsigned int q = 0; q < s->n_neigh[i]; q++)
        {
            unsigned int j = row[q];
            PF dx = (PF)VN(min_image)(pos[3 * i] - pos[3 * j], L, Linv);
            PF dy = (PF)VN(min_image)(pos[3 * i + 1] - pos[3 * j + 1], L, Linv);
            PF dz = (PF)VN(min_image)(pos[3 * i + 2] - pos[3 * j + 2], L, Linv);
            PF rsq = dx * dx + dy * dy + dz * dz;
            if (rsq < rcsq)
            {
                PF r2inv = (PF)1 / rsq;
                PF r6inv = r2inv * r2inv * r2inv;
                PF fdr = r2inv * r6inv * ((PF)12 * lj1 * r6inv - (PF)6 * lj2);
                fx += dx * fdr;
                fy += dy * fdr;
                fz += dz * fdr;
            }
        }
        accel[3 * i] = (PX)fx; /* mass 1 */
        accel[3 * i + 1] = (PX)fy;
        accel[3 * i + 2] = (PX)fz;
    }
}

/* NeighborList::needsUpdating with rebuild_check_delay 1 and check_dist (NeighborList.cc:1325-
 * 1418, distanceCheck :1238-1300): rebuild when a particle moved r_buff / 2 since the last build */
static int VN(needs_build)(nvp_state* s, const PX* pos, const PX* last)
{
    const PX L = (PX)s->L, Linv = (PX)1 / (PX)s->L;
    const PX lim = (PX)(0.25 * s->rbuff * s->rbuff);
    int hit = 0;
#pragma omp parallel for num_threads(s->threads) schedule(static) reduction(| : hit)
    for (size_t i = 0; i < s->N; i++)
    {
        PX dx = VN(min_image)(pos[3 * i] - last[3 * i], L, Linv);
        PX dy = VN(min_image)(pos[3 * i + 1] - last[3 * i + 1], L, Linv);
        PX dz = VN(min_image)(pos[3 * i + 2] - last[3 * i + 2], L, Linv);
        if (dx * dx + dy * dy + dz * dz >= lim)
            hit = 1;
    }
    return hit;
}

static int VN(list)(nvp_state* s, const PX* pos, PX* last)
{
    for (;;)
    {
        unsigned int need = VN(build)(s, pos);
        if (need <= s->nmax)
            break;
        s->nmax = (need + 7) & ~7u;
        free(s->nlist);
        s->nlist = (unsigned int*)malloc(sizeof(unsigned int) * s->N * s->nmax);
        if (!s->nlist)
            return -1;
    }
    memcpy(last, pos, sizeof(PX) * 3 * s->N);
    s->n_builds++;
    return 0;
}

/* kinetic and shifted LJ potential energy of the state, both summed in FP64 from the variant's
 * values (the measurement is the same for every variant) */
static void VN(energy)(nvp_state* s, const PX* pos, const PX* vel, double* ke, double* pe)
{
    const double L = s->L, Linv = 1.0 / s->L;
    const double rcsq = s->rcut * s->rcut;
    const double r6c = 1.0 / (rcsq * rcsq * rcsq);
    const double uc = 4.0 * r6c * (r6c - 1.0);
    double k = 0.0, p = 0.0;
#pragma omp parallel for num_threads(s->threads) schedule(static) reduction(+ : k, p)
    for (size_t i = 0; i < s->N; i++)
    {
        for (int d = 0; d < 3; d++)
            k += 0.5 * (double)vel[3 * i + d] * (double)vel[3 * i + d];
        const unsigned int* row = s->nlist + i * s->nmax;
        for (unsigned int q = 0; q < s->n_neigh[i]; q++)
        {
            unsigned int j = row[q];
            double r[3];
            for (int d = 0; d < 3; d++)
            {
                double w = (double)pos[3 * i + d] - (double)pos[3 * j + d];
                r[d] = w - L * rint(w * Linv);
            }
            double rsq = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
            if (rsq < rcsq)
            {
                double r6 = 1.0 / (rsq * rsq * rsq);
                p += 0.5 * (4.0 * r6 * (r6 - 1.0) - uc);
            }
        }
    }
    *ke = k;
    *pe = p;
}

/* TwoStepConstantVolume without a thermostat (TwoStepConstantVolume.cc:7-80, 213-260) */
static int VN(run)(nvp_state* s, const double* pos0, const double* vel0, uint64_t nsteps,
                   uint64_t every, double* ke_out, double* pe_out, double* pos_out,
                   double* vel_out)
{
    const size_t N = s->N;
    const PX dt = (PX)s->dt, half = (PX)0.5, L = (PX)s->L;
    PX* pos = (PX*)malloc(sizeof(PX) * 3 * N);
    PX* vel = (PX*)malloc(sizeof(PX) * 3 * N);
    PX* acc = (PX*)malloc(sizeof(PX) * 3 * N);
    PX* last = (PX*)malloc(sizeof(PX) * 3 * N);
    int rc = -1;
    if (!pos || !vel || !acc || !last)
        goto done;
    for (size_t q = 0; q < 3 * N; q++)
    {
        pos[q] = (PX)pos0[q];
        vel[q] = (PX)vel0[q];
    }
    if (VN(list)(s, pos, last))
        goto done;
    VN(forces)(s, pos, acc);
    VN(energy)(s, pos, vel, &ke_out[0], &pe_out[0]);
    for (uint64_t t = 1; t <= nsteps; t++)
    {
#pragma omp parallel for num_threads(s->threads) schedule(static)
        for (size_t q = 0; q < 3 * N; q++)
        {
            PX v = vel[q] + half * acc[q] * dt;
            pos[q] += dt * v;
            vel[q] = v;
            VN(wrap)(&pos[q], L);
        }
        if (VN(needs_build)(s, pos, last) && VN(list)(s, pos, last))
            goto done;
        VN(forces)(s, pos, acc);
#pragma omp parallel for num_threads(s->threads) schedule(static)
        for (size_t q = 0; q < 3 * N; q++)
            vel[q] += half * dt * acc[q];
        if (t % every == 0)
            VN(energy)(s, pos, vel, &ke_out[t / every], &pe_out[t / every]);
    }
    if (pos_out)
        for (size_t q = 0; q < 3 * N; q++)
            pos_out[q] = (double)pos[q];
    if (vel_out)
        for (size_t q = 0; q < 3 * N; q++)
            vel_out[q] = (double)vel[q];
    rc = 0;
done:
    free(pos);
    free(vel);
    free(acc);
    free(last);
    return rc;
}
