// capi_util.h -- helpers shared by the C-ABI drivers and the system engine.
#pragma once

#include <cstring>

#include "kernels.h"

namespace hm
{

int fail(int code, const char* what);
int check(hipError_t e, const char* what);

// NeighborList::distanceCheck threshold for an unchanged box (hoomd/md/NeighborList.cc:1243-1252):
// delta_max = ((r_cut_max + r_buff) * 1 - r_cut_max) / 2, squared when positive.
inline double distance_check_maxsq(double rcut_max, double r_buff)
{
    double rmax = rcut_max + r_buff;
    double delta_max = (rmax * 1.0 - rcut_max) / 2.0;
    return delta_max > 0 ? delta_max * delta_max : 0.0;
}

inline int ilog2(unsigned int x)
{
    int r = 0;
    while ((1u << r) < x)
        r++;
    return r;
}

inline void fill_pair_args(PairKernelArgs& k, unsigned int N, const double4* pos, const hm_box& box,
                           const unsigned int* nlist, unsigned int nmax,
                           const unsigned int* n_neigh, unsigned int ntypes, const double* r_cut,
                           const double* r_on, int shift_mode, int compute_virial, float4* force,
                           float* virial)
{
    memset(&k, 0, sizeof(k));
    k.N = N;
    k.pos = pos;
    k.box = make_devbox(box);
    k.nlist = nlist;
    k.nmax = nmax;
    k.n_neigh = n_neigh;
    k.ntypes = ntypes;
    for (unsigned int q = 0; q < ntypes * ntypes; q++)
    {
        double rc = r_cut[q] > 0.0 ? r_cut[q] : 0.0;
        double ro = r_on ? r_on[q] : 0.0;
        k.rcutsq[q] = rc * rc;
        k.ronsq[q] = ro * ro;
        k.rcutsq_f[q] = (float)(rc * rc);
        k.ronsq_f[q] = (float)(ro * ro);
    }
    k.shift_mode = shift_mode;
    k.compute_virial = compute_virial;
    k.force = force;
    k.virial = virial;
}

} // namespace hm
