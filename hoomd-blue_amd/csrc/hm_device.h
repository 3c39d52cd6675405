// hm_device.h -- device-side building blocks shared by the gfx950 kernels.
//
// Box geometry (BoxDim subset), Philox4x32-10 with HOOMD's Seed/Counter packing and
// distributions, wavefront-64 helpers and the XCD-aware block remap. Everything here is
// HOST+DEVICE so the host-side Bussi draw and the device kernels share one implementation.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "hoomd_amd.h"

#define HM_HD __host__ __device__ __forceinline__

namespace hm
{

// ------------------------------------------------------------------------------------ box
// BoxDim::setL (hoomd/BoxDim.h:186-200): hi = L/2, lo = -hi, Linv = 1/L.
struct DevBox
{
    double L[3];
    double lo[3];
    double hi[3];
    double Linv[3];
};

inline DevBox make_devbox(const hm_box& b)
{
    DevBox d;
    for (int k = 0; k < 3; k++)
    {
        d.L[k] = b.L[k];
        d.hi[k] = b.L[k] / 2.0;
        d.lo[k] = -d.hi[k];
        d.Linv[k] = 1.0 / b.L[k];
    }
    return d;
}

// Minimum image, device form of BoxDim::minImage (hoomd/BoxDim.h:322-342): w -= L*rint(w/L).
// For |w| < L this equals the host branch form bit for bit (the shift is 0 or exactly L).
HM_HD void min_image(const DevBox& b, double& x, double& y, double& z)
{
#pragma clang fp contract(off)
    z -= b.L[2] * rint(z * b.Linv[2]);
    y -= b.L[1] * rint(y * b.Linv[1]);
    x -= b.L[0] * rint(x * b.Linv[0]);
}

// BoxDim::wrap (hoomd/BoxDim.h:407-475) for zero tilt: at most one image shift per axis.
HM_HD void wrap_axis(double& w, int& img, double lo, double hi, double L)
{
#pragma clang fp contract(off)
    if (w >= hi)
    {
        w -= L;
        img++;
    }
    else if (w < lo)
    {
        w += L;
        img--;
    }
}

// ------------------------------------------------------------------------------------ Philox
// Philox4x32-10 (Salmon et al., SC'11): multipliers 0xD2511F53 / 0xCD9E8D57, key increments
// 0x9E3779B9 / 0xBB67AE85, ten rounds. HOOMD packs Seed/Counter as in
// hoomd/RandomNumbers.h:150-188 and advances ctr[0] after each draw (:264-270).
struct Philox
{
    uint32_t key[2];
    uint32_t ctr[4];

    HM_HD Philox(uint8_t id, uint64_t timestep, uint16_t seed, uint32_t a, uint32_t b = 0,
                 uint32_t c = 0, uint16_t d = 0)
    {
        key[0] = (uint32_t(id) << 24) | (uint32_t(seed) << 8)
                 | uint32_t((timestep & 0x000000ff00000000ull) >> 32);
        key[1] = uint32_t(timestep & 0x00000000ffffffffull);
        ctr[0] = uint32_t(d) << 16;
        ctr[1] = c;
        ctr[2] = b;
        ctr[3] = a;
    }

    HM_HD static void round(uint32_t c[4], uint32_t k0, uint32_t k1)
    {
        uint64_t p0 = uint64_t(0xD2511F53u) * c[0];
        uint64_t p1 = uint64_t(0xCD9E8D57u) * c[2];
        uint32_t n0 = uint32_t(p1 >> 32) ^ c[1] ^ k0;
        uint32_t n2 = uint32_t(p0 >> 32) ^ c[3] ^ k1;
        c[0] = n0;
        c[1] = uint32_t(p1);
        c[2] = n2;
        c[3] = uint32_t(p0);
    }

    HM_HD void next(uint32_t out[4])
    {
        uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
        uint32_t k0 = key[0], k1 = key[1];
#pragma unroll
        for (int r = 0; r < 10; r++)
        {
            if (r > 0)
            {
                k0 += 0x9E3779B9u;
                k1 += 0xBB67AE85u;
            }
            round(c, k0, k1);
        }
        out[0] = c[0];
        out[1] = c[1];
        out[2] = c[2];
        out[3] = c[3];
        ctr[0] += 1;
    }

    HM_HD uint64_t u64()
    {
        uint32_t u[4];
        next(u);
        return (uint64_t(u[0]) << 32) | u[1];
    }
};

// Random123 u01 / uneg11 for 64-bit inputs (hoomd/RandomNumbers.h:72-114).
HM_HD double u01(uint64_t in)
{
#pragma clang fp contract(off)
    const double factor = 1.0 / 18446744073709551616.0;
    return double(in) * factor + 0.5 * factor;
}

HM_HD double uneg11(uint64_t in)
{
#pragma clang fp contract(off)
    const double factor = 1.0 / 9223372036854775808.0;
    return double(int64_t(in)) * factor + 0.5 * factor;
}

// UniformDistribution (RandomNumbers.h:302-330)
HM_HD double uniform(Philox& g, double a, double b)
{
#pragma clang fp contract(off)
    double w = b - a;
    return a + w * u01(g.u64());
}

// NormalDistribution, one value per draw (RandomNumbers.h:353-365). sincospi(x) is evaluated
// as sin(pi x), the form the reference host build uses (HOOMDMath.h:330-338).
HM_HD double normal(Philox& g, double sigma, double mu)
{
#pragma clang fp contract(off)
    uint32_t u[4];
    g.next(u);
    uint64_t u0 = (uint64_t(u[0]) << 32) | u[1];
    uint64_t u1 = (uint64_t(u[2]) << 32) | u[3];
    double x = sin(M_PI * uneg11(u0));
    double r = sqrt(-2.0 * log(u01(u1)));
    x *= r;
    return x * sigma + mu;
}

// GammaDistribution, Marsaglia-Tsang with squeeze (RandomNumbers.h:452-515). The rejection
// loop is capped; the acceptance rate is above 0.95 per iteration for alpha >= 1.
HM_HD double gamma_dist(Philox& g, double m_alpha, double m_b)
{
#pragma clang fp contract(off)
    if (m_alpha <= 0)
        return 0.0;
    double alpha = m_alpha;
    if (m_alpha < 1.0)
        alpha += 1.0;
    double d = alpha - 1.0 / 3.0;
    double c = (1.0 / sqrt(d)) / 3.0;
    double v = 1.0;
    for (int it = 0; it < 4096; it++)
    {
        double x;
        int inner = 0;
        do
        {
            x = normal(g, 1.0, 0.0);
            v = 1.0 + c * x;
        } while (v <= 0.0 && ++inner < 4096);
        v = v * v * v;
        const double x2 = x * x;
        double u = u01(g.u64());
        if (u < 1.0 - 0.0331 * x2 * x2)
            break;
        if (log(u) < 0.5 * x2 + d * (1.0 - v + log(v)))
            break;
    }
    v *= d * m_b;
    if (m_alpha < 1.0)
        v *= pow(u01(g.u64()), 1.0 / m_alpha);
    return v;
}

// BussiThermostat::compute_rescale_factor (hoomd/md/Thermostat.h:416-446).
HM_HD double bussi_alpha(double K, double ndof, double dt, double tau, double kT,
                         uint64_t timestep, uint16_t seed, uint32_t instance)
{
#pragma clang fp contract(off)
    if (dt == 0.0 || ndof == 0)
        return 1.0;
    Philox g(45, timestep, seed, instance);
    double c = 0.0;
    if (tau != 0.0)
        c = exp(-dt / tau);
    double r1 = normal(g, 1.0, 0.0);
    double rg = 0.0;
    if (ndof > 1.0)
        rg = 2.0 * gamma_dist(g, (ndof - 1.0) / 2.0, 1.0);
    double v = kT / 2.0 / K;
    double term1 = v * (1.0 - c) * (rg + r1 * r1);
    double term2 = 2.0 * r1 * sqrt(v * (1.0 - c) * c);
    return sqrt(c + term1 + term2);
}

// ------------------------------------------------------------------------------------ wave64
__device__ __forceinline__ unsigned int lane_id()
{
    return __lane_id();
}

// number of set bits of mask below this lane
// Wave64 ballot of a lane predicate, taken straight from the compare mask (HIP's __ballot goes
// through an int and costs two extra VALU instructions per call)
__device__ __forceinline__ uint64_t wave_ballot(bool p)
{
    return __builtin_amdgcn_ballot_w64(p);
}

__device__ __forceinline__ unsigned int lanes_below(uint64_t mask)
{
    return __builtin_amdgcn_mbcnt_hi(uint32_t(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo(uint32_t(mask), 0u));
}

// XCD-aware block remap: hardware dispatches workgroup b to XCD (b % 8); give each XCD a
// contiguous range of logical blocks so neighbouring particles share one L2. Requires the
// grid to be a multiple of 8.
__device__ __forceinline__ unsigned int xcd_block(unsigned int b, unsigned int nblocks)
{
    return (b & 7u) * (nblocks >> 3) + (b >> 3);
}

inline unsigned int round_up(unsigned int x, unsigned int m)
{
    return (x + m - 1) / m * m;
}

} // namespace hm
