// pair_ref.h -- the pair-force kernel template behind the reference-layout driver
// (include/hoomd_amd_pair.h) and the macro that instantiates it for one evaluator.
//
// A plugin adds an evaluator without touching the library: in its own .hip file
//
//     #include "pair_ref.h"
//     template<typename Real> struct EvaluatorPairMine { ... };   // the evaluator concept
//     HM_PAIR_REF_EXPORT(mine, EvaluatorPairMine)
//
// and builds a shared library of that one file (hipcc --offload-arch=gfx950 -shared -fPIC with
// -I pointing at include/ and csrc/). This is the role of the explicit instantiation
//     template hipError_t gpu_compute_pair_forces<EvaluatorPairExample>(...)
// in the reference's example_plugins/pair_plugin/PotentialPairExampleGPUKernel.cu and
// hoomd/md/PotentialPairGPUKernel.cu.inc:19-21.
//
// The evaluator concept is the reference's (hoomd/md/EvaluatorPairLJ.h:35-94), templated on the
// arithmetic type Real (float here):
//   struct param_type { ... };                      the layout of d_params on the device (the
//                                                   reference's param_type, e.g. Scalar members)
//   Evaluator(Real rsq, Real rcutsq, const param_type&)
//   static bool needsCharge(); void setCharge(Real qi, Real qj);
//   bool evalForceAndEnergy(Real& force_divr, Real& pair_eng, bool energy_shift)
//
// Kernel (one group of TPP lanes per particle, as gpu_compute_pair_forces_shared_kernel,
// hoomd/md/PotentialPairGPU.cuh:140-404): separation and minimum image in FP64 (BoxDim::minImage,
// hoomd/BoxDim.h:317-342, tilt factors included), the evaluator in FP32, energy shift when
// shift_mode is 1 or when it is 2 with r_on > r_cut, XPLOR smoothing for r_on <= r < r_cut,
// FP32 sums of force, energy and virial (virial with force / 2r), halved energy, group sums by
// wavefront shuffles, the FP32 sums widened to Scalar4 / Scalar at the store.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "hoomd_amd.h"
#include "hoomd_amd_pair.h"

namespace hm
{
namespace pref
{

struct PrefBox
{
    double L[3], Linv[3];
    double xy, xz, yz;
    int px, py, pz;
};

inline PrefBox make_pref_box(const hm_box_ref& b)
{
    PrefBox r;
    for (int k = 0; k < 3; k++)
    {
        r.L[k] = b.L[k];
        r.Linv[k] = 1.0 / b.L[k];
    }
    r.xy = b.xy;
    r.xz = b.xz;
    r.yz = b.yz;
    r.px = b.periodic[0];
    r.py = b.periodic[1];
    r.pz = b.periodic[2];
    return r;
}

// BoxDim::minImage, device form (hoomd/BoxDim.h:322-342)
__device__ __forceinline__ void min_image_ref(const PrefBox& b, double& x, double& y, double& z)
{
#pragma clang fp contract(off)
    if (b.pz)
    {
        const double img = rint(z * b.Linv[2]);
        z -= b.L[2] * img;
        y -= b.L[2] * b.yz * img;
        x -= b.L[2] * b.xz * img;
    }
    if (b.py)
    {
        const double img = rint(y * b.Linv[1]);
        y -= b.L[1] * img;
        x -= b.L[1] * b.xy * img;
    }
    if (b.px)
        x -= b.L[0] * rint(x * b.Linv[0]);
}

__device__ __forceinline__ int type_bits(double w)
{
    // __scalar_as_int (hoomd/HOOMDMath.h:205-215): the int lies in the low 32 bits
    return (int)(uint32_t)((unsigned long long)__double_as_longlong(w));
}

template<class Eval, int SHIFT, bool VIRIAL, unsigned int TPP>
__global__ __launch_bounds__(1024) void pair_ref_kernel(hm_pair_ref_args a,
                                                        const typename Eval::param_type* params,
                                                        PrefBox box)
{
#pragma clang fp contract(off)
    const unsigned int g = blockIdx.x * blockDim.x + threadIdx.x;
    const unsigned int i = g / TPP;
    const unsigned int sub = g % TPP;
    float fx = 0.f, fy = 0.f, fz = 0.f, en = 0.f;
    float vxx = 0.f, vxy = 0.f, vxz = 0.f, vyy = 0.f, vyz = 0.f, vzz = 0.f;
    const bool active = i < a.N;
    if (active)
    {
        const double4 pi = reinterpret_cast<const double4*>(a.d_pos)[i];
        const unsigned int ti = (unsigned int)type_bits(pi.w);
        float qi = 0.f;
        if (Eval::needsCharge())
            qi = (float)a.d_charge[i];
        const size_t head = a.d_head_list[i];
        const unsigned int nn = a.d_n_neigh[i];
        for (unsigned int k = sub; k < nn; k += TPP)
        {
            const unsigned int j = a.d_nlist[head + k];
            const double4 pj = reinterpret_cast<const double4*>(a.d_pos)[j];
            double dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
            min_image_ref(box, dx, dy, dz);
            const double rsq = dx * dx + dy * dy + dz * dz;
            const unsigned int tp = ti * a.ntypes + (unsigned int)type_bits(pj.w);
            const double rcutsq = a.d_rcutsq[tp];
            const double ronsq = SHIFT == 2 ? a.d_ronsq[tp] : 0.0;
            const bool energy_shift = SHIFT == 1 || (SHIFT == 2 && ronsq > rcutsq);
            float fdr = 0.f, pe = 0.f;
            // the cutoff test in FP64 (the evaluators test rsq < rcutsq; FP32 rounding of a
            // separation near the cutoff must not move a pair across it)
            if (rsq < rcutsq)
            {
                const float rsqf = (float)rsq, rcf = (float)rcutsq;
                Eval ev(rsqf, rcf, params[tp]);
                if (Eval::needsCharge())
                    ev.setCharge(qi, (float)a.d_charge[j]);
                ev.evalForceAndEnergy(fdr, pe, energy_shift);
                if (SHIFT == 2 && rsq >= ronsq)
                {
                    // XPLOR smoothing (PotentialPairGPU.cuh:326-345)
                    const float ron = (float)ronsq;
                    const float den = 1.f / ((rcf - ron) * (rcf - ron) * (rcf - ron));
                    const float rmc = rsqf - rcf;
                    const float s = rmc * rmc * (rcf + 2.f * rsqf - 3.f * ron) * den;
                    const float ds = 12.f * (rsqf - ron) * rmc * den;
                    const float e0 = pe;
                    pe = e0 * s;
                    fdr = s * fdr - ds * e0;
                }
            }
            const float ddx = (float)dx, ddy = (float)dy, ddz = (float)dz;
            if (VIRIAL)
            {
                const float h = 0.5f * fdr;
                vxx += ddx * ddx * h;
                vxy += ddx * ddy * h;
                vxz += ddx * ddz * h;
                vyy += ddy * ddy * h;
                vyz += ddy * ddz * h;
                vzz += ddz * ddz * h;
            }
            fx += ddx * fdr;
            fy += ddy * fdr;
            fz += ddz * fdr;
            en += pe;
        }
        en *= 0.5f;
    }
    for (unsigned int o = TPP / 2; o > 0; o >>= 1)
    {
        fx += __shfl_xor(fx, o, 64);
        fy += __shfl_xor(fy, o, 64);
        fz += __shfl_xor(fz, o, 64);
        en += __shfl_xor(en, o, 64);
        if (VIRIAL)
        {
            vxx += __shfl_xor(vxx, o, 64);
            vxy += __shfl_xor(vxy, o, 64);
            vxz += __shfl_xor(vxz, o, 64);
            vyy += __shfl_xor(vyy, o, 64);
            vyz += __shfl_xor(vyz, o, 64);
            vzz += __shfl_xor(vzz, o, 64);
        }
    }
    if (active && sub == 0)
    {
        reinterpret_cast<double4*>(a.d_force)[i] = make_double4(fx, fy, fz, en);
        if (VIRIAL)
        {
            const size_t p = a.virial_pitch;
            a.d_virial[0 * p + i] = vxx;
            a.d_virial[1 * p + i] = vxy;
            a.d_virial[2 * p + i] = vxz;
            a.d_virial[3 * p + i] = vyy;
            a.d_virial[4 * p + i] = vyz;
            a.d_virial[5 * p + i] = vzz;
        }
    }
}

template<class Eval, int SHIFT, bool VIRIAL>
hipError_t launch_tpp(const hm_pair_ref_args& a, const void* params, hipStream_t s)
{
    const PrefBox box = make_pref_box(a.box);
    const auto* p = static_cast<const typename Eval::param_type*>(params);
    const unsigned int tpp = a.threads_per_particle;
    const dim3 grid((unsigned int)(((size_t)a.N * tpp + a.block_size - 1) / a.block_size));
    const dim3 block(a.block_size);
    switch (tpp)
    {
    case 1: hipLaunchKernelGGL((pair_ref_kernel<Eval, SHIFT, VIRIAL, 1>), grid, block, 0, s, a, p, box); break;
    case 2: hipLaunchKernelGGL((pair_ref_kernel<Eval, SHIFT, VIRIAL, 2>), grid, block, 0, s, a, p, box); break;
    case 4: hipLaunchKernelGGL((pair_ref_kernel<Eval, SHIFT, VIRIAL, 4>), grid, block, 0, s, a, p, box); break;
    case 8: hipLaunchKernelGGL((pair_ref_kernel<Eval, SHIFT, VIRIAL, 8>), grid, block, 0, s, a, p, box); break;
    case 16: hipLaunchKernelGGL((pair_ref_kernel<Eval, SHIFT, VIRIAL, 16>), grid, block, 0, s, a, p, box); break;
    case 32: hipLaunchKernelGGL((pair_ref_kernel<Eval, SHIFT, VIRIAL, 32>), grid, block, 0, s, a, p, box); break;
    default: hipLaunchKernelGGL((pair_ref_kernel<Eval, SHIFT, VIRIAL, 64>), grid, block, 0, s, a, p, box); break;
    }
    return hipGetLastError();
}

// gpu_compute_pair_forces (PotentialPairGPU.cuh:579-656): shift mode and virial flag select the
// kernel instance. Returns HM_OK, HM_ERR_ARGS or HM_ERR_LAUNCH.
template<class Eval>
int compute_pair_forces(const hm_pair_ref_args* a, const void* params, void* stream)
{
    if (!a || !params || !a->d_force || (a->N && (!a->d_pos || !a->d_n_neigh || !a->d_nlist
                                                   || !a->d_head_list || !a->d_rcutsq))
        || (a->shift_mode == 2 && !a->d_ronsq) || a->shift_mode > 2 || a->ntypes == 0
        || (a->compute_virial && (!a->d_virial || a->virial_pitch < a->N))
        || (Eval::needsCharge() && a->N && !a->d_charge))
        return HM_ERR_ARGS;
    const unsigned int tpp = a->threads_per_particle;
    if (tpp == 0 || tpp > 64 || (tpp & (tpp - 1)) || a->block_size == 0 || a->block_size > 1024
        || a->block_size % 64)
        return HM_ERR_ARGS;
    for (int k = 0; k < 3; k++)
        if (!(a->box.L[k] > 0.0))
            return HM_ERR_ARGS;
    if (a->N == 0)
        return HM_OK;
    hipStream_t s = (hipStream_t)stream;
    hipError_t e;
    if (a->compute_virial)
        e = a->shift_mode == 0   ? launch_tpp<Eval, 0, true>(*a, params, s)
            : a->shift_mode == 1 ? launch_tpp<Eval, 1, true>(*a, params, s)
                                 : launch_tpp<Eval, 2, true>(*a, params, s);
    else
        e = a->shift_mode == 0   ? launch_tpp<Eval, 0, false>(*a, params, s)
            : a->shift_mode == 1 ? launch_tpp<Eval, 1, false>(*a, params, s)
                                 : launch_tpp<Eval, 2, false>(*a, params, s);
    return e == hipSuccess ? HM_OK : HM_ERR_LAUNCH;
}

} // namespace pref
} // namespace hm

// One exported driver per evaluator template: hm_compute_pair_forces_<name>.
#define HM_PAIR_REF_EXPORT(name, EvalTemplate)                                                    \
    extern "C" __attribute__((visibility("default"))) int hm_compute_pair_forces_##name(        \
        const hm_pair_ref_args* args, const void* d_params, void* stream)                        \
    {                                                                                             \
        return hm::pref::compute_pair_forces<EvalTemplate<float>>(args, d_params, stream);       \
    }
