// evaluators.h -- pair and bond evaluators (the PotentialPair / PotentialBond plugin concept).
//
// A pair evaluator has the shape the reference defines in hoomd/md/EvaluatorPairLJ.h:35-94:
//   * a POD param_type built on the host from user parameters,
//   * a constructor (rsq, rcutsq, const param_type&),
//   * needsCharge() / setCharge(qi, qj),
//   * bool evalForceAndEnergy(Real& force_divr, Real& energy, bool energy_shift),
//   * getName().
// The pair kernel template (pair_kernel.h) is instantiated once per evaluator, as
// PotentialPairGPUKernel.cu.inc:19-21 does per evaluator name. Real is the arithmetic type
// (float on the device hot path).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace hm
{

// 1 / x. On the device in FP32 this is the hardware reciprocal v_rcp_f32 (at most 1 ulp from
// the correctly rounded quotient), one instruction instead of the IEEE division sequence.
__host__ __device__ __forceinline__ double recip(double x)
{
    return 1.0 / x;
}
__host__ __device__ __forceinline__ float recip(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rcpf(x);
#else
    return 1.0f / x;
#endif
}

template<typename Real> struct EvaluatorPairLJ
{
    // EvaluatorPairLJ::param_type (hoomd/md/EvaluatorPairLJ.h:99-155): (sigma^6, 4 epsilon)
    struct param_type
    {
        Real sigma_6;
        Real epsilon_x_4;

        static param_type make(double sigma, double epsilon)
        {
            param_type p;
            double s6 = sigma * sigma * sigma * sigma * sigma * sigma;
            p.sigma_6 = Real(s6);
            p.epsilon_x_4 = Real(4.0 * epsilon);
            return p;
        }
    };

    __host__ __device__ __forceinline__ EvaluatorPairLJ(Real _rsq, Real _rcutsq,
                                                         const param_type& p)
        : rsq(_rsq), rcutsq(_rcutsq), lj1(p.epsilon_x_4 * p.sigma_6 * p.sigma_6),
          lj2(p.epsilon_x_4 * p.sigma_6)
    {
    }

    __host__ __device__ static constexpr bool needsCharge()
    {
        return false;
    }
    __host__ __device__ void setCharge(Real, Real) { }

    // EvaluatorPairLJ::evalForceAndEnergy (hoomd/md/EvaluatorPairLJ.h:190-211); the caller has
    // already applied the exact (FP64) cutoff test, lj1 == 0 disables the pair.
    __host__ __device__ __forceinline__ bool evalForceAndEnergy(Real& force_divr, Real& pair_eng,
                                                                bool energy_shift) const
    {
        if (lj1 == Real(0))
            return false;
        Real r2inv = recip(rsq);
        Real r6inv = r2inv * r2inv * r2inv;
        force_divr = r2inv * r6inv * (Real(12) * lj1 * r6inv - Real(6) * lj2);
        pair_eng = r6inv * (lj1 * r6inv - lj2);
        if (energy_shift)
        {
            Real rc2inv = recip(rcutsq);
            Real rc6inv = rc2inv * rc2inv * rc2inv;
            pair_eng -= rc6inv * (lj1 * rc6inv - lj2);
        }
        return true;
    }

    static const char* getName()
    {
        return "lj";
    }

    static constexpr int shift_modes = 3; // none, shift, xplor

    Real rsq, rcutsq, lj1, lj2;
};

// Tabulated pair potential (hoomd/md/EvaluatorPairTable.h:30-196): U and F = -dU/dr on width
// points r_min + k dr, dr = (r_cut - r_min) / width, interpolated linearly; the point after the
// last is zero, and the pair does not interact below r_min or from r_cut on. The tables are
// device arrays (the reference keeps them in ManagedArrays); param_type carries their address.
template<typename Real> struct EvaluatorPairTable
{
    struct param_type
    {
        const Real* U;
        const Real* F;
        Real rmin;
        unsigned int width;

        static param_type make(const Real* U, const Real* F, double rmin, unsigned int width)
        {
            param_type p;
            p.U = U;
            p.F = F;
            p.rmin = Real(rmin);
            p.width = width;
            return p;
        }
    };

    __host__ __device__ __forceinline__ EvaluatorPairTable(Real _rsq, Real _rcutsq,
                                                            const param_type& p)
        : rsq(_rsq), rcutsq(_rcutsq), rmin(p.rmin), width(p.width), U(p.U), F(p.F)
    {
    }

    __host__ __device__ static constexpr bool needsCharge()
    {
        return false;
    }
    __host__ __device__ void setCharge(Real, Real) { }

    // EvaluatorPairTable::evalForceAndEnergy (EvaluatorPairTable.h:152-196); tables do not shift.
    // The index stays below width: rounding can put r just under r_cut at width dr, where the
    // reference would read one element past its table.
    __host__ __device__ __forceinline__ bool evalForceAndEnergy(Real& force_divr, Real& pair_eng,
                                                                bool) const
    {
        const Real r = sqrt(rsq);
        if (!(rsq < rcutsq) || r < rmin || width == 0)
            return false;
        const Real rcut = sqrt(rcutsq);
        const Real delta_r = (rcut - rmin) / Real(width);
        const Real value_f = (r - rmin) / delta_r;
        unsigned int k = (unsigned int)floor(value_f);
        k = k < width ? k : width - 1;
        const Real U0 = U[k], F0 = F[k];
        Real U1 = Real(0), F1 = Real(0);
        if (k + 1 < width)
        {
            U1 = U[k + 1];
            F1 = F[k + 1];
        }
        const Real f = value_f - Real(k);
        const Real Ur = U0 + f * (U1 - U0);
        const Real Fr = F0 + f * (F1 - F0);
        if (rsq > Real(0))
            force_divr = Fr / r;
        pair_eng = Ur;
        return true;
    }

    static const char* getName()
    {
        return "table";
    }

    static constexpr int shift_modes = 1; // none only

    Real rsq, rcutsq, rmin;
    unsigned int width;
    const Real *U, *F;
};

// Fast form of the pair-evaluator concept for the FP32 gather kernels. An entry holds the
// per-type-pair constants, built once on the host from param_type; eval returns force / r and
// half the pair energy, both exactly zero when the pair is not `in` range. The generic form
// wraps the standard concept (constructor + evalForceAndEnergy), so any evaluator works;
// evaluators may specialise it with the constants hoisted out of the per-pair work.
template<class Eval> struct FastPair
{
    struct entry
    {
        typename Eval::param_type p;
        float rcutsq;
        int shift;
    };
    static entry make(const typename Eval::param_type& p, double rcutsq, int shift)
    {
        entry e;
        e.p = p;
        e.rcutsq = (float)rcutsq;
        e.shift = shift;
        return e;
    }
    template<bool SHIFTED = true>
    __device__ __forceinline__ static void eval(float rsq, bool in, const entry& t, float& fdr,
                                                float& ehalf)
    {
        Eval ev(in ? rsq : 1.0f, t.rcutsq, t.p);
        float f = 0.f, e = 0.f;
        const bool ok = ev.evalForceAndEnergy(f, e, t.shift != 0);
        fdr = (in && ok) ? f : 0.f;
        ehalf = (in && ok) ? 0.5f * e : 0.f;
    }
};

// LJ: lj1 = 4 eps sigma^12, lj2 = 4 eps sigma^6 (EvaluatorPairLJ.h:162-166) pre-multiplied by
// 12, 6 and 1/2, and the energy shift at r_cut (EvaluatorPairLJ.h:203-209) precomputed. A pair
// with lj1 == 0 never contributes (evalForceAndEnergy returns false); the caller gives it a zero
// cutoff. The reciprocal of an out-of-range pair is replaced by 0 before use, which zeroes
// force and energy without a division by zero.
template<> struct FastPair<EvaluatorPairLJ<float>>
{
    struct entry
    {
        float c12, c6, h1, h2, hcut;
    };
    static entry make(const EvaluatorPairLJ<float>::param_type& p, double rcutsq, int shift)
    {
        const double lj1 = (double)p.epsilon_x_4 * (double)p.sigma_6 * (double)p.sigma_6;
        const double lj2 = (double)p.epsilon_x_4 * (double)p.sigma_6;
        entry e;
        e.c12 = (float)(12.0 * lj1);
        e.c6 = (float)(6.0 * lj2);
        e.h1 = (float)(0.5 * lj1);
        e.h2 = (float)(0.5 * lj2);
        e.hcut = 0.f;
        if (shift && rcutsq > 0.0)
        {
            const double rc6inv = 1.0 / (rcutsq * rcutsq * rcutsq);
            e.hcut = (float)(0.5 * rc6inv * (lj1 * rc6inv - lj2));
        }
        return e;
    }
    template<bool SHIFTED = true>
    __device__ __forceinline__ static void eval(float rsq, bool in, const entry& t, float& fdr,
                                                float& ehalf)
    {
        const float r2inv = in ? recip(rsq) : 0.f;
        const float r6inv = r2inv * r2inv * r2inv;
        fdr = r2inv * r6inv * (t.c12 * r6inv - t.c6);
        ehalf = r6inv * (t.h1 * r6inv - t.h2);
        if (SHIFTED)
            ehalf -= in ? t.hcut : 0.f;
    }
};

// FENE + WCA bond evaluator (hoomd/md/EvaluatorBondFENE.h:30-156).
template<typename Real> struct EvaluatorBondFENE
{
    struct param_type
    {
        Real k, r_0, epsilon_x_4, sigma_6, delta;

        static param_type make(double k, double r0, double epsilon, double sigma, double delta)
        {
            param_type p;
            double s6 = sigma * sigma * sigma * sigma * sigma * sigma;
            p.k = Real(k);
            p.r_0 = Real(r0);
            p.epsilon_x_4 = Real(4.0 * epsilon);
            p.sigma_6 = Real(s6);
            p.delta = Real(delta);
            return p;
        }
    };

    __host__ __device__ __forceinline__ EvaluatorBondFENE(Real _rsq, const param_type& p)
        : rsq(_rsq), K(p.k), r_0(p.r_0), lj1(p.epsilon_x_4 * p.sigma_6 * p.sigma_6),
          lj2(p.epsilon_x_4 * p.sigma_6), delta(p.delta)
    {
    }

    __host__ __device__ __forceinline__ bool evalForceAndEnergy(Real& force_divr, Real& bond_eng)
    {
        Real r = sqrt(rsq);
        Real rtemp = r - delta;
        Real rmdoverr = rtemp / r;
        Real rsq2 = rtemp * rtemp;
        Real r2inv = Real(1) / rsq2;
        Real r6inv = r2inv * r2inv * r2inv;
        Real wca = Real(0);
        Real pair_eng = Real(0);
        Real sigma6inv = lj2 / lj1;
        Real epsilon = lj2 * lj2 / Real(4) / lj1;
        if (lj1 != Real(0) && r6inv > sigma6inv / Real(2))
        {
            wca = r2inv * r6inv * (Real(12) * lj1 * r6inv - Real(6) * lj2);
            pair_eng = r6inv * (lj1 * r6inv - lj2) + epsilon;
        }
        force_divr = wca * rmdoverr;
        bond_eng = pair_eng;
        if (K != Real(0))
        {
            Real r0sq = r_0 * r_0;
            if (rsq2 >= r0sq)
                return false;
            force_divr += -K / (Real(1) - rsq2 / r0sq) * rmdoverr;
            bond_eng += -Real(0.5) * K * r0sq * log(Real(1) - rsq2 / r0sq);
        }
        return true;
    }

    static const char* getName()
    {
        return "fene";
    }

    Real rsq, K, r_0, lj1, lj2, delta;
};

// Harmonic bond evaluator (hoomd/md/EvaluatorBondHarmonic.h:37-140): U = k (r - r0)^2 / 2,
// force / r = k (r0 / r - 1), set to 0 when not finite (r = 0).
template<typename Real> struct EvaluatorBondHarmonic
{
    struct param_type
    {
        Real k, r_0;

        static param_type make(double k, double r0)
        {
            param_type p;
            p.k = Real(k);
            p.r_0 = Real(r0);
            return p;
        }
    };

    __host__ __device__ __forceinline__ EvaluatorBondHarmonic(Real _rsq, const param_type& p)
        : rsq(_rsq), K(p.k), r_0(p.r_0)
    {
    }

    __host__ __device__ __forceinline__ bool evalForceAndEnergy(Real& force_divr, Real& bond_eng)
    {
        Real r = sqrt(rsq);
        force_divr = K * (r_0 / r - Real(1));
        if (!isfinite(force_divr))
            force_divr = Real(0);
        bond_eng = Real(0.5) * K * (r_0 - r) * (r_0 - r);
        return true;
    }

    static const char* getName()
    {
        return "harmonic";
    }

    Real rsq, K, r_0;
};

} // namespace hm
