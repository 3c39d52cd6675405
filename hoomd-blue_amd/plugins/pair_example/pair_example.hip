// pair_example.hip -- an evaluator added from outside the library, the way the reference's
// example_plugins/pair_plugin adds EvaluatorPairExample: the evaluator and one instantiation line,
// built into its own shared library (libpair_example.so) that exports
//     int hm_compute_pair_forces_example(const hm_pair_ref_args*, const void* d_params, void*)
// with d_params = {Scalar k, Scalar sigma} [ntypes * ntypes] on the device.
//
// The potential is the reference example's (example_plugins/pair_plugin/EvaluatorPairExample.h):
// a harmonic repulsion U = k (sigma - r)^2 / 2 for r < r_cut, shifted to zero at r_cut on request.
#include <cmath>

#include "pair_ref.h"

template<typename Real> struct EvaluatorPairExample
{
    struct param_type
    {
        double k;
        double sigma;
    } __attribute__((aligned(16)));

    __host__ __device__ EvaluatorPairExample(Real _rsq, Real _rcutsq, const param_type& p)
        : rsq(_rsq), rcutsq(_rcutsq), k(Real(p.k)), sigma(Real(p.sigma))
    {
    }

    __host__ __device__ static bool needsCharge()
    {
        return false;
    }
    __host__ __device__ void setCharge(Real, Real) { }

    __host__ __device__ bool evalForceAndEnergy(Real& force_divr, Real& pair_eng,
                                                bool energy_shift) const
    {
        if (!(rsq < rcutsq))
            return false;
        const Real r = sqrt(rsq);
        const Real overlap = sigma - r;
        force_divr = k * overlap / r;
        pair_eng = Real(0.5) * k * overlap * overlap;
        if (energy_shift)
        {
            const Real cut = sigma - sqrt(rcutsq);
            pair_eng -= Real(0.5) * k * cut * cut;
        }
        return true;
    }

    Real rsq, rcutsq, k, sigma;
};

HM_PAIR_REF_EXPORT(example, EvaluatorPairExample)
