// pair_ref.hip -- the library's instantiation of the reference-layout pair driver for LJ
// (hoomd/md/PotentialPairGPUKernel.cu.inc:19-21 instantiates gpu_compute_pair_forces per
// evaluator; plugins do the same with HM_PAIR_REF_EXPORT in their own library).
#include "evaluators.h"
#include "pair_ref.h"

namespace
{
// EvaluatorPairLJ over the reference's device parameter layout (hoomd/md/EvaluatorPairLJ.h:99-155:
// Scalar sigma_6, Scalar epsilon_x_4), arithmetic in Real
template<typename Real> struct RefPairLJ : hm::EvaluatorPairLJ<Real>
{
    typedef hm_lj_param_ref param_type;

    __host__ __device__ static typename hm::EvaluatorPairLJ<Real>::param_type
    narrow(const param_type& p)
    {
        typename hm::EvaluatorPairLJ<Real>::param_type q;
        q.sigma_6 = Real(p.sigma_6);
        q.epsilon_x_4 = Real(p.epsilon_x_4);
        return q;
    }

    __host__ __device__ RefPairLJ(Real rsq, Real rcutsq, const param_type& p)
        : hm::EvaluatorPairLJ<Real>(rsq, rcutsq, narrow(p))
    {
    }
};
} // namespace

HM_PAIR_REF_EXPORT(lj, RefPairLJ)
