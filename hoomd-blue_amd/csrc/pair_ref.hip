// pair_ref.hip -- the library's instantiation of the reference-layout pair driver for LJ
// (hoomd/md/PotentialPairGPUKernel.cu.inc:19-21 instantiates gpu_compute_pair_forces per
// evaluator; plugins do the same with HM_PAIR_REF_EXPORT in their own library).
#include "capi_util.h"
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

// the HM_PAIR_REF_EXPORT body, with the reason of a failure kept for hm_last_error
extern "C" int hm_compute_pair_forces_lj(const hm_pair_ref_args* args, const void* d_params,
                                         void* stream)
{
    const int r = hm::pref::compute_pair_forces<RefPairLJ<float>>(args, d_params, stream);
    if (r == HM_ERR_ARGS)
        return hm::fail(r, "hm_compute_pair_forces_lj: invalid arguments");
    if (r != HM_OK)
        return hm::fail(r, "hm_compute_pair_forces_lj: kernel launch failed");
    return HM_OK;
}
