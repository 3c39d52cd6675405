// ref_driver.cc -- thin C entry points around the REFERENCE's own evaluator and box headers.
//
// TEST INFRASTRUCTURE ONLY. Compiled by oracle/Makefile (target `ref`) against the unmodified
// HOOMD-blue sources under $(REF) into oracle/_ref/libhoomd_ref.so. It pins the oracle's
// restatements of:
//   hoomd/BoxDim.h            makeFraction (:276-283), minImage host path (:317-389), wrap (:407-475)
//   hoomd/md/EvaluatorPairLJ.h      param_type(sigma, epsilon) (:128-133), evalForceAndEnergy (:190-211)
//   hoomd/md/EvaluatorBondFENE.h    evaluator ctor (:84-90), evalForceAndEnergy (:116-156)
// The reference headers are used exactly as they lie; nothing here stands in for a missing
// header. The library is loaded into a Python process, which provides libpython's symbols.

#include "hoomd/BoxDim.h"
#include "hoomd/HOOMDMath.h"
#include "hoomd/md/EvaluatorBondFENE.h"
#include "hoomd/md/EvaluatorPairLJ.h"

using namespace hoomd;

extern "C"
    {
    int ref_lj(double rsq,
               double rcutsq,
               double sigma,
               double epsilon,
               int shift,
               double* force_divr,
               double* energy)
        {
        md::EvaluatorPairLJ::param_type p(sigma, epsilon);
        md::EvaluatorPairLJ eval(rsq, rcutsq, p);
        Scalar f = 0, e = 0;
        bool ok = eval.evalForceAndEnergy(f, e, shift != 0);
        *force_divr = f;
        *energy = e;
        return ok ? 1 : 0;
        }

    int ref_fene(double rsq,
                 double k,
                 double r0,
                 double sigma,
                 double epsilon,
                 double delta,
                 double* force_divr,
                 double* energy)
        {
        md::fene_params p;
        p.k = k;
        p.r_0 = r0;
        p.delta = delta;
        p.sigma_6 = sigma * sigma * sigma * sigma * sigma * sigma;
        p.epsilon_x_4 = Scalar(4.0) * epsilon;
        md::EvaluatorBondFENE eval(rsq, p);
        Scalar f = 0, e = 0;
        bool ok = eval.evalForceAndEnergy(f, e);
        *force_divr = f;
        *energy = e;
        return ok ? 1 : 0;
        }

    void ref_make_fraction(double Lx, double Ly, double Lz, const double* v, double* f)
        {
        BoxDim box(Lx, Ly, Lz);
        Scalar3 r = box.makeFraction(make_scalar3(v[0], v[1], v[2]));
        f[0] = r.x;
        f[1] = r.y;
        f[2] = r.z;
        }

    void ref_min_image(double Lx, double Ly, double Lz, double* w)
        {
        BoxDim box(Lx, Ly, Lz);
        Scalar3 r = box.minImage(make_scalar3(w[0], w[1], w[2]));
        w[0] = r.x;
        w[1] = r.y;
        w[2] = r.z;
        }

    void ref_wrap(double Lx, double Ly, double Lz, double* w, int* img)
        {
        BoxDim box(Lx, Ly, Lz);
        Scalar3 p = make_scalar3(w[0], w[1], w[2]);
        int3 im = make_int3(img[0], img[1], img[2]);
        box.wrap(p, im);
        w[0] = p.x;
        w[1] = p.y;
        w[2] = p.z;
        img[0] = im.x;
        img[1] = im.y;
        img[2] = im.z;
        }
    }
