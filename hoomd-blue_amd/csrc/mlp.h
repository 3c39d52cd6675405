// mlp.h -- launch interfaces of the learned coarse-grained site-energy kernels (mlp.hip).
//
// The potential is a two-pass many-body form, the structure of the reference's EAM
// (hoomd/metal/EAMForceGPU.cu:20 density pass, :147 force pass):
//   pass 1  G_k(i) = sum_j phi_k(r_ij)                 16 Gaussian radial-basis densities
//   site    E_i = MLP(x_i), x_i = (G(i) - gmean) gscale  16 -> 64 -> 64 -> 1 SiLU on bf16 MFMA
//           A(i) = dE_i/dG(i)                           the analytic backward of the MLP
//   pass 2  F_i = -sum_j sum_k (A_k(i) + A_k(j)) phi_k'(r_ij) r_ij / |r_ij|
//           (+ an optional tabulated pair term of the same pairs, e.g. a structural prior)
// Training adds the adjoint pass C(i) = -gscale sum_j phi'(r_ij) (e_i - e_j) . r_ij / |r_ij|
// (e = dLoss/dF) and the per-site parameter gradient of C(i) . dE_i/dx, whose matrix products
// also run on bf16 MFMA.
#pragma once

#include "kernels.h"

namespace hm
{

constexpr int MLP_NK = 16;
constexpr int MLP_H = 64;
// flat parameter vector: W1[64][16], b1[64], W2[64][64] (row = output unit), b2[64], w3[64], b3
constexpr int MP_W1 = 0;
constexpr int MP_B1 = MP_W1 + MLP_H * MLP_NK;
constexpr int MP_W2 = MP_B1 + MLP_H;
constexpr int MP_B2 = MP_W2 + MLP_H * MLP_H;
constexpr int MP_W3 = MP_B2 + MLP_H;
constexpr int MP_B3 = MP_W3 + MLP_H;
constexpr int MLP_NPARAM = MP_B3 + 1;
// bf16 operand copies of the weights, one buffer: W1 [64][16], W1^T [16][64], W2 [64][64],
// W2^T [64][64]
constexpr int MB_W1 = 0;
constexpr int MB_W1T = MB_W1 + MLP_H * MLP_NK;
constexpr int MB_W2 = MB_W1T + MLP_H * MLP_NK;
constexpr int MB_W2T = MB_W2 + MLP_H * MLP_H;
constexpr int MLP_NBF16 = MB_W2T + MLP_H * MLP_H;
// workgroups of the gradient kernel (each writes one row of partial gradients)
constexpr unsigned int MLP_GRAD_BLOCKS = 256;
// workgroups of the force pass in training (each writes one partial loss)
constexpr unsigned int MLP_PAIR_BLOCK = 256;

// Descriptor constants: centres mu0 + k dmu, exp(-(r - mu)^2 inv2w2), cosine cutoff at rcut.
struct MlpDescr
{
    float rcut, rcutsq, inv_rcut, mu0, dmu, inv2w2, invw2;
    float gmean[MLP_NK], gscale[MLP_NK];
    // the Gaussians by recurrence from the middle centre (three exponentials per pair) when
    // every g_MID stays far above the FP32 underflow over [0, rcut); else sixteen exponentials
    int rebase;
    float inv_dmu, b, q;   // 1 / dmu, inv2w2 dmu^2, exp(-2 b)
};

struct MlpArgs
{
    unsigned int N;
    const double4* pos;            // FP64 positions
    DevBox box;
    const unsigned int* nlist;     // row-major full list, rows of nmax particle indices
    unsigned int nmax;
    const unsigned int* n_neigh;
    MlpDescr d;
    const float* master;           // FP32 parameters [MLP_NPARAM]
    const uint16_t* wb;            // bf16 weight copies [MLP_NBF16]
    float4* G;                     // [4N] descriptor (16 floats per site)
    float4* A;                     // [4N] dE/dG
    float* Esite;                  // [N] site energies
    float4* force;                 // [N] (fx, fy, fz, site energy)
    float* virial;                 // [6N] rows of pitch N, or NULL
    // optional tabulated pair term on [ptab_rmin, rcut) evaluated in the force pass like
    // EvaluatorPairTable (U then F, ptab_width points each); ptab_width 0: none
    const float* ptab;
    float ptab_rmin;
    unsigned int ptab_width;
    // training
    const float4* target;          // [N] target forces (xyz)
    float4* err;                   // [N] dLoss/dF
    float4* C;                     // [4N] adjoint, already scaled by gscale
    float* loss_partial;           // [ceil(N * tpp / MLP_PAIR_BLOCK)]
    float* gpart;                  // [MLP_GRAD_BLOCKS][MLP_NPARAM]
    float loss_scale;              // 2 / (3N)
    // bf16 copy of A (4 values per uint2, 16 per site) written by the site kernel and gathered
    // by the force pass for both partners, or NULL (the force pass then reads A)
    uint2* Ab;
    // FP32 copy of the positions (x, y, z) gathered by the density and force passes, or NULL
    // (they then read the FP64 positions)
    const float4* pos32;
};
// FP32 copy of the FP64 positions (x, y, z, 0), round to nearest
hipError_t launch_pos_f32(unsigned int N, const double4* pos, float4* pos32, hipStream_t s);

MlpDescr make_mlp_descr(const hm_mlp_cg_model& m);
unsigned int mlp_pair_blocks(unsigned int N);
// pass 1: G
hipError_t launch_mlp_descriptor(const MlpArgs& a, hipStream_t s);
// site: A and Esite from G
hipError_t launch_mlp_site(const MlpArgs& a, hipStream_t s);
// pass 2: force (and virial); with train, err and loss_partial from target
hipError_t launch_mlp_force(const MlpArgs& a, bool train, hipStream_t s);
// training: adjoint pass, then the parameter-gradient kernel (gpart rows)
hipError_t launch_mlp_backward(const MlpArgs& a, hipStream_t s);
// sum of the partial losses into out[0] (FP64)
hipError_t launch_mlp_loss(const float* loss_partial, unsigned int n, double* out, hipStream_t s);
// Adam step t (from 1) on master with moments m1, m2 from the gpart rows; refreshes wb
hipError_t launch_mlp_adam(float* master, float* m1, float* m2, const float* gpart,
                           uint16_t* wb, float lr, float beta1, float beta2, float eps,
                           unsigned int t, hipStream_t s);
// minibatches: gacc (+)= the sum of the gpart rows; Adam step t on gacc * scale
hipError_t launch_mlp_grad_accum(const float* gpart, float* gacc, bool first, hipStream_t s);
hipError_t launch_mlp_adam_acc(float* master, float* m1, float* m2, const float* gacc,
                               float scale, uint16_t* wb, float lr, float beta1, float beta2,
                               float eps, unsigned int t, hipStream_t s);
// gradient only: sum of the gpart rows into grad (FP32)
hipError_t launch_mlp_grad_sum(const float* gpart, float* grad, hipStream_t s);
// bf16 copies wb from master
hipError_t launch_mlp_pack(const float* master, uint16_t* wb, hipStream_t s);

// Frame f of a training set (FP32 xyz in tag order) into the system's FP64 positions (type
// kept, wrapped into the box) and the target forces, both in the current index order.
hipError_t launch_mlp_load_frame(unsigned int N, const unsigned int* rtag, const float* fpos,
                                 const float* fforce, const DevBox& box, double4* pos,
                                 float4* target, hipStream_t s);

// Coarse-grained mapping: site I holds the particles with tags tags[row[I]] .. tags[row[I+1]-1];
// its position is their mass-weighted centre (each particle unwrapped to the minimum image of
// the first one, the centre wrapped into the box) and its force the sum of their forces over
// nforces force arrays (at most 4). Outputs are FP32 xyz per site.
struct SiteMapArgs
{
    unsigned int nsites;
    const unsigned int* row;
    const unsigned int* tags;
    const unsigned int* rtag;
    const double4* pos;
    const double4* vel;           // mass in w
    const float4* f[4];
    unsigned int nforces;
    DevBox box;
    float* out_pos;               // [3 nsites]
    float* out_force;             // [3 nsites] or NULL
};
hipError_t launch_map_sites(const SiteMapArgs& a, hipStream_t s);

} // namespace hm
