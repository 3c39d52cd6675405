// mlp.hip -- learned coarse-grained site energy: descriptor pass, MLP forward/backward on bf16
// MFMA, force pass, and the training kernels (adjoint pass, parameter gradient, Adam).
//
// Matrix products use v_mfma_f32_16x16x16_bf16 (bf16 operands, FP32 accumulation). One
// wavefront evaluates the MLP for a tile of 16 sites: the sites are the N dimension, the hidden
// units the M dimension. Lane l holds, for site s = l % 16 and group g = l / 16,
//   A operand (16x16, M x K):  rows 16m + s,       columns k0 + 4g .. 4g + 3
//   B operand (16x16, K x N):  rows 4g .. 4g + 3,  column s
//   C result  (16x16, M x N):  rows 4g .. 4g + 3,  column s
// so a layer's result, rounded to bf16, is already the next layer's B operand: no shuffles
// between layers. Weights (and their transposes, for the backward) are staged in LDS as bf16.
#include "mlp.h"

#include <cmath>

namespace hm
{

namespace
{
typedef short bf4 __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));

constexpr int PAIR_TPP = 4;  // lanes per site in the pair passes
constexpr int GMID = MLP_NK / 2;  // base centre of the Gaussian recurrence

typedef __bf16 hb4 __attribute__((ext_vector_type(4)));

// FP32 -> bf16 by the hardware conversion (v_cvt_pk_bf16_f32, round to nearest even)
__device__ __forceinline__ unsigned short f2bf(float f)
{
    return __builtin_bit_cast(unsigned short, (__bf16)f);
}

__device__ __forceinline__ bf4 pack4(const f4v& x)
{
    return __builtin_bit_cast(bf4, __builtin_convertvector(x, hb4));
}

__device__ __forceinline__ bf4 pack4(float a, float b, float c, float d)
{
    f4v x = {a, b, c, d};
    return pack4(x);
}

__device__ __forceinline__ f4v mfma(bf4 a, bf4 b, f4v c)
{
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// A operand of the 16x16 tile (m, kb) of a row-major matrix with K columns
__device__ __forceinline__ bf4 ldsA(const unsigned short* W, int K, int m, int kb, int s, int g)
{
    return *(const bf4*)(W + (16 * m + s) * K + 16 * kb + 4 * g);
}

__device__ __forceinline__ float sigm(float h)
{
    // hardware reciprocal (1 ulp): the result feeds bf16 operands and FP32 sums
    return __builtin_amdgcn_rcpf(1.0f + __expf(-h));
}
// silu'(h) and silu''(h) from s = sigm(h)
__device__ __forceinline__ float silu_d_s(float h, float s)
{
    return s * (1.0f + h * (1.0f - s));
}
__device__ __forceinline__ float silu_dd_s(float h, float s)
{
    return s * (1.0f - s) * (2.0f + h * (1.0f - 2.0f * s));
}
__device__ __forceinline__ f4v zero4()
{
    f4v z = {0.f, 0.f, 0.f, 0.f};
    return z;
}

// FP32 pair vector from FP64 positions with the FP64 minimum image
__device__ __forceinline__ void pair_vec(const DevBox& box, const double4& pi, const double4& pj,
                                         float& fx, float& fy, float& fz)
{
    double dx, dy, dz;
    {
#pragma clang fp contract(off)
        dx = pi.x - pj.x;
        dy = pi.y - pj.y;
        dz = pi.z - pj.z;
        min_image(box, dx, dy, dz);
    }
    fx = (float)dx;
    fy = (float)dy;
    fz = (float)dz;
}

// Positions of a pair pass: the FP64 positions with the FP64 minimum image, or the FP32 copy.
// With FP32 positions the difference of two FP32 values is exact or within half an ulp of the
// difference, and the image shift subtracts the box length in two FP32 parts (L = Lh + Ll, the
// shift by Lh exact by Sterbenz's lemma), so the pair vector keeps the FP32 positions' accuracy.
template<bool P32> struct PairPos;
template<> struct PairPos<false>
{
    double4 p;
    __device__ __forceinline__ void load(const MlpArgs& a, unsigned int j) { p = a.pos[j]; }
    __device__ __forceinline__ void vec(const MlpArgs& a, const PairPos& pj, float& dx, float& dy,
                                        float& dz) const
    {
        pair_vec(a.box, p, pj.p, dx, dy, dz);
    }
};
template<> struct PairPos<true>
{
    float4 p;
    __device__ __forceinline__ void load(const MlpArgs& a, unsigned int j) { p = a.pos32[j]; }
    __device__ __forceinline__ static float axis(float d, double L, double Linv)
    {
        const float Lh = (float)L, Ll = (float)(L - (double)Lh);
        const float s = rintf(d * (float)Linv);
        d = __builtin_fmaf(-s, Lh, d);
        return __builtin_fmaf(-s, Ll, d);
    }
    __device__ __forceinline__ void vec(const MlpArgs& a, const PairPos& pj, float& dx, float& dy,
                                        float& dz) const
    {
        dx = axis(p.x - pj.p.x, a.box.L[0], a.box.Linv[0]);
        dy = axis(p.y - pj.p.y, a.box.L[1], a.box.Linv[1]);
        dz = axis(p.z - pj.p.z, a.box.L[2], a.box.Linv[2]);
    }
};

// cosine cutoff f_c(r) = (cos(pi r / rc) + 1) / 2 and its derivative. The hardware sine and
// cosine take their argument in turns (v_sin_f32: sin(2 pi x)); x = r / (2 rc) lies in [0, 1/2].
__device__ __forceinline__ void cutoff(const MlpDescr& d, float r, float& fc, float& dfc)
{
    const float x = 0.5f * r * d.inv_rcut;
    const float sn = __builtin_amdgcn_sinf(x), cs = __builtin_amdgcn_cosf(x);
    fc = 0.5f * (cs + 1.0f);
    dfc = -0.5f * 3.14159265358979f * d.inv_rcut * sn;
}

// g[k] = exp(-(r - mu_k)^2 inv2w2), mu_k = mu0 + k dmu. With rebase: t = (r - mu0) / dmu and
// g_k = exp(-b (t - k)^2); from the middle centre GMID the ratio g_{k+1} / g_k = exp(2 b (t - k) - b)
// and g_{k-1} / g_k = exp(-2 b (t - k) - b) each change by the factor exp(-2 b) per step, so three
// exponentials and thirty products give all sixteen (about 1e-6 relative error after 8 steps).
__device__ __forceinline__ void gaussians(const MlpDescr& d, float r, float (&g)[MLP_NK])
{
    if (d.rebase)
    {
        const float s = (r - d.mu0) * d.inv_dmu - (float)GMID;
        g[GMID] = __expf(-d.b * s * s);
        float up = __expf(2.f * d.b * s - d.b);    // g_{GMID+1} / g_GMID
        float dn = __expf(-2.f * d.b * s - d.b);   // g_{GMID-1} / g_GMID
#pragma unroll
        for (int k = GMID; k + 1 < MLP_NK; k++)
        {
            g[k + 1] = g[k] * up;
            up *= d.q;
        }
#pragma unroll
        for (int k = GMID; k > 0; k--)
        {
            g[k - 1] = g[k] * dn;
            dn *= d.q;
        }
    }
    else
    {
#pragma unroll
        for (int k = 0; k < MLP_NK; k++)
        {
            const float dd = r - (d.mu0 + (float)k * d.dmu);
            g[k] = __expf(-dd * dd * d.inv2w2);
        }
    }
}

// sum_k c_k phi_k'(r) with c_k = ai_k + aj_k, phi_k = g_k f_c, phi_k' = g_k (dfc - fc (r - mu_k)
// / w^2). With the recurrence the Gaussians are consumed as they are formed:
// S0 = sum c_k g_k, Sm = sum (k - GMID) c_k g_k and sum c_k g_k (r - mu_k) = dmu (s S0 - Sm),
// s = (r - mu0) / dmu - GMID, so no array of sixteen values is held per pair.
__device__ __forceinline__ float basis_deriv_dot(const MlpDescr& d, float r, float fc, float dfc,
                                                 const float (&ai)[MLP_NK],
                                                 const float (&aj)[MLP_NK])
{
    if (d.rebase)
    {
        const float s = (r - d.mu0) * d.inv_dmu - (float)GMID;
        const float g0 = __expf(-d.b * s * s);
        float up = __expf(2.f * d.b * s - d.b);
        float dn = __expf(-2.f * d.b * s - d.b);
        float S0 = (ai[GMID] + aj[GMID]) * g0, Sm = 0.f;
        float g = g0;
#pragma unroll
        for (int k = GMID + 1; k < MLP_NK; k++)
        {
            g *= up;
            up *= d.q;
            const float t = (ai[k] + aj[k]) * g;
            S0 += t;
            Sm += (float)(k - GMID) * t;
        }
        g = g0;
#pragma unroll
        for (int k = GMID - 1; k >= 0; k--)
        {
            g *= dn;
            dn *= d.q;
            const float t = (ai[k] + aj[k]) * g;
            S0 += t;
            Sm += (float)(k - GMID) * t;
        }
        return dfc * S0 - fc * d.invw2 * d.dmu * (s * S0 - Sm);
    }
    float sdot = 0.f;
#pragma unroll
    for (int k = 0; k < MLP_NK; k++)
    {
        const float dd = r - (d.mu0 + (float)k * d.dmu);
        const float g = __expf(-dd * dd * d.inv2w2);
        sdot += (ai[k] + aj[k]) * g * (dfc - fc * dd * d.invw2);
    }
    return sdot;
}

template<int T> __device__ __forceinline__ float group_sum(float v)
{
#pragma unroll
    for (int m = T / 2; m > 0; m >>= 1)
        v += __shfl_xor(v, m);
    return v;
}

// ------------------------------------------------------------------------------ pass 1
template<int U, bool P32> __global__ __launch_bounds__(256) void mlp_descriptor_kernel(MlpArgs a)
{
    // XCD-aware: each XCD's workgroups take a contiguous eighth of the (Hilbert-ordered) sites,
    // so the neighbour gathers of one XCD stay in its own L2
    const unsigned int gid = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    const unsigned int i = gid / PAIR_TPP, sub = gid % PAIR_TPP;
    const bool valid = i < a.N;
    float acc[MLP_NK];
#pragma unroll
    for (int k = 0; k < MLP_NK; k++)
        acc[k] = 0.f;
    if (valid)
    {
        PairPos<P32> pi;
        pi.load(a, i);
        const unsigned int n = min(a.n_neigh[i], a.nmax);
        const unsigned int* row = a.nlist + (size_t)i * a.nmax;
        // two neighbours per iteration, and the row entries of the next iteration loaded one
        // iteration ahead, so the list loads and the position gathers overlap
        unsigned int jn[U];
#pragma unroll
        for (int u = 0; u < U; u++)
        {
            const unsigned int q = sub + u * PAIR_TPP;
            jn[u] = q < n ? row[q] : i;
        }
        for (unsigned int q0 = sub; q0 < n; q0 += PAIR_TPP * U)
        {
            PairPos<P32> pj[U];
            bool live[U];
#pragma unroll
            for (int u = 0; u < U; u++)
            {
                live[u] = q0 + u * PAIR_TPP < n;
                pj[u].load(a, jn[u]);
            }
#pragma unroll
            for (int u = 0; u < U; u++)
            {
                const unsigned int q = q0 + (U + u) * PAIR_TPP;
                jn[u] = q < n ? row[q] : i;
            }
#pragma unroll
            for (int u = 0; u < U; u++)
            {
                float dx, dy, dz;
                pi.vec(a, pj[u], dx, dy, dz);
                const float rsq = dx * dx + dy * dy + dz * dz;
                if (!live[u] || !(rsq < a.d.rcutsq))
                    continue;
                const float r = rsq * __builtin_amdgcn_rsqf(rsq);
                float fc, dfc;
                cutoff(a.d, r, fc, dfc);
                float gk[MLP_NK];
                gaussians(a.d, r, gk);
#pragma unroll
                for (int k = 0; k < MLP_NK; k++)
                    acc[k] += gk[k] * fc;
            }
        }
    }
#pragma unroll
    for (int k = 0; k < MLP_NK; k++)
        acc[k] = group_sum<PAIR_TPP>(acc[k]);
    if (valid)
    {
        // lane sub of the group stores features 4 sub .. 4 sub + 3
        float4 o;
        o.x = acc[0];
        o.y = acc[1];
        o.z = acc[2];
        o.w = acc[3];
#pragma unroll
        for (int q = 1; q < 4; q++)
            if (sub == (unsigned int)q)
            {
                o.x = acc[4 * q];
                o.y = acc[4 * q + 1];
                o.z = acc[4 * q + 2];
                o.w = acc[4 * q + 3];
            }
        a.G[(size_t)i * 4 + sub] = o;
    }
}

// ------------------------------------------------------------------------------ site MLP
struct MlpLds
{
    unsigned short w[MLP_NBF16];
    float b1[MLP_H], b2[MLP_H], w3[MLP_H];
};

__device__ void stage_weights(MlpLds& S, const MlpArgs& a)
{
    for (unsigned int t = threadIdx.x; t < MLP_NBF16 / 4; t += blockDim.x)
        ((uint2*)S.w)[t] = ((const uint2*)a.wb)[t];
    for (unsigned int t = threadIdx.x; t < MLP_H; t += blockDim.x)
    {
        S.b1[t] = a.master[MP_B1 + t];
        S.b2[t] = a.master[MP_B2 + t];
        S.w3[t] = a.master[MP_W3 + t];
    }
    __syncthreads();
}

// Forward and input backward of one 16-site tile, values in the C/B layout of the lane
struct SiteTile
{
    bf4 xb;           // normalised descriptor (B operand, features 4g..4g+3)
    float x[4];
    f4v h1[4], h2[4];
    f4v s1[4];        // sigm(h1), kept for the backward
    bf4 a1b[4];
    bf4 d2b[4];       // w3 * silu'(h2)
    f4v u[4];         // W2^T d2
    bf4 d1b[4];       // u * silu'(h1)
    f4v dx;           // dE/dx (features 4g..4g+3)
    float e;          // site energy (every lane of the site)
};

__device__ __forceinline__ void site_eval(const MlpLds& S, const MlpArgs& a, const float4 gv,
                                          int s, int g, float b3, SiteTile& T)
{
    const float gin[4] = {gv.x, gv.y, gv.z, gv.w};
#pragma unroll
    for (int r = 0; r < 4; r++)
        T.x[r] = (gin[r] - a.d.gmean[4 * g + r]) * a.d.gscale[4 * g + r];
    T.xb = pack4(T.x[0], T.x[1], T.x[2], T.x[3]);
    const unsigned short* W1 = S.w + MB_W1;
    const unsigned short* W1T = S.w + MB_W1T;
    const unsigned short* W2 = S.w + MB_W2;
    const unsigned short* W2T = S.w + MB_W2T;
#pragma unroll
    for (int m = 0; m < 4; m++)
    {
        T.h1[m] = mfma(ldsA(W1, MLP_NK, m, 0, s, g), T.xb, zero4());
        f4v act;
#pragma unroll
        for (int r = 0; r < 4; r++)
        {
            T.h1[m][r] += S.b1[16 * m + 4 * g + r];
            T.s1[m][r] = sigm(T.h1[m][r]);
            act[r] = T.h1[m][r] * T.s1[m][r];
        }
        T.a1b[m] = pack4(act);
    }
    float e = 0.f;
#pragma unroll
    for (int m = 0; m < 4; m++)
    {
        f4v acc = zero4();
#pragma unroll
        for (int kb = 0; kb < 4; kb++)
            acc = mfma(ldsA(W2, MLP_H, m, kb, s, g), T.a1b[kb], acc);
        f4v d2;
#pragma unroll
        for (int r = 0; r < 4; r++)
        {
            const int o = 16 * m + 4 * g + r;
            acc[r] += S.b2[o];
            const float sg = sigm(acc[r]);
            e += S.w3[o] * (acc[r] * sg);
            d2[r] = S.w3[o] * silu_d_s(acc[r], sg);
        }
        T.h2[m] = acc;
        T.d2b[m] = pack4(d2);
    }
    e += __shfl_xor(e, 16);
    e += __shfl_xor(e, 32);
    T.e = e + b3;
#pragma unroll
    for (int m = 0; m < 4; m++)
    {
        f4v acc = zero4();
#pragma unroll
        for (int kb = 0; kb < 4; kb++)
            acc = mfma(ldsA(W2T, MLP_H, m, kb, s, g), T.d2b[kb], acc);
        T.u[m] = acc;
        f4v d1;
#pragma unroll
        for (int r = 0; r < 4; r++)
            d1[r] = acc[r] * silu_d_s(T.h1[m][r], T.s1[m][r]);
        T.d1b[m] = pack4(d1);
    }
    f4v dx = zero4();
#pragma unroll
    for (int kb = 0; kb < 4; kb++)
        dx = mfma(ldsA(W1T, MLP_H, 0, kb, s, g), T.d1b[kb], dx);
    T.dx = dx;
}

constexpr unsigned int SITE_BLOCK = 256;  // 1024 (16 wavefronts per staging) measured 2.4x slower

__global__ __launch_bounds__(SITE_BLOCK) void mlp_site_kernel(MlpArgs a)
{
    __shared__ MlpLds S;
    const int lane = threadIdx.x & 63;
    const int s = lane & 15, g = lane >> 4;
    const unsigned int ntiles = (a.N + 15) / 16;
    const unsigned int nwaves = gridDim.x * (blockDim.x / 64);
    const float4 zero = make_float4(0.f, 0.f, 0.f, 0.f);
    // the descriptor of the wavefront's next tile is loaded one tile ahead (the first one
    // while the weights are staged), so each tile's load latency hides behind the previous
    // tile's matrix work
    unsigned int t = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
    float4 gnext = (t < ntiles && 16 * t + s < a.N) ? a.G[(size_t)(16 * t + s) * 4 + g] : zero;
    stage_weights(S, a);
    const float b3 = a.master[MP_B3];
    for (; t < ntiles; t += nwaves)
    {
        const unsigned int site = 16 * t + s;
        const bool valid = site < a.N;
        const float4 gv = gnext;
        const unsigned int tn = t + nwaves;
        if (tn < ntiles)
            gnext = 16 * tn + s < a.N ? a.G[(size_t)(16 * tn + s) * 4 + g] : zero;
        SiteTile T;
        site_eval(S, a, gv, s, g, b3, T);
        if (valid)
        {
            float4 o;
            o.x = a.d.gscale[4 * g] * T.dx[0];
            o.y = a.d.gscale[4 * g + 1] * T.dx[1];
            o.z = a.d.gscale[4 * g + 2] * T.dx[2];
            o.w = a.d.gscale[4 * g + 3] * T.dx[3];
            if (a.A)
                a.A[(size_t)site * 4 + g] = o;
            if (a.Ab)
            {
                const bf4 ob = pack4(o.x, o.y, o.z, o.w);
                a.Ab[(size_t)site * 4 + g] = __builtin_bit_cast(uint2, ob);
            }
            if (g == 0)
                a.Esite[site] = T.e;
        }
    }
}

// ------------------------------------------------------------------------------ pass 2
// Tables up to this many points are staged in LDS for the force pass
constexpr unsigned int MLP_TAB_LDS = 2048;

// the 16 dE/dG values of site j: FP32 (four 16-byte loads) or the bf16 copy (two)
template<bool BF> struct AjLoad;
template<> struct AjLoad<false>
{
    float4 v[4];
    __device__ __forceinline__ void load(const MlpArgs& a, unsigned int j)
    {
#pragma unroll
        for (int w = 0; w < 4; w++)
            v[w] = a.A[(size_t)j * 4 + w];
    }
    __device__ __forceinline__ void get(float (&o)[MLP_NK]) const
    {
#pragma unroll
        for (int w = 0; w < 4; w++)
        {
            o[4 * w] = v[w].x;
            o[4 * w + 1] = v[w].y;
            o[4 * w + 2] = v[w].z;
            o[4 * w + 3] = v[w].w;
        }
    }
};
template<> struct AjLoad<true>
{
    uint4 v[2];
    __device__ __forceinline__ void load(const MlpArgs& a, unsigned int j)
    {
        const uint4* p = (const uint4*)a.Ab + (size_t)j * 2;
        v[0] = p[0];
        v[1] = p[1];
    }
    __device__ __forceinline__ void get(float (&o)[MLP_NK]) const
    {
        const unsigned int u[8] = {v[0].x, v[0].y, v[0].z, v[0].w,
                                   v[1].x, v[1].y, v[1].z, v[1].w};
#pragma unroll
        for (int h = 0; h < 8; h++)
        {
            o[2 * h] = __uint_as_float(u[h] << 16);
            o[2 * h + 1] = __uint_as_float(u[h] & 0xffff0000u);
        }
    }
};

template<bool TRAIN, int U, bool BF, bool P32> __global__ __launch_bounds__(MLP_PAIR_BLOCK) void mlp_force_kernel(MlpArgs a)
{
    __shared__ float stab[2 * MLP_TAB_LDS];
    // XCD-aware: each XCD's workgroups take a contiguous eighth of the (Hilbert-ordered) sites,
    // so the neighbour gathers of one XCD stay in its own L2
    const unsigned int gid = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    const unsigned int i = gid / PAIR_TPP, sub = gid % PAIR_TPP;
    const bool valid = i < a.N;
    float fx = 0.f, fy = 0.f, fz = 0.f, ep = 0.f;
    float v[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    // the pair table (EvaluatorPairTable.h:152-196 interpolation, the division by the spacing
    // taken as a product with its reciprocal), from LDS when it fits
    const unsigned int tw = a.ptab_width;
    const float* tU = a.ptab;
    if (tw && tw <= MLP_TAB_LDS)
    {
        for (unsigned int t = threadIdx.x; t < 2 * tw; t += blockDim.x)
            stab[t] = a.ptab[t];
        __syncthreads();
        tU = stab;
    }
    const float* tF = tU + tw;
    const float t_inv = tw ? (float)tw / (a.d.rcut - a.ptab_rmin) : 0.f;
    if (valid)
    {
        float ai[MLP_NK];
        {
            AjLoad<BF> li;
            li.load(a, i);
            li.get(ai);
        }
        PairPos<P32> pi;
        pi.load(a, i);
        const unsigned int n = min(a.n_neigh[i], a.nmax);
        const unsigned int* row = a.nlist + (size_t)i * a.nmax;
        // two neighbours per iteration: both position and both dE/dG gathers are issued before
        // the arithmetic of the first
        // the row entries of the next iteration are loaded one iteration ahead
        unsigned int jn[U];
#pragma unroll
        for (int u = 0; u < U; u++)
        {
            const unsigned int q = sub + u * PAIR_TPP;
            jn[u] = q < n ? row[q] : i;
        }
        for (unsigned int q0 = sub; q0 < n; q0 += PAIR_TPP * U)
        {
            unsigned int jj[U];
            bool live[U];
#pragma unroll
            for (int u = 0; u < U; u++)
            {
                live[u] = q0 + u * PAIR_TPP < n;
                jj[u] = jn[u];
            }
            PairPos<P32> pj[U];
            AjLoad<BF> lj[U];
#pragma unroll
            for (int u = 0; u < U; u++)
            {
                pj[u].load(a, jj[u]);
                lj[u].load(a, jj[u]);
            }
#pragma unroll
            for (int u = 0; u < U; u++)
            {
                const unsigned int q = q0 + (U + u) * PAIR_TPP;
                jn[u] = q < n ? row[q] : i;
            }
#pragma unroll
            for (int u = 0; u < U; u++)
            {
                float dx, dy, dz;
                pi.vec(a, pj[u], dx, dy, dz);
                const float rsq = dx * dx + dy * dy + dz * dz;
                if (!live[u] || !(rsq < a.d.rcutsq))
                    continue;
                // r and 1/r from the hardware reciprocal square root (1 ulp): the correctly
                // rounded sqrtf expands to a dozen instructions per pair
                const float rinv = __builtin_amdgcn_rsqf(rsq);
                const float r = rsq * rinv;
                float fc, dfc;
                cutoff(a.d, r, fc, dfc);
                float aj[MLP_NK];
                lj[u].get(aj);
                const float sdot = basis_deriv_dot(a.d, r, fc, dfc, ai, aj);
                float fdr = -sdot * rinv;
                if (tw && r >= a.ptab_rmin)
                {
                    const float vf = (r - a.ptab_rmin) * t_inv;
                    unsigned int k = (unsigned int)floorf(vf);
                    k = k < tw ? k : tw - 1;
                    const float U0 = tU[k], F0 = tF[k];
                    const float U1 = k + 1 < tw ? tU[k + 1] : 0.f;
                    const float F1 = k + 1 < tw ? tF[k + 1] : 0.f;
                    const float f = vf - (float)k;
                    fdr += (F0 + f * (F1 - F0)) * rinv;
                    ep += 0.5f * (U0 + f * (U1 - U0));
                }
                fx += fdr * dx;
                fy += fdr * dy;
                fz += fdr * dz;
                if (a.virial)
                {
                    const float h = 0.5f * fdr;
                    v[0] += h * dx * dx;
                    v[1] += h * dx * dy;
                    v[2] += h * dx * dz;
                    v[3] += h * dy * dy;
                    v[4] += h * dy * dz;
                    v[5] += h * dz * dz;
                }
            }
        }
    }
    fx = group_sum<PAIR_TPP>(fx);
    fy = group_sum<PAIR_TPP>(fy);
    fz = group_sum<PAIR_TPP>(fz);
    if (a.ptab_width)
        ep = group_sum<PAIR_TPP>(ep);
    if (a.virial)
#pragma unroll
        for (int m = 0; m < 6; m++)
            v[m] = group_sum<PAIR_TPP>(v[m]);
    float sq = 0.f;
    if (valid && sub == 0)
    {
        a.force[i] = make_float4(fx, fy, fz, a.Esite[i] + ep);
        if (a.virial)
            for (int m = 0; m < 6; m++)
                a.virial[(size_t)m * a.N + i] = v[m];
        if (TRAIN)
        {
            const float4 t = a.target[i];
            const float rx = fx - t.x, ry = fy - t.y, rz = fz - t.z;
            a.err[i] = make_float4(a.loss_scale * rx, a.loss_scale * ry, a.loss_scale * rz, 0.f);
            sq = rx * rx + ry * ry + rz * rz;
        }
    }
    if (TRAIN)
    {
        // block sum of the squared residuals, fixed order
        __shared__ float part[MLP_PAIR_BLOCK / 64];
#pragma unroll
        for (int m = 32; m > 0; m >>= 1)
            sq += __shfl_xor(sq, m);
        if ((threadIdx.x & 63) == 0)
            part[threadIdx.x / 64] = sq;
        __syncthreads();
        if (threadIdx.x == 0)
        {
            float t = 0.f;
            for (unsigned int w = 0; w < MLP_PAIR_BLOCK / 64; w++)
                t += part[w];
            a.loss_partial[blockIdx.x] = t;
        }
    }
}

// ------------------------------------------------------------------------------ adjoint pass
__global__ __launch_bounds__(256) void mlp_adjoint_kernel(MlpArgs a)
{
    // XCD-aware: each XCD's workgroups take a contiguous eighth of the (Hilbert-ordered) sites,
    // so the neighbour gathers of one XCD stay in its own L2
    const unsigned int gid = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    const unsigned int i = gid / PAIR_TPP, sub = gid % PAIR_TPP;
    const bool valid = i < a.N;
    float c[MLP_NK];
#pragma unroll
    for (int k = 0; k < MLP_NK; k++)
        c[k] = 0.f;
    if (valid)
    {
        const float4 ei = a.err[i];
        const double4 pi = a.pos[i];
        const unsigned int n = min(a.n_neigh[i], a.nmax);
        const unsigned int* row = a.nlist + (size_t)i * a.nmax;
        for (unsigned int q = sub; q < n; q += PAIR_TPP)
        {
            const unsigned int j = row[q];
            float dx, dy, dz;
            pair_vec(a.box, pi, a.pos[j], dx, dy, dz);
            const float rsq = dx * dx + dy * dy + dz * dz;
            if (!(rsq < a.d.rcutsq))
                continue;
            const float r = sqrtf(rsq);
            float fc, dfc;
            cutoff(a.d, r, fc, dfc);
            const float4 ej = a.err[j];
            const float de = ((ei.x - ej.x) * dx + (ei.y - ej.y) * dy + (ei.z - ej.z) * dz) / r;
            float gk[MLP_NK];
            gaussians(a.d, r, gk);
#pragma unroll
            for (int k = 0; k < MLP_NK; k++)
            {
                const float dd = r - (a.d.mu0 + (float)k * a.d.dmu);
                c[k] -= gk[k] * (dfc - fc * dd * a.d.invw2) * de;
            }
        }
    }
#pragma unroll
    for (int k = 0; k < MLP_NK; k++)
        c[k] = group_sum<PAIR_TPP>(c[k]);
    if (valid)
    {
        float o[4] = {c[0], c[1], c[2], c[3]};
#pragma unroll
        for (int q = 1; q < 4; q++)
            if (sub == (unsigned int)q)
                for (int r = 0; r < 4; r++)
                    o[r] = c[4 * q + r];
        a.C[(size_t)i * 4 + sub] = make_float4(a.d.gscale[4 * sub] * o[0],
                                               a.d.gscale[4 * sub + 1] * o[1],
                                               a.d.gscale[4 * sub + 2] * o[2],
                                               a.d.gscale[4 * sub + 3] * o[3]);
    }
}

// ------------------------------------------------------------------------------ gradient
// Per 16-site tile: the forward/backward of site_eval, the tangent along C, its reverse pass,
// and the weight gradients as MFMA outer products over the tile's sites (K = sites). The
// outer-product operands are transposed through a per-wave LDS buffer: stored [feature][site]
// in bf16, an A operand (rows = features, K = sites) and a B operand (K = sites, columns =
// features) are both 4 consecutive bf16 of one feature row.
struct GradWaveLds
{
    unsigned short d2[MLP_H * 16], th2[MLP_H * 16], ta1[MLP_H * 16], a1[MLP_H * 16];
    unsigned short d1[MLP_H * 16], th1[MLP_H * 16], ct[MLP_NK * 16], x[MLP_NK * 16];
};
struct GradLds
{
    MlpLds m;
    GradWaveLds w[4];
    float red[MLP_NPARAM];
};

__device__ __forceinline__ void put_tile(unsigned short* buf, int m, int s, int g, bf4 v)
{
#pragma unroll
    for (int r = 0; r < 4; r++)
        buf[(16 * m + 4 * g + r) * 16 + s] = (unsigned short)v[r];
}

__device__ __forceinline__ bf4 get_row4(const unsigned short* buf, int row, int g)
{
    return *(const bf4*)(buf + row * 16 + 4 * g);
}

__global__ __launch_bounds__(256) void mlp_grad_kernel(MlpArgs a)
{
    __shared__ GradLds S;
    stage_weights(S.m, a);
    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x / 64;
    const int s = lane & 15, g = lane >> 4;
    const float b3 = a.master[MP_B3];
    const unsigned short* W1 = S.m.w + MB_W1;
    const unsigned short* W2 = S.m.w + MB_W2;
    const unsigned short* W2T = S.m.w + MB_W2T;
    GradWaveLds& B = S.w[wv];

    f4v gW2[4][4], gW1[4], gb1[4], gb2[4], gw3[4];
#pragma unroll
    for (int m = 0; m < 4; m++)
    {
        gW1[m] = zero4();
        gb1[m] = zero4();
        gb2[m] = zero4();
        gw3[m] = zero4();
#pragma unroll
        for (int n = 0; n < 4; n++)
            gW2[m][n] = zero4();
    }
    const unsigned int ntiles = (a.N + 15) / 16;
    const unsigned int nwaves = gridDim.x * 4;
    for (unsigned int t = blockIdx.x * 4 + wv; t < ntiles; t += nwaves)
    {
        const unsigned int site = 16 * t + s;
        const bool valid = site < a.N;
        const float4 gv = valid ? a.G[(size_t)site * 4 + g] : make_float4(0.f, 0.f, 0.f, 0.f);
        // a padding site carries a zero adjoint, so it adds nothing to any gradient
        const float4 cv = valid ? a.C[(size_t)site * 4 + g] : make_float4(0.f, 0.f, 0.f, 0.f);
        SiteTile T;
        site_eval(S.m, a, gv, s, g, b3, T);
        const bf4 cb = pack4(cv.x, cv.y, cv.z, cv.w);
        // tangent: t1 = W1 c, ta1 = silu'(h1) t1, t2 = W2 ta1
        f4v t1[4];
        bf4 ta1b[4];
#pragma unroll
        for (int m = 0; m < 4; m++)
        {
            t1[m] = mfma(ldsA(W1, MLP_NK, m, 0, s, g), cb, zero4());
            f4v v;
#pragma unroll
            for (int r = 0; r < 4; r++)
                v[r] = silu_d_s(T.h1[m][r], T.s1[m][r]) * t1[m][r];
            ta1b[m] = pack4(v);
        }
        bf4 th2b[4];
#pragma unroll
        for (int m = 0; m < 4; m++)
        {
            f4v t2 = zero4();
#pragma unroll
            for (int kb = 0; kb < 4; kb++)
                t2 = mfma(ldsA(W2, MLP_H, m, kb, s, g), ta1b[kb], t2);
            f4v th2;
#pragma unroll
            for (int r = 0; r < 4; r++)
            {
                const int o = 16 * m + 4 * g + r;
                const float h = T.h2[m][r];
                const float sg = sigm(h);
                gw3[m][r] += silu_d_s(h, sg) * t2[r];
                th2[r] = S.m.w3[o] * silu_dd_s(h, sg) * t2[r];
                gb2[m][r] += th2[r];
            }
            th2b[m] = pack4(th2);
        }
        // v = W2^T Th2; Th1 = u silu''(h1) t1 + v silu'(h1)
        bf4 th1b[4];
#pragma unroll
        for (int m = 0; m < 4; m++)
        {
            f4v vv = zero4();
#pragma unroll
            for (int kb = 0; kb < 4; kb++)
                vv = mfma(ldsA(W2T, MLP_H, m, kb, s, g), th2b[kb], vv);
            f4v th1;
#pragma unroll
            for (int r = 0; r < 4; r++)
            {
                const float h = T.h1[m][r], sg = T.s1[m][r];
                th1[r] = T.u[m][r] * silu_dd_s(h, sg) * t1[m][r] + vv[r] * silu_d_s(h, sg);
                gb1[m][r] += th1[r];
            }
            th1b[m] = pack4(th1);
        }
        // outer products over the tile's sites, operands transposed through LDS
#pragma unroll
        for (int m = 0; m < 4; m++)
        {
            put_tile(B.d2, m, s, g, T.d2b[m]);
            put_tile(B.th2, m, s, g, th2b[m]);
            put_tile(B.ta1, m, s, g, ta1b[m]);
            put_tile(B.a1, m, s, g, T.a1b[m]);
            put_tile(B.d1, m, s, g, T.d1b[m]);
            put_tile(B.th1, m, s, g, th1b[m]);
        }
        put_tile(B.ct, 0, s, g, cb);
        put_tile(B.x, 0, s, g, T.xb);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const bf4 bct = get_row4(B.ct, s, g);
        const bf4 bx = get_row4(B.x, s, g);
#pragma unroll
        for (int m = 0; m < 4; m++)
        {
            const bf4 ad2 = get_row4(B.d2, 16 * m + s, g);
            const bf4 ath2 = get_row4(B.th2, 16 * m + s, g);
#pragma unroll
            for (int n = 0; n < 4; n++)
            {
                gW2[m][n] = mfma(ad2, get_row4(B.ta1, 16 * n + s, g), gW2[m][n]);
                gW2[m][n] = mfma(ath2, get_row4(B.a1, 16 * n + s, g), gW2[m][n]);
            }
            gW1[m] = mfma(get_row4(B.d1, 16 * m + s, g), bct, gW1[m]);
            gW1[m] = mfma(get_row4(B.th1, 16 * m + s, g), bx, gW1[m]);
        }
        // the next tile overwrites the buffers: all lanes are done reading
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    // vectors: sum over the 16 sites (lanes s) of each group
#pragma unroll
    for (int m = 0; m < 4; m++)
#pragma unroll
        for (int r = 0; r < 4; r++)
        {
            gb1[m][r] = group_sum<16>(gb1[m][r]);
            gb2[m][r] = group_sum<16>(gb2[m][r]);
            gw3[m][r] = group_sum<16>(gw3[m][r]);
        }
    // workgroup sum in wave order into S.red, then one row of gpart
    for (int w = 0; w < 4; w++)
    {
        if (wv == w)
        {
#pragma unroll
            for (int m = 0; m < 4; m++)
#pragma unroll
                for (int r = 0; r < 4; r++)
                {
                    const int row = 16 * m + 4 * g + r;
#pragma unroll
                    for (int n = 0; n < 4; n++)
                    {
                        float& d = S.red[MP_W2 + row * MLP_H + 16 * n + s];
                        d = (w == 0 ? 0.f : d) + gW2[m][n][r];
                    }
                    float& d1 = S.red[MP_W1 + row * MLP_NK + s];
                    d1 = (w == 0 ? 0.f : d1) + gW1[m][r];
                    if (s == 0)
                    {
                        float& e1 = S.red[MP_B1 + row];
                        e1 = (w == 0 ? 0.f : e1) + gb1[m][r];
                        float& e2 = S.red[MP_B2 + row];
                        e2 = (w == 0 ? 0.f : e2) + gb2[m][r];
                        float& e3 = S.red[MP_W3 + row];
                        e3 = (w == 0 ? 0.f : e3) + gw3[m][r];
                    }
                }
            if (threadIdx.x == 64 * w)
                S.red[MP_B3] = 0.f;
        }
        __syncthreads();
    }
    float* out = a.gpart + (size_t)blockIdx.x * MLP_NPARAM;
    for (unsigned int p = threadIdx.x; p < (unsigned int)MLP_NPARAM; p += blockDim.x)
        out[p] = S.red[p];
}

__global__ __launch_bounds__(256) void mlp_loss_kernel(const float* part, unsigned int n,
                                                       double* out)
{
    __shared__ double red[256];
    double t = 0.0;
    for (unsigned int k = threadIdx.x; k < n; k += 256)
        t += (double)part[k];
    red[threadIdx.x] = t;
    __syncthreads();
    for (unsigned int w = 128; w > 0; w >>= 1)
    {
        if (threadIdx.x < w)
            red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        out[0] = red[0];
}

__device__ __forceinline__ void write_bf16_copy(uint16_t* wb, unsigned int p, float v)
{
    const unsigned short b = f2bf(v);
    if (p < (unsigned int)MP_B1)
    {
        const unsigned int j = p / MLP_NK, k = p % MLP_NK;
        wb[MB_W1 + j * MLP_NK + k] = b;
        wb[MB_W1T + k * MLP_H + j] = b;
    }
    else if (p >= (unsigned int)MP_W2 && p < (unsigned int)MP_B2)
    {
        const unsigned int q = p - MP_W2, o = q / MLP_H, i = q % MLP_H;
        wb[MB_W2 + o * MLP_H + i] = b;
        wb[MB_W2T + i * MLP_H + o] = b;
    }
}

__global__ __launch_bounds__(256) void mlp_adam_kernel(float* master, float* m1, float* m2,
                                                       const float* gpart, uint16_t* wb, float lr,
                                                       float beta1, float beta2, float eps,
                                                       float c1, float c2)
{
    const unsigned int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (unsigned int)MLP_NPARAM)
        return;
    float gsum = 0.f;
    for (unsigned int w = 0; w < MLP_GRAD_BLOCKS; w++)
        gsum += gpart[(size_t)w * MLP_NPARAM + p];
    const float mm = beta1 * m1[p] + (1.f - beta1) * gsum;
    const float vv = beta2 * m2[p] + (1.f - beta2) * gsum * gsum;
    m1[p] = mm;
    m2[p] = vv;
    const float v = master[p] - lr * (mm / c1) / (sqrtf(vv / c2) + eps);
    master[p] = v;
    write_bf16_copy(wb, p, v);
}

// minibatch: gacc += the sum of one frame's gradient rows (gacc cleared first when first)
__global__ __launch_bounds__(256) void mlp_grad_accum_kernel(const float* gpart, float* gacc,
                                                             int first)
{
    const unsigned int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (unsigned int)MLP_NPARAM)
        return;
    float g = first ? 0.f : gacc[p];
    for (unsigned int w = 0; w < MLP_GRAD_BLOCKS; w++)
        g += gpart[(size_t)w * MLP_NPARAM + p];
    gacc[p] = g;
}

// Adam on a summed gradient, scaled (1 / frames of the minibatch)
__global__ __launch_bounds__(256) void mlp_adam_acc_kernel(float* master, float* m1, float* m2,
                                                           const float* gacc, float scale,
                                                           uint16_t* wb, float lr, float beta1,
                                                           float beta2, float eps, float c1,
                                                           float c2)
{
    const unsigned int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (unsigned int)MLP_NPARAM)
        return;
    const float gsum = gacc[p] * scale;
    const float mm = beta1 * m1[p] + (1.f - beta1) * gsum;
    const float vv = beta2 * m2[p] + (1.f - beta2) * gsum * gsum;
    m1[p] = mm;
    m2[p] = vv;
    const float v = master[p] - lr * (mm / c1) / (sqrtf(vv / c2) + eps);
    master[p] = v;
    write_bf16_copy(wb, p, v);
}

__global__ __launch_bounds__(256) void mlp_grad_sum_kernel(const float* gpart, float* grad)
{
    const unsigned int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (unsigned int)MLP_NPARAM)
        return;
    float gsum = 0.f;
    for (unsigned int w = 0; w < MLP_GRAD_BLOCKS; w++)
        gsum += gpart[(size_t)w * MLP_NPARAM + p];
    grad[p] = gsum;
}

__global__ __launch_bounds__(256) void mlp_pack_kernel(const float* master, uint16_t* wb)
{
    const unsigned int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < (unsigned int)MLP_NPARAM)
        write_bf16_copy(wb, p, master[p]);
}

__global__ __launch_bounds__(256) void mlp_load_frame_kernel(unsigned int N,
                                                             const unsigned int* rtag,
                                                             const float* fpos,
                                                             const float* fforce, DevBox box,
                                                             double4* pos, float4* target)
{
    const unsigned int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= N)
        return;
    const unsigned int i = rtag[t];
    double4 p = pos[i];
    double x[3] = {(double)fpos[3 * (size_t)t], (double)fpos[3 * (size_t)t + 1],
                   (double)fpos[3 * (size_t)t + 2]};
    for (int d = 0; d < 3; d++)
    {
        // FP32 rounding may land a coordinate on the upper face
        if (x[d] >= box.hi[d])
            x[d] -= box.L[d];
        if (x[d] < box.lo[d])
            x[d] += box.L[d];
    }
    p.x = x[0];
    p.y = x[1];
    p.z = x[2];
    pos[i] = p;
    if (target)
        target[i] = make_float4(fforce[3 * (size_t)t], fforce[3 * (size_t)t + 1],
                                fforce[3 * (size_t)t + 2], 0.f);
}

__global__ __launch_bounds__(256) void map_sites_kernel(SiteMapArgs a)
{
    const unsigned int I = blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= a.nsites)
        return;
    const unsigned int b0 = a.row[I], b1 = a.row[I + 1];
    const double4 p0 = a.pos[a.rtag[a.tags[b0]]];
    double cx = 0.0, cy = 0.0, cz = 0.0, mt = 0.0;
    float fx = 0.f, fy = 0.f, fz = 0.f;
    for (unsigned int b = b0; b < b1; b++)
    {
        const unsigned int i = a.rtag[a.tags[b]];
        const double4 p = a.pos[i];
        const double m = a.vel[i].w;
        double dx = p.x - p0.x, dy = p.y - p0.y, dz = p.z - p0.z;
        min_image(a.box, dx, dy, dz);
        cx += m * dx;
        cy += m * dy;
        cz += m * dz;
        mt += m;
        for (unsigned int q = 0; q < a.nforces; q++)
        {
            const float4 f = a.f[q][i];
            fx += f.x;
            fy += f.y;
            fz += f.z;
        }
    }
    double c[3] = {p0.x + cx / mt, p0.y + cy / mt, p0.z + cz / mt};
    for (int d = 0; d < 3; d++)
    {
        c[d] -= a.box.L[d] * floor((c[d] - a.box.lo[d]) * a.box.Linv[d]);
        if (c[d] >= a.box.hi[d])
            c[d] -= a.box.L[d];
        a.out_pos[3 * (size_t)I + d] = (float)c[d];
    }
    if (a.out_force)
    {
        a.out_force[3 * (size_t)I] = fx;
        a.out_force[3 * (size_t)I + 1] = fy;
        a.out_force[3 * (size_t)I + 2] = fz;
    }
}

// Frame forces minus harmonic bond forces, per (frame, site): out_i = F_i - sum over the bonded
// partners j of k (r0 / r - 1) d, d = minimum image of p_i - p_j (PotentialBond with
// EvaluatorBondHarmonic, hoomd/md/EvaluatorBondHarmonic.h:105-123, force on the first endpoint;
// a zero distance gives no force, as the evaluator zeroes a non-finite force_divr).
__global__ __launch_bounds__(256) void frames_sub_harmonic_kernel(
    unsigned int nframes, unsigned int nsites, const float* __restrict__ fpos,
    const float* fforce, float* out, const unsigned int* __restrict__ row,
    const unsigned int* __restrict__ partner, double k, double r0, DevBox box)
{
    const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (size_t)nframes * nsites)
        return;
    const size_t f0 = g - g % nsites;  // first site of this frame
    const unsigned int i = (unsigned int)(g % nsites);
    const float* pf = fpos + 3 * f0;
    double fx = fforce[3 * g], fy = fforce[3 * g + 1], fz = fforce[3 * g + 2];
    for (unsigned int q = row[i]; q < row[i + 1]; q++)
    {
        const unsigned int j = partner[q];
        double dx = (double)pf[3 * (size_t)i] - pf[3 * (size_t)j];
        double dy = (double)pf[3 * (size_t)i + 1] - pf[3 * (size_t)j + 1];
        double dz = (double)pf[3 * (size_t)i + 2] - pf[3 * (size_t)j + 2];
        min_image(box, dx, dy, dz);
        const double r = sqrt(dx * dx + dy * dy + dz * dz);
        double fdr = k * (r0 / r - 1.0);
        if (!isfinite(fdr))
            fdr = 0.0;
        fx -= fdr * dx;
        fy -= fdr * dy;
        fz -= fdr * dz;
    }
    out[3 * g] = (float)fx;
    out[3 * g + 1] = (float)fy;
    out[3 * g + 2] = (float)fz;
}

unsigned int site_blocks(unsigned int N)
{
    // one workgroup of 4 wavefronts per CU, each wavefront walking its tiles of 16 sites: the
    // kernel's time is its matrix and SiLU work per SIMD, which this spreads evenly (a grid of
    // 1.5 wavefronts per SIMD measured 15 % slower, and every extra workgroup stages the weights
    // again); never more wavefronts than tiles
    static int cus[16] = {0};
    int dev = 0;
    (void)hipGetDevice(&dev);
    int& c = cus[dev & 15];
    if (c <= 0)
    {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess
            || v <= 0)
            v = 256;
        c = v;
    }
    const unsigned int tiles = (N + 15) / 16;
    const unsigned int want = (tiles + 3) / 4;
    return want < 1 ? 1 : (want > (unsigned int)c ? (unsigned int)c : want);
}

} // namespace

unsigned int mlp_pair_blocks(unsigned int N)
{
    // a multiple of 8 for the XCD-aware block order
    return round_up((unsigned int)(((size_t)N * PAIR_TPP + MLP_PAIR_BLOCK - 1) / MLP_PAIR_BLOCK),
                    8u);
}

__global__ __launch_bounds__(256) void pos_f32_kernel(unsigned int N, const double4* pos,
                                                      float4* pos32)
{
    const unsigned int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n < N)
    {
        const double4 p = pos[n];
        pos32[n] = make_float4((float)p.x, (float)p.y, (float)p.z, 0.f);
    }
}

hipError_t launch_pos_f32(unsigned int N, const double4* pos, float4* pos32, hipStream_t s)
{
    if (N == 0)
        return hipSuccess;
    hipLaunchKernelGGL(pos_f32_kernel, dim3((N + 255) / 256), dim3(256), 0, s, N, pos, pos32);
    return hipGetLastError();
}

hipError_t launch_mlp_descriptor(const MlpArgs& a, hipStream_t s)
{
    if (a.N == 0)
        return hipSuccess;
    // two neighbours per iteration in flight (one: 44 us at 200 000 sites; three or four: no
    // faster than two)
    if (a.pos32)
        hipLaunchKernelGGL((mlp_descriptor_kernel<2, true>), dim3(mlp_pair_blocks(a.N)),
                           dim3(MLP_PAIR_BLOCK), 0, s, a);
    else
        hipLaunchKernelGGL((mlp_descriptor_kernel<2, false>), dim3(mlp_pair_blocks(a.N)),
                           dim3(MLP_PAIR_BLOCK), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_mlp_site(const MlpArgs& a, hipStream_t s)
{
    if (a.N == 0)
        return hipSuccess;
    hipLaunchKernelGGL(mlp_site_kernel, dim3(site_blocks(a.N)), dim3(SITE_BLOCK), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_mlp_force(const MlpArgs& a, bool train, hipStream_t s)
{
    if (a.N == 0)
        return hipSuccess;
    // two neighbours per iteration in flight (three: 141 VGPRs, 3 wavefronts per SIMD, slower)
    const dim3 g(mlp_pair_blocks(a.N)), b(MLP_PAIR_BLOCK);
    const int v = (train ? 4 : 0) | (a.Ab ? 2 : 0) | (a.pos32 ? 1 : 0);
    switch (v)
    {
    case 0: hipLaunchKernelGGL((mlp_force_kernel<false, 2, false, false>), g, b, 0, s, a); break;
    case 1: hipLaunchKernelGGL((mlp_force_kernel<false, 2, false, true>), g, b, 0, s, a); break;
    case 2: hipLaunchKernelGGL((mlp_force_kernel<false, 2, true, false>), g, b, 0, s, a); break;
    case 3: hipLaunchKernelGGL((mlp_force_kernel<false, 2, true, true>), g, b, 0, s, a); break;
    case 4: hipLaunchKernelGGL((mlp_force_kernel<true, 2, false, false>), g, b, 0, s, a); break;
    case 5: hipLaunchKernelGGL((mlp_force_kernel<true, 2, false, true>), g, b, 0, s, a); break;
    case 6: hipLaunchKernelGGL((mlp_force_kernel<true, 2, true, false>), g, b, 0, s, a); break;
    default: hipLaunchKernelGGL((mlp_force_kernel<true, 2, true, true>), g, b, 0, s, a); break;
    }
    return hipGetLastError();
}

hipError_t launch_mlp_backward(const MlpArgs& a, hipStream_t s)
{
    if (a.N == 0)
        return hipSuccess;
    hipLaunchKernelGGL(mlp_adjoint_kernel, dim3(mlp_pair_blocks(a.N)), dim3(MLP_PAIR_BLOCK), 0, s,
                       a);
    hipLaunchKernelGGL(mlp_grad_kernel, dim3(MLP_GRAD_BLOCKS), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_mlp_loss(const float* loss_partial, unsigned int n, double* out, hipStream_t s)
{
    hipLaunchKernelGGL(mlp_loss_kernel, dim3(1), dim3(256), 0, s, loss_partial, n, out);
    return hipGetLastError();
}

hipError_t launch_mlp_adam(float* master, float* m1, float* m2, const float* gpart,
                           uint16_t* wb, float lr, float beta1, float beta2, float eps,
                           unsigned int t, hipStream_t s)
{
    const float c1 = (float)(1.0 - std::pow((double)beta1, (double)t));
    const float c2 = (float)(1.0 - std::pow((double)beta2, (double)t));
    hipLaunchKernelGGL(mlp_adam_kernel, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, s, master,
                       m1, m2, gpart, wb, lr, beta1, beta2, eps, c1, c2);
    return hipGetLastError();
}

hipError_t launch_mlp_grad_accum(const float* gpart, float* gacc, bool first, hipStream_t s)
{
    hipLaunchKernelGGL(mlp_grad_accum_kernel, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, s,
                       gpart, gacc, first ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_mlp_adam_acc(float* master, float* m1, float* m2, const float* gacc,
                               float scale, uint16_t* wb, float lr, float beta1, float beta2,
                               float eps, unsigned int t, hipStream_t s)
{
    const float c1 = (float)(1.0 - std::pow((double)beta1, (double)t));
    const float c2 = (float)(1.0 - std::pow((double)beta2, (double)t));
    hipLaunchKernelGGL(mlp_adam_acc_kernel, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, s,
                       master, m1, m2, gacc, scale, wb, lr, beta1, beta2, eps, c1, c2);
    return hipGetLastError();
}

hipError_t launch_mlp_grad_sum(const float* gpart, float* grad, hipStream_t s)
{
    hipLaunchKernelGGL(mlp_grad_sum_kernel, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, s,
                       gpart, grad);
    return hipGetLastError();
}

hipError_t launch_mlp_pack(const float* master, uint16_t* wb, hipStream_t s)
{
    hipLaunchKernelGGL(mlp_pack_kernel, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, s, master,
                       wb);
    return hipGetLastError();
}

hipError_t launch_mlp_load_frame(unsigned int N, const unsigned int* rtag, const float* fpos,
                                 const float* fforce, const DevBox& box, double4* pos,
                                 float4* target, hipStream_t s)
{
    if (N == 0)
        return hipSuccess;
    hipLaunchKernelGGL(mlp_load_frame_kernel, dim3((N + 255) / 256), dim3(256), 0, s, N, rtag,
                       fpos, fforce, box, pos, target);
    return hipGetLastError();
}

hipError_t launch_map_sites(const SiteMapArgs& a, hipStream_t s)
{
    if (a.nsites == 0)
        return hipSuccess;
    hipLaunchKernelGGL(map_sites_kernel, dim3((a.nsites + 255) / 256), dim3(256), 0, s, a);
    return hipGetLastError();
}

} // namespace hm

// ====================================================================== C ABI (kernel drivers)
#include "capi_util.h"

namespace hm
{
MlpDescr make_mlp_descr(const hm_mlp_cg_model& m)
{
    MlpDescr d;
    d.rcut = (float)m.r_cut;
    d.rcutsq = (float)(m.r_cut * m.r_cut);
    d.inv_rcut = (float)(1.0 / m.r_cut);
    d.mu0 = (float)m.mu0;
    d.dmu = (float)m.dmu;
    d.inv2w2 = (float)(1.0 / (2.0 * m.width * m.width));
    d.invw2 = (float)(1.0 / (m.width * m.width));
    for (int k = 0; k < MLP_NK; k++)
    {
        d.gmean[k] = (float)m.gmean[k];
        d.gscale[k] = (float)m.gscale[k];
    }
    // recurrence from centre GMID: g_GMID = exp(-b (t - GMID)^2) over t = (r - mu0) / dmu for r in
    // [0, rcut] must stay above exp(-80) (FP32 normals end near exp(-87)); the ratios then stay
    // below exp(2 sqrt(80 b)) <= exp(80), and every product is a Gaussian value <= 1
    const double b = m.dmu * m.dmu / (2.0 * m.width * m.width);
    const double t0 = -m.mu0 / m.dmu, t1 = (m.r_cut - m.mu0) / m.dmu;
    const double far = std::max(std::fabs(t0 - GMID), std::fabs(t1 - GMID));
    d.rebase = (m.dmu > 0.0 && b * far * far <= 80.0 && 2.0 * std::sqrt(80.0 * b) <= 80.0) ? 1 : 0;
    d.inv_dmu = m.dmu > 0.0 ? (float)(1.0 / m.dmu) : 0.f;
    d.b = (float)b;
    d.q = (float)std::exp(-2.0 * b);
    return d;
}
} // namespace hm

namespace
{
using namespace hm;

bool model_ok(const hm_mlp_cg_model* m)
{
    return m && m->r_cut > 0.0 && m->width > 0.0 && std::isfinite(m->r_cut)
           && std::isfinite(m->width);
}

MlpArgs to_args(const hm_mlp_cg_args* x, const hm_mlp_cg_model* m)
{
    MlpArgs a;
    memset(&a, 0, sizeof(a));
    a.N = x->N;
    a.pos = (const double4*)x->pos;
    a.box = make_devbox(x->box);
    a.nlist = x->nlist;
    a.nmax = x->nmax;
    a.n_neigh = x->n_neigh;
    a.d = make_mlp_descr(*m);
    a.master = x->params;
    a.wb = x->params_bf16;
    a.G = (float4*)x->G;
    a.A = (float4*)x->A;
    a.Esite = x->Esite;
    a.force = (float4*)x->force;
    a.virial = x->virial;
    a.target = (const float4*)x->target;
    a.err = (float4*)x->err;
    a.C = (float4*)x->C;
    a.loss_partial = x->loss_partial;
    a.gpart = x->grad_partial;
    a.loss_scale = x->N ? (float)(2.0 / (3.0 * (double)x->N)) : 0.f;
    a.Ab = (uint2*)x->A_bf16;
    a.pos32 = (const float4*)x->pos32;
    return a;
}
} // namespace

extern "C"
{

unsigned int hm_mlp_cg_loss_blocks(unsigned int N)
{
    return mlp_pair_blocks(N);
}

int hm_mlp_cg_pack(const float* params, uint16_t* params_bf16, hm_stream_t stream)
{
    if (!params || !params_bf16)
        return fail(HM_ERR_ARGS, "hm_mlp_cg_pack: invalid arguments");
    return check(launch_mlp_pack(params, params_bf16, (hipStream_t)stream), "hm_mlp_cg_pack");
}

int hm_mlp_cg_descriptor(const hm_mlp_cg_args* x, const hm_mlp_cg_model* m, hm_stream_t stream)
{
    if (!x || !model_ok(m) || (x->N && (!x->pos || !x->nlist || !x->n_neigh || !x->G)))
        return fail(HM_ERR_ARGS, "hm_mlp_cg_descriptor: invalid arguments");
    return check(launch_mlp_descriptor(to_args(x, m), (hipStream_t)stream),
                 "hm_mlp_cg_descriptor");
}

int hm_mlp_cg_forward(const hm_mlp_cg_args* x, const hm_mlp_cg_model* m, hm_stream_t stream)
{
    if (!x || !model_ok(m)
        || (x->N && (!x->G || (!x->A && !x->A_bf16) || !x->Esite || !x->params
                      || !x->params_bf16)))
        return fail(HM_ERR_ARGS, "hm_mlp_cg_forward: invalid arguments");
    return check(launch_mlp_site(to_args(x, m), (hipStream_t)stream), "hm_mlp_cg_forward");
}

int hm_mlp_cg_forces(const hm_mlp_cg_args* x, const hm_mlp_cg_model* m, int train,
                     hm_stream_t stream)
{
    if (!x || !model_ok(m)
        || (x->N
            && (!x->pos || !x->nlist || !x->n_neigh || (!x->A && !x->A_bf16) || !x->Esite
                || !x->force))
        || (train && x->N && (!x->target || !x->err || !x->loss_partial)))
        return fail(HM_ERR_ARGS, "hm_mlp_cg_forces: invalid arguments");
    return check(launch_mlp_force(to_args(x, m), train != 0, (hipStream_t)stream),
                 "hm_mlp_cg_forces");
}

int hm_mlp_cg_backward(const hm_mlp_cg_args* x, const hm_mlp_cg_model* m, hm_stream_t stream)
{
    if (!x || !model_ok(m)
        || (x->N && (!x->pos || !x->nlist || !x->n_neigh || !x->G || !x->err || !x->C
                     || !x->params || !x->params_bf16))
        || !x->grad_partial)
        return fail(HM_ERR_ARGS, "hm_mlp_cg_backward: invalid arguments");
    MlpArgs a = to_args(x, m);
    if (a.N == 0)
        return check(hipMemsetAsync(x->grad_partial, 0,
                                    sizeof(float) * MLP_NPARAM * MLP_GRAD_BLOCKS,
                                    (hipStream_t)stream),
                     "hm_mlp_cg_backward");
    return check(launch_mlp_backward(a, (hipStream_t)stream), "hm_mlp_cg_backward");
}

int hm_mlp_cg_loss(const float* loss_partial, unsigned int N, double* out, hm_stream_t stream)
{
    if (!loss_partial || !out)
        return fail(HM_ERR_ARGS, "hm_mlp_cg_loss: invalid arguments");
    return check(launch_mlp_loss(loss_partial, mlp_pair_blocks(N), out, (hipStream_t)stream),
                 "hm_mlp_cg_loss");
}

int hm_mlp_cg_grad(const float* grad_partial, float* grad, hm_stream_t stream)
{
    if (!grad_partial || !grad)
        return fail(HM_ERR_ARGS, "hm_mlp_cg_grad: invalid arguments");
    return check(launch_mlp_grad_sum(grad_partial, grad, (hipStream_t)stream), "hm_mlp_cg_grad");
}

int hm_mlp_cg_adam(float* params, float* m1, float* m2, const float* grad_partial,
                   uint16_t* params_bf16, double lr, double beta1, double beta2, double eps,
                   unsigned int t, hm_stream_t stream)
{
    if (!params || !m1 || !m2 || !grad_partial || !params_bf16 || t == 0)
        return fail(HM_ERR_ARGS, "hm_mlp_cg_adam: invalid arguments");
    return check(launch_mlp_adam(params, m1, m2, grad_partial, params_bf16, (float)lr,
                                 (float)beta1, (float)beta2, (float)eps, t, (hipStream_t)stream),
                 "hm_mlp_cg_adam");
}

int hm_frames_sub_harmonic(unsigned int nframes, unsigned int nsites, const float* frame_pos,
                           const float* frame_force, float* out_force, const unsigned int* row,
                           const unsigned int* partner, double k, double r0, hm_box box,
                           hm_stream_t stream)
{
    if ((nframes && nsites) && (!frame_pos || !frame_force || !out_force || !row || !partner))
        return fail(HM_ERR_ARGS, "hm_frames_sub_harmonic: invalid arguments");
    const size_t n = (size_t)nframes * nsites;
    if (n == 0)
        return HM_OK;
    hipLaunchKernelGGL(frames_sub_harmonic_kernel, dim3((unsigned int)((n + 255) / 256)),
                       dim3(256), 0, (hipStream_t)stream, nframes, nsites, frame_pos,
                       frame_force, out_force, row, partner, k, r0, make_devbox(box));
    return check(hipGetLastError(), "hm_frames_sub_harmonic");
}

} // extern "C"
