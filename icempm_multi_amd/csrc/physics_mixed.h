// physics_mixed.h -- FP32 device math of the mixed-precision form.
//
// Same model as physics_fp64.h (simulation/kernels.cuh:24-44, 547-619), with
// the state kept where FP32 keeps its precision:
//   * F is stored as D = F - I.  det F - 1 = tr D + det D and F F^T - I =
//     D + D^T + D D^T are formed from the small D, so stress and (p, q) keep
//     FP32 relative accuracy even when the strain is 1e-5;
//   * positions are fixed point, x/h = xq * 2^-frac_bits - origin, so the
//     resolution is uniform over the grid (2^-19 cells at 4096 columns) and
//     the B-spline offset fx is exact;
//   * the 2x2 SVD and the yield-surface projection of crushed points, and
//     the grid velocity division, run in FP64 (physics_fp64.h).
#pragma once

#include <hip/hip_runtime.h>
#include "mpm_common.h"

namespace icempm {
namespace mixed {

constexpr float kCoeff1f = 1.41421356237309505f;

// Float constants of a scene, converted once per tile from SceneConsts.
struct ConstsF {
    float h, hinv, dt, dt_vol_Dpinv, pmass, wmass, mu, kappa;
    float ics, its, iss, its2, beta, grain_var, s4;   // s4 = 4/h
    float dq_scale;                                   // dt * hinv * 2^frac_bits
    int fb, origin, origin_y;
};

__device__ __forceinline__ ConstsF to_float(const SceneConsts &c)
{
    ConstsF f;
    f.h = (float)c.h;
    f.hinv = (float)c.hinv;
    f.dt = (float)c.dt;
    f.dt_vol_Dpinv = (float)c.dt_vol_Dpinv;
    f.pmass = (float)c.pmass;
    f.wmass = (float)c.wmass;
    f.mu = (float)c.mu;
    f.kappa = (float)c.kappa;
    f.ics = (float)c.ics;
    f.its = (float)c.its;
    f.iss = (float)c.iss;
    f.its2 = (float)c.its2;
    f.beta = (float)c.beta;
    f.grain_var = (float)c.grain_var;
    f.s4 = (float)(4. * c.hinv);
    f.dq_scale = (float)(c.dt * c.hinv * (double)(1LL << c.frac_bits));
    f.fb = c.frac_bits;
    f.origin = c.origin;
    f.origin_y = c.origin_y;
    return f;
}

// Base node and offset from a fixed-point coordinate: base = (int)(x/h - 0.5)
// (C truncation, as kernels.cuh:90), fx = x/h - base, both exact.
__device__ __forceinline__ void base_fx_fixed(unsigned q, int fb, int origin, int &base, float &fx)
{
    const long long v = (long long)q - ((long long)origin << fb) - (1LL << (fb - 1));
    base = v >= 0 ? (int)(v >> fb) : -(int)((-v) >> fb);
    const long long frac = (long long)q - ((long long)(base + origin) << fb);
    fx = (float)frac * __int_as_float((127 - fb) << 23);   // * 2^-fb
}

__device__ __forceinline__ void weights(float fx, float w[3])
{
    const float v0 = 1.5f - fx, v1 = fx - 1.0f, v2 = fx - 0.5f;
    w[0] = 0.5f * v0 * v0;
    w[1] = 0.75f - v1 * v1;
    w[2] = 0.5f * v2 * v2;
}

// J - 1 and dev(F F^T) from D = F - I; dev = [[d0, d1], [d1, -d0]].
__device__ __forceinline__ void strain_measures(const float D[4], float &Jm1, float &d0, float &d1)
{
    Jm1 = (D[0] + D[3]) + (D[0] * D[3] - D[2] * D[1]);
    const float e00 = 2.f * D[0] + D[0] * D[0] + D[1] * D[1];
    const float e11 = 2.f * D[3] + D[2] * D[2] + D[3] * D[3];
    const float e01 = D[1] + D[2] + D[0] * D[2] + D[1] * D[3];
    d0 = 0.5f * (e00 - e11);
    d1 = e01;
}

// Kirchhoff stress (kernels.cuh:24-34): mu/J dev(F F^T) + kappa/2 (J^2-1) I
__device__ __forceinline__ void kirchhoff(const ConstsF &c, const float D[4], float tau[4])
{
    float Jm1, d0, d1;
    strain_measures(D, Jm1, d0, d1);
    const float J = 1.f + Jm1;
    const float a = c.mu / J;
    const float k = 0.5f * c.kappa * Jm1 * (J + 1.f);
    tau[0] = a * d0 + k;
    tau[1] = a * d1;
    tau[2] = a * d1;
    tau[3] = -a * d0 + k;
}

__device__ __forceinline__ void water(const ConstsF &c, float J, float tau[4])
{
    const float k = c.kappa * (1.f - __powf(J, -5.f));
    tau[0] = k; tau[1] = 0; tau[2] = 0; tau[3] = k;
}

// ComputePQ (kernels.cuh:547-557)
__device__ __forceinline__ void compute_pq(const ConstsF &c, const float D[4], float &p, float &q)
{
    float Jm1, d0, d1;
    strain_measures(D, Jm1, d0, d1);
    const float J = 1.f + Jm1;
    p = -0.5f * c.kappa * Jm1 * (J + 1.f);
    q = kCoeff1f * c.mu / J * sqrtf(2.f * (d0 * d0 + d1 * d1));
}

// CheckIfPointIsInsideFailureSurface (kernels.cuh:560-619) in FP32
__device__ __forceinline__ bool failure(const ConstsF &c, int grain, float p, float q)
{
    const float var2 = 1.0f + c.grain_var * 0.033f * (float)(-15 + (grain + 3) % 30);
    const float var3 = 1.0f + c.grain_var * 0.1f * (float)(-10 + (grain + 4) % 11);
    const float pmax = c.ics, pmin = -c.its;
    const float qmax = c.iss * var3;
    const float pmin2 = -c.its2 * var2;
    const float beta = c.beta;
    const float dp = pmax - pmin;
    if (p < 0) {
        if (p < pmin2) return true;
        const float q0 = 2.f * sqrtf(-pmax * pmin) * qmax / dp;
        const float q_limit = -q0 / pmin2 * (p - pmin2);
        return q > q_limit;
    }
    const float mSq = 4.f * qmax * qmax * (1.f + 2.f * beta) / (dp * dp);
    const float y = (1.f + 2.f * beta) * q * q + mSq * (p + beta * pmax) * (p - pmax);
    return y > 0;
}

}  // namespace mixed
}  // namespace icempm
