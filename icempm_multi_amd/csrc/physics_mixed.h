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

// Float constants of a scene (set by the host).
__device__ __forceinline__ const ConstsF &to_float(const SceneConsts &c) { return c.f; }

// Base node and offset from a fixed-point coordinate: base = (int)(x/h - 0.5)
// (C truncation, as kernels.cuh:90), fx = x/h - base, both exact.
// In 32-bit arithmetic: q - off is formed on whichever side is non-negative,
// and frac = x/h - base (in [-0.5, 1.5) * 2^fb) is exact modulo 2^32.
__device__ __forceinline__ void base_fx_fixed(unsigned q, int fb, int origin, int &base, float &fx)
{
    const unsigned off = ((unsigned)origin << fb) + (1u << (fb - 1));   // (origin + 1/2) * 2^fb
    base = q >= off ? (int)((q - off) >> fb) : -(int)((off - q) >> fb);
    const int frac = (int)(q - ((unsigned)(base + origin) << fb));
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

// Stress terms of an intact point, shared by ComputePQ in G2P and the
// Kirchhoff stress in P2G when the constitutive step leaves D unchanged:
// a = mu/J, dev(F F^T) = [[d0, d1], [d1, -d0]], k = kappa/2 (J^2-1) = -p.
struct StressTerms {
    float a, d0, d1, k;
    bool valid;
};

__device__ __forceinline__ void stress_terms(const ConstsF &c, const float D[4], StressTerms &st)
{
    float Jm1;
    strain_measures(D, Jm1, st.d0, st.d1);
    const float J = 1.f + Jm1;
    // v_rcp_f32 (1 ulp) in place of the IEEE divide's ten-instruction sequence
    st.a = c.mu * __builtin_amdgcn_rcpf(J);
    st.k = 0.5f * c.kappa * Jm1 * (J + 1.f);
    st.valid = true;
}

// Kirchhoff stress (kernels.cuh:24-34): mu/J dev(F F^T) + kappa/2 (J^2-1) I
__device__ __forceinline__ void kirchhoff(const StressTerms &st, float tau[4])
{
    tau[0] = st.a * st.d0 + st.k;
    tau[1] = st.a * st.d1;
    tau[2] = st.a * st.d1;
    tau[3] = -st.a * st.d0 + st.k;
}

__device__ __forceinline__ void kirchhoff(const ConstsF &c, const float D[4], float tau[4])
{
    StressTerms st;
    stress_terms(c, D, st);
    kirchhoff(st, tau);
}

// Water (kernels.cuh:37-44) from q = J - 1: 1 - J^-5 = -expm1(-5 log1p(q)),
// which keeps its relative precision for the small q of a liquid point.
__device__ __forceinline__ void water(const ConstsF &c, float q, float tau[4])
{
    const float k = -c.kappa * expm1f(-5.f * log1pf(q));
    tau[0] = k; tau[1] = 0; tau[2] = 0; tau[3] = k;
}

// ComputePQ (kernels.cuh:547-557) from the stress terms: p = -k,
// q = sqrt(2) (mu/J) |dev(F F^T)|, with the 1-ulp v_sqrt_f32 (q only meets
// the failure test, which FP32 rounding already moves by more than an ulp).
__device__ __forceinline__ void compute_pq(const StressTerms &st, float &p, float &q)
{
    p = -st.k;
    q = kCoeff1f * st.a * __builtin_amdgcn_sqrtf(2.f * (st.d0 * st.d0 + st.d1 * st.d1));
}

// CheckIfPointIsInsideFailureSurface (kernels.cuh:560-619) in FP32
__device__ __forceinline__ bool failure(const ConstsF &c, int grain, float p, float q)
{
    const float var2 = 1.0f + c.gv033 * (float)(-15 + (grain + 3) % 30);
    const float var3 = 1.0f + c.gv01 * (float)(-10 + (grain + 4) % 11);
    const float qmax = c.iss * var3;
    const float pmin2 = -c.its2 * var2;
    if (p < 0) {
        if (p < pmin2) return true;
        const float q0 = c.q0c * qmax;
        // q > -q0/pmin2 (p - pmin2), multiplied through by -pmin2 > 0
        return q * -pmin2 > q0 * (p - pmin2);
    }
    const float mSq = c.msq_c * (qmax * qmax);
    const float y = c.b2 * q * q + mSq * (p + c.bpmax) * (p - c.ics);
    return y > 0;
}

}  // namespace mixed
}  // namespace icempm
