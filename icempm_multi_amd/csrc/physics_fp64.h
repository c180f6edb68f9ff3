// physics_fp64.h -- FP64 device math of one MPM point, restated from the
// reference's kernels (simulation/kernels.cuh, helper_math.cuh, givens.cuh).
//
// Matrices are row-major double[4] (a00, a01, a10, a11).  Operation order
// follows the reference's Eigen expressions.
#pragma once

#include <hip/hip_runtime.h>
#include "mpm_common.h"

namespace icempm {
namespace fp64 {

constexpr double kPi = 3.14159265358979323846;       // parameters_sim.h:22
constexpr double kCoeff1 = 1.4142135623730950;       // sqrt((6-d)/2), kernels.cuh:10
constexpr double kCoeff1Inv = 0.7071067811865475;    // kernels.cuh:11

__device__ __forceinline__ double det2(const double m[4]) { return m[0] * m[3] - m[2] * m[1]; }

__device__ __forceinline__ void mul2(const double a[4], const double b[4], double c[4])
{
    const double r0 = a[0] * b[0] + a[1] * b[2];
    const double r1 = a[0] * b[1] + a[1] * b[3];
    const double r2 = a[2] * b[0] + a[3] * b[2];
    const double r3 = a[2] * b[1] + a[3] * b[3];
    c[0] = r0; c[1] = r1; c[2] = r2; c[3] = r3;
}

// deviatoric part of F F^T (kernels.cuh:31, 591-594)
__device__ __forceinline__ void dev_ffT(const double F[4], double d[4])
{
    const double b0 = F[0] * F[0] + F[1] * F[1];
    const double b1 = F[0] * F[2] + F[1] * F[3];
    const double b2 = F[2] * F[0] + F[3] * F[1];
    const double b3 = F[2] * F[2] + F[3] * F[3];
    const double hh = (b0 + b3) / 2;
    d[0] = b0 - hh; d[1] = b1; d[2] = b2; d[3] = b3 - hh;
}

// quadratic B-spline weights (kernels.cuh:99-103)
__device__ __forceinline__ void weights(double fx, double w[3])
{
    const double v0 = 1.5 - fx, v1 = fx - 1.0, v2 = fx - 0.5;
    w[0] = 0.5 * v0 * v0;
    w[1] = 0.75 - v1 * v1;
    w[2] = 0.5 * v2 * v2;
}

// Terms that ComputePQ (G2P) and the Kirchhoff stress (P2G) share for an
// intact point, whose F the constitutive step leaves unchanged: 1/Je,
// dev(F F^T) and p = -kappa/2 (Je^2-1).
struct StressTerms {
    double inv_je, d[4], p;
    bool valid;
};

// Kirchhoff stress of Wolper 2019 (kernels.cuh:24-34) from the shared terms:
// the same operations on the same values as the form below.
__device__ __forceinline__ void kirchhoff(const SceneConsts &c, const StressTerms &st, double tau[4])
{
    const double a = c.mu * st.inv_je;
    const double k = -st.p;
    tau[0] = a * st.d[0] + k;
    tau[1] = a * st.d[1];
    tau[2] = a * st.d[2];
    tau[3] = a * st.d[3] + k;
}

// Kirchhoff stress of Wolper 2019 (kernels.cuh:24-34)
__device__ __forceinline__ void kirchhoff(const SceneConsts &c, const double F[4], double tau[4])
{
    const double Je = det2(F);
    double d[4];
    dev_ffT(F, d);
    const double a = c.mu * (1 / Je);
    const double k = c.kappa_half * (Je * Je - 1.);
    tau[0] = a * d[0] + k;
    tau[1] = a * d[1];
    tau[2] = a * d[2];
    tau[3] = a * d[3] + k;
}

// water (kernels.cuh:37-44): kappa (1 - J^-5) I, with J^-5 formed as
// 1/(J^2 J^2 J) (within a few ulps of pow(J, -5.0), whose exp/log polynomial
// constants would occupy vector registers in the tile kernels)
__device__ __forceinline__ void water(const SceneConsts &c, double J, double tau[4])
{
    const double j2 = J * J;
    const double k = c.kappa * (1. - 1. / (j2 * j2 * J));
    tau[0] = k; tau[1] = 0; tau[2] = 0; tau[3] = k;
}

// Base node (truncating cast) and fractional offset (kernels.cuh:90-92, 344-346).
__device__ __forceinline__ void base_fx(const SceneConsts &c, double x, int &b, double &fx)
{
    const double s = x * c.hinv;
    b = (int)(s - 0.5);
    fx = s - (double)b;
}

// Error byte for a stencil leaving the grid (kernels.cuh:94-97); bxc = base.x + halo.
__device__ __forceinline__ unsigned stencil_error(const SceneConsts &c, int bxc, int by)
{
    unsigned e = 0;
    if (bxc < 0) e = 70;
    if (by < 0) e = 71;
    if (bxc > c.ncols - 3) e = 72;
    if (by > c.gridY - 3) e = 73;
    return e;
}

// Velocity of one grid node with mass (kernels.cuh:178-300, live lines): p/m,
// gravity, the vmax clamp, the moving indenter (its reaction force binned by
// angle into `force`, unless force is null), and the walls.  gix, giy are
// global node indices.
__device__ __forceinline__ void node_velocity(const SceneConsts &c, const SceneClock &clk, double *force, int gix,
                                              int giy, double mass, double px, double py, double &vx_out,
                                              double &vy_out)
{
    const double gnx = (double)gix * c.h, gny = (double)giy * c.h;
    double vx = px / mass, vy = py / mass;
    vy -= c.dt_gravity;
    const double sq = vx * vx + vy * vy;
    if (sq > c.vmax_sq) {
        const double r = sqrt(sq);
        vx = vx / r * c.vmax;
        vy = vy / r * c.vmax;
    }
    double n0 = gnx - clk.ind_x, n1 = gny - clk.ind_y;
    if (n0 * n0 + n1 * n1 < c.ind_rsq) {
        const double vr0 = vx - c.ind_vel, vr1 = vy;
        const double z = n0 * n0 + n1 * n1;
        if (z > 0) { const double r = sqrt(z); n0 = n0 / r; n1 = n1 / r; }
        const double vn = vr0 * n0 + vr1 * n1;
        if (vn < 0) {
            const double vt0 = vr0 - n0 * vn, vt1 = vr1 - n1 * vn;
            const double pv0 = vx, pv1 = vy;
            vx = c.ind_vel + vt0;
            vy = 0 + vt1;
            if (force) {
                const double f0 = (pv0 - vx) * mass / c.dt, f1 = (pv1 - vy) * mass / c.dt;
                float angle = atan2f((float)n0, (float)n1);
                angle += kPi;
                angle *= c.n_ind_sub / (2 * kPi);
                const int index = min(max((int)angle, 0), c.n_ind_sub - 1);
                atomicAdd(&force[2 * index], f0);
                atomicAdd(&force[2 * index + 1], f1);
            }
        }
    }
    if (giy <= 2) { vx = 0; vy = 0; }
    if (giy >= c.gridY - 3 && vy > 0) vy = 0;
    if (gix <= 2 && vx < 0) vx = 0;
    else if (gix >= c.gridXTotal - 3 && vx > 0) vx = 0;
    vx_out = vx;
    vy_out = vy;
}

// ---- constitutive update of partition_kernel_g2p (kernels.cuh:383-390) ----

struct Givens { double c, s; };

// polar_decomposition (helper_math.cuh:26-46), column-major input
__device__ __forceinline__ void polar(const double a[4], Givens &r, double s[4])
{
    const double x0 = a[0] + a[3], x1 = a[1] - a[2];
    const double den = sqrt(x0 * x0 + x1 * x1);
    r.c = 1; r.s = 0;
    if (den != 0) { r.c = x0 / den; r.s = -x1 / den; }
    for (int d = 0; d < 2; d++) {
        const double t1 = a[0 + 2 * d], t2 = a[1 + 2 * d];
        s[0 + 2 * d] = r.c * t1 - r.s * t2;
        s[1 + 2 * d] = r.s * t1 + r.c * t2;
    }
}

// 2x2 SVD (helper_math.cuh:80-140 via kernels.cuh:521-538): F = U diag(S) V^T,
// U = [[u.c, u.s], [-u.s, u.c]], V likewise, S[0] >= S[1].
__device__ __forceinline__ void svd2x2(const double F[4], double U[4], double S[2], double V[4])
{
    const double a[4] = {F[0], F[2], F[1], F[3]};
    Givens u, v;
    double ss[4];
    polar(a, u, ss);
    const double x = ss[0], y = ss[2], z = ss[3];
    const double y2 = y * y;
    double cs, sn;
    if (y2 == 0) {
        cs = 1; sn = 0; S[0] = x; S[1] = z;
    } else {
        const double tau = 0.5 * (x - z);
        const double w = sqrt(tau * tau + y2);
        const double t = tau > 0 ? y / (tau + w) : y / (tau - w);
        cs = 1.0 / sqrt(t * t + 1.0);
        sn = -t * cs;
        const double c2 = cs * cs, csy = 2 * cs * sn * y, s2 = sn * sn;
        S[0] = c2 * x - csy + s2 * z;
        S[1] = s2 * x + csy + c2 * z;
    }
    if (S[0] < S[1]) {
        const double tmp = S[0]; S[0] = S[1]; S[1] = tmp;
        v.c = -sn; v.s = cs;
    } else {
        v.c = cs; v.s = sn;
    }
    const double nc = u.c * v.c - u.s * v.s;
    const double ns = u.s * v.c + u.c * v.s;
    U[0] = nc; U[1] = ns; U[2] = -ns; U[3] = nc;
    V[0] = v.c; V[1] = v.s; V[2] = -v.s; V[3] = v.c;
}

// (U diag(s)) V^T (kernels.cuh:642, 671)
__device__ __forceinline__ void usvt(const double U[4], double s0, double s1, const double V[4], double F[4])
{
    const double UD[4] = {U[0] * s0, U[1] * s1, U[2] * s0, U[3] * s1};
    const double Vt[4] = {V[0], V[2], V[1], V[3]};
    mul2(UD, Vt, F);
}

// ComputePQ (kernels.cuh:547-557); also leaves the terms the stress shares.
__device__ __forceinline__ void compute_pq(const SceneConsts &c, const double F[4], double &Je, double &p,
                                           double &q, StressTerms &st)
{
    Je = det2(F);
    p = -c.kappa_half * (Je * Je - 1.);
    dev_ffT(F, st.d);
    const double *d = st.d;
    const double nrm = sqrt(d[0] * d[0] + d[2] * d[2] + d[1] * d[1] + d[3] * d[3]);
    st.inv_je = 1. / Je;
    q = c.coeff1_mu * st.inv_je * nrm;
    st.p = p;
    st.valid = true;
}

__device__ __forceinline__ void compute_pq(const SceneConsts &c, const double F[4], double &Je, double &p,
                                           double &q)
{
    StressTerms st;
    compute_pq(c, F, Je, p, q, st);
}

// CheckIfPointIsInsideFailureSurface + GetParametersForGrain (kernels.cuh:560-619).
// The grain-dependent mSq, q0 and pmin2 come from the scene's table `tab`
// (SceneDev::fail_tab, computed on the host with the same operations; an LDS
// copy in the tile kernels), indexed by the two remainders of the grain id.
template <class Tab>
__device__ __forceinline__ bool failure(const SceneConsts &c, Tab tab, int grain, double p, double q)
{
    const int r11 = (grain + 4) % 11 + 10, r30 = (grain + 3) % 30 + 29;
    const double pmin2 = tab[2 * kFailTab11 + r30];
    if (p < 0) {
        if (p < pmin2) return true;
        const double q0 = tab[2 * r11 + 1];
        const double k = -q0 / pmin2;
        const double q_limit = k * (p - pmin2);
        return q > q_limit;
    }
    const double mSq = tab[2 * r11];
    const double y = c.b2 * q * q + mSq * (p + c.bpmax) * (p - c.ics);
    return y > 0;
}

// ComputeSVD + Wolper_Drucker_Prager (kernels.cuh:540-545, 623-674)
__device__ __forceinline__ void drucker_prager(const SceneConsts &c, double F[4], double &Jp_inv, double Je,
                                               double p, double q)
{
    double U[4], S[2], V[4];
    svd2x2(F, U, S, V);
    const double s2[2] = {S[0] * S[0], S[1] * S[1]};
    const double half = (s2[0] + s2[1]) / 2;
    const double f = c.mu / Je;
    const double sh[2] = {f * (s2[0] - half), f * (s2[1] - half)};

    if (p < -c.dp_threshold_p || Jp_inv < 1) {
        // p_new = -DP_threshold_p: Je_new and its root are scene constants
        const double Je_new = c.je_tension;
        const double sq = c.sq_je_tension;
        usvt(U, sq, sq, V, F);
        Jp_inv *= Je_new / Je;
    } else {
        const double pmax = c.ics, qmax = c.iss;
        double q_n_1;
        if (p > pmax) {
            q_n_1 = 0;
        } else {
            const double q_from_dp = (p + c.dp_threshold_p) * c.tan_phi;
            const double pmin = -c.its;
            const double q_fs = 2 * sqrt((pmax - p) * (p - pmin)) * qmax / c.dp;
            q_n_1 = fmin(q_fs, q_from_dp);
        }
        if (q >= q_n_1) {
            const double s_norm = q_n_1 * kCoeff1Inv;
            double n0 = sh[0], n1 = sh[1];
            const double z = n0 * n0 + n1 * n1;
            if (z > 0) { const double r = sqrt(z); n0 = n0 / r; n1 = n1 / r; }
            const double g = s_norm * (Je / c.mu);
            const double tr = (s2[0] + s2[1]) / 2;
            usvt(U, sqrt(g * n0 + tr), sqrt(g * n1 + tr), V, F);
        }
    }
}

// The constitutive branch of partition_kernel_g2p for one point (kernels.cuh:378-391).
// A = I + dt*C has been applied to F already.
// The terms the stress shares are left in st (st.valid false when the
// projection changed F or the point is liquid).
template <class Tab>
__device__ __forceinline__ void constitutive(const SceneConsts &c, Tab ftab, double F[4], double &Jp_inv,
                                             long long &ud, const double A[4], double &p_out, double &q_out,
                                             StressTerms &st)
{
    st.valid = false;
    if (ud & kLiquid) {
        Jp_inv *= det2(A);
        p_out = 0; q_out = 0;
        return;
    }
    double Je, p, q;
    compute_pq(c, F, Je, p, q, st);
    p_out = p; q_out = q;
    if (!(ud & kCrushed))
        if (failure(c, ftab, (short)ud, p, q)) ud |= kCrushed;
    if (ud & kCrushed) {
        st.valid = false;
        drucker_prager(c, F, Jp_inv, Je, p, q);
    }
}

}  // namespace fp64
}  // namespace icempm
