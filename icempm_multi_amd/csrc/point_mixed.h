// point_mixed.h -- per-point G2P and P2G of the mixed-precision form.
//
// Point SoA: 16 lines of 4-byte elements: 0 utility (low 32 bits), 1 P, 2 Q,
// 3 Jp_inv - 1, 4-5 fixed-point x/y, 6-7 v, 8-11 D = F - I, 12-15 C (FP32).
// Jp_inv and F are stored as their difference from 1 and I: a liquid point's
// Jp_inv moves by about dt * div v per substep, a few FP32 ulps of 1.
// Grid buffer: 3 FP64 lines (m, px, py), as in the FP64 form: the P2G sums in
// FP64 (LDS windows, then FP64 atomics into the grid); after the grid update
// lines 1 and 2 hold the node velocity.
#pragma once

#include <hip/hip_runtime.h>

#include "kernels.h"
#include "physics_fp64.h"
#include "physics_mixed.h"

namespace icempm {
namespace mixed {

typedef __attribute__((address_space(3))) float lds_f;
typedef __attribute__((address_space(3))) double lds_d;
typedef __attribute__((address_space(1))) double glb_d;
typedef __attribute__((address_space(1))) unsigned glb_u;

constexpr int kBlock = 256;
constexpr int kMaxTX = 32;
constexpr int kLineM = 0, kLinePx = 1, kLinePy = 2;   // grid lines

__device__ __forceinline__ void lds_add(lds_d *p, double v)
{
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Non-returning FP64 atomic add into the grid.
__device__ __forceinline__ void glb_add(glb_d *p, double v)
{
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Adds one node's (m, px, py) into grid g (line pitch gp).
__device__ __forceinline__ void grid_add(glb_d *g, long long gp, long long gi, double m, double px, double py)
{
    glb_add(g + kLineM * gp + gi, m);
    glb_add(g + kLinePx * gp + gi, px);
    glb_add(g + kLinePy * gp + gi, py);
}

// SoA access: uniform line base + 32-bit byte offset (saddr + voffset form).
struct PtRef {
    unsigned *P;
    long long pitch;
    __device__ __forceinline__ glb_u &u(int line, long long i) const
    {
        typedef __attribute__((address_space(1))) char glb_c;
        glb_c *base = (glb_c *)(P + line * pitch);
        return *(glb_u *)(base + ((unsigned)i << 2));
    }
    __device__ __forceinline__ float f(int line, long long i) const { return __uint_as_float(u(line, i)); }
    __device__ __forceinline__ void setf(int line, long long i, float v) const { u(line, i) = __float_as_uint(v); }
};

// Two uniforms in [0, 1) (16 bits each) from a point index and a substep
// (murmur3 finalizer).  Used to round the fixed-point position increments
// stochastically: rounding to nearest biases slowly varying increments of a
// few units every substep the same way.
__device__ __forceinline__ void unit_hash2(unsigned a, unsigned b, float &u0, float &u1)
{
    unsigned h = a * 0x9E3779B1u ^ (b * 0x85EBCA77u + 0x7F4A7C15u);
    h ^= h >> 16;
    h *= 0x85EBCA6Bu;
    h ^= h >> 13;
    h *= 0xC2B2AE35u;
    h ^= h >> 16;
    u0 = (float)(h & 0xffffu) * (1.0f / 65536.0f);
    u1 = (float)(h >> 16) * (1.0f / 65536.0f);
}

__device__ __forceinline__ unsigned stencil_error(const SceneConsts &c, int bxc, int by)
{
    unsigned e = 0;
    if (bxc < 0) e = 70;
    if (by < 0) e = 71;
    if (bxc > c.ncols - 3) e = 72;
    if (by > c.gridY - 3) e = 73;
    return e;
}

// The constitutive branch of partition_kernel_g2p (kernels.cuh:378-391):
// intact points in FP32; crushed points go through the 2x2 SVD and the
// Drucker-Prager projection in FP64.
// The stress terms of the updated D are left in st (st.valid false when the
// projection changed D or the point is liquid).
__device__ __forceinline__ void constitutive(const SceneConsts &c, const ConstsF &cf, float D[4], float &Jm1,
                                             unsigned &ud, const float A[4], float &p_out, float &q_out,
                                             StressTerms &st)
{
    st.valid = false;
    if (ud & (unsigned)kLiquid) {
        // Jp_inv *= det(I + A): (1 + q)(1 + a) - 1 = q + a + q a, a = tr A + det A
        const float a = (A[0] + A[3]) + (A[0] * A[3] - A[2] * A[1]);
        Jm1 = Jm1 + a + Jm1 * a;
        p_out = 0; q_out = 0;
        return;
    }
    float p, q;
    stress_terms(cf, D, st);
    compute_pq(st, p, q);
    p_out = p; q_out = q;
    if (!(ud & (unsigned)kCrushed))
        if (failure(cf, (short)(ud & 0xffff), p, q)) ud |= (unsigned)kCrushed;
    if (ud & (unsigned)kCrushed) {
        st.valid = false;
        double F[4] = {1.0 + (double)D[0], (double)D[1], (double)D[2], 1.0 + (double)D[3]};
        double Jpd = 1.0 + (double)Jm1, Je, pd, qd;
        fp64::compute_pq(c, F, Je, pd, qd);
        fp64::drucker_prager(c, F, Jpd, Je, pd, qd);
        D[0] = (float)(F[0] - 1.0);
        D[1] = (float)F[1];
        D[2] = (float)F[2];
        D[3] = (float)(F[3] - 1.0);
        Jm1 = (float)(Jpd - 1.0);
        p_out = (float)pd;
        q_out = (float)qd;
    }
}

// G2P of one point from an LDS window of node velocities (win_w x win_h nodes
// from grid node (wc0, wr0), node (i, j) at vw[2*(i*win_s + j)]), or from
// global memory if the stencil left the window: the velocity lines, or (vclk
// not null) node velocities formed from the momenta with that clock.
// defer_outside: if the stencil leaves the window, return false with e = 0
// and change nothing (the caller handles the point later).
__device__ __forceinline__ bool g2p_point(const SceneConsts &c, const ConstsF &cf, const glb_d *g, long long gp,
                                          const SceneClock *vclk, const lds_f *vw, int wc0, int wr0, int win_w,
                                          int win_h, int win_s, bool defer_outside, unsigned &xq, unsigned &yq, float v[2],
                                          float C[4], float D[4], float &Jm1, unsigned &ud, float &p_out,
                                          float &q_out, unsigned &e, unsigned point, unsigned step, StressTerms &st)
{
    int bx, by;
    float fx, fy;
    base_fx_fixed(xq, cf.fb, cf.origin, bx, fx);
    base_fx_fixed(yq, cf.fb, cf.origin_y, by, fy);
    const int bxc = bx - c.xoff + c.halo;
    e = stencil_error(c, bxc, by);
    if (e) return false;
    float wx[3], wy[3];
    weights(fx, wx);
    weights(fy, wy);
    const int lx = bxc - wc0, ly = by - wr0;
    const bool inwin = vw && lx >= 0 && lx + 2 < win_w && ly >= 0 && ly + 2 < win_h;
    if (!inwin && defer_outside) return false;   // e = 0
    // factored sums, as in the FP64 form: per column i, a = sum_j wy[j] v_ij
    // and b = sum_j wy[j] (j - fy) v_ij; v = sum_i wx[i] a and
    // C = (4/h) sum_i (wx[i] (i - fx) a, wx[i] b)
    const float wyd[3] = {wy[0] * (0.f - fy), wy[1] * (1.f - fy), wy[2] * (2.f - fy)};
    float vx = 0, vy = 0, c00 = 0, c01 = 0, c10 = 0, c11 = 0;
    auto column = [&](int i, float ax, float ay, float bx_, float by_) {
        const float w = wx[i], wd = wx[i] * ((float)i - fx);
        vx += w * ax;
        vy += w * ay;
        c00 += wd * ax;
        c01 += w * bx_;
        c10 += wd * ay;
        c11 += w * by_;
    };
    // separate LDS and global loops, so that the LDS gather does not wait for
    // the global loads in flight (the prefetch of the next point); the LDS
    // loop is unrolled (its 9 reads in flight together)
    if (inwin) {
#pragma unroll
        for (int i = 0; i < 3; i++) {
            float ax = 0, ay = 0, bx_ = 0, by_ = 0;
#pragma unroll
            for (int j = 0; j < 3; j++) {
                const int li = (lx + i) * win_s + ly + j;
                const float nvx = vw[2 * li], nvy = vw[2 * li + 1];   // (vx, vy) interleaved: one 64-bit LDS read
                ax += wy[j] * nvx;
                ay += wy[j] * nvy;
                bx_ += wyd[j] * nvx;
                by_ += wyd[j] * nvy;
            }
            column(i, ax, ay, bx_, by_);
        }
    } else if (vclk) {
        // momenta in global memory (the points a launch set aside, handled
        // at its end): the loads of the 9 nodes issued together, then each
        // node's velocity formed here
        double gl[9][3];
#pragma unroll
        for (int n = 0; n < 9; n++) {
            const long long gi = (long long)(bxc + n / 3) * c.gridY + by + n % 3;
#pragma unroll
            for (int l = 0; l < 3; l++) gl[n][l] = g[l * gp + gi];
        }
#pragma unroll
        for (int i = 0; i < 3; i++) {
            float ax = 0, ay = 0, bx_ = 0, by_ = 0;
#pragma unroll
            for (int j = 0; j < 3; j++) {
                const double *q = gl[3 * i + j];
                float nvx = 0, nvy = 0;
                if (q[kLineM] != 0) {
                    double dvx, dvy;
                    fp64::node_velocity(c, *vclk, nullptr, bxc + i - c.halo + c.xoff, by + j, q[kLineM],
                                        q[kLinePx], q[kLinePy], dvx, dvy);
                    nvx = (float)dvx;
                    nvy = (float)dvy;
                }
                ax += wy[j] * nvx;
                ay += wy[j] * nvy;
                bx_ += wyd[j] * nvx;
                by_ += wyd[j] * nvy;
            }
            column(i, ax, ay, bx_, by_);
        }
    } else {
#pragma unroll 1
        for (int i = 0; i < 3; i++) {
            float ax = 0, ay = 0, bx_ = 0, by_ = 0;
#pragma unroll
            for (int j = 0; j < 3; j++) {
                const long long gi = (long long)(bxc + i) * c.gridY + by + j;
                const float nvx = (float)g[kLinePx * gp + gi], nvy = (float)g[kLinePy * gp + gi];
                ax += wy[j] * nvx;
                ay += wy[j] * nvy;
                bx_ += wyd[j] * nvx;
                by_ += wyd[j] * nvy;
            }
            column(i, ax, ay, bx_, by_);
        }
    }
    v[0] = vx;
    v[1] = vy;
    C[0] = cf.s4 * c00;
    C[1] = cf.s4 * c01;
    C[2] = cf.s4 * c10;
    C[3] = cf.s4 * c11;
    // x += v dt in fixed point, rounded stochastically (unbiased)
    float ux, uy;
    unit_hash2(point, step, ux, uy);
    xq += (unsigned)__float2int_rd(v[0] * cf.dq_scale + ux);
    yq += (unsigned)__float2int_rd(v[1] * cf.dq_scale + uy);
    const float A[4] = {cf.dt * C[0], cf.dt * C[1], cf.dt * C[2], cf.dt * C[3]};
    // D <- (I + A)(I + D) - I = D + A + A D
    const float n0 = D[0] + A[0] + (A[0] * D[0] + A[1] * D[2]);
    const float n1 = D[1] + A[1] + (A[0] * D[1] + A[1] * D[3]);
    const float n2 = D[2] + A[2] + (A[2] * D[0] + A[3] * D[2]);
    const float n3 = D[3] + A[3] + (A[2] * D[1] + A[3] * D[3]);
    D[0] = n0; D[1] = n1; D[2] = n2; D[3] = n3;
    constitutive(c, cf, D, Jm1, ud, A, p_out, q_out, st);
    return true;
}

// P2G stencil of one point: base node, weights and the affine momentum
// terms.  nodes() calls f(i, j, dm, dpx, dpy) for the 9 stencil nodes, in the
// factored order of the FP64 form, widened to FP64 for the accumulators.
struct P2GStencil {
    int bxc, by;
    float fx, fy, wx[3], wy[3], S[4], mv0, mv1, mass;

    template <class F>
    __device__ __forceinline__ void nodes(const ConstsF &cf, F &&f) const
    {
        float wyd[3];
#pragma unroll
        for (int j = 0; j < 3; j++) wyd[j] = wy[j] * (((float)j - fy) * cf.h);
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const float d0 = ((float)i - fx) * cf.h;
            const float w = wx[i];
            const float wm = w * mass;
            const float px = w * (mv0 + S[0] * d0), qx = w * S[1];
            const float py = w * (mv1 + S[2] * d0), qy = w * S[3];
#pragma unroll
            for (int j = 0; j < 3; j++)
                f(i, j, (double)(wm * wy[j]), (double)(px * wy[j] + qx * wyd[j]), (double)(py * wy[j] + qy * wyd[j]));
        }
    }
};

// Stress, base node and weights of a point; returns the error byte (0 = ok).
// terms: the stress terms of D if the G2P of this launch left them valid.
__device__ __forceinline__ unsigned p2g_stencil(const SceneConsts &c, const ConstsF &cf, unsigned xq, unsigned yq,
                                                const float v[2], const float C[4], const float D[4], float Jm1,
                                                unsigned ud, const StressTerms &terms, P2GStencil &st)
{
    const bool liquid = (ud & (unsigned)kLiquid) != 0;
    st.mass = liquid ? cf.wmass : cf.pmass;
    float tau[4];
    if (liquid) water(cf, Jm1, tau);
    else if (terms.valid) kirchhoff(terms, tau);
    else kirchhoff(cf, D, tau);
#pragma unroll
    for (int k = 0; k < 4; k++) st.S[k] = st.mass * C[k] - cf.dt_vol_Dpinv * tau[k];
    int bx;
    base_fx_fixed(xq, cf.fb, cf.origin, bx, st.fx);
    base_fx_fixed(yq, cf.fb, cf.origin_y, st.by, st.fy);
    st.bxc = bx - c.xoff + c.halo;
    const unsigned e = stencil_error(c, st.bxc, st.by);
    if (e) return e;
    weights(st.fx, st.wx);
    weights(st.fy, st.wy);
    st.mv0 = v[0] * st.mass;
    st.mv1 = v[1] * st.mass;
    return 0;
}

// P2G contributions of one point; calls sink(col, row, dm, dpx, dpy) (FP64 terms).
template <class Sink>
__device__ __forceinline__ unsigned p2g_point(const SceneConsts &c, const ConstsF &cf, unsigned xq, unsigned yq,
                                              const float v[2], const float C[4], const float D[4], float Jm1,
                                              unsigned ud, Sink &&sink)
{
    P2GStencil st;
    StressTerms none;
    none.valid = false;
    const unsigned e = p2g_stencil(c, cf, xq, yq, v, C, D, Jm1, ud, none, st);
    if (e) return e;
    st.nodes(cf, [&](int i, int j, double dm, double dpx, double dpy) { sink(st.bxc + i, st.by + j, dm, dpx, dpy); });
    return 0;
}

__device__ __forceinline__ void load_state(const PtRef &R, long long pt, unsigned &xq, unsigned &yq, float D[4],
                                           float &Jm1)
{
    xq = R.u(kLinePosX, pt);
    yq = R.u(kLinePosX + 1, pt);
#pragma unroll
    for (int k = 0; k < 4; k++) D[k] = R.f(kLineFe00 + k, pt);
    Jm1 = R.f(kLineJp, pt);
}

__device__ __forceinline__ void store_state(const PtRef &R, long long pt, unsigned xq, unsigned yq, const float v[2],
                                            const float D[4], const float C[4], float Jm1, unsigned ud, bool record,
                                            float p, float q)
{
    R.u(kLinePosX, pt) = xq;
    R.u(kLinePosX + 1, pt) = yq;
    R.setf(kLineVelX, pt, v[0]);
    R.setf(kLineVelX + 1, pt, v[1]);
#pragma unroll
    for (int k = 0; k < 4; k++) {
        R.setf(kLineFe00 + k, pt, D[k]);
        R.setf(kLineBp00 + k, pt, C[k]);
    }
    R.setf(kLineJp, pt, Jm1);
    R.u(kLineUtility, pt) = ud;
    if (record) {
        R.setf(kLineP, pt, p);
        R.setf(kLineQ, pt, q);
    }
}

// State stores of the fused launch (see the FP64 form): x, y and D; Jp_inv
// and the utility word of crushed and liquid points; P, Q on recording substeps.
__device__ __forceinline__ void store_state_fused(const PtRef &R, long long pt, unsigned xq, unsigned yq,
                                                  const float D[4], float Jm1, unsigned ud, bool record, float p,
                                                  float q)
{
    R.u(kLinePosX, pt) = xq;
    R.u(kLinePosX + 1, pt) = yq;
#pragma unroll
    for (int k = 0; k < 4; k++) R.setf(kLineFe00 + k, pt, D[k]);
    if (ud & (unsigned)(kLiquid | kCrushed)) {
        R.setf(kLineJp, pt, Jm1);
        R.u(kLineUtility, pt) = ud;
    }
    if (record) {
        R.setf(kLineP, pt, p);
        R.setf(kLineQ, pt, q);
    }
}

__device__ __forceinline__ int xcd_block()
{
    const int G = gridDim.x;
    if ((G & 7) != 0) return blockIdx.x;
    return (blockIdx.x & 7) * (G >> 3) + (blockIdx.x >> 3);
}

}  // namespace mixed
}  // namespace icempm
