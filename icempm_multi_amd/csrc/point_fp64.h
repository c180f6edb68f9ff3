// point_fp64.h -- FP64 per-point G2P and P2G steps shared by the binned and
// unbinned kernels (simulation/kernels.cuh:46-122, 305-414).
#pragma once

#include <hip/hip_runtime.h>

#include "kernels.h"
#include "physics_fp64.h"

namespace icempm {
namespace fp64 {

// Explicit address spaces: LDS window accesses and global grid accesses stay
// distinct ds_* / global_* instructions (a generic pointer that may point to
// either compiles to slower flat_* instructions).
typedef __attribute__((address_space(3))) double lds_d;
typedef __attribute__((address_space(1))) double glb_d;

__device__ __forceinline__ void lds_add(lds_d *p, double v)
{
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void glb_add(glb_d *p, double v)
{
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

constexpr int kBlock = 256;
constexpr int kMaxTX = 32;

__device__ __forceinline__ long long as_ll(double d) { return __double_as_longlong(d); }
__device__ __forceinline__ double as_d(long long v) { return __longlong_as_double(v); }

// SoA access: a uniform (scalar) line base plus a 32-bit byte offset, so loads
// and stores use the saddr + voffset form (points per launch < 2^29).
struct PtRef {
    double *P;
    long long pitch;
    __device__ __forceinline__ glb_d &at(int line, long long i) const
    {
        typedef __attribute__((address_space(1))) char glb_c;
        glb_c *base = (glb_c *)(P + line * pitch);
        return *(glb_d *)(base + ((unsigned)i << 3));
    }
};

// Logical block index that keeps consecutive work items on one XCD: workgroups
// are dispatched round-robin over the 8 XCDs, so block b runs on XCD b%8.
__device__ __forceinline__ int xcd_block()
{
    const int G = gridDim.x;
    if ((G & 7) != 0) return blockIdx.x;
    return (blockIdx.x & 7) * (G >> 3) + (blockIdx.x >> 3);
}

// G2P of one point: gathers v and C (kernels.cuh:344-372), advects x, updates F
// (375-376) and runs the constitutive branch (378-391).  Returns false (and
// sets the error byte) if the stencil leaves the grid.
//
// vw: the LDS window of node velocities (null: every node from global memory),
// win_w x win_h nodes from grid node (wc0, wr0), node (i, j) at vw[2*(i*win_s + j)].
// vclk: null if grid g holds node velocities (a grid update ran); otherwise g
// holds momenta and a node read from global memory gets its velocity formed
// here with that clock (node_velocity, no force).
// defer_outside: if the stencil leaves the window, return false with e = 0
// and change nothing (the caller handles the point later).
// ftab: the scene's failure-surface table (SceneDev::fail_tab or a copy).
template <class Tab>
__device__ __forceinline__ bool g2p_point(const SceneConsts &c, Tab ftab, const glb_d *g, long long gpitch,
                                          const SceneClock *vclk, const lds_d *vw, int wc0, int wr0, int win_w,
                                          int win_h, int win_s, bool defer_outside, double &x, double &y, double v[2],
                                          double B[4], double F[4], double &Jp, long long &ud, double &p_out,
                                          double &q_out, unsigned &e, StressTerms &st)
{
    int bx, by;
    double fx, fy;
    base_fx(c, x, bx, fx);
    base_fx(c, y, by, fy);
    const int bxc = bx - c.xoff + c.halo;
    e = stencil_error(c, bxc, by);
    if (e) return false;
    double wx[3], wy[3];
    weights(fx, wx);
    weights(fy, wy);
    const int lx = bxc - wc0, ly = by - wr0;
    const bool inwin = vw && lx >= 0 && lx + 2 < win_w && ly >= 0 && ly + 2 < win_h;
    if (!inwin && defer_outside) return false;   // e = 0
    // The weight of node (i, j) is wx[i] * wy[j], so the sums separate: per
    // column i, a = sum_j wy[j] v_ij and b = sum_j wy[j] (j - fy) v_ij, then
    // v = sum_i wx[i] a, C = (4/h) sum_i (wx[i] (i - fx) a, wx[i] b)
    // (kernels.cuh:357-372 in factored order).
    const double wyd[3] = {wy[0] * (0 - fy), wy[1] * (1 - fy), wy[2] * (2 - fy)};
    double vx = 0, vy = 0, c00 = 0, c01 = 0, c10 = 0, c11 = 0;
    auto column = [&](int i, double ax, double ay, double bx, double by) {
        const double w = wx[i], wd = wx[i] * (i - fx);
        vx += w * ax;
        vy += w * ay;
        c00 += wd * ax;
        c01 += w * bx;
        c10 += wd * ay;
        c11 += w * by;
    };
    // Two separate loops: if the LDS and the global reads fed one loop, the
    // merged value would make every gather wait for all outstanding global
    // loads, including the prefetch of the next point.  The LDS loop is
    // unrolled (its 9 reads in flight together, one wait), the global one is
    // not (18 node values live at once would not fit).
    if (inwin) {
#pragma unroll
        for (int i = 0; i < 3; i++) {
            double ax = 0, ay = 0, bx = 0, by_ = 0;
#pragma unroll
            for (int j = 0; j < 3; j++) {
                const int li = (lx + i) * win_s + ly + j;
                const double nvx = vw[2 * li], nvy = vw[2 * li + 1];   // one 128-bit LDS read
                ax += wy[j] * nvx;
                ay += wy[j] * nvy;
                bx += wyd[j] * nvx;
                by_ += wyd[j] * nvy;
            }
            column(i, ax, ay, bx, by_);
        }
    } else if (vclk) {
        // momenta in global memory (the points a launch set aside, handled
        // at its end): the 27 loads of the 9 nodes issued together, then
        // each node's velocity formed here
        double gm[9], gpx[9], gpy[9];
#pragma unroll
        for (int n = 0; n < 9; n++) {
            const long long gi = (long long)(bxc + n / 3) * c.gridY + by + n % 3;
            gm[n] = g[gi];
            gpx[n] = g[gpitch + gi];
            gpy[n] = g[2 * gpitch + gi];
        }
#pragma unroll
        for (int i = 0; i < 3; i++) {
            double ax = 0, ay = 0, bx = 0, by_ = 0;
#pragma unroll
            for (int j = 0; j < 3; j++) {
                const int n = 3 * i + j;
                double nvx = 0, nvy = 0;
                if (gm[n] != 0)
                    node_velocity(c, *vclk, nullptr, bxc + i - c.halo + c.xoff, by + j, gm[n], gpx[n], gpy[n], nvx,
                                  nvy);
                ax += wy[j] * nvx;
                ay += wy[j] * nvy;
                bx += wyd[j] * nvx;
                by_ += wyd[j] * nvy;
            }
            column(i, ax, ay, bx, by_);
        }
    } else {
#pragma unroll 1
        for (int i = 0; i < 3; i++) {
            double ax = 0, ay = 0, bx = 0, by_ = 0;
#pragma unroll
            for (int j = 0; j < 3; j++) {
                const long long gi = (long long)(bxc + i) * c.gridY + by + j;
                const double nvx = g[gpitch + gi], nvy = g[2 * gpitch + gi];
                ax += wy[j] * nvx;
                ay += wy[j] * nvy;
                bx += wyd[j] * nvx;
                by_ += wyd[j] * nvy;
            }
            column(i, ax, ay, bx, by_);
        }
    }
    const double s4 = c.s4;
    v[0] = vx;
    v[1] = vy;
    B[0] = s4 * c00;
    B[1] = s4 * c01;
    B[2] = s4 * c10;
    B[3] = s4 * c11;
    x += v[0] * c.dt;
    y += v[1] * c.dt;
    const double A[4] = {1.0 + c.dt * B[0], c.dt * B[1], c.dt * B[2], 1.0 + c.dt * B[3]};
    mul2(A, F, F);
    constitutive(c, ftab, F, Jp, ud, A, p_out, q_out, st);
    return true;
}

// P2G stencil of one point (kernels.cuh:59-121): base node, weights and the
// affine momentum terms.  nodes() calls f(i, j, dm, dpx, dpy) for the 9
// stencil nodes (base + (i, j)).
struct P2GStencil {
    int bxc, by;
    double fx, fy, wx[3], wy[3], S[4], mv0, mv1, mass;

    // Node (i, j) gets W = wx[i] wy[j] times (m, mv + S (d0_i, d1_j)), with
    // d0_i = (i - fx) h and d1_j = (j - fy) h (kernels.cuh:105-121), formed as
    // W m = (wx[i] m) wy[j] and W (mv0 + S00 d0_i + S01 d1_j)
    // = (wx[i] (mv0 + S00 d0_i)) wy[j] + (wx[i] S01) (wy[j] d1_j).
    template <class F>
    __device__ __forceinline__ void nodes(const SceneConsts &c, F &&f) const
    {
        double wyd[3];
#pragma unroll
        for (int j = 0; j < 3; j++) wyd[j] = wy[j] * ((j - fy) * c.h);
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const double d0 = (i - fx) * c.h;
            const double w = wx[i];
            const double wm = w * mass;
            const double px = w * (mv0 + S[0] * d0), qx = w * S[1];
            const double py = w * (mv1 + S[2] * d0), qy = w * S[3];
#pragma unroll
            for (int j = 0; j < 3; j++) f(i, j, wm * wy[j], px * wy[j] + qx * wyd[j], py * wy[j] + qy * wyd[j]);
        }
    }
};

// Stress, base node and weights of a point; returns the error byte (0 = ok).
// terms: the stress terms of F if the G2P of this launch left them valid.
__device__ __forceinline__ unsigned p2g_stencil(const SceneConsts &c, double x, double y, const double v[2],
                                                const double B[4], const double F[4], double Jp, long long ud,
                                                const StressTerms &terms, P2GStencil &st)
{
    const bool liquid = (ud & kLiquid) != 0;
    st.mass = liquid ? c.wmass : c.pmass;
    double tau[4];
    if (liquid) water(c, Jp, tau);
    else if (terms.valid) kirchhoff(c, terms, tau);
    else kirchhoff(c, F, tau);
#pragma unroll
    for (int k = 0; k < 4; k++) st.S[k] = st.mass * B[k] - c.dt_vol_Dpinv * tau[k];
    int bx;
    base_fx(c, x, bx, st.fx);
    base_fx(c, y, st.by, st.fy);
    st.bxc = bx - c.xoff + c.halo;
    const unsigned e = stencil_error(c, st.bxc, st.by);
    if (e) return e;
    weights(st.fx, st.wx);
    weights(st.fy, st.wy);
    st.mv0 = v[0] * st.mass;
    st.mv1 = v[1] * st.mass;
    return 0;
}

// P2G contributions of one point.  Calls sink(col, row, dm, dpx, dpy) for
// the 9 stencil nodes; returns the error byte (0 = ok).
template <class Sink>
__device__ __forceinline__ unsigned p2g_point(const SceneConsts &c, double x, double y, const double v[2],
                                              const double B[4], const double F[4], double Jp, long long ud,
                                              Sink &&sink)
{
    P2GStencil st;
    StressTerms none;
    none.valid = false;
    const unsigned e = p2g_stencil(c, x, y, v, B, F, Jp, ud, none, st);
    if (e) return e;
    st.nodes(c, [&](int i, int j, double dm, double dpx, double dpy) { sink(st.bxc + i, st.by + j, dm, dpx, dpy); });
    return 0;
}

__device__ __forceinline__ void load_state(const PtRef &R, long long pt, double &x, double &y, double F[4],
                                           double &Jp)
{
    x = R.at(kLinePosX, pt);
    y = R.at(kLinePosX + 1, pt);
#pragma unroll
    for (int k = 0; k < 4; k++) F[k] = R.at(kLineFe00 + k, pt);
    Jp = R.at(kLineJp, pt);
}

__device__ __forceinline__ void store_state(const PtRef &R, long long pt, double x, double y, const double v[2],
                                            const double F[4], const double B[4], double Jp, long long ud,
                                            bool record, double p, double q)
{
    R.at(kLinePosX, pt) = x;
    R.at(kLinePosX + 1, pt) = y;
    R.at(kLineVelX, pt) = v[0];
    R.at(kLineVelX + 1, pt) = v[1];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        R.at(kLineFe00 + k, pt) = F[k];
        R.at(kLineBp00 + k, pt) = B[k];
    }
    R.at(kLineJp, pt) = Jp;
    R.at(kLineUtility, pt) = as_d(ud);
    if (record) {
        R.at(kLineP, pt) = p;
        R.at(kLineQ, pt) = q;
    }
}

// State stores of the fused G2P(n)+P2G(n+1) launch.  v and C go to P2G in
// registers, and the next launch's G2P computes them afresh from the grid:
// they are read from memory only by a call's first P2G and by the download,
// both of which follow a plain G2P launch that stores them.  So the fused
// launch stores x and F, Jp_inv and the utility word only for crushed and
// liquid points (the only ones whose Jp_inv or flags change), and P, Q on
// recording substeps.
__device__ __forceinline__ void store_state_fused(const PtRef &R, long long pt, double x, double y, const double F[4],
                                                  double Jp, long long ud, bool record, double p, double q)
{
    R.at(kLinePosX, pt) = x;
    R.at(kLinePosX + 1, pt) = y;
#pragma unroll
    for (int k = 0; k < 4; k++) R.at(kLineFe00 + k, pt) = F[k];
    if (ud & (kLiquid | kCrushed)) {   // the only points whose Jp_inv or flags can change
        R.at(kLineJp, pt) = Jp;
        R.at(kLineUtility, pt) = as_d(ud);
    }
    if (record) {
        R.at(kLineP, pt) = p;
        R.at(kLineQ, pt) = q;
    }
}

}  // namespace fp64
}  // namespace icempm
