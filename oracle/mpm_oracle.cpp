/*
 * mpm_oracle.cpp -- TEST INFRASTRUCTURE ONLY: the FP64 CPU oracle.
 *
 * A restatement, in plain C++17 (+ optional OpenMP), of the reference's MPM
 * substep (igorg520b/iceMPM_multi, read as text).  Each function cites the
 * reference lines it follows and keeps the reference's branch structure and
 * operation order; Eigen expressions are expanded by hand with Eigen 3.4
 * semantics (normalized() of a zero vector returns it unchanged, norm() =
 * sqrt(squaredNorm()), cast<int> truncates, 2x2 determinant a00*a11-a10*a01).
 * It is built with -ffp-contract=off, so no operation is fused.
 *
 * Determinism: P2G scatters in two colour phases over 4-column strips, each
 * strip in point-index order, so results do not depend on the thread count.
 *
 * What pins each function (tests/test_oracle_*.py):
 *   orc_params_*          hand-computed constants for the reference's p1.json
 *   orc_weights           partition of unity, first and second moments
 *   orc_kirchhoff_stress  F=I -> 0; F=sI closed form; pure shear closed form
 *   orc_water_stress      J=1 -> 0; closed form kappa(1-J^-5)
 *   orc_compute_pq        closed forms for F=sI and simple shear
 *   orc_grain_params / orc_failure_check  known-answer table over grains 0..29
 *   orc_polar2x2 / orc_svd2x2  reconstruction, orthogonality, det=+1, sorting,
 *                         special cases, numpy.linalg.svd singular values
 *   orc_point_plasticity  tension branch: det F' = Je', p(F') = -p_t;
 *                         shear branch: q(F') = q*, det F' = Je; hydrostatic case
 *   orc_p2g               mass and momentum conservation, brute-force double
 *                         loop written independently in the test
 *   orc_update_nodes      free fall, vmax clamp, each wall rule, indenter
 *   orc_g2p               uniform field, affine field (APIC exactness)
 *   orc_substeps          mass conservation, free-fall momentum balance
 *   orc_cell_key / orc_sort  brute-force integer keys, numpy stable argsort
 */
#include "mpm_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <numeric>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr double kPi = 3.14159265358979323846;        // parameters_sim.h:22
constexpr double kCoeff1 = 1.4142135623730950;        // kernels.cuh:10
constexpr double kCoeff1Inv = 0.7071067811865475;     // kernels.cuh:11
constexpr int64_t kCrushed = ICEMPM_STATUS_CRUSHED;
constexpr int64_t kDisabled = ICEMPM_STATUS_DISABLED;
constexpr int64_t kLiquid = ICEMPM_STATUS_LIQUID;

inline int64_t as_i64(double d) { int64_t v; std::memcpy(&v, &d, 8); return v; }
inline double as_f64(int64_t v) { double d; std::memcpy(&d, &v, 8); return d; }

// row-major 2x2: m[0]=a00 m[1]=a01 m[2]=a10 m[3]=a11
inline double det2(const double m[4]) { return m[0] * m[3] - m[2] * m[1]; }

// A*B
inline void mul2(const double a[4], const double b[4], double c[4])
{
    double r[4];
    r[0] = a[0] * b[0] + a[1] * b[2];
    r[1] = a[0] * b[1] + a[1] * b[3];
    r[2] = a[2] * b[0] + a[3] * b[2];
    r[3] = a[2] * b[1] + a[3] * b[3];
    std::memcpy(c, r, sizeof r);
}

// F*F^T (Eigen lazy product, k summed 0 then 1)
inline void ffT(const double F[4], double b[4])
{
    b[0] = F[0] * F[0] + F[1] * F[1];
    b[1] = F[0] * F[2] + F[1] * F[3];
    b[2] = F[2] * F[0] + F[3] * F[1];
    b[3] = F[2] * F[2] + F[3] * F[3];
}

// dev(A) = A - tr(A)/2 I   (kernels.cuh:591-594)
inline void dev2(const double A[4], double D[4])
{
    const double h = (A[0] + A[3]) / 2;
    D[0] = A[0] - h; D[1] = A[1]; D[2] = A[2]; D[3] = A[3] - h;
}

// U * diag(s) * V^T, evaluated as (U*diag(s))*V^T (kernels.cuh:642,671)
inline void usvt(const double U[4], const double s[2], const double V[4], double F[4])
{
    const double UD[4] = {U[0] * s[0], U[1] * s[1], U[2] * s[0], U[3] * s[1]};
    const double Vt[4] = {V[0], V[2], V[1], V[3]};
    mul2(UD, Vt, F);
}

struct PointState {
    double pos[2], vel[2], Bp[4], Fe[4];
    double Jp_inv;
    int grain;   // (short)utility_data, kernels.cuh:341
    double p_tr, q_tr, Je_tr;
    double U[4], V[4], vSigma[2], vSigmaSquared[2], v_s_hat_tr[2];
    int64_t utility_data;
};

// kernels.cuh:547-557
void compute_pq(const icempm_simparams &P, PointState &p)
{
    const double *F = p.Fe;
    p.Je_tr = det2(F);
    p.p_tr = -(P.kappa / 2.) * (p.Je_tr * p.Je_tr - 1.);
    const double J = p.Je_tr;
    double b[4], d[4];
    ffT(F, b);
    dev2(b, d);
    const double nrm = std::sqrt(d[0] * d[0] + d[2] * d[2] + d[1] * d[1] + d[3] * d[3]);
    p.q_tr = kCoeff1 * P.mu * (1. / J) * nrm;
}

// kernels.cuh:560-576
void grain_params(const icempm_simparams &P, int grain, double &pmin, double &pmax, double &qmax,
                  double &beta, double &mSq, double &pmin2)
{
    const double var2 = 1.0 + P.GrainVariability * 0.033 * (-15 + (grain + 3) % 30);
    const double var3 = 1.0 + P.GrainVariability * 0.1 * (-10 + (grain + 4) % 11);
    pmax = P.IceCompressiveStrength;
    pmin = -P.IceTensileStrength;
    qmax = P.IceShearStrength * var3;
    pmin2 = -P.IceTensileStrength2 * var2;
    beta = P.NACC_beta;
    mSq = (4 * qmax * qmax * (1 + 2 * beta)) / ((pmax - pmin) * (pmax - pmin));
}

// kernels.cuh:598-619
void failure_check(const icempm_simparams &P, PointState &p)
{
    double beta, M_sq, pmin, pmax, qmax, pmin2;
    grain_params(P, p.grain, pmin, pmax, qmax, beta, M_sq, pmin2);
    if (p.p_tr < 0) {
        if (p.p_tr < pmin2) { p.utility_data |= kCrushed; return; }
        const double q0 = 2 * std::sqrt(-pmax * pmin) * qmax / (pmax - pmin);
        const double k = -q0 / pmin2;
        const double q_limit = k * (p.p_tr - pmin2);
        if (p.q_tr > q_limit) { p.utility_data |= kCrushed; return; }
    } else {
        const double y = (1. + 2. * beta) * p.q_tr * p.q_tr + M_sq * (p.p_tr + beta * pmax) * (p.p_tr - pmax);
        if (y > 0) p.utility_data |= kCrushed;
    }
}

// GivensRotation (givens.cuh:30-176): only c, s are carried.
struct Givens { double c = 1, s = 0; };

// polar_decomposition (helper_math.cuh:26-46) on a column-major 2x2 `a`.
void polar(const double a[4], Givens &r, double s[4])
{
    const double x0 = a[0] + a[3], x1 = a[1] - a[2];
    const double denominator = std::sqrt(x0 * x0 + x1 * x1);
    r.c = 1; r.s = 0;
    if (denominator != 0) { r.c = x0 / denominator; r.s = -x1 / denominator; }
    for (int i = 0; i < 4; ++i) s[i] = a[i];
    // mat_rotation<2> with rowi=0, rowk=1 (givens.cuh:120-128)
    for (int d = 0; d < 2; d++) {
        const double tau1 = s[0 + d * 2], tau2 = s[1 + d * 2];
        s[0 + d * 2] = r.c * tau1 - r.s * tau2;
        s[1 + d * 2] = r.s * tau1 + r.c * tau2;
    }
}

// singular_value_decomposition (helper_math.cuh:80-140); aa row-major.
void svd_givens(const double aa[4], Givens &u, double sigma[2], Givens &v)
{
    double s_sym[4];
    const double a[4] = {aa[0], aa[2], aa[1], aa[3]};
    polar(a, u, s_sym);
    double cosine, sine;
    const double x = s_sym[0], y = s_sym[2], z = s_sym[3];
    const double y2 = y * y;
    if (y2 == 0) {
        cosine = 1; sine = 0;
        sigma[0] = x; sigma[1] = z;
    } else {
        const double tau = 0.5 * (x - z);
        const double w = std::sqrt(tau * tau + y2);
        double t;
        if (tau > 0) t = y / (tau + w);
        else t = y / (tau - w);
        cosine = 1.0 / std::sqrt(t * t + 1.0);
        sine = -t * cosine;
        const double c2 = cosine * cosine;
        const double csy = 2 * cosine * sine * y;
        const double s2 = sine * sine;
        sigma[0] = c2 * x - csy + s2 * z;
        sigma[1] = s2 * x + csy + c2 * z;
    }
    if (sigma[0] < sigma[1]) {
        std::swap(sigma[0], sigma[1]);
        v.c = -sine; v.s = cosine;
    } else {
        v.c = cosine; v.s = sine;
    }
    // u *= v (givens.cuh:161-166)
    const double nc = u.c * v.c - u.s * v.s;
    const double ns = u.s * v.c + u.c * v.s;
    u.c = nc; u.s = ns;
}

// svd2x2 (kernels.cuh:521-538): fill<2> writes {c, s, -s, c} and the Eigen
// comma initializer reads it row-major, so U = [[c, s], [-s, c]].
void svd2x2(const double F[4], double U[4], double S[2], double V[4])
{
    Givens gu, gv;
    svd_givens(F, gu, S, gv);
    U[0] = gu.c; U[1] = gu.s; U[2] = -gu.s; U[3] = gu.c;
    V[0] = gv.c; V[1] = gv.s; V[2] = -gv.s; V[3] = gv.c;
}

// ComputeSVD (kernels.cuh:540-545)
void compute_svd(const icempm_simparams &P, PointState &p)
{
    svd2x2(p.Fe, p.U, p.vSigma, p.V);
    p.vSigmaSquared[0] = p.vSigma[0] * p.vSigma[0];
    p.vSigmaSquared[1] = p.vSigma[1] * p.vSigma[1];
    // dev_d (kernels.cuh:586-589)
    const double half = (p.vSigmaSquared[0] + p.vSigmaSquared[1]) / 2;
    const double f = P.mu / p.Je_tr;
    p.v_s_hat_tr[0] = f * (p.vSigmaSquared[0] - half);
    p.v_s_hat_tr[1] = f * (p.vSigmaSquared[1] - half);
}

// Wolper_Drucker_Prager (kernels.cuh:623-674)
void drucker_prager(const icempm_simparams &P, PointState &p)
{
    const double mu = P.mu, kappa = P.kappa;
    const double tan_phi = P.DP_tan_phi, DP_threshold_p = P.DP_threshold_p;
    const double pmax = P.IceCompressiveStrength, qmax = P.IceShearStrength;

    if (p.p_tr < -DP_threshold_p || p.Jp_inv < 1) {
        const double p_new = -DP_threshold_p;
        const double Je_new = std::sqrt(-2. * p_new / kappa + 1.);
        const double sqrt_Je_new = std::sqrt(Je_new);
        const double s[2] = {sqrt_Je_new, sqrt_Je_new};
        usvt(p.U, s, p.V, p.Fe);
        p.Jp_inv *= Je_new / p.Je_tr;
    } else {
        double q_n_1;
        if (p.p_tr > pmax) {
            q_n_1 = 0;
        } else {
            const double q_from_dp = (p.p_tr + DP_threshold_p) * tan_phi;
            const double pmin = -P.IceTensileStrength;
            const double q_from_failure_surface =
                2 * std::sqrt((pmax - p.p_tr) * (p.p_tr - pmin)) * qmax / (pmax - pmin);
            q_n_1 = std::fmin(q_from_failure_surface, q_from_dp);   // CUDA device min(double,double)
        }
        if (p.q_tr >= q_n_1) {
            const double s_hat_n_1_norm = q_n_1 * kCoeff1Inv;
            double nh[2] = {p.v_s_hat_tr[0], p.v_s_hat_tr[1]};
            const double z = nh[0] * nh[0] + nh[1] * nh[1];
            if (z > 0) { const double r = std::sqrt(z); nh[0] = nh[0] / r; nh[1] = nh[1] / r; }
            const double f = s_hat_n_1_norm * (p.Je_tr / mu);
            const double tr = (p.vSigmaSquared[0] + p.vSigmaSquared[1]) / 2;
            const double s[2] = {std::sqrt(f * nh[0] + tr), std::sqrt(f * nh[1] + tr)};
            usvt(p.U, s, p.V, p.Fe);
        }
    }
}

// kernels.cuh:99-103 / 348-352
inline void weights(double fx, double w[3])
{
    const double v0 = 1.5 - fx, v1 = fx - 1.0, v2 = fx - 0.5;
    w[0] = 0.5 * v0 * v0;
    w[1] = 0.75 - v1 * v1;
    w[2] = 0.5 * v2 * v2;
}

// KirchhoffStress_Wolper (kernels.cuh:24-34)
void kirchhoff(const icempm_simparams &P, const double F[4], double tau[4])
{
    const double Je = det2(F);
    double b[4], d[4];
    ffT(F, b);
    dev2(b, d);
    const double a = P.mu * (1 / Je);
    const double c = P.kappa * 0.5 * (Je * Je - 1.);
    tau[0] = a * d[0] + c;
    tau[1] = a * d[1];
    tau[2] = a * d[2];
    tau[3] = a * d[3] + c;
}

// Water (kernels.cuh:37-44)
void water(const icempm_simparams &P, double J, double tau[4])
{
    const double c = P.kappa * (1. - std::pow(J, -5.0));
    tau[0] = c; tau[1] = 0; tau[2] = 0; tau[3] = c;
}

// error checks of kernels.cuh:94-97 (single partition: gridX_offset = 0, gridX = GridXTotal)
inline uint8_t stencil_error(const icempm_simparams &P, int bx, int by)
{
    uint8_t e = 0;
    const int halo = P.GridHaloSize;
    if (bx < -halo) e = 70;
    if (by < 0) e = 71;
    if (bx > P.GridXTotal + halo - 3) e = 72;
    if (by > P.GridY - 3) e = 73;
    return e;
}

inline void base_and_fx(const icempm_simparams &P, const double pos[2], int base[2], double fx[2])
{
    for (int k = 0; k < 2; k++) {
        base[k] = (int)(pos[k] * P.cellsize_inv - 0.5);
        fx[k] = pos[k] * P.cellsize_inv - (double)base[k];
    }
}

}  // namespace

extern "C" {

void orc_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

int orc_get_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

// parameters_sim.cpp:128-133
void orc_params_compute_lame(icempm_simparams *P)
{
    P->lambda = P->YoungsModulus * P->PoissonsRatio / ((1 + P->PoissonsRatio) * (1 - 2 * P->PoissonsRatio));
    P->mu = P->YoungsModulus / (2 * (1 + P->PoissonsRatio));
    P->kappa = P->mu * 2. / 3. + P->lambda;
}

// parameters_sim.cpp:155-165
void orc_params_camclay(icempm_simparams *P)
{
    orc_params_compute_lame(P);
    P->NACC_beta = P->IceTensileStrength / P->IceCompressiveStrength;
    const double beta = P->NACC_beta, q = P->IceShearStrength, p0 = P->IceCompressiveStrength;
    P->NACC_M = (2 * q * std::sqrt(1 + 2 * beta)) / (p0 * (1 + beta));
    P->NACC_Msq = P->NACC_M * P->NACC_M;
}

// parameters_sim.cpp:135-153
void orc_params_helper(icempm_simparams *P)
{
    P->ParticleMass = P->ParticleVolume * P->Density;
    P->WaterParticleMass = P->ParticleVolume * P->WaterDensity;
    P->UpdateEveryNthStep = (int)(1.f / (200 * P->InitialTimeStep));
    P->cellsize = P->GridXDimension / P->GridXTotal;
    P->cellsize_inv = 1. / P->cellsize;
    P->Dp_inv = 4. / (P->cellsize * P->cellsize);
    P->IndRSq = P->IndDiameter * P->IndDiameter / 4.;
    P->dt_vol_Dpinv = P->InitialTimeStep * P->ParticleVolume * P->Dp_inv;
    P->dt_Gravity = P->InitialTimeStep * P->Gravity;
    P->vmax = 0.5 * P->cellsize / P->InitialTimeStep;
    P->vmax_squared = P->vmax * P->vmax;
    P->VectorCapacity_transfer =
        P->nPartitions > 0 ? (int)(P->nPtsTotal / P->nPartitions * P->PointsTransferBufferFraction) : 0;
    P->gbOffset = P->GridY * P->GridHaloSize;
}

// parameters_sim.cpp:5-60
void orc_params_reset(icempm_simparams *P)
{
    std::memset(P, 0, sizeof(*P));
    P->nPtsTotal = 0;
    P->n_indenter_subdivisions = 100;
    P->InitialTimeStep = 3.e-5;
    P->YoungsModulus = 5.e8;
    P->GridXTotal = 128;
    P->GridY = 55;
    P->ParticleViewSize = 2.5f;
    P->GridXDimension = 3.33;
    P->SimulationEndTime = 12;
    P->PoissonsRatio = 0.3;
    P->Gravity = 9.81;
    P->Density = 916;
    P->WaterDensity = 1000;
    P->IndDiameter = 0.324;
    P->IndVelocity = 0.2;
    P->IndDepth = 0.25;
    P->SimulationStep = 0;
    P->SimulationTime = 0;
    P->IceCompressiveStrength = 100e6;
    P->IceTensileStrength = 10e6;
    P->IceShearStrength = 6e6;
    P->IceTensileStrength2 = 5e6;
    P->DP_tan_phi = std::tan(30 * kPi / 180.);
    P->DP_threshold_p = 1e3;
    P->tpb_P2G = 256;
    P->tpb_Upd = 512;
    P->tpb_G2P = 128;
    P->indenter_x = P->indenter_x_initial = P->indenter_y = P->indenter_y_initial = 0;
    P->SetupType = 0;
    P->GrainVariability = 0.50;
    P->GridHaloSize = 5;
    P->ExtraSpaceForIncomingPoints = 0.2;
    P->PointsTransferBufferFraction = 0.05;
    P->nPartitions = 3;
    P->RebalanceThresholdFreeSpaceRemaining = 0.15;
    P->RebalanceThresholdDisabledPercentage = 0.01;
    P->PointTransferFrequency = 2;
    orc_params_compute_lame(P);
    orc_params_camclay(P);
    orc_params_helper(P);
}

void orc_weights(double fx, double w[3]) { weights(fx, w); }

void orc_kirchhoff_stress(const icempm_simparams *P, const double F[4], double tau[4]) { kirchhoff(*P, F, tau); }

void orc_water_stress(const icempm_simparams *P, double J, double tau[4]) { water(*P, J, tau); }

void orc_compute_pq(const icempm_simparams *P, const double F[4], double *Je, double *p, double *q)
{
    PointState s{};
    std::memcpy(s.Fe, F, sizeof s.Fe);
    compute_pq(*P, s);
    *Je = s.Je_tr; *p = s.p_tr; *q = s.q_tr;
}

void orc_grain_params(const icempm_simparams *P, int grain, double out[6])
{
    grain_params(*P, (short)grain, out[0], out[1], out[2], out[3], out[4], out[5]);
}

int orc_failure_check(const icempm_simparams *P, int grain, double p, double q)
{
    PointState s{};
    s.grain = (short)grain; s.p_tr = p; s.q_tr = q; s.utility_data = 0;
    failure_check(*P, s);
    return (s.utility_data & kCrushed) ? 1 : 0;
}

void orc_polar2x2(const double F[4], double R[4], double S[4])
{
    // input row-major, returned R row-major as the Eigen reading of fill<2>,
    // S row-major (symmetric).
    const double a[4] = {F[0], F[2], F[1], F[3]};
    Givens g;
    double s[4];
    polar(a, g, s);
    R[0] = g.c; R[1] = g.s; R[2] = -g.s; R[3] = g.c;
    S[0] = s[0]; S[1] = s[2]; S[2] = s[1]; S[3] = s[3];
}

void orc_svd2x2(const double F[4], double U[4], double S[2], double V[4]) { svd2x2(F, U, S, V); }

// The constitutive part of partition_kernel_g2p (kernels.cuh:383-390), solid points.
void orc_point_plasticity(const icempm_simparams *P, double F[4], double *Jp_inv, int64_t *utility,
                          double *p_out, double *q_out)
{
    PointState s{};
    std::memcpy(s.Fe, F, sizeof s.Fe);
    s.Jp_inv = *Jp_inv;
    s.utility_data = *utility;
    s.grain = (short)s.utility_data;
    compute_pq(*P, s);
    if (!(s.utility_data & kCrushed)) failure_check(*P, s);
    if (s.utility_data & kCrushed) {
        compute_svd(*P, s);
        drucker_prager(*P, s);
    }
    std::memcpy(F, s.Fe, sizeof s.Fe);
    *Jp_inv = s.Jp_inv;
    *utility = s.utility_data;
    *p_out = s.p_tr;
    *q_out = s.q_tr;
}

// partition_kernel_p2g (kernels.cuh:46-122), single partition.
int orc_p2g(const icempm_simparams *Pp, const double *soa, size_t pitch, int n, double *grid, size_t gpitch,
            uint8_t *err)
{
    const icempm_simparams &P = *Pp;
    const int halo = P.GridHaloSize, GY = P.GridY;
    const int ncols = P.GridXTotal + 2 * halo;
    const int W = 4;
    const int nstrips = (ncols + W - 1) / W;
    std::vector<int> strip_of(n, -1);
    std::vector<uint8_t> perr(n, 0);

#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) {
        const int64_t ud = as_i64(soa[i + pitch * ICEMPM_IDX_UTILITY_DATA]);
        if (ud & kDisabled) continue;
        const double pos[2] = {soa[i + pitch * ICEMPM_IDX_POSX], soa[i + pitch * (ICEMPM_IDX_POSX + 1)]};
        int base[2]; double fx[2];
        base_and_fx(P, pos, base, fx);
        const uint8_t e = stencil_error(P, base[0], base[1]);
        if (e) { perr[i] = e; continue; }
        strip_of[i] = (base[0] + halo) / W;
    }
    uint8_t e_last = 0;
    for (int i = 0; i < n; i++) if (perr[i]) e_last = perr[i];
    if (e_last && err) *err = e_last;

    // bucket points by strip, index order kept
    std::vector<int> cnt(nstrips + 1, 0);
    for (int i = 0; i < n; i++) if (strip_of[i] >= 0) cnt[strip_of[i] + 1]++;
    for (int s = 0; s < nstrips; s++) cnt[s + 1] += cnt[s];
    std::vector<int> order(cnt[nstrips]);
    {
        std::vector<int> fill(cnt.begin(), cnt.end() - 1);
        for (int i = 0; i < n; i++) if (strip_of[i] >= 0) order[fill[strip_of[i]]++] = i;
    }

    double *gm = grid, *gx = grid + gpitch, *gy = grid + 2 * gpitch;
    for (int colour = 0; colour < 2; colour++) {
#pragma omp parallel for schedule(dynamic, 1)
        for (int s = colour; s < nstrips; s += 2) {
            for (int k = cnt[s]; k < cnt[s + 1]; k++) {
                const int pt = order[k];
                const int64_t ud = as_i64(soa[pt + pitch * ICEMPM_IDX_UTILITY_DATA]);
                const int isLiquid = (int)((ud & kLiquid) >> 18);
                const double particle_mass = isLiquid ? P.WaterParticleMass : P.ParticleMass;
                double pos[2], vel[2], Fe[4], Bp[4];
                for (int a = 0; a < 2; a++) {
                    pos[a] = soa[pt + pitch * (ICEMPM_IDX_POSX + a)];
                    vel[a] = soa[pt + pitch * (ICEMPM_IDX_VELX + a)];
                }
                for (int a = 0; a < 4; a++) {
                    Fe[a] = soa[pt + pitch * (ICEMPM_IDX_FE00 + a)];
                    Bp[a] = soa[pt + pitch * (ICEMPM_IDX_BP00 + a)];
                }
                double PFt[4];
                if (isLiquid) water(P, soa[pt + pitch * ICEMPM_IDX_JP_INV], PFt);
                else kirchhoff(P, Fe, PFt);
                double S[4];
                for (int a = 0; a < 4; a++) S[a] = particle_mass * Bp[a] - P.dt_vol_Dpinv * PFt[a];
                int base[2]; double fx[2];
                base_and_fx(P, pos, base, fx);
                double wx[3], wy[3];
                weights(fx[0], wx);
                weights(fx[1], wy);
                const double h = P.cellsize;
                for (int i = 0; i < 3; i++)
                    for (int j = 0; j < 3; j++) {
                        const double Wip = wx[i] * wy[j];
                        const double d0 = (i - fx[0]) * h, d1 = (j - fx[1]) * h;
                        const double t0 = vel[0] * particle_mass + (S[0] * d0 + S[1] * d1);
                        const double t1 = vel[1] * particle_mass + (S[2] * d0 + S[3] * d1);
                        const double incM = Wip * particle_mass;
                        const size_t idx = (size_t)(base[0] + i + halo) * GY + (base[1] + j);
                        gm[idx] += incM;
                        gx[idx] += Wip * t0;
                        gy[idx] += Wip * t1;
                    }
            }
        }
    }
    return e_last;
}

// partition_kernel_update_nodes (kernels.cuh:164-301), live lines only.
void orc_update_nodes(const icempm_simparams *Pp, double *grid, size_t gpitch, double t, double ind_x,
                      double ind_y, double *ind_force)
{
    (void)t;
    const icempm_simparams &P = *Pp;
    const int halo = P.GridHaloSize, GY = P.GridY;
    const int nNodes = GY * (P.GridXTotal + 2 * halo);
    const double dt = P.InitialTimeStep;
    const double vco[2] = {P.IndVelocity, 0};
    const int nsub = P.n_indenter_subdivisions;
    std::vector<int> contact_bin(nNodes, -1);
    std::vector<double> contact_f(ind_force ? 2 * (size_t)nNodes : 0);

#pragma omp parallel for schedule(static)
    for (int idx = 0; idx < nNodes; idx++) {
        const double mass = grid[idx];
        double vx = grid[gpitch + idx], vy = grid[2 * gpitch + idx];
        if (mass == 0) continue;
        const int gix = idx / GY - halo, giy = idx % GY;
        const double gnpos[2] = {(double)gix * P.cellsize, (double)giy * P.cellsize};
        double v[2] = {vx / mass, vy / mass};
        v[1] -= P.dt_Gravity;
        const double sq = v[0] * v[0] + v[1] * v[1];
        if (sq > P.vmax_squared) {
            const double r = std::sqrt(sq);
            v[0] = v[0] / r * P.vmax;
            v[1] = v[1] / r * P.vmax;
        }
        double nrm[2] = {gnpos[0] - ind_x, gnpos[1] - ind_y};
        if (nrm[0] * nrm[0] + nrm[1] * nrm[1] < P.IndRSq) {
            const double vrel[2] = {v[0] - vco[0], v[1] - vco[1]};
            const double z = nrm[0] * nrm[0] + nrm[1] * nrm[1];
            if (z > 0) { const double r = std::sqrt(z); nrm[0] = nrm[0] / r; nrm[1] = nrm[1] / r; }
            const double vn = vrel[0] * nrm[0] + vrel[1] * nrm[1];
            if (vn < 0) {
                const double vt[2] = {vrel[0] - nrm[0] * vn, vrel[1] - nrm[1] * vn};
                const double prev[2] = {v[0], v[1]};
                v[0] = vco[0] + vt[0];
                v[1] = vco[1] + vt[1];
                const double f0 = (prev[0] - v[0]) * mass / dt;
                const double f1 = (prev[1] - v[1]) * mass / dt;
                float angle = std::atan2((float)nrm[0], (float)nrm[1]);
                angle += kPi;
                angle *= nsub / (2 * kPi);
                const int index = std::min(std::max((int)angle, 0), nsub - 1);
                if (ind_force) {
                    contact_bin[idx] = index;
                    contact_f[2 * (size_t)idx] = f0;
                    contact_f[2 * (size_t)idx + 1] = f1;
                }
            }
        }
        if (giy <= 2) { v[0] = 0; v[1] = 0; }
        if (giy >= GY - 3 && v[1] > 0) v[1] = 0;
        if (gix <= 2 && v[0] < 0) v[0] = 0;
        else if (gix >= P.GridXTotal - 3 && v[0] > 0) v[0] = 0;
        grid[gpitch + idx] = v[0];
        grid[2 * gpitch + idx] = v[1];
    }
    if (ind_force)
        for (int idx = 0; idx < nNodes; idx++)
            if (contact_bin[idx] >= 0) {
                ind_force[2 * contact_bin[idx]] += contact_f[2 * (size_t)idx];
                ind_force[2 * contact_bin[idx] + 1] += contact_f[2 * (size_t)idx + 1];
            }
}

// partition_kernel_g2p (kernels.cuh:305-414); migration (416-461) not restated.
int orc_g2p(const icempm_simparams *Pp, double *soa, size_t pitch, int n, const double *grid, size_t gpitch,
            int recordPQ, uint8_t *err)
{
    const icempm_simparams &P = *Pp;
    const int halo = P.GridHaloSize, GY = P.GridY;
    const double h_inv = P.cellsize_inv, dt = P.InitialTimeStep;
    std::vector<uint8_t> perr(n, 0);

#pragma omp parallel for schedule(static)
    for (int pt = 0; pt < n; pt++) {
        PointState p{};
        p.utility_data = as_i64(soa[pt + pitch * ICEMPM_IDX_UTILITY_DATA]);
        if (p.utility_data & kDisabled) continue;
        const bool isLiquid = p.utility_data & kLiquid;
        for (int a = 0; a < 2; a++) p.pos[a] = soa[pt + pitch * (ICEMPM_IDX_POSX + a)];
        for (int a = 0; a < 4; a++) p.Fe[a] = soa[pt + pitch * (ICEMPM_IDX_FE00 + a)];
        p.Jp_inv = soa[pt + pitch * ICEMPM_IDX_JP_INV];
        p.grain = (short)p.utility_data;

        int base[2]; double fx[2];
        base_and_fx(P, p.pos, base, fx);
        const uint8_t e = stencil_error(P, base[0], base[1]);
        if (e) { perr[pt] = e; continue; }
        double wx[3], wy[3];
        weights(fx[0], wx);
        weights(fx[1], wy);

        p.vel[0] = p.vel[1] = 0;
        for (int a = 0; a < 4; a++) p.Bp[a] = 0;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                const double dpos[2] = {i - fx[0], j - fx[1]};
                const double weight = wx[i] * wy[j];
                const size_t idx = (size_t)(base[0] + i + halo) * GY + (base[1] + j);
                const double nv[2] = {grid[gpitch + idx], grid[2 * gpitch + idx]};
                p.vel[0] += weight * nv[0];
                p.vel[1] += weight * nv[1];
                const double s = (4. * h_inv) * weight;
                p.Bp[0] += s * (nv[0] * dpos[0]);
                p.Bp[1] += s * (nv[0] * dpos[1]);
                p.Bp[2] += s * (nv[1] * dpos[0]);
                p.Bp[3] += s * (nv[1] * dpos[1]);
            }

        p.pos[0] += p.vel[0] * dt;
        p.pos[1] += p.vel[1] * dt;
        const double A[4] = {1.0 + dt * p.Bp[0], dt * p.Bp[1], dt * p.Bp[2], 1.0 + dt * p.Bp[3]};
        mul2(A, p.Fe, p.Fe);

        if (isLiquid) {
            p.Jp_inv *= det2(A);
        } else {
            compute_pq(P, p);
            if (!(p.utility_data & kCrushed)) failure_check(P, p);
            if (p.utility_data & kCrushed) {
                compute_svd(P, p);
                drucker_prager(P, p);
            }
        }

        for (int a = 0; a < 2; a++) {
            soa[pt + pitch * (ICEMPM_IDX_POSX + a)] = p.pos[a];
            soa[pt + pitch * (ICEMPM_IDX_VELX + a)] = p.vel[a];
        }
        for (int a = 0; a < 4; a++) {
            soa[pt + pitch * (ICEMPM_IDX_FE00 + a)] = p.Fe[a];
            soa[pt + pitch * (ICEMPM_IDX_BP00 + a)] = p.Bp[a];
        }
        soa[pt + pitch * ICEMPM_IDX_JP_INV] = p.Jp_inv;
        soa[pt + pitch * ICEMPM_IDX_UTILITY_DATA] = as_f64(p.utility_data);
        if (recordPQ) {
            soa[pt + pitch * ICEMPM_IDX_P] = p.p_tr;
            soa[pt + pitch * ICEMPM_IDX_Q] = p.q_tr;
        }
    }
    uint8_t e_last = 0;
    for (int i = 0; i < n; i++) if (perr[i]) e_last = perr[i];
    if (e_last && err) *err = e_last;
    return e_last;
}

// icy::Model::Step substep loop (model.cpp:38-54) with one partition:
// t += dt; indenter_x = x0 + t*V; reset_grid; p2g; update_nodes(t); g2p(recordPQ).
int orc_substeps(const icempm_simparams *P, double *soa, size_t pitch, int n, int nsub, double *t, int *step,
                 double *ind_force, uint8_t *err)
{
    const size_t nNodes = (size_t)P->GridY * (P->GridXTotal + 2 * P->GridHaloSize);
    std::vector<double> grid(3 * nNodes);
    int rc = 0;
    for (int k = 0; k < nsub; k++) {
        *t += P->InitialTimeStep;
        const double ind_x = P->indenter_x_initial + *t * P->IndVelocity;
        std::fill(grid.begin(), grid.end(), 0.0);
        int e = orc_p2g(P, soa, pitch, n, grid.data(), nNodes, err);
        if (e) rc = e;
        orc_update_nodes(P, grid.data(), nNodes, *t, ind_x, P->indenter_y, ind_force);
        const int recordPQ = (*step % P->UpdateEveryNthStep) == 0;
        e = orc_g2p(P, soa, pitch, n, grid.data(), nNodes, recordPQ, err);
        if (e) rc = e;
        (*step)++;
    }
    return rc;
}

// ProxyPoint::getCellIndex (proxypoint2d.cpp:164-169)
int orc_cell_key(const icempm_simparams *P, double x, double y)
{
    const int ix = (int)(x * P->cellsize_inv + 0.5);
    const int iy = (int)(y * P->cellsize_inv + 0.5);
    return ix * P->GridY + iy;
}

// HostSideSOA::RemoveDisabledAndSort (host_side_soa.cpp:50-61) with the order
// made total: std::remove_if keeps the relative order of the kept points, and
// ties of the (unstable) std::sort are broken by that order, i.e. a stable sort.
// Returns the new size; keys_out/perm_out (may be null) get the sorted keys and
// the previous index of each sorted point.
int orc_sort(const icempm_simparams *P, double *soa, size_t pitch, int n, int32_t *keys_out, int32_t *perm_out)
{
    std::vector<int> kept;
    kept.reserve(n);
    for (int i = 0; i < n; i++)
        if (!(as_i64(soa[i + pitch * ICEMPM_IDX_UTILITY_DATA]) & kDisabled)) kept.push_back(i);
    std::vector<int> key(n);
    for (int i : kept)
        key[i] = orc_cell_key(P, soa[i + pitch * ICEMPM_IDX_POSX], soa[i + pitch * (ICEMPM_IDX_POSX + 1)]);
    std::stable_sort(kept.begin(), kept.end(), [&](int a, int b) { return key[a] < key[b]; });
    const int m = (int)kept.size();
    std::vector<double> tmp(m);
    for (int line = 0; line < ICEMPM_N_PTS_ARRAYS; line++) {
        for (int k = 0; k < m; k++) tmp[k] = soa[kept[k] + pitch * line];
        for (int k = 0; k < m; k++) soa[k + pitch * line] = tmp[k];
    }
    for (int k = 0; k < m; k++) {
        if (keys_out) keys_out[k] = key[kept[k]];
        if (perm_out) perm_out[k] = kept[k];
    }
    return m;
}

// Cell key of a point whose coordinates are stored in fixed point,
// x/h = xq * 2^-frac_bits - origin_x (y likewise): (int)(x/h + 0.5) with C
// truncation, evaluated exactly in integers.
int orc_cell_key_fixed(uint32_t xq, uint32_t yq, int frac_bits, int origin_x, int origin_y, int GridY)
{
    const int64_t half = (int64_t)1 << (frac_bits - 1);
    int c[2];
    const uint32_t q[2] = {xq, yq};
    const int org_cells[2] = {origin_x, origin_y};
    for (int k = 0; k < 2; k++) {
        const int64_t org = (int64_t)org_cells[k] << frac_bits;
        const int64_t v = (int64_t)q[k] + half - org;
        c[k] = v >= 0 ? (int)(v >> frac_bits) : -(int)((-v) >> frac_bits);
    }
    return c[0] * GridY + c[1];
}

}  // extern "C"

extern "C" {
// Byte offset of every icempm_simparams member, in declaration order, so tests
// can check a foreign mirror of the struct (ctypes) against the C layout.
int orc_simparams_offsets(size_t *out, int n)
{
#define OFF(f) offsetof(icempm_simparams, f)
    const size_t offs[] = {
        OFF(n_indenter_subdivisions), OFF(tpb_P2G), OFF(tpb_Upd), OFF(tpb_G2P), OFF(nPtsTotal),
        OFF(GridXTotal), OFF(GridY), OFF(GridTotal), OFF(GridXDimension), OFF(InitialTimeStep),
        OFF(SimulationEndTime), OFF(UpdateEveryNthStep), OFF(SimulationStep), OFF(SimulationTime),
        OFF(Gravity), OFF(Density), OFF(WaterDensity), OFF(PoissonsRatio), OFF(YoungsModulus),
        OFF(lambda), OFF(mu), OFF(kappa), OFF(IceCompressiveStrength), OFF(IceTensileStrength),
        OFF(IceShearStrength), OFF(IceTensileStrength2), OFF(NACC_beta), OFF(NACC_M), OFF(NACC_Msq),
        OFF(DP_tan_phi), OFF(DP_threshold_p), OFF(IndDiameter), OFF(IndRSq), OFF(IndVelocity),
        OFF(IndDepth), OFF(cellsize), OFF(cellsize_inv), OFF(Dp_inv), OFF(xmin), OFF(xmax), OFF(ymin),
        OFF(ymax), OFF(nxmin), OFF(nxmax), OFF(nymin), OFF(nymax), OFF(ParticleVolume),
        OFF(ParticleMass), OFF(WaterParticleMass), OFF(ParticleViewSize), OFF(indenter_x),
        OFF(indenter_x_initial), OFF(indenter_y), OFF(indenter_y_initial), OFF(Volume), OFF(SetupType),
        OFF(GrainVariability), OFF(GridHaloSize), OFF(ExtraSpaceForIncomingPoints),
        OFF(PointsTransferBufferFraction), OFF(RebalanceThresholdFreeSpaceRemaining),
        OFF(RebalanceThresholdDisabledPercentage), OFF(PointTransferFrequency), OFF(nPartitions),
        OFF(VectorCapacity_transfer), OFF(dt_vol_Dpinv), OFF(dt_Gravity), OFF(vmax), OFF(vmax_squared),
        OFF(gbOffset), sizeof(icempm_simparams)};
#undef OFF
    const int m = (int)(sizeof(offs) / sizeof(offs[0]));
    for (int i = 0; i < m && i < n; i++) out[i] = offs[i];
    return m;
}
}
