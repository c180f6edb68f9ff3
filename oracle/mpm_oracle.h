/*
 * mpm_oracle.h -- TEST INFRASTRUCTURE ONLY (the checker, never the product).
 *
 * FP64 CPU restatement of the reference's MPM substep, written from the
 * reference's CUDA sources read as text (igorg520b/iceMPM_multi):
 *   simulation/kernels.cuh, simulation/helper_math.cuh, simulation/givens.cuh,
 *   simulation/parameters_sim.cpp, simulation/hssoa/host_side_soa.cpp, proxypoint2d.cpp, simulation/model.cpp.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * Layout contracts shared with the product:
 *   points: HostSideSOA, ICEMPM_N_PTS_ARRAYS lines of `pitch` doubles;
 *   grid:   3 lines (mass, momentum/velocity x, y) of `gpitch` doubles; node
 *           (gx, gy), gx in [-halo, GridX+halo), stored at (gx+halo)*GridY + gy,
 *           i.e. the reference's single-partition layout (kernels.cuh:114-118).
 *
 * Every function restates one reference function; its citation is at the
 * definition.  What pins each one is listed in the header of mpm_oracle.cpp.
 */
#ifndef MPM_ORACLE_H
#define MPM_ORACLE_H

#include <stddef.h>
#include <stdint.h>
#include "../include/icempm.h"

#ifdef __cplusplus
extern "C" {
#endif

void orc_set_threads(int n);
int  orc_get_threads(void);

/* parameters_sim.cpp */
void orc_params_reset(icempm_simparams *P);
void orc_params_compute_lame(icempm_simparams *P);
void orc_params_camclay(icempm_simparams *P);
void orc_params_helper(icempm_simparams *P);

/* device math (kernels.cuh, helper_math.cuh, givens.cuh) */
void orc_weights(double fx, double w[3]);
void orc_kirchhoff_stress(const icempm_simparams *P, const double F[4], double tau[4]);
void orc_water_stress(const icempm_simparams *P, double J, double tau[4]);
void orc_compute_pq(const icempm_simparams *P, const double F[4], double *Je, double *p, double *q);
void orc_grain_params(const icempm_simparams *P, int grain, double out[6]);
int  orc_failure_check(const icempm_simparams *P, int grain, double p, double q);
void orc_polar2x2(const double F[4], double R[4], double S[4]);
void orc_svd2x2(const double F[4], double U[4], double S[2], double V[4]);
void orc_point_plasticity(const icempm_simparams *P, double F[4], double *Jp_inv, int64_t *utility,
                          double *p_out, double *q_out);

/* substep phases */
int  orc_p2g(const icempm_simparams *P, const double *soa, size_t pitch, int n, double *grid, size_t gpitch,
             uint8_t *err);
void orc_update_nodes(const icempm_simparams *P, double *grid, size_t gpitch, double t, double ind_x,
                      double ind_y, double *ind_force);
int  orc_g2p(const icempm_simparams *P, double *soa, size_t pitch, int n, const double *grid, size_t gpitch,
             int recordPQ, uint8_t *err);
/* Model::Step's substep loop; *t and *step advance; ind_force accumulates. */
int  orc_substeps(const icempm_simparams *P, double *soa, size_t pitch, int n, int nsub, double *t,
                  int *step, double *ind_force, uint8_t *err);

/* binning */
int  orc_cell_key(const icempm_simparams *P, double x, double y);
int  orc_sort(const icempm_simparams *P, double *soa, size_t pitch, int n, int32_t *keys_out,
              int32_t *perm_out);
int  orc_cell_key_fixed(uint32_t xq, uint32_t yq, int frac_bits, int origin_x, int origin_y, int GridY);
int  orc_simparams_offsets(size_t *out, int n);

#ifdef __cplusplus
}
#endif
#endif
