/* CPU oracle of the RK3 split-explicit dynamics step: TEST INFRASTRUCTURE ONLY.
 *
 * A plain C++/OpenMP restatement, kernel by kernel, of the reference CPU FP64 build's dry dynamics
 * (reference src/dyn/{rk3,large_step,small_step}.cpp, src/core/halo.cpp), used by the tests to check
 * the HIP path bit for bit and by bench.py as the timed CPU baseline.  Nothing in the library
 * (cppwrf_amd/) may call it.  It shares only the configuration structs of include/cppwrf_amd.h.
 *
 * Transcendental functions are the host glibc's (libm pow and sin, as the reference CPU build
 * compiled with CPPWRF_NO_LIBMVEC, src/core/backend.hpp:116); orc_set_libmvec(1) switches
 * calc_p_rho_phi's pow to glibc libmvec's 2-lane _ZGVbN2vv_pow lanes exactly as the reference's
 * default CPU build does (backend.hpp:144,169-175), for comparisons with its golden files.
 */
#ifndef CPPWRF_AMD_ORACLE_H
#define CPPWRF_AMD_ORACLE_H

#include "../include/cppwrf_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_domain orc_domain;

const char* orc_last_error(void);
int orc_domain_create(const mi_grid* g, const mi_params* p, const mi_levels* lv, orc_domain** out);
int orc_domain_destroy(orc_domain* d);
int orc_domain_n_sound(const orc_domain* d);
int orc_set_field(orc_domain* d, int field, const double* host, int64_t n);
int orc_get_field(orc_domain* d, int field, double* host, int64_t n);
/* 1: calc_p_rho_phi uses libmvec's 2-lane pow lanes (reference default build); 0: scalar pow */
int orc_set_libmvec(orc_domain* d, int on);
/* OpenMP threads for the oracle's loops (0: the OpenMP default) */
int orc_set_threads(int n);
int orc_step(orc_domain* d, int nsteps);
int orc_step_begin(orc_domain* d);
int orc_rk_prep(orc_domain* d, int rk);
int orc_rk_tendency(orc_domain* d, int rk);
int orc_small_steps(orc_domain* d, int rk);
int orc_scalar_transport(orc_domain* d, int rk); /* ScalarTransport::run_stage */
int orc_cfl_max(orc_domain* d, double* vert, double* horiz); /* LargeStep::cfl_max */
int orc_set_dt(orc_domain* d, double dt, int n_sound);        /* RK3::set_dt + LateralBC::set_dt (rk3.cpp:79-84) */
int orc_calc_p_rho_phi(orc_domain* d, int rk);
int orc_calc_p_rho_phi_interior(orc_domain* d, int spec_zone);
/* the advection flux operators alone (kernels_common.hpp:13-33, 83-88), for closed-form checks */
int orc_flux_eval(int which, const double* q, double ua, double up, double* out);
int orc_step_end(orc_domain* d); /* after_rk, spec_bdy_final, set_w_surface, after_all_rk_steps */
/* a column physics plugin (reference ColumnPhysics, src/dyn/column_physics.hpp:21-34): the hooks of
 * mi_domain_set_column_physics with the oracle's own user pointer; NULL detaches */
typedef void (*orc_phys_fn)(void* user);
typedef struct orc_column_physics {
  void* user;
  orc_phys_fn part1, part2, after_rk, after_all_rk_steps;
  int32_t th_forcing, qv_forcing;
} orc_column_physics;
int orc_set_column_physics(orc_domain* d, const orc_column_physics* cp);
int orc_set_diffusion(orc_domain* d, int diff_6th_opt, double diff_6th_factor);
/* specified boundaries: as mi_bdy_set / mi_bdy_set_time / mi_bdy_get_time */
int orc_bdy_set(orc_domain* d, int var, int side, const double* values, const double* tends, int64_t n);
int orc_bdy_set_time(orc_domain* d, double dtbc);
int orc_bdy_get_time(orc_domain* d, double* dtbc);

#ifdef __cplusplus
}
#endif
#endif
