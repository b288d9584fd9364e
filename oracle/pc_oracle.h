/*
 * TEST INFRASTRUCTURE ONLY.  The CPU oracle: a restatement of the reference
 * Godunov path, used by tests/, the smoke check and the bench's cpu_baseline
 * leg as the checker.  It is never linked into the product library.
 *
 * Same entry points as include/pc_hydro.h with a pc_ref_ prefix, operating
 * on host memory, with the EOS chosen per call (1 = GammaLaw, 9 = LiDryer).
 */
#ifndef PC_ORACLE_H
#define PC_ORACLE_H
#include "../include/pc_hydro.h"

#ifdef __cplusplus
extern "C" {
#endif

void pc_ref_params_default(pc_params_t* p);
int pc_ref_ctoprim(int eos, pc_box_t qbx, pc_array4_t u, pc_array4_t q,
                   pc_array4_t qaux);
int pc_ref_srctoprim(int eos, pc_box_t qbx, pc_array4_t q, pc_array4_t qaux,
                     pc_array4_t src, pc_array4_t srcq);
int pc_ref_umdrv(int eos, double time, pc_box_t bx, const int* domlo,
                 const int* domhi, const int* bclo, const int* bchi,
                 pc_array4_t uin, pc_array4_t uout, pc_array4_t q,
                 pc_array4_t qaux, pc_array4_t src_q, const double* dx,
                 double dt, int ppm_type, int plm_iorder, int use_flattening,
                 int use_hybrid_weno, int weno_scheme, double difmag,
                 const pc_array4_t* flx, const pc_array4_t* a, pc_array4_t vol,
                 double* cflLoc);
int pc_ref_hydro_source_box(int eos, pc_box_t bx, const int* domlo,
                            const int* domhi, const int* bclo, const int* bchi,
                            pc_array4_t sarr, pc_array4_t src,
                            pc_array4_t hyd_src, const double* dx, double dt,
                            const pc_params_t* prm, const pc_array4_t* flx,
                            const pc_array4_t* a, pc_array4_t vol);

/* Level on the host: same arena layout as the device level (box b's FAB over
 * grow(box,4), NVAR comps, at arena + offset[b]). */
long long pc_ref_level_arena_size(int eos, const pc_level_desc_t* desc,
                                  long long* offsets /* nboxes or NULL */);
int pc_ref_level_fill_ghosts(int eos, const pc_level_desc_t* desc,
                             double* arena);
int pc_ref_level_estdt(int eos, const pc_level_desc_t* desc,
                       const pc_params_t* prm, const double* arena,
                       double* dt_out);
/* nsteps Godunov steps; *dt is the dt of the first step on entry and the dt
 * of the next step on return (computeNewDt, with the stop_time limit at the
 * time each step ended).  *time (may be NULL: 0) is the level time, advanced
 * by each step.  nthreads > 0 sets the OpenMP thread count. */
int pc_ref_level_step(int eos, const pc_level_desc_t* desc,
                      const pc_params_t* prm, double* arena, int nsteps,
                      double* dt, double* time, int nthreads);
/* computeInitialDt for one level at `time`: init_shrink * estTimeStep (or
 * fixed_dt), limited by stop_time. */
int pc_ref_level_init_dt(int eos, const pc_level_desc_t* desc,
                         const pc_params_t* prm, const double* arena,
                         double time, double* dt_out);

/* Conservative state from primitive fields (pc_initdata pattern).
 * mode 0: prim = (rho, p, u, v, w, Y[NS]); mode 1: prim = (p, T, u, v, w, Y). */
int pc_ref_prim_to_cons(int eos, pc_box_t bx, pc_array4_t prim, int mode,
                        pc_array4_t u);

/* Point functions for known-answer tests. */
/* which: 0 RYP2E(r,Y,p)->e, 1 RPY2Cs(r,p,Y)->c, 2 REY2T(r,e,Y;T0)->T,
 *        3 RTY2P(r,T,Y)->p, 4 RTY2Cs(r,T,Y)->c, 5 RTY2G(r,T,Y)->G,
 *        6 Y2WBAR(Y)->W, 7 T2Ei(T)->e_k (NS outputs), 8 CKCVBS(T,Y)
 *        (species sum), 9 h_k/(R T) (NS outputs), 10 RTY2dpde_dpdre ->
 *        (dpde, dpdr_e), 11 cv(T,Y) from mixture coefficients */
int pc_ref_eos(int eos, int which, double a, double b, double c,
               const double* Y, double* out);
/* riemann() of Source/Riemann.H:9-200: out[0..] = ustar, uflx_rho,
 * uflx_u, uflx_v, uflx_w, uflx_eden, uflx_eint, qint_iu, qint_iv1,
 * qint_iv2, qint_gdpres, qint_gdgame */
int pc_ref_riemann(int eos, const double* left5, const double* spl,
                   const double* right5, const double* spr, double cav,
                   double* out12);
/* pc_transdo / pc_transdd / pc_pdivu on one cell, for known-answer tests
 * (argument meaning in pc_oracle.cpp: transdo_point, transdd_point). */
int pc_ref_transdo_point(int eos, int dir, int t, const double* qplus,
                         const double* qminus, const double* fl_lo,
                         const double* fl_hi, const double* qi_lo,
                         const double* qi_hi, double gamc, double cdtdx,
                         double* out_p, double* out_m);
int pc_ref_transdd_point(int eos, int dir, const double* qplus,
                         const double* qminus, const double* fl0_lo,
                         const double* fl0_hi, const double* fl1_lo,
                         const double* fl1_hi, const double* qi0_lo,
                         const double* qi0_hi, const double* qi1_lo,
                         const double* qi1_hi, double gamc, const double* srcq,
                         double hdt, double c0, double c1, double* out_p,
                         double* out_m);
double pc_ref_pdivu_point(const double* pg, const double* un,
                          const double* area, double vol);
int pc_ref_ppm_reconstruct(const double* s5, double flatn, double* smsp);
int pc_ref_ppm_int_profile(double sm, double sp, double sc, double u, double c,
                           double dtdx, double* Ip3, double* Im3);
double pc_ref_plm_slope_line(const double* q5, double flat, int order);

/* getMOLSrcTerm for one box, hydro only: molsrc (NVAR comps on bx) =
 * -div(F)/vol of the MOL fluxes of sarr (>= 4 ghosts); flx[d] (NULL, or on
 * surroundingNodes(bx,d)) receives the extensive fluxes.  Slopes iff
 * prm->plm_iorder != 1; prm->use_laxf_flux selects Lax-Friedrichs. */
int pc_ref_mol_src_box(int eos, pc_box_t bx, pc_array4_t sarr,
                       pc_array4_t molsrc, const pc_params_t* prm,
                       const pc_array4_t* flx, const pc_array4_t* a,
                       pc_array4_t vol);
/* Plot variables of Source/Derive.cpp on the cells of bx (u covering
 * grow(bx,1)): codes 0 state comp `comps[m]`, 1 comps[m] / rho, 2 magvel,
 * 3 kineng, 4 pressure, 5 soundspeed, 6 MachNumber, 7 magvort, 8 logden. */
int pc_ref_derive(int eos, pc_box_t bx, pc_array4_t u, const int* codes,
                  const int* comps, int nplot, pc_array4_t out,
                  const double* dx);
/* Coarse/fine flux register and average-down (the pc_fluxreg_* and
 * pc_average_down_v1 semantics), dense restatement; boxes identify
 * themselves by their index ranges. */
void* pc_ref_fluxreg_create(int nc, const pc_box_t* crse_boxes, int ncrse,
                            const pc_box_t* fine_boxes, int nfine,
                            const int* domlo, const int* domhi,
                            const int* periodic, int ratio);
void pc_ref_fluxreg_destroy(void* reg);
int pc_ref_fluxreg_nfaces(void* reg);
int pc_ref_fluxreg_crse_add(void* reg, pc_box_t bx, const pc_array4_t* flx,
                            double scale);
int pc_ref_fluxreg_fine_add(void* reg, pc_box_t bx, const pc_array4_t* flx,
                            double scale);
int pc_ref_fluxreg_reflux(void* reg, pc_box_t bx, pc_array4_t state, double vol);
int pc_ref_average_down(int nc, pc_box_t crse_region, pc_array4_t crse,
                        pc_array4_t fine, int ratio);
/* laxfriedrich_flux of Source/Riemann.H:205-307: out[0..] = ustar,
 * maxeigval, uflx_rho, uflx_u, uflx_v, uflx_w, uflx_eden, uflx_eint,
 * uflx_rhoY[NS] */
int pc_ref_laxf(int eos, const double* left5, const double* spl,
                const double* right5, const double* spr, double* out);

#ifdef __cplusplus
}
#endif
#endif
