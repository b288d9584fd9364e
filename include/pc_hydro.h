/*
 * pc_hydro.h -- C ABI of the MI355X (gfx950) Godunov hydro engine.
 *
 * The engine stands behind the reference's hydro path
 *   PeleC::construct_hydro_source   (Source/Hydro.cpp:5-301)
 *     -> pc_ctoprim / pc_srctoprim  (Source/Hydro.cpp:106-173)
 *     -> pc_umdrv                   (Source/Hydro.cpp:304-377, decl Source/Hydro.H:910-935)
 *        -> pc_umeth_3D             (Source/Godunov.cpp:60-476)
 *        -> pc_adjust_fluxes        (Source/Hydro.cpp:380-413)
 *        -> pc_consup               (Source/Hydro.cpp:545-557)
 * and the step glue around it (construct_Snew, computeTemp, estTimeStep:
 * Source/Advance.cpp:397-416, Source/PeleC.cpp:829-1034,2222-2307).
 *
 * Plain C: pointers, sizes and POD structs only.  Every entry point returns
 * an int status (PC_OK = 0, negative on error) and never aborts; the message
 * of the last failure is available from pc_last_error().  Device pointers are
 * HIP device (or managed) memory on the current device; `stream` is a
 * hipStream_t passed as void* (NULL = the null stream).  Work is enqueued
 * asynchronously on that stream unless the function says otherwise.
 */
#ifndef PC_HYDRO_H
#define PC_HYDRO_H

#ifdef __cplusplus
extern "C" {
#endif

/* Non-owning view of one FAB: the fields of amrex::Array4<double>
 * (AMReX_Array4.H: p, jstride, kstride, nstride, begin, end, ncomp).
 * Element (i,j,k,n) is p[(i-begin[0]) + (j-begin[1])*jstride
 *                       + (k-begin[2])*kstride + n*nstride]. */
typedef struct pc_array4 {
  double* p;
  long long jstride;
  long long kstride;
  long long nstride;
  int begin[3]; /* inclusive lower corner of the index box */
  int end[3];   /* exclusive upper corner */
  int ncomp;
} pc_array4_t;

/* Cell-centred index box, amrex::Box smallEnd/bigEnd (both inclusive). */
typedef struct pc_box {
  int lo[3];
  int hi[3];
} pc_box_t;

/* Status codes */
enum {
  PC_OK = 0,
  PC_ERR_ARG = -1,         /* null pointer, bad box, wrong ncomp, ...       */
  PC_ERR_GROW = -2,        /* input FAB has fewer than 4 ghost cells        */
  PC_ERR_PPM_TYPE = -3,    /* ppm_type not 0 or 1 (Source/Godunov.cpp:203)  */
  PC_ERR_PLM_IORDER = -4,  /* plm_iorder not 1, 2 or 4                      */
  PC_ERR_HIP = -5,         /* a HIP runtime call failed                     */
  PC_ERR_NAN = -6,         /* NaN/Inf found in the state (okToContinue)     */
  PC_ERR_UNSUPPORTED = -7, /* option outside what this build implements     */
  PC_ERR_NOMEM = -8,       /* device allocation failed                      */
  PC_ERR_DT = -9,          /* dt fell below dt_cutoff (okToContinue)        */
  PC_ERR_MOL_IORDER = -10, /* do_mol with mol_iorder not 1 or 2             */
  PC_ERR_IO = -11          /* a file could not be written                   */
};

/* Equation of state / mechanism (compile-time NUM_SPECIES in the reference:
 * Exec/RegTests/Sod/GNUmakefile selects GammaLaw + Null, Exec/RegTests/PMF
 * selects Fuego + LiDryer).  The value is NUM_SPECIES. */
enum { PC_EOS_GAMMALAW = 1, PC_EOS_LIDRYER = 9 };

/* Physical boundary codes, PCPhysBCType (Source/PeleC.H:48-58). */
enum {
  PC_BC_INTERIOR = 0,
  PC_BC_INFLOW = 1,
  PC_BC_OUTFLOW = 2,
  PC_BC_SYMMETRY = 3,
  PC_BC_SLIPWALL = 4,
  PC_BC_NOSLIPWALL = 5,
  PC_BC_USERBC = 6
};

/* Hot-path runtime parameters, named and defaulted as the reference's
 * pelec.* keys (Source/Params/_cpp_parameters:53,96,103,113,118-130,
 * 267-289).  pc_params_default() fills the reference defaults. */
typedef struct pc_params {
  int ppm_type;        /* 0 = PLM, 1 = PPM                       (default 0) */
  int plm_iorder;      /* 1, 2 or 4                              (4)         */
  int use_flattening;  /*                                        (1)         */
  int use_hybrid_weno; /* must be 0 in this build                (0)         */
  int weno_scheme;     /*                                        (1)         */
  double difmag;       /* artificial viscosity coefficient       (0.1)       */
  double cfl;          /*                                        (0.8)       */
  double init_shrink;  /*                                        (1.0)       */
  double change_max;   /*                                        (1.1)       */
  double max_dt;       /*                                        (1e200)     */
  double fixed_dt;     /* > 0 overrides the CFL estimate         (-1)        */
  double dt_cutoff;    /*                                        (0)         */
  int allow_small_energy;          /* must be 1 in this build    (1)         */
  int allow_negative_energy;       /* must be 1 in this build    (1)         */
  int dual_energy_update_E_from_e; /*                            (1)         */
  double dual_energy_eta2;         /*                            (1e-4)      */
  /* Method-of-lines hydro (Source/Params/_cpp_parameters:62,106,109): a
   * level with do_mol = 1 advances by do_mol_advance's two-stage Heun step
   * on PLM-MOL fluxes instead of the CTU step.  As in the reference,
   * mol_iorder is only validated (1 or 2); the MOL slopes are taken when
   * plm_iorder != 1 (Source/Diffusion.cpp:371-373 passes plm_iorder). */
  int do_mol;          /*                                        (0)         */
  int mol_iorder;      /* 1 or 2                                 (2)         */
  int use_laxf_flux;   /* 1: Lax-Friedrichs flux in MOL          (0)         */
  /* computeNewDt / computeInitialDt's stop_time limit (Source/PeleC.cpp:
   * 1018-1027, 1061-1068; amr.stop_time): a step that would end past
   * stop_time - 0.001 dt is shortened to end on stop_time.  < 0: no limit.
   * Steps issued once the level time has reached stop_time take dt = 0. */
  double stop_time;    /*                                        (-1)        */
} pc_params_t;

void pc_params_default(pc_params_t* p);

/* ------------------------------------------------------------ context --- */
typedef struct pc_ctx pc_ctx_t;

/* A context binds an EOS (PC_EOS_*) to the current HIP device and owns the
 * scratch workspace of the box drivers (stream-ordered reuse, grown on
 * demand).  Replaces the reference's The_Async_Arena temporaries
 * (Source/Hydro.cpp:84-97,333-345; Source/Godunov.cpp:113-388). */
int pc_ctx_create(int eos_kind, pc_ctx_t** out);
int pc_ctx_destroy(pc_ctx_t* ctx);
int pc_ctx_num_species(const pc_ctx_t* ctx); /* NUM_SPECIES */
int pc_ctx_nvar(const pc_ctx_t* ctx);        /* NVAR = 7 + NUM_SPECIES */
int pc_ctx_qvar(const pc_ctx_t* ctx);        /* QVAR = 8 + NUM_SPECIES */
const char* pc_last_error(const pc_ctx_t* ctx);

/* ------------------------------------------------------- box drivers --- */

/* K1: pc_ctoprim on every cell of qbx (Source/Hydro.cpp:106-115 ->
 * Source/PeleCUtilities.H:242-320).  u: NVAR comps, q: QVAR, qaux: NQAUX=6. */
int pc_ctoprim_v1(pc_ctx_t* ctx, pc_box_t qbx, pc_array4_t u, pc_array4_t q,
                  pc_array4_t qaux, void* stream);

/* K2: pc_srctoprim on every cell of qbx (Source/Hydro.cpp:164-173 ->
 * Source/Hydro.H:14-68). */
int pc_srctoprim_v1(pc_ctx_t* ctx, pc_box_t qbx, pc_array4_t q,
                    pc_array4_t qaux, pc_array4_t src, pc_array4_t srcq,
                    void* stream);

/* pc_umdrv (Source/Hydro.H:910-935, Source/Hydro.cpp:304-377): the CTU
 * unsplit fluxes (PPM or PLM trace, transverse corrections, Riemann),
 * artificial viscosity, species-flux normalisation, area scaling and the
 * flux divergence.  Same arguments in the same order; GpuArrays become
 * pointers to 3 elements.  Box conventions as the reference: uin, q, qaux,
 * src_q cover grow(bx,4); flx[d] and a[d] cover surroundingNodes(bx,d);
 * vol covers bx.  On return (stream-ordered): flx[d] holds the extensive
 * fluxes, uout += dU/dt on bx.  src_q.p may be NULL for identically zero
 * sources.  cflLoc is unused, as in the reference (may be NULL).
 * use_hybrid_weno must be 0. */
int pc_umdrv_v1(pc_ctx_t* ctx, double time, pc_box_t bx, const int* domlo,
                const int* domhi, const int* bclo, const int* bchi,
                pc_array4_t uin, pc_array4_t uout, pc_array4_t q,
                pc_array4_t qaux, pc_array4_t src_q, const double* dx,
                double dt, int ppm_type, int plm_iorder, int use_flattening,
                int use_hybrid_weno, int weno_scheme, double difmag,
                const pc_array4_t* flx, const pc_array4_t* a, pc_array4_t vol,
                double* cflLoc, void* stream);

/* getMOLSrcTerm for one regular box with diffusion off (Source/Diffusion.cpp:
 * 90-98,136-148,226-240,349-374,509-519,654): primitives of sarr (>= 4 ghost
 * cells), the MOL hydro fluxes of pc_compute_hyp_mol_flux (Source/MOL.cpp:
 * 6-190; slopes when prm->plm_iorder != 1, Lax-Friedrichs when
 * prm->use_laxf_flux) and molsrc = -div(F)/vol on bx (overwritten).  flx[d]
 * (surroundingNodes(bx,d), NVAR comps) receives the extensive fluxes, or
 * flx may be NULL; a[d] and vol as pc_umdrv_v1.  prm->mol_iorder must be 1
 * or 2 (PC_ERR_MOL_IORDER); do_mol itself is not read. */
int pc_mol_src_box_v1(pc_ctx_t* ctx, pc_box_t bx, pc_array4_t sarr,
                      pc_array4_t molsrc, const pc_params_t* prm,
                      const pc_array4_t* flx, const pc_array4_t* a,
                      pc_array4_t vol, void* stream);

/* One box of construct_hydro_source (Source/Hydro.cpp:73-252): K1 ctoprim
 * and K2 srctoprim on grow(bx,4) into context scratch, then pc_umdrv.
 * sarr = state with >= 4 ghost cells; src = sources_for_hydro (NVAR comps,
 * p may be NULL for zero sources); hyd_src: hydro_source (+= on bx);
 * flx[d] on surroundingNodes(bx,d) are overwritten; a[d], vol as pc_umdrv. */
int pc_hydro_source_box_v1(pc_ctx_t* ctx, pc_box_t bx, const int* domlo,
                           const int* domhi, const int* bclo, const int* bchi,
                           pc_array4_t sarr, pc_array4_t src,
                           pc_array4_t hyd_src, const double* dx, double dt,
                           const pc_params_t* prm, const pc_array4_t* flx,
                           const pc_array4_t* a, pc_array4_t vol, void* stream);

/* pc_hydro_source_box_v1 without any device allocation, for callers that
 * cannot allocate (code under hipGraph capture, tight serving loops, hosts
 * that forbid hipMalloc).  No counterpart in the reference.  Same arguments
 * and results; the workspace (q, qaux, srcq, the CTU intermediates and the
 * job descriptor) lives in one static device array of the library, 112 MiB
 * per GPU, held until the process ends: any box of up to 32^3 cells fits
 * (9 species, with sources); a larger one returns PC_ERR_UNSUPPORTED before
 * any device work.  Every call on a GPU shares that GPU's array, so calls
 * that overlap on one GPU (different streams, or another thread) race;
 * calls in order on one stream do not.  Nothing in the array needs clearing
 * (each call writes every element it reads first). */
int pc_hydro_source_box_noalloc_v1(pc_ctx_t* ctx, pc_box_t bx,
                                   const int* domlo, const int* domhi,
                                   const int* bclo, const int* bchi,
                                   pc_array4_t sarr, pc_array4_t src,
                                   pc_array4_t hyd_src, const double* dx,
                                   double dt, const pc_params_t* prm,
                                   const pc_array4_t* flx,
                                   const pc_array4_t* a, pc_array4_t vol,
                                   void* stream);

/* -------------------------------------------------------- level (boxes) --- */
/* A single AMR level as a table of boxes (a BoxArray on one GPU), each with
 * its own FAB of NVAR comps and 4 ghost cells in one device arena, uniform
 * Cartesian geometry (area = dx*dx, vol = dx^3 as amrex::Geometry on a
 * Cartesian grid, Source/PeleC.cpp:600-617).  The level owns the state and
 * runs the whole step on the device: ghost values (intra-level copy,
 * periodic, FOExtrap/reflect by the BC map of Source/Setup.cpp:42-121),
 * hydro source, S_new = S_old + dt*hydro_source, reset_internal_energy +
 * computeTemp, and the next dt (estTimeStep + computeNewDt limits), with dt
 * kept in device memory so consecutive steps need no host synchronisation.
 * The Godunov step reads each ghost value from its source cell inside the
 * primitives kernel and stores only the first ghost layer back into the FAB
 * (the artificial viscosity reads it); the method-of-lines step reads them
 * the same way and stores none.  The FAB's other ghost cells therefore hold
 * stale values after a step; pc_level_fill_ghosts_v1 fills all four layers
 * on request (pc_level_write_plotfile_v1 fills them itself for magvort). */
typedef struct pc_level pc_level_t;

/* Ghost cells beyond Inflow and UserBC domain faces: the reference fills
 * them by ext_dir, i.e. the problem's bcnormal (Source/Setup.cpp:44-60),
 * which is problem code outside this path.  The level and batch paths fill
 * them by first-order extrapolation instead (the Outflow rule); the face
 * fluxes there still come from pc_low_order_boundary's first-order states
 * (Source/Godunov.cpp:7-55), as in the reference.  A caller that needs its
 * own inflow state drives the box entry points (pc_hydro_source_box_v1,
 * pc_umdrv_v1) with ghost cells it has filled. */
typedef struct pc_level_desc {
  int nboxes;
  const pc_box_t* boxes; /* host array of nboxes valid boxes            */
  int domlo[3], domhi[3];
  int bclo[3], bchi[3];    /* PC_BC_* per direction                       */
  int periodic_self;       /* 1: every box is its own periodic domain     */
  double dx[3];
  double problo[3];
} pc_level_desc_t;

int pc_level_create_v1(pc_ctx_t* ctx, const pc_level_desc_t* desc,
                       const pc_params_t* prm, pc_level_t** out);
int pc_level_destroy_v1(pc_level_t* lev);
/* Number of doubles in box b's FAB ((n+8)^3 * NVAR) and its offset in the
 * level arena (the FAB layout is Array4 over grow(box,4), comp slowest). */
long long pc_level_box_offset_v1(const pc_level_t* lev, int b);
long long pc_level_box_size_v1(const pc_level_t* lev, int b);
double* pc_level_state_ptr_v1(pc_level_t* lev); /* device arena base */
long long pc_level_state_size_v1(const pc_level_t* lev);
/* Copy box b's whole FAB (grow(box,4) x NVAR doubles, Array4 layout)
 * between the level arena and `buf` (host or device memory; synchronous). */
int pc_level_write_box_v1(pc_level_t* lev, int b, const double* buf);
int pc_level_read_box_v1(pc_level_t* lev, int b, double* buf);
/* Set / get the device-resident dt (synchronous). */
int pc_level_set_dt_v1(pc_level_t* lev, double dt);
int pc_level_get_dt_v1(pc_level_t* lev, double* dt);
/* Set / get the level time (state[State_Type].curTime(), 0 at creation),
 * which every step advances by its dt on the device and against which the
 * stop_time limit applies (synchronous). */
int pc_level_set_time_v1(pc_level_t* lev, double time);
int pc_level_get_time_v1(pc_level_t* lev, double* time);
/* dt = init_shrink * estTimeStep(state) (or fixed_dt), limited by stop_time
 * at the level time (computeInitialDt, Source/PeleC.cpp:1041-1075), on the
 * device. */
int pc_level_init_dt_v1(pc_level_t* lev, void* stream);
/* Advance nsteps steps (asynchronous on `stream`): Godunov steps, or with
 * prm->do_mol = 1 at creation the two-stage MOL steps of do_mol_advance
 * (Source/Advance.cpp:49-227, hydro only: no reactions or other sources). */
int pc_level_step_v1(pc_level_t* lev, int nsteps, void* stream);
/* Fill the 4 ghost cells of every box (asynchronous). */
int pc_level_fill_ghosts_v1(pc_level_t* lev, void* stream);
/* estTimeStep over the valid cells of the level: cfl*min(dx/(c+|u|)),
 * limited by max_dt (Source/PeleC.cpp:829-961).  Synchronous. */
int pc_level_estdt_v1(pc_level_t* lev, double* dt_out, void* stream);
/* The level as an AMReX plotfile directory `dir` (Header, Level_0/Cell_H,
 * Level_0/Cell_D_00000; the format PeleCAmr::writePlotFileDoit writes
 * through amrex::WriteMultiLevelPlotfile, Source/PeleCAmr.cpp:225-267), for
 * fcompare or any plotfile reader.  vars: space-separated names, NULL/"" for
 * every state component (amr.plot_vars' default); state names density xmom
 * ymom zmom rho_E rho_e Temp rho_<species>; derived (Source/Derive.cpp)
 * pressure kineng soundspeed MachNumber magvort logden Y(<species>) (or
 * massfrac for all) x_velocity y_velocity z_velocity magvel; ALL for every
 * one.  Variables are written in the reference's order (state components,
 * then derived in derive-list order).  Species: H2 O2 H2O H O OH HO2 H2O2 N2
 * (LiDryer), N2 (gamma law).  prob_lo: physical corner of the domain (NULL:
 * origin); time and step label the plotfile.  Synchronous (derives on the
 * device on `stream`, then writes). */
int pc_level_write_plotfile_v1(pc_level_t* lev, const char* dir,
                               const char* vars, double time, int step,
                               const double* prob_lo, void* stream);

/* The same format from host arrays: data[b] holds nvars components of box
 * boxes[b] (AMReX FAB layout, i fastest, component slowest). */
int pc_write_plotfile_v1(const char* dir, int nvars, const char* const* names,
                         int nboxes, const pc_box_t* boxes,
                         const double* const* data, const int* domlo,
                         const int* domhi, const double* prob_lo,
                         const double* dx, double time, int step);

/* Per-kernel device time: with timing on, every launch of a step is
 * bracketed by HIP events on the step's stream and the step waits for its
 * last event (so timed steps do not overlap the host).  Kernel ids, in
 * launch order: */
enum {
  /* the Godunov step has no ghost-fill launch (its primitives kernel reads
   * ghost values from their sources): PC_K_GHOST reads 0 there */
  PC_K_GHOST = 0, PC_K_CTOPRIM, PC_K_DIVU, PC_K_TRACE_X, PC_K_TRACE_Y,
  PC_K_TRACE_Z, PC_K_FLUX1_X, PC_K_FLUX1_Y, PC_K_FLUX1_Z,
  /* the transverse fluxes of d: both (d|t) in one launch, timed under the
   * first t (PC_K_TRANS_XY, _YX, _ZX); _XZ, _YZ and _ZY read 0 */
  PC_K_TRANS_XY,
  PC_K_TRANS_XZ, PC_K_TRANS_YX, PC_K_TRANS_YZ, PC_K_TRANS_ZX, PC_K_TRANS_ZY,
  PC_K_FINAL_X, PC_K_FINAL_Y, PC_K_FINAL_Z, PC_K_UPDATE, PC_K_DT,
  /* method-of-lines step (do_mol = 1): ghost fill and primitives count
   * under PC_K_GHOST / PC_K_CTOPRIM (both stages), then */
  PC_K_MOL_X, PC_K_MOL_Y, PC_K_MOL_Z, PC_K_MOL_UPD1, PC_K_MOL_UPD2,
  PC_NKERN
};
/* on = 1 also resets the accumulated times. */
int pc_level_enable_timing_v1(pc_level_t* lev, int on);
/* Average ms per timed step of each kernel id (n <= PC_NKERN entries). */
int pc_level_kernel_times_v1(pc_level_t* lev, double* ms, int n);
/* Launch shapes.  Every stage of the level step is one launch over all
 * boxes; a launch is a set of logical blocks of 256 threads, each a run of
 * 256 cells of the stage's region or a 2^lx x 2^ly x 2^lz tile, dealt to
 * workgroups (one per logical block, or fewer that loop: persistent).
 * pc_level_autotune_v1 times a fixed candidate set per stage on the level as
 * it stands (`reps` timed launches each, the update stage on a saved copy of
 * the state) and keeps the fastest; state and dt are unchanged on return.
 * Synchronous.  Replaces the tiling choice of amrex::ParallelFor /
 * MFIter tiling in the reference (Source/Hydro.cpp:60-73); no counterpart
 * to the timing itself. */
int pc_level_autotune_v1(pc_level_t* lev, int reps, void* stream);
/* shape[5] = {lx, ly, lz, workgroups, logical blocks} of stage `stage`
 * (0..21: ctoprim, divu, trace, flux1 x/y/z, trans xy/xz/yx/yz/zx/zy, final
 * x/y/z, update, ghost, MOL primitives, MOL fluxes x/y/z, MOL update); lx =
 * -1: runs, -2 / -3: the PPM trace's LDS tiles / pencils, -4: the FP64 MOL
 * fluxes' LDS face tiles (those stages always run on their tiles; a plan set
 * for them changes only the persistence); *ms = its tuned time per launch
 * (0 if never tuned; may be NULL). */
int pc_level_plan_v1(const pc_level_t* lev, int stage, int* shape, double* ms);
/* Set stage `stage`'s plan by hand (e.g. a shape saved from an earlier
 * autotune): lx = -1 for runs, else a tile with lx + ly + lz = 8; wg_per_cu
 * > 0 caps the workgroups at that many per compute unit (persistent).
 * Synchronous. */
int pc_level_set_plan_v1(pc_level_t* lev, int stage, int lx, int ly, int lz,
                         int wg_per_cu);
/* Arithmetic of the level step.  0 (default): FP64 throughout, the
 * reference's arithmetic.  1 (mixed): the primitives and every stage
 * intermediate are stored in FP32 and the PPM/PLM reconstruction, limiters,
 * characteristic tracing, transverse corrections and Riemann solves run in
 * FP32 (with the EOS calls they make); the e -> T Newton inversion (ctoprim
 * and computeTemp), the total-minus-kinetic internal energy, the final flux
 * assembly (artificial viscosity, species normalisation, area) and the
 * conservative update stay FP64, so sums of fluxes telescope exactly and
 * conservation holds to FP64 round-off.  No reference counterpart (the
 * reference computes in FP64 only).  Synchronous. */
int pc_level_set_precision_v1(pc_level_t* lev, int mixed);
/* steps_per_graph > 0: pc_level_step_v1 captures that many consecutive
 * steps into a hipGraph once (re-captured after autotuning) and replays it;
 * leftover steps and timed steps launch directly.  0 turns graphs off. */
int pc_level_set_graph_v1(pc_level_t* lev, int steps_per_graph);
/* on = 1: the stages of the three directions that do not depend on each
 * other (trace + first fluxes, the two transverse fluxes, the final fluxes
 * of a direction; the MOL fluxes) run on three streams of the level, forked
 * from and joined back to the stream a step is issued on, so kernels of
 * different directions overlap on the device.  Results are unchanged (each
 * kernel writes only its own direction's arrays).  Steps that time their
 * kernels (pc_level_enable_timing_v1) stay on one stream.  A hipGraph of
 * pc_level_set_graph_v1 is captured on one stream and its chain of kernel
 * nodes then given the same dependencies, so replays run the directions
 * concurrently too.  No reference counterpart (the reference issues its
 * kernels on one stream).  Synchronous. */
int pc_level_set_branches_v1(pc_level_t* lev, int on);
/* The level's current hipGraph: *nodes = its kernel nodes (0 before the
 * first capture), *branched = 1 when it carries the branch-stream
 * dependencies.  No reference counterpart. */
int pc_level_graph_info_v1(const pc_level_t* lev, int* nodes, int* branched);
/* okToContinue (Source/PeleC.cpp:1310-1340): synchronous; returns
 * PC_ERR_NAN if any step produced NaN/Inf, PC_ERR_DT if dt < dt_cutoff. */
int pc_level_status_v1(pc_level_t* lev, int* bad, int* small_dt);
/* Conservative state of box b's valid cells from primitive fields, the
 * pc_initdata pattern (Exec/RegTests/Sod/prob.H:25-117 with
 * Exec/RegTests/Sod/prob.cpp:215-241).  prim covers the valid box with
 * 5 + NUM_SPECIES comps: mode 0 = (rho, p, u, v, w, Y), T from e = RYP2E
 * by EY2T; mode 1 = (p, T, u, v, w, Y), rho = p Wbar / (R T). */
int pc_level_prim_to_cons_v1(pc_level_t* lev, int b, pc_array4_t prim,
                             int mode, void* stream);

/* ---------------------------------------- flux register, average-down --- */
/* Refluxing between a coarse level and a level refined by `ratio`, for
 * regular Cartesian boxes: the operations PeleC::update_flux_registers
 * (CrseAdd / FineAdd per box, Source/PeleC.cpp:1343-1394, called from
 * Source/Hydro.cpp:280-288) and PeleC::reflux / avgDown (:1396-1437) perform
 * through AMReX's flux register.  Register faces are the coarse faces on the
 * boundary of the refined region whose other ("outside") coarse cell is not
 * refined (faces on a non-periodic domain boundary are not register faces;
 * across a periodic one the outside cell is the periodic image).  Each face
 * accumulates crse_add's scale * coarse flux plus fine_add's scale * (sum of
 * its ratio^2 fine fluxes); fluxes are extensive (x area), as pc_umdrv_v1 and
 * pc_hydro_source_box_v1 return them.  With crse scale -dt_c and fine scale
 * +dt_f (per fine step), reflux adds -reg / V to an outside cell below the
 * face and +reg / V to one above it.  Fine boxes must be aligned to ratio.
 * All calls are stream-ordered; a box id is the index in the create call's
 * list, bx that box (its flux views cover surroundingNodes(bx, d)). */
typedef struct pc_fluxreg pc_fluxreg_t;
int pc_fluxreg_create_v1(pc_ctx_t* ctx, const pc_box_t* crse_boxes, int ncrse,
                         const pc_box_t* fine_boxes, int nfine,
                         const int* crse_domlo, const int* crse_domhi,
                         const int* periodic, int ratio, pc_fluxreg_t** out);
int pc_fluxreg_destroy_v1(pc_fluxreg_t* reg);
/* Number of register faces (or -1). */
int pc_fluxreg_nfaces_v1(const pc_fluxreg_t* reg);
/* Zero the register (a fill kernel: safe inside a captured graph). */
int pc_fluxreg_reset_v1(pc_fluxreg_t* reg, void* stream);
int pc_fluxreg_crse_add_v1(pc_fluxreg_t* reg, int crse_box, pc_box_t bx,
                           const pc_array4_t* flx, double scale, void* stream);
int pc_fluxreg_fine_add_v1(pc_fluxreg_t* reg, int fine_box, pc_box_t bx,
                           const pc_array4_t* flx, double scale, void* stream);
/* state: coarse box crse_box's state (covering bx); vol.p = NULL: uniform
 * volume vol_u. */
int pc_fluxreg_reflux_v1(pc_fluxreg_t* reg, int crse_box, pc_box_t bx,
                         pc_array4_t state, pc_array4_t vol, double vol_u,
                         void* stream);
/* crse = (1 / ratio^3) * sum of the ratio^3 fine cells over crse_region
 * (conserved state, uniform volumes; the fine view covers the refined
 * region). */
int pc_average_down_v1(pc_ctx_t* ctx, pc_box_t crse_region, pc_array4_t crse,
                       pc_array4_t fine, int ratio, void* stream);

/* ------------------------------------------------- batch (ensembles) --- */
/* Independent requests resident on the device: each is one box that is its
 * own domain, with its own boundary codes, cell size and dt (a shock tube or
 * a premixed slab of an ensemble).  Memory comes from one arena of
 * `budget_bytes` allocated at creation (admission control: pc_batch_add_v1
 * fails with PC_ERR_NOMEM, without side effects, when the request's state and
 * stage workspace do not fit).  pc_batch_step_v1 advances a set of requests
 * by one step each with one launch per stage over all their boxes (the
 * batched, continuous form of pc_level_step_v1); every request keeps its own
 * dt (estTimeStep/computeNewDt per request).  The arithmetic of a request
 * does not depend on which others share its steps.  No reference counterpart
 * (the reference steps one AMR hierarchy per process). */
typedef struct pc_batch pc_batch_t;
/* mixed: as pc_level_set_precision_v1.  share = NULL: a new arena of
 * budget_bytes; else draw on share's arena (one budget for batches of
 * different mechanisms), which lives until its last batch is destroyed.
 * max_slots: most requests alive in this batch.  prm->do_mol = 1 makes the
 * batch step the method-of-lines step of a level (both Heun stages). */
int pc_batch_create_v1(pc_ctx_t* ctx, const pc_params_t* prm, int mixed,
                       long long budget_bytes, pc_batch_t* share,
                       int max_slots, pc_batch_t** out);
int pc_batch_destroy_v1(pc_batch_t* batch);
/* device bytes a request on box bx takes from the arena */
long long pc_batch_slot_bytes_v1(const pc_batch_t* batch, pc_box_t bx);
int pc_batch_memory_v1(const pc_batch_t* batch, long long* total,
                       long long* used, long long* largest_free);
/* The arena's placement rule, for callers that plan admissions ahead: a
 * request takes one block of its bytes rounded up to a multiple of 256, at
 * the start of the free range of lowest offset that holds it (first fit);
 * a removed request's block merges with free neighbours.  No reference
 * counterpart (admission planning of an ensemble server). */
/* The free ranges in ascending offset order: the first min(cap, *count)
 * into offsets[] / sizes[] (bytes from the arena's base); *count = how many
 * there are. */
int pc_batch_free_ranges_v1(const pc_batch_t* batch, long long* offsets,
                            long long* sizes, int cap, int* count);
/* The block of a live request: offset and bytes (rounded as above). */
int pc_batch_slot_extent_v1(const pc_batch_t* batch, int slot,
                            long long* offset, long long* bytes);
/* Admit a request: box bx (its own domain), PC_BC_* codes per direction,
 * cell size dx[3], primitive fields over bx as pc_level_prim_to_cons_v1
 * (device memory, prim_mode 0 or 1).  Builds the conservative state and the
 * request's initial dt (init_shrink * estTimeStep); synchronous.  *slot
 * names the request in the other calls. */
int pc_batch_add_v1(pc_batch_t* batch, pc_box_t bx, const int* bclo,
                    const int* bchi, const double* dx, pc_array4_t prim,
                    int prim_mode, void* stream, int* slot);
/* One step of each listed request (distinct live slots); asynchronous.  When
 * the set of listed slots changes, its job tables go to the device on
 * `stream` through the next of four rotating table sets; the host waits only
 * if that set's last step is still queued. */
int pc_batch_step_v1(pc_batch_t* batch, const int* slots, int n, void* stream);
/* on = 1: the batch's steps run the three directions' independent stages on
 * three streams, as pc_level_set_branches_v1; results are unchanged.  No
 * reference counterpart.  Synchronous. */
int pc_batch_set_branches_v1(pc_batch_t* batch, int on);
/* The request's whole state FAB (grow(bx,4) x NVAR, Array4 layout) into buf
 * (host or device memory); synchronous. */
int pc_batch_read_v1(pc_batch_t* batch, int slot, double* buf);
/* dt of its next step and the NaN / dt_cutoff flags; synchronous. */
int pc_batch_status_v1(pc_batch_t* batch, int slot, double* dt, int* bad,
                       int* small_dt);
/* The same without waiting for the device (no reference counterpart; for a
 * request loop that keeps steps queued ahead of the host): snapshot queues a
 * copy of the request's dt state to pinned host memory on `stream` (after
 * its last step); snapshot_status returns what the copy brought once the
 * caller has seen the stream pass that point (an event). */
int pc_batch_snapshot_v1(pc_batch_t* batch, int slot, void* stream);
int pc_batch_snapshot_status_v1(const pc_batch_t* batch, int slot, double* dt,
                                int* bad, int* small_dt);
/* Retire a request and return its memory.  stream: waited for first (the
 * stream of the request's last step or read); NULL: no wait, the caller has
 * seen the request's work complete. */
int pc_batch_remove_v1(pc_batch_t* batch, int slot, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* PC_HYDRO_H */
