/*
 * fesomc_gpu.h -- C-ABI of the MI355X-native FESOM-C internal time step.
 *
 * The reference (rakowsk/fesom-c) has no operator API: its hot path is a set
 * of argument-less Fortran subroutines that mutate module globals
 * (o_MESH / o_ARRAYS / o_PARAM, src/fv_var.F90:303-438).  This header is
 * the boundary a Fortran host binds with ISO_C_BINDING: one entry point per
 * call site of oce_timestep (src/fv_main.F90:1996-2076) plus set-up,
 * state transfer, diagnostics and teardown.  All arguments are plain
 * pointers and sizes; arrays use the reference's Fortran layout
 * (column-major, e.g. X(nsigma-1, nod2D) has the level index fastest) and
 * 1-based node/element/edge ids exactly as the Fortran arrays hold them.
 *
 * Every function returns 0 on success or a negative FCG_E* code; the text
 * of the last error is available from fcg_last_error().  Calls on one
 * context are stream-ordered and asynchronous unless stated otherwise; a
 * context is not re-entrant (one host thread per context), matching the
 * reference's single-threaded call sites.
 */
#ifndef FESOMC_GPU_H
#define FESOMC_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define FCG_OK 0
#define FCG_E_ARG (-1)        /* invalid argument / unsupported option      */
#define FCG_E_MESH (-2)       /* mesh check failed (test_elem, edges)       */
#define FCG_E_HIP (-3)        /* HIP runtime error                          */
#define FCG_E_NAN (-4)        /* non-finite elevation or tracer (no reference STOP fired) */
#define FCG_E_SALINITY (-5)   /* "Negative salinity ..." STOP (fv_tracer.F90:1469-1473) */
#define FCG_E_TEMP (-6)       /* "Negative temperature ..." STOP (fv_tracer.F90:1479-1482) */
#define FCG_E_NOGPU (-7)      /* no HIP device: the library does not fall back */

/* ---- precision / closure selectors ------------------------------------ */
#define FCG_FP64 0            /* plain FP64 drop-in                          */
#define FCG_MIXED 1           /* FP32 fluxes/tendencies, FP64 elevation, wet/dry, BPG */
#define FCG_CLOSURE_D3END 0   /* reference closure d3_end (fv_mixing.F90:73-347) */
#define FCG_CLOSURE_MLP 1     /* learned per-column closure with d3_end fallback */

/*
 * Run parameters: the o_PARAM values READ_DATA_RUN derives from fv_run.dat
 * (src/fv_read_run.F90).  Units and meaning follow the reference; time
 * scalars in days (time_jd0) and seconds (dt, dt_2D).
 */
typedef struct fcg_params {
    double dt, dt_2D;          /* baroclinic / barotropic steps [s]          */
    double time_jd0;           /* JD0                                        */
    double Dmin;               /* minimal depth (wet/dry threshold)          */
    double C_d;                /* bottom drag Cd                             */
    double z0b_min, z0s_min;   /* roughness lengths (d3_end)                 */
    double snul;               /* background viscosity                       */
    double Cd_bz, Bp_bz;       /* buffer-zone drag / BPG factors             */
    double clim_relax;         /* [1/s], already inverted as fv_read_run.F90:884 */
    double relax2clim_ac;      /* [1/s], fv_read_run.F90:894                 */
    int nsigma;                /* number of sigma levels (layers = nsigma-1) */
    int Mt;                    /* barotropic substeps per internal step      */
    int tracer_adv;            /* 1 upwind, 2 MUSCL, 4 Miura                 */
    int muscl_call;            /* 1: tracer_adv=2 runs solve_tracer_muscl    */
    int mom_adv_2D, mom_adv_3D, vert_adv;
    int im_vert_visc;          /* 1: momentum_vert_impl_visc (fv_viscosity.F90:1266-1371) */
    int im_vert_visc_fic;      /* 1: momentum_vert_impl_visc_tmp; with both set the
                                  _tmp solve follows on the first's result
                                  (fv_3D_vel.F90:183-186)                   */
    int fic_point;             /* 0,1,2 (fv_viscosity.F90:1449-1460)         */
    int wet_dry_on;
    int tf_presence;           /* tidal open-boundary elevation              */
    int restart;               /* hot start: no 10-day tidal ramp            */
    int cl_relax;              /* OB climatology relaxation                  */
    int precision;             /* FCG_FP64 | FCG_MIXED                       */
    int closure;               /* FCG_CLOSURE_D3END | FCG_CLOSURE_MLP        */
    /* surface forcing (key_atm): the fields qsr, qns, emp, mslp are set by
     * the host (fcg_set_state) from its own sbc_do; the step applies them
     * as the reference does (fv_average_dynamic.F90:278-285, fv_3D_vel.F90:
     * 40-53, fv_tracer.F90:1376-1413) */
    int key_atm;
    int type_swr_body;         /* 1 Paulson-Simpson (constant R, a, b), 2 Soloviev */
    double Rjc, ajc, bjc;      /* fv_run.dat keys R, a, b (jer_const = T)    */
    /* rivers (key_rivers): the field "runoff" (m3/s per node, 0 away from
     * river mouths) is set by the host from its own river table; the step
     * adds it to the elevation's right-hand side (fv_average_dynamic.F90:
     * 177-183) and dilutes the mouth's salinity towards 0.2
     * (fv_main.F90:1805-1822) */
    int key_rivers;
    /* neighbour-difference filters (fv_viscosity.F90:769-939): filt_2D on
     * the barotropic velocity, filt_3D on the 3-D velocity and its
     * pulsation part in the barotropic forcing, both at the rate
     * tau_inv_filt [1/s] */
    int filt_2D, filt_3D;
    double tau_inv_filt;
    /* the twice-applied filters (viscosity_filt2x_2D/_3D, fv_viscosity.F90:
     * 570-767; switched at fv_average_dynamic.F90:347 and fv_3D_vel.F90:112):
     * filt_bi_2D on U_n_2D into UV2_rhs of update_2D_vel, filt_bi_3D on
     * U_n*Je into UV_rhs of compute_vel_rhs, at the rate tau_c [1/day]
     * (fv_read_run.F90:336-341) scaled by sqrt(scale_area/elem_area) */
    int filt_bi_2D, filt_bi_3D;
    double tau_c;
    /* biharmonic viscosity (fv_viscosity.F90:104-568, 964-1220): bih_2D on
     * UAB into U_rhs_2D of update_2D_vel (fv_average_dynamic.F90:348),
     * bih_3D on U_n into U_rhs of compute_vel_rhs (fv_3D_vel.F90:113) and on
     * the pulsation velocity into the barotropic forcing of the first substep
     * (fv_main.F90:1501), with the coefficient Abh0 (fv_read_run.F90:320-325) */
    int bih_2D, bih_3D;
    double Abh0;
} fcg_params;

/*
 * Mesh after the reference's own preprocessing (read_mesh_ser, test_elem,
 * find_edges, find_elem_neighbors, mesh_arrays1/2, SET_SIGMA,
 * initial_state constants).  Arrays are the reference module arrays with
 * their Fortran shapes; ids 1-based; -999 marks a missing neighbour.
 */
typedef struct fcg_mesh {
    int nod2D, elem2D, edge2D, edge2D_in, nsigma, nobn;
    const int *elem2D_nodes;     /* (4, elem2D)                              */
    const int *edge_nodes;       /* (2, edge2D)                              */
    const int *edge_tri;         /* (2, edge2D)                              */
    const int *elem_edges;       /* (4, elem2D)                              */
    const int *elem_neighbors;   /* (4, elem2D)                              */
    const int *index_nod2D;      /* (nod2D)                                  */
    const int *in_obn;           /* (nobn) open-boundary nodes, ascending    */
    const double *depth;         /* (nod2D)                                  */
    const double *sigma;         /* (nsigma)                                 */
    const double *area;          /* (nod2D)                                  */
    const double *elem_area;     /* (elem2D)                                 */
    const double *elem_cos;      /* (elem2D)                                 */
    const double *coriolis;      /* (elem2D)                                 */
    const double *w_cv;          /* (4, elem2D)                              */
    const double *gradient_sca;  /* (8, elem2D)                              */
    const double *gradient_vec;  /* (8, elem2D)                              */
    const double *edge_dxdy;     /* (2, edge2D)                              */
    const double *edge_cross_dxdy; /* (4, edge2D)                            */
    const double *ac;            /* (nod2D) sponge coefficient               */
    const double *a_tp;          /* (nod2D, 2)                               */
    const double *ampt, *fazt;   /* (nobn, 12) tidal amplitude / phase [rad] */
    const double *amp_factor, *phy_factor; /* (12)                           */
    const double *coord_nod2D;   /* (2, nod2D) radians; optional: used only to
                                    order the device arrays for locality     */
    double density_0;            /* fv_ini.F90:187-199                        */
    double scale_area;           /* fv_ini.F90:81                             */
} fcg_mesh;

typedef struct fcg_ctx fcg_ctx;
typedef struct fcg_mlp fcg_mlp;

/* Diagnostics gathered on the device (deterministic reductions). */
typedef struct fcg_diag {
    double energy;       /* energy(eout): fv_average_dynamic.F90:476-554     */
    double volume;       /* sum area*(depth+eta_n)                            */
    double heat;         /* sum area*Jc*TF  (tracer content)                  */
    double salt;         /* sum area*Jc*SF                                    */
    double eta_min, eta_max, T_min, T_max, S_min, S_max;
    double time_jd;      /* model clock after the last step                   */
    int n_dt;            /* internal steps taken                              */
    int n_wet_elem;      /* sum of mask_wd                                    */
    long long mlp_fallback; /* column-steps sent back to d3_end by the OOD guard */
    int status;          /* first device-side STOP condition (FCG_E_* or 0)   */
} fcg_diag;

/* ---- set-up -------------------------------------------------------------
 * fcg_mesh_build: the project's own C++ restatement of the reference mesh
 * preprocessing (read_mesh_ser ordering src/fv_main.F90:853-893, test_elem
 * src/fv_mesh_array.F90:1-82, find_edges :395-718, find_elem_neighbors
 * :791-946, mesh_arrays1/2 :948-1386, SET_SIGMA src/fv_vert_coord.F90:174-195,
 * initial_state constants src/fv_ini.F90:81-107,187-240), for hosts that do
 * not run the reference's Fortran set-up.  Inputs are the raw file contents:
 *   xy(2,nod2D) as nod2d.out holds them, ind(nod2D), elem(4,elem2D) 1-based
 *   as elem2d.out holds them, depth(nod2D), ob(2*n_harm, nobn) m2.out rows.
 * cfg carries the scalar switches below (cartesian, Coriolis, sigma law ...).
 * The mesh object owns its arrays; fcg_mesh_view exposes them as fcg_mesh.
 */
typedef struct fcg_mesh_cfg {
    int cartesian, coriolis_tf, nsigma, ind_sigma, is_LE, T_potential, tf_presence;
    double lat_cartesian_deg, cyclic_length_deg, pow_sig, lev_bot, lev_sur, T_const, S_const;
    int harmonics[12];           /* 1 if harmonic j (M2,S2,...,M4) is active  */
    int SL_obn;                  /* sponge layer at the open boundary: ac < 1 */
    double SL_radius;            /* its width in metres (fv_var.F90:207-209)  */
} fcg_mesh_cfg;
typedef struct fcg_mesh_obj fcg_mesh_obj;

int fcg_mesh_build(const fcg_mesh_cfg *cfg, int nod2D, const double *xy, const int *ind, int elem2D,
                   const int *elem, const double *depth, int nobn, const double *ob, fcg_mesh_obj **out);
int fcg_mesh_view(const fcg_mesh_obj *m, fcg_mesh *view);
void fcg_mesh_free(fcg_mesh_obj *m);

/* Replaces the device part of array_setup + jacobian + initial_state
 * (src/fv_main.F90:157-168, 320-362): uploads topology/geometry, builds the
 * CSR gather lists, allocates the device state and initialises it exactly as
 * the reference does on a cold start.  Fails with FCG_E_NOGPU when no HIP
 * device is present. */
int fcg_create(const fcg_mesh *mesh, const fcg_params *par, fcg_ctx **out);
void fcg_destroy(fcg_ctx *ctx);

/* Ensemble (no counterpart in the reference, which runs one model per
 * executable): nmemb models with their own meshes (uneven sizes) and forcing
 * advanced together, every stage one launch over all members.  The meshes
 * are joined into one disjoint mesh: nodes and elements concatenated in
 * member order, internal edges of all members first, then their boundary
 * edges, each member keeping its own order, so each member evolves exactly
 * as it would alone (T_aver/S_aver and diagnostics are per member).  The
 * members must share nsigma, sigma, density_0 and the run parameters; state
 * transfers address the joined arrays (member b's nodes follow those of
 * members 0..b-1).  scale_area (Miura only) is taken from member 0. */
int fcg_create_ensemble(const fcg_mesh *members, int nmemb, const fcg_params *par, fcg_ctx **out);
int fcg_ensemble_size(const fcg_ctx *ctx);

/* ---- state transfer (host mirrors <-> device), synchronous -------------
 * name is the reference array name: eta_n, eta_n_1, eta_n_2, U_n_2D, U_n_1,
 * U_n_2, U_n, V_n, U_rhsAB, V_rhsAB, Wvel, bt, snu, Kv, TF, SF, T_old,
 * S_old, Tclim, Sclim, taux, tauy, windx, windy, mask_wd, C_d_el, Je,
 * Bar_pru_3D, Bar_prv_3D, Bar_pr_2D, hpressure, U_n_filt, V_n_filt,
 * U_filt_2D, V_filt_2D, U_rhs, V_rhs, ssh_rhs, etaAB, dmean_n, qsr, qns,
 * emp, mslp, runoff; tau_node (2, nod2D: taux_node, tauy_node of the bulk
 * formulas, held once fcg_sbc_set_coef has been called); and, for
 * fcg_get_state only, Jc, Jc_old, Z, zbar (formed from the column depth and
 * sigma as the jacobian forms them; the device does not store them).
 * count is the number of doubles the Fortran array holds; the library
 * transposes between Fortran layout and its device layout. */
int fcg_set_state(fcg_ctx *ctx, const char *name, const double *host, long count);
int fcg_get_state(fcg_ctx *ctx, const char *name, double *host, long count);
/* Scalar state: time_jd (days), lfirst, n_dt. */
int fcg_set_clock(fcg_ctx *ctx, double time_jd, int lfirst, int n_dt);
/* Hot start, after the restart records (eta_n ... Sclim, src/fv_restart.F90:
 * 518-600) have been set with fcg_set_state: the set-up jacobian from the
 * restored elevation (the call at src/fv_main.F90:362 that follows
 * read_restart_separate at 336; Jc_old keeps the array set-up value), the
 * clock at the restart's time_jd and lfirst = .false. (fv_restart.F90:611).
 * The context must have been created with restart = 1 (no tidal ramp). */
int fcg_hot_start(fcg_ctx *ctx, double time_jd);

/* ---- surface boundary conditions from atmospheric fields (key_atm) -----
 * The reference's sbc_do (src/fv_sbc.F90:1012-1046, called at
 * src/fv_main.F90:419 after the clock advance) reads NetCDF records and
 * interpolates them onto the nodes (getcoeffld, 665-841), then forms each
 * field at the clock and applies the CORE bulk formulas.  The reading stays
 * on the host; fcg_sbc_set_coef hands over getcoeffld's result: coef_a,
 * coef_b (8, nod2D), Fortran layout, mesh node order, fields x-wind, y-wind,
 * specific humidity, shortwave, longwave, 2 m air temperature [K],
 * precipitation, sea-level pressure (the reference's i_xwind..i_mslp); the
 * field at time t [days] is t*coef_a + coef_b.  Call again whenever the host
 * reads the next records.  FCG_E_ARG without key_atm. */
int fcg_sbc_set_coef(fcg_ctx *ctx, const double *coef_a, const double *coef_b);
/* data_timeinterp + apply_atm_fluxes at the context's clock (src/fv_sbc.F90:
 * 843-867, 368-482) for task type 3 without sea ice: windx/y, the element
 * stress taux/y, qns, emp, qsr and mslp, from the wind relative to the top
 * layer's node velocity and the top layer's temperature (core_coeff_2z, Large
 * & Yeager 2004, 485-663).  nm_prec_coef: precipitation divisor (1 NCEP,
 * 3600 CoastDat); nm_net_flux: 0 downward longwave, 1 net longwave. */
int fcg_sbc_do(fcg_ctx *ctx, double nm_prec_coef, int nm_net_flux);

/* ---- open-boundary T/S climatology (key_obc) ---------------------------
 * obc_do (src/fv_obc.F90:577-623, called at src/fv_main.F90:431 after the
 * clock advance) reads the file of T and S profiles at the open-boundary
 * nodes and keeps time coefficients per node and depth (getcoeffld,
 * 418-520).  fcg_obc_set_coef hands them over: ndepth ascending depths [m,
 * positive down], coef_a, coef_b (2, nobn, ndepth) Fortran layout (T, S;
 * open-boundary nodes in in_obn order); call again when the host reads the
 * next records.  fcg_obc_do forms the profiles at time_jd (the clock, or the
 * time of the last interpolation once the reference's one_field holds,
 * 610-616) and interpolates them linearly to every layer's mid-depth Z
 * (constant below the last depth) into Tclim/Sclim of the open-boundary
 * nodes (apply_obc, profile_interp, 376-419).  FCG_E_ARG for a layer above
 * the first depth, where the reference has no record to read. */
int fcg_obc_set_coef(fcg_ctx *ctx, int ndepth, const double *depth, const double *coef_a, const double *coef_b);
int fcg_obc_do(fcg_ctx *ctx, double time_jd);

/* ---- the hot path: one entry per call site of oce_timestep ------------
 * fcg_advance_clock:        time_jd += dt/86400; time = time_jd*86400 (src/fv_main.F90:400-401)
 * fcg_pressure:             call pressure                  (src/fv_main.F90:1996)
 * fcg_vertical_mixing:      call compute_el2nodes_3D + d3_end (src/fv_main.F90:2006,2025)
 * fcg_compute_vel_rhs:      call compute_vel_rhs           (src/fv_main.F90:2034)
 * fcg_barotropic_timestep:  call oce_barotropic_timestep   (src/fv_main.F90:2041), the
 *                           Mt-substep loop, replayed as one captured HIP graph
 * fcg_jacobian:             call jacobian                  (src/fv_main.F90:2050)
 * fcg_update_3d_vel:        call update_3D_vel             (src/fv_main.F90:2057)
 * fcg_vert_vel_sigma:       call vert_vel_sigma            (src/fv_main.F90:2064)
 * fcg_solve_tracers:        call solve_tracers             (src/fv_main.F90:2076)
 * fcg_oce_timestep:         the clock update followed by all of the above, in order
 */
int fcg_advance_clock(fcg_ctx *ctx);
int fcg_pressure(fcg_ctx *ctx);
int fcg_vertical_mixing(fcg_ctx *ctx);
int fcg_compute_vel_rhs(fcg_ctx *ctx);
int fcg_barotropic_timestep(fcg_ctx *ctx);
int fcg_jacobian(fcg_ctx *ctx);
int fcg_update_3d_vel(fcg_ctx *ctx);
int fcg_vert_vel_sigma(fcg_ctx *ctx);
int fcg_solve_tracers(fcg_ctx *ctx);
int fcg_oce_timestep(fcg_ctx *ctx);
/* nsteps internal steps back to back (no host round trips). */
int fcg_run(fcg_ctx *ctx, int nsteps);

/* ---- diagnostics / control --------------------------------------------- */
int fcg_diagnostics(fcg_ctx *ctx, fcg_diag *out);  /* synchronous */
/* The same for one ensemble member (0-based). */
int fcg_diagnostics_member(fcg_ctx *ctx, int member, fcg_diag *out);
int fcg_sync(fcg_ctx *ctx);
/* Copy of the last error text, NUL-terminated, truncated to len. */
int fcg_last_error(const fcg_ctx *ctx, char *buf, int len);
/* Kernel timing probe (measurement hook; no counterpart in the reference):
 * bracket every launch of the named device kernel (e.g. "k_d3end") with HIP
 * events on the context's stream.  Kernels inside the captured barotropic
 * graph are not probed.  name NULL or "" disables.  fcg_probe_read is
 * synchronous, returns the summed event time and launch count, and resets. */
int fcg_probe(fcg_ctx *ctx, const char *kernel);
int fcg_probe_read(fcg_ctx *ctx, double *total_ms, int *launches);
/* ---- learned vertical closure ------------------------------------------
 * No counterpart in the reference: these stand in for d3_end
 * (src/fv_mixing.F90:73-347, called at src/fv_main.F90:2025) when a context
 * is created with closure = FCG_CLOSURE_MLP.  The network maps five features
 * of each (column, level) -- log10 shear^2, asinh(1e5 N^2), log10 depth,
 * log10 bottom stress, sigma -- to log10 snu and log10 Kv (5 -> 128 -> 128 ->
 * 128 -> 2, ReLU, BF16 matrix-core products with FP32 accumulation).  Columns whose
 * normalised features leave the box [x_lo, x_hi], and columns of the
 * open-boundary sponge layer (ac < 1), fall back to d3_end; the number of
 * such columns summed over all steps is fcg_diag.mlp_fallback.
 *
 * Training uses device buffers (X: n x 5 normalised features, T: n x 2
 * normalised targets, float, row-major) owned by the caller. */
int fcg_mlp_create(unsigned seed, fcg_mlp **out);        /* Glorot init */
void fcg_mlp_destroy(fcg_mlp *net);
/* Flat FP32 parameters W1[128][32] b1[128] W2[128][128] b2[128] W3[128][128]
 * b3[128] W4[16][128] b4[16] (count 39312); padding entries stay zero. */
int fcg_mlp_get_params(fcg_mlp *net, float *host, long count);
int fcg_mlp_set_params(fcg_mlp *net, const float *host, long count);
/* Feature/target normalisation x' = (x - mean)/std, y = y'*std + mean, and
 * the accepted feature box in normalised units (5, 5, 2, 2, 5, 5 floats). */
int fcg_mlp_set_norm(fcg_mlp *net, const float *x_mean, const float *x_std, const float *y_mean,
                     const float *y_std, const float *x_lo, const float *x_hi);
int fcg_mlp_forward(fcg_mlp *net, const float *X_dev, float *Y_dev, long n);   /* synchronous */
/* Gradient of 0.5*mean_i |f(X_i) - T_i|^2 with respect to the parameters,
 * copied to grad_host (count 39312); synchronous. */
int fcg_mlp_gradient(fcg_mlp *net, const float *X_dev, const float *T_dev, long n, float *grad_host);
/* The same gradient without any device allocation: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under HIP
 * graph capture, tight training loops).  Writes the 39312 gradient entries to
 * the caller's device buffer grad_dev, launching on `stream` (a hipStream_t;
 * NULL = the network's own stream) without synchronising.  The per-wave
 * partial sums live in a file-scope device array of the library (128 x 39312
 * floats, 20 MB per GPU, held until the process ends); it serves any batch
 * since each wave sums several 16-sample tiles.  n must lie in
 * 1..FCG_MLP_STATIC_MAX_SAMPLES (FCG_E_ARG otherwise).  Calls that overlap
 * on one GPU race on that array.  The result equals fcg_mlp_gradient's bit
 * for bit. */
#define FCG_MLP_STATIC_MAX_SAMPLES 2147483647L
int fcg_mlp_gradient_static(fcg_mlp *net, const float *X_dev, const float *T_dev, long n, float *grad_dev,
                            void *stream);
/* One Adam step (beta 0.9/0.999, eps 1e-8) on that loss; synchronous. */
int fcg_mlp_train_step(fcg_mlp *net, const float *X_dev, const float *T_dev, long n, float lr);
/* Attach a trained network to a context created with FCG_CLOSURE_MLP (the
 * network must outlive the context's steps). */
int fcg_set_mlp(fcg_ctx *ctx, fcg_mlp *net);
/* Training samples from the current state, to be called right after
 * fcg_vertical_mixing: raw features X_dev[i][5] and targets T_dev[i][2] =
 * log10 snu, log10 Kv for all nsigma levels of every stride-th column
 * (column-major sample order), at most cap samples; *count = samples written. */
int fcg_mlp_collect(fcg_ctx *ctx, float *X_dev, float *T_dev, long cap, int stride, long *count);

/* Launch shapes (no counterpart in the reference; used by the autotuner,
 * fesomc_amd/autotune.py): block size 64, 128 or 256 threads for the named
 * device kernel ("k_d3end", "k_tr_upwind", ...; layer-parallel kernels then
 * cover 64 entities x block/64 layers per block); 256, 512 or 1024 threads of
 * whole columns for "k_velrhs" (FCG_E_ARG when its LDS would not fit).
 * Default 256.  "<kernel>:levels" sets the levels each lane owns in kernels
 * with a level-vector variant: 1, or W = 4 (FP32 storage) / 2 (FP64) for one
 * 16-byte access per W levels, which needs nsigma to be a multiple of W (the
 * default then).  Results do not depend on the shape. */
int fcg_set_launch(fcg_ctx *ctx, const char *kernel, int block);
int fcg_get_launch(const fcg_ctx *ctx, const char *kernel);

/* Library build identification ("fesomc_gpu <ver> gfx950 ..."). */
const char *fcg_version(void);

#ifdef __cplusplus
}
#endif
#endif /* FESOMC_GPU_H */
