/*
 * adpres_gpu.h -- C-ABI of the MI355X (gfx950) static k-eigenvalue engine
 * that replaces ADPRES's CMFD + SANM/PNM hot path.
 *
 * The reference has no plugin or FFI layer: its hot path is the Fortran
 * module procedures cmfd::outer / cmfd::outer_ad (src/mod_cmfd.f90:415,602)
 * working on the global arrays of module sdata (src/mod_data.f90).  This
 * header is the seam a Fortran host crosses through iso_c_binding instead
 * (the interface module is shown in INTEGRATION.md).  Every entry point
 * takes plain pointers to host arrays laid out exactly as the reference
 * holds them: Fortran column-major, float64, 1-based integer indices.
 *
 *   f0(N,G), fs0(N), D/sigr/nuf/sigf/exsrc(N,G), sigs(N,G,G) = sigs(n,from,to),
 *   chi(nmat,G), dc/dn/df(N,G,6) with faces 1..6 = x+, x-, y+, y-, z+, z-.
 *
 * A context owns all device memory; the host keeps ownership of its arrays
 * (they are copied in by the set calls and copied out by the get calls).
 * A context is used from one host thread at a time; independent contexts
 * may be used concurrently.  All functions return an ADG status code.
 */
#ifndef ADPRES_GPU_H
#define ADPRES_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  The reference prints a message and STOPs in these cases;
 * the engine returns the code and the host shim prints the message. */
enum adg_status {
  ADG_OK = 0,
  ADG_E_MAX_OUTER = 1,      /* src/mod_cmfd.f90:498-505  NOUT reached              */
  ADG_E_NODAL_UNSTABLE = 2, /* src/mod_nodal.f90:131-142 max |d(dn)| > 1e3          */
  ADG_E_LU_PIVOT = 3,       /* src/mod_nodal.f90:856-862 |a_ii| < 1e-4 (node in loc) */
  ADG_E_POWER_NONPOS = 4,   /* src/mod_cmfd.f90:1322-1326 total power <= 0          */
  ADG_E_ARG = 5,            /* invalid argument / unsupported geometry              */
  ADG_E_HIP = 6,            /* HIP runtime error (adg_last_error has the text)       */
  ADG_E_NOT_SET = 7,        /* cross sections not set yet                            */
  ADG_E_XS = 8,             /* src/mod_xsec.f90:215,86-150 the reference STOPs:      */
                            /* sigtr < 1e-5, D ~ 0 or negative sigr, nuf or sigs     */
  ADG_E_DEVICE_SYNC = 9     /* no reference counterpart: a device-wide wait of the   */
                            /* persistent BiCGSTAB kernel timed out (its blocks were */
                            /* not resident together, e.g. another process's         */
                            /* persistent kernel on the same GPU)                    */
};

enum adg_kern { ADG_KERN_FDM = 0, ADG_KERN_PNM = 1, ADG_KERN_SANM = 2 };   /* %KERN */
enum adg_mode { ADG_FORWARD = 0, ADG_ADJOINT = 1, ADG_FIXED_SOURCE = 2, ADG_TRANSIENT = 3,
                ADG_ADJOINT_FROZEN = 4 };  /* %MODE */
enum adg_precision {
  ADG_FP64 = 0,          /* plain FP64 drop-in                                        */
  ADG_FP32_REFINE = 1    /* FP32 BiCGSTAB on an FP64 defect (residual) each group solve */
};

/* Geometry, as built by the reference's mod_io (misc, src/mod_io.f90:1293). */
typedef struct adg_geom {
  int nnod, ng, nmat;
  int nxx, nyy, nzz;
  const int *ix, *iy, *iz;        /* (N) 1-based node coordinates, numbering of misc() */
  const double *xdel, *ydel, *zdel;
  int bc[6];                      /* xeast, xwest, ynorth, ysouth, zbott, ztop: 0 zero flux,
                                     1 zero incoming current, 2 reflective */
  int precision;                  /* enum adg_precision */
} adg_geom;

/* Iteration control: the %ITER and %KERN cards (src/mod_io.f90:1522,1593). */
typedef struct adg_iter {
  int nout, nin, nac, nupd;
  int kern;                       /* enum adg_kern */
  double serc, ferc;
} adg_iter;

/* Per-outer trace, i.e. what outer prints (src/mod_cmfd.f90:331,367-374,491). */
typedef struct adg_trace {
  int cap;                        /* capacity of every array (>= nout)            */
  int n;                          /* outers performed                              */
  double *k, *ser, *fer, *ndmax;  /* (cap)                                         */
  int *flags;                     /* (cap) bit0 extrapolated, bit1 nodal update    */
  int *loc;                       /* (3*cap) i,j,k of the max D-hat change         */
} adg_trace;

typedef struct adg_ctx adg_ctx;

/* Replaces the allocation/geometry part of matrix_setup's first call and
 * set_ind (src/mod_cmfd.f90:217-241,141-213). */
int adg_create(const adg_geom *geom, adg_ctx **ctx);
void adg_destroy(adg_ctx *ctx);

/* Node-wise cross sections after xsec::XS_updt (src/mod_xsec.f90:11-46):
 * D and sigr as made by Dsigr_updt (:199-226); chi per material with the
 * node material map mat(N); exsrc may be NULL (zero). */
int adg_set_xs(adg_ctx *ctx, const double *D, const double *sigr, const double *sigs,
               const double *nuf, const double *sigf, const double *chi, const int *mat,
               const double *dc, const double *exsrc);

/* Flux, fission source, k and D-hat, i.e. the SAVEd state the reference
 * keeps between outer calls (src/mod_cmfd.f90:439-454, :25-33).  With
 * first != 0 the next outer call starts from f0 = 1, k = 1 as the
 * reference's first call does (the other arguments may be NULL). */
int adg_set_state(adg_ctx *ctx, const double *f0, const double *fs0, double Ke,
                  const double *dn, int first);
int adg_get_state(adg_ctx *ctx, double *f0, double *fs0, double *Ke, double *dn, double *df);

/* cmfd::outer (src/mod_cmfd.f90:415-509) for mode ADG_FORWARD,
 * cmfd::outer_ad (:602-699) for ADG_ADJOINT and cmfd::outer_fs (:513-598)
 * for ADG_FIXED_SOURCE (the forward outer with the external source of
 * adg_set_xs and k held at its value: 1 on the first call), cmfd::outer_tr
 * (:800-880) for ADG_TRANSIENT (the external source is the caller's
 * get_exsrc result, passed to adg_set_xs; the fission source is weighted by
 * adg_set_transient's 1 - beta + dfis and not divided by k; the nodal update
 * runs in its transient form, cmode 2; ADG_E_MAX_OUTER is outer_tr's
 * maxi = .TRUE.) and cmfd::outer_ad called with popt = 0 for
 * ADG_ADJOINT_FROZEN (the adjoint outer with the forward D-hats kept, no
 * nodal update, :677-679: rod_eject's adjoint).  trace may be NULL. */
int adg_outer(adg_ctx *ctx, int mode, const adg_iter *iter, adg_trace *trace);

/* The feedback tables of xsec::XS_updt as sdata holds them
 * (src/mod_data.f90), filled by the %XSEC, %CBCS/%BCON, %FTEM, %MTEM, %CDEN
 * and %CROD readers of src/mod_io.f90.  tab[kind][xs]: kind ADG_XS_BASE
 * (x*), ADG_XS_BORON (c*), ADG_XS_FTEM (f*), ADG_XS_MTEM (m*), ADG_XS_CDEN
 * (l*), ADG_XS_ROD (d*); xs 0 sigtr, 1 siga, 2 nuf, 3 sigf as (nmat,G) and
 * 4 sigs as (nmat,G,G) (from, to), column-major.  tab[kind][0] == NULL:
 * that feedback is off (its b* flag is 0; boron is on for bbcon or bcbcs).
 * fbmap (nxx,nyy): the rod bank of each node column, 0 none.  The reference
 * values rbcon, rftem, rmtem, rcden and the rod geometry coreh, pos0, ssize
 * are sdata's. */
enum adg_xs_kind { ADG_XS_BASE = 0, ADG_XS_BORON = 1, ADG_XS_FTEM = 2, ADG_XS_MTEM = 3, ADG_XS_CDEN = 4,
                   ADG_XS_ROD = 5 };
typedef struct adg_xs_tables {
  int nmat, nb;
  double rbcon, rftem, rmtem, rcden, coreh, pos0, ssize;
  const double *tab[6][5];
  const int *fbmap;
} adg_xs_tables;

/* Installs the tables on the device, with the node material map mat(N),
 * the fission spectrum chi(nmat,G) and discontinuity factors dc(N,G,6) (what
 * adg_set_xs takes besides the node cross sections; the external source is
 * set to zero).  The node cross sections are then made by adg_update_xs. */
int adg_set_xs_tables(adg_ctx *ctx, const adg_xs_tables *t, const int *mat, const double *chi,
                      const double *dc);
/* Branch tables of a %XTAB deck (sdata's m(:), src/mod_data.f90:176-192,
 * read by src/mod_io.f90), flattened.  dims (nmat,5) row-major: nd, nb, nf,
 * nm (branch points of coolant density, boron, fuel and moderator
 * temperature) and trod (1: rodded set present) per material; points: per
 * material pd, pb, pf, pm concatenated; xsec: per material its
 * nd*nb*nf*nm branch records (density fastest, then boron, fuel, moderator)
 * of E = 4G + G^2 + 6G doubles: sigtr, siga, nuf, sigf (G), sigs (G,G)
 * (from, to), dc (G,6); rxsec: the same for the rodded sets of the trod = 1
 * materials (NULL without); fbmap (nxx,nyy) or NULL without rods. */
typedef struct adg_xtab_tables {
  int nmat, nb;
  const int *dims;
  const double *points, *xsec, *rxsec;
  double coreh, pos0, ssize;
  const int *fbmap;
} adg_xtab_tables;

/* Installs branch tables instead of feedback tables: adg_update_xs then
 * runs xsec::XStab_updt (src/mod_xsec.f90:48-84: brInterp per node,
 * :520-788, and crod_tab_updt, :291-376, with the same parameters) and
 * also makes the node discontinuity factors dc; ADG_E_XS also where a
 * parameter leaves a branch range by more than brInterp extrapolates or a
 * rodded node's material has no rodded set. */
int adg_set_xtab_tables(adg_ctx *ctx, const adg_xtab_tables *t, const int *mat, const double *chi);

/* xsec::XS_updt (src/mod_xsec.f90:11-46) on the device: the node D, sigr,
 * nuf, sigf and sigs for boron concentration bcon, node fuel temperatures
 * ftem(N), moderator temperatures mtem(N), coolant densities cden(N) and
 * bank positions bpos(nb) (steps), bit for bit the reference's; arrays
 * whose feedback is off may be NULL.  Replaces the host XS_updt call before
 * each outer of the T-H iteration and the boron searches (src/mod_th.f90:48,
 * 774-794; src/mod_control.f90:28-107).  ADG_E_XS where the reference would
 * stop; the context's cross sections are then not to be used. */
int adg_update_xs(adg_ctx *ctx, double bcon, const double *ftem, const double *mtem, const double *cden,
                  const double *bpos);
/* The node cross sections the context holds (after adg_set_xs or
 * adg_update_xs), in node order: D, sigr, nuf, sigf (N,G), sigs (N,G,G)
 * (from, to), dc (N,G,6); NULL skips an array. */
int adg_get_xs(adg_ctx *ctx, double *D, double *sigr, double *nuf, double *sigf, double *sigs, double *dc);

/* cmfd::PowDis (src/mod_cmfd.f90:1290-1333): node powers normalised to 1. */
int adg_powdis(adg_ctx *ctx, double *p);
/* The numbers of the radial and axial power maps the reference prints,
 * io::AsmPow (src/mod_io.f90:3267-3403) and io::AxiPow (:3407-3498), made on
 * the device from the context's PowDis node powers: fasm (nx,ny) the
 * axially averaged assembly powers normalised to average 1 over the
 * assemblies with power, pmax the (i, j) of their maximum; faxi (nz) the
 * plane powers normalised the same way, amax the plane of their maximum
 * (1-based; pmax, amax may be NULL).  nx, ny, nz and the node divisions
 * xdiv, ydiv, zdiv are the %GEOM card's.  Same sums in the same order as
 * the reference: bit for bit its numbers for the same node powers. */
int adg_power_maps(adg_ctx *ctx, int nx, int ny, int nz, const int *xdiv, const int *ydiv, const int *zdiv,
                   double *fasm, int *pmax, double *faxi, int *amax);
/* adg_power_maps without any device allocation, for callers that cannot
 * allocate (code under stream capture, tight serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  Same arguments,
 * results and statuses as adg_power_maps (bit for bit); the map kernels'
 * tables and intermediate sums go to a fixed per-GPU scratch array held
 * until the process ends instead of buffers allocated per call.  Limits:
 * nxx*nyy + nx*ny + nz <= 131072 and 2*nyy + nx + ny + nz + 6 <= 8192;
 * larger layouts return ADG_E_ARG.  Calls on one GPU share that scratch, so
 * calls that overlap in time on one GPU (from several host threads) race. */
int adg_power_maps_noalloc(adg_ctx *ctx, int nx, int ny, int nz, const int *xdiv, const int *ydiv,
                           const int *zdiv, double *fasm, int *pmax, double *faxi, int *amax);
/* The numbers of the radial flux map the reference prints, io::AsmFlux
 * (src/mod_io.f90:3502-3644), made on the device from the context's flux:
 * fasm (nx,ny,G) the axially averaged assembly fluxes, normalised per group
 * to norm (norm <= 0: no normalisation, the reference's call without it);
 * negf (may be NULL) 1 where an assembly flux is negative.  Bit for bit the
 * reference's numbers for the same flux. */
int adg_flux_map(adg_ctx *ctx, int nx, int ny, const int *xdiv, const int *ydiv, double norm, double *fasm,
                 int *negf);
/* The per-node factor (1 - tbeta(mat(n)) + dfis(n)) of TSrcTr
 * (src/mod_cmfd.f90:1068-1096) and of the transient get_B
 * (src/mod_nodal.f90:1379-1391), for ADG_TRANSIENT outers: wtr(N). */
int adg_set_transient(adg_ctx *ctx, const double *wtr);
int adg_set_transient_core(adg_ctx *ctx, int core, const double *wtr);
/* nodal::Lxyz (src/mod_nodal.f90:901-1005) of every node and group of the
 * context's current flux and D-tilde/D-hat: L(N,G,3), the x, y and z zeroth
 * transverse leakages (single-core context). */
int adg_leakage(adg_ctx *ctx, double *L);

/* Kernel-level entry points, each replacing one reference routine. */
/* nodal::nodal_update / nodal_update_pnm (src/mod_nodal.f90:18,151) followed
 * by the matrix rebuild of cmfd::nodal_upd (src/mod_cmfd.f90:339-383);
 * cmode 1 forward, 0 adjoint.  Uses the context's f0, Ke and dn. */
int adg_nodal_update(adg_ctx *ctx, int kern, int cmode, double *ndmax, int *loc3);
/* cmfd::bicg (src/mod_cmfd.f90:1203-1243) on group g (0-based): exactly nin
 * iterations from the initial guess x (in/out, length N). */
int adg_bicg(adg_ctx *ctx, int g, int nin, const double *b, double *x);
/* cmfd::sp_matvec (src/mod_cmfd.f90:1247-1268) on group g (0-based). */
int adg_matvec(adg_ctx *ctx, int g, const double *x, double *y);

/* Timers, the counterpart of the reference's CPU-time buckets fdm_time and
 * nod_time (src/mod_data.f90:204-217, accumulated at src/mod_cmfd.f90:442-489
 * and :356-378), measured with HIP events on the context's stream.
 * adg_last_timing: device time (ms) of the last adg_outer call,
 *   [0] total, [1] group solves (TSrc + BiCGSTAB), [2] nodal updates, [3] rest;
 *   without profiling, [2] is estimated from events around the replays of
 *   the nodal-update outers only (their time beyond the mean of the other
 *   outers), [1] is the rest of the call and [3] zero;
 *   after adg_update_xs, [0] its kernel's time and [1..3] zero.
 * With profiling on (adg_set_profiling(ctx, 1)), adg_outer also brackets the
 * phases and one steady-state BiCGSTAB K1 launch per group solve with events;
 * adg_read_profile returns averages over the last adg_outer call:
 *   [0] K1 kernel ms, [1] K1 samples, [2] BiCGSTAB group solve ms (setup + nin
 *   iterations), [3] samples, [4] nodal update ms, [5] samples. */
int adg_last_timing(adg_ctx *ctx, double *ms4);
/* Launch-shape autotuner (no counterpart in the reference): times one
 * BiCGSTAB group solve of nin iterations on the context's matrices for each
 * candidate shape of the flat kernels (one core: the largest grid, 256 to
 * 2048 blocks; a batch: 256, 512 or 1024 cells per block) and, for a single
 * core, the persistent kernel with each of its brick plans (up to four, the
 * cheapest by the planner's cost model, each checked on the host), and
 * keeps the fastest.
 * The solution state is left as it was; results stay within the parity bars
 * for any shape (only the summation order of the reductions depends on
 * it).  out4: chosen shape (0: the persistent kernel), its ms per group
 * solve, grid size, candidates tried (may be NULL). */
int adg_autotune(adg_ctx *ctx, int nin, double *out4);
int adg_set_profiling(adg_ctx *ctx, int on);
/* How the context's group solves run (no counterpart in the reference):
 * info8[0] 1 when one persistent launch per group solve keeps the
 * BiCGSTAB state on chip for all nin iterations (single-core contexts whose
 * brick plan fits the device; ADG_PERSIST=0 disables it), 0 for the flat
 * two-kernels-per-iteration form; [1] blocks and [2] threads per block of
 * the BiCGSTAB launches; with the persistent form [3..5] the brick extents
 * BX, BY, BZ in nodes, [6] nodes per thread and [7] the largest halo of a
 * brick. */
int adg_solver_info(adg_ctx *ctx, int *info8);
/* The persistent kernel's brick plan for an nxx x nyy x nzz node box on a
 * device with ncu CUs, made and checked on the host (no device is touched):
 * plan8[0] 1 when a plan exists, [1] blocks, [2..4] BX, BY, BZ, [5] nodes
 * per thread, [6] largest halo, [7] 0 when every index the kernel forms is
 * inside its buffers, every node is owned once and every stencil neighbour
 * is on the tile (1 otherwise).  No counterpart in the reference. */
int adg_persistent_plan(int nxx, int nyy, int nzz, int precision, int ncu, int *plan8);
int adg_read_profile(adg_ctx *ctx, double *out6);

/* Batching form.  The reference solves one core per process; these entry
 * points have no counterpart there and serve callers that solve many
 * independent cores at once (perturbed or differently sized cores, same
 * group count and precision form): all cores of a batch advance together
 * through the same kernels, one hipGraph replay per outer iteration covering
 * every core, and each core keeps its own k, schedule, convergence test and
 * status exactly as cmfd::outer has them for one core.
 * adg_create_batch takes an array of ncore geometries; the *_core calls are
 * the single-core calls above for core `core` (0-based).  The single-core
 * calls act on core 0; adg_nodal_update, adg_bicg and adg_matvec need a
 * single-core context. */
int adg_create_batch(const adg_geom *geoms, int ncore, adg_ctx **ctx);
int adg_core_count(adg_ctx *ctx);
int adg_set_xs_core(adg_ctx *ctx, int core, const double *D, const double *sigr, const double *sigs,
                    const double *nuf, const double *sigf, const double *chi, const int *mat,
                    const double *dc, const double *exsrc);
int adg_set_state_core(adg_ctx *ctx, int core, const double *f0, const double *fs0, double Ke,
                       const double *dn, int first);
int adg_get_state_core(adg_ctx *ctx, int core, double *f0, double *fs0, double *Ke, double *dn,
                       double *df);
/* cmfd::outer / outer_ad for every core of the batch.  iter holds niter_in
 * entries: 1 (one %ITER for every core) or ncore (each core's own; nout, nin
 * and kern must agree, nac, nupd, serc and ferc are per core).  Per core:
 * status (enum adg_status), outers performed and k (arrays of ncore, each
 * may be NULL).  Returns the first non-OK core status. */
int adg_outer_batch(adg_ctx *ctx, int mode, const adg_iter *iter, int niter_in, int *status, int *niter,
                    double *keff);
int adg_trace_core(adg_ctx *ctx, int core, adg_trace *trace);
int adg_set_xs_tables_core(adg_ctx *ctx, int core, const adg_xs_tables *t, const int *mat,
                           const double *chi, const double *dc);
int adg_set_xtab_tables_core(adg_ctx *ctx, int core, const adg_xtab_tables *t, const int *mat,
                             const double *chi);
int adg_update_xs_core(adg_ctx *ctx, int core, double bcon, const double *ftem, const double *mtem,
                       const double *cden, const double *bpos);
int adg_get_xs_core(adg_ctx *ctx, int core, double *D, double *sigr, double *nuf, double *sigf,
                    double *sigs, double *dc);
int adg_powdis_core(adg_ctx *ctx, int core, double *p);
int adg_power_maps_core(adg_ctx *ctx, int core, int nx, int ny, int nz, const int *xdiv, const int *ydiv,
                        const int *zdiv, double *fasm, int *pmax, double *faxi, int *amax);
int adg_flux_map_core(adg_ctx *ctx, int core, int nx, int ny, const int *xdiv, const int *ydiv, double norm,
                      double *fasm, int *negf);

/* Differentiation form.  The reference's only adjoint is cmfd::outer_ad's
 * physical adjoint (src/mod_cmfd.f90:602-699: the same matrices, transposed
 * group coupling); these entry points have no counterpart there.  For a
 * single-core context holding a converged forward solution (adg_outer):
 * adg_adjoint solves the discrete adjoint M^T psi = (1/k) F^T psi of the
 * converged CMFD operator (transposed 7-point matrices, D-hat held at its
 * converged value, no nodal update) with the given %ITER, returns its
 * eigenvalue (equal to k) and leaves the forward state untouched.
 * adg_backward returns, per node, the gradient of J = dJdk * k + (response
 * of the PowDis node powers with gradient dJdp, which may be NULL) with
 * respect to the node's D_g, sigma_a_g, nu-sigma_f_g (sigma_f_g keeping its
 * ratio to nu-sigma_f_g) and sigma_s(g->h), g != h, in that order: grad is
 * (N, 3G + G(G-1)) column-major.  The node-power response needs the
 * generalized adjoint, a fixed-source adjoint outer with k fixed (the
 * reference's outer_fs, src/mod_cmfd.f90:513-598, in adjoint form) whose
 * fundamental-mode component is projected out after every outer; iter
 * controls it (serc, ferc bound the relative 2-norm change per outer).
 * D-hat is held fixed: these are the gradients of the converged CMFD
 * eigenproblem, not of the nodal D-hat update. */
int adg_adjoint(adg_ctx *ctx, const adg_iter *iter, double *kadj);
int adg_get_adjoint(adg_ctx *ctx, double *psi, double *psifphi);
/* Install an adjoint flux psi(N,G) of the caller's (for instance the
 * reference's physical adjoint from adg_outer in ADG_ADJOINT mode), so that
 * adg_backward with dJdk = 1 and no dJdp gives the first-order perturbation
 * estimate of dk with it; the context must hold the forward flux. */
int adg_set_adjoint(adg_ctx *ctx, const double *psi);
int adg_backward(adg_ctx *ctx, const adg_iter *iter, double dJdk, const double *dJdp, double *grad);
/* Transpose of one nodal D-hat update (nodal::nodal_update / _pnm,
 * src/mod_nodal.f90:18-147 with 282-371 and 483-1451, forward static form)
 * at the context's state (flux, k, D-hat as adg_set_state or adg_outer
 * left them, cross sections): for a cotangent w(N, G, 6) of the new D-hat
 * (laid out as adg_get_state's dn, one value per face side), the
 * cotangents of the flux phib(N, G), of k (kb), of the old D-hat dnb(N, G,
 * 6) and of the node parameters grad (N, 3G + G(G-1), adg_backward's
 * columns; D through the coupling coefficients as well).  Any output may
 * be NULL.  The state is left as it was. */
int adg_nodal_vjp(adg_ctx *ctx, int kern, const double *w, double *phib, double *kb, double *dnb, double *grad);
/* Gradient of the full nodal model (adg_backward's response and columns):
 * for a context holding a converged solution of the outer iteration with
 * nodal updates (a fixed point of both the CMFD eigenproblem and the D-hat
 * update, kern as solved) and its adjoint (adg_adjoint), the D-hat
 * cotangent w solves w = dJ/d(D-hat) + A^T w, A the Jacobian of "D-hat ->
 * CMFD eigenpair -> nodal update", by fixed-point iteration: each step
 * applies adg_nodal_vjp's transpose and a few outers of the generalized
 * adjoint, continued from the previous step (16; environment ADG_ADJ_INNER,
 * 0 = to iter's tolerance every step).  Stops when the relative 2-norm
 * change of w is at most tol and the last outers met iter's tolerance, or
 * after maxit steps (then ADG_E_MAX_OUTER, grad still filled); nit and
 * resid (may be NULL) return the steps and the last change. */
int adg_backward_nodal(adg_ctx *ctx, const adg_iter *iter, int kern, double dJdk, const double *dJdp, double tol,
                       int maxit, double *grad, int *nit, double *resid);

const char *adg_status_string(int status);
const char *adg_last_error(adg_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif
