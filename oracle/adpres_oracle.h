/*
 * CPU restatement of ADPRES's static CMFD + SANM/PNM hot path, in FP64.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: tests, the smoke
 * entry and the benchmark's CPU-baseline leg call it; the product library
 * never links it.
 *
 * Every routine follows the reference's loop order and operation order
 * (src/mod_cmfd.f90, src/mod_nodal.f90) so that, compiled without FP
 * contraction, it reproduces the reference's FP64 results bit for bit on
 * the same inputs.  Arrays are Fortran column-major with 1-based indices
 * stored in the integer arrays, exactly as the reference holds them in its
 * `sdata` module (src/mod_data.f90).
 */
#ifndef ADPRES_ORACLE_H
#define ADPRES_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
  ORC_OK = 0,
  ORC_E_MAX_OUTER = 1,      /* src/mod_cmfd.f90:498-505 */
  ORC_E_NODAL_UNSTABLE = 2, /* src/mod_nodal.f90:131-142 */
  ORC_E_LU_PIVOT = 3,       /* src/mod_nodal.f90:856-862 */
  ORC_E_POWER_NONPOS = 4,   /* src/mod_cmfd.f90:1322-1326 */
  ORC_E_ARG = 5,            /* bad argument / unsupported geometry */
  ORC_E_XS = 6,             /* src/mod_xsec.f90:215,86-150: the reference STOPs */
};

typedef struct {
  int nnod, ng, nxx, nyy, nzz, nmat;
  const int *ix, *iy, *iz, *mat;          /* (N), 1-based */
  const int *ystag_min, *ystag_max;       /* (nyy) */
  const int *xstag_min, *xstag_max;       /* (nxx) */
  const double *xdel, *ydel, *zdel;
  int bc[6];                              /* east west north south bottom top */
  const double *D, *sigr, *nuf, *sigf;    /* (N,G) */
  const double *sigs;                     /* (N,G,G): sigs(n, from, to) */
  const double *chi;                      /* (nmat,G) */
  const double *dc;                       /* (N,G,6) */
  const double *exsrc;                    /* (N,G) */
  const double *wtr;                      /* (N) transient weight 1 - tbeta + dfis (mode 3, cmode 2) */
} orc_problem;

typedef struct {
  int nout, nin, nac, nupd;
  int kern;                               /* 0 FDM, 1 PNM, 2 SANM */
  double serc, ferc;
} orc_iter;

typedef struct {
  double Ke;
  double *f0;      /* (N,G) */
  double *fs0;     /* (N) */
  double *dn;      /* (N,G,6) nodal (D-hat) coupling coefficients */
  double *df;      /* (N,G,6) FDM (D-tilde) coupling coefficients (output) */
  double ndmax;
  int im, jm, km;
  int first;       /* 1: initialise f0 = 1, Ke = 1, fs0 from f0 */
} orc_state;

typedef struct {
  int cap;         /* capacity of the arrays below */
  int n;           /* outers performed */
  double *k, *ser, *fer, *ndmax;
  int *flags;      /* bit0 extrapolated, bit1 nodal update */
  int *loc;        /* 3 per outer: im jm km after a nodal update */
} orc_trace;

/* Whole static solve: mode 0 outer (src/mod_cmfd.f90:415), 1 outer_ad
   (:602), 2 outer_fs (:513, fixed source: k is not updated), 3 outer_tr
   (:800, one transient step: TSrcTr source with pb->wtr, no k update,
   nodal update in cmode 2; exsrc and wtr are the step's get_exsrc
   result), 4 outer_ad called with popt = 0 (:677-679: no nodal update,
   the D-hats stay those of the forward state). */
int orc_outer(const orc_problem *pb, orc_state *st, const orc_iter *it,
              int mode, orc_trace *tr);

/* The feedback tables XS_updt reads (sdata, src/mod_data.f90; filled by
   src/mod_io.f90's %XSEC, %CBCS/%BCON, %FTEM, %MTEM, %CDEN and %CROD
   readers).  tab[kind][xs]: kind 0 base x*, 1 boron c*, 2 fuel temperature
   f*, 3 moderator temperature m*, 4 coolant density l*, 5 control rod d*;
   xs 0 sigtr, 1 siga, 2 nuf, 3 sigf as (nmat,G), 4 sigs as (nmat,G,G)
   (from, to).  tab[kind][0] == NULL: that feedback is off (its b* flag 0).
   fbmap (nxx,nyy): the rod bank of each node column, 0 none. */
typedef struct {
  int nmat, nb;
  double rbcon, rftem, rmtem, rcden, coreh, pos0, ssize;
  const double *tab[6][5];
  const int *fbmap;
} orc_xs_tables;

/* xsec::XS_updt (src/mod_xsec.f90:11-46): base_updt, bcon/ftem/mtem/cden
   updates, crod_updt, Dsigr_updt and check_xs, at boron bcon, node fuel
   and moderator temperatures and coolant densities (N) and bank positions
   bpos (nb, steps).  Arrays the tables leave off may be NULL.  Outputs are
   (N,G) and sigs (N,G,G).  ORC_E_XS where the reference would STOP. */
int orc_xs_update(const orc_problem *pb, const orc_xs_tables *t, double bcon, const double *ftem,
                  const double *mtem, const double *cden, const double *bpos, double *sigtr,
                  double *siga, double *nuf, double *sigf, double *sigs, double *D, double *sigr);

/* Branch tables of a %XTAB deck (sdata's m(:), src/mod_data.f90:176-192,
   read by src/mod_io.f90), flattened.  dims (nmat,5) row-major: nd, nb, nf,
   nm (points of coolant density, boron, fuel and moderator temperature)
   and trod (1: rodded set present) per material; points: per material
   pd, pb, pf, pm concatenated; xsec: per material its nd*nb*nf*nm branch
   records (s fastest, then t, u, v) of E = 4G + G^2 + 6G doubles: sigtr,
   siga, nuf, sigf (G), sigs (G,G) (from, to), dc (G,6); rxsec: the same for
   the rodded sets of the trod = 1 materials. */
typedef struct {
  int nmat, nb;
  const int *dims;
  const double *points, *xsec, *rxsec;
  double coreh, pos0, ssize;
  const int *fbmap;                       /* (nxx,nyy) or NULL: no rods */
} orc_xtab_tables;

/* xsec::XStab_updt (src/mod_xsec.f90:48-84): brInterp (:520-788) per
   node, crod_tab_updt (:291-376) with the same parameters, Dsigr_updt and
   check_xs.  Outputs as orc_xs_update plus dc (N,G,6).  ORC_E_XS where the
   reference would STOP (also a parameter outside a branch range, or a
   rodded node of a material without rodded data). */
int orc_xtab_update(const orc_problem *pb, const orc_xtab_tables *t, double bcon, const double *ftem,
                    const double *mtem, const double *cden, const double *bpos, double *sigtr,
                    double *siga, double *nuf, double *sigf, double *sigs, double *D, double *sigr,
                    double *dc);

/* The numbers of the radial and axial power maps, io::AsmPow
   (src/mod_io.f90:3267-3403) and io::AxiPow (:3407-3498), from node powers
   pow(N): assembly powers fasm (nx,ny) normalised to average 1 over the
   assemblies with power, the position of the maximum pmax (i, j, 1-based),
   plane powers faxi (nz) normalised the same way and amax.  nx, ny, nz and
   xdiv, ydiv, zdiv are the %GEOM card's assemblies, planes and their node
   divisions. */
int orc_power_maps(const orc_problem *pb, int nx, int ny, int nz, const int *xdiv, const int *ydiv,
                   const int *zdiv, const double *pow, double *fasm, int *pmax, double *faxi, int *amax);

/* The numbers of the radial flux map, io::AsmFlux (src/mod_io.f90:
   3502-3644), of the flux f0 (N,G): fasm (nx,ny,G) axially averaged
   assembly fluxes, normalised per group to norm (a norm <= 0: absent, as
   in the adjoint call); negf = 1 where an assembly flux is negative. */
int orc_flux_map(const orc_problem *pb, int nx, int ny, const int *xdiv, const int *ydiv, const double *f0,
                 double norm, double *fasm, int *negf);

/* Lxyz (src/mod_nodal.f90:901-1005) of the state: L(N,G,3). */
int orc_leakage(const orc_problem *pb, orc_state *st, double *L);

/* coup_coef (src/mod_cmfd.f90:11): fills st->df. */
void orc_coup_coef(const orc_problem *pb, orc_state *st);

/* One nodal update (src/mod_nodal.f90:18 SANM, :151 PNM), cmode 1 forward
 * or 0 adjoint.  Reads st->f0, st->Ke, st->df, st->dn; writes st->dn,
 * st->ndmax (reset to 0 first) and the location. */
int orc_nodal_update(const orc_problem *pb, orc_state *st, int kern, int cmode);

/* nin BiCGSTAB iterations for group g (0-based) on the CMFD matrix built
 * from st->df and st->dn (src/mod_cmfd.f90:1203).  x is in/out (N). */
int orc_bicg(const orc_problem *pb, const orc_state *st, int g, int nin,
             const double *b, double *x);

/* y = A_g x with the CMFD matrix (src/mod_cmfd.f90:1247). */
int orc_matvec(const orc_problem *pb, const orc_state *st, int g,
               const double *x, double *y);

/* SANM coefficients A,B,E,F,G,H of one node and group
 * (get_ABEFGH, src/mod_nodal.f90:1409-1451). */
void orc_sanm_coef(double sigr, double D, double h, double *out6);

/* PowDis (src/mod_cmfd.f90:1290). */
int orc_powdis(const orc_problem *pb, const double *f0, double *p);

#ifdef __cplusplus
}
#endif
#endif
