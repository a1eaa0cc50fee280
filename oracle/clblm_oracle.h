/*
 * CLBLM oracle: a sequential CPU restatement of the reference's layer optical-depth
 * path (ODLAY -> OPDPTH -> LINF4Q/SHRINQ/LBLF4Q -> HIRACQ/LNCORQ/CNVFNQ/panelOut).
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only to check results.
 * The product (clblm_amd/) never links, loads or calls it.
 *
 * All arithmetic is IEEE double with no FMA contraction (build with
 * -ffp-contract=off), matching the reference's -fdefault-real-8 build, whose
 * default REALs are 8 bytes.
 */
#ifndef CLBLM_ORACLE_H
#define CLBLM_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* TAPE3-equivalent line list (one record = 250 consecutive lines). */
typedef struct {
    int64_t n;
    const double *nu;      /* line centre, cm-1, sorted ascending over the lines */
    const float *s;        /* intensity as stored in TAPE3 */
    const float *a_air;    /* air-broadened HWHM at 296 K, 1 atm */
    const float *e_lo;     /* lower-state energy, cm-1 */
    const int32_t *mol;    /* m + 100*iso */
    const float *a_self;   /* self-broadened HWHM */
    const float *tmpalf;   /* TAPE3 temperature exponent of the width */
    const float *shift;    /* pressure shift, cm-1/atm */
    const int32_t *iflg;   /* 0 plain line; 1 / 3 coupled (coefficients in the next
                              record); 2 plain; < 0 not a line (coefficients, skipped) */
    /* species broadening BRD_MOL_FLG and BRD_MOL_DAT (widths, exponents, shifts) of
     * broadener j = 1..7 at 7 i + j - 1; NULL when the list has none */
    const int32_t *brd_flg;
    const float *brd_hw, *brd_tmp, *brd_shft;
} orc_lines;

typedef struct {
    double pave, tave, wtot; /* hPa, K, molec/cm2 */
    double w[48];            /* column of molecule m at w[m], m = 1..47 */
    int sel[48];             /* nonzero if molecule m is a line molecule of this layer */
} orc_layer;

typedef struct {
    double v1, v2;           /* spectral range, cm-1 */
    double dvnormal, dvnarrow, sample;
    double dvout;            /* <= 0: native grid */
    int jrad;                /* state in/out (reference SAVE variable) */
    int contnm;              /* nonzero: absrb[] holds the continuum on the OPDPTH 1 cm-1 grid */
    const double *absrb;
    int nptabs;
    int species_broad;       /* odCtrl%speciesBroad */
} orc_params;

/* Builds the line-shape tables (SHAPEL/SHAPEG/VERFN).  Arrays are 1-based
 * (element 0 unused) with 3002 slots. */
void orc_tables(double *f1, double *f2, double *f3, double *fg, double *xver);

/* Layer optical depth.  out may be NULL to query the number of points.
 * Returns 0 on success, nonzero on error (message via orc_error()). */
int orc_odlay(const orc_lines *lines, const orc_layer *layer, orc_params *par,
              double *out, int64_t cap, int64_t *nsamp, double *dv_used);

const char *orc_error(void);

/* A direction of differentiation: the layer temperature moves by dt (K), the
 * broadening column by the factor (1 + dlnwtot), each column W(m) by (1 + dlnw[m]). */
typedef struct {
    double dt, dlnwtot;
    double dlnw[48];
} orc_dir;

/* Layer OD of the smooth model (the reference's ODLAY with CNVFNQ's shape tables
 * replaced by their closed forms) and its derivatives along ndir (<= 16) directions.
 * out holds 1 + ndir channels of cap points each (channel 0 the OD); dabsrb, when the
 * continuum is on, holds ndir arrays of nptabs derivatives of ABSRB.  Species
 * broadening is not differentiated (error). */
int orc_odlay_tan(const orc_lines *lines, const orc_layer *layer, orc_params *par, int ndir, const orc_dir *dirs,
                  const double *dabsrb, double *out, int64_t cap, int64_t *nsamp, double *dv_used);
double orc_dradfn_dt(double v, double t);
double orc_tips_dlogq(int mol, int iso, double t);

/* Helpers exposed for unit tests. */
double orc_radfn(double vi, double xkt);
double orc_radfni(double vi, double dvi, double xkt, double *vinew, double *rdel, double *rdlast);
int orc_tips_scor(int mol, int iso, double t, double *scor);
void orc_xint(double v1a, double v2a, double dva, const double *a1, double afact,
              double vft, double dvr3, double *r3_1, int n1r3, int n2r3);
void orc_interp(const double *spin, double dvin, double v1in, int nin,
                double *spout, double dvout, double v1out, int nout);
double orc_calculate_dv(double vbar, double pave, double tave, double frh2o,
                        double alfal0, double sample);
double orc_avrat(int iz);
int orc_molec(const orc_layer *layer, double *scor /*[48*11]*/, double *rhoslf /*[48]*/,
              double *alfd1 /*[48*11]*/);

/* Radiative transfer and Jacobians (upwelling nadir, merge-up). */
double orc_planck(double v, double t);
void orc_planck_drv(double v, double t, double *b, double *dbdt);
void orc_rlin(double tau, double *w, double *dw);
void orc_layer_emis(int64_t npts, const double *od, double v1, double dv, double ta, double tb, double tave,
                    int lin, double *out);
void orc_rt_down(int nlay, int64_t npts, double v1, double dv, const double *od, const double *t_lay,
                 const double *t_lev, int lin, const double *sol, double *rad, double *tx);
int orc_interp_to_finest(int nspec, double v1, const double *dv, const int64_t *nlim, const double *const *in,
                         double *const *out);
int64_t orc_rt_up_grids(int nlay, double v1, const double *dv, const int64_t *nlim, const int64_t *off,
                        const double *od, const double *t_lay, const double *t_lev, double emis, double tskin,
                        double surf_dv, int64_t surf_nlim, int lin, double sec_dn, double *rad, int64_t cap,
                        double *dv_out);
void orc_rt_up(int nlay, int64_t npts, double v1, double dv, const double *od, const double *t_lay,
               const double *t_lev, double emis, double tskin, int lin, double sec_dn, double *rad);

/* Level Jacobians: level2layer, the per-point Jacobians and drv_LevelTemp/drv_LevelConc. */
void orc_level2layer(int nlay, int nmol, const double *plev, const double *tlev, const double *tave,
                     const double *qlev, double *dt_dtup, double *dt_dtlo, double *dw_dqup, double *dw_dqlo);
void orc_jacobians(int nlay, int64_t npts, double v1, double dv, const double *od, const double *t_lay,
                   const double *t_lev, double emis, double tskin, int lin, double sec_dn, const double *dt_dtup,
                   const double *dt_dtlo, double *rad, double *dr_dtau, double *dr_dtlev, double *dr_dtskin,
                   double *dr_demis);
void orc_jacobians_paths(int nlay, int64_t npts, double v1, double dv, const double *od, const double *amf_view,
                         const double *amf_down, const double *amf_sun, const double *t_lay, const double *t_lev,
                         double emis, double tskin, int solar, const double *sol_toa, double rsfc_sun, int obs,
                         int uplook, int lin, const double *dt_dtup, const double *dt_dtlo, double *rad_o,
                         double *dr_dtau, double *dr_dtlev, double *dr_dtskin, double *dr_demis, double *dr_drsfc);
void orc_level_jac(int nlay, int nmol, int64_t npts, const double *od_var, const double *dr_dtau,
                   const double *dr_dtlev, const double *wtot, const double *tave, double dt, double dw_frac,
                   const double *dt_dtup, const double *dt_dtlo, const double *dw_dqup, const double *dw_dqlo,
                   double *jac_t, double *jac_q);

/* Instrument scan (FFT), see clblm_oracle.c. */
typedef struct {
    int jfn;
    double hwhm, a;            /* A = C(JFN)*HWHM */
    int pre_interp;            /* input re-gridded to dvn before the transform */
    double dvn;
    double v1z, v2z, dv;       /* scan grid (after the optional re-gridding) */
    int64_t n1load, ltotal;    /* first point (1-based) of the scan-grid spectrum loaded, count */
    int64_t lrec, nfft;        /* records of LPTSMX and transform length */
    int64_t n1, ntotal;        /* wrtSpect: first point (1-based) and count on the scan grid */
    double v1s;
    int ivx;                   /* 1 analytic apodization, -1 transform of the scanning function, 0 none */
    int fft;                   /* 0: no transform (boxcar scan without deconvolution) */
    int prebox, m;             /* boxcar pre-scan over m points (m >= 2) */
    int64_t ltotal0, lrec0;    /* points and records as loaded, before the boxcar */
    double v1z0, dv0;          /* scan grid as loaded, before the boxcar */
    int deconv;                /* divide by the boxcar's transform, width parameter deconv_a */
    double deconv_a;
    int deconv_only;
} orc_scan_plan;
int64_t orc_scan_interp(const double *x, int64_t n, double v1, double dv, double vv1, double vv2, double dvo,
                        double *out, int64_t cap, double *v1res);
int orc_mtckd_h2o(const double *wvn, const double *self_ref, const double *for_ref, const double *texp, int ncoeffin,
                  double ref_temp, double ref_press, double pave, double tave, double wk1, double wtot, double v1,
                  double v2, int jrad, double xself, double xfrgn, double *absrb, int cap);
int orc_rt_flux(int nlay, int64_t npts, double v1, double v2, double dv, const double *od, const double *t_lay,
                const double *t_lev, const double *p_lev, int lin, double emis, double tskin, int nang,
                double dv_flux, int lambertian, double *up, double *dn, double *net, double *htr);
int orc_contnm_ir(double pave, double tave, double wtot, const double *wk, double v1, double v2, int jrad,
                  const double *scale, double v1abs, int nptabs, double *absrb);
int64_t orc_rt_mode1(int nlay, double v1, double v2, const double *dv, const int64_t *nlim, const int64_t *off,
                     const double *od_up, const double *od_dn, const double *od_sun, const double *fac_up,
                     const double *fac_dn, const double *fac_sun, const double *t_lay, const double *t_lev,
                     int obs_level, int lin, int thermal, int solar, double emis, double tskin, double refl,
                     const double *sol, int64_t sol_n, double sol_dv, double *rad, double *tx, int64_t cap,
                     double *dv_out);
int orc_fltrfn(const double *x, int64_t n, double v1, double dv, double v1f, double dvf, int nptf, const double *xf,
               double *out);
void orc_scan_shape(int jfn, double *xf);
int64_t orc_scanfn_conv(const double *x, int64_t n, double v1, double dv, double v1out, double v2out, double dvout,
                        int funct_id, double hwhm, double *out, int64_t cap, double *v1res, double *dvres);
int64_t orc_scanfn_rect(const double *x, int64_t n, double v1, double dv, double v1out, double v2out, double dvout,
                        double hwhm, double *out, int64_t cap, double *v1res, double *dvres);
int orc_fftscn_plan2(double v1c, double dv, int64_t n, double v1out, double v2out, int funct_id, double hwhm,
                     double fill, double pre_box_hw, int deconv_pre_box, orc_scan_plan *p);
void orc_scan_boxcar(const orc_scan_plan *p, const double *loaded, double fill, double *out);
void orc_fft_deconv(const orc_scan_plan *p, double *h);
void orc_scan_sfn(const orc_scan_plan *p, double *f);
int orc_fftscn_plan(double v1c, double dv, int64_t n, double v1out, double v2out, int funct_id, double hwhm,
                    orc_scan_plan *p);
int orc_fftscn_finish(orc_scan_plan *p, double v1p, int64_t np, double v1out, double v2out);
void orc_fft_apod(const orc_scan_plan *p, double *h);

#ifdef __cplusplus
}
#endif
#endif
