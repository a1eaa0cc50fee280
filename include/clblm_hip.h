/*
 * clblm_hip.h -- C-ABI of the MI355X layer optical-depth library.
 *
 * The library replaces the reference's per-layer OD entry point and the line
 * file it reads.  Every entry point returns 0 on success and a nonzero status on
 * error; clblm_last_error() returns the message of the last failure on the
 * calling thread.  Plain pointers and sizes only; no torch types.
 *
 * Molecules 1-7 (H2O, CO2, O3, N2O, CO, CH4, O2) and all their isotopologues
 * are supported (TIPS data and masses are built in).
 */
#ifndef CLBLM_HIP_H
#define CLBLM_HIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CLBLM_OK 0
#define CLBLM_ERR_ARG 1
#define CLBLM_ERR_UNSUPPORTED 2
#define CLBLM_ERR_HIP 3
#define CLBLM_ERR_CAPACITY 4

#define CLBLM_FP64 0         /* plain path: FP64 everywhere, 1e-10 of the oracle */
#define CLBLM_FP32C 1        /* FP64 offsets/intensities, FP32 Kahan line-shape sums */

#define CLBLM_MAXMOL 47

typedef struct clblm_ctx clblm_ctx;

/* One atmospheric layer: the arguments Pav, Tav, W(:), Wtot, Zbot, Ztop and the
 * molecule list of ODLAY_nonStruct (src/clblm_src/clblm_ODLAY.f90:129-176). */
typedef struct {
    double pave;                 /* hPa */
    double tave;                 /* K */
    double wtot;                 /* broadening column, molec/cm2 */
    double zbot, ztop;           /* km (informational) */
    double w[CLBLM_MAXMOL + 1];  /* w[m] = column of HITRAN molecule m, molec/cm2 */
    int32_t sel[CLBLM_MAXMOL + 1]; /* sel[m] != 0: molecule m is a line molecule of the layer */
} clblm_layer;

/* Spectral grid and control record: V1, V2, DVnormal, DVnarrow, SAMPLE, DVout
 * of ODLAY_nonStruct plus the CLBLM_OD_Ctrl fields this path honours
 * (src/clblm_src/clblm_Config.f90:84-118). */
typedef struct {
    double v1, v2;               /* cm-1 */
    double dvnormal, dvnarrow;   /* per-layer DV (setDV) */
    double sample;               /* SAMPLE, 4 by default */
    double dvout;                /* > 0: re-grid the OD to this DV (interpSpectrum); <= 0: native */
    int32_t precision;           /* CLBLM_FP64 or CLBLM_FP32C */
    int32_t contnm;              /* nonzero: continuum[] supplied on the 1 cm-1 ABSRB grid */
    const double *continuum;     /* host pointer, nptabs values from V1ABS = INT(V1)-3 */
    int32_t nptabs;
    int32_t flags;               /* CLBLM_OD_SPECIES_BROAD: odCtrl%speciesBroad */
} clblm_od_params;
/* Species-by-species broadening and shifts (clblm_Config.f90:89, used by LNCORQ
 * clblm_ODLAY.f90:1853-1883 and LINF4Q clblm_ODLAY_LineF4.f90:996-1054) from the data
 * attached by clblm_linelist_broadening. */
#define CLBLM_OD_SPECIES_BROAD 1
/* The MT_CKD water-vapour continuum (self and foreign, CONTNM's H2O part,
 * clblm_ODLAY_Continuum.f90:356-404 with mt_ckd_h2o_absco, mt_ckd_h2o_module.f90:27-103)
 * from the table given to clblm_mtckd_load, added to ABSRB (after the caller's
 * continuum[] when contnm is also set) and carried into the OD as the reference's
 * CONTNM output is (XINT into R3). */
#define CLBLM_OD_MTCKD 2
/* CONTNM's continua other than water vapour (clblm_ODLAY_Continuum.f90:408-1060): CO2
 * (FRNCO2 with the XFACCO2 scaling); O3 (Chappuis and Wulf, Hartley-Huggins with its
 * temperature terms, Hartley-Huggins in the UV); O2 (collision-induced fundamental,
 * 1.27 um, 1.06 um, A-band, visible, Herzberg and far-UV continua); N2 (collision-induced
 * pure rotation, fundamental and first overtone); Rayleigh extinction; scaled by the
 * context's XCO2C, XO3CN, XO2CN, XN2CN, XRAYL (clblm_contnm_scales) and added to ABSRB
 * after the water vapour in the reference's order.  A layer whose ABSRB is longer than
 * the reference's 5050 points (N_ABSRB) returns CLBLM_ERR_ARG. */
#define CLBLM_OD_CONTNM 4

/* ---- context ---------------------------------------------------------------- */
int clblm_ctx_create(int device, clblm_ctx **out);
int clblm_ctx_destroy(clblm_ctx *ctx);
const char *clblm_last_error(void);
int clblm_version(void);

/* ---- line list --------------------------------------------------------------
 * Replaces the TAPE3 reads of RDLIN (clblm_ODLAY.f90:2571-2798) and RDLNFL
 * (clblm_ODLAY_LineF4.f90:1355-1498).  Host arrays of n records in TAPE3 order;
 * consecutive groups of 250 form the TAPE3 records.  mol = m + 100*iso.
 * iflg: 0 for a plain line; 1 for a line-coupled line (Y and G) and 3 for one with
 * reduced widths (GAMMA1, GAMMA2), each followed in the same 250-record by its
 * coefficient record; 2 for a line treated as plain (LNCORQ reads no coefficients
 * for it); < 0 for a record that is not a line (a coefficient record, or one the
 * readers skip).  A coefficient record holds the coefficients at 200, 250, 296 and
 * 340 K: Y / GAMMA1 in its nu, a_air, mol (the bits of a float) and tmpalf fields,
 * G / GAMMA2 in its s, e_lo, a_self and shift fields (clblm_ODLAY.f90:1821-1829).
 * The centres of the lines (iflg >= 0) must be sorted; those of the other records
 * are free.  Returns CLBLM_ERR_ARG for unsorted lines or a coupled line without its
 * coefficient record. */
int clblm_linelist_upload(clblm_ctx *ctx, int64_t n, const double *nu, const float *s,
                          const float *a_air, const float *a_self, const float *e_lo,
                          const float *tmpalf, const float *shift, const int32_t *mol,
                          const int32_t *iflg, int *handle);
/* Frees a line list once the batches that read it are done; CLBLM_ERR_ARG while a
 * captured graph (clblm_od_graph_create) holds it. */
int clblm_linelist_free(clblm_ctx *ctx, int handle);
/* ---- TAPE3 line files ---------------------------------------------------------
 * The binary line file LNFL writes, read as CLBLM_LineData_init
 * (clblm_LineData.f90:99-256), RDLIN (clblm_ODLAY.f90:2571-2798) and RDLNFL
 * (clblm_ODLAY_LineF4.f90:1355-1498) read it: the header record(s), then per panel
 * its header (VMIN, VMAX, NREC, NWDS) and its 250-line block (clblm_LineData.f90:52-70),
 * up to the end of the file or BUFIN_sgl's -99 end mark.  The panels overlapping
 * [vlo, vhi] are read (every panel when vlo > vhi), and the first panel beyond vhi;
 * each becomes a 250-entry record, padded with entries that are not lines (IFLG -1)
 * when it holds fewer.  O2 widths and shifts are corrected on use, as RDLIN does.
 *
 * clblm_tape3_read: host arrays in clblm_linelist_upload's layout plus the
 * species-broadening fields (n x 7 each).  Call with nu == NULL to get *n, then with
 * arrays of *n records (*n x 7 for the broadening fields; any array may be NULL).
 * clblm_linelist_load_tape3: reads and uploads in one call (the broadening fields too
 * when any flag is set); *n (if non-NULL) receives the record count. */
int clblm_tape3_read(const char *path, double vlo, double vhi, int64_t *n, double *nu, float *s,
                     float *a_air, float *a_self, float *e_lo, float *tmpalf, float *shift,
                     int32_t *mol, int32_t *iflg, int32_t *brd_flg, float *brd_hw, float *brd_tmp,
                     float *brd_shft);
int clblm_linelist_load_tape3(clblm_ctx *ctx, const char *path, double vlo, double vhi, int *handle,
                              int64_t *n);

/* The species-broadening fields of the TAPE3 records (BRD_MOL_FLG, and BRD_MOL_DAT
 * split into widths, temperature exponents and shifts: clblm_LineData.f90:67-69,
 * RDLIN clblm_ODLAY.f90:2697-2706) for the n records of a list already uploaded:
 * arrays of n x 7 values, record-major (entry 7 i + j - 1 for broadener j = 1..7).
 * RDLIN's read-time correction of an O2 line's shift by its own species data
 * (:2730-2738) applies from then on; the broadening itself applies to batches whose
 * parameters set CLBLM_OD_SPECIES_BROAD.  CLBLM_ERR_ARG while a captured graph holds
 * the list (the graph reads the fields in place). */
int clblm_linelist_broadening(clblm_ctx *ctx, int handle, const int32_t *flg, const float *hw,
                              const float *tmp, const float *shft);

/* ---- one layer, host buffers -------------------------------------------------
 * Drop-in for ODLAY_nonStruct (clblm_ODLAY.f90:129-611).  Two-phase sizing:
 * call with od_out == NULL to get *nsamp, then with a buffer of >= *nsamp
 * doubles.  *jrad_state carries the reference's SAVEd JRAD across calls
 * (clblm_ODLAY.f90:182-183,349-352; OPDPTH :877-878): set it to the control
 * record's JRAD (1) before the first call of a run and pass it back unchanged.
 * *dv_used returns the DV of the returned grid (the DVout argument). */
int clblm_odlay(clblm_ctx *ctx, int handle, const clblm_layer *layer, const clblm_od_params *par,
                int32_t *jrad_state, int64_t *nsamp, double *dv_used, double *od_out);

/* ---- many layers, device buffers, asynchronous --------------------------------
 * nlay layers (layers[], pars[] host arrays) in one set of launches on `stream`
 * (a hipStream_t; NULL = the context's stream).  Query with d_od == NULL:
 * offsets[k] receives the first index of layer k in d_od and offsets[nlay] the
 * total.  Then d_od (device memory, offsets[nlay] doubles) is filled.  jrad is
 * applied to all layers with the reference's first-call rule (layers are treated
 * as successive calls, in order).  A query followed by the fill call with the same
 * inputs plans the batch once.
 * The fill call returns once the work is enqueued: d_od is ready in stream order,
 * and the host arrays may be reused at once.  Argument errors are returned at
 * once; errors found on the device (line coupling records, capacity) are returned
 * by the next clblm_od_batch, clblm_odlay or clblm_sync on the context.  Batches
 * of one context run in call order, also across streams. */
int clblm_od_batch(clblm_ctx *ctx, int handle, int nlay, const clblm_layer *layers,
                   const clblm_od_params *pars, int32_t *jrad_state, int64_t *offsets,
                   double *d_od, void *stream);

/* ---- layer ODs with their derivatives (the adjoint leg) ------------------------
 * No single counterpart in the reference: its Jacobians take dTau/dT and dTau/dW
 * from ODLAY re-runs at Tave + 1 K, Wtot + 1 % and each molecule alone
 * (drv_LevelTemp, src/clblm_src/clblm_noScattJac.f90:1005-1034; drv_LevelConc
 * :1156-1171).  This entry point returns, in one batch, the layer ODs and their
 * exact derivatives along up to 16 directions.  The model differentiated is
 * ODLAY's (clblm_ODLAY.f90:757-2545, clblm_ODLAY_LineF4.f90:25-1647) with CNVFNQ's
 * shape tables replaced by the closed forms they sample (SHAPEL/SHAPEG/VERFN,
 * clblm_ODLAY.f90:2804-3016); its OD differs from clblm_od_batch's FP64 OD by the
 * tables' spacing (below 1e-3 of the peak).
 * A direction moves the layer temperature by dt (K), the broadening column Wtot by
 * the factor (1 + dlnwtot) and each column W(m) by (1 + dlnw[m]), to first order.
 * Arguments as clblm_od_batch (FP64 only; no CLBLM_OD_MTCKD / CLBLM_OD_CONTNM flags
 * and no species broadening: pass the continuum as pars[l].continuum and its
 * derivatives in dcontinuum, ndir arrays of nptabs values per layer that has a
 * continuum, in layer order; NULL = zero).  Query with d_out == NULL (offsets).
 * d_out: (1 + ndir) channels of offsets[nlay] doubles each: channel 0 the OD, channel
 * 1 + k the derivative along dirs[k].  Synchronous. */
typedef struct {
    double dt;                     /* K */
    double dlnwtot;                /* relative change of Wtot */
    double dlnw[CLBLM_MAXMOL + 1]; /* relative change of W(m) */
} clblm_od_dir;
int clblm_od_tangent(clblm_ctx *ctx, int handle, int nlay, const clblm_layer *layers,
                     const clblm_od_params *pars, int32_t *jrad_state, int ndir, const clblm_od_dir *dirs,
                     const double *dcontinuum, int64_t *offsets, double *d_out, void *stream);

/* ---- a batch captured in a hipGraph ----------------------------------------------
 * (no counterpart in the reference, whose layer loop calls ODLAY one layer at a
 * time; for callers that compute the same layer set again and again, e.g. a
 * retrieval iterating on one scene geometry).  clblm_od_graph_create plans the
 * batch exactly as clblm_od_batch would (same arguments; d_od must be given and
 * offsets/jrad_state are filled as there) and records all of its launches, on the
 * F4 wing stream too, into one hipGraph with its own device scratch.  Nothing runs
 * yet.  clblm_od_graph_launch replays it on `stream` (NULL = the context's stream)
 * and returns once it is enqueued; each replay rewrites d_od from the layer inputs
 * and the line list as they were at creation (tuning changes made later do not
 * apply); until the graph is destroyed, clblm_linelist_free and
 * clblm_linelist_broadening refuse its line list.  Replays of one graph run in launch
 * order, also across streams; replays and other batches writing d_od must be
 * ordered by the caller's streams.  clblm_od_graph_check waits for the last replay
 * and returns the device error of clblm_od_batch (line coupling, capacity), if any. */
typedef struct clblm_od_graph clblm_od_graph;
int clblm_od_graph_create(clblm_ctx *ctx, int handle, int nlay, const clblm_layer *layers,
                          const clblm_od_params *pars, int32_t *jrad_state, int64_t *offsets,
                          double *d_od, clblm_od_graph **graph);
int clblm_od_graph_launch(clblm_ctx *ctx, clblm_od_graph *graph, void *stream);
int clblm_od_graph_check(clblm_ctx *ctx, clblm_od_graph *graph);
int clblm_od_graph_destroy(clblm_ctx *ctx, clblm_od_graph *graph);

/* Launch-shape tuning of a context (no counterpart in the reference, whose
 * loops are sequential).  key selects the parameter; value 0 only queries it;
 * *old_value (if non-NULL) receives the setting before the call.
 *   CLBLM_TUNE_F4_XCD_CHUNK  consecutive F4 tiles given to one XCD (1-4096).
 *   CLBLM_TUNE_F4_FAR_FIELD  1: CLBLM_FP32C evaluates far-wing records (centre at
 *                            least half a tile outside the tile) at 16 Chebyshev
 *                            nodes per F4 tile (default); 2: at every point; 3: as
 *                            1 from a third of a tile.  CLBLM_FP64 always sums
 *                            every point.
 *   CLBLM_TUNE_F4_STREAM     1: the F4 wing chain of a batch runs on a stream of
 *                            the context beside pass 1 (default); 2: every kernel
 *                            of the batch on the batch's stream, one after another.
 *   CLBLM_TUNE_F4_POINTS_FP32C, CLBLM_TUNE_F4_POINTS_FP64  R4 points per lane of the
 *                            F4 gather (2-4; a tile is 64 times that), per precision.
 *   CLBLM_TUNE_CONV_UNROLL   R1 points per step of the pass-1 walk (2, 4 or 8).
 *   CLBLM_TUNE_CONV_SPLIT    the most parts (1, 2 or 4) a pass-1 line walk is split into
 *                            in panels with fewer lines than lanes. */
#define CLBLM_TUNE_F4_XCD_CHUNK 1
#define CLBLM_TUNE_F4_FAR_FIELD 2
#define CLBLM_TUNE_F4_STREAM 3
#define CLBLM_TUNE_F4_POINTS_FP32C 4
#define CLBLM_TUNE_F4_POINTS_FP64 5
#define CLBLM_TUNE_CONV_UNROLL 6
#define CLBLM_TUNE_CONV_SPLIT 7
int clblm_tuning(clblm_ctx *ctx, int key, int value, int *old_value);

/* Autotuner of the launch shapes (no counterpart in the reference).  Runs the batch
 * (layers, pars as for clblm_od_batch; the precision is pars[0].precision) with each
 * candidate value of each key in turn and keeps the fastest, so later batches of the
 * context use it.  keys = NULL (nkeys 0) tunes the F4 points per lane of that
 * precision, CLBLM_TUNE_CONV_UNROLL, CLBLM_TUNE_F4_XCD_CHUNK and CLBLM_TUNE_F4_STREAM.
 * reps timed runs per candidate after one untimed run; *ms_best (if non-NULL) receives
 * the time per batch of the chosen settings.  Query them with clblm_tuning(key, 0, &v).
 * Allocates one output buffer for the duration of the call.  Synchronous. */
int clblm_autotune(clblm_ctx *ctx, int handle, int nlay, const clblm_layer *layers,
                   const clblm_od_params *pars, const int *keys, int nkeys, int reps, double *ms_best);

/* Waits for every batch enqueued on the context and returns the first device-found
 * error among them (0 if none). */
int clblm_sync(clblm_ctx *ctx);

/* Down-welling radiance at the observer of a look-up path (mode3_lookUp_mergeUp,
 * clblm_noScattRT.f90:843-1108): nlay layer ODs on one grid (d_od[l*npts + i], l = 0
 * at the observer), layer and level temperatures as for clblm_rt_up, layerEmis with
 * linInTau 0/1/2 (clblm_EMLAY.f90:104-178).  d_sol_toa (nullable): the solar
 * radiance at the top on the same grid, added through the whole path.  d_tx
 * (nullable) receives the path transmittance.  Enqueued on `stream`; returns once
 * launched. */
int clblm_rt_down(clblm_ctx *ctx, int nlay, int64_t npts, double v1, double dv, const double *d_od,
                  const double *d_tlay, const double *d_tlev, int lin_in_tau, const double *d_sol_toa,
                  double *d_rad, double *d_tx, void *stream);

/* ---- layer grids (setDV, clblm_setDV.f90:114-382) ---------------------------------
 * Host only.  Layer l has pressure pave[l] (hPa), temperature tave[l] (K), broadening
 * column wtot[l] and water column w_h2o[l] (molec/cm2).
 * clblm_calculate_dv: calculateDV (:221-271), the exact DV of every layer for the mean
 * wavenumber vbar, ALFAL0 and SAMPLE.
 * clblm_set_dv: setDV_struct (:114-215).  grid_type CLBLM_GRID_EXACT (-1): the exact
 * DVs; CLBLM_GRID_FIXED_RATIO (0): fixDVRatio (:277-349), the first DV to three
 * significant figures with an even last digit and each next one the previous times
 * 1 or n/(n+1) so that adjacent grids nest (CLBLM_ERR_ARG where the exact DVs of two
 * adjacent layers differ by more than 2.5x, where the reference stops);
 * CLBLM_GRID_UNIFORM (1): the exact DVs and a uniform output DV.  For types 0 and 1
 * a user_dv > 0 within 20 % of the last layer's DV rescales all of them (DVSET,
 * :356-382).  Outputs: dvnormal[nlay] (ALFAL0), dvnarrow[nlay] (alfal0_fine) and
 * *dvout: user_dv (> 0) or the smallest dvnarrow for type 1, -999.9 otherwise. */
#define CLBLM_GRID_EXACT (-1)
#define CLBLM_GRID_FIXED_RATIO 0
#define CLBLM_GRID_UNIFORM 1
int clblm_calculate_dv(int nlay, const double *pave, const double *tave, const double *wtot, const double *w_h2o,
                       double vbar, double alfal0, double sample, double *dv);
int clblm_set_dv(int nlay, const double *pave, const double *tave, const double *wtot, const double *w_h2o,
                 double v1, double v2, double alfal0, double alfal0_fine, double sample, int grid_type,
                 double user_dv, double *dvnormal, double *dvnarrow, double *dvout);

/* ---- spectra on different grids (clblm_Spectrum.f90) -----------------------------
 * clblm_interp_spectrum: interpSpectrum_array2array (:326-590) of the device spectrum
 * d_in[0..nin) on v1 + i dvin onto d_out[0..nout) on v1 + i dvout (nin >= 2; the
 * reference's linear extrapolation for the two points past either end, zero beyond).
 * clblm_interp_to_finest: interpToFinestGrid (:733-795) over nspec device spectra from
 * one v1, spectrum k on dv[k] with nlim[k] points: *dv_out and *nlim_out receive the
 * grid of the first spectrum of smallest DV (query with d_out == NULL); then every
 * spectrum is written to d_out[k] on that grid (nlim_out points; d_out[k] may be
 * d_in[k] for spectra already on it).  Enqueued on `stream`. */
int clblm_interp_spectrum(clblm_ctx *ctx, const double *d_in, int64_t nin, double v1, double dvin, double *d_out,
                          int64_t nout, double dvout, void *stream);
int clblm_interp_to_finest(clblm_ctx *ctx, int nspec, double v1, const double *dv, const int64_t *nlim,
                           const double *const *d_in, double *const *d_out, double *dv_out, int64_t *nlim_out,
                           void *stream);

/* Upwelling radiance at the top of nlay layers each on its own grid, as the reference
 * merges them (mode2_lookDn_mergeUp / layerMerge_mode2, clblm_noScattRT.f90:1287-1296,
 * 1700-1794): layer l's OD (d_od + od_off[l], nlim[l] points on v1 + i dv[l], layer 0
 * at the surface, e.g. from clblm_od_batch without DVout) gives its transmittances and
 * emissions (layerEmis, the down path sec_dn times the view OD) on its grid; before
 * each merge interpToFinestGrid brings the merged radiance and reflectance, or the
 * layer's four arrays, onto the finer grid.  The surface, eps B(t_skin) and 1 - eps,
 * starts on v1 + i surf_dv, surf_nlim points.  t_lay[nlay], t_lev[nlay+1] are HOST
 * arrays (K).  Query with d_rad == NULL: *dv_out, *nlim_out give the final grid; then
 * d_rad (device, *nlim_out doubles) is filled.  Synchronous. */
int clblm_rt_up_grids(clblm_ctx *ctx, int nlay, double v1, const double *dv, const int64_t *nlim,
                      const int64_t *od_off, const double *d_od, const double *t_lay, const double *t_lev,
                      double surf_emis, double t_skin, double surf_dv, int64_t surf_nlim, int lin_in_tau,
                      double sec_dn, double *d_rad, double *dv_out, int64_t *nlim_out, void *stream);

/* ---- radiance and its derivatives ---------------------------------------------
 * Upwelling radiance at the top of nlay layers (layer 0 at the surface) on a
 * common grid v1 + i*dv, i < npts: the merge-up recursion of mode2_lookDn_mergeUp /
 * layerMerge_mode2 (clblm_noScattRT.f90:1287-1296, 1704-1794) with layerEmis
 * (clblm_EMLAY.f90:104-178).  d_od[l*npts + i] is the layer OD along the view path;
 * the downwelling path uses sec_dn * OD (1 for a specular surface, 1.67 for the
 * diffusivity approximation).  lin_in_tau: 0, 1 (LBLRTM) or 2.  Temperatures are
 * device arrays: d_tlay[nlay], d_tlev[nlay+1].  All outputs are device arrays.
 * d_rad[npts]; when d_dr_dod is non-NULL the reverse sweep also writes dR/dOD_l and
 * dR/dT_l through the layer source (nlay*npts each), and, if non-NULL,
 * dR/dTskin and dR/deps (npts each): the adjoint of this recursion, for callers that
 * differentiate the radiance-only path.  Synchronous. */
int clblm_rt_up(clblm_ctx *ctx, int nlay, int64_t npts, double v1, double dv, const double *d_od,
                const double *d_tlay, const double *d_tlev, double surf_emis, double t_skin, int lin_in_tau,
                double sec_dn, double *d_rad, double *d_dr_dod, double *d_dr_dtlay, double *d_dr_dtskin,
                double *d_dr_demis, void *stream);

/* ---- level Jacobians (noScattJacob, clblm_noScattJac.f90:20-336) ---------------
 * level2layer (clblm_noScattJac.f90:1208-1279), host arrays, no device work:
 * derivatives of each layer's mean temperature with respect to its upper and lower
 * level temperatures, and of ln(layer amount) with respect to ln(level number
 * density) of each molecule.  p_lev, t_lev [nlay+1] (level 0 at the surface, hPa, K),
 * t_lay [nlay], q_lev[m*(nlay+1) + k] the density of molecule m at level k (the
 * reference's path%Q); outputs dt_dtup, dt_dtlo [nlay], dlnw_dlnq_up/lo[m*nlay + l]. */
int clblm_level2layer(int nlay, int nmol, const double *p_lev, const double *t_lev, const double *t_lay,
                      const double *q_lev, double *dt_dtup, double *dt_dtlo, double *dlnw_dlnq_up,
                      double *dlnw_dlnq_lo);

/* Jacobians (clblm_noScattJac.f90:346-716) at every point of v1 + i*dv: radiance and
 * its analytic derivatives for a down-looking sensor above the top layer, thermal
 * only.  d_od[l*npts + i] is the layer OD along the view path (nadir: the vertical
 * OD, amfView 1); the down path uses sec_dn times it (the air-mass-scaled path,
 * amfDown = sec_dn).  Device inputs d_tlay[nlay], d_tlev[nlay+1], and for lin_in_tau
 * 0 and 1 the level2layer factors d_dt_dtup, d_dt_dtlo[nlay].  Device outputs:
 * d_rad[npts]; d_dr_dtau[l*npts + i] = dR/d(layer OD); d_dr_dtlev[k*npts + i]
 * (nlay+1 levels) = dR/dT_lev through the level Planck terms and the layer-mean
 * terms, the reference's dRad_dTlev; d_dr_dtskin, d_dr_demis (npts each, optional).
 * As in the reference, the reflected downwelling term is dropped where the total
 * view transmittance is 1e-6 or less.  Synchronous. */
int clblm_rt_jac(clblm_ctx *ctx, int nlay, int64_t npts, double v1, double dv, const double *d_od,
                 const double *d_tlay, const double *d_tlev, const double *d_dt_dtup, const double *d_dt_dtlo,
                 double surf_emis, double t_skin, int lin_in_tau, double sec_dn, double *d_rad, double *d_dr_dtau,
                 double *d_dr_dtlev, double *d_dr_dtskin, double *d_dr_demis, void *stream);

/* clblm_rt_jac without any allocation.  No counterpart in the reference: for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same arguments, kernel and results as
 * clblm_rt_jac, but k_rt_jac's per-level states live in a static device array of
 * CLBLM_RT_JAC_NOALLOC_BYTES that the library holds on each GPU until the process
 * ends, instead of context memory grown with hipMalloc.  It needs
 * 8 (3 nlay + 2) npts <= CLBLM_RT_JAC_NOALLOC_BYTES (e.g. 60 layers x 46 000 points)
 * and returns CLBLM_ERR_CAPACITY beyond that.  All calls on one GPU share the array,
 * which keeps the last call's states: the library serializes them (a per-GPU lock held
 * until the kernel has finished, like clblm_rt_jac synchronous), so calls made at
 * the same time from several threads or contexts wait for each other instead of
 * racing.  Calls no hipMalloc, hipFree or allocating routine. */
#define CLBLM_RT_JAC_NOALLOC_BYTES ((int64_t)64 << 20)
int clblm_rt_jac_noalloc(clblm_ctx *ctx, int nlay, int64_t npts, double v1, double dv, const double *d_od,
                         const double *d_tlay, const double *d_tlev, const double *d_dt_dtup,
                         const double *d_dt_dtlo, double surf_emis, double t_skin, int lin_in_tau, double sec_dn,
                         double *d_rad, double *d_dr_dtau, double *d_dr_dtlev, double *d_dr_dtskin,
                         double *d_dr_demis, void *stream);

/* drv_LevelTemp and drv_LevelConc (clblm_noScattJac.f90:899-1203): level Jacobians
 * from clblm_rt_jac's derivatives and the layer-OD variants the reference computes.
 * d_od_var holds 3 + nmol device blocks of nlay*npts: the layer OD, the OD at
 * Tave + dt, the OD at Wtot + dw_frac*Wtot, then the OD of each molecule alone (all
 * other amounts zero).  dtau/dT = (OD(T+dt) - OD)/dt - dtau/dWtot * Wtot/Tave (the
 * reference: dt = 1 K, dw_frac = 0.01).  Device arrays d_wtot, d_tlay, d_dt_dtup,
 * d_dt_dtlo [nlay], d_dw_dqup/lo[m*nlay + l].  Outputs (device):
 * d_jac_t[k*npts + i] = dR/dT_lev (nlay+1 levels) and, if non-NULL,
 * d_jac_q[(m*(nlay+1) + k)*npts + i] = dR/dln q_lev.  Enqueued on `stream`. */
int clblm_level_jac(clblm_ctx *ctx, int nlay, int nmol, int64_t npts, const double *d_od_var,
                    const double *d_dr_dtau, const double *d_dr_dtlev, const double *d_wtot, const double *d_tlay,
                    double dt, double dw_frac, const double *d_dt_dtup, const double *d_dt_dtlo,
                    const double *d_dw_dqup, const double *d_dw_dqlo, double *d_jac_t, double *d_jac_q,
                    void *stream);

/* Spectral-domain scan: clblm_SCANFN_array (src/clblm_src/clblm_ScanFilter.f90:33-67).
 * d_in: nrows spectra of n device doubles on v1 + i dv, row r at d_in + r ld_in;
 * output row r at d_out + r ld_out (ld_out >= *nout); output limits v1out..v2out.
 * The input is first interpolated onto hwhm/12 when dv is coarser (:331-334).
 *   funct_id 1  the rectangle of adjacent boxes of width 2 hwhm every dvout (CNVRCT,
 *            :1478-1698), which output_JacArray applies as the boxcar pre-average of
 *            Jacobian arrays (clblm_noScattJac.f90:1341-1409): the output step is dvout
 *            rounded to a whole number of input steps; with fewer than two input steps
 *            per box the (interpolated) input is returned unchanged.  dvout != 0.
 *   2..5     triangle, Gaussian, sinc**2, sinc of half width hwhm, tabulated and
 *            convolved as SHRKSC / CONVSC do (:551-955), bit for bit the reference's
 *            sums; -4 and -5 take hwhm as the first zero crossing (:272-280).  Output
 *            step |dvout|, or hwhm / (the function's default sampling) when dvout is 0.
 *            CLBLM_ERR_ARG where the reference would overrun its 5000-point R1 buffer
 *            or its table (dvout too fine or too coarse).  When the input ends before
 *            v2out + the function's bound, the reference writes panels of 2400 points
 *            only up to the one holding its last nonzero input: then the count depends
 *            on the data, the query reads d_in (and waits for it on `stream`), and rows
 *            whose counts differ return CLBLM_ERR_ARG.
 * Two-phase: d_out == NULL returns *nout, *v1res, *dvres only (d_in may be NULL for
 * functID 1 and whenever the count does not depend on the data). */
int clblm_scanfn(clblm_ctx *ctx, const double *d_in, int64_t n, int64_t nrows, int64_t ld_in, double v1,
                 double dv, double v1out, double v2out, double dvout, int funct_id, double hwhm, double *d_out,
                 int64_t ld_out, int64_t *nout, double *v1res, double *dvres, void *stream);

/* ---- continuum ------------------------------------------------------------------
 * The MT_CKD water-vapour continuum coefficients mt_ckd_h2o_absco reads from
 * absco-ref_wv-mt-ckd.nc (mt_ckd_h2o_module.f90:34-41, read_module.f90:50-101): n
 * wavenumbers (increasing, uniform), the self and foreign coefficients at ref_temp K and
 * ref_press hPa, the self temperature exponent; xself and xfrgn are CONTNM's XSELF and
 * XFRGN scale factors (1 by default).  Host arrays, copied.  Layers whose params set
 * CLBLM_OD_MTCKD then carry the continuum. */
int clblm_mtckd_load(clblm_ctx *ctx, int n, const double *wavenumber, const double *self_absco_ref,
                     const double *for_absco_ref, const double *self_texp, double ref_temp, double ref_press,
                     double xself, double xfrgn);

/* The table read from the reference's absco-ref_wv-mt-ckd.nc (netCDF classic, CDF-1 or
 * CDF-2), as read_module.f90's getData reads it (:50-101), with no netCDF library:
 * clblm_mtckd_read_nc returns *n with wavenumber == NULL, then fills the four arrays of
 * *n values and the two scalars; clblm_mtckd_load_nc reads and loads in one call. */
int clblm_mtckd_read_nc(const char *path, int *n, double *wavenumber, double *self_absco_ref,
                        double *for_absco_ref, double *self_texp, double *ref_temp, double *ref_press);
int clblm_mtckd_load_nc(clblm_ctx *ctx, const char *path, double xself, double xfrgn);

/* CONTNM's scale factors XCO2C, XO3CN, XO2CN, XN2CN, XRAYL (CLBLM_OD_Ctrl,
 * clblm_Config.f90:98-102) for the layers that set CLBLM_OD_CONTNM; 1 each by default.
 * XSELF and XFRGN are those of clblm_mtckd_load. */
int clblm_contnm_scales(clblm_ctx *ctx, double xco2c, double xo3cn, double xo2cn, double xn2cn, double xrayl);

/* CONTNM's ABSRB alone (clblm_ODLAY.f90:880-896) for the continua flags selects
 * (CLBLM_OD_MTCKD, CLBLM_OD_CONTNM), on the grid and with the JRAD rule of
 * clblm_mtckd_absrb below. */
int clblm_contnm_absrb(clblm_ctx *ctx, int nlay, const clblm_layer *layers, const double *v1, const double *v2,
                       const int32_t *jrad, int32_t flags, double *d_absrb, int64_t ld, int32_t *nptabs,
                       void *stream);

/* The same continuum alone, as CONTNM adds it to a zeroed ABSRB (clblm_ODLAY.f90:880-896):
 * for layer l, on V1ABS = INT(v1[l]) - 3 at 1 cm-1 (nptabs[l] points, returned), with the
 * radiation term unless jrad[l] is 0 (jrad: OPDPTH's JRAD after its flip,
 * clblm_ODLAY.f90:877-878).  d_absrb (device) row l at d_absrb + l ld, zero past the
 * continuum's points; d_absrb == NULL returns nptabs only.  Enqueued on `stream`, waited
 * for before returning. */
int clblm_mtckd_absrb(clblm_ctx *ctx, int nlay, const clblm_layer *layers, const double *v1, const double *v2,
                      const int32_t *jrad, double *d_absrb, int64_t ld, int32_t *nptabs, void *stream);

/* ---- merge-down radiance -----------------------------------------------------------
 * mode1_lookDn_mergeDn with layerMerge_mode1 (src/clblm_src/clblm_noScattRT.f90:42-842),
 * addThermalBoundary and addSolarBoundary (:1945-2148), flux off: the radiance at an
 * observer above the surface, merged from the top layer down, every layer's ODs on its
 * own grid (nlim[l] points on v1 + i dv[l], device arrays at od_off[l]; layer 0 at the
 * surface).  The view, down and solar path ODs of layer l are d_od_up, d_od_dn and
 * d_od_sun times fac_up[l], fac_dn[l], fac_sun[l] (the reference's airmass-scaled form
 * scalUpFac, scalDownFac, scalSolarFac; NULL factors are 1, for ODs computed along each
 * path).  Layers l < obs_level lie below the observer and enter the view path; the
 * others only the down and solar paths.  Before each merge the reference's
 * interpToFinestGrid (clblm_Spectrum.f90:733-795) brings the spectra it combines onto
 * the finest of their grids; the merged spectra start on the two-point grid v1, v2.
 * Surface: constant emissivity (reflecting below 1, which needs thermal), skin
 * temperature, solar reflectance; d_sol: the solar radiance at the top (S0 umu0/pi) on
 * v1 + i sol_dv, covering no more than the layers' grids do (past them the reference's
 * interpolation reads outside its extended input; zeros here).  Outputs (device, on the final grid *dv_out, *nlim_out points): d_rad
 * the radiance and, if non-NULL, d_tx the total view transmittance.  d_rad == NULL
 * returns the final grid only.  Waits for the stream before returning. */
typedef struct {
    int32_t nlay;
    int32_t obs_level;           /* levels from 0 (surface) to nlay (top) */
    int32_t lin_in_tau;          /* 0, 1 or 2 (layerEmis) */
    int32_t thermal, solar;      /* sources */
    double v1, v2;
    double surf_emis, t_skin, surf_refl;
} clblm_mode1_params;
int clblm_rt_mode1(clblm_ctx *ctx, const clblm_mode1_params *p, const double *dv, const int64_t *nlim,
                   const int64_t *od_off, const double *d_od_up, const double *d_od_dn, const double *d_od_sun,
                   const double *fac_up, const double *fac_dn, const double *fac_sun, const double *t_lay,
                   const double *t_lev, const double *d_sol, int64_t sol_n, double sol_dv, double *d_rad,
                   double *d_tx, double *dv_out, int64_t *nlim_out, void *stream);

/* Jacobians (clblm_noScattJac.f90:346-716) with every branch of the reference: the
 * layer's path ODs are its vertical OD d_od[l npts + i] times the air-mass factors
 * amf_view, amf_down, amf_sun (host arrays [nlay], NULL: 1; pathOD_up_dn_sun's scaled
 * paths :826-838); the observer at level obs_level (0 the surface .. nlay the top);
 * up-looking (the down-welling radiance from the layers above the observer, :436-494)
 * or down-looking (the reflected downwelling, the surface emission, the solar term
 * d_sol[i] (solar radiance at the top) times rsfc_sun through the sun and view paths
 * when solar is set, and the upwelling of the layers below the observer, :496-678).
 * Outputs as clblm_rt_jac's, dR/dtau with respect to the vertical OD, plus
 * d_dr_drsfc (dR/d solar reflectance); the surface outputs are zero when up-looking.
 * Device arrays; d_dr_dtskin, d_dr_demis, d_dr_drsfc may be NULL.  Synchronous. */
typedef struct {
    int32_t nlay, obs_level, uplook, lin_in_tau, solar, pad;
    double surf_emis, t_skin, rsfc_sun;
    const double *amf_view, *amf_down, *amf_sun;
} clblm_jac_paths_params;
int clblm_rt_jac_paths(clblm_ctx *ctx, const clblm_jac_paths_params *p, int64_t npts, double v1, double dv,
                       const double *d_od, const double *d_tlay, const double *d_tlev, const double *d_dt_dtup,
                       const double *d_dt_dtlo, const double *d_sol, double *d_rad, double *d_dr_dtau,
                       double *d_dr_dtlev, double *d_dr_dtskin, double *d_dr_demis, double *d_dr_drsfc,
                       void *stream);

/* ---- fluxes -------------------------------------------------------------------------
 * The flux branches of mode1_lookDn_mergeDn (src/clblm_src/clblm_noScattRT.f90:133-422,
 * 482-507) and mode2_lookDn_mergeUp (:1214-1450) with the driver's net flux and heating
 * rate (clblm_Drivers.f90:418-466): every layer's OD on one grid (d_od[l npts + i],
 * layer 0 at the surface, device; v1 + i dv, i < npts), temperatures d_tlay[nlay],
 * d_tlev[nlay + 1] (device), level pressures p_lev[nlay + 1] (hPa, host).  nang = 1..3
 * quadrature angles (secants and weights of clblm_ConstParam.f90:31-38 and the RT
 * routines :155-172); per angle the ODs are scaled by the secant, the down radiance is
 * merged from the top (layerEmis toward the lower level), the surface radiance is
 * eps B(t_skin) plus (1 - eps) times the downwelling (lambertian: the quadrature flux
 * over pi, the reference's ThmReflMode 0; else the angle's own radiance), and the up
 * radiance is merged from the surface with no reflected term.  The radiances of each
 * level are summed in groups of int(dv_flux / dv) points from v1 and weighted:
 * flux = (sum_a w_a S_a) dv 1e4 2 pi (W/m2 per group).  Outputs (host, level-major,
 * (nlay + 1) x *nbin, level 0 at the surface): flux_up (level 0 the surface upwelling
 * flux), flux_dn (0 at the top), net_flux = up - dn, heating = 8.4391
 * (net_j - net_j+1) / (p_j - p_j+1) K/day (0 at the top).  *nbin = NOUT = (v2 - v1) /
 * dv_flux, which must be whole (to 1e-4) as the reference requires; the reference's
 * arrays carry one more group, always zero.  flux_up == NULL returns *nbin only.
 * Synchronous. */
typedef struct {
    int32_t nlay, nang, lin_in_tau, lambertian;
    double v1, v2, dv, dv_flux, surf_emis, t_skin;
} clblm_flux_params;
int clblm_rt_flux(clblm_ctx *ctx, const clblm_flux_params *p, const double *d_od, int64_t npts, const double *d_tlay,
                  const double *d_tlev, const double *p_lev, double *flux_up, double *flux_dn, double *net_flux,
                  double *heating, int32_t *nbin, void *stream);

/* Spectral response function filter: clblm_FLTRFN with CNVFLT
 * (src/clblm_src/clblm_ScanFilter.f90:1703-1900), once per channel as
 * postProc_filter_singleStruct does (clblm_PostProc.f90:289-336).  d_in: the spectrum (n
 * device doubles on v1 + i dv); channel c: nptf[c] filter values at xf + xf_off[c] (host)
 * on v1f[c] + i dvf[c] (dvf < 0: v1f is the centre, nptf odd).  d_out[c] (device): the
 * spectrum integrated against the filter (end values halved, linear between values)
 * over the filter's own integral, bit for bit the reference's sum; 0 for v1f < 0, where
 * the reference returns no value.  A channel whose end falls where CNVFLT's subscript
 * check stops the reference (:1830-1841) gets NaN and the call returns CLBLM_ERR_ARG
 * after filling the others.  Waits for `stream` before returning. */
int clblm_filter(clblm_ctx *ctx, const double *d_in, int64_t n, double v1, double dv, int nchan, const double *v1f,
                 const double *dvf, const int32_t *nptf, const int64_t *xf_off, const double *xf, double *d_out,
                 void *stream);

/* ---- instrument scan ------------------------------------------------------------
 * Drop-in for clblm_FFTSCN_array (clblm_FFT.f90:308-349): the spectrum d_in[0..n)
 * (device) on v1 + i*dv is scanned with function funct_id of half width hwhm between
 * v1out and v2out, then interpolated (4-point) onto dvout when dvout > 0.
 *   funct_id  1 boxcar; 2 triangle, 3 Gaussian, 4 sinc**2, 5 sinc, 6 Beer, 7 Hamming,
 *             8 Hanning, 9/10/11 Norton-Beer weak/medium/strong, convolved by FFT over
 *             the reference's transform length with the apodization evaluated
 *             analytically above CRATIO half widths of scanned range and as the
 *             normalised transform of the scanning function below (2..8 only);
 *             102..111 the same for a spectrum pre-scanned elsewhere with a boxcar
 *             of half width pre_box_hw, which is deconvolved; -1 deconvolution of
 *             such a boxcar only.  12 (Brault) and 13 (Kaiser-Bessel) need an optical
 *             path difference that clblm_FFTSCN does not take, and 14 (Kiruna) is not
 *             supported: these return an error.
 *   fill      the transform padding, in [0, 1].
 *   pre_box_hw  > 0: boxcar pre-scan of that half width (functions 2..11), with the
 *             boxcar deconvolved when deconv_pre_box != 0; <= 0: none.
 * Query with d_out == NULL: *nout, *v1res and *dvres give the output grid; then d_out
 * (device, *nout doubles) is filled on `stream`.  precision CLBLM_FP32C transforms in
 * FP32. */
int clblm_fft_scan(clblm_ctx *ctx, const double *d_in, int64_t n, double v1, double dv, double v1out,
                   double v2out, double dvout, int funct_id, double hwhm, double fill, double pre_box_hw,
                   int deconv_pre_box, int precision, double *d_out, int64_t *nout, double *v1res, double *dvres,
                   void *stream);

/* ---- device memory for callers without a HIP binding ------------------------------
 * No counterpart in the reference: Fortran callers of the device-array entry points
 * (clblm_od_batch, clblm_rt_jac, clblm_level_jac ...) get their device arrays here.
 * clblm_dev_put / clblm_dev_get copy after the device is idle (synchronous);
 * clblm_dev_free waits for the context's batches. */
int clblm_dev_alloc(clblm_ctx *ctx, int64_t bytes, void **d);
int clblm_dev_free(clblm_ctx *ctx, void *d);
int clblm_dev_put(clblm_ctx *ctx, void *d, const void *h, int64_t bytes);
int clblm_dev_get(clblm_ctx *ctx, void *h, const void *d, int64_t bytes);

/* ---- timing -------------------------------------------------------------------
 * Average device time (ms) of each pipeline stage over the launches since the
 * last reset, measured with events on the stream the stages run on (waits for the
 * enqueued batches). */
int clblm_timing(clblm_ctx *ctx, int reset, double *ms_prep, double *ms_conv, double *ms_f4,
                 double *ms_panel, int64_t *launches);

/* Work done per launch, averaged since the last reset: line-shape evaluations of
 * the R1/R2/R3 convolution (points of each line's R1 support) and of the F4 wing
 * sum (line, R4 point) pairs, with the average device time of those two kernels.
 * reset != 0 also clears clblm_timing's accumulators. */
int clblm_work(clblm_ctx *ctx, int reset, int64_t *conv_evals, int64_t *f4_evals, double *ms_conv_kernel,
               double *ms_f4_kernel);

#ifdef __cplusplus
}
#endif
#endif
