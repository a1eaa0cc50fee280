// Layer spectral grids: setDV_struct, calculateDV, fixDVRatio and DVSET
// (src/clblm_src/clblm_setDV.f90:114-382), host code of the library.  The reference
// computes these in REAL promoted to double (-fdefault-real-8); every expression here
// rounds in the same order.
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/clblm_hip.h"

extern "C" __attribute__((visibility("hidden"))) int clblm_set_error(int rc, const char *msg);

namespace {

constexpr double P0 = 1013.25;       // Press1013, clblm_ConstParam.f90:110
constexpr double T0 = 296.0;         // Temp296 :112
constexpr double AVMASS = 36.0;      // AVMWT :118
constexpr double FILLREAL = -999.9;  // :48

// 10.**k for an integer k, as the reference's build evaluates it: the exact power
// for k >= 0 and the correctly rounded reciprocal of it for k < 0.
double pow10i(int k) {
    double p = 1.0;
    for (int i = 0; i < (k < 0 ? -k : k); i++) p *= 10.0;
    return k < 0 ? 1.0 / p : p;
}

// calculateDV (:221-271): the exact DV of each layer for the mean wavenumber vbar.
void calculate_dv(int nlay, const double *pave, const double *tave, const double *wtot, const double *w_h2o,
                  double vbar, double alfal0, double sample, double *dv) {
    for (int l = 0; l < nlay; l++) {
        const double frh2o = w_h2o[l] / wtot[l];
        const double alfcor = (pave[l] / P0) * std::sqrt(T0 / tave[l]);
        const double h2oslf = (1. - frh2o + 5. * frh2o);   // water self-broadening
        const double albar = alfal0 * alfcor * h2oslf;
        const double adbar = 3.58115E-07 * vbar * std::sqrt(tave[l] / AVMASS);
        const double avbar = 0.5 * (albar + std::sqrt(albar * albar + 4. * adbar * adbar));
        dv[l] = avbar / sample;
    }
}

// fixDVRatio (:277-349): the first layer's DV rounded to three significant figures
// with an even last digit, then each next DV the previous one times 1 (ratio below
// 1.2) or ITYPE/(ITYPE+1), ITYPE = INT(1/(ratio-1) + 0.5) with 3 taken as 2, so that
// adjacent layer grids nest.  A ratio above 2.5 stops the reference.
int fix_dv_ratio(double *out, const double *in, int nlay) {
    const double TYPMAX = 2.5;
    double olddv = 0.;
    for (int l = 0; l < nlay; l++) {
        double dv = in[l];
        if (l == 0) {
            const int iscal = (int)(std::log10(dv) - 3.);
            const double scal = pow10i(iscal);
            int idv = (int)((dv / scal) + 0.5);
            if (idv % 2 > 0) idv = idv + 1;
            dv = scal * (double)idv;
        } else {
            const double type = olddv / dv;
            if (type > TYPMAX) {
                char msg[160];
                snprintf(msg, sizeof msg, "fixDVRatio: DV ratio %.6g between layers %d and %d exceeds %.2f", type, l,
                         l + 1, TYPMAX);
                return clblm_set_error(CLBLM_ERR_ARG, msg);
            } else if (type >= 1.2) {
                int itype = (int)(1. / (type - 1.) + 0.5);
                if (itype == 3) itype = 2;
                dv = olddv * (double)itype / (double)(itype + 1);
            } else {
                dv = olddv;
            }
        }
        olddv = dv;
        out[l] = dv;
    }
    return 0;
}

// DVSET (:356-382): every DV scaled by user_dv / DV of the last layer.
void dvset(double *out, const double *in, double user_dv, int nlay) {
    const double ratio = user_dv / in[nlay - 1];
    for (int l = 0; l < nlay; l++) out[l] = in[l] * ratio;
}

}  // namespace

extern "C" {

int clblm_calculate_dv(int nlay, const double *pave, const double *tave, const double *wtot, const double *w_h2o,
                       double vbar, double alfal0, double sample, double *dv) {
    if (nlay <= 0 || !pave || !tave || !wtot || !w_h2o || !dv)
        return clblm_set_error(CLBLM_ERR_ARG, "null argument or nlay <= 0");
    if (!(sample > 0.)) return clblm_set_error(CLBLM_ERR_ARG, "sample must be > 0");
    calculate_dv(nlay, pave, tave, wtot, w_h2o, vbar, alfal0, sample, dv);
    return 0;
}

int clblm_set_dv(int nlay, const double *pave, const double *tave, const double *wtot, const double *w_h2o,
                 double v1, double v2, double alfal0, double alfal0_fine, double sample, int grid_type,
                 double user_dv, double *dvnormal, double *dvnarrow, double *dvout) {
    if (nlay <= 0 || !pave || !tave || !wtot || !w_h2o || !dvnormal || !dvnarrow || !dvout)
        return clblm_set_error(CLBLM_ERR_ARG, "null argument or nlay <= 0");
    if (grid_type < -1 || grid_type > 1) return clblm_set_error(CLBLM_ERR_ARG, "grid_type must be -1, 0 or 1");
    if (!(sample > 0.)) return clblm_set_error(CLBLM_ERR_ARG, "sample must be > 0");
    double vbar = 0.5 * (v1 + v2);
    if (v2 < v1) vbar = v1;
    std::vector<double> dvc(nlay), dvc_fine(nlay), dvl(nlay), dvl_fine(nlay);
    calculate_dv(nlay, pave, tave, wtot, w_h2o, vbar, alfal0, sample, dvc.data());
    calculate_dv(nlay, pave, tave, wtot, w_h2o, vbar, alfal0_fine, sample, dvc_fine.data());
    dvl = dvc;
    dvl_fine = dvc_fine;
    if (grid_type >= 0) {
        if (grid_type == 0) {
            int rc = fix_dv_ratio(dvl.data(), dvc.data(), nlay);
            if (rc) return rc;
            rc = fix_dv_ratio(dvl_fine.data(), dvc_fine.data(), nlay);
            if (rc) return rc;
        }
        if (user_dv > 0.) {   // a user DV within 20 % of the last layer's rescales all of them
            if (std::fabs(1. - user_dv / dvl[nlay - 1]) < 0.2) dvset(dvl.data(), dvl.data(), user_dv, nlay);
            if (std::fabs(1. - user_dv / dvl_fine[nlay - 1]) < 0.2)
                dvset(dvl_fine.data(), dvl_fine.data(), user_dv, nlay);
        }
    }
    double mn = dvl_fine[0];
    for (int l = 0; l < nlay; l++) {
        dvnormal[l] = dvl[l];
        dvnarrow[l] = dvl_fine[l];
        if (dvl_fine[l] < mn) mn = dvl_fine[l];
    }
    if (grid_type == 1) *dvout = user_dv > 0. ? user_dv : mn;
    else *dvout = FILLREAL;
    return 0;
}

}  // extern "C"
