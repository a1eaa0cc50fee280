// Upwelling radiance through the layer stack and its reverse-mode derivatives.
//
// Forward: the merge-up recursion of mode2_lookDn_mergeUp / layerMerge_mode2
// (src/clblm_src/clblm_noScattRT.f90:1287-1296, 1704-1794) with the layer source of
// layerEmis (src/clblm_src/clblm_EMLAY.f90:104-178):
//     R_0 = eps B(Tskin),  rho_0 = 1 - eps
//     R_{l+1} = Eup_l + ta_l (R_l + Edn_l rho_l),  rho_{l+1} = ta_l tb_l rho_l
// with ta = exp(-tau_l), tb = exp(-sec_dn tau_l) and, for linear-in-tau 1,
// E = (1 - t) (B(Ta) + 2 (B(Tave) - B(Ta)) h(od)),  h = 1/od - 1/(e^od - 1) (rlin).
// Reverse: the adjoint of that recursion, one lane per spectral point, giving
// dR/dtau_l (view-path OD; the down path follows it through sec_dn), dR/dT_l through
// the layer source B(T_l), dR/dTskin and dR/deps.  The per-layer forward states are
// kept in a scratch buffer between the two sweeps.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <string>

#include "../../include/clblm_hip.h"
#include "spectrum_interp.h"

namespace {
constexpr double RADCN1 = 1.191042722E-12;
constexpr double RADCN2 = 1.4387752;

// planck (clblm_EMLAY.f90:18-40) with its temperature derivative (planckFnDrv :45-72)
__device__ inline double planck_x(double v, double xkt) {
    if (xkt <= 1.e-6) return 0.;
    double cv3 = RADCN1 * v * v * v;
    double evt = exp(v / xkt);
    return cv3 / (evt - 1.);
}
__device__ inline void planck_drv(double v, double t, double *b, double *db) {
    double xkt = t / RADCN2;
    if (xkt <= 1.e-6) { *b = 0.; *db = 0.; return; }
    double cv3 = RADCN1 * v * v * v;
    double evt = exp(v / xkt);
    *b = cv3 / (evt - 1.);
    *db = *b * *b * (evt * v) / (RADCN2 * cv3 * xkt * xkt);
}
// rlin (clblm_EMLAY.f90:187-223)
__device__ inline void rlin(double od, double *hlf, double *dhlf) {
    const double c = -1. / 12;
    if (od < 1.e-03) { *hlf = .5 + od * c; *dhlf = c; }
    else if (od > 20.) { *hlf = 1. / od; *dhlf = -1. / (od * od); }
    else {
        double tex = exp(od);
        *hlf = 1. / (od) - 1. / (tex - 1.);
        *dhlf = -1. / (od * od) + tex / ((tex - 1) * (tex - 1));
    }
}

struct Src {  // one layer source and its partials
    double e, de_dod, de_dba, de_dbb;  // d/d(od) at fixed t passed separately
};

// layerEmis at one point: E = (1-tx) * S(od) and the partials of S.
__device__ inline void layer_src(int lin, double od, double ba, double bb, double bbar, double *S,
                                 double *dS_dod, double *dS_dba, double *dS_dbb, double *dS_dbbar) {
    *dS_dba = *dS_dbb = *dS_dbbar = 0.;
    if (lin == 0) { *S = bbar; *dS_dod = 0.; *dS_dbbar = 1.; return; }
    double h, dh;
    rlin(od, &h, &dh);
    if (lin == 1) {
        *S = ba + 2 * (bbar - ba) * h;
        *dS_dod = 2 * (bbar - ba) * dh;
        *dS_dba = 1. - 2. * h;
        *dS_dbbar = 2. * h;
    } else {
        *S = ba + (bb - ba) * h;
        *dS_dod = (bb - ba) * dh;
        *dS_dba = 1. - h;
        *dS_dbb = h;
    }
}

__global__ void k_rt_up(int nlay, int64_t npts, double v1, double dv, const double *__restrict__ od,
                        const double *__restrict__ t_lay, const double *__restrict__ t_lev, double emis,
                        double tskin, int lin, double sec_dn, double *__restrict__ rad, double *__restrict__ dr_dod,
                        double *__restrict__ dr_dt, double *__restrict__ dr_dtskin, double *__restrict__ dr_demis,
                        double *__restrict__ state) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npts) return;
    double v = v1 + (double)i * dv;
    double bs, dbs;
    planck_drv(v, tskin, &bs, &dbs);
    double r = emis * bs;
    double rho = 1. - emis;
    const bool adj = dr_dod != nullptr;
    for (int l = 0; l < nlay; l++) {
        double oda = od[(int64_t)l * npts + i];
        double odb = oda * sec_dn;
        double txa = exp(-oda), txb = exp(-odb);
        double tbot = t_lev[l], ttop = t_lev[l + 1], tave = t_lay[l];
        if (adj) {
            state[(2 * (int64_t)l) * npts + i] = r;
            state[(2 * (int64_t)l + 1) * npts + i] = rho;
        }
        double xkt = tave / RADCN2;
        double bbar = planck_x(v, xkt);
        double emdn, emup;
        if (lin == 0) {
            emdn = (1 - txb) * bbar;
            emup = emdn;
        } else if (lin == 1) {
            double h;
            double dh;
            rlin(odb, &h, &dh);
            double ba = planck_x(v, tbot / RADCN2);
            emdn = (1 - txb) * (ba + 2 * (bbar - ba) * h);
            rlin(oda, &h, &dh);
            double bu = planck_x(v, ttop / RADCN2);
            emup = (1 - txa) * (bu + 2 * (bbar - bu) * h);
        } else {
            double h, dh;
            rlin(odb, &h, &dh);
            double ba = planck_x(v, tbot / RADCN2), bb = planck_x(v, ttop / RADCN2);
            emdn = (1 - txb) * (ba + (bb - ba) * h);
            rlin(oda, &h, &dh);
            emup = (1 - txa) * (bb + (ba - bb) * h);
        }
        r = emup + txa * (r + emdn * rho);
        rho = txa * txb * rho;
    }
    rad[i] = r;
    if (!adj) return;
    // reverse sweep
    double rb = 1., rhob = 0.;
    for (int l = nlay - 1; l >= 0; l--) {
        double oda = od[(int64_t)l * npts + i];
        double odb = oda * sec_dn;
        double txa = exp(-oda), txb = exp(-odb);
        double tbot = t_lev[l], ttop = t_lev[l + 1], tave = t_lay[l];
        double rl = state[(2 * (int64_t)l) * npts + i];
        double rhol = state[(2 * (int64_t)l + 1) * npts + i];
        double bbar, dbbar;
        planck_drv(v, tave, &bbar, &dbbar);
        double ba_dn, bb_dn, ba_up, bb_up;   // boundary Planck values for each direction
        double xbot = planck_x(v, tbot / RADCN2), xtop = planck_x(v, ttop / RADCN2);
        ba_dn = xbot; bb_dn = xtop; ba_up = xtop; bb_up = xbot;
        double Sdn, Sdn_od, Sdn_ba, Sdn_bb, Sdn_bbar, Sup, Sup_od, Sup_ba, Sup_bb, Sup_bbar;
        layer_src(lin, odb, ba_dn, bb_dn, bbar, &Sdn, &Sdn_od, &Sdn_ba, &Sdn_bb, &Sdn_bbar);
        layer_src(lin, oda, ba_up, bb_up, bbar, &Sup, &Sup_od, &Sup_ba, &Sup_bb, &Sup_bbar);
        double emdn = (1 - txb) * Sdn;
        double emup = lin != 0 ? (1 - txa) * Sup : emdn;
        // R_{l+1} = emup + txa (rl + emdn rhol);  rho_{l+1} = txa txb rhol
        double emupb = rb;
        double txab = rb * (rl + emdn * rhol) + rhob * txb * rhol;
        double emdnb = rb * txa * rhol;
        double txbb = rhob * txa * rhol;
        double rlb = rb * txa;
        double rholb = rb * txa * emdn + rhob * txa * txb;
        // sources: E = (1 - t) S(od)
        double odab = -txa * txab, odbb = -txb * txbb, bbarb = 0.;
        if (lin == 0) {
            // emup == emdn == (1 - txb) bbar (reference: lyrEmUp = lyrEmDn)
            double eb = emupb + emdnb;
            odbb += txb * bbar * eb;
            bbarb += (1 - txb) * eb;
        } else {
            odbb += emdnb * (txb * Sdn + (1 - txb) * Sdn_od);
            odab += emupb * (txa * Sup + (1 - txa) * Sup_od);
            bbarb += emdnb * (1 - txb) * Sdn_bbar + emupb * (1 - txa) * Sup_bbar;
        }
        dr_dod[(int64_t)l * npts + i] = odab + sec_dn * odbb;
        dr_dt[(int64_t)l * npts + i] = bbarb * dbbar;
        rb = rlb;
        rhob = rholb;
    }
    // R_0 = emis B(tskin), rho_0 = 1 - emis
    if (dr_dtskin) dr_dtskin[i] = rb * emis * dbs;
    if (dr_demis) dr_demis[i] = rb * bs - rhob;
}

// Down-welling radiance at the lowest level of a look-up path: mode3_lookUp_mergeUp
// (clblm_noScattRT.f90:843-1108).  Layers from the observer up: each layer's
// downward emission (layerEmis with Tbot, Ttop, :972-974) is added through the
// transmittance of the layers below it, then that transmittance takes the layer's
// (:975-990); the top-of-atmosphere solar radiance, if given, comes through the
// whole path (:1021-1027).  One lane per spectral point.
__global__ void __launch_bounds__(256) k_rt_down(int nlay, int64_t npts, double v1, double dv,
                                                 const double *__restrict__ od, const double *__restrict__ t_lay,
                                                 const double *__restrict__ t_lev, int lin,
                                                 const double *__restrict__ sol, double *__restrict__ rad,
                                                 double *__restrict__ tx_out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npts) return;
    const double v = v1 + (double)i * dv;
    double r = 0., tacc = 1.;
    for (int l = 0; l < nlay; l++) {
        const double o = od[(int64_t)l * npts + i];
        const double tx = exp(-o);
        const double tbot = t_lev[l], ttop = t_lev[l + 1], tave = t_lay[l];
        const double bbar = planck_x(v, tave / RADCN2);
        double em;
        if (lin == 0) {
            em = (1 - tx) * bbar;
        } else {
            double h, dh;
            rlin(o, &h, &dh);
            const double ba = planck_x(v, tbot / RADCN2);
            em = lin == 1 ? (1 - tx) * (ba + 2 * (bbar - ba) * h)
                          : (1 - tx) * (ba + (planck_x(v, ttop / RADCN2) - ba) * h);
        }
        r = r + em * tacc;
        tacc = tacc * tx;
    }
    if (sol) r = r + tacc * sol[i];
    rad[i] = r;
    if (tx_out) tx_out[i] = tacc;
}

extern "C" __attribute__((visibility("hidden"))) int clblm_rt_down_impl(int nlay, int64_t npts, double v1, double dv,
                                                                        const double *d_od, const double *d_tlay,
                                                                        const double *d_tlev, int lin,
                                                                        const double *d_sol, double *d_rad,
                                                                        double *d_tx, void *stream) {
    hipLaunchKernelGGL(k_rt_down, dim3((unsigned)((npts + 255) / 256)), dim3(256), 0, (hipStream_t)stream, nlay,
                       npts, v1, dv, d_od, d_tlay, d_tlev, lin, d_sol, d_rad, d_tx);
    return hipGetLastError() == hipSuccess ? 0 : CLBLM_ERR_HIP;
}

// ---------------------------------------------------------------------------------
// Jacobians (src/clblm_src/clblm_noScattJac.f90:346-716) per spectral point, for a
// down-looking sensor above the top layer (obsLev = nlay + 1), thermal only: the view
// OD of layer l is od[l*n+i] (amfView = 1) and the down-path OD is sec_dn times it
// (amfDown = sec_dn, the air-mass-scaled path of pathOD_up_dn_sun :826-838).  The
// reference's cumulative-transmittance form with its summation order, including the
// rule that the reflected downwelling term is dropped when the total view
// transmittance is 1e-6 or less (:480).  Layer 0 is at the surface; level k is the
// bottom of layer k.  Per-layer states the loops pass to each other live in `sa`
// (view transmittance to the top, nlay+1 levels), `sb` (two-path transmittance,
// nlay+1) and `sbar` (layer-mean source derivative of the downwelling pass, nlay),
// all [k][n]; dr_dtlev holds the downwelling level terms until the upwelling pass
// folds them in.
__global__ void __launch_bounds__(256) k_rt_jac(int L, int64_t n, double v1, double dv, const double *__restrict__ od,
                                                const double *__restrict__ t_lay, const double *__restrict__ t_lev,
                                                double emis, double tskin, int lin, double sec,
                                                const double *__restrict__ dt_dtup, const double *__restrict__ dt_dtlo,
                                                double *__restrict__ rad_o, double *__restrict__ dtau,
                                                double *__restrict__ dtlev, double *__restrict__ dts,
                                                double *__restrict__ de, double *__restrict__ sa,
                                                double *__restrict__ sb, double *__restrict__ sbar) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double vn = v1 + (double)i * dv;
    auto at = [n, i](int k) { return (int64_t)k * n + i; };
    // 0- view-path transmittance from the top down to the surface
    double cumA = 0.;
    sa[at(L)] = 1.;
    for (int l = L - 1; l >= 0; l--) {
        cumA = cumA + od[at(l)];
        sa[at(l)] = exp(-cumA);
    }
    const double totODUp = cumA, totTxUp = sa[at(0)];
    double rad = 0.;
    // 1- downwelling thermal radiance, reflected at the surface
    const bool dn = totTxUp > 1.e-06;
    if (dn) {
        sb[at(0)] = totTxUp;
        double cumB = 0.;
        for (int l = 0; l < L; l++) {
            cumB = cumB + od[at(l)] * sec;
            sb[at(l + 1)] = exp(-(totODUp + cumB));
        }
        double bhi = 0., dbhi = 0., tdn_hi = 0.;   // level l+1 values carried down
        if (lin != 0) planck_drv(vn, t_lev[L], &bhi, &dbhi);
        double bhi_tx = sb[at(L)];
        for (int l = L - 1; l >= 0; l--) {
            const double blo_tx = sb[at(l)];
            const double dTran = blo_tx - bhi_tx;
            const double odd = od[at(l)] * sec;
            double bbar, dbbar;
            planck_drv(vn, t_lay[l], &bbar, &dbbar);
            double Blt;
            if (lin != 0) {
                double blo, dblo, h, dh;
                planck_drv(vn, t_lev[l], &blo, &dblo);
                rlin(odd, &h, &dh);
                if (lin == 2) {
                    Blt = blo + (bhi - blo) * h;
                    dtau[at(l)] = (bhi_tx * Blt - rad + dTran * dh * (bhi - blo)) * sec;
                    tdn_hi = tdn_hi + dTran * dbhi * h;
                    dtlev[at(l + 1)] = tdn_hi;                 // level l+1 complete
                    tdn_hi = dTran * dblo * (1. - h);
                } else {
                    Blt = 2.0 * bbar * h + blo * (1.0 - 2.0 * h);
                    dtau[at(l)] = (bhi_tx * Blt - rad + 2.0 * dTran * dh * (bbar - blo)) * sec;
                    dtlev[at(l)] = dTran * dblo * (1.0 - 2.0 * h);
                    sbar[at(l)] = dTran * 2.0 * dbbar * h;
                }
                bhi = blo;
                dbhi = dblo;
            } else {
                Blt = bbar;
                dtau[at(l)] = (bhi_tx * bbar - rad) * sec;
                sbar[at(l)] = dTran * dbbar;
            }
            rad = rad + dTran * Blt;
            bhi_tx = blo_tx;
        }
        if (lin == 2) dtlev[at(0)] = tdn_hi;
    }
    // 2-3- surface reflectance and emission
    const double rsfc = 1. - emis;
    double bs, dbs;
    planck_drv(vn, tskin, &bs, &dbs);
    if (de) de[i] = totTxUp * bs - rad;
    if (dts) dts[i] = emis * totTxUp * dbs;
    rad = rad * rsfc + emis * totTxUp * bs;
    // 5- upwelling thermal radiance; level derivatives are completed on the way up
    double blo = 0., dblo = 0.;
    if (lin != 0) planck_drv(vn, t_lev[0], &blo, &dblo);
    double tup_lo = 0.;          // dRad_dT_up of the current layer's bottom level
    double tbar_prev = 0.;       // layer-mean derivative of the layer below
    double alo = sa[at(0)];
    for (int l = 0; l < L; l++) {
        const double ahi = sa[at(l + 1)];
        const double dTran = ahi - alo;
        const double oda = od[at(l)];
        double bbar, dbbar;
        planck_drv(vn, t_lay[l], &bbar, &dbbar);
        double t_l = dn ? dtau[at(l)] * rsfc : 0. * rsfc;
        double tb_l = (dn && lin != 2) ? sbar[at(l)] * rsfc : 0.;
        double Blt, tup_hi = 0., bhi = 0., dbhi = 0.;
        if (lin != 0) {
            double h, dh;
            planck_drv(vn, t_lev[l + 1], &bhi, &dbhi);
            rlin(oda, &h, &dh);
            if (lin == 2) {
                Blt = bhi * (1 - h) + blo * h;
                t_l = t_l + (alo * Blt - rad + dTran * dh * (blo - bhi));
                tup_hi = dTran * dbhi * (1 - h);
                tup_lo = tup_lo + dTran * dblo * h;
            } else {
                Blt = 2.0 * bbar * h + bhi * (1.0 - 2.0 * h);
                t_l = t_l + (alo * Blt - rad + dTran * dh * 2.0 * (bbar - bhi));
                tup_hi = dTran * dbhi * (1.0 - 2.0 * h);
                tb_l = tb_l + dTran * 2.0 * dbbar * h;
            }
        } else {
            Blt = bbar;
            t_l = t_l + (alo * bbar - rad);
            tb_l = tb_l + dTran * dbbar;
        }
        rad = rad + dTran * Blt;
        dtau[at(l)] = t_l;
        // level l: emission terms, then the layer-mean terms of layers l and l-1
        double lev = 0.;
        if (lin != 0) {
            const bool has_dn = dn && (lin == 2 || l < L);
            lev = tup_lo + (has_dn ? dtlev[at(l)] : 0.) * rsfc;
        }
        if (lin != 2) {
            lev = lev + tb_l * dt_dtlo[l];
            if (l > 0) lev = lev + tbar_prev * dt_dtup[l - 1];
        }
        dtlev[at(l)] = lev;
        tbar_prev = tb_l;
        tup_lo = tup_hi;
        blo = bhi;
        dblo = dbhi;
        alo = ahi;
    }
    // top level L
    double lev = 0.;
    if (lin != 0) lev = tup_lo + ((dn && lin == 2) ? dtlev[at(L)] : 0.) * rsfc;
    if (lin != 2) lev = lev + tbar_prev * dt_dtup[L - 1];
    dtlev[at(L)] = lev;
    rad_o[i] = rad;
}

// Jacobians (src/clblm_src/clblm_noScattJac.f90:346-716) with every branch, one lane
// per point: the path ODs are the vertical OD od[l*n+i] times the per-layer air-mass
// factors av, ad, as (view, down, sun; pathOD_up_dn_sun's scaled paths :826-838); the
// observer at level obs (0 = surface .. L = top); up-looking (:436-494) or
// down-looking with the reflected downwelling, the surface terms, the solar term
// (sol[i] at the top, surface reflectance rs) and the upwelling of the layers below the
// observer (:496-678); then the level assembly (:681-715).  The reference's loops and
// summation order.  States passed between the loops live in `st`, five arrays of
// (L+1) x n: cumulative view transmittance, cumulative down (two-path) transmittance,
// the down and up level terms and the layer-mean terms.
__global__ void __launch_bounds__(256) k_rt_jac_paths(
    int L, int64_t n, double v1, double dv, const double *__restrict__ od, const double *__restrict__ av,
    const double *__restrict__ ad, const double *__restrict__ as, const double *__restrict__ t_lay,
    const double *__restrict__ t_lev, double emis, double tskin, int solar, const double *__restrict__ sol, double rs,
    int obs, int uplook, int lin, const double *__restrict__ dt_dtup, const double *__restrict__ dt_dtlo,
    double *__restrict__ rad_o, double *__restrict__ dtau, double *__restrict__ dtlev, double *__restrict__ dts,
    double *__restrict__ de, double *__restrict__ drs, double *__restrict__ st) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double vn = v1 + (double)i * dv;
    const int64_t S = (int64_t)(L + 1) * n;
    double *cA = st, *cB = st + S, *tdn = st + 2 * S, *tup = st + 3 * S, *tbar = st + 4 * S;
    auto at = [n, i](int k) { return (int64_t)k * n + i; };
    auto vod = [&](int j) { return od[at(j)] * av[j]; };
    auto dod = [&](int j) { return od[at(j)] * ad[j]; };
    auto blev = [&](int k, double *b, double *db) { planck_drv(vn, t_lev[k], b, db); };
    for (int j = 0; j < L; j++) {
        dtau[at(j)] = 0.;
        tbar[at(j)] = 0.;
    }
    for (int k = 0; k <= L; k++) {
        tdn[at(k)] = 0.;
        tup[at(k)] = 0.;
        dtlev[at(k)] = 0.;
    }
    double rad = 0., dEmis = 0., dTsfc = 0., dRsfc = 0., rsfc = 0.;
    int obsLay = obs;   // the first layer above the observer
    if (uplook) {
        for (int k = 0; k <= obs; k++) cB[at(k)] = 1.0;
        double cum = 0.0;
        if (obsLay < 0) obsLay = 0;
        for (int j = obsLay; j < L; j++) {
            cum = cum + vod(j);
            cB[at(j + 1)] = exp(-cum);
        }
        for (int j = L - 1; j >= obsLay; j--) {
            const double dTran = cB[at(j)] - cB[at(j + 1)];
            double bbar, dbbar, Blt;
            planck_drv(vn, t_lay[j], &bbar, &dbbar);
            if (lin != 0) {
                double h, dh, blo, dblo, bhi, dbhi;
                rlin(vod(j), &h, &dh);
                blev(j, &blo, &dblo);
                blev(j + 1, &bhi, &dbhi);
                if (lin == 2) {
                    Blt = blo * (1 - h) + bhi * h;
                    dtau[at(j)] = (cB[at(j + 1)] * Blt - rad + dTran * dh * (bhi - blo)) * av[j];
                    tdn[at(j + 1)] = tdn[at(j + 1)] + dTran * dbhi * h;
                    tdn[at(j)] = dTran * dblo * (1.0 - h);
                    tbar[at(j)] = 0.0;
                } else {
                    Blt = 2.0 * bbar * h + blo * (1.0 - 2.0 * h);
                    dtau[at(j)] = (cB[at(j + 1)] * Blt - rad + 2.0 * dTran * dh * (bbar - blo)) * av[j];
                    tdn[at(j)] = dTran * dblo * (1.0 - 2.0 * h);
                    tbar[at(j)] = dTran * 2.0 * dbbar * h;
                }
                rad = rad + dTran * Blt;
            } else {
                dtau[at(j)] = (cB[at(j + 1)] * bbar - rad) * av[j];
                rad = dTran * bbar + rad;
                tbar[at(j)] = dTran * dbbar;
            }
        }
    } else {
        // 0- view-path transmittance from the observer down to the surface
        double cumA = 0.;
        for (int k = obs; k <= L; k++) cA[at(k)] = 1.;
        for (int j = obsLay - 1; j >= 0; j--) {
            cumA = cumA + vod(j);
            cA[at(j)] = exp(-cumA);
        }
        const double totODUp = cumA, totTxUp = cA[at(0)];
        // 1- downwelling thermal radiance
        if (totTxUp > 1.e-06) {
            cB[at(0)] = totTxUp;
            double cumB = 0.;
            for (int j = 0; j < L; j++) {
                cumB = cumB + dod(j);
                cB[at(j + 1)] = exp(-(totODUp + cumB));
            }
            for (int j = L - 1; j >= 0; j--) {
                const double dTran = cB[at(j)] - cB[at(j + 1)];
                double bbar, dbbar, Blt;
                planck_drv(vn, t_lay[j], &bbar, &dbbar);
                if (lin != 0) {
                    double h, dh, blo, dblo, bhi, dbhi;
                    rlin(dod(j), &h, &dh);
                    blev(j, &blo, &dblo);
                    blev(j + 1, &bhi, &dbhi);
                    if (lin == 2) {
                        Blt = blo + (bhi - blo) * h;
                        dtau[at(j)] = (cB[at(j + 1)] * Blt - rad + dTran * dh * (bhi - blo)) * ad[j];
                        tdn[at(j + 1)] = tdn[at(j + 1)] + dTran * dbhi * h;
                        tdn[at(j)] = dTran * dblo * (1. - h);
                        tbar[at(j)] = 0.0;
                    } else {
                        Blt = 2.0 * bbar * h + blo * (1.0 - 2.0 * h);
                        dtau[at(j)] = (cB[at(j + 1)] * Blt - rad + 2.0 * dTran * dh * (bbar - blo)) * ad[j];
                        tdn[at(j)] = dTran * dblo * (1.0 - 2.0 * h);
                        tbar[at(j)] = dTran * 2.0 * dbbar * h;
                    }
                    rad = rad + dTran * Blt;
                } else {
                    dtau[at(j)] = (cB[at(j + 1)] * bbar - rad) * ad[j];
                    rad = rad + dTran * bbar;
                    tbar[at(j)] = dTran * dbbar;
                }
            }
        }
        // 2- scale by the surface reflectance
        rsfc = (1. - emis);
        for (int j = 0; j < L; j++) {
            dtau[at(j)] = dtau[at(j)] * rsfc;
            tbar[at(j)] = tbar[at(j)] * rsfc;
        }
        // 3- surface terms
        double bs, dbs;
        planck_drv(vn, tskin, &bs, &dbs);
        dEmis = totTxUp * bs - rad;
        dTsfc = emis * totTxUp * dbs;
        const double sfcEmit = emis * totTxUp * bs;
        rad = rad * rsfc + sfcEmit;
        // 4- solar
        if (solar) {
            double totODSun = 0.;
            for (int j = 0; j < L; j++) totODSun = totODSun + od[at(j)] * as[j];
            const double totTxSun2 = exp(-(totODUp + totODSun));
            dRsfc = totTxSun2 * sol[i];
            const double radsun = rs * dRsfc;
            for (int j = 0; j < L; j++) dtau[at(j)] = dtau[at(j)] - radsun * as[j];
            rad = rad + radsun;
        }
        // 5- upwelling thermal radiance of the layers below the observer
        for (int j = 0; j <= obsLay - 1; j++) {
            const double dTran = cA[at(j + 1)] - cA[at(j)];
            double bbar, dbbar, Blt;
            planck_drv(vn, t_lay[j], &bbar, &dbbar);
            if (lin != 0) {
                double h, dh, blo, dblo, bhi, dbhi;
                rlin(vod(j), &h, &dh);
                blev(j, &blo, &dblo);
                blev(j + 1, &bhi, &dbhi);
                if (lin == 2) {
                    Blt = bhi * (1 - h) + blo * h;
                    dtau[at(j)] = dtau[at(j)] + (cA[at(j)] * Blt - rad + dTran * dh * (blo - bhi)) * av[j];
                    tup[at(j + 1)] = dTran * dbhi * (1 - h);
                    tup[at(j)] = tup[at(j)] + dTran * dblo * h;
                } else {
                    Blt = 2.0 * bbar * h + bhi * (1.0 - 2.0 * h);
                    dtau[at(j)] = dtau[at(j)] + (cA[at(j)] * Blt - rad + dTran * dh * 2.0 * (bbar - bhi)) * av[j];
                    tup[at(j + 1)] = dTran * dbhi * (1.0 - 2.0 * h);
                    tbar[at(j)] = tbar[at(j)] + dTran * 2.0 * dbbar * h;
                }
                rad = rad + dTran * Blt;
            } else {
                dtau[at(j)] = dtau[at(j)] + (cA[at(j)] * bbar - rad) * av[j];
                rad = rad + dTran * bbar;
                tbar[at(j)] = tbar[at(j)] + dTran * dbbar;
            }
        }
    }
    // level temperature derivatives
    if (lin != 0) {
        if (uplook)
            for (int k = obs; k <= L; k++) dtlev[at(k)] = tdn[at(k)];
        else
            for (int k = 0; k <= L; k++) dtlev[at(k)] = tup[at(k)] + tdn[at(k)] * rsfc;
    }
    if (lin == 0 || lin == 1) {
        const int j2 = uplook ? obsLay : 0;
        for (int j = L - 1; j >= j2; j--) {
            dtlev[at(j + 1)] = dtlev[at(j + 1)] + tbar[at(j)] * dt_dtup[j];
            dtlev[at(j)] = dtlev[at(j)] + tbar[at(j)] * dt_dtlo[j];
        }
    }
    rad_o[i] = rad;
    if (dts) dts[i] = dTsfc;
    if (de) de[i] = dEmis;
    if (drs) drs[i] = dRsfc;
}

extern "C" __attribute__((visibility("hidden"))) int clblm_rt_jac_paths_impl(
    int L, int64_t n, double v1, double dv, const double *d_od, const double *d_av, const double *d_ad,
    const double *d_as, const double *d_tlay, const double *d_tlev, double emis, double tskin, int solar,
    const double *d_sol, double rs, int obs, int uplook, int lin, const double *d_dt_dtup, const double *d_dt_dtlo,
    double *d_rad, double *d_dtau, double *d_dtlev, double *d_dts, double *d_de, double *d_drs, double *d_state,
    void *stream) {
    hipLaunchKernelGGL(k_rt_jac_paths, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, L, n,
                       v1, dv, d_od, d_av, d_ad, d_as, d_tlay, d_tlev, emis, tskin, solar, d_sol, rs, obs, uplook, lin,
                       d_dt_dtup, d_dt_dtlo, d_rad, d_dtau, d_dtlev, d_dts, d_de, d_drs, d_state);
    return hipGetLastError() == hipSuccess ? 0 : CLBLM_ERR_HIP;
}

// drv_LevelTemp and drv_LevelConc (src/clblm_src/clblm_noScattJac.f90:899-1203): one
// thread per (point, quantity); quantity 0 is temperature, q = 1..nmol the molecules.
// od_var holds 3 + nmol blocks of [L][n]: the layer OD, the OD at Tave + dt, the OD at
// Wtot + dw_frac*Wtot, then each molecule alone.
__global__ void __launch_bounds__(256) k_level_jac(int L, int nmol, int64_t n, const double *__restrict__ od_var,
                                                   const double *__restrict__ dtau,
                                                   const double *__restrict__ dtlev,
                                                   const double *__restrict__ wtot, const double *__restrict__ tave,
                                                   double dt, double dw_frac, const double *__restrict__ dt_dtup,
                                                   const double *__restrict__ dt_dtlo,
                                                   const double *__restrict__ dw_dqup,
                                                   const double *__restrict__ dw_dqlo, double *__restrict__ jac_t,
                                                   double *__restrict__ jac_q) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int q = blockIdx.y;
    if (i >= n) return;
    const int64_t LN = (int64_t)L * n;
    double lo = 0., up = 0.;
    if (q == 0) {
        const double *od = od_var, *odp = od_var + LN, *odw = od_var + 2 * LN;
        for (int l = 0; l < L; l++) {
            const int64_t j = (int64_t)l * n + i;
            const double base = od[j];
            double dTau_dT = (odp[j] - base) / dt;
            const double dTau_dW = (odw[j] - base) / (dw_frac * wtot[l]);
            dTau_dT = dTau_dT - dTau_dW * (wtot[l] / tave[l]);
            const double dRaddT = dtau[j] * dTau_dT;
            up = dRaddT * dt_dtup[l];
            lo = lo + dRaddT * dt_dtlo[l] + dtlev[j];
            jac_t[j] = lo;
            lo = up;
        }
        jac_t[LN + i] = lo + dtlev[LN + i];
    } else {
        const int m = q - 1;
        const double *odm = od_var + (3 + (int64_t)m) * LN;
        double *jq = jac_q + (int64_t)m * (L + 1) * n;
        for (int l = 0; l < L; l++) {
            const int64_t j = (int64_t)l * n + i;
            const double dRad_dW = dtau[j] * odm[j];
            up = dRad_dW * dw_dqup[(int64_t)m * L + l];
            lo = lo + dRad_dW * dw_dqlo[(int64_t)m * L + l];
            jq[j] = lo;
            lo = up;
        }
        jq[LN + i] = lo;
    }
}

thread_local std::string g_rt_err;
}  // namespace

extern "C" __attribute__((visibility("hidden"))) int clblm_rt_up_impl(int nlay, int64_t npts, double v1, double dv, const double *d_od,
                                const double *d_tlay, const double *d_tlev, double emis, double tskin, int lin,
                                double sec_dn, double *d_rad, double *d_dr_dod, double *d_dr_dt,
                                double *d_dr_dtskin, double *d_dr_demis, double *d_state, void *stream) {
    if (nlay <= 0 || npts <= 0 || !d_od || !d_tlay || !d_tlev || !d_rad) return CLBLM_ERR_ARG;
    if (lin < 0 || lin > 2) return CLBLM_ERR_ARG;
    if (d_dr_dod && (!d_dr_dt || !d_state)) return CLBLM_ERR_ARG;
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(k_rt_up, dim3((unsigned)((npts + 255) / 256)), dim3(256), 0, st, nlay, npts, v1, dv, d_od,
                       d_tlay, d_tlev, emis, tskin, lin, sec_dn, d_rad, d_dr_dod, d_dr_dt, d_dr_dtskin, d_dr_demis,
                       d_state);
    return hipGetLastError() == hipSuccess ? 0 : CLBLM_ERR_HIP;
}

// level2layer (src/clblm_src/clblm_noScattJac.f90:1208-1279) on the host: layer-mean
// temperature and ln(layer amount) derivatives with respect to the bounding levels,
// for a profile with exponential density and linear temperature between levels.
// q_lev[m*(nlay+1)+k] is the number density of molecule m at level k.
extern "C" __attribute__((visibility("hidden"))) void clblm_level2layer_impl(
    int nlay, int nmol, const double *p_lev, const double *t_lev, const double *t_lay, const double *q_lev,
    double *dt_dtup, double *dt_dtlo, double *dw_dqup, double *dw_dqlo) {
    const double BOLTZ_HPA = 1.3806503E-19;   // hPa cm3 / K (clblm_noScattJac.f90:1240)
    for (int l = 0; l < nlay; l++) {
        const double pup = p_lev[l + 1], plo = p_lev[l], tup = t_lev[l + 1], tlo = t_lev[l], tbar = t_lay[l];
        const double rhoU = pup / (tup * BOLTZ_HPA), rhoL = plo / (tlo * BOLTZ_HPA);
        const double alpha = rhoU / rhoL;
        const double alphaT = -(tup - tlo) / std::log(alpha);
        dt_dtlo[l] = ((tbar - alphaT) / tlo) * (rhoL / (rhoL - rhoU)) + (1.0 - alphaT / tlo) / std::log(alpha);
        dt_dtup[l] = ((tbar - alphaT) / tup) * (-rhoU / (rhoL - rhoU)) - (1.0 - alphaT / tup) / std::log(alpha);
        for (int m = 0; m < nmol; m++) {
            const double dup = q_lev[(int64_t)m * (nlay + 1) + l + 1], dlo = q_lev[(int64_t)m * (nlay + 1) + l];
            double &wu = dw_dqup[(int64_t)m * nlay + l], &wl = dw_dqlo[(int64_t)m * nlay + l];
            if (dlo != 0.0) {
                const double ratU = dup / rhoU, ratL = dlo / rhoL;
                wl = (ratL / (ratL - alpha * ratU)) + 1.0 / std::log(alpha * ratU / ratL);
                wu = ((-alpha * ratU) / (ratL - alpha * ratU)) - 1.0 / std::log(alpha * ratU / ratL);
            } else {
                wl = 0.0;
                wu = 0.0;
            }
        }
    }
}

extern "C" __attribute__((visibility("hidden"))) int clblm_rt_jac_impl(
    int nlay, int64_t npts, double v1, double dv, const double *d_od, const double *d_tlay, const double *d_tlev,
    const double *d_dt_dtup, const double *d_dt_dtlo, double emis, double tskin, int lin, double sec_dn,
    double *d_rad, double *d_dr_dtau, double *d_dr_dtlev, double *d_dr_dtskin, double *d_dr_demis, double *d_state,
    void *stream) {
    hipStream_t st = (hipStream_t)stream;
    const int64_t n = npts;
    double *sa = d_state, *sb = d_state + (int64_t)(nlay + 1) * n, *sbar = d_state + 2 * (int64_t)(nlay + 1) * n;
    hipLaunchKernelGGL(k_rt_jac, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, nlay, n, v1, dv, d_od, d_tlay,
                       d_tlev, emis, tskin, lin, sec_dn, d_dt_dtup, d_dt_dtlo, d_rad, d_dr_dtau, d_dr_dtlev,
                       d_dr_dtskin, d_dr_demis, sa, sb, sbar);
    return hipGetLastError() == hipSuccess ? 0 : CLBLM_ERR_HIP;
}

// Scratch of clblm_rt_jac_noalloc: k_rt_jac's per-level states (sa, sb, sbar) in an
// array the library owns for the life of the process, one copy on each GPU it runs
// on, so that the entry point allocates nothing.  Host code uses the per-device
// address that hipGetSymbolAddress returns, never the array's name (which in host
// code is a host-side stand-in, not a device address).  Every state a lane reads is
// written by that lane earlier in the same launch, so nothing needs clearing.
static __device__ __align__(256) double scratch[CLBLM_RT_JAC_NOALLOC_BYTES / sizeof(double)];

namespace {
constexpr int SCRATCH_MAX_DEV = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[SCRATCH_MAX_DEV] = {};
std::mutex g_scratch_call_mu[SCRATCH_MAX_DEV];
}  // namespace

// Once per device, with that device current (clblm_ctx_create): the device address of
// `scratch` on it.
extern "C" __attribute__((visibility("hidden"))) int clblm_rt_scratch_init(int device) {
    if (device < 0 || device >= SCRATCH_MAX_DEV) return CLBLM_ERR_ARG;
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (g_scratch_ptr[device]) return 0;
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return CLBLM_ERR_HIP;
    g_scratch_ptr[device] = (double *)p;
    return 0;
}

// clblm_rt_jac with its states in `scratch`: the caller must have checked the size.
// Calls on one device are serialized by a per-device lock held until the kernel has
// finished, because they share the one array.
extern "C" __attribute__((visibility("hidden"))) int clblm_rt_jac_scratch(
    int device, int nlay, int64_t npts, double v1, double dv, const double *d_od, const double *d_tlay,
    const double *d_tlev, const double *d_dt_dtup, const double *d_dt_dtlo, double emis, double tskin, int lin,
    double sec_dn, double *d_rad, double *d_dr_dtau, double *d_dr_dtlev, double *d_dr_dtskin, double *d_dr_demis,
    void *stream) {
    if (device < 0 || device >= SCRATCH_MAX_DEV) return CLBLM_ERR_ARG;
    double *ptr;
    {
        std::lock_guard<std::mutex> g(g_scratch_mu);
        ptr = g_scratch_ptr[device];
    }
    if (!ptr) return CLBLM_ERR_HIP;
    std::lock_guard<std::mutex> call(g_scratch_call_mu[device]);
    int rc = clblm_rt_jac_impl(nlay, npts, v1, dv, d_od, d_tlay, d_tlev, d_dt_dtup, d_dt_dtlo, emis, tskin, lin,
                               sec_dn, d_rad, d_dr_dtau, d_dr_dtlev, d_dr_dtskin, d_dr_demis, ptr, stream);
    if (rc) return rc;
    return hipStreamSynchronize((hipStream_t)stream) == hipSuccess ? 0 : CLBLM_ERR_HIP;
}

extern "C" __attribute__((visibility("hidden"))) int clblm_level_jac_impl(
    int nlay, int nmol, int64_t npts, const double *d_od_var, const double *d_dr_dtau, const double *d_dr_dtlev,
    const double *d_wtot, const double *d_tlay, double dt, double dw_frac, const double *d_dt_dtup,
    const double *d_dt_dtlo, const double *d_dw_dqup, const double *d_dw_dqlo, double *d_jac_t, double *d_jac_q,
    void *stream) {
    hipStream_t st = (hipStream_t)stream;
    const int nq = 1 + (d_jac_q ? nmol : 0);   // quantity 0 (temperature) always, then the molecules
    hipLaunchKernelGGL(k_level_jac, dim3((unsigned)((npts + 255) / 256), (unsigned)nq), dim3(256), 0, st, nlay, nmol,
                       npts, d_od_var, d_dr_dtau, d_dr_dtlev, d_wtot, d_tlay, dt, dw_frac, d_dt_dtup, d_dt_dtlo,
                       d_dw_dqup, d_dw_dqlo, d_jac_t, d_jac_q);
    return hipGetLastError() == hipSuccess ? 0 : CLBLM_ERR_HIP;
}

// ---------------------------------------------------------------------------------
// Merge-up on the layers' own grids: mode2_lookDn_mergeUp / layerMerge_mode2
// (clblm_noScattRT.f90:1287-1296, 1700-1794), where every layer's OD comes on its own
// DV and interpToFinestGrid (clblm_Spectrum.f90:733-795) brings the merged radiance and
// reflectance and the layer's transmittances and emissions onto the finest grid of
// the set before each merge.  All grids start at V1.

// interpSpectrum of X[0..nin) on v1 + i dvin onto o[0..nout) on v1 + i dvout.
__global__ void __launch_bounds__(256) k_interp_spec(const double *__restrict__ X, int64_t nin, double v1,
                                                     double dvin, double dvout, double *__restrict__ o,
                                                     int64_t nout, int64_t nunits) {
    const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= nunits) return;
    clblm_interp::unit(X - 1, nin, v1, dvin, dvout, o - 1, nout, u);
}

// The surface (calcSurfThmEmis_array, clblm_Scene.f90): R = B(Tskin) eps, rho = 1 - eps.
__global__ void __launch_bounds__(256) k_grid_surface(double *__restrict__ r, double *__restrict__ rho, int64_t n,
                                                      double v1, double dv, double emis, double tskin) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double v = v1 + (double)i * dv;
    r[i] = planck_x(v, tskin / RADCN2) * emis;
    rho[i] = 1.0 - emis;
}

// One layer's view and down transmittances and its upward and downward emission
// (layerEmis, clblm_EMLAY.f90:104-178) at point i of its own grid.
__device__ inline void grid_layer_point(double v, double oda, double sec, double tbot, double ttop, double tave,
                                        int lin, double *txa, double *txb, double *emup, double *emdn) {
    const double odb = oda * sec;
    *txa = exp(-oda);
    *txb = exp(-odb);
    const double bbar = planck_x(v, tave / RADCN2);
    if (lin == 0) {
        *emdn = (1 - *txb) * bbar;
        *emup = *emdn;            // lyrEmUp = lyrEmDn (:1752)
    } else if (lin == 1) {
        double h, dh;
        rlin(odb, &h, &dh);
        const double ba = planck_x(v, tbot / RADCN2);
        *emdn = (1 - *txb) * (ba + 2 * (bbar - ba) * h);
        rlin(oda, &h, &dh);
        const double bu = planck_x(v, ttop / RADCN2);
        *emup = (1 - *txa) * (bu + 2 * (bbar - bu) * h);
    } else {
        double h, dh;
        rlin(odb, &h, &dh);
        const double ba = planck_x(v, tbot / RADCN2), bb = planck_x(v, ttop / RADCN2);
        *emdn = (1 - *txb) * (ba + (bb - ba) * h);
        rlin(oda, &h, &dh);
        *emup = (1 - *txa) * (bb + (ba - bb) * h);
    }
}

// The layer's quantities on its grid, merged at once (the merged arrays are on it).
__global__ void __launch_bounds__(256) k_grid_layer_merge(const double *__restrict__ od, int64_t n, double v1,
                                                          double dv, double sec, double tbot, double ttop,
                                                          double tave, int lin, double *__restrict__ r,
                                                          double *__restrict__ rho) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double txa, txb, emup, emdn;
    grid_layer_point(v1 + (double)i * dv, od[i], sec, tbot, ttop, tave, lin, &txa, &txb, &emup, &emdn);
    r[i] = emup + txa * (r[i] + emdn * rho[i]);
    rho[i] = txa * txb * rho[i];
}

// The layer's quantities on its grid, stored for interpolation onto a finer one.
__global__ void __launch_bounds__(256) k_grid_layer(const double *__restrict__ od, int64_t n, double v1, double dv,
                                                    double sec, double tbot, double ttop, double tave, int lin,
                                                    double *__restrict__ q) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    grid_layer_point(v1 + (double)i * dv, od[i], sec, tbot, ttop, tave, lin, q + i, q + n + i, q + 2 * n + i,
                     q + 3 * n + i);
}

// R = EmUp + TxA (R + EmDn rho), rho = TxA TxB rho (:1773-1786) with the layer's
// quantities already on the merged grid (q: TxA, TxB, EmUp, EmDn, n each).
__global__ void __launch_bounds__(256) k_grid_merge(const double *__restrict__ q, int64_t n, double *__restrict__ r,
                                                    double *__restrict__ rho) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double txa = q[i], txb = q[n + i], emup = q[2 * n + i], emdn = q[3 * n + i];
    r[i] = emup + txa * (r[i] + emdn * rho[i]);
    rho[i] = txa * txb * rho[i];
}

namespace {
inline unsigned blocks_of(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }
void launch_interp(const double *X, int64_t nin, double v1, double dvin, double dvout, double *o, int64_t nout,
                   hipStream_t st) {
    const int64_t nu = clblm_interp::units(dvin, dvout, nout);
    hipLaunchKernelGGL(k_interp_spec, dim3(blocks_of(nu, 256)), dim3(256), 0, st, X, nin, v1, dvin, dvout, o, nout,
                       nu);
}
}  // namespace

// interpSpectrum of one device spectrum onto another grid from the same V1.
extern "C" __attribute__((visibility("hidden"))) int clblm_interp_spectrum_impl(const double *d_in, int64_t nin,
                                                                                double v1, double dvin,
                                                                                double *d_out, int64_t nout,
                                                                                double dvout, void *stream) {
    launch_interp(d_in, nin, v1, dvin, dvout, d_out, nout, (hipStream_t)stream);
    return hipGetLastError() == hipSuccess ? 0 : CLBLM_ERR_HIP;
}

// The finest grid of a set (interpToFinestGrid: the first spectrum of smallest DV
// gives DV and NLIM).
extern "C" __attribute__((visibility("hidden"))) int clblm_finest_grid(int nspec, const double *dv,
                                                                       const int64_t *nlim) {
    int loc = 0;
    for (int k = 1; k < nspec; k++)
        if (dv[k] < dv[loc]) loc = k;
    return loc;
}

// The grids the merge passes through: the surface grid, then before layer l the
// finest of the merged grid and the layer's.  Returns the final grid.
extern "C" __attribute__((visibility("hidden"))) void clblm_rt_grids_final(int nlay, double v1, const double *dv,
                                                                          const int64_t *nlim, double surf_dv,
                                                                          int64_t surf_nlim, double *dv_out,
                                                                          int64_t *nlim_out, int64_t *max_n) {
    double cdv = surf_dv;
    int64_t cn = surf_nlim, mx = surf_nlim;
    for (int l = 0; l < nlay; l++) {
        mx = std::max(mx, nlim[l]);
        if (dv[l] < cdv && !clblm_interp::same_grid(v1, dv[l], nlim[l], cdv, cn)) {   // the layer's TxA leads the set
            cdv = dv[l];
            cn = nlim[l];
        }
    }
    *dv_out = cdv;
    *nlim_out = cn;
    *max_n = mx;
}

// scratch: 12 * max_n doubles (merged R and rho, their refined copies, the layer's
// four quantities on its grid and on the merged grid).
extern "C" __attribute__((visibility("hidden"))) int clblm_rt_up_grids_impl(
    int nlay, double v1, const double *dv, const int64_t *nlim, const int64_t *od_off, const double *d_od,
    const double *t_lay, const double *t_lev, double emis, double tskin, double surf_dv, int64_t surf_nlim, int lin,
    double sec, double *scratch, int64_t max_n, double *d_rad, void *stream) {
    hipStream_t st = (hipStream_t)stream;
    double *r = scratch, *rho = scratch + max_n, *r2 = scratch + 2 * max_n, *rho2 = scratch + 3 * max_n;
    double *q = scratch + 4 * max_n, *qi = scratch + 8 * max_n;
    double cdv = surf_dv;
    int64_t cn = surf_nlim;
    hipLaunchKernelGGL(k_grid_surface, dim3(blocks_of(cn, 256)), dim3(256), 0, st, r, rho, cn, v1, cdv, emis, tskin);
    for (int l = 0; l < nlay; l++) {
        const double ldv = dv[l];
        const int64_t ln = nlim[l];
        const double *od = d_od + od_off[l];
        const double tbot = t_lev[l], ttop = t_lev[l + 1], tave = t_lay[l];
        if (clblm_interp::same_grid(v1, ldv, ln, cdv, cn)) {
            // one grid: no interpolation (the merged arrays keep their DV)
            hipLaunchKernelGGL(k_grid_layer_merge, dim3(blocks_of(cn, 256)), dim3(256), 0, st, od, cn, v1, ldv, sec,
                               tbot, ttop, tave, lin, r, rho);
        } else if (ldv < cdv) {
            // the layer's grid is the finest: the merged arrays are interpolated onto it
            launch_interp(r, cn, v1, cdv, ldv, r2, ln, st);
            launch_interp(rho, cn, v1, cdv, ldv, rho2, ln, st);
            std::swap(r, r2);
            std::swap(rho, rho2);
            cdv = ldv;
            cn = ln;
            hipLaunchKernelGGL(k_grid_layer_merge, dim3(blocks_of(cn, 256)), dim3(256), 0, st, od, cn, v1, cdv, sec,
                               tbot, ttop, tave, lin, r, rho);
        } else {
            // the merged grid is the finest (or as fine, with another length): the
            // layer's four quantities are interpolated onto it
            hipLaunchKernelGGL(k_grid_layer, dim3(blocks_of(ln, 256)), dim3(256), 0, st, od, ln, v1, ldv, sec, tbot,
                               ttop, tave, lin, q);
            for (int k = 0; k < 4; k++) launch_interp(q + k * ln, ln, v1, ldv, cdv, qi + k * cn, cn, st);
            hipLaunchKernelGGL(k_grid_merge, dim3(blocks_of(cn, 256)), dim3(256), 0, st, qi, cn, r, rho);
        }
    }
    if (hipMemcpyAsync(d_rad, r, sizeof(double) * (size_t)cn, hipMemcpyDeviceToDevice, st) != hipSuccess)
        return CLBLM_ERR_HIP;
    return hipGetLastError() == hipSuccess ? 0 : CLBLM_ERR_HIP;
}


// ---------------------------------------------------------------------------------
// Merge-down on the layers' own grids: mode1_lookDn_mergeDn / layerMerge_mode1
// (clblm_noScattRT.f90:42-842) with addThermalBoundary and addSolarBoundary
// (:1945-2148).  From the top layer down: the down radiance gains the layer's
// downward emission under the merged down-path radiance (mrgRadDn = EmDn + mrgRadDn
// TxB), the up radiance gains the layer's upward emission under the merged view
// transmittance (mrgRadUp = mrgRadUp + mrgTxA EmUp), then mrgTxA and the solar
// transmittance take the layer's factors.  Each step is preceded by the
// reference's interpToFinestGrid over the spectra it combines; the grids involved
// depend on the DVs alone, so the host walks them (also to size the output) and the
// device does the arithmetic.

// The layer's view and down transmittances, emissions and solar transmittance on its
// grid (the up emission only for linInTau 1 and 2: for 0 it is the down one).
__global__ void __launch_bounds__(256) k_m1_layer(const double *__restrict__ od_up, const double *__restrict__ od_dn,
                                                  const double *__restrict__ od_sun, int64_t n, double v1, double dv,
                                                  double fu, double fd, double fs, double tbot, double ttop,
                                                  double tave, int lin, int below, int thermal, int solar,
                                                  double *__restrict__ txa, double *__restrict__ txb,
                                                  double *__restrict__ emdn, double *__restrict__ emup,
                                                  double *__restrict__ sol) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double v = v1 + (double)i * dv;
    double oda = 0., ta = 1.;
    if (below) {
        oda = od_up[i] * fu;
        ta = exp(-oda);
        txa[i] = ta;
    }
    if (thermal) {
        const double odb = od_dn[i] * fd;
        const double tb = exp(-odb);
        txb[i] = tb;
        const double bbar = planck_x(v, tave / RADCN2);
        double h, dh;
        if (lin == 0) {
            emdn[i] = (1 - tb) * bbar;
        } else if (lin == 1) {
            rlin(odb, &h, &dh);
            const double ba = planck_x(v, tbot / RADCN2);
            emdn[i] = (1 - tb) * (ba + 2 * (bbar - ba) * h);
            if (below) {
                rlin(oda, &h, &dh);
                const double bu = planck_x(v, ttop / RADCN2);
                emup[i] = (1 - ta) * (bu + 2 * (bbar - bu) * h);
            }
        } else {
            rlin(odb, &h, &dh);
            const double ba = planck_x(v, tbot / RADCN2), bb = planck_x(v, ttop / RADCN2);
            emdn[i] = (1 - tb) * (ba + (bb - ba) * h);
            if (below) {
                rlin(oda, &h, &dh);
                emup[i] = (1 - ta) * (bb + (ba - bb) * h);
            }
        }
    }
    if (solar) sol[i] = exp(-(od_sun[i] * fs));
}

__global__ void __launch_bounds__(256) k_m1_fill(double *__restrict__ x, int64_t n, double v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = v;
}

// mrgRadDn = EmDn + mrgRadDn TxB
__global__ void __launch_bounds__(256) k_m1_dn(const double *__restrict__ em, const double *__restrict__ tb,
                                               double *__restrict__ r, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) r[i] = em[i] + r[i] * tb[i];
}

// mrgRadUp = mrgRadUp + mrgTxA EmUp
__global__ void __launch_bounds__(256) k_m1_up(const double *__restrict__ ta, const double *__restrict__ em,
                                               double *__restrict__ r, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) r[i] = r[i] + ta[i] * em[i];
}

// a = a b
__global__ void __launch_bounds__(256) k_m1_mul(double *__restrict__ a, const double *__restrict__ b, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = a[i] * b[i];
}

// addThermalBoundary: emission plus the reflected down radiance (refl) or emission only.
__global__ void __launch_bounds__(256) k_m1_surface(const double *__restrict__ up, const double *__restrict__ ta,
                                                    const double *__restrict__ dn, int64_t n, double v1, double dv,
                                                    double emis, double tskin, int refl, double *__restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double v = v1 + (double)i * dv;
    const double srad = planck_x(v, tskin / RADCN2) * emis;
    if (refl) out[i] = up[i] + ta[i] * (srad + (1.0 - emis) * dn[i]);
    else out[i] = up[i] + ta[i] * srad;
}

// addSolarBoundary: tot = tot + TxA refl Tsun S0
__global__ void __launch_bounds__(256) k_m1_solar(double *__restrict__ tot, const double *__restrict__ ta,
                                                  const double *__restrict__ ts, const double *__restrict__ s0,
                                                  int64_t n, double refl) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) tot[i] = tot[i] + ta[i] * refl * ts[i] * s0[i];
}

namespace {
struct M1Spec {
    double *d = nullptr;   // device values (nullptr in the sizing walk)
    double dv = 0.;
    int64_t n = 0;
};

// The walk: run == false follows the grids only.
struct M1Walk {
    double v1 = 0.;
    bool run = false;
    hipStream_t st = nullptr;
    std::vector<double *> free_bufs;
    double *spare = nullptr;   // handed out if the pool ran dry (then the call fails)
    bool dry = false;
    double *take() {
        if (!run) return nullptr;
        if (free_bufs.empty()) {
            dry = true;
            return spare;
        }
        double *p = free_bufs.back();
        free_bufs.pop_back();
        return p;
    }
    void give(double *p) {
        if (run && p) free_bufs.push_back(p);
    }
    // interpToFinestGrid over s[0..ns)
    void finest(int ns, M1Spec **s) {
        int loc = 0;
        for (int k = 1; k < ns; k++)
            if (s[k]->dv < s[loc]->dv) loc = k;
        const double dvl = s[loc]->dv;
        const int64_t nl = s[loc]->n;
        for (int k = 0; k < ns; k++) {
            if (k == loc || clblm_interp::same_grid(v1, s[k]->dv, s[k]->n, dvl, nl)) continue;
            double *o = take();
            if (run) launch_interp(s[k]->d, s[k]->n, v1, s[k]->dv, dvl, o, nl, st);
            give(s[k]->d);
            s[k]->d = o;
            s[k]->dv = dvl;
            s[k]->n = nl;
        }
    }
};
}  // namespace

// scratch: 12 * max_n doubles, max_n >= every layer's points, sol_n and 2.
extern "C" __attribute__((visibility("hidden"))) int clblm_rt_mode1_impl(
    int nlay, double v1, double v2, const double *dv, const int64_t *nlim, const int64_t *od_off,
    const double *d_od_up, const double *d_od_dn, const double *d_od_sun, const double *fac_up, const double *fac_dn,
    const double *fac_sun, const double *t_lay, const double *t_lev, int obs_level, int lin, int thermal, int solar,
    double emis, double tskin, double refl, const double *d_sol, int64_t sol_n, double sol_dv, double *scratch,
    int64_t max_n, double *d_rad, double *d_tx, double *dv_out, int64_t *n_out, void *stream) {
    M1Walk W;
    W.v1 = v1;
    W.run = d_rad != nullptr;
    W.st = (hipStream_t)stream;
    if (W.run) {
        for (int k = 11; k >= 0; k--) W.free_bufs.push_back(scratch + (size_t)k * max_n);
        W.spare = scratch;
    }
    auto fill = [&](M1Spec &s, double v) {
        s.d = W.take();
        if (W.run) hipLaunchKernelGGL(k_m1_fill, dim3(blocks_of(s.n, 256)), dim3(256), 0, W.st, s.d, s.n, v);
    };
    M1Spec TxA{nullptr, v2 - v1, 2}, RadUp{nullptr, v2 - v1, 2}, RadDn{nullptr, v2 - v1, 2}, SolTx{nullptr, v2 - v1, 2};
    fill(TxA, 1.);
    fill(RadUp, 0.);
    if (thermal) fill(RadDn, 0.);
    if (solar) fill(SolTx, 1.);
    for (int l = nlay - 1; l >= 0; l--) {
        const int below = l < obs_level;
        const double ldv = dv[l];
        const int64_t ln = nlim[l];
        M1Spec lTxA{nullptr, ldv, ln}, lTxB{nullptr, ldv, ln}, lEmDn{nullptr, ldv, ln}, lEmUp{nullptr, ldv, ln},
            lSol{nullptr, ldv, ln};
        if (below) lTxA.d = W.take();
        if (thermal) {
            lTxB.d = W.take();
            lEmDn.d = W.take();
            if (below && lin != 0) lEmUp.d = W.take();
        }
        if (solar) lSol.d = W.take();
        if (W.run)
            hipLaunchKernelGGL(k_m1_layer, dim3(blocks_of(ln, 256)), dim3(256), 0, W.st, d_od_up + od_off[l],
                               d_od_dn + od_off[l], d_od_sun + od_off[l], ln, v1, ldv, fac_up[l], fac_dn[l], fac_sun[l],
                               t_lev[l], t_lev[l + 1], t_lay[l], lin, below, thermal, solar, lTxA.d, lTxB.d, lEmDn.d,
                               lEmUp.d, lSol.d);
        if (thermal) {
            M1Spec *a[3] = {&lEmDn, &lTxB, &RadDn};
            W.finest(3, a);
            if (W.run)
                hipLaunchKernelGGL(k_m1_dn, dim3(blocks_of(RadDn.n, 256)), dim3(256), 0, W.st, lEmDn.d, lTxB.d,
                                   RadDn.d, RadDn.n);
            if (below) {
                if (lin == 0) {   // lyrEmUp = lyrEmDn on its grid after the merge above
                    lEmUp = M1Spec{W.take(), lEmDn.dv, lEmDn.n};
                    if (W.run &&
                        hipMemcpyAsync(lEmUp.d, lEmDn.d, sizeof(double) * (size_t)lEmDn.n, hipMemcpyDeviceToDevice,
                                       W.st) != hipSuccess)
                        return CLBLM_ERR_HIP;
                }
                M1Spec *b[3] = {&lEmUp, &TxA, &RadUp};
                W.finest(3, b);
                if (W.run)
                    hipLaunchKernelGGL(k_m1_up, dim3(blocks_of(RadUp.n, 256)), dim3(256), 0, W.st, TxA.d, lEmUp.d,
                                       RadUp.d, RadUp.n);
            }
        }
        if (below) {
            M1Spec *c[2] = {&lTxA, &TxA};
            W.finest(2, c);
            if (W.run)
                hipLaunchKernelGGL(k_m1_mul, dim3(blocks_of(TxA.n, 256)), dim3(256), 0, W.st, TxA.d, lTxA.d, TxA.n);
        }
        if (solar) {
            M1Spec *d[2] = {&lSol, &SolTx};
            W.finest(2, d);
            if (W.run)
                hipLaunchKernelGGL(k_m1_mul, dim3(blocks_of(SolTx.n, 256)), dim3(256), 0, W.st, SolTx.d, lSol.d,
                                   SolTx.n);
        }
        W.give(lTxA.d);
        W.give(lTxB.d);
        W.give(lEmDn.d);
        W.give(lEmUp.d);
        W.give(lSol.d);
    }
    // addThermalBoundary (flux off)
    const bool reflecting = emis < 1.;
    M1Spec Tot;
    if (reflecting) {
        M1Spec *a[3] = {&TxA, &RadUp, &RadDn};
        W.finest(3, a);
    } else {
        M1Spec *a[2] = {&RadUp, &TxA};
        W.finest(2, a);
    }
    Tot = M1Spec{W.take(), RadUp.dv, RadUp.n};
    if (W.run)
        hipLaunchKernelGGL(k_m1_surface, dim3(blocks_of(Tot.n, 256)), dim3(256), 0, W.st, RadUp.d, TxA.d, RadDn.d,
                           Tot.n, v1, Tot.dv, emis, tskin, reflecting ? 1 : 0, Tot.d);
    if (solar) {   // addSolarBoundary
        M1Spec S0{nullptr, sol_dv, sol_n};
        if (W.run) {
            S0.d = W.take();
            if (hipMemcpyAsync(S0.d, d_sol, sizeof(double) * (size_t)sol_n, hipMemcpyDeviceToDevice, W.st) !=
                hipSuccess)
                return CLBLM_ERR_HIP;
        }
        M1Spec *a[4] = {&TxA, &Tot, &SolTx, &S0};
        W.finest(4, a);
        if (W.run)
            hipLaunchKernelGGL(k_m1_solar, dim3(blocks_of(Tot.n, 256)), dim3(256), 0, W.st, Tot.d, TxA.d, SolTx.d,
                               S0.d, Tot.n, refl);
    }
    *dv_out = Tot.dv;
    *n_out = Tot.n;
    if (!W.run) return 0;
    if (W.dry) return CLBLM_ERR_CAPACITY;   // more spectra alive than the scratch holds
    if (hipMemcpyAsync(d_rad, Tot.d, sizeof(double) * (size_t)Tot.n, hipMemcpyDeviceToDevice, W.st) != hipSuccess)
        return CLBLM_ERR_HIP;
    if (d_tx && hipMemcpyAsync(d_tx, TxA.d, sizeof(double) * (size_t)TxA.n, hipMemcpyDeviceToDevice, W.st) != hipSuccess)
        return CLBLM_ERR_HIP;
    return hipGetLastError() == hipSuccess ? 0 : CLBLM_ERR_HIP;
}

// ---------------------------------------------------------------------------------
// Fluxes: the flux branches of mode1_lookDn_mergeDn (clblm_noScattRT.f90:133-422,
// 482-507) and mode2_lookDn_mergeUp (:1214-1450), every layer on one grid.  Per
// quadrature angle the ODs are scaled by its secant; the down radiance is merged from
// the top, the surface term formed, the up radiance merged from the surface.  The
// points are taken in blocks of whole output groups: k_flux_cols writes every level's
// down and up radiance of every angle for the block's points (one lane per point),
// k_flux_groups sums each (level, group) over its points in order (one lane each) and
// weights the angles, k_flux_surface sums the surface upwelling flux of each group.

struct FluxArgs {
    int nlay, nang, lin, lambertian;
    double v1, dv, emis, tskin, factor, pi;
    double sec[3], w[3];
    int64_t npts, p0, nblk;   // the block: points p0 .. p0 + nblk - 1
    int outinrat, nout, g0, ngrp;
};

__global__ void __launch_bounds__(256) k_flux_cols(FluxArgs A, const double *__restrict__ od,
                                                   const double *__restrict__ t_lay, const double *__restrict__ t_lev,
                                                   double *__restrict__ dn, double *__restrict__ up,
                                                   double *__restrict__ sfc) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= A.nblk) return;
    const int64_t i = A.p0 + k;
    const double v = A.v1 + (double)i * A.dv;
    const int64_t LB = (int64_t)A.nlay * A.nblk;
    double r0[3] = {0., 0., 0.};
    for (int a = 0; a < A.nang; a++) {
        double r = 0.;
        for (int l = A.nlay - 1; l >= 0; l--) {
            const double odb = od[(int64_t)l * A.npts + i] * A.sec[a];
            double txa, txb, emup, emdn;   // on the down path alone: txa = txb
            grid_layer_point(v, odb, 1., t_lev[l], t_lev[l + 1], t_lay[l], A.lin, &txa, &txb, &emup, &emdn);
            r = emdn + r * txb;
            dn[a * LB + (int64_t)l * A.nblk + k] = r;
        }
        r0[a] = r;
    }
    double d;   // the downwelling flux density at the surface (:396-403)
    if (A.nang == 1) d = A.factor * A.w[0] * r0[0];
    else if (A.nang == 2) d = A.factor * (A.w[0] * r0[0] + A.w[1] * r0[1]);
    else d = A.factor * (A.w[0] * r0[0] + A.w[1] * r0[1] + A.w[2] * r0[2]);
    const double srf = planck_x(v, A.tskin / RADCN2) * A.emis;
    sfc[k] = srf * A.dv * 1.E04 * A.pi + (1. - A.emis) * d;
    for (int a = 0; a < A.nang; a++) {
        // the lambertian form for points 1 .. nlim - 1 (:488-495), the angle's own
        // reflected radiance otherwise (addThermalBoundary's flux branch :2020-2024)
        double r = (A.lambertian && i < A.npts - 1) ? srf + ((1. - A.emis) * (d / (A.pi * A.dv * 1.E04)))
                                                    : srf + (1. - A.emis) * r0[a];
        for (int l = 0; l < A.nlay; l++) {
            const double oda = od[(int64_t)l * A.npts + i] * A.sec[a];
            double txa, txb, emup, emdn;
            grid_layer_point(v, oda, 1., t_lev[l], t_lev[l + 1], t_lay[l], A.lin, &txa, &txb, &emup, &emdn);
            r = emup + txa * (r + emdn * 0.);
            up[a * LB + (int64_t)l * A.nblk + k] = r;
        }
    }
}

// (pass, layer, group): pass 0 the down radiance of layer l (the flux at level l),
// pass 1 its up radiance (the flux at level l + 1).
__global__ void __launch_bounds__(256) k_flux_groups(FluxArgs A, const double *__restrict__ dn,
                                                     const double *__restrict__ up, double *__restrict__ fdn,
                                                     double *__restrict__ fup) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t per = (int64_t)A.nlay * A.ngrp;
    if (t >= 2 * per) return;
    const int pass = (int)(t / per);
    const int l = (int)((t % per) / A.ngrp);
    const int gl = (int)(t % A.ngrp);
    const int g = A.g0 + gl;
    if (g >= A.nout) return;
    const double *R = pass == 0 ? dn : up;
    const int64_t LB = (int64_t)A.nlay * A.nblk;
    const int64_t k0 = (int64_t)gl * A.outinrat;
    double S[3] = {0., 0., 0.};
    for (int a = 0; a < A.nang; a++) {
        const double *r = R + a * LB + (int64_t)l * A.nblk;
        double s = 0.;
        for (int64_t k = k0; k < k0 + A.outinrat && k < A.nblk; k++) s = s + r[k];
        S[a] = s;
    }
    double f;
    if (A.nang == 1) f = A.w[0] * S[0] * A.factor;
    else if (A.nang == 2) f = (A.w[0] * S[0] + A.w[1] * S[1]) * A.factor;
    else f = (A.w[0] * S[0] + A.w[1] * S[1] + A.w[2] * S[2]) * A.factor;
    if (pass == 0) fdn[(int64_t)l * A.nout + g] = f;
    else fup[(int64_t)(l + 1) * A.nout + g] = f;
}

// the surface upwelling flux of each group, over points 1 .. nlim - 1 (:482-507)
__global__ void __launch_bounds__(256) k_flux_surface(FluxArgs A, const double *__restrict__ sfc,
                                                      double *__restrict__ fup) {
    const int gl = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (gl >= A.ngrp) return;
    const int g = A.g0 + gl;
    if (g >= A.nout) return;
    const int64_t k0 = (int64_t)gl * A.outinrat;
    double s = 0.;
    for (int64_t k = k0; k < k0 + A.outinrat && k < A.nblk && A.p0 + k < A.npts - 1; k++) s = s + sfc[k];
    fup[g] = s;
}

// fdn, fup: device arrays of (nlay + 1) * nout, level-major; scratch: device array of
// (2 nang nlay + 1) * blk_pts doubles.
extern "C" __attribute__((visibility("hidden"))) int clblm_rt_flux_impl(
    int nlay, int64_t npts, double v1, double dv, const double *d_od, const double *d_tlay, const double *d_tlev,
    int lin, double emis, double tskin, int nang, int outinrat, int nout, int lambertian, double *scratch,
    int64_t blk_pts, double *fdn, double *fup, void *stream) {
    hipStream_t st = (hipStream_t)stream;
    FluxArgs A{};
    A.nlay = nlay;
    A.nang = nang;
    A.lin = lin;
    A.lambertian = lambertian;
    A.v1 = v1;
    A.dv = dv;
    A.emis = emis;
    A.tskin = tskin;
    A.pi = 3.1415926535898;
    A.factor = dv * 1.E04 * 2. * A.pi;
    if (nang == 1) {
        A.sec[0] = 1. / 0.66666666667;
        A.w[0] = 0.50;
    } else if (nang == 2) {
        A.sec[0] = 1.0 / 0.84494897;
        A.sec[1] = 1.0 / 0.35505103;
        A.w[0] = 0.31804138;
        A.w[1] = 0.18195862;
    } else {
        A.sec[0] = 1.0 / 0.91141204;
        A.sec[1] = 1.0 / 0.59053314;
        A.sec[2] = 1.0 / 0.21234054;
        A.w[0] = 0.20093191;
        A.w[1] = 0.22924111;
        A.w[2] = 0.06982698;
    }
    A.npts = npts;
    A.outinrat = outinrat;
    A.nout = nout;
    const int gpb = (int)(blk_pts / outinrat);   // whole groups per block
    double *dn = scratch, *up = scratch + (int64_t)nang * nlay * blk_pts, *sfc = up + (int64_t)nang * nlay * blk_pts;
    for (int64_t p0 = 0, g0 = 0; p0 < npts && g0 < nout; p0 += (int64_t)gpb * outinrat, g0 += gpb) {
        A.p0 = p0;
        A.nblk = std::min<int64_t>((int64_t)gpb * outinrat, npts - p0);
        A.g0 = (int)g0;
        A.ngrp = (int)((A.nblk + outinrat - 1) / outinrat);
        hipLaunchKernelGGL(k_flux_cols, dim3((unsigned)((A.nblk + 255) / 256)), dim3(256), 0, st, A, d_od, d_tlay,
                           d_tlev, dn, up, sfc);
        const int64_t nt = 2 * (int64_t)nlay * A.ngrp;
        hipLaunchKernelGGL(k_flux_groups, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, st, A, dn, up, fdn, fup);
        hipLaunchKernelGGL(k_flux_surface, dim3((unsigned)((A.ngrp + 255) / 256)), dim3(256), 0, st, A, sfc, fup);
    }
    return hipGetLastError() == hipSuccess ? 0 : CLBLM_ERR_HIP;
}
