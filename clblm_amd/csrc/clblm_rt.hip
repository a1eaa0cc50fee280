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

#include <cmath>
#include <string>

#include "../../include/clblm_hip.h"

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
