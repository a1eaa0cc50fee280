// MI355X layer optical depth: the reference's ODLAY nested-grid Voigt line sum
// (HIRACQ/LNCORQ/CNVFNQ/panelOut and LINF4Q/SHRINQ/LBLF4Q/CNVF4Q), restated as a
// set of batched HIP kernels for gfx950 behind the C-ABI of include/clblm_hip.h.
//
// Pipeline for a batch of layers (every stage is one launch for all layers):
//   host     plan_layer: panel and F4-block schedules (pure functions of V1, V2, DV)
//   k_prep   LNCORQ per line and the first panel each line fits (CNVFNQ JMAX1 test)
//   k_home_* running max of that panel over the line order = the panel that convolves it
//   k_ranges line range homed in each panel
//   k_conv   pass 1: each panel's own R1/R2/R3 contributions (lane per line, FP64 LDS sums)
//   k_f4*    LINF4Q per line, SHRINQ chunk simulation, CNVF4Q gather on R4 blocks
//   k_pass2  pass 2: panel assembly with the carry of the previous panel, XINT R4->R3,
//            R3->R2->R1 Lagrange interpolation, RADFNI radiation term, store
//   k_regrid interpSpectrum to DVout
// Device FP64 code is compiled with -ffp-contract=off so every expression rounds as
// the reference's does.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/clblm_hip.h"
#include "clblm_common.h"
#include "seq_ramp.h"
#include "spectrum_interp.h"

using namespace clblm;

#include "clblm_data.inc"

// ------------------------------------------------------------------------------------
// errors
// ------------------------------------------------------------------------------------
static thread_local std::string g_err;
static int set_err(int code, const std::string &m) {
    g_err = m;
    return code;
}
// set_err for the library's other translation units (not exported)
extern "C" __attribute__((visibility("hidden"))) int clblm_set_error(int code, const char *msg) {
    return set_err(code, std::string(msg));
}
#define HIPCHK(x)                                                                          \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess)                                                              \
            return set_err(CLBLM_ERR_HIP, std::string(#x ": ") + hipGetErrorString(e_));   \
    } while (0)

// ------------------------------------------------------------------------------------
// device constants
// ------------------------------------------------------------------------------------
__constant__ double c_avrat[102], c_cgauss[102], c_cf1[102], c_cf2[102], c_cf3[102], c_cer[102];
__constant__ int c_maxiso[NMOL];
// SHAPEL's constants (clblm_ODLAY.f90:2822-2833): 1/pi and A(z), B(z) of the quadratic
// continuations at z = 4, 16, 64: {recpi, a1, b1, a2, b2, a3, b3}
__constant__ double c_shape[8];

struct Tables {
    const double *f1, *f2, *f3, *fg, *xver;   // 1-based, NFMX+1 slots
    const float *f1s, *f2s, *f3s, *fgs, *xvers;
    // {F1, FG, XVER, 0} per slot: the three shapes of one R1 point in one load
    const double4 *g1;
    const float4 *g1s;
};

__device__ __forceinline__ int ftrunc(double x) { return (int)x; }

// RADFN, clblm_ODLAY_CommonSub.f90:293-328
__host__ __device__ inline double radfn(double vi, double xkt) {
    double xvi = vi;
    if (xkt > 0.0) {
        double xviokt = xvi / xkt;
        if (xviokt <= 0.01) return 0.5 * xviokt * xvi;
        if (xviokt <= 10.0) {
            double e = exp(-xviokt);
            return xvi * (1. - e) / (1. + e);
        }
        return xvi;
    }
    return xvi;
}

// d/dT of RADFN(v, T / RADCN2), branch by branch: -RADFN/T on the small-x branch,
// -2 v x e / ((1 + e)^2 T) with e = exp(-x) on the middle one, 0 above x = 10.
__host__ __device__ inline double dradfn_dt(double v, double t) {
    const double xkt = t / RADCN2;
    if (!(xkt > 0.)) return 0.;
    const double x = v / xkt;
    if (x <= 0.01) return -0.5 * x * v / t;
    if (x <= 10.0) {
        const double e = exp(-x);
        return -2. * v * x * e / ((1. + e) * (1. + e) * t);
    }
    return 0.;
}

// RADFNI, clblm_ODLAY_CommonSub.f90:379-525 (one step of the piecewise-linear
// radiation term; *vinew stays >= 0 on this path)
__host__ __device__ inline double radfni(double vi, double dvi, double xkt, double *vinew,
                                         double *rdel, double *rdlast) {
    const double FACT1 = 3.0E-03;
    double xvi = vi;
    if (*rdlast < 0.) *rdlast = radfn(vi, xkt);
    double r = *rdlast;
    int intvls = 1;
    double rdnext;
    if (xkt > 0.0) {
        double xviokt = xvi / xkt;
        if (xviokt <= 0.01) {
            if (*vinew >= 0.0) {
                *vinew = vi + FACT1 * 0.5 * xvi;
                intvls = (int)((*vinew - vi) / dvi);
                if (intvls < 1) intvls = 1;
                *vinew = vi + dvi * (double)intvls;
            } else {
                *vinew = fabs(*vinew);
                intvls = (int)((*vinew - vi) / dvi);
                if (intvls < 1) intvls = 1;
            }
            double xn = *vinew;
            double xnk = xn / xkt;
            rdnext = 0.5 * xnk * xn;
        } else if (xviokt <= 10.0) {
            double expvkt = exp(-xviokt);
            double xminus = 1. - expvkt;
            double xplus = 1. + expvkt;
            if (*vinew >= 0.0) {
                double cvikt = xviokt * expvkt;
                *vinew = vi + FACT1 * xvi / (1. + (cvikt / xminus + cvikt / xplus));
                intvls = (int)((*vinew - vi) / dvi);
                if (intvls < 1) intvls = 1;
                *vinew = vi + dvi * (double)intvls;
            } else {
                *vinew = fabs(*vinew);
                intvls = (int)((*vinew - vi) / dvi);
                if (intvls < 1) intvls = 1;
            }
            double xn = *vinew;
            double xnk = xn / xkt;
            double e = exp(-xnk);
            rdnext = xn * (1. - e) / (1. + e);
        } else {
            if (*vinew >= 0.0) {
                *vinew = vi + (FACT1 * xvi);
                intvls = (int)((*vinew - vi) / dvi);
                if (intvls < 1) intvls = 1;
                *vinew = vi + dvi * (double)intvls;
            } else {
                *vinew = fabs(*vinew);
                intvls = (int)((*vinew - vi) / dvi);
                if (intvls < 1) intvls = 1;
            }
            rdnext = *vinew;
        }
    } else {
        if (*vinew >= 0.0) {
            *vinew = vi + (FACT1 * xvi);
            intvls = (int)((*vinew - vi) / dvi);
            if (intvls < 1) intvls = 1;
            *vinew = vi + dvi * (double)intvls;
        } else {
            *vinew = fabs(*vinew);
            intvls = (int)((*vinew - vi) / dvi);
            if (intvls < 1) intvls = 1;
        }
        rdnext = *vinew;
    }
    *rdel = (rdnext - r) / (double)intvls;
    *rdlast = rdnext;
    return r;
}

// CNVFNQ's coupling factors CLC = HWF / (NX - 1), NX = 2001 (clblm_ODLAY.f90:2117-2119)
constexpr double CLC1 = 4. / 2000., CLC2 = 16. / 2000., CLC3 = 64. / 2000.;

// IZ = int(100 zeta + ONEPL) (clblm_ODLAY.f90:1893), kept within the 101 table
// intervals when zeta is not in [0, 1] (only for widths the line shape does not define)
__device__ __forceinline__ int zeta_index(double fzeta) {
    const double x = fzeta + ONEPL;
    return x >= 1. && x < 102. ? (int)x : (x >= 102. ? 101 : 1);
}

__device__ __forceinline__ double avrat_lerp(const double *tab, int iz, double zetdif) {
    return tab[iz - 1] + zetdif * (tab[iz] - tab[iz - 1]);
}

// First panel (>= 0) in which a line at vnu with 1/alpha_V = recalf fits below MAX1:
// CNVFNQ clblm_ODLAY.f90:2150-2159.
__device__ inline int first_fit_panel(const Panel *P, int np, double vnu, double recalf,
                                      double dv, double wavdxf, double hwdxf, int max1) {
    double zslope = recalf * wavdxf;
    double bhwdxf = hwdxf / zslope;
    double vft0 = P[0].vft;
    double est = ((vnu - vft0) / dv + bhwdxf + 1.5 - (double)max1) / 2400.0;
    int p = est > 0 ? (int)est - 1 : 0;
    if (p < 0) p = 0;
    if (p > np) p = np;
    auto fits = [&](int q) { return ftrunc((vnu - P[q].vft) / dv + bhwdxf + 1.5) <= max1; };
    while (p > 0 && p - 1 < np && fits(p - 1)) p--;
    while (p < np && !fits(p)) p++;
    return p;
}

// ------------------------------------------------------------------------------------
// k_prep: LNCORQ (clblm_ODLAY.f90:1566-2041) for each line of the RDLIN range.
// ------------------------------------------------------------------------------------
struct LinesDev {
    int64_t n;
    const double *nu;   // centres; a record with IFLG < 0 carries its predecessor's (sorted)
    const float *s, *a_air, *a_self, *e_lo, *tmpalf, *shift;
    const int32_t *mol, *iflg;
    const double *nu_rec;   // the VNU field as given (for IFLG < 0 records: coupling Y(200 K))
    // species broadening, 7 per record (broadener j at 7 i + j - 1), or null
    const int32_t *brd_flg;
    const float *brd_hw, *brd_tmp, *brd_shft;
};

// PSHIFT as RDLIN / RDLNFL read it: an O2 line with its own species data gets its
// foreign shift (clblm_ODLAY.f90:2730-2738, clblm_ODLAY_LineF4.f90:1463-1468).
__device__ inline double read_pshift(const LinesDev &L, int64_t i, int m) {
    double ps = (double)L.shift[i];
    if (m == 7 && L.brd_flg && L.brd_flg[7 * i + 6] > 0) {
        const double rvmr = 0.21;
        ps = (ps - rvmr * (double)L.brd_shft[7 * i + 6]) / (1.0 - rvmr);
    }
    return ps;
}
// Species broadening applies to line i: the layer asks for it and the flags sum > 0.
__device__ inline bool brd_sum(const Job &J, const LinesDev &L, int64_t i) {
    if (!J.species_broad || !L.brd_flg) return false;
    int t = 0;
#pragma unroll 1
    for (int j = 0; j < 7; j++) t += L.brd_flg[7 * i + j];
    return t > 0;
}
// The line centre after the pressure shifts: VNU + RHORAT PSHIFT, plus the
// species-weighted shift differences (LNCORQ :1849-1856, LINF4Q :994-999).
__device__ inline double brd_shift_sum(const Job &J, const LinesDev &L, int64_t i, double pshift) {
    double acc = 0.;
#pragma unroll 1
    for (int j = 0; j < 7; j++)
        acc = acc + (J.rhoslf[j + 1] * (double)L.brd_flg[7 * i + j]) * ((double)L.brd_shft[7 * i + j] - pshift);
    return acc;
}
template <bool BRD>
__device__ inline double shifted_vnu(const Job &J, const LinesDev &L, int64_t i, double vnu, double pshift) {
    double v = vnu + J.rhorat * pshift;
    if (BRD && brd_sum(J, L, i)) v = v + brd_shift_sum(J, L, i, pshift);
    return v;
}
// The Lorentz width: air and self parts, or with species broadening the
// species-weighted widths (LNCORQ :1861-1883, LINF4Q :1032-1054).
__device__ inline double brd_width(const Job &J, const LinesDev &L, int64_t i, int m, double alfa0i,
                                         double hwhmsi) {
    double alfl;
    {
        double alfsum = 0., rs = 0.;
#pragma unroll 1
        for (int j = 0; j < 7; j++) {
            const double at = (double)L.brd_hw[7 * i + j] * pow(J.tratio, (double)L.brd_tmp[7 * i + j]);
            alfsum = alfsum + (J.rhoslf[j + 1] * (double)L.brd_flg[7 * i + j]) * at;
        }
#pragma unroll 1
        for (int j = 0; j < 7; j++) rs = rs + J.rhoslf[j + 1] * (double)L.brd_flg[7 * i + j];
        alfl = (J.rhorat - rs) * alfa0i + alfsum;
        if (L.brd_flg[7 * i + m - 1] == 0) alfl = alfl + J.rhoslf[m] * (hwhmsi - alfa0i);
    }
    return alfl;
}
template <bool BRD>
__device__ inline double lorentz_width(const Job &J, const LinesDev &L, int64_t i, int m, double alfa0i,
                                       double hwhmsi) {
    double alfl = alfa0i * (J.rhorat - J.rhoslf[m]) + hwhmsi * J.rhoslf[m];
    if (!BRD) return alfl;
    if (brd_sum(J, L, i)) alfl = brd_width(J, L, i, m, alfa0i, hwhmsi);
    if (J.species_broad && L.brd_flg && L.brd_flg[7 * i + m - 1] == -1) alfl = hwhmsi;
    return alfl;
}

// Line-coupling coefficients of the IFLG 1 / 3 line at i at the layer temperature.
// The record after the line (IFLG = -1) holds them at 200, 250, 296 and 340 K
// (clblm_ODLAY.f90:1821-1829, clblm_ODLAY_LineF4.f90:971-978):
//   A(1..4) = its VNU, ALFA, MOL (the bits of a REAL*4), TMPALF: Y (IFLAG 1) or GAMMA1 (3);
//   B(1..4) = its SP, EPP, HWHM, PSHIFT: G (IFLAG 1) or GAMMA2 (3).
// Linear in temperature on one bracket: value = X(anchor) + ((X(lo+1) - X(lo)) rect) dif,
// where LNCORQ anchors at the lower node and LINF4Q at the upper one.
struct LcPair {
    double a, b;
};
__device__ inline double lc_node_a(const LinesDev &L, int64_t k, int t) {
    return t == 0 ? L.nu_rec[k]
                  : t == 1 ? (double)L.a_air[k] : t == 2 ? (double)__int_as_float(L.mol[k]) : (double)L.tmpalf[k];
}
__device__ inline double lc_node_b(const LinesDev &L, int64_t k, int t) {
    return t == 0 ? (double)L.s[k] : t == 1 ? (double)L.e_lo[k] : t == 2 ? (double)L.a_self[k] : (double)L.shift[k];
}
__device__ inline LcPair lc_coeffs(const LinesDev &L, int64_t i, int lo, int anchor, double rect, double dif) {
    const int64_t k = i + 1;
    const double a0 = lc_node_a(L, k, lo), a1 = lc_node_a(L, k, lo + 1);
    const double b0 = lc_node_b(L, k, lo), b1 = lc_node_b(L, k, lo + 1);
    LcPair r;
    r.a = (anchor == lo ? a0 : a1) + ((a1 - a0) * rect) * dif;
    r.b = (anchor == lo ? b0 : b1) + ((b1 - b0) * rect) * dif;
    return r;
}

struct LinePrep {
    double *vnu, *sp, *recalf, *zeta, *sppsp;
    int *iz, *pmin, *home;
};

// BRD: the batch uses species broadening (a separate instantiation, so that its
// registers stay out of the common one)
template <bool BRD>
__global__ void k_prep(const Job *jobs, const Panel *panels, LinesDev L, LinePrep lp, int *err) {
    const Job &J = jobs[blockIdx.y];
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= J.line_n) return;
    int64_t i = J.line_lo + k, o = J.lp_off + k;
    double sp = 0., vnu = L.nu[i], recalf = 0., zeta = 0., sppsp = 0.;
    int iz = 1;
    int pm = 0x3fffffff;
    const int fl = L.iflg[i];   // IFLG < 0: not a line (RDLIN's IOUT skips it)
    int mol = L.mol[i];
    int m = mol % 100, iso = (mol % 1000) / 100;
    if (iso == 0) iso = 10;
    bool ok = fl >= 0 && m >= 1 && m <= NMOL && J.sel[m] && iso <= c_maxiso[m - 1];
    if (ok) {
        double sui = (double)L.s[i] * J.w[m];
        if (J.jrad == 1) sui = sui * radfn(vnu, J.xkt);
        if (sui != 0.) {
            // Y, G (IFLAG 1) or GAMMA1, GAMMA2 (IFLAG 3); IFLAG 2 carries none
            LcPair lc{0., 0.};
            if (fl == 1 || fl == 3) lc = lc_coeffs(L, i, J.lcq_lo, J.lcq_anchor, J.lcq_rect, J.lcq_dif);
            vnu = shifted_vnu<BRD>(J, L, i, vnu, read_pshift(L, i, m));
            double tmpcor = pow(J.tratio, (double)L.tmpalf[i]);
            double a_air = (double)L.a_air[i];
            if (m == 7) {
                double rvmr = 0.21;
                a_air = (a_air - rvmr * (double)L.a_self[i]) / (1.0 - rvmr);
            }
            double alfa0i = a_air * tmpcor;
            double hwhmsi = (double)L.a_self[i] * tmpcor;
            double alfl = lorentz_width<BRD>(J, L, i, m, alfa0i, hwhmsi);
            if (fl == 3) alfl = alfl * (1.0 - lc.a * J.pavp0 - lc.b * J.pavp2);   // reduced width
            double alfad = vnu * J.alfd1[m][iso];
            zeta = alfl / (alfl + alfad);
            // a Lorentz width <= 0 leaves zeta outside [0, 1], where the line shape is
            // undefined (the reference indexes past its tables): reported, and the
            // index kept inside the tables
            if (!(alfl > 0.)) atomicOr(err, 32);
            double fzeta = 100. * zeta;
            iz = zeta_index(fzeta);
            double zetdif = fzeta - (double)(iz - 1);
            double alfv = avrat_lerp(c_avrat, iz, zetdif) * (alfl + alfad);
            if (alfv < J.dv) alfv = J.dv;
            if (alfv > J.alfmax) alfv = J.alfmax;
            if (!(64. * alfv + vnu < J.vft0)) {
                recalf = 1. / alfv;
                double srad = (double)L.e_lo[i];
                if (srad <= -1.001) srad = fabs(srad);
                if (srad > -0.999)
                    sui = sui * J.scor[m][iso] * exp(-(srad * J.betacr)) * (1. + exp(-(vnu / J.xkt)));
                // SP = SUI (1 + G P^2), SPPSP = SUI Y P / SP (clblm_ODLAY.f90:1937-1940); only
                // uncoupled lines (IFLAG 0) pass the depth test (DPTMIN = 0, :1992-2013)
                const double spl = (fl == 1) ? sui * (1. + lc.b * J.pavp2) : sui;
                const double spp = (fl == 1) ? sui * lc.a * J.pavp0 : 0.;
                if (fl != 0 ? spl != 0. : sui * recalf > 0.) {
                    sp = spl;
                    sppsp = spp / spl;
                    const double DXF1 = 0.002, HWF1 = 4.;
                    pm = first_fit_panel(panels + J.panel_off, J.npanel, vnu, recalf, J.dv, J.dv / DXF1,
                                         HWF1 / DXF1, J.max1);
                }
            }
        }
    }
    lp.vnu[o] = vnu;
    lp.sp[o] = sp;
    lp.sppsp[o] = sppsp;
    lp.recalf[o] = recalf;
    lp.zeta[o] = zeta;
    lp.iz[o] = iz;
    lp.pmin[o] = sp != 0. ? pm : -1;
}

// Inclusive prefix max of pmin over the line order (one workgroup per layer).
// Per-layer running max (or min, or a min scan from the end) of one array, in two
// launches over chunks of SCAN_C elements: each workgroup scans its chunk and keeps
// the chunk's total, then each workgroup combines the totals of the chunks before
// it.  Max and min are exact, so the result does not depend on the split.
constexpr int SCAN_T = 256, SCAN_V = 16, SCAN_C = SCAN_T * SCAN_V;

template <typename T, bool MAX>
__device__ __forceinline__ T scan_op(T a, T b) { return MAX ? (a > b ? a : b) : (a < b ? a : b); }

// element p of the scan order of an n-element run at off (reversed when REV)
template <bool REV>
__device__ __forceinline__ int64_t scan_at(int64_t off, int64_t n, int64_t p) { return off + (REV ? n - 1 - p : p); }

template <typename T, bool MAX, bool REV>
__device__ void scan_local(const T *src, T *dst, int64_t off, int64_t n, T ident, T *tot) {
    __shared__ T wtot[SCAN_T / 64];
    const int64_t c = blockIdx.x;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int64_t p0 = c * SCAN_C + (int64_t)tid * SCAN_V;
    T v[SCAN_V];
    T run = ident;
#pragma unroll
    for (int u = 0; u < SCAN_V; u++) {
        const int64_t p = p0 + u;
        T x = p < n ? src[scan_at<REV>(off, n, p)] : ident;
        run = scan_op<T, MAX>(run, x);
        v[u] = run;
    }
    T incl = run;  // inclusive scan of the thread totals across the wave, then the waves
    for (int d = 1; d < 64; d <<= 1) {
        T o = __shfl_up(incl, d, 64);
        if (lane >= d) incl = scan_op<T, MAX>(incl, o);
    }
    if (lane == 63) wtot[wid] = incl;
    __syncthreads();
    T pre = __shfl_up(incl, 1, 64);
    if (lane == 0) pre = ident;
    for (int w = 0; w < wid; w++) pre = scan_op<T, MAX>(wtot[w], pre);
#pragma unroll
    for (int u = 0; u < SCAN_V; u++) {
        const int64_t p = p0 + u;
        if (p < n) dst[scan_at<REV>(off, n, p)] = scan_op<T, MAX>(pre, v[u]);
    }
    if (tid == SCAN_T - 1) tot[c] = scan_op<T, MAX>(pre, incl);
}

template <typename T, bool MAX, bool REV>
__device__ void scan_fix(T *dst, int64_t off, int64_t n, T ident, const T *tot) {
    __shared__ T pre_s;
    const int64_t c = blockIdx.x;
    if (c == 0) return;
    if (threadIdx.x == 0) {
        T pre = ident;
        for (int64_t q = 0; q < c; q++) pre = scan_op<T, MAX>(pre, tot[q]);
        pre_s = pre;
    }
    __syncthreads();
    const T pre = pre_s;
    for (int64_t p = c * SCAN_C + threadIdx.x; p < (c + 1) * SCAN_C && p < n; p += SCAN_T) {
        const int64_t i = scan_at<REV>(off, n, p);
        dst[i] = scan_op<T, MAX>(pre, dst[i]);
    }
}

// home = running max of pmin over the line order (the panel that convolves each line)
__global__ void __launch_bounds__(SCAN_T) k_home_local(const Job *jobs, LinePrep lp, int *tot, int64_t cap) {
    const Job &J = jobs[blockIdx.y];
    if ((int64_t)blockIdx.x * SCAN_C >= J.line_n) return;
    scan_local<int, true, false>(lp.pmin, lp.home, J.lp_off, J.line_n, -1, tot + blockIdx.y * cap);
}
__global__ void __launch_bounds__(SCAN_T) k_home_fix(const Job *jobs, LinePrep lp, const int *tot, int64_t cap) {
    const Job &J = jobs[blockIdx.y];
    if ((int64_t)blockIdx.x * SCAN_C >= J.line_n) return;
    scan_fix<int, true, false>(lp.home, J.lp_off, J.line_n, -1, tot + blockIdx.y * cap);
}

// Line range homed in each panel (home is non-decreasing).
__global__ void k_ranges(const Job *jobs, Panel *panels, int64_t npanel_tot, LinePrep lp) {
    int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= npanel_tot) return;
    Panel &P = panels[q];
    const Job &J = jobs[P.layer];
    int p = (int)(q - J.panel_off);
    const int *h = lp.home + J.lp_off;
    auto lb = [&](int key) {
        int64_t lo = 0, hi = J.line_n;
        while (lo < hi) {
            int64_t mid = (lo + hi) >> 1;
            if (h[mid] < key) lo = mid + 1; else hi = mid;
        }
        return lo;
    };
    P.lstart = J.lp_off + lb(p);
    P.lend = J.lp_off + lb(p + 1);
}

// ------------------------------------------------------------------------------------
// Pass 1: CNVFNQ (clblm_ODLAY.f90:2067-2258) for the lines homed in one panel.  Each
// lane walks part of one line's support in the reference's order (the table argument
// ZF is accumulated exactly as the reference does); contributions land in LDS with
// FP64 atomics and the panel's arrays go to HBM.  Local index 1..MAX is kept; lower
// indices are never read by the reference's panel logic.
// Work items.  A panel of fewer lines than lanes splits every line's walk into 2 or 4
// consecutive parts, one per lane, so that its waves are filled (the upper layers'
// panels hold about 80 lines for 256 lanes); a part starting inside the walk first
// replays the ZF sums up to its first point, so the table indices stay the
// reference's.  The parts of one line sit in neighbouring lanes, with equal lengths.
// ------------------------------------------------------------------------------------
// CONV_U: R1 points per step of the walk (a launch-shape setting, 2, 4 or 8)
// CPL: the line list holds coupled lines (IFLG 1); the coupling walk is compiled
// only into that instantiation, so that the common one keeps its occupancy

// F2[iz] and F3[iz] of SHAPEL (clblm_ODLAY.f90:2822-2833, the tables of build_tables)
// from their closed forms in FP32, for the FP32 line shapes of pass 1: on x = (iz-1) dx,
// the difference of two quadratic continuations up to the inner boundary (iz <= lim),
// then the Lorentzian minus the outer quadratic; zero from iz = NFPTS on.  Evaluating
// them costs a few VALU operations against a gather from the tables (pass 1 waits on
// those loads).
struct Shape23 {
    float recpi, pa, pb, qa, qb, dx;
    int lim;
};
__device__ __forceinline__ float shape23(const Shape23 &c, int iz) {
    const float x = (float)(iz - 1) * c.dx, xs = x * x;
    const float in = iz <= c.lim ? __fmaf_rn(c.pb, xs, c.pa) : __builtin_amdgcn_rcpf(1.f + xs);
    const float v = c.recpi * (in - __fmaf_rn(c.qb, xs, c.qa));
    return iz < NFPTS ? v : 0.f;
}

// One line's CNVFNQ constants in the panel (:2120-2158).
struct ConvLine {
    double s1, s2, s3, sd, sv, zslope, zf1l, zf2l, zf3l;
    int jmin1, jmax1, j2shft, j3shft;
};
__device__ __forceinline__ void conv_line(ConvLine &c, const LinePrep &lp, int64_t o, double sp, double DV,
                                          double wavdxf, double hwdxf, double conf2, double conf3, double vft) {
    const double recalf = lp.recalf[o];
    const double depth = sp * recalf;
    const int izm = lp.iz[o];
    const double zetdif = 100. * lp.zeta[o] - (double)(izm - 1);
    c.s1 = depth * avrat_lerp(c_cf1, izm, zetdif);
    c.s2 = depth * avrat_lerp(c_cf2, izm, zetdif);
    c.s3 = depth * avrat_lerp(c_cf3, izm, zetdif);
    c.sd = depth * avrat_lerp(c_cgauss, izm, zetdif);
    c.sv = depth * avrat_lerp(c_cer, izm, zetdif);
    c.zslope = recalf * wavdxf;
    const double zint = (lp.vnu[o] - vft) / DV;
    const double bhwdxf = hwdxf / c.zslope;
    c.jmax1 = ftrunc(zint + bhwdxf + 1.5);
    c.jmin1 = ftrunc(zint - bhwdxf + 1.5);
    const double rshft = zint < 0.0 ? -0.5 : 0.5;
    c.j2shft = ftrunc(zint * (1. - conf2) + rshft);
    c.j3shft = ftrunc(zint * (1. - conf3) + rshft);
    const int jmin2 = c.jmin1 - c.j2shft, jmin3 = c.jmin1 - c.j3shft;
    c.zf1l = ((double)(c.jmin1 - 2) - zint) * c.zslope;
    c.zf2l = ((double)(jmin2 - 2) - zint * conf2) * c.zslope;
    c.zf3l = ((double)(jmin3 - 2) - zint * conf3) * c.zslope;
}

// log2 of the parts per line of a panel with nl lines for nt lanes, at most 2^maxlg.
__device__ __forceinline__ int conv_parts_lg(int nl, int nt, int maxlg) {
    const int lg = nl >= nt ? 0 : (2 * nl >= nt ? 1 : 2);
    return lg < maxlg ? lg : maxlg;
}

template <int CONV_U, bool CPL>
__global__ void __launch_bounds__(256) k_conv64(const Job *jobs, const Panel *panels, LinePrep lp,
                                                Tables T, double *A, unsigned long long *work, int split_lg) {
    extern __shared__ double sm[];
    const Panel &P = panels[blockIdx.x];
    const Job &J = jobs[P.layer];
    const int max1 = J.max1, max2 = J.max2, max3 = J.max3, stride = J.stride;
    double *a1 = sm - 1, *a2 = sm + max1 - 1, *a3 = sm + max1 + max2 - 1;  // 1-based
    for (int t = threadIdx.x; t < stride; t += blockDim.x) sm[t] = 0.;
    __syncthreads();
    const double DXF1 = 0.002, HWF1 = 4.;
    const double DV = J.dv;
    const double wavdxf = DV / DXF1, hwdxf = HWF1 / DXF1;
    const double conf2 = DV / J.dvr2, conf3 = DV / J.dvr3;
    const double vft = P.vft;
    unsigned long long nev = 0;
    const int nl = (int)(P.lend - P.lstart);
    const int lg = conv_parts_lg(nl, blockDim.x, split_lg), parts = 1 << lg;
    for (int it = threadIdx.x; it < (nl << lg); it += blockDim.x) {
        const int64_t o = P.lstart + (it >> lg);
        const int part = it & (parts - 1);
        const double sp = lp.sp[o];
        if (sp == 0.) continue;
        ConvLine c;
        conv_line(c, lp, o, sp, DV, wavdxf, hwdxf, conf2, conf3, vft);
        const double s1 = c.s1, s2 = c.s2, s3 = c.s3, sd = c.sd, sv = c.sv, zslope = c.zslope;
        const int j2shft = c.j2shft, j3shft = c.j3shft;
        const int len = c.jmax1 - c.jmin1 + 1;
        const int ja = c.jmin1 + (part * len) / parts, jb = c.jmin1 + ((part + 1) * len) / parts - 1;
        double zf1 = c.zf1l, zf2 = c.zf2l, zf3 = c.zf3l;
        for (int j = c.jmin1; j < ja; j++) {   // replay the sums up to this part
            zf3 = zf3 + zslope;
            zf2 = zf2 + zslope;
            zf1 = zf1 + zslope;
        }
        const double y1s = zf1, y2s = zf2, y3s = zf3;
        nev += (unsigned long long)(jb >= ja ? jb - ja + 1 : 0);
        // CONV_U R1 points at a time (see k_conv32)
        int j1 = ja;
        for (; j1 + (CONV_U - 1) <= jb; j1 += CONV_U) {
            double4 g[CONV_U];
            double v2[CONV_U], v3[CONV_U];
#pragma unroll
            for (int u = 0; u < CONV_U; u++) {
                zf3 = zf3 + zslope;
                zf2 = zf2 + zslope;
                zf1 = zf1 + zslope;
                g[u] = T.g1[min(ftrunc(fabs(zf1) + 1.5), NFMX)];
                v2[u] = T.f2[min(ftrunc(fabs(zf2) + 1.5), NFMX)];
                v3[u] = T.f3[min(ftrunc(fabs(zf3) + 1.5), NFMX)];
            }
#pragma unroll
            for (int u = 0; u < CONV_U; u++) {
                const int jj = j1 + u, j2 = jj - j2shft, j3 = jj - j3shft;
                if (j3 >= 1 && j3 <= max3) atomicAdd(&a3[j3], s3 * v3[u]);
                if (j2 >= 1 && j2 <= max2) atomicAdd(&a2[j2], s2 * v2[u]);
                if (jj >= 1 && jj <= max1) atomicAdd(&a1[jj], s1 * g[u].x + sd * g[u].y + sv * g[u].z);
            }
        }
        for (; j1 <= jb; j1++) {
            int j2 = j1 - j2shft, j3 = j1 - j3shft;
            zf3 = zf3 + zslope;
            zf2 = zf2 + zslope;
            zf1 = zf1 + zslope;
            int iz3 = ftrunc(fabs(zf3) + 1.5);
            int iz2 = ftrunc(fabs(zf2) + 1.5);
            int iz1 = ftrunc(fabs(zf1) + 1.5);
            if (j3 >= 1 && j3 <= max3) atomicAdd(&a3[j3], s3 * T.f3[iz3]);
            if (j2 >= 1 && j2 <= max2) atomicAdd(&a2[j2], s2 * T.f2[iz2]);
            if (j1 >= 1 && j1 <= max1) {
                const double4 g = T.g1[iz1];
                atomicAdd(&a1[j1], s1 * g.x + sd * g.y + sv * g.z);
            }
        }
        const double dptrat = CPL ? lp.sppsp[o] : 0.;
        if (CPL && dptrat != 0.) {
            // line coupling (clblm_ODLAY.f90:2195-2235): the same walk again, each
            // term scaled by CLC DPTRAT and by its table argument ZF
            const double c1 = s1 * CLC1 * dptrat, c2 = s2 * CLC2 * dptrat, c3 = s3 * CLC3 * dptrat;
            const double cd = sd * CLC1 * dptrat, cv = sv * CLC1 * dptrat;
            double y1 = y1s, y2 = y2s, y3 = y3s;
            for (int j1 = ja; j1 <= jb; j1++) {
                const int j2 = j1 - j2shft, j3 = j1 - j3shft;
                y3 = y3 + zslope;
                y2 = y2 + zslope;
                y1 = y1 + zslope;
                const int iz3 = ftrunc(fabs(y3) + 1.5), iz2 = ftrunc(fabs(y2) + 1.5), iz1 = ftrunc(fabs(y1) + 1.5);
                if (j3 >= 1 && j3 <= max3) atomicAdd(&a3[j3], c3 * T.f3[iz3] * y3);
                if (j2 >= 1 && j2 <= max2) atomicAdd(&a2[j2], c2 * T.f2[iz2] * y2);
                if (j1 >= 1 && j1 <= max1) {
                    const double4 g = T.g1[iz1];
                    atomicAdd(&a1[j1], (c1 * g.x + cd * g.y + cv * g.z) * y1);
                }
            }
        }
    }
    for (int d = 32; d > 0; d >>= 1) nev += __shfl_down(nev, d, 64);
    if ((threadIdx.x & 63) == 0 && nev) atomicAdd(work, nev);
    __syncthreads();
    double *dst = A + P.a_off;
    for (int t = threadIdx.x; t < stride; t += blockDim.x) dst[t] = sm[t];
}

// ------------------------------------------------------------------------------------
// Pass 1, FP32 line shapes: same walk and the same FP64 table argument ZF (so the
// table indices are the reference's), FP32 tables and FP32 line-shape products, and
// the per-point sums accumulated in FP64 in LDS.  On gfx950 a no-return ds_add_f64 is
// cheaper than a compensated FP32 sum, which needs the returned old value of every
// add; an FP64 accumulator of FP32 terms is at least as accurate as a Kahan sum.
// ------------------------------------------------------------------------------------
template <int CONV_U, bool CPL>
__global__ void __launch_bounds__(256) k_conv32(const Job *jobs, const Panel *panels, LinePrep lp, Tables T,
                                                double *A, unsigned long long *work, int split_lg) {
    extern __shared__ double sm[];
    const Panel &P = panels[blockIdx.x];
    const Job &J = jobs[P.layer];
    const int max1 = J.max1, max2 = J.max2, max3 = J.max3, stride = J.stride;
    double *a1 = sm - 1, *a2 = sm + max1 - 1, *a3 = sm + max1 + max2 - 1;
    const float recpi = (float)c_shape[0];
    const Shape23 sh2{recpi, (float)c_shape[1], (float)c_shape[2], (float)c_shape[3], (float)c_shape[4], 0.008f, 501};
    const Shape23 sh3{recpi, (float)c_shape[3], (float)c_shape[4], (float)c_shape[5], (float)c_shape[6], 0.032f, 501};
    for (int t = threadIdx.x; t < stride; t += blockDim.x) sm[t] = 0.;
    __syncthreads();
    const double DXF1 = 0.002, HWF1 = 4.;
    const double DV = J.dv;
    const double wavdxf = DV / DXF1, hwdxf = HWF1 / DXF1;
    const double conf2 = DV / J.dvr2, conf3 = DV / J.dvr3;
    const double vft = P.vft;
    unsigned long long nev = 0;
    const int nl = (int)(P.lend - P.lstart);
    const int lg = conv_parts_lg(nl, blockDim.x, split_lg), parts = 1 << lg;
    for (int it = threadIdx.x; it < (nl << lg); it += blockDim.x) {
        const int64_t o = P.lstart + (it >> lg);
        const int part = it & (parts - 1);
        const double sp = lp.sp[o];
        if (sp == 0.) continue;
        ConvLine c;
        conv_line(c, lp, o, sp, DV, wavdxf, hwdxf, conf2, conf3, vft);
        const float s1 = (float)c.s1, s2 = (float)c.s2, s3 = (float)c.s3, sd = (float)c.sd, sv = (float)c.sv;
        const double zslope = c.zslope;
        const int j2shft = c.j2shft, j3shft = c.j3shft;
        const int len = c.jmax1 - c.jmin1 + 1;
        const int ja = c.jmin1 + (part * len) / parts, jb = c.jmin1 + ((part + 1) * len) / parts - 1;
        double zf1 = c.zf1l, zf2 = c.zf2l, zf3 = c.zf3l;
        for (int j = c.jmin1; j < ja; j++) {   // replay the sums up to this part
            zf3 = zf3 + zslope;
            zf2 = zf2 + zslope;
            zf1 = zf1 + zslope;
        }
        const double y1s = zf1, y2s = zf2, y3s = zf3;
        nev += (unsigned long long)(jb >= ja ? jb - ja + 1 : 0);
        // CONV_U R1 points at a time, so that their table loads are in flight together;
        // the ZF arguments are still accumulated one step at a time.  Indices of points
        // outside the R2/R3 buffers are clamped into the tables (their values unused).
        int j1 = ja;
        for (; j1 + (CONV_U - 1) <= jb; j1 += CONV_U) {
            float4 g[CONV_U];
            float v2[CONV_U], v3[CONV_U];
#pragma unroll
            for (int u = 0; u < CONV_U; u++) {
                zf3 = zf3 + zslope;
                zf2 = zf2 + zslope;
                zf1 = zf1 + zslope;
                g[u] = T.g1s[min(ftrunc(fabs(zf1) + 1.5), NFMX)];
                v2[u] = shape23(sh2, ftrunc(fabs(zf2) + 1.5));
                v3[u] = shape23(sh3, ftrunc(fabs(zf3) + 1.5));
            }
#pragma unroll
            for (int u = 0; u < CONV_U; u++) {
                const int jj = j1 + u, j2 = jj - j2shft, j3 = jj - j3shft;
                if (j3 >= 1 && j3 <= max3) atomicAdd(&a3[j3], (double)(s3 * v3[u]));
                if (j2 >= 1 && j2 <= max2) atomicAdd(&a2[j2], (double)(s2 * v2[u]));
                if (jj >= 1 && jj <= max1)
                    atomicAdd(&a1[jj], (double)__fmaf_rn(sv, g[u].z, __fmaf_rn(sd, g[u].y, s1 * g[u].x)));
            }
        }
        for (; j1 <= jb; j1++) {
            int j2 = j1 - j2shft, j3 = j1 - j3shft;
            zf3 = zf3 + zslope;
            zf2 = zf2 + zslope;
            zf1 = zf1 + zslope;
            int iz3 = ftrunc(fabs(zf3) + 1.5);
            int iz2 = ftrunc(fabs(zf2) + 1.5);
            int iz1 = ftrunc(fabs(zf1) + 1.5);
            if (j3 >= 1 && j3 <= max3) atomicAdd(&a3[j3], (double)(s3 * shape23(sh3, iz3)));
            if (j2 >= 1 && j2 <= max2) atomicAdd(&a2[j2], (double)(s2 * shape23(sh2, iz2)));
            if (j1 >= 1 && j1 <= max1) {
                const float4 g = T.g1s[iz1];
                atomicAdd(&a1[j1], (double)__fmaf_rn(sv, g.z, __fmaf_rn(sd, g.y, s1 * g.x)));
            }
        }
        const double dptrat = CPL ? lp.sppsp[o] : 0.;
        if (CPL && dptrat != 0.) {   // line coupling, as in k_conv64
            const float c1 = (float)((double)s1 * CLC1 * dptrat), c2 = (float)((double)s2 * CLC2 * dptrat);
            const float c3 = (float)((double)s3 * CLC3 * dptrat), cd = (float)((double)sd * CLC1 * dptrat);
            const float cv = (float)((double)sv * CLC1 * dptrat);
            double y1 = y1s, y2 = y2s, y3 = y3s;
            for (int j1 = ja; j1 <= jb; j1++) {
                const int j2 = j1 - j2shft, j3 = j1 - j3shft;
                y3 = y3 + zslope;
                y2 = y2 + zslope;
                y1 = y1 + zslope;
                const int iz3 = ftrunc(fabs(y3) + 1.5), iz2 = ftrunc(fabs(y2) + 1.5), iz1 = ftrunc(fabs(y1) + 1.5);
                if (j3 >= 1 && j3 <= max3) atomicAdd(&a3[j3], (double)(c3 * shape23(sh3, iz3) * (float)y3));
                if (j2 >= 1 && j2 <= max2) atomicAdd(&a2[j2], (double)(c2 * shape23(sh2, iz2) * (float)y2));
                if (j1 >= 1 && j1 <= max1) {
                    const float4 g = T.g1s[iz1];
                    atomicAdd(&a1[j1], (double)(__fmaf_rn(cv, g.z, __fmaf_rn(cd, g.y, c1 * g.x)) * (float)y1));
                }
            }
        }
    }
    for (int d = 32; d > 0; d >>= 1) nev += __shfl_down(nev, d, 64);
    if ((threadIdx.x & 63) == 0 && nev) atomicAdd(work, nev);
    __syncthreads();
    double *dst = A + P.a_off;
    for (int t = threadIdx.x; t < stride; t += blockDim.x) dst[t] = sm[t];
}

// ------------------------------------------------------------------------------------
// F4 far wings.  LINF4Q per-line physics (clblm_ODLAY_LineF4.f90:951-1170).
// ------------------------------------------------------------------------------------
struct F4Prep {
    int8_t *acc;        // LINF4Q accepts the line
    int *stop;          // per layer: first accepted line with vnu > V2L4 (relative), else f4_n
    int *rec_acc;       // per layer record: accepted lines (before the stop is applied)
    int *G;             // per layer record: emitted count of a chunk starting there, by length
    int *nshr;          // per layer: shrunk lines
    int *nchunk, *nblk9;
    int4 *chunks;       // (start record, length, output offset, layer)
    int64_t *blk9_end;  // TAPE9 block ends (relative shrunk index), per layer capacity = nrec
};

struct Shrunk {
    double *vnu, *sp, *alfa, *epp, *spp, *pmax, *smin, *bvmax;
    int8_t *co2;
};

// Which lines LINF4Q accepts (RDLNFL's molecule and strength tests, the F4 band's
// lower edge), the stop line (the first accepted line beyond V2L4, where reading
// stops), and the accepted count of each TAPE3 record before the stop is applied.
// One workgroup per record.
__device__ __forceinline__ bool f4_accept(const Job &J, const LinesDev &L, int64_t i, int *m_out, int *iso_out) {
    const int mol = L.mol[i];
    int m = mol % 100, iso = (mol % 1000) / 100;
    if (iso == 0) iso = 10;
    *m_out = m;
    *iso_out = iso;
    if (!(L.iflg[i] >= 0 && m >= 1 && m <= NMOL && J.sel[m] && iso <= c_maxiso[m - 1])) return false;
    const double sui = (double)L.s[i] * J.w[m];
    return sui != 0. && !(L.nu[i] < J.v1l4);
}
template <bool BRD>
__global__ void __launch_bounds__(256) k_f4acc(const Job *jobs, LinesDev L, F4Prep f, int *err) {
    const Job &J = jobs[blockIdx.y];
    const int64_t r = blockIdx.x;
    if (r >= J.nrec) return;
    __shared__ int wcnt[4];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    int64_t a = (J.rec0 + r) * REC - J.f4_lo, b = a + REC;
    if (a < 0) a = 0;
    if (b > J.f4_n) b = J.f4_n;
    const int64_t k = a + tid;
    bool on = false;
    if (k < b) {
        const int64_t i = J.f4_lo + k;
        int m, iso;
        on = f4_accept(J, L, i, &m, &iso);
        if (on && shifted_vnu<BRD>(J, L, i, L.nu[i], read_pshift(L, i, m)) > J.v2l4)
            atomicMin(&f.stop[blockIdx.y], (int)k);
        f.acc[J.f4_off + k] = on ? 1 : 0;
    }
    const unsigned long long bal = __ballot(on);
    if (lane == 0) wcnt[wid] = __popcll(bal);
    __syncthreads();
    if (tid == 0) f.rec_acc[J.rec_off + r] = wcnt[0] + wcnt[1] + wcnt[2] + wcnt[3];
}

// SHRINQ (clblm_ODLAY_LineF4.f90:1516-1647) and LINF4Q's buffer logic (:1172-1243).
//
// The accepted lines of each layer up to the stop line are first compacted, in
// order, into positions (k_recpos, k_f4prep).  SHRINQ over a chunk of positions
// [p0, p1) is greedy: the first line opens a group with VLMT = v + DVR4/2; each
// following line joins the open group while v <= VLMT; otherwise the group closes
// (VLMT restarts from the closing line, VLMT = v_close + DVR4/2) and the line opens
// the next one; the last line of the chunk closes its group.  A group yields one
// effective line for its non-CO2 lines and one for its CO2 lines, each only if its
// strength sum is positive: strengths below 1e-35 are zeroed (they still belong to
// the group), so a part is emitted iff it holds a line of strength >= 1e-35.
struct Comp {
    double *v, *s, *al, *ad, *sp2;   // per layer at f4_off + position (sp2: coupling SPP)
    uint8_t *fl;               // bit 0: CO2; bit 1: strength >= 1e-35; bit 2: that and SPP != 0
    int *recpos;               // per layer record: position of its first accepted line
    int *nacc;                 // per layer: accepted lines up to the stop line
};
// A coupled line (CF_CPL) is not shrunk: SHRINQ copies it out as it comes (MOL 0
// unless CO2) and it joins no sum, but it takes part in the VLMT chain like any line
// (clblm_ODLAY_LineF4.f90:1563-1574).
constexpr uint8_t CF_CO2 = 1, CF_LIVE = 2, CF_CPL = 4;

// Exclusive prefix of the per-record accepted counts truncated after the stop line
// (the stop record counts its lines up to the stop line, later records none), one
// workgroup per layer.
__global__ void __launch_bounds__(256) k_recpos(const Job *jobs, F4Prep f, Comp c) {
    const Job &J = jobs[blockIdx.x];
    __shared__ int wsum[4];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int64_t stop = f.stop[blockIdx.x];
    int64_t stop_rec = J.nrec;   // no truncation
    int stop_cnt = 0;
    if (stop < J.f4_n) {
        stop_rec = (J.f4_lo + stop) / REC - J.rec0;
        int64_t a = (J.rec0 + stop_rec) * REC - J.f4_lo;
        if (a < 0) a = 0;
        const int64_t k = a + tid;
        const bool on = tid < REC && k <= stop && f.acc[J.f4_off + k] != 0;
        const unsigned long long bal = __ballot(on);
        if (lane == 0) wsum[wid] = __popcll(bal);
        __syncthreads();
        stop_cnt = wsum[0] + wsum[1] + wsum[2] + wsum[3];
        __syncthreads();
    }
    int carry = 0;
    for (int64_t b = 0; b < J.nrec; b += 256) {
        const int64_t r = b + tid;
        const int x = r < stop_rec ? f.rec_acc[J.rec_off + r] : (r == stop_rec ? stop_cnt : 0);
        int incl = x;
        for (int d = 1; d < 64; d <<= 1) {
            const int o = __shfl_up(incl, d, 64);
            if (lane >= d) incl += o;
        }
        if (lane == 63) wsum[wid] = incl;
        __syncthreads();
        int pre = carry;
        for (int w = 0; w < wid; w++) pre += wsum[w];
        if (r < J.nrec) c.recpos[J.rec_off + r] = pre + incl - x;
        carry += wsum[0] + wsum[1] + wsum[2] + wsum[3];
        __syncthreads();
    }
    if (tid == 0) c.nacc[blockIdx.x] = carry;
}

// LINF4Q per-line physics (clblm_ODLAY_LineF4.f90:951-1170) for the accepted lines of
// one TAPE3 record up to the stop line, written at their positions.
template <bool BRD>
__global__ void __launch_bounds__(256) k_f4prep(const Job *jobs, LinesDev L, F4Prep f, Comp c, int *err) {
    const Job &J = jobs[blockIdx.y];
    const int64_t r = blockIdx.x;
    if (r >= J.nrec) return;
    __shared__ int wcnt[4];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    int64_t a = (J.rec0 + r) * REC - J.f4_lo, b = a + REC;
    if (a < 0) a = 0;
    if (b > J.f4_n) b = J.f4_n;
    const int64_t stop = f.stop[blockIdx.y];
    if (b > stop + 1) b = stop + 1;
    const int64_t k = a + tid;
    const bool on = k < b && f.acc[J.f4_off + k] != 0;
    const unsigned long long bal = __ballot(on);
    if (lane == 0) wcnt[wid] = __popcll(bal);
    __syncthreads();
    if (!on) return;
    int pos = c.recpos[J.rec_off + r] + __popcll(bal & ((1ull << lane) - 1ull));
    for (int w = 0; w < wid; w++) pos += wcnt[w];
    const int64_t i = J.f4_lo + k, o = J.f4_off + pos;
    int m, iso;
    (void)f4_accept(J, L, i, &m, &iso);
    double sui = (double)L.s[i] * J.w[m];
    const double vnu = shifted_vnu<BRD>(J, L, i, L.nu[i], read_pshift(L, i, m));
    double alfa = (double)L.a_air[i];
    if (m == 7) {
        const double rvmr = 0.21;
        alfa = (alfa - rvmr * (double)L.a_self[i]) / (1.0 - rvmr);
    }
    double epp = (double)L.e_lo[i];
    const int fl = L.iflg[i];
    LcPair lc{0., 0.};   // Y, G (IFLAG 1) or GAMMA1, GAMMA2 (IFLAG 3), LINF4Q's bracket
    if (fl == 1 || fl == 3) lc = lc_coeffs(L, i, J.lc4_lo, J.lc4_anchor, J.lc4_rect, J.lc4_dif);
    if (vnu == 0.) sui = 2. * sui;
    if (epp <= -1.001) epp = fabs(epp);
    if (epp > -0.999) sui = sui * J.scor[m][iso] * exp(-(epp * J.betacr4)) * (1. + exp(-(vnu * J.beta4)));
    const double tmpcor = pow(J.tratio, (double)L.tmpalf[i]);
    const double alfa0i = alfa * tmpcor;
    const double hwhmsi = (double)L.a_self[i] * tmpcor;
    alfa = lorentz_width<BRD>(J, L, i, m, alfa0i, hwhmsi);
    if (fl == 3) alfa = alfa * (1.0 - lc.a * J.pavp0 - lc.b * J.pavp2);
    // SP = SUI (1 + G P^2), SPP = SUI Y P (:1130-1131); Y = G = 0 unless IFLAG 1
    const double sp = fl == 1 ? sui * (1. + lc.b * J.pavp2) : sui;
    const double spp = fl == 1 ? sui * lc.a * J.pavp0 : 0.;
    const bool live = !(sp < 1.0e-35);
    c.v[o] = vnu;
    c.s[o] = sp;
    c.sp2[o] = spp;
    c.al[o] = alfa;
    c.ad[o] = vnu * J.alfd1[m][iso];
    if (!(alfa > 0.)) atomicOr(err, 32);
    c.fl[o] = (uint8_t)((m == 2 ? CF_CO2 : 0) | (live ? CF_LIVE : 0) | (live && spp != 0. ? CF_CPL : 0));
}

// record end e (relative record index) of a layer as a position
__device__ __forceinline__ int rec_end_pos(const Comp &c, const Job &J, int l, int64_t e) {
    return e < J.nrec ? c.recpos[J.rec_off + e] : c.nacc[l];
}

constexpr int LCAP = 64;  // longest shrink chunk, in records, with a precomputed count

// One step of the SHRINQ grouping at position q with centre x and flags fl: the open
// group closes before q when x lies beyond VLMT (its emissions are counted), then
// q joins the open group.  fa / fb: the open group holds a non-CO2 / CO2 line of
// strength >= 1e-35.
struct ShrCount {
    double vlmt, vprev;
    int emitted, fa, fb;
};
__device__ __forceinline__ void shr_count_step(ShrCount &w, bool first, double x, int fl, double dvh) {
    if (first) {
        w.vlmt = x + dvh;
    } else if (!(x <= w.vlmt)) {
        w.emitted += w.fa + w.fb;
        w.fa = 0; w.fb = 0;
        w.vlmt = w.vprev + dvh;
    }
    fl &= 7;
    w.emitted += (fl & CF_CPL) ? 1 : 0;
    w.fa |= fl == CF_LIVE;
    w.fb |= fl == (CF_LIVE | CF_CO2);
    w.vprev = x;
}

// Lines SHRINQ emits for the chunk of positions [p0, p1), one lane (rare path).
__device__ int shrinq_count_lane(const Comp &c, const Job &J, int p0, int p1) {
    const double *vv = c.v + J.f4_off;
    const uint8_t *ff = c.fl + J.f4_off;
    const double dvh = J.dvr4 / 2.;
    ShrCount w{0., 0., 0, 0, 0};
    for (int q = p0; q < p1; q++) shr_count_step(w, q == p0, vv[q], ff[q], dvh);
    return w.emitted + w.fa + w.fb;
}

// dst[i] = src[i] for i < n by NT threads, U loads in flight per thread
template <int NT, int U, typename T>
__device__ __forceinline__ void stage(T *dst, const T *src, int n, int tid) {
    for (int b = 0; b < n; b += NT * U) {
        T t[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int i = b + u * NT + tid;
            t[u] = i < n ? src[i] : T(0);
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int i = b + u * NT + tid;
            if (i < n) dst[i] = t[u];
        }
    }
}

// G(r, L): lines a SHRINQ call emits for the chunk of the L records starting at
// record r (L = 1..LCAP), one lane per record r walking the positions once (CB
// positions loaded per batch).  A chunk never takes LIMIN4 new lines or more
// before its last record, so the walk stops at the first record end after LIMIN4
// lines (the entries beyond it are never read).
constexpr int CB = 16;
__global__ void __launch_bounds__(64) k_chunkG(const Job *jobs, Comp c, int *G) {
    const int l = blockIdx.y;
    const Job &J = jobs[l];
    const int64_t r = (int64_t)blockIdx.x * 64 + threadIdx.x;
    if (r >= J.nrec) return;
    const int p0 = c.recpos[J.rec_off + r];
    const int pend = rec_end_pos(c, J, l, r + LCAP);
    int *gout = G + (J.rec_off + r) * LCAP;
    const double *vv = c.v + J.f4_off;
    const uint8_t *ff = c.fl + J.f4_off;
    const double dvh = J.dvr4 / 2.;
    ShrCount w{0., 0., 0, 0, 0};
    int L = 0, pe = rec_end_pos(c, J, l, r + 1);
    for (int base = p0; base < pend; base += CB) {
        double vb[CB];
        int fbits[CB];
#pragma unroll
        for (int u = 0; u < CB; u++) {
            const int p = base + u < pend ? base + u : pend - 1;
            vb[u] = vv[p];
            fbits[u] = ff[p];
        }
#pragma unroll
        for (int u = 0; u < CB; u++) {
            const int q = base + u;
            if (q >= pend) break;
            bool wrote = false;
            while (pe <= q) {  // record ends between the previous line and this one
                gout[L++] = w.emitted + w.fa + w.fb;
                wrote = true;
                if (L >= LCAP) return;
                pe = rec_end_pos(c, J, l, r + 1 + L);
            }
            if (wrote && q - p0 >= LIMIN4) return;
            shr_count_step(w, q == p0, vb[u], fbits[u], dvh);
        }
    }
    for (; L < LCAP; L++) gout[L] = w.emitted + w.fa + w.fb;
}

// LINF4Q buffer logic (clblm_ODLAY_LineF4.f90:1205-1243): records are appended until
// the buffer holds LIMIN lines, the new lines are shrunk, and the buffer is written
// as one TAPE9 block once it holds LIMIN shrunk lines.  A chunk starting at record
// `start` with S shrunk lines in the buffer therefore ends at the first record end
// where it holds LIMIN - S accepted lines, or after the stop record.  One workgroup
// per layer stages the record positions and the chunk counts G(r, 1..SIMG) of a
// window of SIMW records in LDS; one lane walks the chunks.
constexpr int SIMW = 1024, SIMG = 8;
__global__ void __launch_bounds__(256) k_sim(const Job *jobs, F4Prep f, Comp c, const int *G) {
    __shared__ int rp[SIMW + 1];
    __shared__ short g8[SIMW * SIMG];
    __shared__ long long st_start, st_out;
    __shared__ int st_S, st_nch, st_nb, st_done;
    const int l = blockIdx.x;
    const Job &J = jobs[l];
    const int tid = threadIdx.x;
    const int64_t stop = f.stop[l];
    const int64_t stop_rec = stop < J.f4_n ? (J.f4_lo + stop) / REC - J.rec0 : J.nrec - 1;
    int4 *ch = f.chunks + J.rec_off;
    int64_t *be = f.blk9_end + J.rec_off;
    if (tid == 0) { st_start = 0; st_out = 0; st_S = 0; st_nch = 0; st_nb = 0; st_done = 0; }
    __syncthreads();
    while (!st_done) {
        const int64_t w0 = st_start;
        {
            int t[5];
#pragma unroll
            for (int u = 0; u < 5; u++) {
                const int i = u * 256 + tid;
                t[u] = i <= SIMW ? rec_end_pos(c, J, l, w0 + i) : 0;
            }
#pragma unroll
            for (int u = 0; u < 5; u++) {
                const int i = u * 256 + tid;
                if (i <= SIMW) rp[i] = t[u];
            }
        }
        for (int b = 0; b < SIMW * SIMG; b += 256 * 8) {
            int t[8];
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const int i = b + u * 256 + tid;
                const int64_t rr = w0 + i / SIMG;
                t[u] = rr < J.nrec ? G[(J.rec_off + rr) * LCAP + i % SIMG] : 0;
            }
#pragma unroll
            for (int u = 0; u < 8; u++) g8[b + u * 256 + tid] = (short)t[u];
        }
        __syncthreads();
        if (tid == 0) {
            int64_t start = st_start, out = st_out;
            int S = st_S, nch = st_nch, nb = st_nb, done = 0;
            while (start - w0 < SIMW - LCAP) {
                const int target = LIMIN4 - S;
                const int pst = rp[start - w0];
                int64_t e = start;
                int pe;
                for (;;) {
                    e++;
                    pe = e - w0 <= SIMW ? rp[e - w0] : rec_end_pos(c, J, l, e);
                    if (e > stop_rec || pe - pst >= target) break;
                }
                const bool eof = e > stop_rec;
                const int64_t L = e - start;
                if (pe > pst) {
                    int g;
                    if (L <= SIMG) g = g8[(start - w0) * SIMG + L - 1];
                    else if (L <= LCAP) g = G[(J.rec_off + start) * LCAP + L - 1];
                    else g = shrinq_count_lane(c, J, pst, pe);
                    ch[nch++] = make_int4((int)start, (int)L, (int)(out + S), l);
                    S += g;
                }
                start = e;
                if (S >= LIMIN4 || eof) {
                    if (S > 0) {
                        out += S;
                        be[nb++] = out;
                    }
                    if (eof) { done = 1; break; }
                    S = 0;
                }
            }
            st_start = start; st_out = out; st_S = S; st_nch = nch; st_nb = nb; st_done = done;
        }
        __syncthreads();
    }
    if (tid == 0) {
        f.nshr[l] = (int)st_out;
        f.nchunk[l] = st_nch;
        f.nblk9[l] = st_nb;
    }
}

// SHRINQ of one chunk, one workgroup per chunk: the centres and flags are staged in
// LDS, one lane walks them to find where the groups start, then the lanes take the
// groups, count their effective lines, place them by a prefix sum, and sum them in
// the reference's order (the non-CO2 part of a group first).
constexpr int EMAX = LIMIN4 + REC;
struct ShrinkSums {
    double v, s, al, ad;
};
// Derivatives of the accepted lines' LINF4Q values (v, s, al, ad, sp2 at
// dcomp[(5 k + q) tot + position] for direction k) and of the shrunk lines (dshr, the
// same layout over shrunk-line indices), for the tangent batch.
struct TanF4 {
    const double *dcomp;
    double *dshr;
    int nd;
    int64_t tot;
};
// TAN: also the derivatives of each emitted line: of a group's strength-weighted
// means (d(sum s v) - v d(sum s)) / sum s and so on, of a coupled line its own.
template <bool TAN>
__global__ void __launch_bounds__(64) k_emit(const Job *jobs, F4Prep f, Comp c, Shrunk sh, int *err, TanF4 tf) {
    __shared__ double sv[EMAX];
    __shared__ uint8_t sf[EMAX];
    __shared__ short gs[EMAX + 1];
    __shared__ int ngs;
    const int64_t q = blockIdx.x;
    const int4 ck = f.chunks[q];
    // chunk slots are per layer at rec_off; only the first nchunk of each layer are valid
    if (ck.w < 0) return;
    const int l = ck.w;
    const Job &J = jobs[l];
    if (q - J.rec_off >= f.nchunk[l]) return;
    const int tid = threadIdx.x;
    const int p0 = rec_end_pos(c, J, l, ck.x), p1 = rec_end_pos(c, J, l, (int64_t)ck.x + ck.y);
    const int n = p1 - p0;
    if (n <= 0) return;
    if (n > EMAX) {  // no LINF4Q buffer produces such a chunk: reported as an internal error
        if (tid == 0) atomicOr(err, 16);
        return;
    }
    const double dvh = J.dvr4 / 2.;
    stage<64, 8>(sv, c.v + J.f4_off + p0, n, tid);
    stage<64, 8>(sf, c.fl + J.f4_off + p0, n, tid);
    __syncthreads();
    if (tid == 0) {  // group starts: the first line, and every line beyond VLMT
        double vlmt = sv[0] + dvh, vprev = sv[0];
        int m = 0;
        gs[m++] = 0;
        for (int i0 = 1; i0 < n; i0 += 8) {
            double xb[8];
#pragma unroll
            for (int u = 0; u < 8; u++) xb[u] = sv[i0 + u < n ? i0 + u : n - 1];
#pragma unroll
            for (int u = 0; u < 8; u++) {
                if (i0 + u >= n) break;
                const double x = xb[u];
                if (!(x <= vlmt)) {
                    gs[m++] = (short)(i0 + u);
                    vlmt = vprev + dvh;
                }
                vprev = x;
            }
        }
        gs[m] = (short)n;
        ngs = m;
    }
    __syncthreads();
    const int m = ngs;
    const double *ss = c.s + J.f4_off + p0, *aa = c.al + J.f4_off + p0, *dd = c.ad + J.f4_off + p0;
    const double *pp = c.sp2 + J.f4_off + p0;
    int base_out = ck.z;
    for (int g0 = 0; g0 < m; g0 += 64) {
        const int g = g0 + tid;
        int ha = 0, hb = 0, hc = 0;
        if (g < m) {
            for (int i = gs[g]; i < gs[g + 1]; i++) {
                const int fl = sf[i] & 7;
                ha |= fl == CF_LIVE;
                hb |= fl == (CF_LIVE | CF_CO2);
                hc += (fl & CF_CPL) ? 1 : 0;
            }
        }
        const int cnt = hc + ha + hb;
        int incl = cnt;
        for (int d = 1; d < 64; d <<= 1) {
            const int o = __shfl_up(incl, d, 64);
            if (tid >= d) incl += o;
        }
        if (g < m && cnt) {
            int64_t o = J.f4_off + base_out + incl - cnt;
            ShrinkSums a{0, 0, 0, 0}, b{0, 0, 0, 0};
            for (int i = gs[g]; i < gs[g + 1]; i++) {
                if (sf[i] & CF_CPL) {   // copied out as it comes, before the group's sums
                    sh.vnu[o] = sv[i]; sh.sp[o] = ss[i]; sh.alfa[o] = aa[i]; sh.epp[o] = dd[i];
                    sh.spp[o] = pp[i]; sh.co2[o] = (sf[i] & CF_CO2) ? 1 : 0;
                    o++;
                    continue;
                }
                if (!(sf[i] & CF_LIVE)) continue;
                const double x = sv[i], sp = ss[i];
                if (sf[i] & CF_CO2) {
                    b.v = b.v + x * sp; b.s = b.s + sp; b.al = b.al + sp * aa[i]; b.ad = b.ad + sp * dd[i];
                } else {
                    a.v = a.v + x * sp; a.s = a.s + sp; a.al = a.al + sp * aa[i]; a.ad = a.ad + sp * dd[i];
                }
            }
            if (ha) {
                sh.vnu[o] = a.v / a.s; sh.sp[o] = a.s; sh.alfa[o] = a.al / a.s; sh.epp[o] = a.ad / a.s;
                sh.spp[o] = 0.; sh.co2[o] = 0;
                o++;
            }
            if (hb) {
                sh.vnu[o] = b.v / b.s; sh.sp[o] = b.s; sh.alfa[o] = b.al / b.s; sh.epp[o] = b.ad / b.s;
                sh.spp[o] = 0.; sh.co2[o] = 1;
            }
            if (TAN) {
                const int64_t p0g = J.f4_off + p0, T5 = tf.tot;
                for (int k = 0; k < tf.nd; k++) {
                    const double *dc = tf.dcomp + (int64_t)5 * k * T5;
                    double *ds = tf.dshr + (int64_t)5 * k * T5;
                    int64_t oo = J.f4_off + base_out + incl - cnt;
                    double ga[5] = {0., 0., 0., 0., 0.}, gb[5] = {0., 0., 0., 0., 0.};
                    for (int i = gs[g]; i < gs[g + 1]; i++) {
                        const int64_t pi = p0g + i;
                        if (sf[i] & CF_CPL) {
#pragma unroll
                            for (int qq = 0; qq < 5; qq++) ds[qq * T5 + oo] = dc[qq * T5 + pi];
                            oo++;
                            continue;
                        }
                        if (!(sf[i] & CF_LIVE)) continue;
                        double *gg = (sf[i] & CF_CO2) ? gb : ga;
                        const double dv = dc[pi], dsv = dc[T5 + pi], dal = dc[2 * T5 + pi], dad = dc[3 * T5 + pi];
                        gg[0] += dv * ss[i] + sv[i] * dsv;
                        gg[1] += dsv;
                        gg[2] += dsv * aa[i] + ss[i] * dal;
                        gg[3] += dsv * dd[i] + ss[i] * dad;
                    }
                    if (ha) {
                        ds[oo] = (ga[0] - sh.vnu[oo] * ga[1]) / a.s;
                        ds[T5 + oo] = ga[1];
                        ds[2 * T5 + oo] = (ga[2] - sh.alfa[oo] * ga[1]) / a.s;
                        ds[3 * T5 + oo] = (ga[3] - sh.epp[oo] * ga[1]) / a.s;
                        ds[4 * T5 + oo] = 0.;
                        oo++;
                    }
                    if (hb) {
                        ds[oo] = (gb[0] - sh.vnu[oo] * gb[1]) / b.s;
                        ds[T5 + oo] = gb[1];
                        ds[2 * T5 + oo] = (gb[2] - sh.alfa[oo] * gb[1]) / b.s;
                        ds[3 * T5 + oo] = (gb[3] - sh.epp[oo] * gb[1]) / b.s;
                        ds[4 * T5 + oo] = 0.;
                    }
                }
            }
        }
        base_out += __shfl(incl, 63, 64);
    }
}

// Prefix max / suffix min of the shrunk centres (the range searches of the R4
// gather), as per-layer scans, and the prefix max of the TAPE9 block maxima.
__global__ void __launch_bounds__(SCAN_T) k_shr_local(const Job *jobs, F4Prep f, Shrunk sh, double *tot,
                                                      int64_t cap) {
    const Job &J = jobs[blockIdx.y];
    const int64_t n = f.nshr[blockIdx.y];
    if ((int64_t)blockIdx.x * SCAN_C >= n) return;
    if (blockIdx.z == 0)
        scan_local<double, true, false>(sh.vnu, sh.pmax, J.f4_off, n, -1e300, tot + (2 * blockIdx.y) * cap);
    else
        scan_local<double, false, true>(sh.vnu, sh.smin, J.f4_off, n, 1e300, tot + (2 * blockIdx.y + 1) * cap);
}
__global__ void __launch_bounds__(SCAN_T) k_shr_fix(const Job *jobs, F4Prep f, Shrunk sh, const double *tot,
                                                    int64_t cap) {
    const Job &J = jobs[blockIdx.y];
    const int64_t n = f.nshr[blockIdx.y];
    if ((int64_t)blockIdx.x * SCAN_C >= n) return;
    if (blockIdx.z == 0)
        scan_fix<double, true, false>(sh.pmax, J.f4_off, n, -1e300, tot + (2 * blockIdx.y) * cap);
    else
        scan_fix<double, false, true>(sh.smin, J.f4_off, n, 1e300, tot + (2 * blockIdx.y + 1) * cap);
}
// Prefix max of the TAPE9 block maxima (the last shrunk centre of each block), one
// workgroup per layer.
__global__ void __launch_bounds__(256) k_shr_blk9(const Job *jobs, F4Prep f, Shrunk sh) {
    __shared__ double wmax[4];
    const int l = blockIdx.x;
    const Job &J = jobs[l];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const double *v = sh.vnu + J.f4_off;
    const int nb = f.nblk9[l];
    double carry = -1e300;
    for (int b0 = 0; b0 < nb; b0 += 256) {
        const int b = b0 + tid;
        const double x = b < nb ? v[f.blk9_end[J.rec_off + b] - 1] : -1e300;
        double incl = x;
        for (int d = 1; d < 64; d <<= 1) {
            const double o = __shfl_up(incl, d, 64);
            if (lane >= d) incl = fmax(incl, o);
        }
        if (lane == 63) wmax[wid] = incl;
        __syncthreads();
        double pre = carry;
        for (int w = 0; w < wid; w++) pre = fmax(pre, wmax[w]);
        if (b < nb) sh.bvmax[J.f4_off + b] = fmax(pre, incl);
        carry = fmax(fmax(carry, wmax[0]), fmax(wmax[1], fmax(wmax[2], wmax[3])));
        __syncthreads();
    }
}

// Line range RDLN4Q/CNVF4Q read for each R4 block (clblm_ODLAY_LineF4.f90:554-638,
// 187-240): from the first TAPE9 block whose last centre reaches V1R4-BOUND4, up to
// the first line at or beyond V2R4+BOUND4.
__global__ void k_r4range(const Job *jobs, R4Blk *blks, int64_t nblk_tot, F4Prep f, Shrunk sh) {
    int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nblk_tot) return;
    R4Blk &B = blks[q];
    const Job &J = jobs[B.layer];
    int64_t n = f.nshr[B.layer];
    int nb = f.nblk9[B.layer];
    const double *bv = sh.bvmax + J.f4_off;
    double vlo = B.v1r4 - J.bound4;
    int lo_b = 0, hi_b = nb;
    while (lo_b < hi_b) {
        int mid = (lo_b + hi_b) >> 1;
        if (bv[mid] < vlo) lo_b = mid + 1; else hi_b = mid;
    }
    int64_t lo = lo_b == 0 ? 0 : (lo_b >= nb ? n : f.blk9_end[J.rec_off + lo_b - 1]);
    double vnulst = B.v2r4 + J.bound4;
    const double *pm = sh.pmax + J.f4_off;
    int64_t a = 0, b = n;
    while (a < b) {
        int64_t mid = (a + b) >> 1;
        if (pm[mid] < vnulst) a = mid + 1; else b = mid;
    }
    B.lo = J.f4_off + lo;
    B.hi = J.f4_off + (a > lo ? a : lo);
}

// ------------------------------------------------------------------------------------
// CNVF4Q (clblm_ODLAY_LineF4.f90:187-494) as a gather: the per-line constants are
// computed once per (shrunk line, R4 block) into records, then each R4 point sums the
// records that reach it, in the reference's line order.
// ------------------------------------------------------------------------------------
struct F4Line {
    double xnui, xnum, sil, siv_a3, siv_b3, f4bnd, alfli2, rec_alfvi2, dptrat;
    int jmin, jmax, jminm, jmaxm, flags;  // flags: 1 co2, 2 spp==0
};

__device__ inline void f4_line_setup(F4Line &o, const Job &J, const R4Blk &B, double vnu, double sp,
                                     double spp, double alfa, double epp, int co2) {
    const double RECPI = 1. / (2. * asin(1.));
    const double ZSQBND = 64. * 64.;
    const double A3 = (1. + 2. * ZSQBND) / ((1. + ZSQBND) * (1. + ZSQBND));
    const double B3 = -1. / ((1. + ZSQBND) * (1. + ZSQBND));
    const double BNDSQ = J.bound4 * J.bound4;
    o.jmin = 1; o.jmax = 0; o.jminm = 1; o.jmaxm = 0; o.flags = 0;
    if (sp == 0. && spp == 0.) return;
    double alfadi = epp, alfali = alfa;
    double zetai = alfali / (alfali + alfadi);
    int iz = zeta_index(100. * zetai);
    double zetdif = 100. * zetai - (double)(iz - 1);
    double alfavi = avrat_lerp(c_avrat, iz, zetdif) * (alfali + alfadi);
    double ralfvi = 1. / alfavi;
    double sil = sp * RECPI * alfali;
    double siv = (alfali * ralfvi) * sp * RECPI * ralfvi;
    double speak = A3 * fabs(siv);
    if (speak <= 0. && spp == 0.) return;
    double xnui = vnu - B.v1r4;
    int jmin = (int)((xnui - J.bound4) / J.dvr4 + 2.);
    if (jmin > B.nptr4) return;  // GO TO 60: no mirror either
    if (jmin < 1) jmin = 1;
    int jmax = (int)((xnui + J.bound4) / J.dvr4 + 1.);
    if (jmax > B.nptr4) jmax = B.nptr4;
    o.xnui = xnui; o.jmin = jmin; o.jmax = jmax;
    if (vnu > 0. && vnu <= 25.) {
        double xm = -vnu - B.v1r4;
        int a = (int)((xm - J.bound4) / J.dvr4 + 2.);
        if (a <= B.nptr4) {
            if (a < 1) a = 1;
            int b = (int)((xm + J.bound4) / J.dvr4 + 1.);
            if (b > B.nptr4) b = B.nptr4;
            o.xnum = xm; o.jminm = a; o.jmaxm = b;
        }
    }
    o.alfli2 = alfali * alfali;
    o.sil = sil;
    o.f4bnd = sil / (o.alfli2 + BNDSQ);
    o.dptrat = spp / (sp * alfavi);
    o.rec_alfvi2 = 1. / (alfavi * alfavi);
    o.siv_a3 = siv * A3;
    o.siv_b3 = siv * B3;
    o.flags = (co2 ? 1 : 0) | (spp == 0. ? 2 : 0);
}

__device__ __forceinline__ int64_t lower_bound_d(const double *a, int64_t lo, int64_t hi, double key) {
    while (lo < hi) {
        int64_t mid = (lo + hi) >> 1;
        if (a[mid] < key) lo = mid + 1; else hi = mid;
    }
    return lo;
}
__device__ __forceinline__ int64_t upper_bound_d(const double *a, int64_t lo, int64_t hi, double key) {
    while (lo < hi) {
        int64_t mid = (lo + hi) >> 1;
        if (a[mid] <= key) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// candidate shrunk-line range for R4 points [j0, j1] of block B (absolute indices)
__device__ inline void f4_candidates(const Job &J, const R4Blk &B, const Shrunk &sh, int j0, int j1,
                                     int64_t *c0, int64_t *c1) {
    double vlo = B.v1r4 + (double)(j0 - 1) * J.dvr4 - J.bound4 - 2. * J.dvr4;
    double vhi = B.v1r4 + (double)(j1 - 1) * J.dvr4 + J.bound4 + 2. * J.dvr4;
    int64_t lo = B.lo, hi = B.hi;
    int64_t a = lower_bound_d(sh.pmax, lo, hi, vlo);
    int64_t b = upper_bound_d(sh.smin, lo, hi, vhi);
    if (vlo < J.bound4 + 2. * J.dvr4) {  // Van Vleck-Weisskopf mirror images can reach
        int64_t m0 = upper_bound_d(sh.pmax, lo, hi, 0.);
        if (m0 < a) a = m0;
        int64_t m1 = upper_bound_d(sh.smin, lo, hi, 25.);
        if (m1 > b) b = m1;
    }
    *c0 = a;
    *c1 = b > a ? b : a;
}

// Per-(line, block) F4 records: the CNVF4Q per-line quantities and the block's
// point range for the line, computed once.  The line shape is written branch-free:
//   base(x)  = (c2 - x^2 rb) f4bnd        (c2 = 2, rb = 1/BNDSQ for CO2 lines; 1, 0 otherwise)
//   shift(f) = f + x dptrat f             (dptrat = 0 when the line has no shift)
// which reproduces CNVF4Q's CO2 and other-molecule loops (clblm_ODLAY_LineF4.f90:388-462).
// [jqlo, jqhi] contains every point where the shape takes its quadratic form
// (zeta^2 <= 64^2, :397-401) with two points of margin; a tile outside it, inside
// [jmin, jmax] and with no shift needs only the Lorentz form.
enum { F4_CO2 = 1, F4_SHIFT = 2, F4_MIRROR = 4 };
struct alignas(16) F4Rec32 {
    float d, sil, alfli2, b0;          // x(j) = (j - jc) dvr4 + d, FP64-anchored; base(x) = b0 + b1 x^2
    int jc, jmin, jmax, kind;
    int jqlo, jqhi;
    float b1, dptrat;
    float rec_alfvi2, siv_a3, siv_b3, dm;
    int jcm, jminm, jmaxm, pad;
};
struct alignas(16) F4Rec64 {
    double xnui, xnum, sil, alfli2, rec_alfvi2, siv_a3, siv_b3, f4bnd, c2, rb, dptrat;
    int jmin, jmax, jminm, jmaxm;
    int jqlo, jqhi, kind, pad;
    double pad2;
};
static_assert(sizeof(F4Rec32) == 80 && sizeof(F4Rec64) == 128, "F4 record layout");
// the gathers read these words as 16-byte vectors: F4Rec32 words 4-9, F4Rec64 words 22-28
static_assert(offsetof(F4Rec32, jc) == 16 && offsetof(F4Rec32, kind) == 28 && offsetof(F4Rec32, jqlo) == 32 &&
                  offsetof(F4Rec32, b1) == 40,
              "F4Rec32 control words");
static_assert(offsetof(F4Rec64, jmin) == 88 && offsetof(F4Rec64, jqlo) == 104 && offsetof(F4Rec64, kind) == 112,
              "F4Rec64 control words");

// CNVF4Q's grid coordinate (clblm_ODLAY_LineF4.f90:356,426,464) is accumulated per
// line: XJJ = (JMIN-1) DVR4, then XJJ = XJJ + DVR4 at every point.  The rounding of
// that sum drifts from (j-1) DVR4 by hundreds of ulps along a 2500-point block.
// Every line's sequence takes the same steps as the one that starts from 0 at
// j = 1 (the step is fixed inside a binade), so XJJ(j) = S(j) + (X0 - S(JMIN))
// to within one ulp, where S is that shared sequence.  S is tabulated per layer
// (XJJ_N values, S(j) at index j-1) by the sequential sum itself, one lane per
// layer; the records carry XNUI - (X0 - S(JMIN)), so that x = S(j) - XNUI' in the
// gathers.
constexpr int XJJ_N = LIMOUT4 + 64 * 4;   // covers every point of the last tile
__global__ void __launch_bounds__(64) k_xjj(const Job *jobs, int nlay, double *xjj) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= nlay) return;
    const double d = jobs[l].dvr4;
    double *S = xjj + (int64_t)l * XJJ_N;
    double t = 0.;
    for (int j = 0; j < XJJ_N; j++) {
        S[j] = t;
        t = t + d;
    }
}

// Record offsets of each block (one lane per layer walks its blocks in order).
__global__ void k_f4recoff(const Job *jobs, int nlay, R4Blk *blks, int *err) {
    int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= nlay) return;
    const Job &J = jobs[l];
    int64_t o = J.f4rec_off;
    for (int b = 0; b < J.nblk; b++) {
        R4Blk &B = blks[J.blk_off + b];
        B.rec = o;
        o += B.hi - B.lo;
    }
    if (o - J.f4rec_off > J.f4rec_cap) atomicOr(err, 8);
}

// the quadratic zone of a line at xnui, widened by two points each side
__device__ inline void f4_qzone(const F4Line &d, const Job &J, int *lo, int *hi) {
    const double R = 64. * sqrt(1. / d.rec_alfvi2);
    *lo = (int)floor((d.xnui - R) / J.dvr4) + 1 - 2;
    *hi = (int)ceil((d.xnui + R) / J.dvr4) + 1 + 2;
}
__device__ inline int f4_kind(const F4Line &d) {
    return ((d.flags & 1) ? F4_CO2 : 0) | ((d.flags & 2) ? 0 : F4_SHIFT) | (d.jmaxm >= d.jminm ? F4_MIRROR : 0);
}

// XNUI' = XNUI - (X0 - S(JMIN)) with X0 = (JMIN-1) DVR4 (see k_xjj)
__device__ inline double xnui_seq(double xnui, int jmin, const double *S, const Job &J) {
    const int i = jmin - 1 < XJJ_N - 1 ? jmin - 1 : XJJ_N - 1;
    return xnui - ((double)(jmin - 1) * J.dvr4 - S[i]);
}

template <typename R>
__device__ inline void f4_fill(R &r, const F4Line &d, const Job &J, const double *S);
template <>
__device__ inline void f4_fill<F4Rec64>(F4Rec64 &r, const F4Line &d, const Job &J, const double *S) {
    const bool co2 = (d.flags & 1) != 0;
    r.xnui = d.jmax >= d.jmin ? xnui_seq(d.xnui, d.jmin, S, J) : d.xnui;
    r.xnum = d.jmaxm >= d.jminm ? xnui_seq(d.xnum, d.jminm, S, J) : d.xnum;
    r.sil = d.sil; r.alfli2 = d.alfli2; r.rec_alfvi2 = d.rec_alfvi2;
    r.siv_a3 = d.siv_a3; r.siv_b3 = d.siv_b3; r.f4bnd = d.f4bnd;
    r.c2 = co2 ? 2. : 1.;
    r.rb = co2 ? 1. / 625. : 0.;
    r.dptrat = (d.flags & 2) ? 0. : d.dptrat;
    r.jmin = d.jmin; r.jmax = d.jmax; r.jminm = d.jminm; r.jmaxm = d.jmaxm;
    r.kind = f4_kind(d);
    if (d.jmax >= d.jmin) f4_qzone(d, J, &r.jqlo, &r.jqhi);
    else { r.jqlo = 1; r.jqhi = 0; }
    r.pad = 0;
    r.pad2 = 0.;
}
template <>
__device__ inline void f4_fill<F4Rec32>(F4Rec32 &r, const F4Line &d, const Job &J, const double *S) {
    const bool co2 = (d.flags & 1) != 0;
    r.jmin = d.jmin; r.jmax = d.jmax; r.jminm = d.jminm; r.jmaxm = d.jmaxm;
    r.kind = f4_kind(d);
    r.pad = 0;
    r.sil = (float)d.sil; r.alfli2 = (float)d.alfli2; r.rec_alfvi2 = (float)d.rec_alfvi2;
    r.siv_a3 = (float)d.siv_a3; r.siv_b3 = (float)d.siv_b3;
    r.b0 = (float)(co2 ? 2. * d.f4bnd : d.f4bnd);
    r.b1 = (float)(co2 ? -d.f4bnd / 625. : 0.);
    r.dptrat = (d.flags & 2) ? 0.f : (float)d.dptrat;
    // nearest grid point and the FP64 offset to it (on the accumulated grid coordinate):
    // x(j) = S(j) - xnui' = (j-jc)*dvr4 + d
    if (d.jmax >= d.jmin) {
        int jc = (int)floor(d.xnui / J.dvr4 + 0.5) + 1;
        const int ic = min(max(jc - 1, 0), XJJ_N - 1);
        r.jc = jc;
        r.d = (float)((S[ic] + (double)(jc - 1 - ic) * J.dvr4) - xnui_seq(d.xnui, d.jmin, S, J));
        f4_qzone(d, J, &r.jqlo, &r.jqhi);
    } else { r.jc = 0; r.d = 0.f; r.jqlo = 1; r.jqhi = 0; }
    if (d.jmaxm >= d.jminm) {
        int jc = (int)floor(d.xnum / J.dvr4 + 0.5) + 1;
        const int ic = min(max(jc - 1, 0), XJJ_N - 1);
        r.jcm = jc;
        r.dm = (float)((S[ic] + (double)(jc - 1 - ic) * J.dvr4) - xnui_seq(d.xnum, d.jminm, S, J));
    } else { r.jcm = 0; r.dm = 0.f; }
}

template <typename R>
__global__ void k_f4setup(const Job *jobs, const R4Blk *blks, int64_t nblk, Shrunk sh, R *recs,
                          const double *xjj) {
    int64_t b = blockIdx.y;
    if (b >= nblk) return;
    const R4Blk &B = blks[b];
    const Job &J = jobs[B.layer];
    const double *S = xjj + (int64_t)B.layer * XJJ_N;
    for (int64_t i = B.lo + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < B.hi;
         i += (int64_t)gridDim.x * blockDim.x) {
        F4Line d;
        f4_line_setup(d, J, B, sh.vnu[i], sh.sp[i], sh.spp[i], sh.alfa[i], sh.epp[i], sh.co2[i]);
        f4_fill<R>(recs[B.rec + (i - B.lo)], d, J, S);
    }
}

// Quotient through the hardware reciprocal refined by two Newton-Raphson steps
// (r <- r + r (1 - b r)); within 2 ulp of a / b for the normal, positive b used here.
__device__ __forceinline__ double div_nr(double a, double b) {
    double r = __builtin_amdgcn_rcp(b);
    double e = fma(-b, r, 1.);
    r = fma(e, r, r);
    e = fma(-b, r, 1.);
    r = fma(e, r, r);
    return a * r;
}

__device__ __forceinline__ double f4_eval64(const F4Rec64 &l, double xm, double dptrat) {
    double xmsq = xm * xm;
    double zvsq = xmsq * l.rec_alfvi2;
    double base = (l.c2 - xmsq * l.rb) * l.f4bnd;
    double q = l.siv_a3 + zvsq * l.siv_b3;
    double lz = div_nr(l.sil, l.alfli2 + xmsq);
    double f = ((zvsq <= 4096.) ? q : lz) - base;
    return f + xm * dptrat * f;
}

__device__ __forceinline__ float f4_eval32(const F4Rec32 &l, float xm, float dptrat) {
    float xmsq = xm * xm;
    float zvsq = xmsq * l.rec_alfvi2;
    float q = __fmaf_rn(zvsq, l.siv_b3, l.siv_a3);
    float lz = l.sil * __builtin_amdgcn_rcpf(l.alfli2 + xmsq);
    float f = ((zvsq <= 4096.f) ? q : lz) - __fmaf_rn(xmsq, l.b1, l.b0);
    return __fmaf_rn(xm * dptrat, f, f);
}

// F4K: points per lane of the F4 gathers (tiles of 64 F4K points), a launch-shape
// setting of the context (defaults 3 for the FP32 shapes, 2 for FP64)
// F4 records summed in FP32 before each FP64 accumulation (FP32 gather)
constexpr int F4CHUNK = 16;
// Far field of the FP32 gather.  A record whose line centre lies at least half a tile
// points outside the tile, and whose shape over the whole tile is the plain Lorentz
// form (covered, no shift, no mirror), is a smooth function there: it is evaluated
// at F4NF Chebyshev nodes of the tile instead of at every point, and the node sums
// are interpolated to the points once per tile (Lagrange weights tabulated in the
// context per tile width).  For 1/(a^2 + x^2) with its centre half a tile outside
// the tile, 16 nodes interpolate to within 1.4e-8 of its maximum over the tile.
constexpr int F4NF = 16;
constexpr int F4KMIN = 2, F4KMAX = 4;   // points per lane the gathers are built for
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
constexpr int RAMP_CAP = 256;  // ramp pieces per panel held in LDS


__device__ __forceinline__ unsigned long long f4_count(int jmin, int jmax, int j0, int j1) {
    int a = jmin > j0 ? jmin : j0, b = jmax < j1 ? jmax : j1;
    return b >= a ? (unsigned long long)(b - a + 1) : 0ull;
}

// Plan of the records staged for a tile, made by the lane that stages each record and
// shared as ballots, so that the walk over the records branches on scalar bits:
//   touch  the record reaches the tile;
//   cover  its range [jmin, jmax] contains the whole tile;
//   lorp   the tile lies outside its quadratic zone and it has no shift, so the
//          Lorentz form alone gives the shape there;
//   all    cover and lorp: no per-point test at all.
struct F4Plan {
    unsigned long long all, cover, lorp, touch, co2, mirror, far;
};
// far (FP32 gather only, dfar > 0): an `all` record without a mirror whose centre
// point jc lies at least dfar points outside the tile; it leaves every other set.
__device__ __forceinline__ void f4_plan(F4Plan &P, bool live, int jmin, int jmax, int jqlo, int jqhi, int kind,
                                        int j0, int j1, int jc = 0, int dfar = 0) {
    const bool mir = live && (kind & F4_MIRROR) != 0;
    P.mirror = __ballot(mir);
    live = live && jmin <= jmax;
    const bool touch = live && jmin <= j1 && jmax >= j0;
    const bool cover = live && jmin <= j0 && jmax >= j1;
    const bool lorq = (jqhi < j0 || jqlo > j1) && (kind & F4_SHIFT) == 0;
    const bool far = dfar > 0 && cover && lorq && !mir && (jc <= j0 - dfar || jc >= j1 + dfar);
    P.touch = __ballot(touch && !far);
    P.cover = __ballot(cover && !far);
    P.lorp = __ballot(touch && lorq && !far);
    P.far = __ballot(far);
    P.all = __ballot(cover && lorq && !far);
    P.co2 = __ballot(live && (kind & F4_CO2) != 0);
}

// F4 gather, FP64: one wave per tile of 64 F4K consecutive R4 points, F4K points per
// lane (j = j0 + lane + 64 k), lines walked in the reference's order so each point's
// sum has the reference's summation order.  Records pass through LDS 64 at a time
// (one coalesced load per lane, the next batch in flight while this one is summed)
// and are read back as broadcasts of plain 16-byte vectors (so that no copy of a
// record is made in private memory).
// XCD-aware tile order.  Workgroups are dispatched to the 8 XCDs round-robin
// (workgroup g runs on XCD g % 8), and each XCD has its own L2.  Neighbouring
// tiles read nearly the same line records, so runs of `chunk` consecutive tiles go
// to one XCD (chunks dealt round-robin, which keeps the XCDs' loads even).  The map
// is a bijection of [0, 8 chunk ceil(n / (8 chunk))); indices >= n exit.
__device__ __forceinline__ int64_t xcd_tile(int64_t g, int chunk) {
    const int64_t x = g & 7, s = g >> 3;
    return ((s / chunk) * 8 + x) * chunk + (s % chunk);
}

template <int F4K>
__global__ void __launch_bounds__(64) k_f4conv64(const Job *jobs, const R4Blk *blks, const int2 *tiles, Shrunk sh,
                                                 const F4Rec64 *__restrict__ recs, double *R4,
                                                 unsigned long long *work, int64_t ntile, int chunk,
                                                 const double *__restrict__ xjj) {
    const int64_t ti = xcd_tile(blockIdx.x, chunk);
    if (ti >= ntile) return;
    const int2 t = tiles[ti];
    const R4Blk &B = blks[t.x];
    const Job &J = jobs[B.layer];
    const int lane = threadIdx.x;
    const int j0 = t.y * 64 * F4K + 1, j1 = j0 + 64 * F4K - 1;
    double acc[F4K], xj[F4K];   // xj: the accumulated grid coordinate S(j) (k_xjj)
    const double *S = xjj + (int64_t)B.layer * XJJ_N;
#pragma unroll
    for (int k = 0; k < F4K; k++) {
        acc[k] = 0.;
        const int i = j0 + lane + 64 * k - 1;
        xj[k] = S[i < XJJ_N ? i : XJJ_N - 1];
    }
    unsigned long long nev = 0;   // (line, point) pairs of this tile, one record per lane
    if (j0 <= B.nptr4) {
        int64_t c0, c1;
        f4_candidates(J, B, sh, j0, min(j1, B.nptr4), &c0, &c1);
        constexpr int W = (int)(sizeof(F4Rec64) / 16);
        const u32x4 *r = reinterpret_cast<const u32x4 *>(recs + B.rec - B.lo);
        __shared__ u32x4 srec[64 * W];
        // the prefetch is unconditional (index clamped to the last record) so that it
        // stays in registers
        u32x4 pre[W];
        const int64_t last = c1 - 1;
        if (c1 > c0) {
            const int64_t k = c0 + lane < last ? c0 + lane : last;
#pragma unroll
            for (int w = 0; w < W; w++) pre[w] = r[k * W + w];
        }
        for (int64_t base = c0; base < c1; base += 64) {
            const int nb = (int)((c1 - base) < 64 ? (c1 - base) : 64);
            F4Plan P;
            f4_plan(P, lane < nb, (int)pre[5].z, (int)pre[5].w, (int)pre[6].z, (int)pre[6].w, (int)pre[7].x, j0, j1);
            if (lane < nb) nev += f4_count((int)pre[5].z, (int)pre[5].w, j0, j1);
            __syncthreads();
            if (lane < nb) {
#pragma unroll
                for (int w = 0; w < W; w++) srec[lane * W + w] = pre[w];
            }
            __syncthreads();
            {
                const int64_t k = base + 64 + lane < last ? base + 64 + lane : last;
#pragma unroll
                for (int w = 0; w < W; w++) pre[w] = r[k * W + w];
            }
            for (int q = 0; q < nb; q++) {
                const unsigned long long bit = 1ull << q;
                const F4Rec64 &l = *reinterpret_cast<const F4Rec64 *>(&srec[q * W]);
                if (P.all & bit) {
                    // the whole tile in the Lorentz zone: f4_eval64 without the branches
                    const double xnui = l.xnui, sil = l.sil, alfli2 = l.alfli2, c2 = l.c2, rb = l.rb, f4bnd = l.f4bnd;
#pragma unroll
                    for (int k = 0; k < F4K; k++) {
                        const double xm = xj[k] - xnui, xmsq = xm * xm;
                        acc[k] += div_nr(sil, alfli2 + xmsq) - (c2 - xmsq * rb) * f4bnd;
                    }
                } else if (P.cover & bit) {
#pragma unroll
                    for (int k = 0; k < F4K; k++) acc[k] += f4_eval64(l, xj[k] - l.xnui, l.dptrat);
                } else if (P.lorp & bit) {
                    // part of the tile, all of it in the Lorentz zone
                    const double xnui = l.xnui, sil = l.sil, alfli2 = l.alfli2, c2 = l.c2, rb = l.rb, f4bnd = l.f4bnd;
                    const int jmn = l.jmin, jmx = l.jmax;
#pragma unroll
                    for (int k = 0; k < F4K; k++) {
                        const int j = j0 + lane + 64 * k;
                        const double xm = xj[k] - xnui, xmsq = xm * xm;
                        const double f = div_nr(sil, alfli2 + xmsq) - (c2 - xmsq * rb) * f4bnd;
                        acc[k] += (j >= jmn && j <= jmx) ? f : 0.;
                    }
                } else if (P.touch & bit) {
                    const int jmn = l.jmin, jmx = l.jmax;
#pragma unroll
                    for (int k = 0; k < F4K; k++) {
                        const int j = j0 + lane + 64 * k;
                        double f = f4_eval64(l, xj[k] - l.xnui, l.dptrat);
                        acc[k] += (j >= jmn && j <= jmx) ? f : 0.;
                    }
                }
                if (P.mirror & bit) {  // Van Vleck-Weisskopf mirror (lines below 25 cm-1)
#pragma unroll
                    for (int k = 0; k < F4K; k++) {
                        const int j = j0 + lane + 64 * k;
                        double f = f4_eval64(l, xj[k] - l.xnum, -l.dptrat);
                        acc[k] += (j >= l.jminm && j <= l.jmaxm) ? f : 0.;
                    }
                }
            }
        }
    }
#pragma unroll
    for (int k = 0; k < F4K; k++) {
        const int j = j0 + lane + 64 * k;
        if (j <= LIMOUT4 + 2) R4[B.r4_off + j - 1] = (j <= B.nptr4) ? acc[k] : 0.;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) nev += __shfl_xor(nev, o);
    if (lane == 0 && nev) atomicAdd(work + 1, nev);
}

// F4 gather, FP32 line shapes (offsets from the FP64 per-record anchor) with FP64 sums
// per point; the same walk as k_f4conv64.
template <int F4K>
__global__ void __launch_bounds__(64) k_f4conv32(const Job *jobs, const R4Blk *blks, const int2 *tiles, Shrunk sh,
                                                 const F4Rec32 *__restrict__ recs, double *R4,
                                                 unsigned long long *work, int64_t ntile, int chunk,
                                                 const double *__restrict__ far_w, const float *__restrict__ far_tau,
                                                 int dfar) {
    const int64_t ti = xcd_tile(blockIdx.x, chunk);
    if (ti >= ntile) return;
    const int2 t = tiles[ti];
    const R4Blk &B = blks[t.x];
    const Job &J = jobs[B.layer];
    const int lane = threadIdx.x;
    const int j0 = t.y * 64 * F4K + 1, j1 = j0 + 64 * F4K - 1;
    const float dvr4f = (float)J.dvr4;
    double acc[F4K];
#pragma unroll
    for (int k = 0; k < F4K; k++) acc[k] = 0.;
    unsigned long long nev = 0;   // (line, point) pairs of this tile, one record per lane
    float node[F4NF];             // this lane's far-field sums at the tile's Chebyshev nodes
#pragma unroll
    for (int i = 0; i < F4NF; i++) node[i] = 0.f;
    bool anyfar = false;
    if (j0 <= B.nptr4) {
        int64_t c0, c1;
        f4_candidates(J, B, sh, j0, min(j1, B.nptr4), &c0, &c1);
        constexpr int W = (int)(sizeof(F4Rec32) / 16);
        const u32x4 *r = reinterpret_cast<const u32x4 *>(recs + B.rec - B.lo);
        __shared__ u32x4 srec[64 * W];
        u32x4 pre[W];
        const int64_t last = c1 - 1;
        if (c1 > c0) {
            const int64_t k = c0 + lane < last ? c0 + lane : last;
#pragma unroll
            for (int w = 0; w < W; w++) pre[w] = r[k * W + w];
        }
        for (int64_t base = c0; base < c1; base += 64) {
            const int nb = (int)((c1 - base) < 64 ? (c1 - base) : 64);
            F4Plan P;
            f4_plan(P, lane < nb, (int)pre[1].y, (int)pre[1].z, (int)pre[2].x, (int)pre[2].y, (int)pre[1].w, j0, j1,
                    (int)pre[1].x, far_tau ? dfar : 0);
            if (lane < nb) nev += f4_count((int)pre[1].y, (int)pre[1].z, j0, j1);
            if (P.far) {
                anyfar = true;
                if ((P.far >> lane) & 1ull) {  // this lane's own record, at the nodes
                    const float d = __uint_as_float(pre[0].x), sil = __uint_as_float(pre[0].y);
                    const float alfli2 = __uint_as_float(pre[0].z), b0 = __uint_as_float(pre[0].w);
                    const float b1 = __uint_as_float(pre[2].z);
                    const float off = (float)(j0 - (int)pre[1].x);   // exact
#pragma unroll
                    for (int i = 0; i < F4NF; i++) {
                        const float xm = __fmaf_rn(off + far_tau[i], dvr4f, d), xmsq = xm * xm;
                        const float f = __fmaf_rn(sil, __builtin_amdgcn_rcpf(alfli2 + xmsq), -b0);
                        node[i] += __fmaf_rn(-xmsq, b1, f);
                    }
                }
            }
            __syncthreads();
            if (lane < nb) {
#pragma unroll
                for (int w = 0; w < W; w++) srec[lane * W + w] = pre[w];
            }
            __syncthreads();
            {
                const int64_t k = base + 64 + lane < last ? base + 64 + lane : last;
#pragma unroll
                for (int w = 0; w < W; w++) pre[w] = r[k * W + w];
            }
            // only the records with work at the points: those that reach the tile and are
            // not far, and the mirror images (a scalar bit scan, so far records and
            // records that miss the tile cost nothing here)
            unsigned long long need = P.touch | P.mirror;
            // whole-tile Lorentz records without the CO2 term or a mirror, taken two at a
            // time: s1/p1 + s2/p2 = (s1 p2 + s2 p1) / (p1 p2), one reciprocal per pair
            // (p <= 630 cm-2 inside the 25 cm-1 range, so p1 p2 stays far from overflow),
            // their bounds b0 subtracted once per chunk
            const unsigned long long pairable = P.all & ~P.co2 & ~P.mirror;
            while (need) {
            // FP32 partial sums over F4CHUNK records, then one FP64 accumulation
            float a32[F4K];
#pragma unroll
            for (int k = 0; k < F4K; k++) a32[k] = 0.f;
            float bsum = 0.f;
            int c = 0;
            for (unsigned long long pr = need & pairable; c + 1 < F4CHUNK && __popcll(pr) >= 2; c += 2) {
                const int q1 = __builtin_ctzll(pr);
                pr &= pr - 1;
                const int q2 = __builtin_ctzll(pr);
                pr &= pr - 1;
                need &= ~((1ull << q1) | (1ull << q2));
                const u32x4 h1 = srec[q1 * W], h2 = srec[q2 * W];
                const float d1 = __uint_as_float(h1.x), s1 = __uint_as_float(h1.y), a1 = __uint_as_float(h1.z);
                const float d2 = __uint_as_float(h2.x), s2 = __uint_as_float(h2.y), a2 = __uint_as_float(h2.z);
                bsum += __uint_as_float(h1.w) + __uint_as_float(h2.w);
                const float fl1 = (float)(j0 + lane - (int)srec[q1 * W + 1].x);
                const float fl2 = (float)(j0 + lane - (int)srec[q2 * W + 1].x);
#pragma unroll
                for (int k = 0; k < F4K; k++) {
                    const float x1 = __fmaf_rn(fl1 + (float)(64 * k), dvr4f, d1);
                    const float x2 = __fmaf_rn(fl2 + (float)(64 * k), dvr4f, d2);
                    const float p1 = __fmaf_rn(x1, x1, a1), p2 = __fmaf_rn(x2, x2, a2);
                    const float num = __fmaf_rn(s1, p2, s2 * p1);
                    a32[k] = __fmaf_rn(num, __builtin_amdgcn_rcpf(p1 * p2), a32[k]);
                }
            }
            for (; c < F4CHUNK && need; c++) {
                const int q = __builtin_ctzll(need);
                need &= need - 1;
                const unsigned long long bit = 1ull << q;
                const u32x4 *rq = &srec[q * W];
                const u32x4 h0 = rq[0];
                const float d = __uint_as_float(h0.x), sil = __uint_as_float(h0.y);
                const float alfli2 = __uint_as_float(h0.z), b0 = __uint_as_float(h0.w);
                const float fl = (float)(j0 + lane - (int)rq[1].x);
                if (P.all & bit) {
                    // the whole tile in the Lorentz zone: sil/(alfli2 + x^2) - base(x)
                    if (P.co2 & bit) {
                        const float b1 = __uint_as_float(rq[2].z);
#pragma unroll
                        for (int k = 0; k < F4K; k++) {
                            const float xm = __fmaf_rn(fl + (float)(64 * k), dvr4f, d), xmsq = xm * xm;
                            const float f = __fmaf_rn(sil, __builtin_amdgcn_rcpf(alfli2 + xmsq), -b0);
                            a32[k] += __fmaf_rn(-xmsq, b1, f);
                        }
                    } else {
#pragma unroll
                        for (int k = 0; k < F4K; k++) {
                            const float xm = __fmaf_rn(fl + (float)(64 * k), dvr4f, d);
                            a32[k] += __fmaf_rn(sil, __builtin_amdgcn_rcpf(__fmaf_rn(xm, xm, alfli2)), -b0);
                        }
                    }
                } else if (P.cover & bit) {
                    const F4Rec32 &l = *reinterpret_cast<const F4Rec32 *>(rq);
#pragma unroll
                    for (int k = 0; k < F4K; k++)
                        a32[k] += f4_eval32(l, __fmaf_rn(fl + (float)(64 * k), dvr4f, d), l.dptrat);
                } else if (P.lorp & bit) {
                    // part of the tile, all of it in the Lorentz zone
                    const F4Rec32 &l = *reinterpret_cast<const F4Rec32 *>(rq);
                    const int jmn = l.jmin, jmx = l.jmax;
                    const float b1 = l.b1;
#pragma unroll
                    for (int k = 0; k < F4K; k++) {
                        const int j = j0 + lane + 64 * k;
                        const float xm = __fmaf_rn(fl + (float)(64 * k), dvr4f, d), xmsq = xm * xm;
                        const float f = __fmaf_rn(-xmsq, b1, __fmaf_rn(sil, __builtin_amdgcn_rcpf(alfli2 + xmsq), -b0));
                        a32[k] += ((j >= jmn && j <= jmx) ? f : 0.f);
                    }
                } else if (P.touch & bit) {
                    const F4Rec32 &l = *reinterpret_cast<const F4Rec32 *>(rq);
                    const int jmn = l.jmin, jmx = l.jmax;
#pragma unroll
                    for (int k = 0; k < F4K; k++) {
                        const int j = j0 + lane + 64 * k;
                        float f = f4_eval32(l, __fmaf_rn(fl + (float)(64 * k), dvr4f, d), l.dptrat);
                        a32[k] += ((j >= jmn && j <= jmx) ? f : 0.f);
                    }
                }
                if (P.mirror & bit) {  // Van Vleck-Weisskopf mirror (lines below 25 cm-1)
                    const F4Rec32 &l = *reinterpret_cast<const F4Rec32 *>(rq);
                    const float flm = (float)(j0 + lane - l.jcm);
#pragma unroll
                    for (int k = 0; k < F4K; k++) {
                        const int j = j0 + lane + 64 * k;
                        float f = f4_eval32(l, __fmaf_rn(flm + (float)(64 * k), dvr4f, l.dm), -l.dptrat);
                        a32[k] += ((j >= l.jminm && j <= l.jmaxm) ? f : 0.f);
                    }
                }
            }
#pragma unroll
            for (int k = 0; k < F4K; k++) acc[k] += (double)(a32[k] - bsum);
            }
        }
    }
    if (anyfar) {  // wave-uniform: sum the nodes over the lanes, interpolate to the points
        double nd[F4NF];
#pragma unroll
        for (int i = 0; i < F4NF; i++) {
            double v = (double)node[i];
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
            nd[i] = v;
        }
#pragma unroll
        for (int k = 0; k < F4K; k++) {
            const double *w = far_w + (lane + 64 * k) * F4NF;
            double s = 0.;
#pragma unroll
            for (int i = 0; i < F4NF; i++) s = fma(w[i], nd[i], s);
            acc[k] += s;
        }
    }
#pragma unroll
    for (int k = 0; k < F4K; k++) {
        const int j = j0 + lane + 64 * k;
        if (j <= LIMOUT4 + 2) R4[B.r4_off + j - 1] = (j <= B.nptr4) ? acc[k] : 0.;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) nev += __shfl_xor(nev, o);
    if (lane == 0 && nev) atomicAdd(work + 1, nev);
}

// LBLF4Q radiation term on R4 when JRAD = 1 (clblm_ODLAY_LineF4.f90:142-177): one
// workgroup per R4 block; one lane turns the RADFNI segments into exact ramp pieces,
// all lanes apply them.
__global__ void __launch_bounds__(256) k_r4rad(const Job *jobs, const R4Blk *blks, double *R4) {
    __shared__ double rbase[RAMP_CAP], rstep[RAMP_CAP];
    __shared__ int rk[RAMP_CAP + 1], nramp;
    const R4Blk &B = blks[blockIdx.x];
    const Job &J = jobs[B.layer];
    if (J.jrad != 1) return;
    double *r = R4 + B.r4_off - 1;
    const int tid = threadIdx.x;
    if (tid == 0) {
        const double xkt = 1. / (RADCN2 / J.tave);
        double vitst = B.v1r4 - J.dvr4, rdlast = -1., rdel;
        int n2 = 0, np = 0;
        do {
            int n1 = n2 + 1;
            double vi = B.v1r4 + J.dvr4 * (double)(n1 - 1);
            double radvi = radfni(vi, J.dvr4, xkt, &vitst, &rdel, &rdlast);
            n2 = (int)((vitst - B.v1r4) / J.dvr4 + 0.001);
            if (n2 > B.nptr4) n2 = B.nptr4;
            int c = seqramp::ramp_pieces(radvi, rdel, n2 - n1 + 1, RAMP_CAP - np, [&](int k, double b, double st) {
                rk[np] = n1 + k; rbase[np] = b; rstep[np] = st; np++;
            });
            if (c < 0) { np = -1; break; }
        } while (n2 < B.nptr4);
        if (np < 0) {  // more pieces than the table holds: the reference's loop, in place
            vitst = B.v1r4 - J.dvr4;
            rdlast = -1.;
            n2 = 0;
            do {
                int n1 = n2 + 1;
                double vi = B.v1r4 + J.dvr4 * (double)(n1 - 1);
                double tt = radfni(vi, J.dvr4, xkt, &vitst, &rdel, &rdlast);
                n2 = (int)((vitst - B.v1r4) / J.dvr4 + 0.001);
                if (n2 > B.nptr4) n2 = B.nptr4;
                for (int i = n1; i <= n2; i++) { r[i] = r[i] * tt; tt = tt + rdel; }
            } while (n2 < B.nptr4);
            np = 0;
        }
        rk[np] = B.nptr4 + 1;
        nramp = np;
    }
    __syncthreads();
    const int np = nramp;
    if (np == 0) return;
    int pc = 0;
    for (int i = 1 + tid; i <= B.nptr4; i += blockDim.x) {
        while (rk[pc + 1] <= i) pc++;
        r[i] = r[i] * (rbase[pc] + (double)(i - rk[pc]) * rstep[pc]);
    }
}

// ------------------------------------------------------------------------------------
// Pass 2: one workgroup per panel rebuilds the reference's R1/R2/R3 panel buffers at
// the moment panelOut runs (clblm_ODLAY.f90:2289-2545): own contributions, the carry
// of the previous panel (shifted by 2400/600/150 points), XINT of the active R4 block
// and of the continuum into R3 (clblm_Spectrum.f90:874-914), then the R3->R2->R1
// Lagrange steps, the RADFNI radiation term and the store.  R3 points 1..5 and R2
// points 8..12 hold values the previous panel produced; they are rebuilt from that
// panel's lines and F4 block only.
// ------------------------------------------------------------------------------------
template <typename TA>
struct PanelArr;
template <>
struct PanelArr<double> {  // FP64: one array of stride doubles per panel
    const double *p;
    int m1, m2;
    __device__ PanelArr(const double *A, int64_t off, int max1, int max2, int) : p(A + off - 1), m1(max1), m2(max2) {}
    __device__ double r1(int j) const { return p[j]; }
    __device__ double r2(int j) const { return p[m1 + j]; }
    __device__ double r3(int j) const { return p[m1 + m2 + j]; }
};
__device__ inline double xint_point(double v1a, double v2a, double dva, const double *A /*1-based*/,
                                    double vft, double dvr3, int I, int n1r3, int n2r3) {
    double recdva = 1. / dva;
    int ilo = (int)((v1a + dva - vft) / dvr3 + 1. + ONEMI);
    if (ilo < n1r3) ilo = n1r3;
    int ihi = (int)((v2a - dva - vft) / dvr3 + ONEMI);
    if (ihi > n2r3) ihi = n2r3;
    if (I < ilo || I > ihi) return 0.;
    double vi = vft + dvr3 * (double)(I - 1);
    int j = (int)((vi - v1a) * recdva + ONEPL);
    double vj = v1a + dva * (double)(j - 1);
    double p = recdva * (vi - vj);
    double c = (3. - 2. * p) * p * p;
    double b = 0.5 * p * (1. - p);
    double b1 = b * (1. - p);
    double b2 = b * p;
    double conti = -A[j - 1] * b1 + A[j] * (1. - c + b2) + A[j + 1] * (c + b1) - A[j + 2] * b2;
    return conti * 1.0;
}

// DRAD: the temperature derivative of the radiation term instead of the term itself,
// scaled by dscale and added to out (JRAD 0 layers only): with the panel sums of the
// line-sum OD this is the product rule's d(RADFNI)/dT part of the OD's temperature
// derivative.  RADFNI's nodes stay where the term puts them; each segment's ramp
// runs from d RADFN/dT at its start node to that at its end node.
template <typename TA, bool DRAD = false>
__global__ void __launch_bounds__(256) k_pass2(const Job *jobs, const Panel *panels, const R4Blk *blks,
                                               const TA *A, const double *R4, const double *absrb,
                                               double *out, double *nat, double dscale = 0., int64_t s_a = 0,
                                               int64_t s_r4 = 0, int64_t s_abs = 0, int64_t s_out = 0,
                                               int64_t s_nat = 0) {
    __shared__ double r3[168], r3e[8], r2[640], r1[2448];
    {   // blockIdx.y: a channel of the tangent batch, its arrays s_* further on
        const int64_t ch = blockIdx.y;
        A += ch * s_a; R4 += ch * s_r4; absrb += ch * s_abs; out += ch * s_out; nat += ch * s_nat;
    }
    __shared__ double rbase[RAMP_CAP], rstep[RAMP_CAP];
    __shared__ int rk[RAMP_CAP + 1], nramp;
    const int64_t q = blockIdx.x;
    const Panel &P = panels[q];
    const Job &J = jobs[P.layer];
    if (DRAD && J.jrad != 0) return;
    const int p = (int)(q - J.panel_off);
    const bool prev = p > 0;
    const Panel &PP = panels[prev ? q - 1 : q];
    const int max1 = J.max1, max2 = J.max2, max3 = J.max3;
    const PanelArr<TA> pa(A, P.a_off, max1, max2, J.stride), pb(A, PP.a_off, max1, max2, J.stride);
    auto a1 = [&](int j) { return pa.r1(j); };
    auto a2 = [&](int j) { return pa.r2(j); };
    auto a3 = [&](int j) { return pa.r3(j); };
    auto b1 = [&](int j) { return pb.r1(j); };
    auto b2 = [&](int j) { return pb.r2(j); };
    auto b3 = [&](int j) { return pb.r3(j); };
    const R4Blk &B = blks[P.blk];
    const R4Blk &BP = blks[PP.blk];
    const int N1R3 = prev ? 6 : 1, N1R3P = (prev && p > 1) ? 6 : 1, N2R3 = 155;
    const double DVR3 = J.dvr3;
    const double *r4 = R4 + B.r4_off - 1, *r4p = R4 + BP.r4_off - 1;
    const double *ab = J.contnm ? absrb + J.absrb_off - 1 : nullptr;
    const double v2abs = J.v1abs + (double)(J.nptabs - 1);
    const int tid = threadIdx.x;
    const int NLO = NSHIFT + 1, NHI = P.nhi;
    // R3
    for (int j = 1 + tid; j <= 160; j += blockDim.x) {
        double v = (j <= max3 ? a3(j) : 0.) + (prev && j + 150 <= max3 ? b3(j + 150) : 0.);
        if (j >= N1R3 && j <= N2R3) {
            v = v + xint_point(B.v1r4, B.v2r4, J.dvr4, r4, P.vft, DVR3, j, N1R3, N2R3);
            if (ab) v = v + xint_point(J.v1abs, v2abs, 1., ab, P.vft, DVR3, j, N1R3, N2R3);
        }
        if (prev && j <= 5) {
            double xp = xint_point(BP.v1r4, BP.v2r4, J.dvr4, r4p, PP.vft, DVR3, j + 150, N1R3P, N2R3);
            if (ab) xp = xp + xint_point(J.v1abs, v2abs, 1., ab, PP.vft, DVR3, j + 150, N1R3P, N2R3);
            r3e[j] = (j + 150 <= max3 ? b3(j + 150) : 0.) + xp;
            v = v + xp;
        }
        r3[j] = v;
    }
    for (int j = 1 + tid; j < 640; j += blockDim.x)
        r2[j] = (j <= max2 ? a2(j) : 0.) + (prev && j + 600 <= max2 ? b2(j + 600) : 0.);
    __syncthreads();
    const double X00 = -7. / 128., X01 = 105. / 128., X02 = 35. / 128., X03 = -5. / 128.;
    const double X10 = -1. / 16., X11 = 9. / 16.;
    // R3 -> R2
    const int LIMLO = prev ? 13 : 5, LIMHI = NHI / 4 + 1;
    const int ngr = LIMLO <= LIMHI ? (LIMHI - LIMLO) / 4 + 1 : 0;
    if (prev && tid < 2) {  // groups J = 605, 609 of the previous panel (local 5..12)
        int Jl = 5 + 4 * tid, J3 = 152 + tid;
        const double *e = r3e - 150;  // e[151..155]
        r2[Jl] = r2[Jl] + e[J3];
        r2[Jl + 1] = r2[Jl + 1] + X00 * e[J3 - 1] + X01 * e[J3] + X02 * e[J3 + 1] + X03 * e[J3 + 2];
        r2[Jl + 2] = r2[Jl + 2] + X10 * (e[J3 - 1] + e[J3 + 2]) + X11 * (e[J3] + e[J3 + 1]);
        r2[Jl + 3] = r2[Jl + 3] + X03 * e[J3 - 1] + X02 * e[J3] + X01 * e[J3 + 1] + X00 * e[J3 + 2];
    }
    for (int g = tid; g < ngr; g += blockDim.x) {
        int Jg = LIMLO + 4 * g, J3 = (Jg - 1) / 4 + 1;
        r2[Jg] = r2[Jg] + r3[J3];
        r2[Jg + 1] = r2[Jg + 1] + X00 * r3[J3 - 1] + X01 * r3[J3] + X02 * r3[J3 + 1] + X03 * r3[J3 + 2];
        r2[Jg + 2] = r2[Jg + 2] + X10 * (r3[J3 - 1] + r3[J3 + 2]) + X11 * (r3[J3] + r3[J3 + 1]);
        r2[Jg + 3] = r2[Jg + 3] + X03 * r3[J3 - 1] + X02 * r3[J3] + X01 * r3[J3 + 1] + X00 * r3[J3 + 2];
    }
    __syncthreads();
    if (P.istop && ngr > 0 && tid == 0) {
        int Jg = LIMLO + 4 * ngr, J3 = (Jg - 1) / 4 + 1;
        r2[Jg] = r2[Jg] + r3[J3];
        r2[Jg + 1] = r2[Jg + 1] + X00 * r3[J3 - 1] + X01 * r3[J3] + X02 * r3[J3 + 1] + X03 * r3[J3 + 2];
    }
    // R1
    for (int j = NLO + tid; j <= NHI + 3; j += blockDim.x)
        r1[j] = (j <= max1 ? a1(j) : 0.) + (prev && j + 2400 <= max1 ? b1(j + 2400) : 0.);
    __syncthreads();
    for (int g = tid; NLO + 4 * g <= NHI; g += blockDim.x) {
        int Jg = NLO + 4 * g, J2 = (Jg - 1) / 4 + 1;
        r1[Jg] = r1[Jg] + r2[J2];
        r1[Jg + 1] = r1[Jg + 1] + X00 * r2[J2 - 1] + X01 * r2[J2] + X02 * r2[J2 + 1] + X03 * r2[J2 + 2];
        r1[Jg + 2] = r1[Jg + 2] + X10 * (r2[J2 - 1] + r2[J2 + 2]) + X11 * (r2[J2] + r2[J2 + 1]);
        r1[Jg + 3] = r1[Jg + 3] + X03 * r2[J2 - 1] + X02 * r2[J2] + X01 * r2[J2 + 1] + X00 * r2[J2 + 2];
    }
    __syncthreads();
    // Radiation term (panelOut, clblm_ODLAY.f90:2447-2483): RADFNI segments whose
    // sequential ramps are turned into exact linear pieces by one lane, then applied
    // by all lanes at the store.
    __shared__ int inplace_from;
    if (DRAD && tid == 0) {
        const double xkt = J.tave / RADCN2, DV = J.dv;
        double vitst = P.vft + (double)(NLO - 1) * DV - DV, rdlast = -1., rdel, drlast = 0.;
        int n2 = NLO - 1, np = 0;
        inplace_from = NHI + 1;
        do {
            int n1 = n2 + 1;
            double vi = P.vft + (double)(n1 - 1) * DV;
            const bool first = rdlast < 0.;
            (void)radfni(vi, DV, xkt, &vitst, &rdel, &rdlast);
            n2 = (int)((vitst - P.vft) / DV + 0.001);
            if (n2 > NHI) n2 = NHI;
            const double d0 = first ? dradfn_dt(vi, J.tave) : drlast;
            int intvls = (int)((vitst - vi) / DV + 0.5);
            if (intvls < 1) intvls = 1;
            const double dn = dradfn_dt(vitst, J.tave);
            drlast = dn;
            const double dstep = (dn - d0) / (double)intvls;
            if (np < RAMP_CAP && inplace_from > NHI) {
                rk[np] = n1; rbase[np] = d0 * dscale; rstep[np] = dstep * dscale; np++;
            } else {   // more segments than the table holds: applied in place
                if (inplace_from > NHI) inplace_from = n1;
                for (int i = n1; i <= n2; i++) r1[i] = r1[i] * ((d0 + (double)(i - n1) * dstep) * dscale);
            }
        } while (n2 < NHI);
        rk[np] = inplace_from;
        nramp = np;
    }
    if (!DRAD && J.jrad == 0) {
        if (tid == 0) {
            const double xkt = J.tave / RADCN2, DV = J.dv;
            double vitst = P.vft + (double)(NLO - 1) * DV - DV, rdlast = -1., rdel;
            int n2 = NLO - 1, np = 0;
            do {
                int n1 = n2 + 1;
                double vi = P.vft + (double)(n1 - 1) * DV;
                double radvi = radfni(vi, DV, xkt, &vitst, &rdel, &rdlast);
                n2 = (int)((vitst - P.vft) / DV + 0.001);
                if (n2 > NHI) n2 = NHI;
                int c = seqramp::ramp_pieces(radvi, rdel, n2 - n1 + 1, RAMP_CAP - np,
                                             [&](int k, double b, double st) {
                                                 rk[np] = n1 + k; rbase[np] = b; rstep[np] = st; np++;
                                             });
                if (c < 0) { np = -1; break; }
            } while (n2 < NHI);
            if (np < 0) {  // more pieces than the table holds: the reference's loop, in place
                vitst = P.vft + (double)(NLO - 1) * DV - DV;
                rdlast = -1.;
                n2 = NLO - 1;
                do {
                    int n1 = n2 + 1;
                    double vi = P.vft + (double)(n1 - 1) * DV;
                    double tt = radfni(vi, DV, xkt, &vitst, &rdel, &rdlast);
                    n2 = (int)((vitst - P.vft) / DV + 0.001);
                    if (n2 > NHI) n2 = NHI;
                    for (int i = n1; i <= n2; i++) { r1[i] = r1[i] * tt; tt = tt + rdel; }
                } while (n2 < NHI);
                np = 0;
            }
            rk[np] = NHI + 1;
            nramp = np;
        }
        __syncthreads();
    }
    if (DRAD) __syncthreads();
    const int np = J.jrad == 0 ? nramp : 0;
    double *o = (J.dvout > 0. ? nat : out) + P.out_off - NLO;
    int pc = 0;
    for (int i = NLO + tid; i <= NHI; i += blockDim.x) {
        double v = r1[i];
        if (np > 0 && (!DRAD || i < inplace_from)) {
            while (rk[pc + 1] <= i) pc++;
            v = v * (rbase[pc] + (double)(i - rk[pc]) * rstep[pc]);
        }
        if (DRAD) o[i] = o[i] + v;
        else o[i] = v;
    }
}

// interpSpectrum_array2array (clblm_Spectrum.f90:326-590).  For the n/(n+1) ratios
// every output point is independent (one lane per point).  The general case walks
// blocks of 2400 output points with a sequentially accumulated fraction P, whose
// rounding the reference's quantised coefficient index depends on; a lane takes a
// piece of a block and replays P from the block's start.  unit_first[l]: the first
// work unit of layer l (nondecreasing; layers without re-grid have none).
// blockIdx.y selects a channel of the tangent batch: nat and out advance by their strides
__global__ void k_regrid(const Job *jobs, int nlay, const int64_t *unit_first, const double *nat, double *out,
                         int64_t nunits, int64_t nat_stride = 0, int64_t out_stride = 0) {
    int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= nunits) return;
    nat += (int64_t)blockIdx.y * nat_stride;
    out += (int64_t)blockIdx.y * out_stride;
    int l = 0, hi = nlay - 1;          // the last layer whose units start at or before u
    while (l < hi) {
        const int mid = (l + hi + 1) / 2;
        if (unit_first[mid] <= u) l = mid;
        else hi = mid - 1;
    }
    const Job &J = jobs[l];
    clblm_interp::unit(nat + J.nat_off - 1, J.nnat, J.v1, J.dv, J.dvout, out + J.out_off - 1, J.npts,
                       u - unit_first[l]);
}

// ====================================================================================
// Host side
// ====================================================================================
namespace {

// Line-shape tables (SHAPEL clblm_ODLAY.f90:2804-2909, SHAPEG :2914-2953, VERFN
// :2959-3016), built on the host in the reference's operation order.
// The constants of build_tables: 1/pi and A, B at z = 4, 16, 64.
void shape_constants(double *sc) {
    auto Af = [](double z) { return (1. + 2. * z * z) / ((1. + z * z) * (1. + z * z)); };
    auto Bf = [](double z) { return -1. / ((1. + z * z) * (1. + z * z)); };
    sc[0] = 1. / (2. * std::asin(1.));
    sc[1] = Af(4.); sc[2] = Bf(4.); sc[3] = Af(16.); sc[4] = Bf(16.); sc[5] = Af(64.); sc[6] = Bf(64.);
    sc[7] = 0.;
}

void build_tables(std::vector<double> &F1, std::vector<double> &F2, std::vector<double> &F3,
                  std::vector<double> &FG, std::vector<double> &XV) {
    const double HWF1 = 4., HWF2 = 16., HWF3 = 64., DXF1 = 0.002, DXF2 = 0.008, DXF3 = 0.032;
    F1.assign(NFMX + 1, 0.); F2.assign(NFMX + 1, 0.); F3.assign(NFMX + 1, 0.);
    FG.assign(NFMX + 1, 0.); XV.assign(NFMX + 1, 0.);
    auto Af = [](double z) { return (1. + 2. * z * z) / ((1. + z * z) * (1. + z * z)); };
    auto Bf = [](double z) { return -1. / ((1. + z * z) * (1. + z * z)); };
    const double recpi = 1. / (2. * std::asin(1.));
    const double a1 = Af(HWF1), b1 = Bf(HWF1), a2 = Af(HWF2), b2 = Bf(HWF2), a3 = Af(HWF3), b3 = Bf(HWF3);
    auto lor = [](double x) { return 1. / (1. + x); };
    F1[1] = recpi * (lor(0.) - (a1 + b1 * 0.));
    for (int j = 2; j <= NFPTS; j++) {
        double x = (double)(j - 1) * DXF1, xs = x * x;
        F1[j] = recpi * (lor(xs) - (a1 + b1 * xs));
    }
    F1[NFPTS] = 0.;
    F2[1] = recpi * ((a1 + b1 * 0.) - (a2 + b2 * 0.));
    int lim = (int)(HWF1 / DXF2 + 1.001);
    for (int j = 2; j <= lim; j++) {
        double x = (double)(j - 1) * DXF2, xs = x * x;
        F2[j] = recpi * ((a1 + b1 * xs) - (a2 + b2 * xs));
    }
    for (int j = lim + 1; j <= NFPTS; j++) {
        double x = (double)(j - 1) * DXF2, xs = x * x;
        F2[j] = recpi * (lor(xs) - (a2 + b2 * xs));
    }
    F2[NFPTS] = 0.;
    F3[1] = recpi * ((a2 + b2 * 0.) - (a3 + b3 * 0.));
    lim = (int)(HWF2 / DXF3 + 1.001);
    for (int j = 2; j <= lim; j++) {
        double x = (double)(j - 1) * DXF3, xs = x * x;
        F3[j] = recpi * ((a2 + b2 * xs) - (a3 + b3 * xs));
    }
    for (int j = lim + 1; j <= NFPTS; j++) {
        double x = (double)(j - 1) * DXF3, xs = x * x;
        F3[j] = recpi * (lor(xs) - (a3 + b3 * xs));
    }
    const double fln2 = std::log(2.);
    const double fgnorm = std::sqrt(fln2 * recpi);
    FG[1] = fgnorm * std::exp(-fln2 * 0.);
    for (int j = 2; j <= NFPTS; j++) {
        double x = (double)(j - 1) * DXF1, xs = x * x;
        FG[j] = fgnorm * std::exp(-(fln2 * xs));
    }
    FG[NFPTS] = 0.;
    const double cexp = 0.45, ce0 = 1., ce2 = -.20737285249, ce4 = -.00872684335747;
    const double se0 = std::sqrt(cexp / PI_REF), ae2 = 2. * cexp, ae4 = ae2 * ae2 / 3.;
    for (int i = 1; i <= NFMX; i++) {
        double z = DXF1 * (double)(i - 1), z2 = z * z;
        double xe0 = se0 * std::exp(-(cexp * z2));
        XV[i] = (1. / (ce0 + ce2 + ce4)) * (ce0 + ce2 * ae2 * z2 + ce4 * ae4 * z2 * z2) * xe0;
    }
}

// TIPS_2003 4-point Lagrange recall (clblm_ODLAY_TIPS.f90:5451-5547).
double tips_interp(double t, const double *q) {
    const double *A = CLBLM_TIPS_TDAT;
    const int n = CLBLM_NT;
    auto a = [&](int i) { return A[i - 1]; };
    auto b = [&](int i) { return q[i - 1]; };
    auto nz = [](double d) { return d == 0. ? 0.0001 : d; };
    for (int i = 2; i <= n; i++) {
        if (a(i) >= t) {
            int j = i;
            if (i < 3 || i == n) {
                if (i < 3) j = 3;
                double a0 = (t - a(j - 1)) * (t - a(j)) / (nz(a(j - 2) - a(j - 1)) * nz(a(j - 2) - a(j)));
                double a1 = (t - a(j - 2)) * (t - a(j)) / (nz(a(j - 1) - a(j - 2)) * nz(a(j - 1) - a(j)));
                double a2 = (t - a(j - 2)) * (t - a(j - 1)) / (nz(a(j) - a(j - 2)) * nz(a(j) - a(j - 1)));
                return a0 * b(j - 2) + a1 * b(j - 1) + a2 * b(j);
            }
            double c0 = (t - a(j - 1)) * (t - a(j)) * (t - a(j + 1));
            c0 = c0 / (nz(a(j - 2) - a(j - 1)) * nz(a(j - 2) - a(j)) * nz(a(j - 2) - a(j + 1)));
            double c1 = (t - a(j - 2)) * (t - a(j)) * (t - a(j + 1));
            c1 = c1 / (nz(a(j - 1) - a(j - 2)) * nz(a(j - 1) - a(j)) * nz(a(j - 1) - a(j + 1)));
            double c2 = (t - a(j - 2)) * (t - a(j - 1)) * (t - a(j + 1));
            c2 = c2 / (nz(a(j) - a(j - 2)) * nz(a(j) - a(j - 1)) * nz(a(j) - a(j + 1)));
            double c3 = (t - a(j - 2)) * (t - a(j - 1)) * (t - a(j));
            c3 = c3 / (nz(a(j + 1) - a(j - 2)) * nz(a(j + 1) - a(j - 1)) * nz(a(j + 1) - a(j)));
            return c0 * b(j - 2) + c1 * b(j - 1) + c2 * b(j) + c3 * b(j + 1);
        }
    }
    return 0.;
}

struct DevArena {
    char *p = nullptr;
    size_t cap = 0, used = 0;
    int ensure(size_t bytes) {
        if (bytes <= cap) return 0;
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        size_t want = bytes + bytes / 8;
        if (hipMalloc(&p, want) != hipSuccess) { p = nullptr; return -1; }
        cap = want;
        return 0;
    }
};

struct Carve {
    size_t off = 0;
    template <typename T>
    size_t take(size_t n) {
        size_t o = (off + 255) & ~(size_t)255;
        off = o + n * sizeof(T);
        return o;
    }
};

}  // namespace

struct LineListH {
    int64_t n = 0;
    std::vector<double> nu;
    void *mem = nullptr;
    void *brd_mem = nullptr;   // species-broadening fields, if attached
    bool has_cpl = false;      // some line has IFLG 1 (Y and G)
    int graphs = 0;            // captured graphs that read this list's device arrays
    LinesDev dev{};
};

// ---- MT_CKD water-vapour continuum ------------------------------------------------
// CONTNM's H2O part (clblm_ODLAY_Continuum.f90:356-404) with mt_ckd_h2o_absco
// (mt_ckd_h2o_module.f90:27-103): per layer, the coefficient window around
// [V1ABS, V2ABS] scaled to the layer (the self coefficients by (T0/T)**texp, both by
// the radiation term unless JRAD is 0) is built on the host with the C library's
// pow and exp, as the reference's build evaluates them; the 4-point interpolation
// onto the 1 cm-1 ABSRB grid (myxint :231-271, bounds by pre_xint :273-292) and the
// sum absrb + W_H2O sh2o xself + W_H2O fh2o xfrgn run on the device.
struct MtckdTable {
    std::vector<double> wvn, self, forr, texp;
    double ref_t = 0., ref_p = 0., xself = 1., xfrgn = 1.;
    double cn_scale[5] = {1., 1., 1., 1., 1.};   // XCO2C, XO3CN, XO2CN, XN2CN, XRAYL
    int64_t gen = 0;             // bumped by every load and every change of the factors
};

struct MtJob {
    int64_t coef;     // its self window in the batch's coefficient array, foreign after it
    int64_t abs_off;  // its ABSRB in the batch's array
    int32_t ncoeff, ilo, ihi, pad;
    double v1a, v2a, dva, vft, wk1, xself, xfrgn;
};

// OPDPTH's ABSRB grid (clblm_ODLAY.f90:880-887)
static void absrb_grid(double V1, double V2, double *v1abs, int *npt) {
    double a = (double)(int)V1;
    if (V1 < 0.) a = a - 1.;
    a = a - 3.;
    const double b = (double)(int)(V2 + 3. + 0.5);
    *v1abs = a;
    *npt = (int)((b - a) / 1. + 1.5);
}

// One layer's MT_CKD job; returns 0 with *has = false where CONTNM adds nothing.
static int mtckd_plan(const MtckdTable &T, const clblm_layer &lay, double V1, double V2, int jrad, MtJob &mj,
                      std::vector<double> &coef, bool *has) {
    *has = false;
    if (T.wvn.size() < 4) return set_err(CLBLM_ERR_ARG, "MT_CKD continuum asked for but no table loaded");
    double V1ABS;
    int NPTABS;
    absrb_grid(V1, V2, &V1ABS, &NPTABS);
    if (NPTABS > 5050) return set_err(CLBLM_ERR_ARG, "MT_CKD continuum: ABSRB longer than 5050 points");
    if (!((V2 > -20.0) && (V1 < 20000.) && ((T.xself > 0.) || (T.xfrgn > 0.)))) return 0;
    const double V2ABS = V1ABS + (double)(NPTABS - 1);
    const double *wvn = T.wvn.data();
    const int n = (int)T.wvn.size();
    const double h2o_fac = lay.w[1] / lay.wtot;
    const bool radflag = jrad != 0;
    const double dvc = wvn[1] - wvn[0];
    int i = 1;
    while (i <= n && wvn[i - 1] <= (V1ABS - 2 * dvc)) i++;
    int i1 = i - 1;
    if (i1 == 0) i1 = 1;
    while (i <= n && wvn[i - 1] < (V2ABS + 2 * dvc)) i++;
    if (i > n) return set_err(CLBLM_ERR_ARG, "MT_CKD continuum: the ABSRB grid reaches past the coefficient table");
    const int i2 = i, ncoeff = i2 - i1 + 1;
    const double xkt = lay.tave / RADCN2;
    const double rho_rat = (lay.pave / T.ref_p) * (T.ref_t / lay.tave);
    mj.coef = (int64_t)coef.size();
    coef.resize(coef.size() + 2 * (size_t)ncoeff);
    double *sh = coef.data() + mj.coef, *fh = sh + ncoeff;
    for (int k = 0; k < ncoeff; k++) {
        const int m = i1 - 1 + k;
        double sv = T.self[m] * std::pow(T.ref_t / lay.tave, T.texp[m]);
        sv = sv * h2o_fac * rho_rat;
        double fv = T.forr[m];
        fv = fv * (1 - h2o_fac) * rho_rat;
        if (radflag) {   // myradfn (:196-226)
            const double xvi = wvn[m];
            double rad = xvi;
            const double xviokt = xvi / xkt;
            if (xviokt <= 0.01) {
                rad = 0.5 * xviokt * xvi;
            } else if (xviokt <= 10.) {
                const double expvkt = std::exp(-xviokt);
                rad = xvi * (1 - expvkt) / (1. + expvkt);
            }
            sv = sv * rad;
            fv = fv * rad;
        }
        sh[k] = sv;
        fh[k] = fv;
    }
    // pre_xint over the whole table, then myxint's own bounds
    const int nptabs = (int)((V2ABS - V1ABS) / 1. + 1);
    const int nbnd_v1c = (int)(2 + (wvn[0] - V1ABS) / 1. + 1.e-5);
    const int ist = std::max(1, nbnd_v1c);
    const int nbnd_v2c = (int)(1 + (wvn[n - 1] - V1ABS) / 1. + 1.e-5);
    const int lst = std::min(nptabs, nbnd_v2c);
    const double V1A = wvn[i1 - 1], V2A = wvn[i2 - 1];
    int ILO = (int)((V1A + dvc - V1ABS) / 1. + 1. + 0.999);
    ILO = std::max(ILO, ist);
    int IHI = (int)((V2A - dvc - V1ABS) / 1. + 0.999);
    IHI = std::min(IHI, lst);
    mj.ncoeff = ncoeff;
    mj.ilo = ILO;
    mj.ihi = IHI;
    mj.pad = 0;
    mj.v1a = V1A;
    mj.v2a = V2A;
    mj.dva = dvc;
    mj.vft = V1ABS;
    mj.wk1 = lay.w[1];
    mj.xself = T.xself;
    mj.xfrgn = T.xfrgn;
    *has = ILO <= IHI;
    if (!*has) coef.resize((size_t)mj.coef);
    return 0;
}

// myxint of both windows at ABSRB point I (1-based) of job blockIdx.y, then CONTNM's
// absrb = absrb + cself + cfor.  Outside [ILO, IHI] the reference adds zeros.
__global__ void k_mtckd(const MtJob *jobs, const double *coef, double *absrb) {
    const MtJob J = jobs[blockIdx.y];
    const int I = J.ilo + (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (I > J.ihi) return;
    const double *A = coef + J.coef - 1, *F = A + J.ncoeff;   // 1-based windows
    const double RECDVA = 1. / J.dva;
    const double VI = J.vft + 1. * (double)(I - 1);
    const int j = (int)((VI - J.v1a) * RECDVA + 1.001);
    const double VJ = J.v1a + J.dva * (double)(j - 1);
    const double P = RECDVA * (VI - VJ);
    const double C = (3. - 2. * P) * P * P;
    const double B = 0.5 * P * (1. - P);
    const double B1 = B * (1. - P);
    const double B2 = B * P;
    const bool in3 = j + 2 <= J.ncoeff;   // one past the window only where B2 = 0
    const double s = -A[j - 1] * B1 + A[j] * (1. - C + B2) + A[j + 1] * (C + B1) - (in3 ? A[j + 2] : 0.) * B2;
    const double f = -F[j - 1] * B1 + F[j] * (1. - C + B2) + F[j + 1] * (C + B1) - (in3 ? F[j + 2] : 0.) * B2;
    const double sh2o = 0. + s * 1.0, fh2o = 0. + f * 1.0;
    double *out = absrb + J.abs_off + (I - 1);
    const double cself = J.wk1 * sh2o * J.xself;
    const double cfor = J.wk1 * fh2o * J.xfrgn;
    *out = *out + cself + cfor;
}

// ---- CONTNM's continua other than water vapour --------------------------------------
// (clblm_ODLAY_Continuum.f90:408-1060).  Each tabulated band is the window of its table
// around [V1ABS, V2ABS] that the reference's routine builds, scaled to the layer as
// CONTNM scales it: CO2 (FRNCO2 :1333-1400, XFACCO2 :415-457); O3 Chappuis (XO3CHP
// :3061-3205), Hartley-Huggins (O3HHT0-2 :5210-7085) and UV (O3HHUV :7090-7175); the O2
// fundamental (o2_ver_1 :7187-7310), 1.27 um (O2INF1 :7327-7488), 1.06 um (O2INF2
// :7494-7560), A-band (O2INF3 :7573-7682), visible (O2_vis :7699-8090), Herzberg (O2HERZ,
// HERTDA, HERPRS :8095-8253) and far-UV (O2FUV :8257-8650) continua; the N2 rotation
// (xn2_r :2533-2600), fundamental (n2_ver_1 :2701-2790) and first overtone
// (n2_overtone1 :2956-3056).  The windows are built on the host (the C library's pow,
// exp and log, as the reference's build evaluates them, and RADFN where JRAD is 1).  The
// device adds the bands in CONTNM's order into ABSRB, each by XINT (clblm_Spectrum.f90
// :874-914) over the ABSRB points [n1, n2] it may touch: pre_xint's bounds for the A-band
// and the far UV, and, for ozone, the points CONTNM's save and restore of ABSRB around
// 40800 cm-1 leave alone (:505-525, :545-565).  Then the Rayleigh extinction of every
// point (:970-1060, no aerosols).
#include "clblm_contnm.inc"

struct CnTerm {
    double v1c, v2c, dvc;
    int64_t coef;      // its window (NPTC values, then one zero) in the batch's array
    int32_t nptc, n1, n2, pad;   // XINT's N1R3, N2R3
};

struct CnJob {
    int64_t abs_off;
    double v1abs, ray_fac;                  // ray_fac: WTOT CONV_CM2MOL
    int32_t nptabs, term0, nterm, jrad, ray, pad;
};

// The table window of CONTNM's band routines around [V1ABS, V2ABS] (the same lines in
// each): returns NPTC.  eps is the offset of the index truncation (0.01; O2FUV 1e-5);
// O2HERZ does not clip NPTC at its table length (clip false).
static int cn_window_g(double V1ABS, double V2ABS, double V1S, double DVS, int NPTS, double eps, bool clip,
                       double *V1C, double *V2C, double *DVC, int *I1) {
    *DVC = DVS;
    double v1c = V1ABS - *DVC, v2c = V2ABS + *DVC;
    const int i1 = v1c < V1S ? -1 : (int)((v1c - V1S) / DVS + eps);
    v1c = V1S + DVS * (double)(i1 - 1);
    const int i2 = (int)((v2c - V1S) / DVS + eps);
    int nptc = i2 - i1 + 3;
    if (clip && nptc > NPTS) nptc = NPTS + 4;
    *V1C = v1c;
    *V2C = v1c + DVS * (double)(nptc - 1);
    *I1 = i1;
    return nptc;
}

static int cn_window(double V1ABS, double V2ABS, const double *grid, double *V1C, double *V2C, double *DVC,
                     int *I1) {
    return cn_window_g(V1ABS, V2ABS, grid[0], grid[2], (int)grid[3], 0.01, true, V1C, V2C, DVC, I1);
}

// pre_xint (clblm_Spectrum.f90:917-944): the ABSRB points XINT may touch for a table on
// [v1ss, v2ss].
static void cn_pre_xint(double v1ss, double v2ss, double v1abs, int nptabs, int32_t *ist, int32_t *last) {
    const double dvabs = 1.;
    const int nbnd_v1c = (int)(2. + (v1ss - v1abs) / dvabs + 1.e-5);
    *ist = std::max(1, nbnd_v1c);
    const int nbnd_v2c = (int)(1. + (v2ss - v1abs) / dvabs + 1.e-5);
    *last = std::min(nptabs, nbnd_v2c);
}

// HERTDA with HERPRS (clblm_ODLAY_Continuum.f90:8164-8253): the Herzberg O2 continuum at
// V with its pressure dependence.
static double cn_herzberg(double V, double T, double P) {
    double HERZ = 0.;
    if (V > 36000.00) {
        double CORR = 0.;
        if (V <= 40000.) CORR = ((40000. - V) / 4000.) * 7.917E-27;
        const double YRATIO = V / 48811.0;
        const double ly = std::log(YRATIO);
        HERZ = 6.884E-04 * (YRATIO)*std::exp(-69.738 * (ly * ly)) - CORR;
    }
    const double PO = 1013., TO = 273.16;
    return HERZ * (1. + .83 * (P / PO) * (TO / T));
}

// One layer's bands.  CLBLM_ERR_ARG where ABSRB is longer than the reference's N_ABSRB
// (5050 points), the length of CONTNM's arrays.
static int cn_plan(const double *scale, const clblm_layer &lay, double V1, double V2, int jrad, CnJob &cj,
                   std::vector<CnTerm> &terms, std::vector<double> &coef, bool *has) {
    *has = false;
    const double xco2c = scale[0], xo3cn = scale[1], xo2cn = scale[2], xn2cn = scale[3], xrayl = scale[4];
    const double v1 = V1, v2 = V2;
    double V1ABS;
    int NPTABS;
    absrb_grid(V1, V2, &V1ABS, &NPTABS);
    if (NPTABS > 5050) return set_err(CLBLM_ERR_ARG, "continuum: ABSRB longer than 5050 points");
    const double V2ABS = V1ABS + (double)(NPTABS - 1);
    const double P0C = 1013., T0C = 296., XLOSMT = 2.68675E+19;
    const double pave = lay.pave, tave = lay.tave, wtot = lay.wtot;
    const double RHOAVE = (pave / P0C) * (T0C / tave);
    const double XKT = tave / RADCN2;
    const double amagat = (pave / P0C) * (273. / tave);
    const double x_vmr_h2o = lay.w[1] / wtot, x_vmr_o2 = lay.w[7] / wtot;
    const double x_vmr_n2 = 1. - x_vmr_h2o - x_vmr_o2;
    const double wn2 = x_vmr_n2 * wtot;
    cj.term0 = (int32_t)terms.size();
    cj.nterm = 0;
    auto begin_at = [&](double v1c, double v2c, double dvc, int nptc) {
        CnTerm t{};
        t.nptc = nptc;
        t.v1c = v1c;
        t.v2c = v2c;
        t.dvc = dvc;
        t.n1 = 1;
        t.n2 = NPTABS;
        t.coef = (int64_t)coef.size();
        return t;
    };
    auto begin = [&](const double *grid, double *V1C, double *DVC, int *I1) {
        double V2C;
        const int nptc = cn_window(V1ABS, V2ABS, grid, V1C, &V2C, DVC, I1);
        return begin_at(*V1C, V2C, *DVC, nptc);
    };
    auto finish = [&](CnTerm &t) {
        coef.push_back(0.);
        terms.push_back(t);
        cj.nterm++;
    };
    double V1C, DVC;
    int I1;
    if ((v2 > -20.0) && (v1 < 10000.) && xco2c > 0) {   // CO2 (:415-457)
        const double WCO2 = lay.w[2] * RHOAVE * 1.0E-20 * xco2c;
        CnTerm t = begin(CN_CO2_GRID, &V1C, &DVC, &I1);
        const int NPTS = (int)CN_CO2_GRID[3];
        const double trat = tave / CN_CO2_TEFF[0];
        for (int J = 1; J <= t.nptc; J++) {
            const int I = I1 + (J - 1);
            double f = 0.;
            if (I >= 1 && I <= NPTS) {
                double tcor = 1.;
                if (I >= 1196 && I <= 1220) tcor = std::pow(trat, CN_CO2_TDEP[I - 1196]);
                f = tcor * CN_CO2_S[I - 1];
            }
            const double VJ = V1C + DVC * (double)(J - 1);
            double CFAC = 1.;
            if (VJ >= 2000. && VJ <= 2998.) CFAC = CN_XFACCO2[(int)((VJ - 1998.) / 2. + 0.00001) - 1];
            f = CFAC * f;
            double c = f * WCO2;
            if (jrad == 1) c = c * radfn(VJ, XKT);
            coef.push_back(c);
        }
        finish(t);
    }
    if (v2 > 8920.0 && v1 <= 24665.0 && xo3cn > 0.) {   // O3 Chappuis and Wulf (:464-481)
        const double WO3 = lay.w[3] * 1.0E-20 * xo3cn;
        CnTerm t = begin(CN_O3C_GRID, &V1C, &DVC, &I1);
        const int NPTS = (int)CN_O3C_GRID[3];
        const double DT = tave - 273.15;
        for (int J = 1; J <= t.nptc; J++) {
            const int I = I1 + (J - 1);
            const double VJ = V1C + DVC * (double)(J - 1);
            double c0 = 0., c1 = 0., c2 = 0.;
            if (I >= 1 && I <= NPTS) {
                c0 = CN_O3C_X[I - 1] / VJ;
                c1 = CN_O3C_Y[I - 1] / VJ;
                c2 = CN_O3C_Z[I - 1] / VJ;
            }
            double c = (c0 + (c1 + c2 * DT) * DT) * WO3;
            if (jrad == 1) c = c * radfn(VJ, XKT);
            coef.push_back(c);
        }
        finish(t);
    }
    if (v2 > 27370. && v1 < 40800. && xo3cn > 0.) {   // O3 Hartley-Huggins (:483-525)
        const double WO3 = lay.w[3] * 1.E-20 * xo3cn;
        const double TC = tave - 273.15;
        CnTerm t = begin(CN_O3T_GRID, &V1C, &DVC, &I1);
        const int NPTS = (int)CN_O3T_GRID[3];
        double VJ = 0.;
        for (int J = 1; J <= t.nptc; J++) {
            const int I = I1 + (J - 1);
            const bool in = I >= 1 && I <= NPTS;
            VJ = V1C + DVC * (double)(J - 1);
            const double c0 = in ? CN_O3T0_S[I - 1] / VJ : 0.;
            const double ct1 = in ? CN_O3T1_S[I - 1] : 0., ct2 = in ? CN_O3T2_S[I - 1] : 0.;
            double c = c0 * WO3;
            if (jrad == 1) c = c * radfn(VJ, XKT);
            c = c * (1. + ct1 * TC + ct2 * TC * TC);
            coef.push_back(c);
        }
        // the points from 40800 cm-1 are saved before XINT and restored after it
        if ((VJ > 40815.) && (v2 > 40800)) t.n2 = std::min(t.n2, (int32_t)((40800. - V1ABS) / 1. + 1.001) - 1);
        finish(t);
    }
    if (v2 > 40800. && v1 < 54000. && xo3cn > 0.) {   // O3 Hartley-Huggins in the UV (:530-566)
        const double WO3 = lay.w[3] * xo3cn;
        CnTerm t = begin(CN_O3UV_GRID, &V1C, &DVC, &I1);
        const int NPTS = (int)CN_O3UV_GRID[3];
        for (int J = 1; J <= t.nptc; J++) {
            const int I = I1 + (J - 1);
            const double VJ = V1C + DVC * (double)(J - 1);
            const double c0 = (I >= 1 && I <= NPTS) ? CN_O3UV_S[I - 1] / VJ : 0.;
            double c = c0 * WO3;
            if (jrad == 1) c = c * radfn(VJ, XKT);
            coef.push_back(c);
        }
        // the points below 40800 cm-1 are saved before XINT and restored after it
        if (v1 < 40800) t.n1 = std::max(t.n1, (int32_t)((40800. - V1ABS) / 1. + 1.001));
        finish(t);
    }
    if ((v2 > 1340.0) && (v1 < 1850.) && xo2cn > 0.) {   // O2 fundamental (:587-628)
        const double tau_fac = xo2cn * lay.w[7] * 1.e-20 * amagat;
        CnTerm t = begin(CN_O2F_GRID, &V1C, &DVC, &I1);
        const int NPTS = (int)CN_O2F_GRID[3];
        const double xktfac = (1. / 296.) - (1. / tave);
        const double factor = (1.e+20 / 2.68675e+19);
        for (int J = 1; J <= t.nptc; J++) {
            const int I = I1 + (J - 1);
            const double VJ = V1C + DVC * (double)(J - 1);
            double c0 = 0.;
            if (I >= 1 && I <= NPTS) c0 = factor * CN_O2F_X[I - 1] * std::exp(CN_O2F_XT[I - 1] * xktfac) / VJ;
            double c = tau_fac * c0;
            if (jrad == 1) c = c * radfn(VJ, XKT);
            coef.push_back(c);
        }
        finish(t);
    }
    if ((v2 > 7536.0) && (v1 < 8500.) && xo2cn > 0.) {   // O2 1.27 um (:633-660)
        const double a_o2 = 1. / 0.446, a_n2 = 0.3 / 0.446, a_h2o = 1.;
        const double tau_fac =
            xo2cn * (lay.w[7] / XLOSMT) * amagat * (a_o2 * x_vmr_o2 + a_n2 * x_vmr_n2 + a_h2o * x_vmr_h2o);
        CnTerm t = begin(CN_O2I1_GRID, &V1C, &DVC, &I1);
        const int NPTS = (int)CN_O2I1_GRID[3];
        for (int J = 1; J <= t.nptc; J++) {
            const int I = I1 + (J - 1);
            const double VJ = V1C + DVC * (double)(J - 1);
            const double c0 = (I >= 1 && I <= NPTS) ? CN_O2I1_X[I - 1] / VJ : 0.;
            double c = tau_fac * c0;
            if (jrad == 1) c = c * radfn(VJ, XKT);
            coef.push_back(c);
        }
        finish(t);
    }
    if ((v2 > 9100.0) && (v1 < 11000.) && xo2cn > 0.) {   // O2 1.06 um (:670-690), O2INF2's own grid
        const double V1_osc = 9375., HW1 = 58.96, V2_osc = 9439., HW2 = 45.04;
        const double S1 = 1.166E-04, S2 = 3.086E-05;
        const double V1S = 9100., v2s = 11000., DVS = 2.;
        DVC = DVS;
        V1C = V1ABS - DVC;
        double V2C = V2ABS + DVC;
        if (V1C < V1S) V1C = V1S - 2. * DVS;
        if (V2C > v2s) V2C = v2s + 2. * DVS;
        const int nptc = (int)((V2C - V1C) / DVC + 3.01);
        V2C = V1C + DVC * (double)(nptc - 1);
        CnTerm t = begin_at(V1C, V2C, DVC, nptc);
        const double WO2 = xo2cn * (lay.w[7] * 1.e-20) * RHOAVE;
        const double ADJWO2 = (lay.w[7] / wtot) * (1. / 0.209) * WO2;
        for (int J = 1; J <= nptc; J++) {
            double c0 = 0.;
            const double VJ = V1C + DVC * (double)(J - 1);
            if ((VJ > V1S) && (VJ < v2s)) {
                const double DV1 = VJ - V1_osc, DV2 = VJ - V2_osc;
                const double DAMP1 = DV1 < 0.0 ? std::exp(DV1 / 176.1) : 1.0;
                const double DAMP2 = DV2 < 0.0 ? std::exp(DV2 / 176.1) : 1.0;
                const double q1 = DV1 / HW1, q2 = DV2 / HW2;
                const double O2INF =
                    0.31831 * (((S1 * DAMP1 / HW1) / (1. + q1 * q1)) + ((S2 * DAMP2 / HW2) / (1. + q2 * q2))) * 1.054;
                c0 = O2INF / VJ;
            }
            double c = c0 * ADJWO2;
            if (jrad == 1) c = c * radfn(VJ, XKT);
            coef.push_back(c);
        }
        finish(t);
    }
    if ((v2 > 12961.5) && (v1 < 13221.5) && xo2cn > 0.) {   // O2 A-band (:700-719)
        const double tau_fac = xo2cn * (lay.w[7] / XLOSMT) * amagat;
        CnTerm t = begin(CN_O2I3_GRID, &V1C, &DVC, &I1);
        const int NPTS = (int)CN_O2I3_GRID[3];
        for (int J = 1; J <= t.nptc; J++) {
            const int I = I1 + (J - 1);
            const double VJ = V1C + DVC * (double)(J - 1);
            const double c0 = (I >= 1 && I <= NPTS) ? CN_O2I3_X[I - 1] / VJ : 0.;
            double c = tau_fac * c0;
            if (jrad == 1) c = c * radfn(VJ, XKT);
            coef.push_back(c);
        }
        cn_pre_xint(CN_O2I3_GRID[0], CN_O2I3_GRID[1], V1ABS, NPTABS, &t.n1, &t.n2);
        finish(t);
    }
    if ((v2 > 15000.0) && (v1 < 29870.) && xo2cn > 0.) {   // O2 visible (:734-758)
        const double WO2 = lay.w[7] * 1.e-20 * ((pave / 1013.) * (273. / tave)) * xo2cn;
        const double CHIO2 = lay.w[7] / wtot;
        const double ADJWO2 = CHIO2 * WO2;
        const double tq = 55. * 273. / 296.;
        const double factor = 1. / ((XLOSMT * 1.e-20 * (tq * tq)) * 89.5);
        CnTerm t = begin(CN_O2V_GRID, &V1C, &DVC, &I1);
        const int NPTS = (int)CN_O2V_GRID[3];
        for (int J = 1; J <= t.nptc; J++) {
            const int I = I1 + (J - 1);
            const double VJ = V1C + DVC * (double)(J - 1);
            const double c0 = (I >= 1 && I <= NPTS) ? factor * CN_O2V_S[I - 1] / VJ : 0.;
            double c = c0 * ADJWO2;
            if (jrad == 1) c = c * radfn(VJ, XKT);
            coef.push_back(c);
        }
        finish(t);
    }
    if (v2 > 36000.0 && xo2cn > 0.) {   // O2 Herzberg (:762-779), no table length
        const double WO2 = lay.w[7] * 1.e-20 * xo2cn;
        double V2C;
        const int nptc = cn_window_g(V1ABS, V2ABS, 36000., 10., 0, 0.01, false, &V1C, &V2C, &DVC, &I1);
        CnTerm t = begin_at(V1C, V2C, DVC, nptc);
        for (int J = 1; J <= nptc; J++) {
            const int I = I1 + (J - 1);
            const double VJ = V1C + DVC * (double)(J - 1);
            const double c0 = I >= 1 ? cn_herzberg(VJ, tave, pave) / VJ : 0.;
            double c = c0 * WO2;
            if (jrad == 1) c = c * radfn(VJ, XKT);
            coef.push_back(c);
        }
        finish(t);
    }
    if (v2 > 56740.0 && xo2cn > 0.) {   // O2 far UV (:784-802)
        const double WO2 = lay.w[7] * 1.e-20 * xo2cn;
        double V2C;
        const int nptc = cn_window_g(V1ABS, V2ABS, CN_O2U_GRID[0], CN_O2U_GRID[2], (int)CN_O2U_GRID[3], 1.e-5, true,
                                     &V1C, &V2C, &DVC, &I1);
        CnTerm t = begin_at(V1C, V2C, DVC, nptc);
        const int NPTS = (int)CN_O2U_GRID[3];
        for (int J = 1; J <= nptc; J++) {
            const int I = I1 + (J - 1);
            const double VJ = V1C + DVC * (double)(J - 1);
            const double c0 = (I >= 1 && I <= NPTS) ? CN_O2U_S[I - 1] / VJ : 0.;
            double c = c0 * WO2;
            if (jrad == 1) c = c * radfn(VJ, XKT);
            coef.push_back(c);
        }
        cn_pre_xint(CN_O2U_GRID[0], CN_O2U_GRID[1], V1ABS, NPTABS, &t.n1, &t.n2);
        finish(t);
    }
    if ((v2 > -10.0) && (v1 < 350.) && xn2cn > 0.) {   // N2 pure rotation (:807-852)
        const double a_h2o = 1.;
        const double tau_fac = xn2cn * (wn2 / XLOSMT) * amagat;
        CnTerm t = begin(CN_N2R_GRID, &V1C, &DVC, &I1);
        const int NPTS = (int)CN_N2R_GRID[3];
        const double tfac = (tave - 296.) / (220. - 296.);
        for (int J = 1; J <= t.nptc; J++) {
            const int I = I1 + (J - 1);
            const double VJ = V1C + DVC * (double)(J - 1);
            double c0 = 0., c1 = 0.;
            if (I >= 1 && I <= NPTS) {
                c0 = CN_N2R_C296[I - 1] * std::pow(CN_N2R_C220[I - 1] / CN_N2R_C296[I - 1], tfac);
                const double sf_T = CN_N2R_SF296[I - 1] * std::pow(CN_N2R_SF220[I - 1] / CN_N2R_SF296[I - 1], tfac);
                c1 = (sf_T - 1.) * (0.79) / (0.21);
            }
            double c = tau_fac * c0 * (x_vmr_n2 + c1 * x_vmr_o2 + a_h2o * x_vmr_h2o);
            if (jrad == 1) c = c * radfn(VJ, XKT);
            coef.push_back(c);
        }
        finish(t);
    }
    if ((v2 > 2001.77) && (v1 < 2897.59) && xn2cn > 0.) {   // N2 fundamental (:856-909)
        const double tau_fac = xn2cn * (wn2 / XLOSMT) * amagat;
        CnTerm t = begin(CN_N2F_GRID, &V1C, &DVC, &I1);
        const int NPTS = (int)CN_N2F_GRID[3];
        const double T = tave;
        const double xtfac = ((1. / T) - (1. / 272.)) / ((1. / 228.) - (1. / 272.));
        const double xt_lin = (T - 272.) / (228. - 272.);
        const double a_o2 = 1.294 - 0.4545 * T / 296.;
        for (int J = 1; J <= t.nptc; J++) {
            const int I = I1 + (J - 1);
            const double VJ = V1C + DVC * (double)(J - 1);
            double c0 = 0., c1 = 0., c2 = 0.;
            if (I >= 1 && I <= NPTS) {
                const double x272 = CN_N2F_X272[I - 1], x228 = CN_N2F_X228[I - 1];
                if (x272 > 0. && x228 > 0.) c0 = x272 * std::pow(x228 / x272, xtfac);
                else c0 = x272 + (x228 - x272) * xt_lin;
                c0 = c0 / VJ;
                c1 = a_o2 * c0;
                c2 = (9. / 7.) * CN_N2F_AH2O[I - 1] * c0;
            }
            double c = tau_fac * (x_vmr_n2 * c0 + x_vmr_o2 * c1 + x_vmr_h2o * c2);
            if (jrad == 1) c = c * radfn(VJ, XKT);
            coef.push_back(c);
        }
        finish(t);
    }
    if ((v2 > 4340.0) && (v1 < 4910.) && xn2cn > 0.) {   // N2 first overtone (:946-987)
        const double a_o2 = 1., a_h2o = 1.;
        const double tau_fac = xn2cn * (wn2 / XLOSMT) * amagat * (x_vmr_n2 + a_o2 * x_vmr_o2 + a_h2o * x_vmr_h2o);
        CnTerm t = begin(CN_N2O_GRID, &V1C, &DVC, &I1);
        const int NPTS = (int)CN_N2O_GRID[3];
        for (int J = 1; J <= t.nptc; J++) {
            const int I = I1 + (J - 1);
            const double VJ = V1C + DVC * (double)(J - 1);
            const double c0 = (I >= 1 && I <= NPTS) ? CN_N2O_X[I - 1] / VJ : 0.;
            double c = tau_fac * c0;
            if (jrad == 1) c = c * radfn(VJ, XKT);
            coef.push_back(c);
        }
        finish(t);
    }
    cj.ray = (v2 >= 820. && xrayl > 0.) ? 1 : 0;   // Rayleigh (:970-1060)
    cj.ray_fac = wtot * (xrayl * 1.E-20 / (2.68675e-1 * 1.e5));
    cj.v1abs = V1ABS;
    cj.nptabs = NPTABS;
    cj.jrad = jrad;
    *has = cj.nterm > 0 || cj.ray;
    return 0;
}

// CONTNM's bands at ABSRB point I (1-based) of job blockIdx.y, added in order: each
// band by XINT where I lies in its range and in [n1, n2], then the Rayleigh extinction.
__global__ void k_contnm(const CnJob *jobs, const CnTerm *terms, const double *coef, double *absrb) {
    const CnJob J = jobs[blockIdx.y];
    const int I = 1 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (I > J.nptabs) return;
    double *out = absrb + J.abs_off + (I - 1);
    double a = *out;
    for (int t = 0; t < J.nterm; t++) {
        const CnTerm T = terms[J.term0 + t];
        int ilo = (int)((T.v1c + T.dvc - J.v1abs) / 1. + 1. + ONEMI);
        if (ilo < T.n1) ilo = T.n1;
        int ihi = (int)((T.v2c - T.dvc - J.v1abs) / 1. + ONEMI);
        if (ihi > T.n2) ihi = T.n2;
        if (I >= ilo && I <= ihi)
            a = a + xint_point(T.v1c, T.v2c, T.dvc, coef + T.coef - 1, J.v1abs, 1., I, T.n1, T.n2);
    }
    if (J.ray) {
        const double vrayleigh = J.v1abs + (double)(I - 1) * 1.;
        const double x = vrayleigh / 1.e4;
        double r = (x * x * x / (9.38076E2 - 10.8426 * (x * x))) * J.ray_fac;
        if (J.jrad == 1) r = r * x;
        else if (J.jrad == 0) r = r / 1.e4;
        a = a + r;
    }
    *out = a;
}

struct Batch {
    std::vector<Job> jobs;
    std::vector<Panel> panels;
    std::vector<R4Blk> blks;
    std::vector<int2> tiles;
    std::vector<double> absrb;
    std::vector<MtJob> mtjobs;   // MT_CKD water-vapour continuum added to absrb on the device
    std::vector<double> mtcoef;
    int64_t mt_maxpts = 0;
    std::vector<CnJob> cnjobs;   // CONTNM's other bands, added after the water vapour
    std::vector<CnTerm> cnterms;
    std::vector<double> cncoef;
    int64_t cn_maxpts = 0;
    std::vector<int> stop0;
    std::vector<int> regrid_lay;
    std::vector<int64_t> regrid_first;
    int64_t max_line = 0, max_f4 = 0, max_rec = 0, tot_line = 0, tot_f4 = 0, tot_rec = 0;
    int64_t tot_a = 0, tot_r4 = 0, tot_nat = 0, tot_out = 0, tot_f4rec = 0;
    int max_stride = 0;
    int f4k = 3;                 // F4 gather points per lane (tile width 64 f4k)
    bool any_jrad1 = false, any_regrid = false;
};

struct clblm_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    // The F4 wing chain (LINF4Q, SHRINQ, CNVF4Q) and pass 1 (CNVFNQ) read disjoint
    // inputs and write disjoint buffers until pass 2 combines them, so a batch forks
    // the F4 chain onto this stream (default priority) and joins it before pass 2: the
    // latency-bound SHRINQ walks then run beside the pass-1 convolution.
    hipStream_t f4stream = nullptr;
    int f4_chunk = 16;       // consecutive F4 tiles per XCD run (xcd_tile)
    int f4_far = 1;          // far-field evaluation in the FP32 gather: 1 on, 2 off, 3 on from a third of a tile
    int f4_stream_mode = 1;  // 1: the F4 chain on its own stream; 2: on the batch's stream
    // far-field nodes per tile width 64 F4K (F4K = F4KMIN..F4KMAX): Lagrange weights
    // [64 F4K][F4NF] and node positions (tile points from the tile's first point)
    double *far_w[F4KMAX + 1] = {};
    float *far_tau[F4KMAX + 1] = {};
    int f4k32 = 3, f4k64 = 2;   // F4 gather points per lane, FP32C / FP64
    int conv_u = 4;             // pass-1 R1 points per walk step
    int conv_split = 0;         // log2 of the most parts a pass-1 line walk is split into
    double *tab = nullptr;   // F1,F2,F3,FG,XVER consecutive, NFMX+1 each
    float *tabs = nullptr;
    double4 *tabg = nullptr;
    float4 *tabgs = nullptr;
    std::vector<LineListH *> lists;
    DevArena arena;
    DevArena tan_arena;   // scratch of the tangent batch (clblm_od_tangent)
    int *d_err = nullptr;
    unsigned long long *d_work = nullptr;   // [conv evaluations, F4 evaluations]
    double *rt_state = nullptr;             // per-layer recursion states for the reverse sweep
    size_t rt_state_cap = 0;
    double *grid_state = nullptr;           // merged and layer arrays of the merge on layer grids
    size_t grid_state_cap = 0;
    // Batches are asynchronous: each call takes the next slot of a ring, which holds
    // its stage events and its device error word (copied to pinned host memory at
    // the end of the call).  A slot is read back when it is reused, by clblm_sync,
    // or by the timing queries; a device-detected error is returned by the next
    // call on the context.
    // Stage events: 0 start, 1 after the line prep, 2 after pass 1 (main stream);
    // 7 start of the F4 chain, 5/6 around the F4 gather, 3 end of the F4 chain (F4
    // stream); 8 after the join, 4 end (main stream).  fork/join order the F4 stream
    // inside the batch.
    struct Slot {
        hipEvent_t ev[9]{};
        hipEvent_t fork = nullptr, join = nullptr;
        hipEvent_t done = nullptr;
        bool pending = false;
    };
    static constexpr int NSLOT = 4;
    Slot slot[NSLOT];
    int *h_err = nullptr;                   // pinned, NSLOT words
    int64_t ncalls = 0;
    hipStream_t last_stream = nullptr;
    // plan uploads go through two alternating pinned buffers
    char *pin[2] = {nullptr, nullptr};
    size_t pin_cap[2] = {0, 0};
    hipEvent_t pin_ev[2]{};
    int pin_next = 0;
    int deferred_rc = 0;
    std::string deferred_msg;
    // the sizing call and the run call of one batch share one plan
    std::vector<char> plan_key;
    Batch plan;
    int32_t plan_jr = 0;
    bool plan_valid = false;
    double ms_conv_k = 0, ms_f4_k = 0;
    double ms[4] = {0, 0, 0, 0};
    int64_t launches = 0;
    void *scan = nullptr;                   // instrument-scan state (FFT plans, buffers)
    MtckdTable mtckd;                       // MT_CKD water-vapour continuum coefficients
    void *mt_buf = nullptr;                 // clblm_mtckd_absrb's jobs and windows
    size_t mt_cap = 0;
    std::mutex mu;
};

// The status code and message for a nonzero device error word.
static int device_error(int herr, std::string &msg) {
    if (herr & 32) {
        msg = "a line's Lorentz width is not positive (its line shape is undefined)";
        return CLBLM_ERR_ARG;
    }
    if (herr & 16) {
        msg = "internal error: a SHRINQ chunk exceeds the LINF4Q buffer";
        return CLBLM_ERR_HIP;
    }
    msg = "F4 record capacity exceeded";
    return CLBLM_ERR_CAPACITY;
}

// Reads back a finished slot: stage times and the device error word.
static void finish_slot(clblm_ctx *C, int k, bool wait) {
    clblm_ctx::Slot &S = C->slot[k];
    if (!S.pending) return;
    if (wait) (void)hipEventSynchronize(S.done);
    else if (hipEventQuery(S.done) != hipSuccess) return;
    S.pending = false;
    float t01 = 0, t12 = 0, t23 = 0, t34 = 0, t56 = 0;
    (void)hipEventElapsedTime(&t01, S.ev[0], S.ev[1]);
    (void)hipEventElapsedTime(&t12, S.ev[1], S.ev[2]);
    (void)hipEventElapsedTime(&t23, S.ev[7], S.ev[3]);
    (void)hipEventElapsedTime(&t34, S.ev[8], S.ev[4]);
    (void)hipEventElapsedTime(&t56, S.ev[5], S.ev[6]);
    C->ms[0] += t01; C->ms[1] += t12; C->ms[2] += t23; C->ms[3] += t34;
    C->ms_conv_k += t12;
    C->ms_f4_k += t56;
    C->launches++;
    int herr = C->h_err[k];
    if (herr && !C->deferred_rc) C->deferred_rc = device_error(herr, C->deferred_msg);
}

// Reads back every finished slot in call order (all of them when wait is set).
static void drain(clblm_ctx *C, bool wait) {
    for (int64_t j = std::max<int64_t>(0, C->ncalls - clblm_ctx::NSLOT); j < C->ncalls; j++) {
        int k = (int)(j % clblm_ctx::NSLOT);
        finish_slot(C, k, wait);
        if (C->slot[k].pending) break;
    }
}

static int take_deferred(clblm_ctx *C) {
    if (!C->deferred_rc) return 0;
    int rc = set_err(C->deferred_rc, C->deferred_msg + " (reported by a batch enqueued earlier)");
    C->deferred_rc = 0;
    C->deferred_msg.clear();
    return rc;
}

static thread_local int g_tls_dummy = 0;

static int ctx_tables(clblm_ctx *c) {
    std::vector<double> f1, f2, f3, fg, xv;
    build_tables(f1, f2, f3, fg, xv);
    std::vector<double> all;
    for (auto *v : {&f1, &f2, &f3, &fg, &xv}) all.insert(all.end(), v->begin(), v->end());
    std::vector<float> alls(all.begin(), all.end());
    HIPCHK(hipMalloc(&c->tab, all.size() * sizeof(double)));
    HIPCHK(hipMalloc(&c->tabs, alls.size() * sizeof(float)));
    HIPCHK(hipMemcpy(c->tab, all.data(), all.size() * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->tabs, alls.data(), alls.size() * sizeof(float), hipMemcpyHostToDevice));
    {
        std::vector<double4> g(NFMX + 1);
        std::vector<float4> gs(NFMX + 1);
        for (int i = 0; i <= NFMX; i++) {
            g[i] = make_double4(f1[i], fg[i], xv[i], 0.);
            gs[i] = make_float4((float)f1[i], (float)fg[i], (float)xv[i], 0.f);
        }
        HIPCHK(hipMalloc(&c->tabg, g.size() * sizeof(double4)));
        HIPCHK(hipMalloc(&c->tabgs, gs.size() * sizeof(float4)));
        HIPCHK(hipMemcpy(c->tabg, g.data(), g.size() * sizeof(double4), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(c->tabgs, gs.data(), gs.size() * sizeof(float4), hipMemcpyHostToDevice));
    }
    HIPCHK(hipMemcpyToSymbol(HIP_SYMBOL(c_avrat), CLBLM_AVRAT, sizeof(CLBLM_AVRAT)));
    HIPCHK(hipMemcpyToSymbol(HIP_SYMBOL(c_cgauss), CLBLM_CGAUSS, sizeof(CLBLM_CGAUSS)));
    HIPCHK(hipMemcpyToSymbol(HIP_SYMBOL(c_cf1), CLBLM_CF1, sizeof(CLBLM_CF1)));
    HIPCHK(hipMemcpyToSymbol(HIP_SYMBOL(c_cf2), CLBLM_CF2, sizeof(CLBLM_CF2)));
    HIPCHK(hipMemcpyToSymbol(HIP_SYMBOL(c_cf3), CLBLM_CF3, sizeof(CLBLM_CF3)));
    HIPCHK(hipMemcpyToSymbol(HIP_SYMBOL(c_cer), CLBLM_CER, sizeof(CLBLM_CER)));
    HIPCHK(hipMemcpyToSymbol(HIP_SYMBOL(c_maxiso), CLBLM_MAXISO, sizeof(CLBLM_MAXISO)));
    {
        double sc[8];
        shape_constants(sc);
        HIPCHK(hipMemcpyToSymbol(HIP_SYMBOL(c_shape), sc, sizeof(sc)));
    }
    for (int fk = F4KMIN; fk <= F4KMAX; fk++) {
        // far-field nodes of the FP32 gather: Chebyshev points of [0, W-1] (rounded to
        // FP32, as the kernel uses them) and the Lagrange weights of every tile point
        const int W = 64 * fk;
        const double PI = 3.14159265358979323846;
        std::vector<float> tau(F4NF);
        std::vector<double> td(F4NF), w((size_t)W * F4NF);
        for (int i = 0; i < F4NF; i++) {
            tau[i] = (float)(0.5 * (W - 1) * (1. + std::cos((2 * i + 1) * PI / (2 * F4NF))));
            td[i] = (double)tau[i];
        }
        for (int t = 0; t < W; t++)
            for (int i = 0; i < F4NF; i++) {
                double v = 1.;
                for (int k = 0; k < F4NF; k++)
                    if (k != i) v *= ((double)t - td[k]) / (td[i] - td[k]);
                w[(size_t)t * F4NF + i] = v;
            }
        HIPCHK(hipMalloc(&c->far_w[fk], w.size() * sizeof(double)));
        HIPCHK(hipMalloc(&c->far_tau[fk], tau.size() * sizeof(float)));
        HIPCHK(hipMemcpy(c->far_w[fk], w.data(), w.size() * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(c->far_tau[fk], tau.data(), tau.size() * sizeof(float), hipMemcpyHostToDevice));
    }
    return 0;
}

static Tables ctx_tabview(clblm_ctx *c) {
    Tables t;
    const size_t s = NFMX + 1;
    t.f1 = c->tab; t.f2 = c->tab + s; t.f3 = c->tab + 2 * s; t.fg = c->tab + 3 * s; t.xver = c->tab + 4 * s;
    t.f1s = c->tabs; t.f2s = c->tabs + s; t.f3s = c->tabs + 2 * s; t.fgs = c->tabs + 3 * s; t.xvers = c->tabs + 4 * s;
    t.g1 = c->tabg;
    t.g1s = c->tabgs;
    return t;
}

// ------------------------------------------------------------------------------------
// Per-layer plan: ODLAY_nonStruct/OPDPTH/HIRACQ scalars, the panel schedule and the
// R4 block schedule (clblm_ODLAY.f90:129-611,757-967,1002-1562).  These depend on
// V1, V2 and DV only.
// ------------------------------------------------------------------------------------
struct Plan {
    Job job;
    std::vector<Panel> panels;
    std::vector<R4Blk> blks;
    std::vector<double> absrb;
    bool mt = false;             // MT_CKD water-vapour continuum: its job and windows
    MtJob mj{};
    std::vector<double> mtc;
    bool cn = false;             // CONTNM's other bands: the job, its bands and windows
    CnJob cj{};
    std::vector<CnTerm> cnt;
    std::vector<double> cnc;
};

static int plan_layer(const LineListH &L, const clblm_layer &lay, const clblm_od_params &par,
                      int32_t *jrad_state, Plan &pl, const MtckdTable *mt) {
    Job &J = pl.job;
    std::memset(&J, 0, sizeof J);
    if (!(par.v2 > par.v1)) return set_err(CLBLM_ERR_ARG, "v2 must exceed v1");
    if (!(par.dvnormal > 0.) || !(par.dvnarrow > 0.) || !(par.sample > 0.))
        return set_err(CLBLM_ERR_ARG, "dvnormal, dvnarrow and sample must be positive");
    if (!(lay.tave >= 70. && lay.tave <= 3000.)) return set_err(CLBLM_ERR_ARG, "TIPS: temperature out of range 70-3000 K");
    if (!(lay.pave > 0.) || !(lay.wtot > 0.)) return set_err(CLBLM_ERR_ARG, "pave and wtot must be positive");
    for (int m = NMOL + 1; m <= CLBLM_MAXMOL; m++)
        if (lay.sel[m]) return set_err(CLBLM_ERR_UNSUPPORTED, "only molecules 1-7 carry TIPS/mass data in this build");
    const double V1 = par.v1, V2 = par.v2, T = lay.tave, Pv = lay.pave;
    double DV = par.dvnormal;
    double ALFAV = par.sample * std::max(par.dvnormal, par.dvnarrow);
    if (!(std::fabs(par.dvnormal - par.dvnarrow) / par.dvnormal > 0.2)) DV = par.dvnarrow;
    double XKT = T / RADCN2;
    if ((V1 / XKT) < 5. && *jrad_state != -1) *jrad_state = 0;
    J.jrad = *jrad_state;
    J.v1 = V1; J.v2 = V2; J.dv = DV; J.tave = T;
    J.dvr2 = (0.008 / 0.002) * DV;
    J.dvr3 = (0.032 / 0.002) * DV;
    double ALFAV4 = 64. * par.sample * DV;
    J.dvr4 = ALFAV4 / par.sample;
    J.bound4 = 25.;
    double V1R4 = V1 - 2. * J.dvr4, V2R4 = V2 + 2. * J.dvr4;
    J.v1l4 = V1R4 - J.bound4 - J.dvr4;
    J.v2l4 = V2R4 + J.bound4 + 2 * J.dvr4;
    J.alfmax = 4. * ALFAV;
    J.nalfmax = (int)(J.alfmax / DV);
    int NBOUND = (int)((2. * 64.) * (double)J.nalfmax + 0.01);
    int MAX1 = NSHIFT + NLIM1 + NSHIFT + NBOUND / 2;
    int MAX2 = MAX1 / 4, MAX3 = MAX1 / 16;
    MAX1 = MAX1 + 4 * J.nalfmax;
    MAX2 = MAX2 + 16 * J.nalfmax / 4 + 1;
    MAX3 = MAX3 + 64 * J.nalfmax / 16 + 1;
    if (MAX1 > 2432 + 3600 || MAX2 < 616 || MAX3 < 160)
        return set_err(CLBLM_ERR_UNSUPPORTED, "DVnormal/DVnarrow ratio gives an unsupported panel size");
    J.max1 = MAX1; J.max2 = MAX2; J.max3 = MAX3; J.stride = MAX1 + MAX2 + MAX3;
    double BOUND = (double)NBOUND * DV / 2.;
    J.vft0 = V1 - (double)NSHIFT * DV;
    if (!(J.vft0 > 0.)) return set_err(CLBLM_ERR_UNSUPPORTED, "panel start <= 0 cm-1 (RSYM) is not supported");
    // LNCORQ constants (clblm_ODLAY.f90:1649-1667)
    double XKT0 = TEMP0 / RADCN2;
    J.xkt = XKT;
    J.betacr = (1. / XKT) - (1. / XKT0);
    J.tratio = T / TEMP0;
    J.rhorat = (Pv / P0) * (TEMP0 / T);
    J.pavp0 = Pv / P0;
    J.pavp2 = J.pavp0 * J.pavp0;
    // line-coupling brackets over TEMPLC = 200, 250, 296, 340 K.  LNCORQ
    // (clblm_ODLAY.f90:1660-1667): the first ILC = 1..3 with TAVE < TEMPLC(ILC+1)
    // (else 3), anchored at TEMPLC(ILC).  LINF4Q (clblm_ODLAY_LineF4.f90:860-866): the
    // first ILC = 1..4 with TAVE <= TEMPLC(ILC), moved into 2..4, anchored at TEMPLC(ILC).
    {
        const double TEMPLC[4] = {200.0, 250.0, 296.0, 340.0};
        int ilc = 1;
        for (int il = 1; il <= 3; il++) {
            ilc = il;
            if (T < TEMPLC[ilc]) break;
        }
        J.lcq_lo = ilc - 1;
        J.lcq_anchor = ilc - 1;
        J.lcq_rect = 1.0 / (TEMPLC[ilc] - TEMPLC[ilc - 1]);
        J.lcq_dif = T - TEMPLC[ilc - 1];
        int k4 = 1;
        while (k4 <= 4 && !(T <= TEMPLC[k4 - 1])) k4++;
        if (k4 == 1) k4 = 2;
        if (k4 == 5) k4 = 4;
        J.lc4_lo = k4 - 2;
        J.lc4_anchor = k4 - 1;
        J.lc4_rect = 1.0 / (TEMPLC[k4 - 1] - TEMPLC[k4 - 2]);
        J.lc4_dif = T - TEMPLC[k4 - 1];
    }
    // LINF4Q constants (clblm_ODLAY_LineF4.f90:945-951)
    J.beta4 = RADCN2 / T;
    J.betacr4 = J.beta4 - RADCN2 / TEMP0;
    // MOLEC (clblm_ODLAY_CommonSub.f90:201-264)
    double FAD = std::log(2.) * 2. * AVOGAD * BOLTZ / (CLIGHT * CLIGHT);
    for (int m = 1; m <= NMOL; m++) {
        J.sel[m] = lay.sel[m] ? 1 : 0;
        J.w[m] = lay.w[m];
        if (!J.sel[m]) continue;
        for (int iso = 1; iso <= CLBLM_MAXISO[m - 1]; iso++) {
            const double *row = CLBLM_TIPS_Q + (size_t)(CLBLM_TIPS_ROW[m - 1] + iso - 1) * CLBLM_NT;
            double q296 = tips_interp(296., row), qt = tips_interp(T, row);
            if (!(qt > 0.) || !(q296 > 0.)) return set_err(CLBLM_ERR_ARG, "TIPS: partition sum less than 0");
            J.scor[m][iso] = q296 / qt;
            J.rhoslf[m] = J.rhorat * lay.w[m] / lay.wtot;
            J.alfd1[m][iso] = std::sqrt(FAD * T / CLBLM_ISO_MASS[CLBLM_ISO_OFFSET[m - 1] + iso - 1]);
        }
    }
    // line ranges (RDLIN: first line >= VBOT to the first line > VTOP inclusive)
    double VBOT = V1 - BOUND, VTOP = V2 + BOUND;
    auto lb = [&](double x) { return (int64_t)(std::lower_bound(L.nu.begin(), L.nu.end(), x) - L.nu.begin()); };
    auto ub = [&](double x) { return (int64_t)(std::upper_bound(L.nu.begin(), L.nu.end(), x) - L.nu.begin()); };
    J.line_lo = lb(VBOT);
    int64_t hi = ub(VTOP);
    if (hi < L.n) hi++;
    J.line_n = std::max<int64_t>(0, hi - J.line_lo);
    J.f4_lo = lb(J.v1l4);
    int64_t fhi = std::min<int64_t>(L.n, ub(J.v2l4) + 4 * REC);
    J.f4_n = std::max<int64_t>(0, fhi - J.f4_lo);
    J.rec0 = J.f4_lo / REC;
    J.nrec = J.f4_n > 0 ? (J.f4_lo + J.f4_n - 1) / REC - J.rec0 + 1 : 0;
    // panels and R4 blocks (HIRACQ panel loop and LBLF4Q refresh)
    auto r4block = [&](double v1r4, double v2r4in) {
        R4Blk b{};
        b.v1r4 = v1r4;
        int np = (int)((v2r4in - v1r4) / J.dvr4 + ONEPL);
        if (np > LIMOUT4) np = LIMOUT4;
        b.nptr4 = np;
        b.v2r4 = v1r4 + J.dvr4 * (double)(np - 1);
        return b;
    };
    const double V2R4ST = V2R4;
    pl.blks.clear();
    pl.panels.clear();
    pl.blks.push_back(r4block(V1R4, V2R4));
    double VFT = J.vft0;
    int NHI = NLIM1 + NSHIFT - 1, NLO = NSHIFT + 1;
    int64_t out = 0;
    for (;;) {
        Panel P{};
        P.vft = VFT;
        P.istop = 0;
        if ((VFT + (double)(NHI - 1) * DV) > V2) {
            NHI = (int)((V2 - VFT) / DV + 1.);
            double V2P = VFT + (double)(NHI - 1) * DV;
            if (V2P < V2) NHI = NHI + 1;
            P.istop = 1;
        }
        P.nhi = NHI;
        P.blk = (int)pl.blks.size() - 1;
        P.out_off = out;
        out += NHI - NLO + 1;
        pl.panels.push_back(P);
        if (P.istop) break;
        VFT = VFT + (double)(NLIM1 - 1) * DV;
        double VF1 = VFT - 2. * J.dvr4;
        VF1 = V1 + (double)(int)std::floor((VF1 - V1) / J.dvr4) * J.dvr4;
        double VF2 = VFT + 2. * J.dvr4 + (double)(155 + 4) * J.dvr3;
        const R4Blk &cur = pl.blks.back();
        if (VF2 > cur.v2r4 && cur.v2r4 != V2R4ST) pl.blks.push_back(r4block(VF1, V2R4ST));
        if (pl.panels.size() > (size_t)1 << 26) return set_err(CLBLM_ERR_ARG, "spectral range too long");
    }
    J.npanel = (int)pl.panels.size();
    J.nblk = (int)pl.blks.size();
    J.npts = out;
    J.nnat = out;
    J.dvout = 0.;
    if (par.dvout > 0. && !(std::fabs(par.dvout - DV) < 2.220446049250313e-16)) {
        J.dvout = par.dvout;
        J.npts = (int64_t)std::ceil((V2 - V1) / par.dvout + 1.);
        if (V1 + (double)(J.npts - 1) * par.dvout > V1 + (double)(J.nnat - 1) * DV + DV)
            return set_err(CLBLM_ERR_ARG, "DVout grid extends beyond the native grid");
    }
    // continuum on the OPDPTH ABSRB grid (clblm_ODLAY.f90:881-896): the caller's
    // array, the MT_CKD water-vapour continuum, or both (the caller's first)
    const bool mtckd = (par.flags & CLBLM_OD_MTCKD) != 0;
    const bool cnband = (par.flags & CLBLM_OD_CONTNM) != 0;
    J.contnm = (par.contnm || mtckd || cnband) ? 1 : 0;
    J.species_broad = (par.flags & CLBLM_OD_SPECIES_BROAD) ? 1 : 0;
    if (J.contnm) {
        double v1abs;
        int npt;
        absrb_grid(V1, V2, &v1abs, &npt);
        J.v1abs = v1abs;
        J.nptabs = npt;
        if (par.contnm) {
            if (par.nptabs != npt || !par.continuum)
                return set_err(CLBLM_ERR_ARG, "continuum needs " + std::to_string(npt) + " points from " +
                                                  std::to_string(v1abs) + " cm-1 at 1 cm-1");
            pl.absrb.assign(par.continuum, par.continuum + npt);
        } else {
            pl.absrb.assign((size_t)npt, 0.);
        }
        if (mtckd) {
            if (!mt) return set_err(CLBLM_ERR_ARG, "MT_CKD continuum asked for but no table loaded");
            int rc = mtckd_plan(*mt, lay, V1, V2, J.jrad, pl.mj, pl.mtc, &pl.mt);
            if (rc) return rc;
        }
        if (cnband) {
            static const double ones[5] = {1., 1., 1., 1., 1.};
            int rc = cn_plan(mt ? mt->cn_scale : ones, lay, V1, V2, J.jrad, pl.cj, pl.cnt, pl.cnc, &pl.cn);
            if (rc) return rc;
        }
    }
    return 0;
}

static int build_batch(const LineListH &L, int nlay, const clblm_layer *layers, const clblm_od_params *pars,
                       int32_t *jrad_state, Batch &B, int precision, int f4k, const MtckdTable *mt) {
    B = Batch{};
    B.f4k = f4k;
    B.jobs.reserve(nlay);
    for (int l = 0; l < nlay; l++) {
        if (pars[l].precision != precision)
            return set_err(CLBLM_ERR_ARG, "all layers of a batch must use the same precision");
        Plan pl;
        int rc = plan_layer(L, layers[l], pars[l], jrad_state, pl, mt);
        if (rc) return rc;
        Job J = pl.job;
        J.panel_off = (int64_t)B.panels.size();
        J.blk_off = (int64_t)B.blks.size();
        J.out_off = B.tot_out;
        B.tot_out += J.npts;
        J.lp_off = B.tot_line;
        B.tot_line += J.line_n;
        J.f4_off = B.tot_f4;
        B.tot_f4 += J.f4_n;
        J.rec_off = B.tot_rec;
        B.tot_rec += J.nrec;
        J.absrb_off = (int64_t)B.absrb.size();
        B.absrb.insert(B.absrb.end(), pl.absrb.begin(), pl.absrb.end());
        if (pl.mt) {
            MtJob mj = pl.mj;
            mj.coef = (int64_t)B.mtcoef.size();
            mj.abs_off = J.absrb_off;
            B.mtcoef.insert(B.mtcoef.end(), pl.mtc.begin(), pl.mtc.end());
            B.mtjobs.push_back(mj);
            B.mt_maxpts = std::max<int64_t>(B.mt_maxpts, mj.ihi - mj.ilo + 1);
        }
        if (pl.cn) {
            CnJob cj = pl.cj;
            cj.abs_off = J.absrb_off;
            cj.term0 = (int32_t)B.cnterms.size();
            const int64_t c0 = (int64_t)B.cncoef.size();
            for (CnTerm t : pl.cnt) {
                t.coef += c0;
                B.cnterms.push_back(t);
            }
            B.cncoef.insert(B.cncoef.end(), pl.cnc.begin(), pl.cnc.end());
            B.cnjobs.push_back(cj);
            B.cn_maxpts = std::max<int64_t>(B.cn_maxpts, cj.nptabs);
        }
        J.nat_off = B.tot_nat;
        if (J.dvout > 0.) {
            B.tot_nat += J.nnat;
            B.any_regrid = true;
        }
        for (auto &P : pl.panels) {
            P.layer = l;
            P.blk += (int)J.blk_off;
            P.out_off += (J.dvout > 0. ? J.nat_off : J.out_off);
            P.a_off = B.tot_a;
            B.tot_a += J.stride;
            B.panels.push_back(P);
        }
        J.r4_off = B.tot_r4;
        // each shrunk line lies in the read range of at most 3 + 50/(2500 DVR4) blocks
        J.f4rec_cap = J.f4_n * (int64_t)(3 + std::ceil(50.0 / (LIMOUT4 * J.dvr4))) + 64;
        J.f4rec_off = B.tot_f4rec;
        B.tot_f4rec += J.f4rec_cap;
        for (auto &b : pl.blks) {
            b.layer = l;
            b.r4_off = B.tot_r4;
            B.tot_r4 += LIMOUT4 + 2;
            int bi = (int)B.blks.size();
            B.blks.push_back(b);
            const int tile = 64 * B.f4k;
            for (int t = 0; t < (LIMOUT4 + 2 + tile - 1) / tile; t++) B.tiles.push_back(make_int2(bi, t));
        }
        B.max_line = std::max(B.max_line, J.line_n);
        B.max_f4 = std::max(B.max_f4, J.f4_n);
        B.max_rec = std::max(B.max_rec, J.nrec);
        B.max_stride = std::max(B.max_stride, J.stride);
        B.any_jrad1 |= J.jrad == 1;
        B.stop0.push_back((int)J.f4_n);
        B.jobs.push_back(J);
    }
    return 0;
}

template <typename T>
static T *at(char *base, size_t off) { return reinterpret_cast<T *>(base + off); }

// A batch captured in a hipGraph (clblm_od_graph_create): it owns its arena, its
// pinned plan staging, its error word and its fork/join events and F4 stream, so
// replays do not touch the context's ring of batches.
struct clblm_od_graph {
    DevArena arena;
    char *pin = nullptr;
    int *d_err = nullptr, *h_err = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    hipStream_t s4 = nullptr, cap = nullptr;
    hipEvent_t done = nullptr;  // recorded after each launch
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool launched = false;
    LineListH *list = nullptr;  // the line list whose device pointers the graph holds
};
static int run_batch(clblm_ctx *C, const LineListH &L, Batch &B, double *d_out, hipStream_t st, int precision,
                     clblm_od_graph *G = nullptr);

static inline unsigned cdiv(int64_t a, int64_t b) { return (unsigned)((a + b - 1) / b); }

// Fills n bytes with the byte v (a kernel, so that it can be captured in a graph).
__global__ void k_fill_bytes(uint8_t *p, size_t n, uint8_t v) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

static int run_batch(clblm_ctx *C, const LineListH &L, Batch &B, double *d_out, hipStream_t st, int precision,
                     clblm_od_graph *G) {
    const int nlay = (int)B.jobs.size();
    if (nlay == 0) return 0;
    // scratch layout
    Carve cv;
    size_t o_jobs = cv.take<Job>(nlay);
    size_t o_pan = cv.take<Panel>(B.panels.size());
    size_t o_blk = cv.take<R4Blk>(B.blks.size());
    size_t o_tiles = cv.take<int2>(B.tiles.size());
    size_t o_abs = cv.take<double>(std::max<size_t>(1, B.absrb.size()));
    size_t o_mtj = cv.take<MtJob>(std::max<size_t>(1, B.mtjobs.size()));
    size_t o_mtc = cv.take<double>(std::max<size_t>(1, B.mtcoef.size()));
    size_t o_cnj = cv.take<CnJob>(std::max<size_t>(1, B.cnjobs.size()));
    size_t o_cnt = cv.take<CnTerm>(std::max<size_t>(1, B.cnterms.size()));
    size_t o_cnc = cv.take<double>(std::max<size_t>(1, B.cncoef.size()));
    size_t o_lp_d = cv.take<double>(5 * (size_t)B.tot_line);
    size_t o_lp_i = cv.take<int>(3 * (size_t)B.tot_line);
    size_t o_A = cv.take<double>((size_t)B.tot_a);
    size_t o_f4_b = cv.take<int8_t>((size_t)B.tot_f4);
    size_t o_sh_d = cv.take<double>(8 * (size_t)B.tot_f4);
    size_t o_sh_b = cv.take<int8_t>((size_t)B.tot_f4);
    size_t o_lay = cv.take<int>(4 * (size_t)nlay);
    size_t o_racc = cv.take<int>((size_t)B.tot_rec);
    const int64_t hcap = (B.max_line + SCAN_C - 1) / SCAN_C, fcap = (B.max_f4 + SCAN_C - 1) / SCAN_C;
    size_t o_htot = cv.take<int>((size_t)std::max<int64_t>(1, hcap * nlay));
    size_t o_ftot = cv.take<double>((size_t)std::max<int64_t>(1, 2 * fcap * nlay));
    size_t o_G = cv.take<int>((size_t)B.tot_rec * LCAP);
    size_t o_cd = cv.take<double>(5 * (size_t)B.tot_f4);
    size_t o_cf = cv.take<uint8_t>((size_t)B.tot_f4);
    size_t o_cr = cv.take<int>((size_t)B.tot_rec);
    size_t o_cn = cv.take<int>((size_t)nlay);
    size_t o_ch = cv.take<int4>((size_t)B.tot_rec);
    size_t o_be = cv.take<int64_t>((size_t)B.tot_rec);
    size_t o_R4 = cv.take<double>((size_t)B.tot_r4);
    size_t o_xjj = cv.take<double>((size_t)nlay * XJJ_N);
    size_t o_rec = precision == CLBLM_FP64 ? cv.take<F4Rec64>((size_t)B.tot_f4rec) : cv.take<F4Rec32>((size_t)B.tot_f4rec);
    size_t o_nat = cv.take<double>(std::max<int64_t>(1, B.tot_nat));
    size_t o_rgf = cv.take<int64_t>(nlay);
    // re-grid work units: points for the n/(n+1) ratios, pieces of 2400-point blocks
    // otherwise; rg_first[l] is layer l's first unit
    std::vector<int64_t> rg_first(nlay, 0);
    int64_t rg_units = 0;
    if (B.any_regrid) {
        for (int l = 0; l < nlay; l++) {
            rg_first[l] = rg_units;
            const Job &Jl = B.jobs[l];
            if (Jl.dvout > 0.) rg_units += clblm_interp::units(Jl.dv, Jl.dvout, Jl.npts);
        }
    }
    // host-built tables, staged in pinned memory: the header region [0, o_lp_d) as it
    // lies in the arena, then the per-layer stop indices and the re-grid units
    const size_t h_stop = (o_lp_d + 255) & ~(size_t)255;
    const size_t h_rgf = (h_stop + sizeof(int) * nlay + 255) & ~(size_t)255;
    const size_t h_size = h_rgf + sizeof(int64_t) * nlay;
    // where this batch lives: the context's arena, staging and slot ring, or (under
    // graph capture) the graph's own, with no stage events
    const int k = (int)(C->ncalls % clblm_ctx::NSLOT);
    clblm_ctx::Slot &S = C->slot[k];
    char *base = nullptr, *hp = nullptr;
    int ps = 0;
    int *d_err = nullptr;
    hipStream_t s4 = st;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    if (!G) {
        // stream order with the previous batch (the arena is shared), then the slot
        finish_slot(C, k, true);
        if (C->last_stream && C->last_stream != st) {
            const int kp = (int)((C->ncalls + clblm_ctx::NSLOT - 1) % clblm_ctx::NSLOT);
            HIPCHK(hipStreamWaitEvent(st, C->slot[kp].done, 0));
        }
        if (cv.off > C->arena.cap) {  // growing frees the old arena: let every batch finish first
            drain(C, true);
            HIPCHK(hipDeviceSynchronize());
        }
        if (C->arena.ensure(cv.off)) return set_err(CLBLM_ERR_HIP, "device scratch allocation failed");
        base = C->arena.p;
        ps = C->pin_next;
        C->pin_next ^= 1;
        HIPCHK(hipEventSynchronize(C->pin_ev[ps]));
        if (C->pin_cap[ps] < h_size) {
            if (C->pin[ps]) HIPCHK(hipHostFree(C->pin[ps]));
            C->pin[ps] = nullptr;
            C->pin_cap[ps] = 0;
            HIPCHK(hipHostMalloc((void **)&C->pin[ps], h_size + h_size / 4, hipHostMallocDefault));
            C->pin_cap[ps] = h_size + h_size / 4;
        }
        hp = C->pin[ps];
        d_err = C->d_err + k;
        s4 = C->f4_stream_mode == 1 ? C->f4stream : st;
        ev_fork = S.fork;
        ev_join = S.join;
    } else {
        if (G->arena.ensure(cv.off)) return set_err(CLBLM_ERR_HIP, "graph scratch allocation failed");
        base = G->arena.p;
        HIPCHK(hipHostMalloc((void **)&G->pin, h_size, hipHostMallocDefault));
        hp = G->pin;
        d_err = G->d_err;
        s4 = C->f4_stream_mode == 1 ? G->s4 : st;
        ev_fork = G->fork;
        ev_join = G->join;
        // everything from here to the end of the batch is recorded into the graph;
        // the F4 stream joins the capture through the fork event
        HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
    }
    auto stage_event = [&](int i, hipStream_t s) -> hipError_t {
        return G ? hipSuccess : hipEventRecord(S.ev[i], s);
    };
    std::memcpy(hp + o_jobs, B.jobs.data(), sizeof(Job) * nlay);
    std::memcpy(hp + o_pan, B.panels.data(), sizeof(Panel) * B.panels.size());
    std::memcpy(hp + o_blk, B.blks.data(), sizeof(R4Blk) * B.blks.size());
    std::memcpy(hp + o_tiles, B.tiles.data(), sizeof(int2) * B.tiles.size());
    if (!B.absrb.empty()) std::memcpy(hp + o_abs, B.absrb.data(), sizeof(double) * B.absrb.size());
    if (!B.mtjobs.empty()) {
        std::memcpy(hp + o_mtj, B.mtjobs.data(), sizeof(MtJob) * B.mtjobs.size());
        std::memcpy(hp + o_mtc, B.mtcoef.data(), sizeof(double) * B.mtcoef.size());
    }
    if (!B.cnjobs.empty()) {
        std::memcpy(hp + o_cnj, B.cnjobs.data(), sizeof(CnJob) * B.cnjobs.size());
        if (!B.cnterms.empty()) std::memcpy(hp + o_cnt, B.cnterms.data(), sizeof(CnTerm) * B.cnterms.size());
        if (!B.cncoef.empty()) std::memcpy(hp + o_cnc, B.cncoef.data(), sizeof(double) * B.cncoef.size());
    }
    std::memcpy(hp + h_stop, B.stop0.data(), sizeof(int) * nlay);
    std::memcpy(hp + h_rgf, rg_first.data(), sizeof(int64_t) * nlay);
    Job *d_jobs = at<Job>(base, o_jobs);
    Panel *d_pan = at<Panel>(base, o_pan);
    R4Blk *d_blk = at<R4Blk>(base, o_blk);
    int2 *d_tiles = at<int2>(base, o_tiles);
    double *d_abs = at<double>(base, o_abs);
    HIPCHK(hipMemcpyAsync(base, hp, o_lp_d, hipMemcpyHostToDevice, st));
    (void)d_jobs; (void)d_pan; (void)d_blk; (void)d_tiles; (void)d_abs;
    if (!B.mtjobs.empty())   // the water-vapour continuum joins ABSRB before pass 2 reads it
        hipLaunchKernelGGL(k_mtckd, dim3(cdiv(B.mt_maxpts, 256), (unsigned)B.mtjobs.size()), dim3(256), 0, st,
                           at<MtJob>(base, o_mtj), at<double>(base, o_mtc), d_abs);
    if (!B.cnjobs.empty())   // then CONTNM's other bands, in the reference's order
        hipLaunchKernelGGL(k_contnm, dim3(cdiv(B.cn_maxpts, 256), (unsigned)B.cnjobs.size()), dim3(256), 0, st,
                           at<CnJob>(base, o_cnj), at<CnTerm>(base, o_cnt), at<double>(base, o_cnc), d_abs);
    LinePrep lp;
    double *lpd = at<double>(base, o_lp_d);
    int *lpi = at<int>(base, o_lp_i);
    lp.vnu = lpd; lp.sp = lpd + B.tot_line; lp.recalf = lpd + 2 * B.tot_line; lp.zeta = lpd + 3 * B.tot_line;
    lp.sppsp = lpd + 4 * B.tot_line;
    lp.iz = lpi; lp.pmin = lpi + B.tot_line; lp.home = lpi + 2 * B.tot_line;
    double *d_A = at<double>(base, o_A);
    F4Prep f;
    f.acc = at<int8_t>(base, o_f4_b);
    int *lay = at<int>(base, o_lay);
    f.stop = lay; f.nshr = lay + nlay; f.nchunk = lay + 2 * nlay; f.nblk9 = lay + 3 * nlay;
    f.rec_acc = at<int>(base, o_racc);
    f.G = at<int>(base, o_G);
    f.chunks = at<int4>(base, o_ch);
    f.blk9_end = at<int64_t>(base, o_be);
    Comp cmp;
    {
        double *cd = at<double>(base, o_cd);
        cmp.v = cd; cmp.s = cd + B.tot_f4; cmp.al = cd + 2 * B.tot_f4; cmp.ad = cd + 3 * B.tot_f4;
        cmp.sp2 = cd + 4 * B.tot_f4;
        cmp.fl = at<uint8_t>(base, o_cf);
        cmp.recpos = at<int>(base, o_cr);
        cmp.nacc = at<int>(base, o_cn);
    }
    Shrunk sh;
    double *shd = at<double>(base, o_sh_d);
    sh.vnu = shd; sh.sp = shd + B.tot_f4; sh.alfa = shd + 2 * B.tot_f4; sh.epp = shd + 3 * B.tot_f4;
    sh.spp = shd + 4 * B.tot_f4; sh.pmax = shd + 5 * B.tot_f4; sh.smin = shd + 6 * B.tot_f4;
    sh.bvmax = shd + 7 * B.tot_f4;
    sh.co2 = at<int8_t>(base, o_sh_b);
    double *d_R4 = at<double>(base, o_R4);
    double *d_nat = at<double>(base, o_nat);
    HIPCHK(hipMemcpyAsync(f.stop, hp + h_stop, sizeof(int) * nlay, hipMemcpyHostToDevice, st));
    {
        const size_t nb = sizeof(int4) * std::max<int64_t>(1, B.tot_rec);
        hipLaunchKernelGGL(k_fill_bytes, dim3(cdiv((int64_t)nb, 256)), dim3(256), 0, st, (uint8_t *)f.chunks, nb,
                           (uint8_t)0xff);
        hipLaunchKernelGGL(k_fill_bytes, dim3(1), dim3(64), 0, st, (uint8_t *)d_err, sizeof(int), (uint8_t)0);
    }
    Tables T = ctx_tabview(C);
    const int64_t npan = (int64_t)B.panels.size();
    const int64_t nblk = (int64_t)B.blks.size();

    // fork: the F4 chain runs on the F4 stream from here to the join before pass 2
    bool brd = false;   // species broadening anywhere in the batch, with the fields attached
    if (L.dev.brd_flg)
        for (const Job &Jb : B.jobs) brd = brd || Jb.species_broad != 0;
    HIPCHK(hipEventRecord(ev_fork, st));
    HIPCHK(hipStreamWaitEvent(s4, ev_fork, 0));
    HIPCHK(stage_event(7, s4));
    if (B.max_f4 > 0) {
        hipLaunchKernelGGL(brd ? k_f4acc<true> : k_f4acc<false>, dim3((unsigned)B.max_rec, nlay), dim3(256), 0, s4, d_jobs, L.dev, f, d_err);
        hipLaunchKernelGGL(k_recpos, dim3(nlay), dim3(256), 0, s4, d_jobs, f, cmp);
        hipLaunchKernelGGL(brd ? k_f4prep<true> : k_f4prep<false>, dim3((unsigned)B.max_rec, nlay), dim3(256), 0, s4, d_jobs, L.dev, f, cmp, d_err);
        hipLaunchKernelGGL(k_chunkG, dim3(cdiv(B.max_rec, 64), nlay), dim3(64), 0, s4, d_jobs, cmp, f.G);
        hipLaunchKernelGGL(k_sim, dim3(nlay), dim3(256), 0, s4, d_jobs, f, cmp, (const int *)f.G);
        hipLaunchKernelGGL(k_emit<false>, dim3((unsigned)std::max<int64_t>(1, B.tot_rec)), dim3(64), 0, s4, d_jobs, f,
                           cmp, sh, d_err, TanF4{});
        double *ftot = at<double>(base, o_ftot);
        if (fcap > 0) {
            hipLaunchKernelGGL(k_shr_local, dim3((unsigned)fcap, nlay, 2), dim3(SCAN_T), 0, s4, d_jobs, f, sh, ftot,
                               fcap);
            hipLaunchKernelGGL(k_shr_fix, dim3((unsigned)fcap, nlay, 2), dim3(SCAN_T), 0, s4, d_jobs, f, sh, ftot,
                               fcap);
        }
        hipLaunchKernelGGL(k_shr_blk9, dim3(nlay), dim3(256), 0, s4, d_jobs, f, sh);
    } else {
        hipLaunchKernelGGL(k_fill_bytes, dim3(cdiv((int64_t)(sizeof(int) * 3 * nlay), 256)), dim3(256), 0, s4,
                           (uint8_t *)f.nshr, sizeof(int) * 3 * (size_t)nlay, (uint8_t)0);
    }
    double *d_xjj = at<double>(base, o_xjj);
    hipLaunchKernelGGL(k_xjj, dim3(cdiv(nlay, 64)), dim3(64), 0, s4, d_jobs, nlay, d_xjj);
    hipLaunchKernelGGL(k_r4range, dim3(cdiv(nblk, 256)), dim3(256), 0, s4, d_jobs, d_blk, nblk, f, sh);
    hipLaunchKernelGGL(k_f4recoff, dim3(cdiv(nlay, 64)), dim3(64), 0, s4, d_jobs, nlay, d_blk, d_err);
    {
        unsigned gx = std::max<unsigned>(1, cdiv(B.max_f4 * 2, 256 * 8));
        if (precision == CLBLM_FP64)
            hipLaunchKernelGGL(k_f4setup<F4Rec64>, dim3(gx, (unsigned)nblk), dim3(256), 0, s4, d_jobs, d_blk, nblk,
                               sh, at<F4Rec64>(base, o_rec), d_xjj);
        else
            hipLaunchKernelGGL(k_f4setup<F4Rec32>, dim3(gx, (unsigned)nblk), dim3(256), 0, s4, d_jobs, d_blk, nblk,
                               sh, at<F4Rec32>(base, o_rec), d_xjj);
    }
    HIPCHK(stage_event(5, s4));
    const int64_t ntile = (int64_t)B.tiles.size();
    const int xchunk = C->f4_chunk > 0 ? C->f4_chunk : 1;
    const unsigned f4grid = (unsigned)(8 * (int64_t)xchunk * ((ntile + 8 * xchunk - 1) / (8 * xchunk)));
    if (ntile > 0) {
        const int fk = B.f4k;
        const float *tau = C->f4_far != 2 ? C->far_tau[fk] : nullptr;
        // far records lie at least half a tile (setting 1) or a third of a tile (3)
        // outside it
        const int dfar = C->f4_far == 3 ? (64 * fk) / 3 : 32 * fk;
#define F4_LAUNCH(K)                                                                                              \
    if (precision == CLBLM_FP64)                                                                                  \
        hipLaunchKernelGGL(k_f4conv64<K>, dim3(f4grid), dim3(64), 0, s4, d_jobs, d_blk, d_tiles, sh,              \
                           at<F4Rec64>(base, o_rec), d_R4, C->d_work, ntile, xchunk, d_xjj);                      \
    else                                                                                                          \
        hipLaunchKernelGGL(k_f4conv32<K>, dim3(f4grid), dim3(64), 0, s4, d_jobs, d_blk, d_tiles, sh,              \
                           at<F4Rec32>(base, o_rec), d_R4, C->d_work, ntile, xchunk, C->far_w[fk], tau, dfar);
        switch (fk) {
        case 2: F4_LAUNCH(2) break;
        case 4: F4_LAUNCH(4) break;
        default: F4_LAUNCH(3) break;
        }
#undef F4_LAUNCH
    }
    HIPCHK(stage_event(6, s4));
    if (B.any_jrad1 && nblk > 0) hipLaunchKernelGGL(k_r4rad, dim3((unsigned)nblk), dim3(256), 0, s4, d_jobs, d_blk, d_R4);
    HIPCHK(stage_event(3, s4));
    HIPCHK(hipEventRecord(ev_join, s4));
    HIPCHK(stage_event(0, st));
    if (B.max_line > 0) {
        hipLaunchKernelGGL(brd ? k_prep<true> : k_prep<false>, dim3(cdiv(B.max_line, 256), nlay), dim3(256), 0, st, d_jobs, d_pan, L.dev, lp, d_err);
        int *htot = at<int>(base, o_htot);
        hipLaunchKernelGGL(k_home_local, dim3((unsigned)hcap, nlay), dim3(SCAN_T), 0, st, d_jobs, lp, htot, hcap);
        hipLaunchKernelGGL(k_home_fix, dim3((unsigned)hcap, nlay), dim3(SCAN_T), 0, st, d_jobs, lp, htot, hcap);
    }
    hipLaunchKernelGGL(k_ranges, dim3(cdiv(npan, 256)), dim3(256), 0, st, d_jobs, d_pan, npan, lp);
    HIPCHK(stage_event(1, st));
    {
        const dim3 g((unsigned)npan), b(256);
        const size_t lds = (size_t)B.max_stride * sizeof(double);
#define CONV_LAUNCH(U)                                                                                          \
    {                                                                                                           \
        auto k64 = L.has_cpl ? k_conv64<U, true> : k_conv64<U, false>;                                         \
        auto k32 = L.has_cpl ? k_conv32<U, true> : k_conv32<U, false>;                                         \
        if (precision == CLBLM_FP64)                                                                            \
            hipLaunchKernelGGL(k64, g, b, lds, st, d_jobs, d_pan, lp, T, d_A, C->d_work, C->conv_split);      \
        else                                                                                                    \
            hipLaunchKernelGGL(k32, g, b, lds, st, d_jobs, d_pan, lp, T, d_A, C->d_work, C->conv_split);       \
    }
        switch (C->conv_u) {
        case 2: CONV_LAUNCH(2) break;
        case 8: CONV_LAUNCH(8) break;
        default: CONV_LAUNCH(4) break;
        }
#undef CONV_LAUNCH
    }
    HIPCHK(stage_event(2, st));
    HIPCHK(hipStreamWaitEvent(st, ev_join, 0));
    HIPCHK(stage_event(8, st));
    hipLaunchKernelGGL(k_pass2<double>, dim3((unsigned)npan), dim3(256), 0, st, d_jobs, d_pan, d_blk, d_A, d_R4,
                       d_abs, d_out, d_nat);
    if (B.any_regrid) {
        int64_t *d_rgf = at<int64_t>(base, o_rgf);
        HIPCHK(hipMemcpyAsync(d_rgf, hp + h_rgf, sizeof(int64_t) * nlay, hipMemcpyHostToDevice, st));
        if (rg_units > 0)
            hipLaunchKernelGGL(k_regrid, dim3((unsigned)cdiv(rg_units, 64)), dim3(64), 0, st, d_jobs, nlay, d_rgf,
                               d_nat, d_out, rg_units, (int64_t)0, (int64_t)0);
    }
    HIPCHK(stage_event(4, st));
    if (G) {
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(G->h_err, d_err, sizeof(int), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamEndCapture(st, &G->graph));
        return 0;
    }
    HIPCHK(hipEventRecord(C->pin_ev[ps], st));
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(C->h_err + k, d_err, sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHK(hipEventRecord(S.done, st));
    S.pending = true;
    C->ncalls++;
    C->last_stream = st;
    return 0;
}

// ====================================================================================
// C-ABI
// ====================================================================================
extern "C" int clblm_rt_down_impl(int nlay, int64_t npts, double v1, double dv, const double *d_od,
                                  const double *d_tlay, const double *d_tlev, int lin, const double *d_sol,
                                  double *d_rad, double *d_tx, void *stream);
extern "C" int clblm_rt_up_impl(int nlay, int64_t npts, double v1, double dv, const double *d_od,
                                const double *d_tlay, const double *d_tlev, double emis, double tskin, int lin,
                                double sec_dn, double *d_rad, double *d_dr_dod, double *d_dr_dt,
                                double *d_dr_dtskin, double *d_dr_demis, double *d_state, void *stream);
extern "C" void clblm_level2layer_impl(int nlay, int nmol, const double *p_lev, const double *t_lev,
                                       const double *t_lay, const double *q_lev, double *dt_dtup, double *dt_dtlo,
                                       double *dw_dqup, double *dw_dqlo);
extern "C" int clblm_rt_jac_impl(int nlay, int64_t npts, double v1, double dv, const double *d_od,
                                 const double *d_tlay, const double *d_tlev, const double *d_dt_dtup,
                                 const double *d_dt_dtlo, double emis, double tskin, int lin, double sec_dn,
                                 double *d_rad, double *d_dr_dtau, double *d_dr_dtlev, double *d_dr_dtskin,
                                 double *d_dr_demis, double *d_state, void *stream);
extern "C" int clblm_rt_scratch_init(int device);
extern "C" int clblm_interp_spectrum_impl(const double *d_in, int64_t nin, double v1, double dvin, double *d_out,
                                          int64_t nout, double dvout, void *stream);
extern "C" int clblm_finest_grid(int nspec, const double *dv, const int64_t *nlim);
extern "C" void clblm_rt_grids_final(int nlay, double v1, const double *dv, const int64_t *nlim, double surf_dv,
                                     int64_t surf_nlim, double *dv_out, int64_t *nlim_out, int64_t *max_n);
extern "C" int clblm_rt_up_grids_impl(int nlay, double v1, const double *dv, const int64_t *nlim,
                                      const int64_t *od_off, const double *d_od, const double *t_lay,
                                      const double *t_lev, double emis, double tskin, double surf_dv,
                                      int64_t surf_nlim, int lin, double sec, double *scratch, int64_t max_n,
                                      double *d_rad, void *stream);
extern "C" int clblm_rt_jac_scratch(int device, int nlay, int64_t npts, double v1, double dv, const double *d_od,
                                    const double *d_tlay, const double *d_tlev, const double *d_dt_dtup,
                                    const double *d_dt_dtlo, double emis, double tskin, int lin, double sec_dn,
                                    double *d_rad, double *d_dr_dtau, double *d_dr_dtlev, double *d_dr_dtskin,
                                    double *d_dr_demis, void *stream);
extern "C" int clblm_level_jac_impl(int nlay, int nmol, int64_t npts, const double *d_od_var,
                                    const double *d_dr_dtau, const double *d_dr_dtlev, const double *d_wtot,
                                    const double *d_tlay, double dt, double dw_frac, const double *d_dt_dtup,
                                    const double *d_dt_dtlo, const double *d_dw_dqup, const double *d_dw_dqlo,
                                    double *d_jac_t, double *d_jac_q, void *stream);
extern "C" void *clblm_scan_state_new();
extern "C" void clblm_scan_state_free(void *s);
extern "C" const char *clblm_scan_error();
extern "C" int clblm_fft_scan_impl(void *state, const double *d_in, int64_t n, double v1, double dv, double v1out,
                                   double v2out, double dvout, int funct_id, double hwhm, double fill,
                                   double pre_box_hw, int deconv_pre_box, int precision, double *d_out, int64_t *nout,
                                   double *v1res, double *dvres, void *stream);

extern "C" int clblm_rt_jac_paths_impl(int L, int64_t n, double v1, double dv, const double *d_od, const double *d_av,
                                       const double *d_ad, const double *d_as, const double *d_tlay,
                                       const double *d_tlev, double emis, double tskin, int solar, const double *d_sol,
                                       double rs, int obs, int uplook, int lin, const double *d_dt_dtup,
                                       const double *d_dt_dtlo, double *d_rad, double *d_dtau, double *d_dtlev,
                                       double *d_dts, double *d_de, double *d_drs, double *d_state, void *stream);
extern "C" int clblm_rt_flux_impl(int nlay, int64_t npts, double v1, double dv, const double *d_od,
                                  const double *d_tlay, const double *d_tlev, int lin, double emis, double tskin,
                                  int nang, int outinrat, int nout, int lambertian, double *scratch, int64_t blk_pts,
                                  double *fdn, double *fup, void *stream);
extern "C" int clblm_rt_mode1_impl(int nlay, double v1, double v2, const double *dv, const int64_t *nlim,
                                   const int64_t *od_off, const double *d_od_up, const double *d_od_dn,
                                   const double *d_od_sun, const double *fac_up, const double *fac_dn,
                                   const double *fac_sun, const double *t_lay, const double *t_lev, int obs_level,
                                   int lin, int thermal, int solar, double emis, double tskin, double refl,
                                   const double *d_sol, int64_t sol_n, double sol_dv, double *scratch, int64_t max_n,
                                   double *d_rad, double *d_tx, double *dv_out, int64_t *n_out, void *stream);
extern "C" int clblm_filter_impl(void *state, const double *d_in, int64_t n, double v1, double dv, int nchan,
                                 const double *v1f, const double *dvf, const int32_t *nptf, const int64_t *xf_off,
                                 const double *xf, double *d_out, void *stream);
extern "C" int clblm_scanfn_impl(void *state, const double *d_in, int64_t n, int64_t nrows, int64_t ld_in,
                                 double v1, double dv, double v1out, double v2out, double dvout, int funct_id,
                                 double hwhm, double *d_out, int64_t ld_out, int64_t *nout, double *v1res,
                                 double *dvres, void *stream);

extern "C" {

int clblm_version(void) { return 1; }
const char *clblm_last_error(void) { return g_err.c_str(); }

int clblm_ctx_create(int device, clblm_ctx **out) {
    if (!out) return set_err(CLBLM_ERR_ARG, "out is NULL");
    int n = 0;
    HIPCHK(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return set_err(CLBLM_ERR_ARG, "no such device");
    HIPCHK(hipSetDevice(device));
    // the device address of clblm_rt_jac_noalloc's scratch array on this device
    if (clblm_rt_scratch_init(device)) return set_err(CLBLM_ERR_HIP, "hipGetSymbolAddress of the scratch array failed");
    clblm_ctx *c = new clblm_ctx;
    c->device = device;
    HIPCHK(hipStreamCreate(&c->stream));
    HIPCHK(hipMalloc(&c->d_err, clblm_ctx::NSLOT * sizeof(int)));
    HIPCHK(hipHostMalloc((void **)&c->h_err, clblm_ctx::NSLOT * sizeof(int), hipHostMallocDefault));
    HIPCHK(hipMalloc(&c->d_work, 2 * sizeof(unsigned long long)));
    HIPCHK(hipMemset(c->d_work, 0, 2 * sizeof(unsigned long long)));
    HIPCHK(hipStreamCreateWithFlags(&c->f4stream, hipStreamNonBlocking));
    for (auto &S : c->slot) {
        for (auto &e : S.ev) HIPCHK(hipEventCreate(&e));
        HIPCHK(hipEventCreateWithFlags(&S.fork, hipEventDisableTiming));
        HIPCHK(hipEventCreateWithFlags(&S.join, hipEventDisableTiming));
        HIPCHK(hipEventCreate(&S.done));
    }
    for (auto &e : c->pin_ev) HIPCHK(hipEventCreate(&e));
    int rc = ctx_tables(c);
    if (rc) { delete c; return rc; }
    *out = c;
    return 0;
}

int clblm_ctx_destroy(clblm_ctx *c) {
    if (!c) return 0;
    (void)hipSetDevice(c->device);
    (void)hipDeviceSynchronize();
    for (auto *l : c->lists)
        if (l) {
            (void)hipFree(l->mem);
            if (l->brd_mem) (void)hipFree(l->brd_mem);
            delete l;
        }
    if (c->arena.p) (void)hipFree(c->arena.p);
    if (c->tan_arena.p) (void)hipFree(c->tan_arena.p);
    (void)hipFree(c->mt_buf);
    (void)hipFree(c->tab);
    (void)hipFree(c->tabs);
    (void)hipFree(c->tabg);
    (void)hipFree(c->tabgs);
    for (int fk = F4KMIN; fk <= F4KMAX; fk++) {
        (void)hipFree(c->far_w[fk]);
        (void)hipFree(c->far_tau[fk]);
    }
    (void)hipFree(c->d_err);
    (void)hipFree(c->d_work);
    if (c->rt_state) (void)hipFree(c->rt_state);
    if (c->grid_state) (void)hipFree(c->grid_state);
    for (auto &S : c->slot) {
        for (auto &e : S.ev) (void)hipEventDestroy(e);
        (void)hipEventDestroy(S.fork);
        (void)hipEventDestroy(S.join);
        (void)hipEventDestroy(S.done);
    }
    for (int i = 0; i < 2; i++) {
        (void)hipEventDestroy(c->pin_ev[i]);
        if (c->pin[i]) (void)hipHostFree(c->pin[i]);
    }
    (void)hipHostFree(c->h_err);
    if (c->scan) clblm_scan_state_free(c->scan);
    (void)hipStreamDestroy(c->f4stream);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return 0;
}

int clblm_linelist_upload(clblm_ctx *c, int64_t n, const double *nu, const float *s, const float *a_air,
                          const float *a_self, const float *e_lo, const float *tmpalf, const float *shift,
                          const int32_t *mol, const int32_t *iflg, int *handle) {
    if (!c || !handle) return set_err(CLBLM_ERR_ARG, "ctx or handle is NULL");
    if (n <= 1 || !nu || !s || !a_air || !a_self || !e_lo || !tmpalf || !shift || !mol || !iflg)
        return set_err(CLBLM_ERR_ARG, "line list needs at least two lines and every field");
    // Records with IFLG < 0 are not lines (the coupling coefficients of the line before
    // them, or records RDLIN skips); their VNU field is kept as given (nu_rec) and the
    // searches see the centre of the line before them, so that the centres are sorted.
    std::vector<double> cen(nu, nu + n);
    int64_t first = n;   // the first line
    double prev = 0.;
    for (int64_t i = 0; i < n; i++) {
        if (iflg[i] == 1 || iflg[i] == 3) {
            if (i + 1 >= n || (i + 1) / REC != i / REC || iflg[i + 1] >= 0)
                return set_err(CLBLM_ERR_ARG, "a line with IFLG 1 or 3 needs its coefficient record (IFLG < 0) "
                                              "right after it in the same 250-line record");
        }
        if (iflg[i] < 0) {
            cen[i] = prev;
            continue;
        }
        if (first < n && !(nu[i] >= prev)) return set_err(CLBLM_ERR_ARG, "line centres must be sorted ascending");
        if (first == n) {
            first = i;
            for (int64_t k = 0; k < i; k++) cen[k] = nu[i];   // leading non-line records
        }
        prev = nu[i];
    }
    HIPCHK(hipSetDevice(c->device));
    LineListH *L = new LineListH;
    L->n = n;
    L->nu = cen;
    for (int64_t i = 0; i < n && !L->has_cpl; i++) L->has_cpl = iflg[i] == 1;
    size_t bytes = (size_t)n * (8 + 8 + 6 * 4 + 2 * 4) + 9 * 256;
    HIPCHK(hipMalloc(&L->mem, bytes));
    Carve cv;
    char *b = (char *)L->mem;
    size_t onu = cv.take<double>(n), os = cv.take<float>(n), oa = cv.take<float>(n), osf = cv.take<float>(n);
    size_t oe = cv.take<float>(n), ot = cv.take<float>(n), osh = cv.take<float>(n), om = cv.take<int32_t>(n);
    size_t oi = cv.take<int32_t>(n), onr = cv.take<double>(n);
    HIPCHK(hipMemcpy(b + onu, cen.data(), 8 * n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(b + onr, nu, 8 * n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(b + os, s, 4 * n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(b + oa, a_air, 4 * n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(b + osf, a_self, 4 * n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(b + oe, e_lo, 4 * n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(b + ot, tmpalf, 4 * n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(b + osh, shift, 4 * n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(b + om, mol, 4 * n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(b + oi, iflg, 4 * n, hipMemcpyHostToDevice));
    L->dev.n = n;
    L->dev.nu = at<double>(b, onu);
    L->dev.s = at<float>(b, os);
    L->dev.a_air = at<float>(b, oa);
    L->dev.a_self = at<float>(b, osf);
    L->dev.e_lo = at<float>(b, oe);
    L->dev.tmpalf = at<float>(b, ot);
    L->dev.shift = at<float>(b, osh);
    L->dev.mol = at<int32_t>(b, om);
    L->dev.iflg = at<int32_t>(b, oi);
    L->dev.nu_rec = at<double>(b, onr);
    std::lock_guard<std::mutex> g(c->mu);
    c->lists.push_back(L);
    *handle = (int)c->lists.size() - 1;
    return 0;
}

int clblm_linelist_broadening(clblm_ctx *c, int handle, const int32_t *flg, const float *hw, const float *tmp,
                              const float *shft) {
    if (!c || !flg || !hw || !tmp || !shft) return set_err(CLBLM_ERR_ARG, "null argument");
    if (handle < 0 || handle >= (int)c->lists.size() || !c->lists[handle])
        return set_err(CLBLM_ERR_ARG, "bad line list handle");
    HIPCHK(hipSetDevice(c->device));
    std::lock_guard<std::mutex> g(c->mu);
    LineListH *L = c->lists[handle];
    if (L->graphs > 0)
        return set_err(CLBLM_ERR_ARG, "line list is read by a captured graph: destroy the graph first");
    drain(c, true);   // batches in flight may read the fields being replaced
    const size_t n7 = (size_t)L->n * 7;
    if (L->brd_mem) {
        HIPCHK(hipDeviceSynchronize());
        (void)hipFree(L->brd_mem);
        L->brd_mem = nullptr;
    }
    Carve cv;
    const size_t of = cv.take<int32_t>(n7), oh = cv.take<float>(n7), ot = cv.take<float>(n7), os = cv.take<float>(n7);
    HIPCHK(hipMalloc(&L->brd_mem, cv.off));
    char *b = (char *)L->brd_mem;
    HIPCHK(hipMemcpy(b + of, flg, 4 * n7, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(b + oh, hw, 4 * n7, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(b + ot, tmp, 4 * n7, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(b + os, shft, 4 * n7, hipMemcpyHostToDevice));
    L->dev.brd_flg = at<int32_t>(b, of);
    L->dev.brd_hw = at<float>(b, oh);
    L->dev.brd_tmp = at<float>(b, ot);
    L->dev.brd_shft = at<float>(b, os);
    return 0;
}

int clblm_linelist_free(clblm_ctx *c, int handle) {
    if (!c) return set_err(CLBLM_ERR_ARG, "null argument");
    std::lock_guard<std::mutex> g(c->mu);
    if (handle < 0 || handle >= (int)c->lists.size() || !c->lists[handle])
        return set_err(CLBLM_ERR_ARG, "bad line list handle");
    if (c->lists[handle]->graphs > 0)
        return set_err(CLBLM_ERR_ARG, "line list is read by a captured graph: destroy the graph first");
    HIPCHK(hipSetDevice(c->device));
    drain(c, true);   // batches in flight may read the list
    (void)hipFree(c->lists[handle]->mem);
    if (c->lists[handle]->brd_mem) (void)hipFree(c->lists[handle]->brd_mem);
    delete c->lists[handle];
    c->lists[handle] = nullptr;
    return 0;
}

int clblm_od_batch(clblm_ctx *c, int handle, int nlay, const clblm_layer *layers, const clblm_od_params *pars,
                   int32_t *jrad_state, int64_t *offsets, double *d_od, void *stream) {
    if (!c || !layers || !pars || !jrad_state || !offsets || nlay <= 0)
        return set_err(CLBLM_ERR_ARG, "null argument or nlay <= 0");
    if (handle < 0 || handle >= (int)c->lists.size() || !c->lists[handle])
        return set_err(CLBLM_ERR_ARG, "bad line list handle");
    int prec = pars[0].precision;
    if (prec != CLBLM_FP64 && prec != CLBLM_FP32C) return set_err(CLBLM_ERR_ARG, "unknown precision");
    HIPCHK(hipSetDevice(c->device));
    std::lock_guard<std::mutex> g(c->mu);
    // the inputs that define the plan: a sizing call followed by the run call with
    // the same inputs plans once
    std::vector<char> key;
    auto put = [&key](const void *p, size_t n) {
        const char *b = (const char *)p;
        key.insert(key.end(), b, b + n);
    };
    put(&handle, sizeof handle);
    put(&nlay, sizeof nlay);
    put(jrad_state, sizeof *jrad_state);
    put(layers, sizeof(clblm_layer) * nlay);
    put(pars, sizeof(clblm_od_params) * nlay);
    put(&c->mtckd.gen, sizeof c->mtckd.gen);
    for (int l = 0; l < nlay; l++)
        if (pars[l].contnm && pars[l].continuum && pars[l].nptabs > 0)
            put(pars[l].continuum, sizeof(double) * (size_t)pars[l].nptabs);
    int32_t jr = *jrad_state;
    Batch B;
    if (c->plan_valid && c->plan_key == key) {
        B = std::move(c->plan);
        jr = c->plan_jr;
        c->plan_valid = false;
    } else {
        c->plan_valid = false;
        int rc = build_batch(*c->lists[handle], nlay, layers, pars, &jr, B, prec,
                             prec == CLBLM_FP64 ? c->f4k64 : c->f4k32, &c->mtckd);
        if (rc) return rc;
    }
    for (int l = 0; l < nlay; l++) offsets[l] = B.jobs[l].out_off;
    offsets[nlay] = B.tot_out;
    if (!d_od) {
        c->plan = std::move(B);
        c->plan_key = std::move(key);
        c->plan_jr = jr;
        c->plan_valid = true;
        return 0;
    }
    drain(c, false);
    int rc = take_deferred(c);
    if (rc) return rc;
    hipStream_t st = stream ? (hipStream_t)stream : c->stream;
    rc = run_batch(c, *c->lists[handle], B, d_od, st, prec);
    if (rc) return rc;
    *jrad_state = jr;
    return 0;
}

static void graph_free(clblm_od_graph *g) {
    if (!g) return;
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    if (g->arena.p) (void)hipFree(g->arena.p);
    if (g->d_err) (void)hipFree(g->d_err);
    if (g->h_err) (void)hipHostFree(g->h_err);
    if (g->pin) (void)hipHostFree(g->pin);
    if (g->fork) (void)hipEventDestroy(g->fork);
    if (g->join) (void)hipEventDestroy(g->join);
    if (g->done) (void)hipEventDestroy(g->done);
    if (g->s4) (void)hipStreamDestroy(g->s4);
    if (g->cap) (void)hipStreamDestroy(g->cap);
    delete g;
}

int clblm_od_graph_create(clblm_ctx *c, int handle, int nlay, const clblm_layer *layers,
                          const clblm_od_params *pars, int32_t *jrad_state, int64_t *offsets, double *d_od,
                          clblm_od_graph **graph) {
    if (!c || !layers || !pars || !jrad_state || !offsets || !d_od || !graph || nlay <= 0)
        return set_err(CLBLM_ERR_ARG, "null argument or nlay <= 0");
    *graph = nullptr;
    if (handle < 0 || handle >= (int)c->lists.size() || !c->lists[handle])
        return set_err(CLBLM_ERR_ARG, "bad line list handle");
    const int prec = pars[0].precision;
    if (prec != CLBLM_FP64 && prec != CLBLM_FP32C) return set_err(CLBLM_ERR_ARG, "unknown precision");
    HIPCHK(hipSetDevice(c->device));
    std::lock_guard<std::mutex> lk(c->mu);
    int32_t jr = *jrad_state;
    Batch B;
    int rc = build_batch(*c->lists[handle], nlay, layers, pars, &jr, B, prec, prec == CLBLM_FP64 ? c->f4k64 : c->f4k32,
                         &c->mtckd);
    if (rc) return rc;
    clblm_od_graph *g = new clblm_od_graph;
    auto fail = [&](int r) {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        if (g->cap && hipStreamIsCapturing(g->cap, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone) {
            hipGraph_t part = nullptr;
            (void)hipStreamEndCapture(g->cap, &part);
            if (part) (void)hipGraphDestroy(part);
        }
        graph_free(g);
        return r;
    };
    if (hipMalloc(&g->d_err, sizeof(int)) != hipSuccess ||
        hipHostMalloc((void **)&g->h_err, sizeof(int), hipHostMallocDefault) != hipSuccess ||
        hipEventCreateWithFlags(&g->fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&g->join, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&g->done, hipEventDisableTiming) != hipSuccess ||
        hipStreamCreateWithFlags(&g->s4, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&g->cap, hipStreamNonBlocking) != hipSuccess)
        return fail(set_err(CLBLM_ERR_HIP, "graph resources could not be allocated"));
    *g->h_err = 0;
    rc = run_batch(c, *c->lists[handle], B, d_od, g->cap, prec, g);
    if (rc) return fail(rc);
    if (hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0) != hipSuccess)
        return fail(set_err(CLBLM_ERR_HIP, "hipGraphInstantiate failed"));
    for (int l = 0; l < nlay; l++) offsets[l] = B.jobs[l].out_off;
    offsets[nlay] = B.tot_out;
    *jrad_state = jr;
    g->list = c->lists[handle];
    g->list->graphs++;
    *graph = g;
    return 0;
}

int clblm_od_graph_launch(clblm_ctx *c, clblm_od_graph *g, void *stream) {
    if (!c || !g) return set_err(CLBLM_ERR_ARG, "ctx or graph is NULL");
    HIPCHK(hipSetDevice(c->device));
    std::lock_guard<std::mutex> lk(c->mu);
    hipStream_t st = stream ? (hipStream_t)stream : c->stream;
    // a launch after an earlier one on another stream runs after it (one arena)
    if (g->launched) HIPCHK(hipStreamWaitEvent(st, g->done, 0));
    HIPCHK(hipGraphLaunch(g->exec, st));
    HIPCHK(hipEventRecord(g->done, st));
    g->launched = true;
    return 0;
}

int clblm_od_graph_check(clblm_ctx *c, clblm_od_graph *g) {
    if (!c || !g) return set_err(CLBLM_ERR_ARG, "ctx or graph is NULL");
    HIPCHK(hipSetDevice(c->device));
    if (!g->launched) return 0;
    HIPCHK(hipEventSynchronize(g->done));
    if (*g->h_err) {
        std::string msg;
        const int rc = device_error(*g->h_err, msg);
        return set_err(rc, msg);
    }
    return 0;
}

int clblm_od_graph_destroy(clblm_ctx *c, clblm_od_graph *g) {
    if (!c || !g) return 0;
    (void)hipSetDevice(c->device);
    if (g->launched) (void)hipEventSynchronize(g->done);
    {
        std::lock_guard<std::mutex> lk(c->mu);
        if (g->list) g->list->graphs--;
    }
    graph_free(g);
    return 0;
}

int clblm_tuning(clblm_ctx *c, int key, int value, int *old_value);
int clblm_sync(clblm_ctx *c);

// Coordinate search over launch-shape settings: each key in turn takes each of its
// candidate values, the batch is run (one untimed run, then reps timed runs on the
// context's stream) and the fastest value is kept.
int clblm_autotune(clblm_ctx *c, int handle, int nlay, const clblm_layer *layers, const clblm_od_params *pars,
                   const int *keys, int nkeys, int reps, double *ms_best) {
    if (!c || !layers || !pars || nlay <= 0 || nkeys < 0 || (nkeys > 0 && !keys))
        return set_err(CLBLM_ERR_ARG, "null argument or bad sizes");
    if (reps < 1) reps = 1;
    const int prec = pars[0].precision;
    std::vector<int> klist;
    if (nkeys > 0) klist.assign(keys, keys + nkeys);
    else klist = {prec == CLBLM_FP64 ? CLBLM_TUNE_F4_POINTS_FP64 : CLBLM_TUNE_F4_POINTS_FP32C, CLBLM_TUNE_CONV_UNROLL,
                  CLBLM_TUNE_CONV_SPLIT, CLBLM_TUNE_F4_XCD_CHUNK, CLBLM_TUNE_F4_STREAM};
    auto candidates = [](int key) -> std::vector<int> {
        switch (key) {
        case CLBLM_TUNE_F4_POINTS_FP32C:
        case CLBLM_TUNE_F4_POINTS_FP64: return {2, 3, 4};
        case CLBLM_TUNE_CONV_UNROLL: return {2, 4, 8};
        case CLBLM_TUNE_CONV_SPLIT: return {1, 2, 4};
        case CLBLM_TUNE_F4_XCD_CHUNK: return {4, 16, 64};
        case CLBLM_TUNE_F4_STREAM: return {1, 2};
        case CLBLM_TUNE_F4_FAR_FIELD: return {1, 2, 3};
        default: return {};
        }
    };
    for (int k : klist)
        if (candidates(k).empty()) return set_err(CLBLM_ERR_ARG, "unknown tuning key");
    HIPCHK(hipSetDevice(c->device));
    std::vector<int64_t> off(nlay + 1);
    int32_t jr0 = 1, jr = jr0;
    int rc = clblm_od_batch(c, handle, nlay, layers, pars, &jr, off.data(), nullptr, nullptr);
    if (rc) return rc;
    double *d_od = nullptr;
    if (hipMalloc(&d_od, sizeof(double) * (size_t)std::max<int64_t>(1, off[nlay])) != hipSuccess)
        return set_err(CLBLM_ERR_HIP, "autotune output allocation failed");
    hipEvent_t e0 = nullptr, e1 = nullptr;
    auto cleanup = [&]() {
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        (void)hipFree(d_od);
    };
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        cleanup();
        return set_err(CLBLM_ERR_HIP, "autotune events");
    }
    auto run = [&](double *ms) -> int {
        jr = jr0;
        int r = clblm_od_batch(c, handle, nlay, layers, pars, &jr, off.data(), d_od, c->stream);
        if (r) return r;
        if (hipEventRecord(e0, c->stream) != hipSuccess) return set_err(CLBLM_ERR_HIP, "autotune event record");
        for (int i = 0; i < reps; i++) {
            jr = jr0;
            r = clblm_od_batch(c, handle, nlay, layers, pars, &jr, off.data(), d_od, c->stream);
            if (r) return r;
        }
        if (hipEventRecord(e1, c->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess)
            return set_err(CLBLM_ERR_HIP, "autotune event wait");
        r = clblm_sync(c);
        if (r) return r;
        float t = 0.f;
        (void)hipEventElapsedTime(&t, e0, e1);
        *ms = t / reps;
        return 0;
    };
    double best = 0.;
    rc = run(&best);
    for (int key : klist) {
        if (rc) break;
        int cur = 0;
        rc = clblm_tuning(c, key, 0, &cur);
        if (rc) break;
        int bestv = cur;
        for (int v : candidates(key)) {
            if (v == cur) continue;
            rc = clblm_tuning(c, key, v, nullptr);
            if (rc) break;
            double t = 0.;
            rc = run(&t);
            if (rc) break;
            if (t < best) { best = t; bestv = v; }
        }
        if (rc) break;
        rc = clblm_tuning(c, key, bestv, nullptr);
    }
    cleanup();
    if (rc) return rc;
    if (ms_best) *ms_best = best;
    return 0;
}

int clblm_tuning(clblm_ctx *c, int key, int value, int *old_value) {
    if (!c) return set_err(CLBLM_ERR_ARG, "ctx is NULL");
    int *slot = nullptr;
    int split_parts = 1 << c->conv_split;   // the key counts parts; the context keeps log2
    int lo = 0, hi = 0;
    switch (key) {
    case CLBLM_TUNE_F4_XCD_CHUNK: slot = &c->f4_chunk; lo = 1; hi = 4096; break;
    case CLBLM_TUNE_F4_FAR_FIELD: slot = &c->f4_far; lo = 1; hi = 3; break;
    case CLBLM_TUNE_F4_STREAM: slot = &c->f4_stream_mode; lo = 1; hi = 2; break;
    case CLBLM_TUNE_F4_POINTS_FP32C: slot = &c->f4k32; lo = F4KMIN; hi = F4KMAX; break;
    case CLBLM_TUNE_F4_POINTS_FP64: slot = &c->f4k64; lo = F4KMIN; hi = F4KMAX; break;
    case CLBLM_TUNE_CONV_UNROLL: slot = &c->conv_u; lo = 2; hi = 8; break;
    case CLBLM_TUNE_CONV_SPLIT: slot = &split_parts; lo = 1; hi = 4; break;
    default: return set_err(CLBLM_ERR_ARG, "unknown tuning key");
    }
    if (old_value) *old_value = *slot;
    if (value == 0) return 0;   // query only
    if (value < lo || value > hi || (key == CLBLM_TUNE_CONV_UNROLL && value != 2 && value != 4 && value != 8) ||
        (key == CLBLM_TUNE_CONV_SPLIT && value == 3))
        return set_err(CLBLM_ERR_ARG, "tuning value out of range");
    std::lock_guard<std::mutex> g(c->mu);
    if (key == CLBLM_TUNE_CONV_SPLIT) c->conv_split = value == 4 ? 2 : value - 1;
    else *slot = value;
    c->plan_valid = false;
    return 0;
}

int clblm_sync(clblm_ctx *c) {
    if (!c) return set_err(CLBLM_ERR_ARG, "ctx is NULL");
    HIPCHK(hipSetDevice(c->device));
    std::lock_guard<std::mutex> g(c->mu);
    drain(c, true);
    return take_deferred(c);
}

int clblm_odlay(clblm_ctx *c, int handle, const clblm_layer *layer, const clblm_od_params *par,
                int32_t *jrad_state, int64_t *nsamp, double *dv_used, double *od_out) {
    if (!c || !layer || !par || !jrad_state || !nsamp) return set_err(CLBLM_ERR_ARG, "null argument");
    int64_t off[2];
    int32_t jr = *jrad_state;
    int rc = clblm_od_batch(c, handle, 1, layer, par, &jr, off, nullptr, nullptr);
    if (rc) return rc;
    *nsamp = off[1];
    if (dv_used) {
        double DV = par->dvnormal;
        if (!(std::fabs(par->dvnormal - par->dvnarrow) / par->dvnormal > 0.2)) DV = par->dvnarrow;
        *dv_used = (par->dvout > 0. && !(std::fabs(par->dvout - DV) < 2.220446049250313e-16)) ? par->dvout : DV;
    }
    if (!od_out) return 0;
    double *d = nullptr;
    HIPCHK(hipMalloc(&d, sizeof(double) * off[1]));
    rc = clblm_od_batch(c, handle, 1, layer, par, jrad_state, off, d, nullptr);
    if (rc == 0) rc = clblm_sync(c);
    if (rc == 0) {
        hipError_t e = hipMemcpy(od_out, d, sizeof(double) * off[1], hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = set_err(CLBLM_ERR_HIP, hipGetErrorString(e));
    }
    (void)hipFree(d);
    return rc;
}

int clblm_mtckd_load(clblm_ctx *c, int n, const double *wavenumber, const double *self_absco_ref,
                     const double *for_absco_ref, const double *self_texp, double ref_temp, double ref_press,
                     double xself, double xfrgn) {
    if (!c || !wavenumber || !self_absco_ref || !for_absco_ref || !self_texp || n < 4)
        return set_err(CLBLM_ERR_ARG, "null argument or fewer than 4 coefficients");
    if (!(ref_temp > 0.) || !(ref_press > 0.)) return set_err(CLBLM_ERR_ARG, "reference T and p must be positive");
    for (int i = 1; i < n; i++)
        if (!(wavenumber[i] > wavenumber[i - 1])) return set_err(CLBLM_ERR_ARG, "wavenumbers must increase");
    std::lock_guard<std::mutex> g(c->mu);
    MtckdTable &T = c->mtckd;
    T.wvn.assign(wavenumber, wavenumber + n);
    T.self.assign(self_absco_ref, self_absco_ref + n);
    T.forr.assign(for_absco_ref, for_absco_ref + n);
    T.texp.assign(self_texp, self_texp + n);
    T.ref_t = ref_temp;
    T.ref_p = ref_press;
    T.xself = xself;
    T.xfrgn = xfrgn;
    T.gen++;
    c->plan_valid = false;
    return 0;
}

int clblm_contnm_scales(clblm_ctx *c, double xco2c, double xo3cn, double xo2cn, double xn2cn, double xrayl) {
    if (!c) return set_err(CLBLM_ERR_ARG, "null context");
    std::lock_guard<std::mutex> g(c->mu);
    const double v[5] = {xco2c, xo3cn, xo2cn, xn2cn, xrayl};
    for (int k = 0; k < 5; k++) {
        if (!(v[k] >= 0.)) return set_err(CLBLM_ERR_ARG, "continuum scale factors must be >= 0");
        c->mtckd.cn_scale[k] = v[k];
    }
    c->mtckd.gen++;
    c->plan_valid = false;
    return 0;
}

int clblm_contnm_absrb(clblm_ctx *c, int nlay, const clblm_layer *layers, const double *v1, const double *v2,
                       const int32_t *jrad, int32_t flags, double *d_absrb, int64_t ld, int32_t *nptabs,
                       void *stream) {
    if (!c || !layers || !v1 || !v2 || !jrad || !nptabs || nlay <= 0 || nlay > 65535)
        return set_err(CLBLM_ERR_ARG, "null argument or nlay outside 1..65535");
    HIPCHK(hipSetDevice(c->device));
    std::lock_guard<std::mutex> g(c->mu);
    std::vector<MtJob> jobs;
    std::vector<double> coef;
    std::vector<CnJob> cjobs;
    std::vector<CnTerm> cterms;
    std::vector<double> ccoef;
    int64_t maxpts = 1, cmax = 1;
    for (int l = 0; l < nlay; l++) {
        if (!(layers[l].tave > 0.) || !(layers[l].wtot > 0.))
            return set_err(CLBLM_ERR_ARG, "tave and wtot must be positive");
        double v1abs;
        absrb_grid(v1[l], v2[l], &v1abs, &nptabs[l]);
        if (d_absrb && ld < nptabs[l]) return set_err(CLBLM_ERR_ARG, "ld is below a layer's ABSRB length");
        if (flags & CLBLM_OD_MTCKD) {
            MtJob mj{};
            bool has = false;
            int rc = mtckd_plan(c->mtckd, layers[l], v1[l], v2[l], jrad[l], mj, coef, &has);
            if (rc) return rc;
            if (has) {
                mj.abs_off = (int64_t)l * ld;
                jobs.push_back(mj);
                maxpts = std::max<int64_t>(maxpts, mj.ihi - mj.ilo + 1);
            }
        }
        if (flags & CLBLM_OD_CONTNM) {
            CnJob cj{};
            bool has = false;
            int rc = cn_plan(c->mtckd.cn_scale, layers[l], v1[l], v2[l], jrad[l], cj, cterms, ccoef, &has);
            if (rc) return rc;
            if (has) {
                cj.abs_off = (int64_t)l * ld;
                cjobs.push_back(cj);
                cmax = std::max<int64_t>(cmax, cj.nptabs);
            }
        }
    }
    if (!d_absrb) return 0;
    hipStream_t st = stream ? (hipStream_t)stream : c->stream;
    auto up = [](size_t o) { return (o + 255) & ~(size_t)255; };
    const size_t oj = 0, oc = up(sizeof(MtJob) * std::max<size_t>(1, jobs.size()));
    const size_t ocj = up(oc + sizeof(double) * std::max<size_t>(1, coef.size()));
    const size_t oct = up(ocj + sizeof(CnJob) * std::max<size_t>(1, cjobs.size()));
    const size_t occ = up(oct + sizeof(CnTerm) * std::max<size_t>(1, cterms.size()));
    const size_t bytes = occ + sizeof(double) * std::max<size_t>(1, ccoef.size());
    HIPCHK(hipStreamSynchronize(st));   // the buffer may still be read by a previous call
    if (bytes > c->mt_cap) {
        (void)hipFree(c->mt_buf);
        c->mt_buf = nullptr;
        c->mt_cap = 0;
        HIPCHK(hipMalloc(&c->mt_buf, bytes));
        c->mt_cap = bytes;
    }
    char *b = (char *)c->mt_buf;
    if (!jobs.empty()) {
        HIPCHK(hipMemcpyAsync(b + oj, jobs.data(), sizeof(MtJob) * jobs.size(), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(b + oc, coef.data(), sizeof(double) * coef.size(), hipMemcpyHostToDevice, st));
    }
    if (!cjobs.empty()) {
        HIPCHK(hipMemcpyAsync(b + ocj, cjobs.data(), sizeof(CnJob) * cjobs.size(), hipMemcpyHostToDevice, st));
        if (!cterms.empty())
            HIPCHK(hipMemcpyAsync(b + oct, cterms.data(), sizeof(CnTerm) * cterms.size(), hipMemcpyHostToDevice, st));
        if (!ccoef.empty())
            HIPCHK(hipMemcpyAsync(b + occ, ccoef.data(), sizeof(double) * ccoef.size(), hipMemcpyHostToDevice, st));
    }
    const size_t nb = sizeof(double) * (size_t)ld * (size_t)nlay;
    hipLaunchKernelGGL(k_fill_bytes, dim3(cdiv((int64_t)nb, 256)), dim3(256), 0, st, (uint8_t *)d_absrb, nb, (uint8_t)0);
    if (!jobs.empty())
        hipLaunchKernelGGL(k_mtckd, dim3(cdiv(maxpts, 256), (unsigned)jobs.size()), dim3(256), 0, st,
                           (const MtJob *)(b + oj), (const double *)(b + oc), d_absrb);
    if (!cjobs.empty())
        hipLaunchKernelGGL(k_contnm, dim3(cdiv(cmax, 256), (unsigned)cjobs.size()), dim3(256), 0, st,
                           (const CnJob *)(b + ocj), (const CnTerm *)(b + oct), (const double *)(b + occ), d_absrb);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(st));   // the host job lists go out of scope
    return 0;
}

int clblm_mtckd_absrb(clblm_ctx *c, int nlay, const clblm_layer *layers, const double *v1, const double *v2,
                      const int32_t *jrad, double *d_absrb, int64_t ld, int32_t *nptabs, void *stream) {
    return clblm_contnm_absrb(c, nlay, layers, v1, v2, jrad, CLBLM_OD_MTCKD, d_absrb, ld, nptabs, stream);
}

int clblm_dev_alloc(clblm_ctx *c, int64_t bytes, void **d) {
    if (!c || !d || bytes <= 0) return set_err(CLBLM_ERR_ARG, "null argument or no bytes");
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipMalloc(d, (size_t)bytes));
    return 0;
}

int clblm_dev_free(clblm_ctx *c, void *d) {
    if (!c) return set_err(CLBLM_ERR_ARG, "null context");
    HIPCHK(hipSetDevice(c->device));
    drain(c, false);   // batches still reading it finish first
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipFree(d));
    return 0;
}

int clblm_dev_put(clblm_ctx *c, void *d, const void *h, int64_t bytes) {
    if (!c || !d || !h || bytes < 0) return set_err(CLBLM_ERR_ARG, "null argument");
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipMemcpy(d, h, (size_t)bytes, hipMemcpyHostToDevice));
    return 0;
}

int clblm_dev_get(clblm_ctx *c, void *h, const void *d, int64_t bytes) {
    if (!c || !d || !h || bytes < 0) return set_err(CLBLM_ERR_ARG, "null argument");
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipMemcpy(h, d, (size_t)bytes, hipMemcpyDeviceToHost));
    return 0;
}

int clblm_filter(clblm_ctx *c, const double *d_in, int64_t n, double v1, double dv, int nchan, const double *v1f,
                 const double *dvf, const int32_t *nptf, const int64_t *xf_off, const double *xf, double *d_out,
                 void *stream) {
    if (!c) return set_err(CLBLM_ERR_ARG, "null context");
    HIPCHK(hipSetDevice(c->device));
    std::lock_guard<std::mutex> g(c->mu);
    if (!c->scan) c->scan = clblm_scan_state_new();
    hipStream_t st = stream ? (hipStream_t)stream : c->stream;
    if (c->last_stream && c->last_stream != st && c->ncalls > 0) {   // after the batches that made d_in
        const int kp = (int)((c->ncalls + clblm_ctx::NSLOT - 1) % clblm_ctx::NSLOT);
        HIPCHK(hipStreamWaitEvent(st, c->slot[kp].done, 0));
    }
    int rc = clblm_filter_impl(c->scan, d_in, n, v1, dv, nchan, v1f, dvf, nptf, xf_off, xf, d_out, st);
    if (rc) return set_err(rc, clblm_scan_error());
    return 0;
}

int clblm_scanfn(clblm_ctx *c, const double *d_in, int64_t n, int64_t nrows, int64_t ld_in, double v1, double dv,
                 double v1out, double v2out, double dvout, int funct_id, double hwhm, double *d_out, int64_t ld_out,
                 int64_t *nout, double *v1res, double *dvres, void *stream) {
    if (!c || !nout || !v1res || !dvres) return set_err(CLBLM_ERR_ARG, "null argument");
    HIPCHK(hipSetDevice(c->device));
    std::lock_guard<std::mutex> g(c->mu);
    if (!c->scan) c->scan = clblm_scan_state_new();
    hipStream_t st = stream ? (hipStream_t)stream : c->stream;
    if (d_in && c->last_stream && c->last_stream != st && c->ncalls > 0) {  // after the batches that made d_in
        const int kp = (int)((c->ncalls + clblm_ctx::NSLOT - 1) % clblm_ctx::NSLOT);
        HIPCHK(hipStreamWaitEvent(st, c->slot[kp].done, 0));
    }
    int rc = clblm_scanfn_impl(c->scan, d_in, n, nrows, ld_in, v1, dv, v1out, v2out, dvout, funct_id, hwhm, d_out,
                               ld_out, nout, v1res, dvres, st);
    if (rc) return set_err(rc, clblm_scan_error());
    return 0;
}

int clblm_fft_scan(clblm_ctx *c, const double *d_in, int64_t n, double v1, double dv, double v1out, double v2out,
                   double dvout, int funct_id, double hwhm, double fill, double pre_box_hw, int deconv_pre_box,
                   int precision, double *d_out, int64_t *nout, double *v1res, double *dvres, void *stream) {
    if (!c || !nout || !v1res || !dvres) return set_err(CLBLM_ERR_ARG, "null argument");
    HIPCHK(hipSetDevice(c->device));
    std::lock_guard<std::mutex> g(c->mu);
    if (!c->scan) c->scan = clblm_scan_state_new();
    hipStream_t st = stream ? (hipStream_t)stream : c->stream;
    if (d_out && c->last_stream && c->last_stream != st && c->ncalls > 0) {  // after the batches that made d_in
        const int kp = (int)((c->ncalls + clblm_ctx::NSLOT - 1) % clblm_ctx::NSLOT);
        HIPCHK(hipStreamWaitEvent(st, c->slot[kp].done, 0));
    }
    int rc = clblm_fft_scan_impl(c->scan, d_in, n, v1, dv, v1out, v2out, dvout, funct_id, hwhm, fill, pre_box_hw,
                                 deconv_pre_box, precision, d_out, nout, v1res, dvres, st);
    if (rc) return set_err(rc, clblm_scan_error());
    return 0;
}

int clblm_timing(clblm_ctx *c, int reset, double *ms_prep, double *ms_conv, double *ms_f4, double *ms_panel,
                 int64_t *launches) {
    if (!c) return set_err(CLBLM_ERR_ARG, "ctx is NULL");
    {
        std::lock_guard<std::mutex> g(c->mu);
        drain(c, true);
    }
    int64_t n = c->launches > 0 ? c->launches : 1;
    if (ms_prep) *ms_prep = c->ms[0] / n;
    if (ms_conv) *ms_conv = c->ms[1] / n;
    if (ms_f4) *ms_f4 = c->ms[2] / n;
    if (ms_panel) *ms_panel = c->ms[3] / n;
    if (launches) *launches = c->launches;
    if (reset) { c->ms[0] = c->ms[1] = c->ms[2] = c->ms[3] = 0; c->launches = 0; }
    return 0;
}

int clblm_work(clblm_ctx *c, int reset, int64_t *conv_evals, int64_t *f4_evals, double *ms_conv_kernel,
               double *ms_f4_kernel) {
    if (!c) return set_err(CLBLM_ERR_ARG, "ctx is NULL");
    {
        std::lock_guard<std::mutex> g(c->mu);
        drain(c, true);
    }
    unsigned long long w[2];
    HIPCHK(hipMemcpy(w, c->d_work, sizeof w, hipMemcpyDeviceToHost));
    int64_t n = c->launches > 0 ? c->launches : 1;
    if (conv_evals) *conv_evals = (int64_t)(w[0] / (unsigned long long)n);
    if (f4_evals) *f4_evals = (int64_t)(w[1] / (unsigned long long)n);
    if (ms_conv_kernel) *ms_conv_kernel = c->ms_conv_k / n;
    if (ms_f4_kernel) *ms_f4_kernel = c->ms_f4_k / n;
    if (reset) {
        HIPCHK(hipMemset(c->d_work, 0, sizeof w));
        c->ms_conv_k = c->ms_f4_k = 0;
        c->ms[0] = c->ms[1] = c->ms[2] = c->ms[3] = 0;
        c->launches = 0;
    }
    return 0;
}

int clblm_level2layer(int nlay, int nmol, const double *p_lev, const double *t_lev, const double *t_lay,
                      const double *q_lev, double *dt_dtup, double *dt_dtlo, double *dlnw_dlnq_up,
                      double *dlnw_dlnq_lo) {
    if (nlay <= 0 || nmol < 0 || !p_lev || !t_lev || !t_lay || !dt_dtup || !dt_dtlo)
        return set_err(CLBLM_ERR_ARG, "null argument or no layers");
    if (nmol > 0 && (!q_lev || !dlnw_dlnq_up || !dlnw_dlnq_lo))
        return set_err(CLBLM_ERR_ARG, "molecular densities and outputs are needed when nmol > 0");
    for (int k = 0; k <= nlay; k++)
        if (!(p_lev[k] > 0.) || !(t_lev[k] > 0.)) return set_err(CLBLM_ERR_ARG, "level pressures and temperatures must be positive");
    for (int k = 0; k < nlay; k++)
        if ((p_lev[k + 1] / t_lev[k + 1]) / (p_lev[k] / t_lev[k]) == 1.0)
            return set_err(CLBLM_ERR_ARG, "a layer has equal air densities at both levels");
    clblm_level2layer_impl(nlay, nmol, p_lev, t_lev, t_lay, q_lev, dt_dtup, dt_dtlo, dlnw_dlnq_up, dlnw_dlnq_lo);
    return 0;
}

int clblm_rt_jac(clblm_ctx *c, int nlay, int64_t npts, double v1, double dv, const double *d_od,
                 const double *d_tlay, const double *d_tlev, const double *d_dt_dtup, const double *d_dt_dtlo,
                 double surf_emis, double t_skin, int lin_in_tau, double sec_dn, double *d_rad, double *d_dr_dtau,
                 double *d_dr_dtlev, double *d_dr_dtskin, double *d_dr_demis, void *stream) {
    if (!c || nlay <= 0 || npts <= 0 || !d_od || !d_tlay || !d_tlev || !d_rad || !d_dr_dtau || !d_dr_dtlev)
        return set_err(CLBLM_ERR_ARG, "null argument or empty grid");
    if (lin_in_tau < 0 || lin_in_tau > 2) return set_err(CLBLM_ERR_ARG, "lin_in_tau must be 0, 1 or 2");
    if (lin_in_tau != 2 && (!d_dt_dtup || !d_dt_dtlo))
        return set_err(CLBLM_ERR_ARG, "lin_in_tau 0 and 1 need dT/dT_lev of the layer means");
    HIPCHK(hipSetDevice(c->device));
    std::lock_guard<std::mutex> g(c->mu);
    const size_t bytes = sizeof(double) * (size_t)(3 * nlay + 2) * (size_t)npts;
    if (c->rt_state_cap < bytes) {
        if (c->rt_state) (void)hipFree(c->rt_state);
        c->rt_state = nullptr;
        c->rt_state_cap = 0;
        HIPCHK(hipMalloc(&c->rt_state, bytes));
        c->rt_state_cap = bytes;
    }
    hipStream_t st = stream ? (hipStream_t)stream : c->stream;
    int rc = clblm_rt_jac_impl(nlay, npts, v1, dv, d_od, d_tlay, d_tlev, d_dt_dtup, d_dt_dtlo, surf_emis, t_skin,
                               lin_in_tau, sec_dn, d_rad, d_dr_dtau, d_dr_dtlev, d_dr_dtskin, d_dr_demis,
                               c->rt_state, st);
    if (rc) return set_err(rc, "Jacobian kernel launch failed");
    HIPCHK(hipStreamSynchronize(st));   // the scratch states are reused by the next call
    return 0;
}

int clblm_interp_spectrum(clblm_ctx *c, const double *d_in, int64_t nin, double v1, double dvin, double *d_out,
                          int64_t nout, double dvout, void *stream) {
    if (!c || !d_in || !d_out || nin < 2 || nout < 1) return set_err(CLBLM_ERR_ARG, "null argument or too few points");
    if (d_in == d_out) return set_err(CLBLM_ERR_ARG, "input and output must not overlap");
    if (!(dvin > 0.) || !(dvout > 0.)) return set_err(CLBLM_ERR_ARG, "DVs must be > 0");
    HIPCHK(hipSetDevice(c->device));
    hipStream_t st = stream ? (hipStream_t)stream : c->stream;
    if (clblm_interp_spectrum_impl(d_in, nin, v1, dvin, d_out, nout, dvout, st))
        return set_err(CLBLM_ERR_HIP, "interpolation kernel launch failed");
    return 0;
}

int clblm_interp_to_finest(clblm_ctx *c, int nspec, double v1, const double *dv, const int64_t *nlim,
                           const double *const *d_in, double *const *d_out, double *dv_out, int64_t *nlim_out,
                           void *stream) {
    if (!c || nspec <= 0 || !dv || !nlim || !dv_out || !nlim_out) return set_err(CLBLM_ERR_ARG, "null argument");
    for (int k = 0; k < nspec; k++)
        if (!(dv[k] > 0.) || nlim[k] < 2) return set_err(CLBLM_ERR_ARG, "every spectrum needs DV > 0 and 2 points");
    const int loc = clblm_finest_grid(nspec, dv, nlim);
    *dv_out = dv[loc];
    *nlim_out = nlim[loc];
    if (!d_out) return 0;
    if (!d_in) return set_err(CLBLM_ERR_ARG, "null input pointers");
    HIPCHK(hipSetDevice(c->device));
    hipStream_t st = stream ? (hipStream_t)stream : c->stream;
    for (int k = 0; k < nspec; k++) {
        if (!d_in[k] || !d_out[k]) return set_err(CLBLM_ERR_ARG, "null spectrum pointer");
        const bool same = k == loc || clblm_interp::same_grid(v1, dv[k], nlim[k], dv[loc], nlim[loc]);
        if (same) {
            // spectraHaveSameGrid: left as it is (copied where the output is elsewhere)
            if (d_out[k] != d_in[k])
                HIPCHK(hipMemcpyAsync(d_out[k], d_in[k], sizeof(double) * (size_t)nlim[k], hipMemcpyDeviceToDevice,
                                      st));
            continue;
        }
        if (d_out[k] == d_in[k]) return set_err(CLBLM_ERR_ARG, "a spectrum to interpolate must have its own output");
        if (clblm_interp_spectrum_impl(d_in[k], nlim[k], v1, dv[k], d_out[k], nlim[loc], dv[loc], st))
            return set_err(CLBLM_ERR_HIP, "interpolation kernel launch failed");
    }
    return 0;
}

int clblm_rt_up_grids(clblm_ctx *c, int nlay, double v1, const double *dv, const int64_t *nlim, const int64_t *od_off,
                      const double *d_od, const double *t_lay, const double *t_lev, double surf_emis, double t_skin,
                      double surf_dv, int64_t surf_nlim, int lin_in_tau, double sec_dn, double *d_rad, double *dv_out,
                      int64_t *nlim_out, void *stream) {
    if (!c || nlay <= 0 || !dv || !nlim || !dv_out || !nlim_out) return set_err(CLBLM_ERR_ARG, "null argument");
    if (lin_in_tau < 0 || lin_in_tau > 2) return set_err(CLBLM_ERR_ARG, "lin_in_tau must be 0, 1 or 2");
    if (!(surf_dv > 0.) || surf_nlim < 2) return set_err(CLBLM_ERR_ARG, "surface grid needs DV > 0 and 2 points");
    for (int l = 0; l < nlay; l++)
        if (!(dv[l] > 0.) || nlim[l] < 2) return set_err(CLBLM_ERR_ARG, "every layer needs DV > 0 and 2 points");
    int64_t max_n = 0;
    clblm_rt_grids_final(nlay, v1, dv, nlim, surf_dv, surf_nlim, dv_out, nlim_out, &max_n);
    if (!d_rad) return 0;
    if (!od_off || !d_od || !t_lay || !t_lev) return set_err(CLBLM_ERR_ARG, "null argument");
    HIPCHK(hipSetDevice(c->device));
    std::lock_guard<std::mutex> g(c->mu);
    const size_t bytes = sizeof(double) * 12 * (size_t)max_n;
    if (c->grid_state_cap < bytes) {
        if (c->grid_state) (void)hipFree(c->grid_state);
        c->grid_state = nullptr;
        c->grid_state_cap = 0;
        HIPCHK(hipMalloc(&c->grid_state, bytes));
        c->grid_state_cap = bytes;
    }
    hipStream_t st = stream ? (hipStream_t)stream : c->stream;
    int rc = clblm_rt_up_grids_impl(nlay, v1, dv, nlim, od_off, d_od, t_lay, t_lev, surf_emis, t_skin, surf_dv,
                                    surf_nlim, lin_in_tau, sec_dn, c->grid_state, max_n, d_rad, st);
    if (rc) return set_err(rc, "radiance kernel launch failed");
    HIPCHK(hipStreamSynchronize(st));   // the scratch arrays are reused by the next call
    return 0;
}

int clblm_rt_mode1(clblm_ctx *c, const clblm_mode1_params *p, const double *dv, const int64_t *nlim,
                   const int64_t *od_off, const double *d_od_up, const double *d_od_dn, const double *d_od_sun,
                   const double *fac_up, const double *fac_dn, const double *fac_sun, const double *t_lay,
                   const double *t_lev, const double *d_sol, int64_t sol_n, double sol_dv, double *d_rad,
                   double *d_tx, double *dv_out, int64_t *nlim_out, void *stream) {
    if (!c || !p || !dv || !nlim || !dv_out || !nlim_out || p->nlay <= 0) return set_err(CLBLM_ERR_ARG, "null argument");
    const int nlay = p->nlay;
    if (p->lin_in_tau < 0 || p->lin_in_tau > 2) return set_err(CLBLM_ERR_ARG, "lin_in_tau must be 0, 1 or 2");
    if (!(p->v2 > p->v1)) return set_err(CLBLM_ERR_ARG, "v2 must exceed v1");
    if (p->obs_level < 0 || p->obs_level > nlay) return set_err(CLBLM_ERR_ARG, "obs_level must lie in 0..nlay");
    if (!p->thermal && p->surf_emis < 1.)
        return set_err(CLBLM_ERR_ARG, "a reflecting surface needs the thermal down path");
    if (!(p->surf_emis < 1.) && !(p->t_skin > 0.))
        return set_err(CLBLM_ERR_ARG, "no surface term: emissivity 1 and no skin temperature");
    if (p->solar && (!(sol_dv > 0.) || sol_n < 2)) return set_err(CLBLM_ERR_ARG, "solar spectrum needs DV > 0 and 2 points");
    int64_t max_n = 2;
    for (int l = 0; l < nlay; l++) {
        if (!(dv[l] > 0.) || nlim[l] < 2) return set_err(CLBLM_ERR_ARG, "every layer needs DV > 0 and 2 points");
        max_n = std::max(max_n, nlim[l]);
    }
    if (p->solar) max_n = std::max(max_n, sol_n);
    std::vector<double> one((size_t)nlay, 1.0);
    const double *fu = fac_up ? fac_up : one.data(), *fd = fac_dn ? fac_dn : one.data();
    const double *fs = fac_sun ? fac_sun : one.data();
    if (!d_rad)
        return clblm_rt_mode1_impl(nlay, p->v1, p->v2, dv, nlim, od_off, d_od_up, d_od_dn, d_od_sun, fu, fd, fs,
                                   t_lay, t_lev, p->obs_level, p->lin_in_tau, p->thermal, p->solar, p->surf_emis,
                                   p->t_skin, p->surf_refl, d_sol, sol_n, sol_dv, nullptr, max_n, nullptr, nullptr,
                                   dv_out, nlim_out, nullptr);
    if (!od_off || !t_lay || !t_lev || !d_od_up || (p->thermal && !d_od_dn) || (p->solar && (!d_od_sun || !d_sol)))
        return set_err(CLBLM_ERR_ARG, "null argument");
    HIPCHK(hipSetDevice(c->device));
    std::lock_guard<std::mutex> g(c->mu);
    const size_t bytes = sizeof(double) * 12 * (size_t)max_n;
    if (c->grid_state_cap < bytes) {
        if (c->grid_state) (void)hipFree(c->grid_state);
        c->grid_state = nullptr;
        c->grid_state_cap = 0;
        HIPCHK(hipMalloc(&c->grid_state, bytes));
        c->grid_state_cap = bytes;
    }
    hipStream_t st = stream ? (hipStream_t)stream : c->stream;
    int rc = clblm_rt_mode1_impl(nlay, p->v1, p->v2, dv, nlim, od_off, d_od_up, d_od_dn, d_od_sun, fu, fd, fs, t_lay,
                                 t_lev, p->obs_level, p->lin_in_tau, p->thermal, p->solar, p->surf_emis, p->t_skin,
                                 p->surf_refl, d_sol, sol_n, sol_dv, c->grid_state, max_n, d_rad, d_tx, dv_out,
                                 nlim_out, st);
    if (rc) return set_err(rc, "merge-down radiance failed");
    HIPCHK(hipStreamSynchronize(st));   // the scratch arrays are reused by the next call
    return 0;
}

int clblm_rt_jac_paths(clblm_ctx *c, const clblm_jac_paths_params *p, int64_t npts, double v1, double dv,
                       const double *d_od, const double *d_tlay, const double *d_tlev, const double *d_dt_dtup,
                       const double *d_dt_dtlo, const double *d_sol, double *d_rad, double *d_dr_dtau,
                       double *d_dr_dtlev, double *d_dr_dtskin, double *d_dr_demis, double *d_dr_drsfc,
                       void *stream) {
    if (!c || !p || p->nlay <= 0 || npts <= 0 || !d_od || !d_tlay || !d_tlev || !d_rad || !d_dr_dtau || !d_dr_dtlev)
        return set_err(CLBLM_ERR_ARG, "null argument or empty grid");
    const int L = p->nlay;
    if (p->lin_in_tau < 0 || p->lin_in_tau > 2) return set_err(CLBLM_ERR_ARG, "lin_in_tau must be 0, 1 or 2");
    if (p->obs_level < 0 || p->obs_level > L) return set_err(CLBLM_ERR_ARG, "obs_level must lie in 0..nlay");
    if (p->lin_in_tau != 2 && (!d_dt_dtup || !d_dt_dtlo))
        return set_err(CLBLM_ERR_ARG, "lin_in_tau 0 and 1 need dT/dT_lev of the layer means");
    if (p->solar && !p->uplook && !d_sol) return set_err(CLBLM_ERR_ARG, "the solar term needs the solar radiance");
    std::vector<double> amf((size_t)3 * L, 1.0);
    const double *src[3] = {p->amf_view, p->amf_down, p->amf_sun};
    for (int k = 0; k < 3; k++)
        if (src[k]) std::copy(src[k], src[k] + L, amf.begin() + (size_t)k * L);
    HIPCHK(hipSetDevice(c->device));
    std::lock_guard<std::mutex> g(c->mu);
    const size_t na = ((size_t)3 * L + 31) & ~(size_t)31;
    const size_t bytes = sizeof(double) * (na + (size_t)5 * (L + 1) * (size_t)npts);
    if (c->grid_state_cap < bytes) {
        if (c->grid_state) (void)hipFree(c->grid_state);
        c->grid_state = nullptr;
        c->grid_state_cap = 0;
        HIPCHK(hipMalloc(&c->grid_state, bytes));
        c->grid_state_cap = bytes;
    }
    hipStream_t st = stream ? (hipStream_t)stream : c->stream;
    double *d_amf = c->grid_state;
    HIPCHK(hipMemcpyAsync(d_amf, amf.data(), sizeof(double) * amf.size(), hipMemcpyHostToDevice, st));
    int rc = clblm_rt_jac_paths_impl(L, npts, v1, dv, d_od, d_amf, d_amf + L, d_amf + 2 * L, d_tlay, d_tlev,
                                     p->surf_emis, p->t_skin, p->solar && !p->uplook ? 1 : 0, d_sol, p->rsfc_sun,
                                     p->obs_level, p->uplook ? 1 : 0, p->lin_in_tau, d_dt_dtup, d_dt_dtlo, d_rad,
                                     d_dr_dtau, d_dr_dtlev, d_dr_dtskin, d_dr_demis, d_dr_drsfc, d_amf + na, st);
    if (rc) return set_err(rc, "Jacobian kernel failed");
    HIPCHK(hipStreamSynchronize(st));   // the host factors and the scratch are reused by the next call
    return 0;
}

int clblm_rt_flux(clblm_ctx *c, const clblm_flux_params *p, const double *d_od, int64_t npts, const double *d_tlay,
                  const double *d_tlev, const double *p_lev, double *flux_up, double *flux_dn, double *net_flux,
                  double *heating, int32_t *nbin, void *stream) {
    if (!c || !p || !nbin || p->nlay <= 0) return set_err(CLBLM_ERR_ARG, "null argument");
    if (p->nang < 1 || p->nang > 3) return set_err(CLBLM_ERR_ARG, "nang must be 1, 2 or 3");
    if (p->lin_in_tau < 0 || p->lin_in_tau > 2) return set_err(CLBLM_ERR_ARG, "lin_in_tau must be 0, 1 or 2");
    if (!(p->v2 > p->v1) || !(p->dv > 0.) || !(p->dv_flux > 0.) || npts < 2)
        return set_err(CLBLM_ERR_ARG, "need v2 > v1, dv > 0, dv_flux > 0 and 2 points");
    // the number of output groups (clblm_noScattRT.f90:197-202, EPS = 1e-4)
    const double OUT = (p->v2 - p->v1) / p->dv_flux;
    const int NOUT = (int)(OUT + 1.E-4);
    if (std::fabs((double)NOUT - OUT) > 1.E-4 || NOUT < 1)
        return set_err(CLBLM_ERR_ARG, "V1, V2 and DV_flux are inconsistent (the reference stops)");
    const int outinrat = (int)(p->dv_flux / p->dv);
    if (outinrat < 1) return set_err(CLBLM_ERR_ARG, "dv_flux is below dv");
    *nbin = NOUT;
    if (!flux_up) return 0;
    if (!d_od || !d_tlay || !d_tlev || !p_lev || !flux_dn || !net_flux || !heating)
        return set_err(CLBLM_ERR_ARG, "null argument");
    const int nlay = p->nlay, nlev = nlay + 1;
    HIPCHK(hipSetDevice(c->device));
    std::lock_guard<std::mutex> g(c->mu);
    const int64_t gpb = std::max<int64_t>(1, ((int64_t)1 << 18) / outinrat);
    const int64_t blk = gpb * outinrat;
    const size_t nscr = (size_t)(2 * p->nang * nlay + 1) * (size_t)blk;
    const size_t nout_arr = 2 * (size_t)nlev * (size_t)NOUT;
    const size_t bytes = sizeof(double) * (nscr + nout_arr);
    if (c->grid_state_cap < bytes) {
        if (c->grid_state) (void)hipFree(c->grid_state);
        c->grid_state = nullptr;
        c->grid_state_cap = 0;
        HIPCHK(hipMalloc(&c->grid_state, bytes));
        c->grid_state_cap = bytes;
    }
    double *fdn = c->grid_state + nscr, *fup = fdn + (size_t)nlev * NOUT;
    hipStream_t st = stream ? (hipStream_t)stream : c->stream;
    const size_t zb = sizeof(double) * nout_arr;   // groups no point reaches stay 0
    hipLaunchKernelGGL(k_fill_bytes, dim3(cdiv((int64_t)zb, 256)), dim3(256), 0, st, (uint8_t *)fdn, zb, (uint8_t)0);
    int rc = clblm_rt_flux_impl(nlay, npts, p->v1, p->dv, d_od, d_tlay, d_tlev, p->lin_in_tau, p->surf_emis,
                                p->t_skin, p->nang, outinrat, NOUT, p->lambertian, c->grid_state, blk, fdn, fup, st);
    if (rc) return set_err(rc, "flux kernels failed");
    HIPCHK(hipMemcpyAsync(flux_dn, fdn, sizeof(double) * (size_t)nlev * NOUT, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(flux_up, fup, sizeof(double) * (size_t)nlev * NOUT, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    // the driver's net flux and heating rate (clblm_Drivers.f90:455-466), HEATFC of
    // clblm_ConstParam.f90:45
    const double HEATFC = 8.4391;
    for (int64_t q = 0; q < (int64_t)nlev * NOUT; q++) net_flux[q] = flux_up[q] - flux_dn[q];
    for (int j = nlev - 1; j >= 0; j--)
        for (int b = 0; b < NOUT; b++) {
            const int64_t q = (int64_t)j * NOUT + b;
            heating[q] = j == nlev - 1 ? 0. : HEATFC * (net_flux[q] - net_flux[q + NOUT]) / (p_lev[j] - p_lev[j + 1]);
        }
    return 0;
}

int clblm_rt_jac_noalloc(clblm_ctx *c, int nlay, int64_t npts, double v1, double dv, const double *d_od,
                         const double *d_tlay, const double *d_tlev, const double *d_dt_dtup,
                         const double *d_dt_dtlo, double surf_emis, double t_skin, int lin_in_tau, double sec_dn,
                         double *d_rad, double *d_dr_dtau, double *d_dr_dtlev, double *d_dr_dtskin,
                         double *d_dr_demis, void *stream) {
    if (!c || nlay <= 0 || npts <= 0 || !d_od || !d_tlay || !d_tlev || !d_rad || !d_dr_dtau || !d_dr_dtlev)
        return set_err(CLBLM_ERR_ARG, "null argument or empty grid");
    if (lin_in_tau < 0 || lin_in_tau > 2) return set_err(CLBLM_ERR_ARG, "lin_in_tau must be 0, 1 or 2");
    if (lin_in_tau != 2 && (!d_dt_dtup || !d_dt_dtlo))
        return set_err(CLBLM_ERR_ARG, "lin_in_tau 0 and 1 need dT/dT_lev of the layer means");
    const double need = 8.0 * (3.0 * (double)nlay + 2.0) * (double)npts;
    if (need > (double)CLBLM_RT_JAC_NOALLOC_BYTES) {
        char msg[200];
        snprintf(msg, sizeof msg, "clblm_rt_jac_noalloc: %d layers x %lld points need %.0f bytes of scratch, "
                 "more than its %lld (8 (3 nlay + 2) npts <= that)", nlay, (long long)npts, need,
                 (long long)CLBLM_RT_JAC_NOALLOC_BYTES);
        return set_err(CLBLM_ERR_CAPACITY, msg);
    }
    HIPCHK(hipSetDevice(c->device));
    hipStream_t st = stream ? (hipStream_t)stream : c->stream;
    int rc = clblm_rt_jac_scratch(c->device, nlay, npts, v1, dv, d_od, d_tlay, d_tlev, d_dt_dtup, d_dt_dtlo,
                                  surf_emis, t_skin, lin_in_tau, sec_dn, d_rad, d_dr_dtau, d_dr_dtlev, d_dr_dtskin,
                                  d_dr_demis, st);
    if (rc) return set_err(rc, "Jacobian kernel launch on the static scratch failed");
    return 0;
}

int clblm_level_jac(clblm_ctx *c, int nlay, int nmol, int64_t npts, const double *d_od_var,
                    const double *d_dr_dtau, const double *d_dr_dtlev, const double *d_wtot, const double *d_tlay,
                    double dt, double dw_frac, const double *d_dt_dtup, const double *d_dt_dtlo,
                    const double *d_dw_dqup, const double *d_dw_dqlo, double *d_jac_t, double *d_jac_q,
                    void *stream) {
    if (!c || nlay <= 0 || nmol < 0 || npts <= 0 || !d_od_var || !d_dr_dtau || !d_dr_dtlev || !d_wtot || !d_tlay ||
        !d_dt_dtup || !d_dt_dtlo || !d_jac_t)
        return set_err(CLBLM_ERR_ARG, "null argument or empty grid");
    if (nmol > 0 && d_jac_q && (!d_dw_dqup || !d_dw_dqlo))
        return set_err(CLBLM_ERR_ARG, "molecular Jacobians need dlnW/dlnq");
    if (!(dt != 0.) || !(dw_frac != 0.)) return set_err(CLBLM_ERR_ARG, "dt and dw_frac must be nonzero");
    if (nmol + 1 > 65535) return set_err(CLBLM_ERR_ARG, "too many molecules");
    HIPCHK(hipSetDevice(c->device));
    hipStream_t st = stream ? (hipStream_t)stream : c->stream;
    int rc = clblm_level_jac_impl(nlay, nmol, npts, d_od_var, d_dr_dtau, d_dr_dtlev, d_wtot, d_tlay, dt, dw_frac,
                                  d_dt_dtup, d_dt_dtlo, d_dw_dqup, d_dw_dqlo, d_jac_t, d_jac_q, st);
    if (rc) return set_err(rc, "level Jacobian kernel launch failed");
    return 0;
}

int clblm_rt_down(clblm_ctx *c, int nlay, int64_t npts, double v1, double dv, const double *d_od,
                  const double *d_tlay, const double *d_tlev, int lin_in_tau, const double *d_sol_toa,
                  double *d_rad, double *d_tx, void *stream) {
    if (!c || nlay <= 0 || npts <= 0 || !d_od || !d_tlay || !d_tlev || !d_rad)
        return set_err(CLBLM_ERR_ARG, "null argument or empty grid");
    if (lin_in_tau < 0 || lin_in_tau > 2) return set_err(CLBLM_ERR_ARG, "lin_in_tau must be 0, 1 or 2");
    HIPCHK(hipSetDevice(c->device));
    hipStream_t st = stream ? (hipStream_t)stream : c->stream;
    int rc = clblm_rt_down_impl(nlay, npts, v1, dv, d_od, d_tlay, d_tlev, lin_in_tau, d_sol_toa, d_rad, d_tx, st);
    if (rc) return set_err(rc, "radiance kernel launch failed");
    return 0;
}

int clblm_rt_up(clblm_ctx *c, int nlay, int64_t npts, double v1, double dv, const double *d_od,
                const double *d_tlay, const double *d_tlev, double surf_emis, double t_skin, int lin_in_tau,
                double sec_dn, double *d_rad, double *d_dr_dod, double *d_dr_dtlay, double *d_dr_dtskin,
                double *d_dr_demis, void *stream) {
    if (!c || nlay <= 0 || npts <= 0 || !d_od || !d_tlay || !d_tlev || !d_rad)
        return set_err(CLBLM_ERR_ARG, "null argument or empty grid");
    if (lin_in_tau < 0 || lin_in_tau > 2) return set_err(CLBLM_ERR_ARG, "lin_in_tau must be 0, 1 or 2");
    if (d_dr_dod && !d_dr_dtlay) return set_err(CLBLM_ERR_ARG, "dR/dOD needs the dR/dT output too");
    HIPCHK(hipSetDevice(c->device));
    std::lock_guard<std::mutex> g(c->mu);
    double *state = nullptr;
    if (d_dr_dod) {
        size_t bytes = sizeof(double) * 2 * (size_t)nlay * (size_t)npts;
        if (c->rt_state_cap < bytes) {
            if (c->rt_state) (void)hipFree(c->rt_state);
            c->rt_state = nullptr;
            c->rt_state_cap = 0;
            HIPCHK(hipMalloc(&c->rt_state, bytes));
            c->rt_state_cap = bytes;
        }
        state = c->rt_state;
    }
    hipStream_t st = stream ? (hipStream_t)stream : c->stream;
    int rc = clblm_rt_up_impl(nlay, npts, v1, dv, d_od, d_tlay, d_tlev, surf_emis, t_skin, lin_in_tau, sec_dn,
                              d_rad, d_dr_dod, d_dr_dtlay, d_dr_dtskin, d_dr_demis, state, st);
    if (rc) return set_err(rc, "radiance kernel launch failed");
    HIPCHK(hipStreamSynchronize(st));
    return 0;
}

}  // extern "C"

// Layer ODs with their derivatives (the tangent batch)
#include "clblm_tan.inc"
