// Instrument scan on the GPU: clblm_FFTSCN_struct (clblm_FFT.f90:352-844) for the
// boxcar scan, the scanning functions 2..11 (triangle, Gaussian, sinc**2, sinc, Beer,
// Hamming, Hanning, Norton-Beer weak/medium/strong), the optional boxcar pre-scan with
// or without deconvolution, the externally pre-scanned and deconvolution-only modes.
//
//   host    the scan bookkeeping (V1Z/V2Z, LoadSpect's LTOTAL and record rounding, the
//           boxcar's new grid, wrtSpect's window), the apodization blocks and the
//           control flow of the 4-point interpolation (Intpdr/clblm_INTERP/panelInput),
//           which depends on the grids only and is cut into segments that restart the
//           accumulated fraction
//   device  load + fill (or the boxcar means), real-to-complex FFT (hipFFT), the
//           apodization evaluated analytically (IVX = 1) or the normalised transform of
//           the scanning function (IVX = -1, below CRATIO half widths), the boxcar
//           deconvolution, complex-to-real FFT, window, and one lane per
//           interpolation segment
//
// The reference transforms LREC*65536 points (LREC = 4 x the next power of two of
// the data records, clblm_FFT.f90:1183-1197); the same length is used here, so the
// wrap-around of the periodic convolution is the reference's.
#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/clblm_hip.h"

namespace {

constexpr int LPTSMX = 65536;
// Block Data Apod_Init (clblm_FFT.f90:120-146)
constexpr double SC_C[14] = {0.0, 2.0, 0.849322, 2.257609, 3.314800, 2.100669, 2.195676,
                             2.0, 2.570274, 2.367714, 2.071759, -3.3, -2.5, 3.314800};
constexpr double SC_CRATIO[14] = {0.0, 40., 10., 40., 160., 20., 20., 20., 40., 40., 20., 100., 10., 1e300};
constexpr double SC_CLIMIT[14] = {1., 2., 3., 40., 160., 20., 20., 20., 40., 40., 20., 100., 10., 160.};

thread_local std::string g_scan_err;
int fail(int code, const std::string &m) {
    g_scan_err = m;
    return code;
}

// ---- interpolation plan ---------------------------------------------------------
// One segment: output points [out0, out0 + count) computed from input panel s (T(5)
// is input point s) with the fraction restarted at (I0, p0).  kind 1 writes zeros.
struct ISeg {
    int64_t out0;
    int64_t s;
    double p0;
    int32_t count, kind, I0, nni, flags;  // flags: 1 first panel, 2 input exhausted
    int32_t pad;
};

struct IPlan {
    double v1 = 0., dvo = 0., vratio = 0.;
    int64_t nout = 0;
    std::vector<ISeg> segs;
};

// Intpdr (clblm_FFT.f90:2392-2443) + clblm_INTERP (clblm_ScanFilter.f90:1975-2300) +
// panelInput (:2302-2370): the input x[0..n) on v1 + i dv, output from vv1 at dvo up to vv2.
int plan_interp(int64_t n, double v1, double dv, double vv1, double vv2, double dvo, IPlan &ip) {
    if (n < 8) return fail(CLBLM_ERR_ARG, "interpolation needs at least 8 input points");
    if (!(dvo > 0.) || !(dv > 0.)) return fail(CLBLM_ERR_ARG, "grid steps must be positive");
    const double V1C = v1, V2C = v1 + dv * (double)(n - 1);
    double VV1 = vv1, VV2 = vv2;
    if (VV1 < V1C) {
        double AN = (V1C - VV1) / dvo;
        VV1 = (AN == (double)(int)AN) ? VV1 + dvo * (double)(int)AN : VV1 + dvo * (double)((int)AN + 1);
    }
    if (VV2 > V2C) {
        double AN = (VV2 - V2C) / dvo;
        VV2 = (AN == (double)(int)AN) ? VV2 - dvo * (double)(int)AN : VV2 - dvo * (double)((int)AN + 1);
    }
    const double V1 = VV1, V2 = VV2;
    const double VBOT = V1 - 4.0 * dvo;
    int64_t done = 0;
    double done_v2 = v1 - dv;
    // next 2400-point panel reaching VBOT: T(5) = input point *s, T(1) at *vmin
    auto panel = [&](int64_t *s, int *nni, double *vmin, double *vmax, double *dvi) {
        for (;;) {
            int nlim = 2400;
            if (done + nlim > n) nlim = (int)(n - done);
            if (nlim == 0) return false;
            const double d = dv;
            const double a = done_v2 + d;
            const double b = a + (double)(nlim - 1) * d;
            done += nlim;
            done_v2 = b;
            if (b >= VBOT) {
                *s = done - nlim;
                *nni = nlim + 4;
                *vmin = a - (double)(4) * d;
                *vmax = b;
                *dvi = d;
                return true;
            }
        }
    };
    int64_t s = 0;
    int NNI = 0;
    double V1I = 0., V2I = 0., DVI = 0.;
    if (!panel(&s, &NNI, &V1I, &V2I, &DVI)) return fail(CLBLM_ERR_ARG, "interpolation range lies beyond the data");
    if (NNI < 6) return fail(CLBLM_ERR_ARG, "interpolation range starts on the last input point");
    int flags = 1;
    const double DVJ = dvo;
    double V1J = V1, VJ = V1J, V2J;
    const double VRATIO = DVJ / DVI;
    int64_t nout = 0;
    auto seg = [&](int64_t out0, int count, int kind, int I0, double p0) {
        if (count <= 0) return;
        ISeg g{};
        g.out0 = out0; g.count = count; g.kind = kind; g.I0 = I0; g.p0 = p0; g.s = s; g.nni = NNI; g.flags = flags;
        ip.segs.push_back(g);
    };
    while (V1J < V1I) {  // zero fill below the data (label 20)
        int J2 = (int)((V1I - V1J) / DVJ) + 1;
        if (J2 > 2400) J2 = 2400;
        seg(nout, J2, 1, 0, 0.);
        nout += J2;
        V2J = V1J + DVJ * (double)(J2 - 1);
        V1J = V2J + DVJ;
        VJ = V1J;
    }
    int J2 = 0;
    bool eof = false, fin = false;
    for (;;) {  // label 40
        int I = (int)((VJ - V1I) / DVI + 1.00001);
        if (I < 1) return fail(CLBLM_ERR_ARG, "interpolation index below the panel");
        const double VI = V1I + DVI * (double)(I - 1);
        double P = (VJ - VI) / DVI;
        const int J1 = (int)((VJ - V1J) / DVJ + 1.001);
        const int ja = (int)((V2 - V1J) / DVJ + 1.001), jb = (int)((V2I - DVI - V1J) / DVJ + 1.);
        J2 = ja < jb ? ja : jb;
        if (J2 > 2400) J2 = 2400;
        if (J2 >= J1) {
            seg(nout + J1 - 1, J2 - J1 + 1, 0, I, P);
            for (int J = J1; J <= J2; J++) {
                P = P + VRATIO;
                if (P >= 1.0) {
                    I = (int)((double)I + P);
                    P = P - (double)(int)P;
                }
            }
        }
        VJ = V1J + DVJ * (double)J2;
        if (J2 >= 2400 || VJ >= V2) {
            nout += J2;
            V2J = V1J + DVJ * (double)(J2 - 1);
            V1J = V2J + DVJ;
            J2 = 0;
        }
        if (VJ >= V2) { fin = true; break; }
        if (eof) break;
        if (I >= NNI - 2) {
            int64_t s2;
            int n2;
            double a1, a2, a3;
            if (panel(&s2, &n2, &a1, &a2, &a3)) {
                s = s2; NNI = n2; V1I = a1; V2I = a2; DVI = a3;
                flags = 0;
            } else {  // the reference keeps the panel, shifts T(1..4) and extends it by one point
                eof = true;
                flags |= 2;
                V2I = V2I + DVI;
            }
        }
    }
    if (!fin) {
        for (;;) {  // zero fill above the data (label 80)
            const int J1 = J2 + 1;
            J2 = (int)((V2 - V1J) / DVJ + 1.0001);
            if (J2 > 2400) J2 = 2400;
            seg(nout + J1 - 1, J2 - J1 + 1, 1, 0, 0.);
            VJ = V1J + DVJ * (double)J2;
            if (J2 >= 2400 || VJ >= V2) {
                nout += J2;
                V2J = V1J + DVJ * (double)(J2 - 1);
                V1J = V2J + DVJ;
                J2 = 0;
            }
            if (!(VJ < V2)) break;
        }
    }
    ip.v1 = V1;
    ip.dvo = dvo;
    ip.vratio = VRATIO;
    ip.nout = nout;
    return 0;
}

struct ICoef {
    double c1[203], c2[203], c3[203], c4[203];
};

// four-point coefficients for 201 fractions plus one at each end (clblm_ScanFilter.f90:2054-2063)
ICoef make_coef() {
    ICoef c;
    const double XNUMCF = 201.;
    for (int ip = 0; ip <= 202; ip++) {
        double P = ((double)ip - 1.0) / (XNUMCF - 1.0);
        double PP = P * P;
        c.c1[ip] = -(P / 2.0 * ((1 - P) * (1 - P)));
        c.c2[ip] = 1.0 - PP * (3.0 - 2.0 * P) + PP / 2.0 * (1.0 - P);
        c.c3[ip] = PP * (3.0 - 2.0 * P) + P / 2.0 * ((1.0 - P) * (1.0 - P));
        c.c4[ip] = -(PP / 2. * (1.0 - P));
    }
    return c;
}

// T(k) of the segment's panel.  T(5..NNI) are input points s.., T(4) of the first
// panel is extrapolated, and once the input is exhausted T(1..4) hold T(NNI-3..NNI),
// T(NNI+1) = 2 T(NNI) - T(NNI-1) and T(NNI+2) = 0.
template <typename TX>
__device__ __forceinline__ double tval(const TX *x, const ISeg &g, int k) {
    const bool eof = (g.flags & 2) != 0;
    if (eof && k <= 4) return (double)x[g.s + g.nni - 9 + k];
    if ((g.flags & 1) && k <= 4) return k == 4 ? 2.0 * (double)x[g.s] - (double)x[g.s + 1] : 0.0;
    if (eof && k == g.nni + 2) return 0.0;
    if (eof && k == g.nni + 1) {
        const double tn = (double)x[g.s + g.nni - 5];
        const int km = g.nni - 1;
        const double tm = km <= 4 ? (double)x[g.s + g.nni - 9 + km] : (double)x[g.s + km - 5];
        return 2.0 * tn - tm;
    }
    // beyond the panel the reference holds stale values, which it reads only with a
    // zero coefficient (the fraction is 0 at the panel's last usable point)
    if (k > g.nni) return 0.0;
    return (double)x[g.s + k - 5];
}

template <typename TX>
__global__ void k_scan_interp(const ISeg *segs, int64_t nseg, const TX *x, const ICoef *cf, double vratio,
                              double *out) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nseg) return;
    const ISeg g = segs[q];
    double *o = out + g.out0;
    if (g.kind == 1) {
        for (int j = 0; j < g.count; j++) o[j] = 0.0;
        return;
    }
    int I = g.I0;
    double P = g.p0;
    for (int j = 0; j < g.count; j++) {
        const int IP = (int)(P * 201. + 1.00001);
        o[j] = cf->c1[IP] * tval(x, g, I - 1) + cf->c2[IP] * tval(x, g, I) + cf->c3[IP] * tval(x, g, I + 1) +
               cf->c4[IP] * tval(x, g, I + 2);
        P = P + vratio;
        if (P >= 1.0) {
            I = (int)((double)I + P);
            P = P - (double)(int)P;
        }
    }
}

// ---- transform, apodization --------------------------------------------------------
template <typename TB>
__global__ void k_scan_load(const double *xs, int64_t first, int64_t ltotal, double fill, TB *buf, int64_t nfft) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nfft) return;
    buf[i] = (TB)(i < ltotal ? xs[first + i] : fill);
}

// Boxcar (clblm_FFT.f90:867-990): the mean of m consecutive loaded points, summed in
// order, then `fill` to the end of the transform buffer.
template <typename TB>
__global__ void k_scan_boxcar(const double *xs, int64_t first, int64_t ktotal, int m, double fill, TB *buf,
                              int64_t nfft) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nfft) return;
    if (i < ktotal) {
        const double *x = xs + first + i * m;
        double sum = 0.0;
        for (int q = 0; q < m; q++) sum = sum + x[q];
        buf[i] = (TB)(sum / (double)m);
    } else {
        buf[i] = (TB)fill;
    }
}

// Scnfnt's apodization branch (clblm_FFT.f90:1411-1958) for coefficient m of block J:
// X = X0(J) + m' DX with X0 accumulated block by block as the reference does.
__device__ inline double apod(int jfn, double X, double A) {
    const double PI = 2.0 * asin(1.0);
    switch (jfn) {
    case 1: {
        if (X == 0) return 1.0;
        const double s = sin(PI * X * A) / (PI * X * A);
        return s * s;
    }
    case 2: {
        const double t = 2.0 * ((PI * X * A) * (PI * X * A));
        return exp(-(t < 34. ? t : 34.));
    }
    case 3: return (X < 1 / A) ? 1.0 - X * A : 0.0;
    case 4: return (X < 1.0 / A) ? 1.0 : 0.0;
    case 5: return (X < 1 / A) ? (1.0 - (X * A) * (X * A)) * (1.0 - (X * A) * (X * A)) : 0.0;
    case 6: return (X < 1 / A) ? 0.53835685 * (1.0 + 0.857504 * cos(PI * X * A)) : 0.0;
    case 7: return (X < 1. / A) ? 0.5 * (1.0 + cos(PI * X * A)) : 0.0;
    default: {  // Norton-Beer weak, medium, strong (clblm_FFT.f90:1758-1798)
        if (!(X < 1. / A)) return 0.0;
        const double U = 1.0 - (X * A) * (X * A);
        if (jfn == 8) return 0.384093 - 0.087577 * U + 0.703484 * (U * U);
        if (jfn == 9) return 0.152442 - 0.136176 * U + 0.983734 * (U * U);
        return 0.045355 + 0.554883 * (U * U) + 0.399782 * ((U * U) * (U * U));
    }
    }
}

// The boxcar deconvolution factor, Scnfnt's function -1 (clblm_FFT.f90:1920-1946).
__device__ inline double deconv_fac(double X, double A) {
    const double PI = 2.0 * asin(1.0);
    const double Z = 2.0 * PI * X * A;
    return Z != 0 ? Z / sin(Z) : 1.0;
}

// Scnfnt's scanning functions (IAPSC = -1) for JFN 1..7 at X >= 0 (clblm_FFT.f90:1519-1755).
__device__ inline double sfn(int jfn, double X, double A) {
    const double Pi = 2.0 * asin(1.0);
    switch (jfn) {
    case 1: return (X > A) ? 0.0 : (1.0 - X / A);
    case 2: {
        const double t = 0.5 * ((X / A) * (X / A));
        return exp(-(t < 34. ? t : 34.));
    }
    case 3:
        if (X != 0) {
            const double s = sin(Pi * X / A) / (Pi * X / A);
            return s * s;
        }
        return 1.0;
    case 4: return (X != 0) ? sin(2.0 * Pi * X / A) / (2.0 * Pi * X / A) : 1.0;
    case 5: {
        const double U = 2.0 * Pi * X / A;
        if (U >= 0.5) {
            const double U2 = U * U, U5 = U2 * U2 * U;
            return 15. * ((3.0 - U2) * sin(U) - 3.0 * U * cos(U)) / U5;
        }
        return (1.0 - 15.0 / 210 * (U * U));
    }
    default: {
        const double c = jfn == 6 ? 0.428752 : 0.5;
        const double U = 2.0 * Pi * X / A;
        if (U != 0 && fabs(U) != Pi) return (sin(U) / U + c * (sin(U + Pi) / (U + Pi) + sin(U - Pi) / (U - Pi)));
        if (U == 0) return 1.0;
        return c;
    }
    }
}

// Scntrn with IVX = -1 (clblm_FFT.f90:2113-2175): the scanning function on the
// transform grid for an even number of records.  Block J < lrec/2 holds f(V_J + q DV);
// block lrec-1-J holds the mirror images, its first element f(V_{J+1}).
// vblk[J] = V_J, accumulated block by block as the reference does (lrec/2 + 1 values).
template <typename TB>
__global__ void k_scan_sfn(TB *f, int64_t nfft, int64_t lrec, const double *vblk, int jfn, double A, double DV) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nfft) return;
    const int64_t J = i / LPTSMX;
    const int64_t q = i % LPTSMX;
    double X;
    if (J < lrec / 2) {
        X = (q == 0) ? vblk[J] : vblk[J] + (double)q * DV;
    } else {
        const int64_t jp = lrec - J;   // the positive block (1-based) this one mirrors
        X = (q == 0) ? vblk[jp] : vblk[jp - 1] + (double)(LPTSMX - q) * DV;
    }
    f[i] = (TB)sfn(jfn, X, A);
}


struct ApodBlock {
    double x0;     // X at the block's first coefficient
    int32_t nend;  // sinc: coefficient whose value is the interpolated rectangle end (-1: none)
    int32_t pad;
    double vend;
};

// Multrn (clblm_FFT.f90:1344-1405) with the analytic apodization of JFN (none for
// JFN 0) and, for deconv_a > 0, the boxcar deconvolution factor; `scale` is the
// inverse transform's 1/N.
template <typename TC, typename TS>
__global__ void k_scan_apod(TC *c, int64_t nfft, const ApodBlock *blk, int jfn, double A, double DX,
                            double deconv_a, double scale) {
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t half = nfft / 2;
    if (m > half) return;
    double h;
    if (m == half) {
        h = 0.0;  // Nyquist: multiplied by the imaginary slot of the first coefficient (0)
    } else {
        const int64_t J = m / (LPTSMX / 2);
        const int kb = (int)(m % (LPTSMX / 2));
        const ApodBlock b = blk[J];
        const double X = b.x0 + (double)kb * DX;
        h = jfn == 0 ? 1.0 : ((kb == b.nend) ? b.vend : apod(jfn, X, A));
        if (deconv_a > 0.) h = h * deconv_fac(X, deconv_a);
    }
    const double f = h * scale;
    c[m].x = (TS)((double)c[m].x * f);
    c[m].y = (TS)((double)c[m].y * f);
}

// Multrn with the transform of the scanning function (IVX = -1), divided by its first
// coefficient (the area, Scntrn :2180-2197); DC and Nyquist are products of reals.
template <typename TC, typename TS>
__global__ void k_scan_cmul(TC *c, const TC *hc, int64_t nfft, const ApodBlock *blk, double DX, double deconv_a,
                            double scale) {
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t half = nfft / 2;
    if (m > half) return;
    const double h0 = (double)hc[0].x;
    const double hr = (double)hc[m].x / h0, hi = (double)hc[m].y / h0;
    const double xr = (double)c[m].x, xi = (double)c[m].y;
    double yr, yi;
    if (m == 0 || m == half) {
        yr = xr * hr;
        yi = 0.0;
    } else {
        yr = xr * hr - xi * hi;
        yi = xr * hi + xi * hr;
    }
    if (deconv_a > 0.) {
        double d = 0.0;
        if (m < half) {
            const ApodBlock b = blk[m / (LPTSMX / 2)];
            d = deconv_fac(b.x0 + (double)(m % (LPTSMX / 2)) * DX, deconv_a);
        }
        yr = yr * d;
        yi = yi * d;
    }
    c[m].x = (TS)(yr * scale);
    c[m].y = (TS)(yi * scale);
}

template <typename TB>
__global__ void k_scan_window(const TB *y, int64_t first, int64_t n, double *out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (double)y[first + i];
}

// ---- per-context state -------------------------------------------------------------
struct ScanState {
    int64_t plan_n = 0;
    int plan_prec = -1;
    hipfftHandle fwd = 0, inv = 0;
    void *buf = nullptr;           // nfft + 2 reals (in-place transforms)
    size_t buf_cap = 0;
    void *hbuf = nullptr;          // the scanning function and its transform (IVX = -1)
    size_t hbuf_cap = 0;
    double *xs = nullptr;          // re-gridded input
    size_t xs_cap = 0;
    double *win = nullptr;         // scan window before the output interpolation
    size_t win_cap = 0;
    void *meta = nullptr;          // segments, apodization blocks, coefficients
    size_t meta_cap = 0;
    void *conv = nullptr;          // spectral-domain scan: tables and per-box arrays
    size_t conv_cap = 0;
    void *filt = nullptr;          // SRF filter: channels, filter values, panel bounds
    size_t filt_cap = 0;
    ~ScanState() {
        if (fwd) hipfftDestroy(fwd);
        if (inv) hipfftDestroy(inv);
        (void)hipFree(buf);
        (void)hipFree(hbuf);
        (void)hipFree(xs);
        (void)hipFree(win);
        (void)hipFree(meta);
        (void)hipFree(conv);
        (void)hipFree(filt);
    }
};

int grow(void **p, size_t *cap, size_t bytes) {
    if (bytes <= *cap) return 0;
    (void)hipFree(*p);
    *p = nullptr;
    *cap = 0;
    if (hipMalloc(p, bytes) != hipSuccess) return fail(CLBLM_ERR_HIP, "scan buffer allocation failed");
    *cap = bytes;
    return 0;
}

// The scan's host-side plan.
struct ScanPlan {
    int jfn = 0;
    double a = 0.;
    bool pre = false;
    double dvn = 0.;
    IPlan pre_ip;          // re-gridding of the input (DV > 2 HWHM / 12)
    double v1p = 0.;       // scan-grid spectrum: start, step, points
    double dv = 0.;
    int64_t np = 0;
    double v1z = 0., v2z = 0.;
    int64_t ltotal = 0, n1load = 0, lrec = 0, nfft = 0;
    int ivx = 0;           // 1 analytic apodization, -1 transform of the scanning function, 0 none
    bool fft = true;       // false: boxcar scan without deconvolution, no transform
    int m = 0;             // boxcar pre-scan over m points (0: none)
    int64_t ktotal = 0;    // boxcar means
    double deconv_a = 0.;  // > 0: divide by the boxcar's transform of this width parameter
    int64_t n1 = 0, ntotal = 0;
    double v1s = 0.;
    bool out_interp = false;
    IPlan out_ip;
    std::vector<ApodBlock> blocks;
    std::vector<double> vblk;   // IVX = -1: start of each positive block of the scanning function
};

int64_t scan_records(int64_t ntot) {  // LoadSpect/Boxcar record count (clblm_FFT.f90:1189-1197, 890-910)
    const double ARDATA = (double)ntot / (double)LPTSMX;
    const int64_t LRDATA = (ARDATA - (double)(int)ARDATA == 0.0) ? (int64_t)(int)ARDATA : (int64_t)(int)(ARDATA + 1);
    const double POWER = std::log((double)LRDATA) / std::log(2.0);
    return (POWER - (double)(int)POWER == 0) ? ((int64_t)1 << (int)(POWER + 2)) : ((int64_t)1 << ((int)POWER + 3));
}

int plan_scan(int64_t n, double v1, double dv, double v1out, double v2out, double dvout, int funct_id, double hwhm,
              double fill, double pre_box_hw, int deconv_pre_box, ScanPlan &p) {
    if (!(fill >= 0. && fill <= 1.)) return fail(CLBLM_ERR_ARG, "fill must lie in [0, 1]");
    if (!((funct_id >= 1 && funct_id <= 14) || (funct_id >= 102 && funct_id <= 114) || funct_id == -1))
        return fail(CLBLM_ERR_ARG, "funct_id must be in [1, 14], [102, 114] or -1");
    if (!(hwhm > 0.) || !(dv > 0.) || n < 8) return fail(CLBLM_ERR_ARG, "hwhm, dv must be positive and n >= 8");
    // the function and the boxcar/deconvolution switches (clblm_FFT.f90:447-474)
    int JFN = funct_id - 1;
    bool prebox = false, deconv = false, deconv_only = false;
    double boxhw = 0., dec_a = 0.;
    if (JFN == 0) {
        prebox = true;
        boxhw = hwhm;
    } else if (JFN >= 1 && JFN <= 13) {
        if (pre_box_hw > 0.) {
            prebox = true;
            deconv = deconv_pre_box != 0;
            boxhw = pre_box_hw;
        }
    } else if (JFN >= 101 && JFN <= 113) {
        JFN -= 100;
        deconv = true;
        dec_a = pre_box_hw;
    } else {  // -2: deconvolution only
        deconv_only = true;
        JFN = 0;
        deconv = true;
        dec_a = pre_box_hw;
    }
    if (JFN == 11 || JFN == 12)
        return fail(CLBLM_ERR_ARG, "the Brault and Kaiser-Bessel functions need the interferometer optical path "
                                   "difference, which clblm_FFTSCN does not take");
    if (JFN == 13)
        return fail(CLBLM_ERR_UNSUPPORTED, "the Kiruna scanning function is not supported");
    if (deconv && !(dec_a > 0.) && !prebox) return fail(CLBLM_ERR_ARG, "deconvolution needs pre_box_hw > 0");
    p.jfn = JFN;
    p.a = JFN > 0 ? SC_C[JFN] * hwhm : 0.;
    const double V1C = v1, V2C = v1 + dv * (double)(n - 1);
    if (v1out >= v2out) return fail(CLBLM_ERR_ARG, "v1out >= v2out");
    if (v1out >= V2C || v2out <= V1C) return fail(CLBLM_ERR_ARG, "v1out..v2out lies outside the data");
    double V1Z, V2Z;
    if (deconv_only) {
        V1Z = V1C;
        V2Z = V2C;
    } else {
        V1Z = v1out - hwhm * SC_CLIMIT[JFN];
        V2Z = v2out + hwhm * SC_CLIMIT[JFN];
        V1Z = V1C + (double)(int)((V1Z - V1C) / dv) * dv;
        V2Z = V1C + (double)(int)((V2Z - V1C) / dv + 1.01) * dv;
    }
    if (V1Z < V1C) V1Z = V1C;
    if (V2Z > V2C) V2Z = V2C;
    if ((V2Z - V1Z) / hwhm < SC_CLIMIT[JFN]) return fail(CLBLM_ERR_ARG, "insufficient data for an accurate scan");
    const double DVN = 2. * hwhm / 12;
    p.dvn = DVN;
    double DV = dv;
    if (DV > DVN && JFN != 0) {  // Intpdr onto DVN first (clblm_FFT.f90:594-614)
        p.pre = true;
        int rc = plan_interp(n, v1, dv, V1Z, V2Z, DVN, p.pre_ip);
        if (rc) return rc;
        DV = DVN;
        p.v1p = p.pre_ip.v1;
        p.np = p.pre_ip.nout;
        V1Z = p.v1p;
        V2Z = p.v1p + DV * (double)(p.np - 1);
    } else {
        p.v1p = v1;
        p.np = n;
    }
    // IVX (clblm_FFT.f90:616-634)
    p.ivx = JFN == 0 ? 0 : (((V2Z - V1Z) / hwhm > SC_CRATIO[JFN]) ? 1 : -1);
    if (p.ivx == -1 && JFN >= 8)
        return fail(CLBLM_ERR_ARG, "Norton-Beer functions have no spectral-domain form: the scan range must exceed "
                                   "CRATIO half widths");
    if (p.v1p > V1Z) return fail(CLBLM_ERR_ARG, "first data frequency above the scan start");
    // LoadSpect (clblm_FFT.f90:1166-1206)
    int64_t LTOTAL = (int64_t)(int)((V2Z - V1Z) / DV + 1.1);
    V2Z = V1Z + DV * (double)(LTOTAL - 1);
    int64_t LREC = scan_records(LTOTAL);
    const int64_t N1L = (int64_t)(int)((V1Z - p.v1p) / DV + 1.00001);
    if (N1L < 1 || N1L + LTOTAL - 1 > p.np) return fail(CLBLM_ERR_ARG, "scan window runs out of data");
    p.n1load = N1L;
    // boxcar pre-scan (clblm_FFT.f90:689-706)
    if (prebox) {
        const int M = (int)(2.0 * boxhw / DV);
        if (M >= 2) {
            V1Z = V1Z + DV * (double)(M - 1) / 2.0;
            const int64_t KTOTAL = LTOTAL / M;
            if (KTOTAL < 1) return fail(CLBLM_ERR_ARG, "boxcar wider than the scan range");
            DV = (double)M * DV;
            V2Z = V1Z + DV * (double)(KTOTAL - 1);
            LREC = scan_records(KTOTAL);
            LTOTAL = KTOTAL;
            p.m = M;
            p.ktotal = KTOTAL;
            dec_a = DV;
        } else {
            deconv = false;
        }
    }
    p.deconv_a = deconv ? dec_a : 0.;
    p.fft = !(JFN == 0 && !deconv);
    p.dv = DV;
    p.v1z = V1Z;
    p.v2z = V2Z;
    p.ltotal = LTOTAL;
    p.lrec = LREC;
    p.nfft = LREC * LPTSMX;
    // wrtSpect window (clblm_FFT.f90:776-796)
    int64_t N1 = (int64_t)(int)((v1out - V1Z) / DV);
    double V1S = V1Z + (double)(N1 - 1) * DV;
    if (V1S < V1Z) { N1 = 1; V1S = V1Z; }
    int64_t N2 = (int64_t)(int)((v2out - V1Z) / DV + 3.01);
    const double V2S = V1Z + (double)(N2 - 1) * DV;
    if (V2S > V2Z) N2 = (int64_t)(int)((V2Z - V1Z) / DV + 1.01);
    p.n1 = N1;
    p.ntotal = N2 - N1 + 1;
    p.v1s = V1S;
    if (p.n1 - 1 + p.ntotal > p.nfft || p.ntotal < 1) return fail(CLBLM_ERR_ARG, "scan window outside the transform");
    if (dvout > 0.) {
        p.out_interp = true;
        int rc = plan_interp(p.ntotal, V1S, DV, v1out, v2out, dvout, p.out_ip);
        if (rc) return rc;
    }
    // Scntrn, IVX = 1 (clblm_FFT.f90:2217-2229): X per block, sinc end point per block;
    // the same positions serve the deconvolution factor
    const double DX = 1.0 / ((double)(LPTSMX * LREC) * DV);
    p.blocks.resize((size_t)LREC);
    double X = 0.0;
    for (int64_t J = 0; J < LREC; J++) {
        ApodBlock &b = p.blocks[(size_t)J];
        b.x0 = X;
        b.nend = -1;
        b.vend = 0.;
        if (JFN == 4 && p.ivx == 1) {
            const double ALEN = 1.0 / p.a, XSTART = X;
            int N = (int)((ALEN - XSTART) / DX) + 1;
            double R = (ALEN - XSTART) / DX + 1 - N;
            if (R <= 0.5) { R = R + 1.0; N = N - 1; }
            if (N >= 1 && N <= LPTSMX / 2 - 1) { b.nend = N; b.vend = R - 0.5; }
        }
        X = X + (double)(LPTSMX) * DX / 2.;
    }
    if (p.ivx == -1) {  // Scntrn, IVX = -1: block starts V_J (clblm_FFT.f90:2113-2134)
        p.vblk.resize((size_t)(LREC / 2 + 1));
        double V = 0.0;
        for (int64_t J = 0; J <= LREC / 2; J++) {
            p.vblk[(size_t)J] = V;
            V = V + (double)LPTSMX * DV;
        }
    }
    return 0;
}

unsigned cdiv(int64_t a, int64_t b) { return (unsigned)((a + b - 1) / b); }

int ensure_plans(ScanState &S, int64_t nfft, int prec, hipStream_t st) {
    if (S.plan_n == nfft && S.plan_prec == prec) {
        if (hipfftSetStream(S.fwd, st) != HIPFFT_SUCCESS || hipfftSetStream(S.inv, st) != HIPFFT_SUCCESS)
            return fail(CLBLM_ERR_HIP, "hipfftSetStream failed");
        return 0;
    }
    if (S.fwd) hipfftDestroy(S.fwd);
    if (S.inv) hipfftDestroy(S.inv);
    S.fwd = S.inv = 0;
    S.plan_n = 0;
    const bool d = prec == CLBLM_FP64;
    if (hipfftPlan1d(&S.fwd, (int)nfft, d ? HIPFFT_D2Z : HIPFFT_R2C, 1) != HIPFFT_SUCCESS ||
        hipfftPlan1d(&S.inv, (int)nfft, d ? HIPFFT_Z2D : HIPFFT_C2R, 1) != HIPFFT_SUCCESS)
        return fail(CLBLM_ERR_HIP, "hipFFT plan creation failed");
    if (hipfftSetStream(S.fwd, st) != HIPFFT_SUCCESS || hipfftSetStream(S.inv, st) != HIPFFT_SUCCESS)
        return fail(CLBLM_ERR_HIP, "hipfftSetStream failed");
    S.plan_n = nfft;
    S.plan_prec = prec;
    return 0;
}

#define SCHK(x)                                                                        \
    do {                                                                               \
        hipError_t e_ = (x);                                                           \
        if (e_ != hipSuccess) return fail(CLBLM_ERR_HIP, std::string(#x ": ") + hipGetErrorString(e_)); \
    } while (0)

// The transform buffer: the loaded spectrum or its boxcar means, padded with `fill`.
template <typename TB>
void fill_buffer(const ScanPlan &p, const double *xs, double fill, TB *buf, hipStream_t st) {
    if (p.m >= 2)
        hipLaunchKernelGGL(k_scan_boxcar<TB>, dim3(cdiv(p.nfft, 256)), dim3(256), 0, st, xs, p.n1load - 1, p.ktotal,
                           p.m, fill, buf, p.nfft);
    else
        hipLaunchKernelGGL(k_scan_load<TB>, dim3(cdiv(p.nfft, 256)), dim3(256), 0, st, xs, p.n1load - 1, p.ltotal,
                           fill, buf, p.nfft);
}

template <typename TB, typename TC>
int run_transform(ScanState &S, const ScanPlan &p, const double *xs, double fill, const ApodBlock *d_blk,
                  const double *d_vblk, hipStream_t st) {
    TB *buf = (TB *)S.buf;
    fill_buffer<TB>(p, xs, fill, buf, st);
    auto fwd = [&](TB *x) {
        hipfftResult r;
        if constexpr (sizeof(TB) == 8)
            r = hipfftExecD2Z(S.fwd, (hipfftDoubleReal *)x, (hipfftDoubleComplex *)x);
        else
            r = hipfftExecR2C(S.fwd, (hipfftReal *)x, (hipfftComplex *)x);
        return r == HIPFFT_SUCCESS ? 0 : fail(CLBLM_ERR_HIP, "forward FFT failed");
    };
    int rc = fwd(buf);
    if (rc) return rc;
    const double DX = 1.0 / ((double)(LPTSMX * p.lrec) * p.dv);
    const unsigned g = cdiv(p.nfft / 2 + 1, 256);
    if (p.ivx == -1) {
        TB *h = (TB *)S.hbuf;
        hipLaunchKernelGGL(k_scan_sfn<TB>, dim3(cdiv(p.nfft, 256)), dim3(256), 0, st, h, p.nfft, p.lrec, d_vblk,
                           p.jfn, p.a, p.dv);
        rc = fwd(h);
        if (rc) return rc;
        hipLaunchKernelGGL((k_scan_cmul<TC, TB>), dim3(g), dim3(256), 0, st, (TC *)buf, (const TC *)h, p.nfft, d_blk,
                           DX, p.deconv_a, 1.0 / (double)p.nfft);
    } else {
        hipLaunchKernelGGL((k_scan_apod<TC, TB>), dim3(g), dim3(256), 0, st, (TC *)buf, p.nfft, d_blk, p.jfn, p.a,
                           DX, p.deconv_a, 1.0 / (double)p.nfft);
    }
    hipfftResult r;
    if constexpr (sizeof(TB) == 8)
        r = hipfftExecZ2D(S.inv, (hipfftDoubleComplex *)buf, (hipfftDoubleReal *)buf);
    else
        r = hipfftExecC2R(S.inv, (hipfftComplex *)buf, (hipfftReal *)buf);
    if (r != HIPFFT_SUCCESS) return fail(CLBLM_ERR_HIP, "inverse FFT failed");
    return 0;
}

}  // namespace

extern "C" __attribute__((visibility("hidden"))) void *clblm_scan_state_new() { return new ScanState; }
extern "C" __attribute__((visibility("hidden"))) void clblm_scan_state_free(void *s) { delete (ScanState *)s; }
extern "C" __attribute__((visibility("hidden"))) const char *clblm_scan_error() { return g_scan_err.c_str(); }

extern "C" __attribute__((visibility("hidden"))) int clblm_fft_scan_impl(
    void *state, const double *d_in, int64_t n, double v1, double dv, double v1out, double v2out, double dvout,
    int funct_id, double hwhm, double fill, double pre_box_hw, int deconv_pre_box, int precision, double *d_out,
    int64_t *nout, double *v1res, double *dvres, void *stream) {
    if (precision != CLBLM_FP64 && precision != CLBLM_FP32C) return fail(CLBLM_ERR_ARG, "unknown precision");
    ScanPlan p;
    int rc = plan_scan(n, v1, dv, v1out, v2out, dvout, funct_id, hwhm, fill, pre_box_hw, deconv_pre_box, p);
    if (rc) return rc;
    if (p.out_interp) {
        *nout = p.out_ip.nout;
        *v1res = p.out_ip.v1;
        *dvres = dvout;
    } else {
        *nout = p.ntotal;
        *v1res = p.v1s;
        *dvres = p.dv;
    }
    if (!d_out) return 0;
    if (!d_in) return fail(CLBLM_ERR_ARG, "input spectrum is NULL");
    ScanState &S = *(ScanState *)state;
    hipStream_t st = (hipStream_t)stream;
    // metadata: coefficients, apodization blocks, block starts, both segment lists
    auto up = [](size_t o) { return (o + 255) & ~(size_t)255; };
    const size_t o_cf = 0;
    const size_t o_blk = up(sizeof(ICoef));
    const size_t o_vb = up(o_blk + sizeof(ApodBlock) * p.blocks.size());
    const size_t o_s1 = up(o_vb + sizeof(double) * p.vblk.size());
    const size_t o_s2 = up(o_s1 + sizeof(ISeg) * p.pre_ip.segs.size());
    const size_t meta = o_s2 + sizeof(ISeg) * p.out_ip.segs.size() + 256;
    // the metadata arrays are read by the kernels below: finish the previous scan first
    SCHK(hipStreamSynchronize(st));
    rc = grow(&S.meta, &S.meta_cap, meta);
    if (rc) return rc;
    std::vector<char> host(meta, 0);
    const ICoef cf = make_coef();
    std::memcpy(host.data() + o_cf, &cf, sizeof cf);
    std::memcpy(host.data() + o_blk, p.blocks.data(), sizeof(ApodBlock) * p.blocks.size());
    if (!p.vblk.empty()) std::memcpy(host.data() + o_vb, p.vblk.data(), sizeof(double) * p.vblk.size());
    if (!p.pre_ip.segs.empty()) std::memcpy(host.data() + o_s1, p.pre_ip.segs.data(), sizeof(ISeg) * p.pre_ip.segs.size());
    if (!p.out_ip.segs.empty()) std::memcpy(host.data() + o_s2, p.out_ip.segs.data(), sizeof(ISeg) * p.out_ip.segs.size());
    SCHK(hipMemcpyAsync(S.meta, host.data(), meta, hipMemcpyHostToDevice, st));
    SCHK(hipStreamSynchronize(st));  // the host copy of the metadata goes out of scope
    char *mb = (char *)S.meta;
    const ICoef *d_cf = (const ICoef *)(mb + o_cf);
    const double *xs = d_in;
    if (p.pre) {
        rc = grow((void **)&S.xs, &S.xs_cap, sizeof(double) * (size_t)p.np);
        if (rc) return rc;
        hipLaunchKernelGGL(k_scan_interp<double>, dim3(cdiv((int64_t)p.pre_ip.segs.size(), 64)), dim3(64), 0, st,
                           (const ISeg *)(mb + o_s1), (int64_t)p.pre_ip.segs.size(), d_in, d_cf, p.pre_ip.vratio,
                           S.xs);
        xs = S.xs;
    }
    // without a transform (boxcar scan) the buffer is FP64 in both precisions
    const bool d64 = precision == CLBLM_FP64 || !p.fft;
    rc = grow(&S.buf, &S.buf_cap, (d64 ? 8 : 4) * (size_t)(p.nfft + 2));
    if (rc) return rc;
    if (p.fft) {
        if (p.ivx == -1) {
            rc = grow(&S.hbuf, &S.hbuf_cap, (d64 ? 8 : 4) * (size_t)(p.nfft + 2));
            if (rc) return rc;
        }
        rc = ensure_plans(S, p.nfft, d64 ? CLBLM_FP64 : CLBLM_FP32C, st);
        if (rc) return rc;
        const ApodBlock *d_blk = (const ApodBlock *)(mb + o_blk);
        const double *d_vb = (const double *)(mb + o_vb);
        rc = d64 ? run_transform<double, hipfftDoubleComplex>(S, p, xs, fill, d_blk, d_vb, st)
                 : run_transform<float, hipfftComplex>(S, p, xs, fill, d_blk, d_vb, st);
        if (rc) return rc;
    } else {
        fill_buffer<double>(p, xs, fill, (double *)S.buf, st);
    }
    if (!p.out_interp) {
        if (d64)
            hipLaunchKernelGGL(k_scan_window<double>, dim3(cdiv(p.ntotal, 256)), dim3(256), 0, st,
                               (const double *)S.buf, p.n1 - 1, p.ntotal, d_out);
        else
            hipLaunchKernelGGL(k_scan_window<float>, dim3(cdiv(p.ntotal, 256)), dim3(256), 0, st,
                               (const float *)S.buf, p.n1 - 1, p.ntotal, d_out);
    } else {
        const ISeg *segs = (const ISeg *)(mb + o_s2);
        const int64_t ns = (int64_t)p.out_ip.segs.size();
        if (d64)
            hipLaunchKernelGGL(k_scan_interp<double>, dim3(cdiv(ns, 64)), dim3(64), 0, st, segs, ns,
                               (const double *)S.buf + (p.n1 - 1), d_cf, p.out_ip.vratio, d_out);
        else
            hipLaunchKernelGGL(k_scan_interp<float>, dim3(cdiv(ns, 64)), dim3(64), 0, st, segs, ns,
                               (const float *)S.buf + (p.n1 - 1), d_cf, p.out_ip.vratio, d_out);
    }
    SCHK(hipGetLastError());
    return 0;
}

// ==================================================================================
// Spectral-domain rectangle scan: clblm_SCANFN_struct (clblm_ScanFilter.f90:70-544)
// for functID 1 (JFN 0): adjacent boxes of width 2 HWHM every DVO, each the weighted
// mean of the points it covers (CNVRCT :1478-1698), the input read in 2400-point
// panels (panelInput :2320-2404, NREN = 0) and the output written in 2400-point
// panels (panelOutput :2411-2506).  It is the boxcar pre-average output_JacArray
// applies to Jacobian arrays (clblm_noScattJac.f90:1341-1409).
//
//   host    CNVRCT's control flow depends on the grids only: it is walked here, and
//           each box becomes its parts (a box split by an input panel edge has two),
//           each part the input range it sums and its two end weights, in the
//           reference's order of additions
//   device  one lane per box: the sum in that order, divided by the weight sum
// ==================================================================================
namespace {

struct RPart {
    int64_t il, ih;    // global input indices of the part's end points
    double wl, wr;     // their weights; the points between have weight 1
};
struct RBox {
    int64_t out;       // output index
    int32_t p0, np;    // its parts
    double rnj;        // sum of the weights
};

struct RectPlan {
    bool scan = true;          // false: IRATIO < 2, the input is the output
    double v1 = 0., dvo = 0.;
    int64_t nout = 0;
    std::vector<RPart> parts;
    std::vector<RBox> boxes;
};

// The reference's walk with the data left out: panelInput's accumulated panel
// frequencies, CNVRCT's boxes and partial sums, panelOutput's panels.
void plan_rect(int64_t ns, double sv1, double sdv, double V1, double V2, double dvout, double HWHM, RectPlan &rp) {
    const double SAMPLE = HWHM / std::fabs(dvout);
    double DVO = HWHM / SAMPLE;
    const int IRATIO = (int)(DVO / sdv + 0.5);
    DVO = (double)IRATIO * sdv;
    if (IRATIO < 2) {
        rp.scan = false;
        rp.v1 = sv1;
        rp.dvo = sdv;
        rp.nout = ns;
        return;
    }
    const int NLIMF = 2401, NSHIFT = 32;
    const int NBOUND = (int)(2. * SAMPLE + 0.01);
    const double BOUND = (double)NBOUND * DVO / 2.;
    int NLO = NSHIFT + 1, NHI = NLIMF + NSHIFT - 1;
    int IEOFSC = 1, IDATA = -1, IPANEL = -1, JFLG = -1, NNI = 0;
    double VFT = V1 - (double)NSHIFT * DVO;
    const double VBOT = V1 - BOUND, VTOP = V2 + BOUND;
    int64_t done = 0, panel0 = 0, nout = 0;
    double done_v2 = sv1 - sdv, V1I = 0., V2I = 0., DVI = sdv;
    // CNVRCT's SAVE state; the current box's parts and weight sum
    int JN = 0, NB = 1, IPC = 1;
    double RNJ = 0.;
    std::vector<RPart> cur;
    int64_t panel_out0 = 0;    // output index of R1(NLO) of the current output panel
    auto cnvrct = [&]() {
        if (IPANEL == -1) {
            cur.clear();
            RNJ = 0.;
            JN = NLO;
            if (JFLG == -1) { NB = 1; IPC = 1; }
        }
        const int NNIV2 = (int)((V2I - V1I) / DVI + 1.0001);
        if (NNI > NNIV2) NNI = NNIV2;
        while (NLO <= NHI) {
            const double VCNT = V1 + (double)(NB - 1) * DVO;
            if (VCNT > V2) { IPANEL = 1; return; }
            const double VLFT = VCNT - HWHM, VRGT = VCNT + HWHM;
            const double RL = (VLFT - V1I) / DVI + 1, RR = (VRGT - V1I) / DVI + 1;
            int IL = (int)(RL + 0.5), IH = (int)(RR + 0.5);
            const double VLBLR = (V1I + (double)(IL - 1) * DVI) + DVI / 2.;
            double WGTL = (VLBLR - VLFT) / DVI;
            const double VLBLL = (V1I + (double)(IH - 1) * DVI) - DVI / 2.;
            double WGTR = (VRGT - VLBLL) / DVI;
            int NEP = 2;
            if (IH > NNI) {
                IH = NNI;
                JFLG = 1;
                WGTR = 1.;
                if (IL == IH) { WGTR = 0.; NEP = 1; }
            }
            if (IL <= 1) {
                IL = 1;
                WGTL = 1.;
                if (IL >= IH) { WGTL = 0.; NEP = 1; }
            }
            if (JFLG == 1 && IDATA == 0 && IL > NNI) { IPANEL = 0; JFLG = 0; return; }
            if (JFLG == 1 && IDATA == 1) return;
            cur.push_back(RPart{panel0 + IL - 1, panel0 + IH - 1, WGTL, WGTR});
            RNJ = RNJ + (double)(IH - IL + 1 - NEP) + WGTL + WGTR;
            if (JFLG == 1 && IDATA == 0) { IPANEL = 0; JFLG = 0; return; }
            RBox b;
            b.out = panel_out0 + (JN - (NSHIFT + 1));
            b.p0 = (int32_t)rp.parts.size();
            b.np = (int32_t)cur.size();
            b.rnj = RNJ;
            rp.parts.insert(rp.parts.end(), cur.begin(), cur.end());
            rp.boxes.push_back(b);
            JN = JN + 1;
            cur.clear();
            RNJ = 0.;
            if (NB == IPC * (NHI - NLO + 1)) {
                IPANEL = 1;
                IPC = IPC + 1;
                NB = NB + 1;
                return;
            }
            NB = NB + 1;
        }
    };
    int state = 40;
    for (;;) {
        if (state == 40) {
            if (IEOFSC <= 0) { state = 60; continue; }
            int nlim;
            for (;;) {
                nlim = 2400;
                if (done + nlim > ns) nlim = (int)(ns - done);
                if (nlim == 0) { IEOFSC = 0; break; }
                DVI = sdv;
                V1I = done_v2 + DVI;
                V2I = V1I + (double)(nlim - 1) * DVI;
                done += nlim;
                done_v2 = V2I;
                if (V2I >= VBOT) break;
            }
            if (IEOFSC <= 0) { state = 60; continue; }
            NNI = nlim;
            panel0 = done - nlim;
            IDATA = V2I > VTOP ? 1 : 0;
            state = 50;
        }
        if (state == 50) {
            cnvrct();
            if (IPANEL == 0) { state = 40; continue; }
        }
        // panelOutput
        int ISTOP = 0;
        const int NNHI = (int)((V2 - VFT) / DVO + 1.5);
        if (NHI >= NNHI) ISTOP = 1;
        if (ISTOP == 1) NHI = NNHI;
        const int64_t cnt = NHI >= NLO ? NHI - NLO + 1 : 0;
        nout += cnt;
        panel_out0 += cnt;
        VFT = VFT + (double)(NLIMF - 1) * DVO;
        if (ISTOP != 1) NLO = NSHIFT + 1;
        IPANEL = -1;
        if (ISTOP != 1 && IEOFSC < 0) { state = 60; continue; }
        if (ISTOP != 1 && IEOFSC > 0) { state = 50; continue; }
        break;
    }
    // boxes past the last output panel's end are never written out
    std::vector<RBox> kept;
    kept.reserve(rp.boxes.size());
    for (const RBox &b : rp.boxes)
        if (b.out >= 0 && b.out < nout) kept.push_back(b);
    rp.boxes.swap(kept);
    rp.v1 = V1;
    rp.dvo = DVO;
    rp.nout = nout;
}

// row blockIdx.y of the arrays: input at x + y ldx, output at out + y ldo
__global__ void k_rect_fill(double *out, int64_t n, int64_t ldo) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[blockIdx.y * ldo + i] = 0.;
}

__global__ void k_rect_boxes(const RBox *boxes, int64_t nbox, const RPart *parts, const double *x, int64_t ldx,
                             double *out, int64_t ldo) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nbox) return;
    x += blockIdx.y * ldx;
    out += blockIdx.y * ldo;
    const RBox b = boxes[q];
    double sum = 0.;
    for (int k = 0; k < b.np; k++) {
        const RPart p = parts[b.p0 + k];
        for (int64_t i = p.il + 1; i <= p.ih - 1; i++) sum = sum + x[i];
        sum = sum + x[p.il] * p.wl;
        sum = sum + x[p.ih] * p.wr;
    }
    out[b.out] = sum / b.rnj;
}

__global__ void k_rect_copy(const double *x, int64_t ldx, int64_t n, double *out, int64_t ldo) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[blockIdx.y * ldo + i] = x[blockIdx.y * ldx + i];
}

// ==================================================================================
// Spectral-domain scan with a tabulated scanning function: clblm_SCANFN_struct
// (clblm_ScanFilter.f90:70-544) for functIDs 2..5 (JFN 1..4: triangle, Gaussian,
// sinc**2, sinc).  The function is tabulated every DXF half widths (SHAPET, SHAPEG,
// SINCSQ, SINC :551-723), the input compressed into boxes of IRATSH points (SHRKSC
// :732-850), and CONVSC (:856-955) adds SI * XF(IT) into R1(JMIN..JMAX) for each box,
// IT = INT(|ZF| + 1.5) with ZF accumulated step by step from JMIN.  R1 is written and
// shifted by 2400 points whenever a box reaches past MAXF (panelOutput :2411-2506).
//
//   host    the walk over input panels (panelInput :2320-2404 with SHRKSC's carried
//           remainder), the output panels' VFT (accumulated as the reference does),
//           and the table, evaluated with the C library's sin/exp as the reference's
//           build does
//   device  one lane per box: its sum, its output panel k (the first whose JMAX fits
//           MAXF: boxes only move forward, so this is the panel the reference is on
//           when it reaches the box), JMIN, JMAX and ZF(JMIN); then one wave per 64
//           outputs, walking the boxes that reach them in the reference's order: each
//           lane replays one box's ZF from JMIN to the tile (the same additions), the
//           table indices go through LDS, and each lane adds its output's terms box by
//           box, so every output is the reference's sum, bit for bit
// Output G = 33 + o of the global R1 index G = JF + 2400 k; a term at JF < 33 falls on a
// position already written (or before V1) and is dropped, as in the reference.
// The reference's output count depends on the data in one case: when the input ends
// before V2 + bound, the last panel written is the one of the last nonzero box (CONVSC
// skips zero boxes, :906); that case reads one integer per row back.
// ==================================================================================
struct ScanFnTab { double hwf, dxf; int nf, nfmax; double smpl, xscal; };
constexpr ScanFnTab SCAN_FN[5] = {   // HWJ, DXJ, NJ, NJMX, SMPLJ, XSCAL (clblm_ScanFilter.f90:182-207)
    {1., 0.0, 0, 0, .5, 0.},
    {2., 0.02, 101, 251, 2., 0.},
    {4., 0.02, 201, 251, 4., 0.},
    {54.1826, 0.02, 2710, 2760, 4., 1.391557377},
    {119.332818, 0.02, 5968, 6018, 4., 1.89549425},
};
constexpr int CONV_NLIMF = 2401, CONV_NSHIFT = 32, CONV_R1 = 5000, CONV_TILE = 64;
constexpr int CONV_DEAD = 0x3fffffff;   // G of a box past the last output panel

// xf[0..nfmax], xf[0] unused: SHAPET, SHAPEG, SINCSQ, SINC (clblm_ScanFilter.f90:551-723)
void scan_shape(int jfn, std::vector<double> &xf) {
    const ScanFnTab &f = SCAN_FN[jfn];
    const double PI = 2. * std::asin(1.);
    xf.assign((size_t)f.nfmax + 1, 0.);
    if (jfn == 1) {
        xf[1] = 0.5;
        for (int i = 2; i <= f.nf; i++) {
            const double X = (double)(i - 1) * f.dxf;
            xf[i] = 0.5 * (1. - 0.5 * X);
        }
    } else if (jfn == 2) {
        const double FLN2 = std::log(2.), RECPI = 1. / (2. * std::asin(1.)), FGNORM = std::sqrt(FLN2 * RECPI);
        xf[1] = FGNORM * std::exp(-FLN2 * 0.);
        for (int j = 2; j <= f.nf; j++) {
            const double X = (double)(j - 1) * f.dxf, XSQ = X * X;
            xf[j] = FGNORM * std::exp(-FLN2 * XSQ);
        }
        xf[f.nf] = 0.;
    } else {
        const double XNORM = f.xscal / PI;
        xf[1] = XNORM;
        for (int i = 2; i <= f.nf; i++) {
            const double X = (double)(i - 1) * f.dxf, T = X * f.xscal, q = std::sin(T) / T;
            xf[i] = (jfn == 3) ? XNORM * (q * q) : XNORM * q;
        }
    }
}

struct CPanel {
    int64_t q0;    // global box index of the panel's box 1
    double v1i;    // V1I after SHRKSC
};

struct ConvPlan {
    int jfn = 0, nfmax = 0;
    double hwhm = 0., dvo = 0., dvi = 0., ratio = 0., zslope = 0., zbound = 0., xnorm = 0., sratio = 1.;
    int maxf = 0, iratsh = 1;
    int64_t raw0 = 0;               // input index of box 0's first point
    int64_t qlo = 0, qhi = -1;      // boxes CONVSC reaches
    std::vector<CPanel> panels;     // the panels holding them, in order
    std::vector<double> vft;        // VFT of output panels 0..kstar+1
    int kstar = 0, nnhi = 0;        // the panel where panelOutput sets ISTOP, and its NNHI
    bool exhausted = false;         // the input ends before V2 + bound
    std::vector<double> xf;
    double v1 = 0.;
    int64_t nout_of(int kex) const {  // output count when the last panel written is kex
        if (!exhausted || kex >= kstar) return (int64_t)2400 * kstar + nnhi - (CONV_NSHIFT);
        return (int64_t)2400 * (kex + 1);
    }
};

int plan_conv(int64_t ns, double sv1, double sdv, double V1, double V2, double dvout, int jfn, double HWHM,
              ConvPlan &cp) {
    const ScanFnTab &F = SCAN_FN[jfn];
    cp.jfn = jfn;
    cp.nfmax = F.nfmax;
    cp.hwhm = HWHM;
    cp.v1 = V1;
    const double SAMPL = -std::fabs(dvout);
    double SAMPLE = F.smpl;
    if (SAMPL < 0.) SAMPLE = HWHM / (-SAMPL);
    const double DVO = HWHM / SAMPLE;
    cp.dvo = DVO;
    const double HWF = F.hwf, DXF = F.dxf;
    const int NBOUND = (int)((2. * HWF) * SAMPLE + 0.01);
    const double BOUND = (double)NBOUND * DVO / 2.;
    cp.maxf = CONV_NLIMF + 2 * NBOUND + CONV_NSHIFT;
    if (cp.maxf > CONV_R1)
        return fail(CLBLM_ERR_ARG, "scan: output spacing too fine (the reference's R1 holds 5000 points)");
    if ((HWF / DVO * HWHM + 0.5) * (DVO / DXF / HWHM) * (1. + 1e-9) + 1.5 >= (double)(F.nfmax + 1))
        return fail(CLBLM_ERR_ARG, "scan: output spacing too coarse for the scanning function's table");
    scan_shape(jfn, cp.xf);
    const double VBOT = V1 - BOUND, VTOP = V2 + BOUND;
    // output panels: VFT accumulated, ISTOP where NNHI <= NHI
    const int NHI = CONV_NLIMF + CONV_NSHIFT - 1;
    double VFT = V1 - (double)CONV_NSHIFT * DVO;
    for (int k = 0;; k++) {
        cp.vft.push_back(VFT);
        const double z = (V2 - VFT) / DVO + 1.5;
        if (z < (double)NHI + 1.) {
            cp.kstar = k;
            cp.nnhi = (int)z;
            break;
        }
        if (k > 400000) return fail(CLBLM_ERR_ARG, "scan: output range too long");
        VFT = VFT + (double)(CONV_NLIMF - 1) * DVO;
    }
    cp.vft.push_back(VFT + (double)(CONV_NLIMF - 1) * DVO);
    if ((int64_t)2400 * cp.kstar + cp.nnhi - CONV_NSHIFT < 1) return fail(CLBLM_ERR_ARG, "scan: empty output range");
    // the input panels (panelInput + SHRKSC), data left out
    static const int JRATIO[24] = {1, 2, 3, 4, 5, 6, 8, 10, 12, 15, 16, 20, 24, 25, 30, 32, 40, 48, 50, 60, 75, 80, 100, 120};
    int64_t done = 0, qnext = 0;
    double done_v2 = sv1 - sdv;
    int NREN = 0, INIT = 0, IRATSH = 1;
    double DVSC = 0., V1SHFT = 0.;
    bool first = true;
    cp.dvi = sdv;
    for (;;) {
        int nlim;
        double VMIN = 0., VMAX = 0.;
        for (;;) {
            nlim = 2400;
            if (done + nlim > ns) nlim = (int)(ns - done);
            if (nlim == 0) break;
            VMIN = done_v2 + sdv;
            VMAX = VMIN + (double)(nlim - 1) * sdv;
            done += nlim;
            done_v2 = VMAX;
            if (VMAX >= VBOT) break;
        }
        if (nlim == 0) { cp.exhausted = true; break; }
        const int NLOW = NREN > 0 ? NREN + 1 : 1;
        double V1I = VMIN - (double)(NLOW - 1) * sdv;
        int NNI = nlim + NLOW - 1;
        const bool IDATA = VMAX > VTOP;
        int ILO = 1, IHI = IDATA ? (int)((VTOP - V1I) / sdv + 1.5) : NNI;
        if (first) cp.raw0 = done - nlim;
        NREN = 0;
        bool skip = false;
        if (INIT == 0) {
            DVSC = HWHM / 12.;
            IRATSH = (int)(DVSC / sdv + 0.5);
            for (int i = 2; i <= 24; i++)
                if (JRATIO[i - 1] > IRATSH) { IRATSH = JRATIO[i - 2]; break; }
            if (IRATSH > JRATIO[23]) IRATSH = JRATIO[23];
            if (IRATSH <= 1) skip = true;
            else {
                DVSC = (double)IRATSH * sdv;
                V1SHFT = (double)(IRATSH - 1) * sdv / 2.;
                INIT = 1;
            }
        }
        int nbox = NNI;
        if (!skip && IRATSH > 1) {
            NREN = NNI - (NNI / IRATSH) * IRATSH;
            V1I = V1I + V1SHFT;
            nbox = NNI / IRATSH;
            ILO = (int)((VBOT - V1I) / DVSC + 1.5);
            if (ILO < 1) ILO = 1;
            IHI = (int)((VTOP - V1I) / DVSC + 1.5);
            if (IHI > nbox) IHI = nbox;
            cp.iratsh = IRATSH;
            cp.sratio = IRATSH;
            cp.dvi = DVSC;
        } else {
            cp.iratsh = 1;
        }
        const int64_t q0 = qnext;
        qnext += nbox;
        if (ILO > IHI) break;     // CONVSC reaches nothing here and writes panels up to ISTOP
        if (!first && ILO != 1)
            return fail(CLBLM_ERR_UNSUPPORTED, "scan: a later input panel starts below the scan bound");
        cp.panels.push_back({q0, V1I});
        if (first) cp.qlo = q0 + ILO - 1;
        cp.qhi = q0 + IHI - 1;
        first = false;
        if (IDATA) break;         // the last panel CONVSC reads
    }
    if (cp.qhi >= cp.qlo && cp.raw0 + (cp.qhi + 1) * (int64_t)cp.iratsh > ns)
        return fail(CLBLM_ERR_UNSUPPORTED, "scan: a box reaches past the input");
    cp.ratio = cp.dvi / DVO;
    cp.zslope = (DVO / DXF) / HWHM;
    cp.zbound = (HWF / DVO) * HWHM;
    cp.xnorm = cp.dvi / HWHM;
    return 0;
}

struct ConvArgs {
    const CPanel *panels;
    int npanel;
    const double *vft;
    int kstar, maxf, iratsh;
    double dvo, ratio, zbound, zslope, xnorm, sratio;
    int64_t qlo, nq, raw0;
};

// One lane per box q = qlo + t: its output panel, JMIN, JMAX and ZF(JMIN).
__global__ void k_conv_boxes(ConvArgs a, int32_t *g0, int32_t *ge, int32_t *kk, double *zf0) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.nq) return;
    const int64_t q = a.qlo + t;
    int lo = 0, hi = a.npanel - 1;            // the panel holding q
    while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if (a.panels[mid].q0 <= q) lo = mid;
        else hi = mid - 1;
    }
    const double V1I = a.panels[lo].v1i;
    const double ip = (double)(q - a.panels[lo].q0) * a.ratio;   // REAL(I-1) * RATIO
    const double lim = (double)a.maxf + 1.;                      // INT(x) <= MAXF  <=>  x < MAXF + 1
    auto top = [&](int k) { return ((ip + (V1I - a.vft[k]) / a.dvo) + a.zbound) + 1.5; };
    int k = (int)fmax(0., ceil((top(0) - lim) / 2400.));
    if (k > a.kstar + 1) k = a.kstar + 1;
    while (k > 0 && top(k - 1) < lim) k--;
    while (k <= a.kstar && !(top(k) < lim)) k++;
    if (k > a.kstar) {                        // reached only after the last panel is written
        g0[t] = CONV_DEAD;
        ge[t] = CONV_DEAD;
        kk[t] = k;
        zf0[t] = 0.;
        return;
    }
    const double ZPEAK = ip + (V1I - a.vft[k]) / a.dvo;
    const int JMAX = (int)((ZPEAK + a.zbound) + 1.5);
    int JMIN = (int)((ZPEAK - a.zbound) + 1.5);
    if (JMIN < 1) JMIN = 1;
    g0[t] = JMIN + 2400 * k;
    ge[t] = JMAX + 2400 * k;
    kk[t] = k;
    zf0[t] = ((double)(JMIN - 1) - ZPEAK) * a.zslope;
}

__global__ void k_conv_zero(int32_t *p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// One lane per box and row: the box sum (SHRKSC), SI = XNORM S, and the last output
// panel of a nonzero box.
__global__ void k_conv_si(ConvArgs a, const double *x, int64_t ldx, const int32_t *kk, double *si, int32_t *kex) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.nq) return;
    const int r = blockIdx.y;
    const double *src = x + r * ldx + a.raw0 + (a.qlo + t) * (int64_t)a.iratsh;
    double S;
    if (a.iratsh > 1) {
        double SUMK = 0.;
        for (int j = 0; j < a.iratsh; j++) SUMK = SUMK + src[j];
        S = SUMK / a.sratio;
    } else {
        S = src[0];
    }
    si[r * a.nq + t] = a.xnorm * S;
    if (S != 0. && kk[t] <= a.kstar) atomicMax(kex + r, kk[t]);
}

// One wave per 64 outputs: the boxes that reach them, in order, 64 at a time.
__global__ __launch_bounds__(64) void k_conv_sum(ConvArgs a, const int32_t *g0, const int32_t *ge,
                                                  const int32_t *kk, const double *zf0, const double *si,
                                                  const double *xf, int nfmax, double *out, int64_t ldo,
                                                  int64_t nout) {
    __shared__ int it[CONV_TILE][CONV_TILE + 1];
    const int lane = threadIdx.x;
    const int r = blockIdx.y;
    const int64_t o0 = (int64_t)blockIdx.x * CONV_TILE;
    const int Ga = (int)(o0 + CONV_NSHIFT + 1);
    const int Gb = Ga + CONV_TILE - 1;
    // first box whose last output reaches Ga (ge is nondecreasing up to 1: a box on a
    // later output panel may round one position lower), last box starting by Gb
    int64_t lo = 0, hi = a.nq;
    while (lo < hi) {
        const int64_t mid = (lo + hi) / 2;
        if (ge[mid] >= Ga - 2) hi = mid;
        else lo = mid + 1;
    }
    const int64_t qa = lo;
    lo = qa - 1;
    hi = a.nq - 1;
    while (lo < hi) {
        const int64_t mid = (lo + hi + 1) / 2;
        if (g0[mid] <= Gb + 2) lo = mid;
        else hi = mid - 1;
    }
    const int64_t qb = lo;
    const double *sir = si + r * a.nq;
    double acc = 0.;
    for (int64_t base = qa; base <= qb; base += CONV_TILE) {
        const int64_t t = base + lane;
        double s = 0.;
        for (int g = 0; g < CONV_TILE; g++) it[lane][g] = -1;
        if (t <= qb) {
            s = sir[t];
            const int G0 = g0[t], GE = ge[t];
            const int Gs = max(G0, 2400 * kk[t] + CONV_NSHIFT + 1);
            const int gl = max(Gs, Ga), gh = min(GE, Gb);
            if (G0 != CONV_DEAD && gl <= gh) {
                const double zs = a.zslope;
                double zf = zf0[t];
                for (int g = G0; g < gl; g++) zf = zf + zs;
                for (int g = gl; g <= gh; g++) {
                    int idx = (int)(fabs(zf) + 1.5);
                    it[lane][g - Ga] = idx < nfmax ? idx : nfmax;
                    zf = zf + zs;
                }
            }
        }
        __syncthreads();
        for (int j = 0; j < CONV_TILE; j++) {
            const double sj = __shfl(s, j);
            const int idx = it[j][lane];
            if (idx >= 0) acc = acc + sj * xf[idx];
        }
        __syncthreads();
    }
    const int64_t o = o0 + lane;
    if (o < nout) out[r * ldo + o] = acc;
}

int run_conv(ScanState &S, const ConvPlan &cp, const double *xs, int64_t ldx, int64_t nrows, bool count_only,
             double *d_out, int64_t ld_out, int64_t *nout, hipStream_t st) {
    const int64_t nq = cp.qhi - cp.qlo + 1;
    if (nq <= 0) {
        *nout = cp.nout_of(0);
        if (count_only) return 0;
        if (ld_out < *nout) return fail(CLBLM_ERR_ARG, "ld_out is below the output length");
        hipLaunchKernelGGL(k_rect_fill, dim3(cdiv(*nout, 256), (unsigned)nrows), dim3(256), 0, st, d_out, *nout,
                           ld_out);
        SCHK(hipGetLastError());
        return 0;
    }
    if (!cp.exhausted) {
        *nout = cp.nout_of(cp.kstar);
        if (count_only) return 0;
    }
    if (!xs) return fail(CLBLM_ERR_ARG, "input spectrum is NULL");
    auto up = [](size_t o) { return (o + 255) & ~(size_t)255; };
    const size_t o_xf = 0, o_vft = up(sizeof(double) * cp.xf.size());
    const size_t o_pan = up(o_vft + sizeof(double) * cp.vft.size());
    const size_t o_g0 = up(o_pan + sizeof(CPanel) * cp.panels.size());
    const size_t o_ge = up(o_g0 + 4 * (size_t)nq), o_kk = up(o_ge + 4 * (size_t)nq);
    const size_t o_zf = up(o_kk + 4 * (size_t)nq), o_si = up(o_zf + 8 * (size_t)nq);
    const size_t o_kex = up(o_si + 8 * (size_t)nq * (size_t)nrows);
    const size_t total = o_kex + 4 * (size_t)nrows + 256;
    SCHK(hipStreamSynchronize(st));   // the buffer may still be read by a previous scan
    int rc = grow(&S.conv, &S.conv_cap, total);
    if (rc) return rc;
    char *b = (char *)S.conv;
    std::vector<char> host(o_g0, 0);
    std::memcpy(host.data() + o_xf, cp.xf.data(), sizeof(double) * cp.xf.size());
    std::memcpy(host.data() + o_vft, cp.vft.data(), sizeof(double) * cp.vft.size());
    std::memcpy(host.data() + o_pan, cp.panels.data(), sizeof(CPanel) * cp.panels.size());
    SCHK(hipMemcpyAsync(b, host.data(), o_g0, hipMemcpyHostToDevice, st));
    ConvArgs a{(const CPanel *)(b + o_pan), (int)cp.panels.size(), (const double *)(b + o_vft), cp.kstar, cp.maxf,
               cp.iratsh, cp.dvo, cp.ratio, cp.zbound, cp.zslope, cp.xnorm, cp.sratio, cp.qlo, nq, cp.raw0};
    int32_t *g0 = (int32_t *)(b + o_g0), *ge = (int32_t *)(b + o_ge), *kk = (int32_t *)(b + o_kk);
    int32_t *kex = (int32_t *)(b + o_kex);
    double *zf0 = (double *)(b + o_zf), *si = (double *)(b + o_si);
    hipLaunchKernelGGL(k_conv_zero, dim3(cdiv(nrows, 64)), dim3(64), 0, st, kex, nrows);
    hipLaunchKernelGGL(k_conv_boxes, dim3(cdiv(nq, 256)), dim3(256), 0, st, a, g0, ge, kk, zf0);
    hipLaunchKernelGGL(k_conv_si, dim3(cdiv(nq, 256), (unsigned)nrows), dim3(256), 0, st, a, xs, ldx, kk, si, kex);
    SCHK(hipGetLastError());
    if (cp.exhausted) {   // the count follows the last nonzero box of each row
        std::vector<int32_t> kx((size_t)nrows);
        SCHK(hipMemcpyAsync(kx.data(), kex, 4 * (size_t)nrows, hipMemcpyDeviceToHost, st));
        SCHK(hipStreamSynchronize(st));
        *nout = cp.nout_of(kx[0]);
        for (int64_t r = 1; r < nrows; r++)
            if (cp.nout_of(kx[r]) != *nout)
                return fail(CLBLM_ERR_ARG, "scan: the rows end on different output counts (the input ends before "
                                           "V2 + bound and trailing zeros differ); scan them separately");
        if (count_only) return 0;
    }
    if (ld_out < *nout) return fail(CLBLM_ERR_ARG, "ld_out is below the output length");
    hipLaunchKernelGGL(k_conv_sum, dim3(cdiv(*nout, CONV_TILE), (unsigned)nrows), dim3(CONV_TILE), 0, st, a, g0, ge,
                       kk, zf0, si, (const double *)(b + o_xf), cp.nfmax, d_out, ld_out, *nout);
    SCHK(hipGetLastError());
    return 0;
}

// clblm_INTERP of each row onto the plan's grid (HWHM/12) into S.xs.
int regrid_rows(ScanState &S, const IPlan &ip, const double *d_in, int64_t nrows, int64_t ld_in, hipStream_t st) {
    const int64_t ns = ip.nout;
    auto up = [](size_t o) { return (o + 255) & ~(size_t)255; };
    const size_t o_seg = up(sizeof(ICoef));
    const size_t meta = o_seg + sizeof(ISeg) * ip.segs.size() + 256;
    SCHK(hipStreamSynchronize(st));  // the metadata buffer may still be read by a previous scan
    int rc = grow(&S.meta, &S.meta_cap, meta);
    if (rc) return rc;
    rc = grow((void **)&S.xs, &S.xs_cap, sizeof(double) * (size_t)ns * (size_t)nrows);
    if (rc) return rc;
    std::vector<char> host(meta, 0);
    const ICoef cf = make_coef();
    std::memcpy(host.data(), &cf, sizeof cf);
    if (!ip.segs.empty()) std::memcpy(host.data() + o_seg, ip.segs.data(), sizeof(ISeg) * ip.segs.size());
    SCHK(hipMemcpyAsync(S.meta, host.data(), meta, hipMemcpyHostToDevice, st));
    SCHK(hipStreamSynchronize(st));  // the host copy goes out of scope
    char *mb = (char *)S.meta;
    for (int64_t r = 0; r < nrows; r++)
        hipLaunchKernelGGL(k_scan_interp<double>, dim3(cdiv((int64_t)ip.segs.size(), 64)), dim3(64), 0, st,
                           (const ISeg *)(mb + o_seg), (int64_t)ip.segs.size(), d_in + r * ld_in, (const ICoef *)mb,
                           ip.vratio, S.xs + r * ns);
    SCHK(hipGetLastError());
    return 0;
}

// functIDs 2..5, -4, -5 (JFN 1..4, clblm_ScanFilter.f90:239-334)
int scanfn_conv(ScanState &S, const double *d_in, int64_t n, int64_t nrows, int64_t ld_in, double v1, double dv,
                double v1out, double v2out, double dvout, int funct_id, double hwhm, double *d_out, int64_t ld_out,
                int64_t *nout, double *v1res, double *dvres, hipStream_t st) {
    int jfn = std::abs(funct_id) - 1;
    if (jfn < 1 || jfn > 4) return fail(CLBLM_ERR_UNSUPPORTED, "scan: functID must be 1..5, -4 or -5");
    double HWHM = hwhm;
    if (funct_id < 0) {   // the first zero crossing in place of the half width (:272-280)
        if (jfn != 3 && jfn != 4) return fail(CLBLM_ERR_ARG, "scan: a negative functID is only for -4 and -5");
        HWHM = HWHM / (2. * std::asin(1.) / SCAN_FN[jfn].xscal);
    }
    const double DVINT = HWHM / 12.;
    IPlan ip;
    const bool pre = dv > DVINT;
    int64_t ns = n;
    double sv1 = v1, sdv = dv;
    if (pre) {
        int rc = plan_interp(n, v1, dv, v1, v1 + dv * (double)(n - 1), DVINT, ip);
        if (rc) return rc;
        ns = ip.nout;
        sv1 = ip.v1;
        sdv = DVINT;
    }
    ConvPlan cp;
    int rc = plan_conv(ns, sv1, sdv, v1out, v2out, dvout, jfn, HWHM, cp);
    if (rc) return rc;
    *v1res = v1out;
    *dvres = cp.dvo;
    const bool count_only = d_out == nullptr;
    const bool need_data = !count_only || (cp.exhausted && cp.qhi >= cp.qlo);
    const double *xs = d_in;
    int64_t ldx = ld_in;
    if (need_data) {
        if (!d_in) return fail(CLBLM_ERR_ARG, "input spectrum is NULL (the output count depends on it here)");
        if (pre) {
            rc = regrid_rows(S, ip, d_in, nrows, ld_in, st);
            if (rc) return rc;
            xs = S.xs;
            ldx = ns;
        }
    }
    return run_conv(S, cp, need_data ? xs : nullptr, ldx, nrows, count_only, d_out, ld_out, nout, st);
}

// ==================================================================================
// Spectral response function filter: clblm_FLTRFN with CNVFLT
// (clblm_ScanFilter.f90:1703-1900) for each channel of postProc_filter_singleStruct
// (clblm_PostProc.f90:289-336).  One lane per channel walks the input panels the
// channel reads (panelInput's accumulated panel bounds come from the host) and sums
// its points in the reference's order, so the result is the reference's to the bit.
// ==================================================================================
struct FChan {
    double v1f, v2f, dvf, sumflt;
    int64_t xf;        // the channel's filter values (1-based at xf + 1, end values halved)
    int32_t npts, bad;
};

__global__ void k_filter(const FChan *ch, int nchan, const double *xfv, const double *x, const double *vmin,
                         const double *vmax, const int32_t *plen, int npanel, double dv, double *out, int *err) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nchan) return;
    const FChan C = ch[c];
    if (C.bad) {          // V1F < 0: the reference returns without a value
        out[c] = 0.;
        return;
    }
    const double *XF = xfv + C.xf;
    const double V1F = C.v1f, V2F = C.v2f, DVF = C.dvf, VBOT = V1F, VTOP = V2F;
    double RFILTR = 0.0;
    int p = 0;
    int64_t start = 0;
    while (p < npanel && vmax[p] < VBOT) start += plen[p++];
    for (; p < npanel; start += plen[p++]) {
        const double VMIN = vmin[p], VMAX = vmax[p];
        const int nlim = plen[p];
        const double *S = x + start - 1;
        int ILO = 1, IHI = nlim, IDATA = 0;
        if (VMAX > VTOP) {
            IHI = (int)((VTOP - VMIN) / dv + 1.5);
            IDATA = 1;
        }
        const double DVI = dv, V1I = VMIN;
        int IMIN = (int)((V1F - V1I) / DVI + 1.0001);
        if (IMIN < ILO) IMIN = ILO;
        int IMAX = (int)((V2F + V1F - V1I) / DVI + 1.0001);
        if (IMAX > IHI) IMAX = IHI;
        const double XIF0 = (V1I - V1F) / DVF + 1.0001;
        const double XDVIF = DVI / DVF;
        if ((int)(XIF0 + XDVIF * (double)IMAX) + 1 > C.npts) {
            for (int i = IMAX - 1; i >= IMIN; i--)
                if ((int)(XIF0 + XDVIF * (double)i) + 1 <= C.npts) {
                    IMAX = i;
                    break;
                }
        }
        if ((int)(XIF0 + XDVIF * (double)IMAX) + 1 > C.npts) {   // the reference stops here
            atomicOr(err, 1);
            out[c] = __longlong_as_double(0x7ff8000000000000LL);
            return;
        }
        for (int I = IMIN; I <= IMAX; I++) {
            const int IFL = (int)(XIF0 + XDVIF * (double)I);
            const double v2s = V1I + (double)I * DVI;
            const double vxf1 = V1F + (double)(IFL - 1) * DVF;
            const double vxf2 = V1F + (double)IFL * DVF;
            const double pp = (vxf2 - v2s) / (vxf2 - vxf1);
            RFILTR = RFILTR + S[I] * (pp * XF[IFL] + (1. - pp) * XF[IFL + 1]);
        }
        if (IDATA == 1) break;
    }
    out[c] = (RFILTR * dv) / C.sumflt;
}

}  // namespace

extern "C" __attribute__((visibility("hidden"))) int clblm_filter_impl(void *state, const double *d_in, int64_t n,
                                                                       double v1, double dv, int nchan,
                                                                       const double *v1f, const double *dvf,
                                                                       const int32_t *nptf, const int64_t *xf_off,
                                                                       const double *xf, double *d_out, void *stream) {
    if (!d_in || !d_out || !v1f || !dvf || !nptf || !xf_off || !xf || n < 2 || nchan < 1 || !(dv > 0.))
        return fail(CLBLM_ERR_ARG, "filter: null argument, fewer than 2 points or no channel");
    ScanState &S = *(ScanState *)state;
    hipStream_t st = (hipStream_t)stream;
    // the channels (the filter values copied with their end values halved) and their sums
    std::vector<FChan> ch((size_t)nchan);
    std::vector<double> vals;
    for (int c = 0; c < nchan; c++) {
        const int np = nptf[c];
        if (np < 2 || np > 3001) return fail(CLBLM_ERR_ARG, "filter: 2..3001 points per channel (NFLTPT)");
        FChan &C = ch[(size_t)c];
        double DVF = dvf[c], V1F = v1f[c];
        C.bad = V1F < 0 ? 1 : 0;
        if (DVF < 0.) {
            DVF = std::fabs(DVF);
            if ((np - 1) % 2 != 0) return fail(CLBLM_ERR_ARG, "filter: a centre frequency needs an odd number of points");
            V1F = V1F - DVF * (double)((np - 1) / 2);
        }
        if (!(DVF > 0.)) return fail(CLBLM_ERR_ARG, "filter: the filter step must be nonzero");
        C.v1f = V1F;
        C.dvf = DVF;
        C.v2f = V1F + DVF * (double)(np - 1);
        C.npts = np;
        C.xf = (int64_t)vals.size();
        vals.push_back(0.);   // 1-based
        const double *f = xf + xf_off[c];
        for (int i = 0; i < np; i++) vals.push_back(f[i]);
        vals[(size_t)C.xf + 1] = 0.5 * vals[(size_t)C.xf + 1];
        vals[(size_t)C.xf + (size_t)np] = 0.5 * vals[(size_t)C.xf + (size_t)np];
        double sum = 0.0;
        for (int i = 1; i <= np; i++) sum = sum + vals[(size_t)C.xf + (size_t)i];
        C.sumflt = sum * DVF;
    }
    // panelInput's panel bounds, accumulated as the reference does
    std::vector<double> vmin, vmax;
    std::vector<int32_t> plen;
    double done_v2 = v1 - dv;
    for (int64_t done = 0; done < n;) {
        const int nlim = (int)std::min<int64_t>(2400, n - done);
        const double a = done_v2 + dv, b = a + (double)(nlim - 1) * dv;
        vmin.push_back(a);
        vmax.push_back(b);
        plen.push_back(nlim);
        done += nlim;
        done_v2 = b;
    }
    auto up = [](size_t o) { return (o + 255) & ~(size_t)255; };
    const size_t o_ch = 0, o_v = up(sizeof(FChan) * ch.size()), o_a = up(o_v + sizeof(double) * vals.size());
    const size_t o_b = up(o_a + sizeof(double) * vmin.size()), o_l = up(o_b + sizeof(double) * vmax.size());
    const size_t o_e = up(o_l + sizeof(int32_t) * plen.size()), total = o_e + 256;
    SCHK(hipStreamSynchronize(st));   // the buffer may still be read by a previous call
    int rc = grow(&S.filt, &S.filt_cap, total);
    if (rc) return rc;
    std::vector<char> host(o_e + sizeof(int), 0);
    std::memcpy(host.data() + o_ch, ch.data(), sizeof(FChan) * ch.size());
    std::memcpy(host.data() + o_v, vals.data(), sizeof(double) * vals.size());
    std::memcpy(host.data() + o_a, vmin.data(), sizeof(double) * vmin.size());
    std::memcpy(host.data() + o_b, vmax.data(), sizeof(double) * vmax.size());
    std::memcpy(host.data() + o_l, plen.data(), sizeof(int32_t) * plen.size());
    char *b = (char *)S.filt;
    SCHK(hipMemcpyAsync(b, host.data(), host.size(), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_filter, dim3(cdiv(nchan, 64)), dim3(64), 0, st, (const FChan *)(b + o_ch), nchan,
                       (const double *)(b + o_v), d_in, (const double *)(b + o_a), (const double *)(b + o_b),
                       (const int32_t *)(b + o_l), (int)plen.size(), dv, d_out, (int *)(b + o_e));
    SCHK(hipGetLastError());
    int herr = 0;
    SCHK(hipMemcpyAsync(&herr, b + o_e, sizeof(int), hipMemcpyDeviceToHost, st));
    SCHK(hipStreamSynchronize(st));   // the host copies go out of scope
    if (herr)
        return fail(CLBLM_ERR_ARG, "filter: a channel reads past its filter (CNVFLT's subscript check, where the "
                                   "reference stops); its output is NaN, the other channels are filled");
    return 0;
}

extern "C" __attribute__((visibility("hidden"))) int clblm_scanfn_impl(
    void *state, const double *d_in, int64_t n, int64_t nrows, int64_t ld_in, double v1, double dv, double v1out,
    double v2out, double dvout, int funct_id, double hwhm, double *d_out, int64_t ld_out, int64_t *nout,
    double *v1res, double *dvres, void *stream) {
    if (!(hwhm > 0.) || !(dv > 0.) || n < 8 || nrows < 1 || nrows > 65535 || ld_in < n)
        return fail(CLBLM_ERR_ARG, "hwhm, dv must be positive, n >= 8, 1 <= nrows <= 65535, ld_in >= n");
    ScanState &S = *(ScanState *)state;
    hipStream_t st = (hipStream_t)stream;
    if (funct_id != 1)
        return scanfn_conv(S, d_in, n, nrows, ld_in, v1, dv, v1out, v2out, dvout, funct_id, hwhm, d_out, ld_out,
                           nout, v1res, dvres, st);
    if (!(dvout != 0.)) return fail(CLBLM_ERR_ARG, "the rectangle scan needs a nonzero dvout");
    // clblm_INTERP onto HWHM/12 when the input is coarser (clblm_ScanFilter.f90:331-334)
    const double DVINT = hwhm / 12.;
    IPlan ip;
    const bool pre = dv > DVINT;
    int64_t ns = n;
    double sv1 = v1, sdv = dv;
    if (pre) {
        int rc = plan_interp(n, v1, dv, v1, v1 + dv * (double)(n - 1), DVINT, ip);
        if (rc) return rc;
        ns = ip.nout;
        sv1 = ip.v1;
        sdv = DVINT;
    }
    RectPlan rp;
    plan_rect(ns, sv1, sdv, v1out, v2out, dvout, hwhm, rp);
    for (const RPart &q : rp.parts)   // every box reads inside the spectrum
        if (q.il < 0 || q.ih < 0 || q.il >= ns || q.ih >= ns)
            return fail(CLBLM_ERR_ARG, "rectangle scan: a box reaches outside the spectrum");
    *nout = rp.nout;
    *v1res = rp.v1;
    *dvres = rp.dvo;
    if (!d_out) return 0;
    if (!d_in) return fail(CLBLM_ERR_ARG, "input spectrum is NULL");
    if (ld_out < (rp.scan ? rp.nout : ns)) return fail(CLBLM_ERR_ARG, "ld_out is below the output length");
    const double *xs = d_in;
    int64_t ldx = ld_in;
    if (pre) {
        int rc = regrid_rows(S, ip, d_in, nrows, ld_in, st);
        if (rc) return rc;
        xs = S.xs;
        ldx = ns;
    }
    auto up = [](size_t o) { return (o + 255) & ~(size_t)255; };
    const size_t o_box = 0;
    const size_t o_part = up(o_box + sizeof(RBox) * rp.boxes.size());
    const size_t meta = o_part + sizeof(RPart) * rp.parts.size() + 256;
    SCHK(hipStreamSynchronize(st));  // the buffer may still be read by a previous scan
    int rc = grow(&S.conv, &S.conv_cap, meta);
    if (rc) return rc;
    std::vector<char> host(meta, 0);
    if (!rp.boxes.empty()) std::memcpy(host.data() + o_box, rp.boxes.data(), sizeof(RBox) * rp.boxes.size());
    if (!rp.parts.empty()) std::memcpy(host.data() + o_part, rp.parts.data(), sizeof(RPart) * rp.parts.size());
    SCHK(hipMemcpyAsync(S.conv, host.data(), meta, hipMemcpyHostToDevice, st));
    SCHK(hipStreamSynchronize(st));  // the host copy goes out of scope
    char *mb = (char *)S.conv;
    const unsigned ny = (unsigned)nrows;
    if (!rp.scan) {
        hipLaunchKernelGGL(k_rect_copy, dim3(cdiv(ns, 256), ny), dim3(256), 0, st, xs, ldx, ns, d_out, ld_out);
    } else {
        hipLaunchKernelGGL(k_rect_fill, dim3(cdiv(std::max<int64_t>(1, rp.nout), 256), ny), dim3(256), 0, st, d_out,
                           rp.nout, ld_out);
        const int64_t nb = (int64_t)rp.boxes.size();
        if (nb > 0)
            hipLaunchKernelGGL(k_rect_boxes, dim3(cdiv(nb, 64), ny), dim3(64), 0, st, (const RBox *)(mb + o_box), nb,
                               (const RPart *)(mb + o_part), xs, ldx, d_out, ld_out);
    }
    SCHK(hipGetLastError());
    return 0;
}
