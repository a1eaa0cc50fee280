// interpSpectrum_array2array (src/clblm_src/clblm_Spectrum.f90:326-590) on the device:
// a spectrum X[1..NIN] on V1 + (i-1) DVin onto V1 + (j-1) DVout, j = 1..NOUT.
//   * DVout/DVin = 1 (within 1e-8): the points themselves (linear extrapolation for
//     the two points past either end, zero beyond, the reference's spInExt);
//   * DVout/DVin = n/(n+1) or (n+1)/n, n = 1..5: the aligned 4-point Lagrange form,
//     one output point per work unit;
//   * any other ratio: the general form, whose fraction P is accumulated point by
//     point in blocks of 2400 output points and quantised to 1/200 for the
//     coefficient index.  A block is cut into SUB pieces of SUBLEN points, one work
//     unit each; the unit of a piece replays the recurrence of P (the same additions
//     in the same order, no memory traffic) from the block's start to its first point,
//     so every point gets the reference's P bit for bit.
// Both the OD re-grid of ODLAY (DVout) and interpToFinestGrid use it.
#pragma once
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdint>

namespace clblm_interp {

__host__ __device__ inline int regrid_type(double DVin, double DVout, double *ratio, int *lenIn, int *lenOut) {
    const double TOL = 1.e-8;
    double dvRatio = DVout / DVin;
    if (fabs(DVout - DVin) < TOL) dvRatio = 1.;
    const double hi2lo[5] = {2. / 1., 3. / 2., 4. / 3., 5. / 4., 6. / 5.};
    const double lo2hi[5] = {1. / 2., 2. / 3., 3. / 4., 4. / 5., 5. / 6.};
    bool anyhi = false, anylo = false;
    for (int k = 0; k < 5; k++) {
        if (fabs(dvRatio - hi2lo[k]) < TOL) anyhi = true;
        if (fabs(dvRatio - lo2hi[k]) < TOL) anylo = true;
    }
    *ratio = dvRatio;
    if (fabs(dvRatio - 1.) < 2.220446049250313e-16) { *lenIn = 1; *lenOut = 1; return 0; }
    if (anyhi) { *lenIn = (int)(DVout / (DVout - DVin) + 0.5); *lenOut = *lenIn - 1; return *lenOut - 1; }
    if (anylo) { *lenIn = (int)(DVout / (DVin - DVout) + 0.5); *lenOut = *lenIn + 1; return *lenOut - 1; }
    *lenIn = 0; *lenOut = 0;
    return -1;
}

// spectraHaveSameGrid (clblm_Spectrum.f90) for two spectra from one V1: DV, V2 =
// V1 + DV (NLIM - 1) (CLBLM_Spectrum_init) and NLIM equal, DV and V2 to within
// epsilon(1d0).
__host__ __device__ inline bool same_grid(double v1, double dva, int64_t na, double dvb, int64_t nb) {
    const double EPS = 2.220446049250313e-16;
    return fabs(dva - dvb) < EPS && na == nb &&
           fabs((v1 + dva * (double)(na - 1)) - (v1 + dvb * (double)(nb - 1))) < EPS;
}

constexpr int64_t BLOCK = 2400;
constexpr int64_t SUB = 64;
constexpr int64_t SUBLEN = (BLOCK + SUB - 1) / SUB;

// Work units of one interpolation: one per output point for the aligned forms, SUB
// per 2400-point block for the general form.
__host__ __device__ inline int64_t units(double DVin, double DVout, int64_t NOUT) {
    double r;
    int a, b;
    return regrid_type(DVin, DVout, &r, &a, &b) >= 0 ? NOUT : (NOUT + BLOCK - 1) / BLOCK * SUB;
}

// Work unit ul (0-based) of the interpolation of X[1..NIN] (1-based pointer) into
// o[1..NOUT].
__device__ inline void unit(const double *X, int64_t NIN, double V1, double DVin, double DVout, double *o,
                            int64_t NOUT, int64_t ul) {
    auto ext = [&](int64_t i) -> double {  // spInExt
        if (i >= 1 && i <= NIN) return X[i];
        if (i == 0 || i == -1) return 2.0 * X[1] - X[2];
        if (i == NIN + 1 || i == NIN + 2) return 2.0 * X[NIN] - X[NIN - 1];
        return 0.;
    };
    double dvRatio;
    int lenIn, lenOut;
    int ITYPE = regrid_type(DVin, DVout, &dvRatio, &lenIn, &lenOut);
    if (ITYPE == 0) {
        int64_t I = ul + 1;
        int64_t step = (int64_t)llround(dvRatio);
        o[I] = ext(1 - step + I * step);
    } else if (ITYPE > 0) {
        int64_t I = ul + 1;
        int64_t nl = (I - 1) % lenOut + 1, m = (I - 1) / lenOut;
        int64_t iIn = nl - lenIn;
        if (nl > 1 && DVin > DVout) iIn = iIn - 1;
        iIn += (m + 1) * lenIn;
        if (nl == 1) {
            o[I] = ext(iIn);
        } else {
            double AP = 1.0 / ((double)ITYPE + 1.0);
            double pp = AP * (double)(nl - 1);
            if (DVin > DVout) pp = 1.0 - pp;
            double B1 = -pp * (pp - 1.0) * (pp - 2.0) / 6.0;
            double B2 = (pp * pp - 1.0) * (pp - 2.0) * 0.5;
            double B3 = -pp * (pp + 1.0) * (pp - 2.0) * 0.5;
            double B4 = pp * (pp * pp - 1.0) / 6.0;
            o[I] = B1 * ext(iIn - 1) + B2 * ext(iIn) + B3 * ext(iIn + 1) + B4 * ext(iIn + 2);
        }
    } else {
        const double XNUMCF = 201.;
        const int64_t blk = ul / SUB;
        int64_t J1 = blk * BLOCK + 1;
        int64_t J2 = J1 + BLOCK - 1 < NOUT ? J1 + BLOCK - 1 : NOUT;
        const int64_t Js = J1 + (ul % SUB) * SUBLEN;
        if (Js > J2) return;
        const int64_t Je = Js + SUBLEN - 1 < J2 ? Js + SUBLEN - 1 : J2;
        double Vo = blk == 0 ? V1 : V1 + (double)(J1 - 1) * DVout;
        int64_t iIn = (int64_t)((Vo - V1) / DVin + 1.00001);
        double Vi = V1 + (double)(iIn - 1) * DVin;
        double P = (Vo - Vi) / DVin;
        // the block's recurrence up to this piece; for P >= 1, (double)(int64_t)P is
        // floor(P), so the whole steps are counted in a double (exact below 2^53)
        double adv = 0.;
        for (int64_t j = J1; j < Js; j++) {
            P = P + dvRatio;
            if (P >= 1.0) {
                const double f = floor(P);
                adv = adv + f;
                P = P - f;
            }
        }
        iIn = iIn + (int64_t)adv;
        for (int64_t j = Js; j <= Je; j++) {
            int ip = (int)(P * XNUMCF + 1.00001);
            double p = ((double)ip - 1.0) / (XNUMCF - 1.0);
            double pp = p * p;
            double A1 = -p / 2.0 * ((1 - p) * (1 - p));
            double A2 = 1.0 - pp * (3.0 - 2.0 * p) + pp / 2.0 * (1.0 - p);
            double A3 = pp * (3.0 - 2.0 * p) + p / 2.0 * ((1.0 - p) * (1.0 - p));
            double A4 = -pp / 2 * (1.0 - p);
            o[j] = A1 * ext(iIn - 1) + A2 * ext(iIn) + A3 * ext(iIn + 1) + A4 * ext(iIn + 2);
            P = P + dvRatio;
            if (P >= 1.0) {
                iIn = iIn + (int64_t)P;
                P = P - (double)(int64_t)P;
            }
        }
    }
}

}  // namespace clblm_interp
