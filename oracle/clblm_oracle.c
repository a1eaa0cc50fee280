/*
 * CLBLM oracle -- sequential CPU restatement of the reference's ODLAY path.
 *
 * TEST INFRASTRUCTURE ONLY (it checks results).  The product never links or calls it.
 *
 * Each routine follows the reference routine named in its comment, statement by
 * statement and in the reference's operation order, because the FP64 GPU path is
 * held to 1e-10 relative against this file and a one-ulp change in a table
 * argument can move a table index.  Build with -O2 -ffp-contract=off.
 *
 * Pinning (see tests/test_oracle_*.py): tables against their closed forms and the
 * reference's own compiled module (oracle/_ref), line sum against area
 * conservation, the Faddeeva Voigt, the Lorentz and Doppler limits and a brute-force
 * re-evaluation of the nested grids on a one-line case.  The exact panel
 * bookkeeping of HIRACQ/panelOut and the SHRINQ grouping are pinned only by those
 * properties: the reference's ODLAY needs netCDF and cannot be built here.
 *
 * Restrictions (error returned): line coupling (IFLG 1..3), NLTE, cross-sections,
 * species broadening, RSYM (panel start <= 0), nonzero DPTMIN/DPTFAC.
 */
#include "clblm_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../clblm_amd/csrc/clblm_data.inc"
#include "../clblm_amd/csrc/clblm_contnm.inc"

/* clblm_ConstParam.f90:49-50,59-71,92,110,112 */
#define ONEPL 1.01
#define ONEMI 0.999
#define RADCN2 1.4387752
#define BOLTZ 1.3806503E-16
#define CLIGHT 2.99792458E+10
#define AVOGAD 6.02214199E+23
#define RADCN1 1.191042722E-12
#define PI_REF 3.1415926535898
#define P0 1013.25
#define TEMP0 296.0
#define NFPTS 2001
#define NFMX 3001
#define MXMOL 47

static char g_err[512];
const char *orc_error(void) { return g_err; }
static int fail(const char *msg) {
    snprintf(g_err, sizeof g_err, "%s", msg);
    return -1;
}

/* Fortran INT(): truncation toward zero. */
static inline int ftrunc(double x) { return (int)x; }

/* ---------------------------------------------------------------------------
 * Line-shape tables.  SHAPEL clblm_ODLAY.f90:2804-2909, SHAPEG :2914-2953,
 * VERFN :2959-3016.  1-based, NFMX+1 slots.
 * ------------------------------------------------------------------------- */
static double Afn(double z0) { return (1. + 2. * z0 * z0) / ((1. + z0 * z0) * (1. + z0 * z0)); }
static double Bfn(double z0) { return -1. / ((1. + z0 * z0) * (1. + z0 * z0)); }

void orc_tables(double *F1, double *F2, double *F3, double *FG, double *XVER) {
    const double HWF1 = 4., HWF2 = 16., HWF3 = 64.;
    const double DXF1 = 0.002, DXF2 = 0.008, DXF3 = 0.032;
    const int NX1 = NFPTS, NX2 = NFPTS, NX3 = NFPTS;
    double RECPI = 1. / (2. * asin(1.));
    double A1 = Afn(HWF1), B1 = Bfn(HWF1), A2 = Afn(HWF2), B2 = Bfn(HWF2);
    double A3 = Afn(HWF3), B3 = Bfn(HWF3);
#define XLORNZ(x) (1. / (1. + (x)))
#define Q1FN(x) (A1 + B1 * (x))
#define Q2FN(x) (A2 + B2 * (x))
#define Q3FN(x) (A3 + B3 * (x))
    for (int i = 0; i <= NFMX; i++) F1[i] = F2[i] = F3[i] = FG[i] = XVER[i] = 0.;
    F1[1] = RECPI * (XLORNZ(0.) - Q1FN(0.));
    for (int jj = 2; jj <= NX1; jj++) {
        double x = (double)(jj - 1) * DXF1, xsq = x * x;
        F1[jj] = RECPI * (XLORNZ(xsq) - Q1FN(xsq));
    }
    F1[NX1] = 0.;
    F2[1] = RECPI * (Q1FN(0.) - Q2FN(0.));
    int j1lim = ftrunc(HWF1 / DXF2 + 1.001);
    for (int jj = 2; jj <= j1lim; jj++) {
        double x = (double)(jj - 1) * DXF2, xsq = x * x;
        F2[jj] = RECPI * (Q1FN(xsq) - Q2FN(xsq));
    }
    for (int jj = j1lim + 1; jj <= NX2; jj++) {
        double x = (double)(jj - 1) * DXF2, xsq = x * x;
        F2[jj] = RECPI * (XLORNZ(xsq) - Q2FN(xsq));
    }
    F2[NX2] = 0.;
    F3[1] = RECPI * (Q2FN(0.) - Q3FN(0.));
    int j2lim = ftrunc(HWF2 / DXF3 + 1.001);
    for (int jj = 2; jj <= j2lim; jj++) {
        double x = (double)(jj - 1) * DXF3, xsq = x * x;
        F3[jj] = RECPI * (Q2FN(xsq) - Q3FN(xsq));
    }
    for (int jj = j2lim + 1; jj <= NX3; jj++) {
        double x = (double)(jj - 1) * DXF3, xsq = x * x;
        F3[jj] = RECPI * (XLORNZ(xsq) - Q3FN(xsq));
    }
#undef XLORNZ
#undef Q1FN
#undef Q2FN
#undef Q3FN
    /* SHAPEG */
    double FLN2 = log(2.);
    double FGNORM = sqrt(FLN2 * RECPI);
    FG[1] = FGNORM * exp(-FLN2 * 0.);
    for (int jj = 2; jj <= NX1; jj++) {
        double x = (double)(jj - 1) * DXF1, xsq = x * x;
        FG[jj] = FGNORM * exp(-(FLN2 * xsq));
    }
    FG[NX1] = 0.;
    /* VERFN */
    const double CEXP = 0.45, CE0 = 1., CE2 = -.20737285249, CE4 = -.00872684335747;
    double SE0 = sqrt(CEXP / PI_REF);
    double AE2 = 2. * CEXP;
    double AE4 = AE2 * AE2 / 3.;
    for (int i = 1; i <= NFMX; i++) {
        double z = DXF1 * (double)(i - 1), z2 = z * z;
        double xe0 = SE0 * exp(-(CEXP * z2));
        XVER[i] = (1. / (CE0 + CE2 + CE4)) * (CE0 + CE2 * AE2 * z2 + CE4 * AE4 * z2 * z2) * xe0;
    }
}

/* Zeta tables, clblm_ODLAY_CommonSub.f90:43-185 (1-based accessors). */
double orc_avrat(int iz) { return CLBLM_AVRAT[iz - 1]; }
#define AVRAT(i) CLBLM_AVRAT[(i)-1]
#define CGAUSS(i) CLBLM_CGAUSS[(i)-1]
#define CF1(i) CLBLM_CF1[(i)-1]
#define CF2(i) CLBLM_CF2[(i)-1]
#define CF3(i) CLBLM_CF3[(i)-1]
#define CER(i) CLBLM_CER[(i)-1]

/* ---------------------------------------------------------------------------
 * TIPS_2003 + AtoB, clblm_ODLAY_TIPS.f90:78-420, :5451-5547.
 * ------------------------------------------------------------------------- */
static double atob(double aa, const double *A, const double *B, int npt) {
    /* A, B 1-based views */
    double bb = 0.;
    for (int i = 2; i <= npt; i++) {
        if (A[i] >= aa) {
            if (i < 3 || i == npt) {
                int j = i;
                if (i < 3) j = 3;
                if (i == npt) j = npt;
                double a0d1 = A[j - 2] - A[j - 1]; if (a0d1 == 0.) a0d1 = 0.0001;
                double a0d2 = A[j - 2] - A[j];     if (a0d2 == 0.) a0d2 = 0.0001;
                double a1d1 = A[j - 1] - A[j - 2]; if (a1d1 == 0.) a1d1 = 0.0001;
                double a1d2 = A[j - 1] - A[j];     if (a1d2 == 0.) a1d2 = 0.0001;
                double a2d1 = A[j] - A[j - 2];     if (a2d1 == 0.) a2d1 = 0.0001;
                double a2d2 = A[j] - A[j - 1];     if (a2d2 == 0.) a2d2 = 0.0001;
                double a0 = (aa - A[j - 1]) * (aa - A[j]) / (a0d1 * a0d2);
                double a1 = (aa - A[j - 2]) * (aa - A[j]) / (a1d1 * a1d2);
                double a2 = (aa - A[j - 2]) * (aa - A[j - 1]) / (a2d1 * a2d2);
                bb = a0 * B[j - 2] + a1 * B[j - 1] + a2 * B[j];
            } else {
                int j = i;
                double a0d1 = A[j - 2] - A[j - 1]; if (a0d1 == 0.) a0d1 = 0.0001;
                double a0d2 = A[j - 2] - A[j];     if (a0d2 == 0.) a0d2 = 0.0001;
                double a0d3 = A[j - 2] - A[j + 1]; if (a0d3 == 0.) a0d3 = 0.0001;
                double a1d1 = A[j - 1] - A[j - 2]; if (a1d1 == 0.) a1d1 = 0.0001;
                double a1d2 = A[j - 1] - A[j];     if (a1d2 == 0.) a1d2 = 0.0001;
                double a1d3 = A[j - 1] - A[j + 1]; if (a1d3 == 0.) a1d3 = 0.0001;
                double a2d1 = A[j] - A[j - 2];     if (a2d1 == 0.) a2d1 = 0.0001;
                double a2d2 = A[j] - A[j - 1];     if (a2d2 == 0.) a2d2 = 0.0001;
                double a2d3 = A[j] - A[j + 1];     if (a2d3 == 0.) a2d3 = 0.0001;
                double a3d1 = A[j + 1] - A[j - 2]; if (a3d1 == 0.) a3d1 = 0.0001;
                double a3d2 = A[j + 1] - A[j - 1]; if (a3d2 == 0.) a3d2 = 0.0001;
                double a3d3 = A[j + 1] - A[j];     if (a3d3 == 0.) a3d3 = 0.0001;
                double a0 = (aa - A[j - 1]) * (aa - A[j]) * (aa - A[j + 1]);
                a0 = a0 / (a0d1 * a0d2 * a0d3);
                double a1 = (aa - A[j - 2]) * (aa - A[j]) * (aa - A[j + 1]);
                a1 = a1 / (a1d1 * a1d2 * a1d3);
                double a2 = (aa - A[j - 2]) * (aa - A[j - 1]) * (aa - A[j + 1]);
                a2 = a2 / (a2d1 * a2d2 * a2d3);
                double a3 = (aa - A[j - 2]) * (aa - A[j - 1]) * (aa - A[j]);
                a3 = a3 / (a3d1 * a3d2 * a3d3);
                bb = a0 * B[j - 2] + a1 * B[j - 1] + a2 * B[j] + a3 * B[j + 1];
            }
            return bb;
        }
    }
    return bb;
}

static int qt(int mol, int iso, double t, double *q) {
    if (t < 70. || t > 3000.) return fail("TIPS: temperature out of range 70-3000 K");
    const double *row = CLBLM_TIPS_Q + (size_t)(CLBLM_TIPS_ROW[mol - 1] + iso - 1) * CLBLM_NT;
    double A[CLBLM_NT + 2], B[CLBLM_NT + 2];
    A[0] = B[0] = A[CLBLM_NT + 1] = B[CLBLM_NT + 1] = 0.;
    for (int k = 0; k < CLBLM_NT; k++) { A[k + 1] = CLBLM_TIPS_TDAT[k]; B[k + 1] = row[k]; }
    *q = atob(t, A, B, CLBLM_NT);
    return 0;
}

int orc_tips_scor(int mol, int iso, double t, double *scor) {
    if (mol < 1 || mol > CLBLM_NMOL_DATA) return fail("TIPS: molecule without data (1..7 supported)");
    if (iso < 1 || iso > CLBLM_MAXISO[mol - 1]) return fail("TIPS: isotopologue out of range");
    double q296, qt_;
    if (qt(mol, iso, 296., &q296)) return -1;
    if (qt(mol, iso, t, &qt_)) return -1;
    if (qt_ <= 0. || q296 <= 0.) return fail("TIPS: partition sum less than 0");
    *scor = q296 / qt_;
    return 0;
}

/* MOLEC(IND=1 then IND=2), clblm_ODLAY_CommonSub.f90:201-264. */
int orc_molec(const orc_layer *L, double *scor, double *rhoslf, double *alfd1) {
    double FLN2 = log(2.);
    double FAD = FLN2 * 2. * AVOGAD * BOLTZ / (CLIGHT * CLIGHT);
    double RHORAT = (L->pave / P0) * (TEMP0 / L->tave);
    for (int m = 1; m <= MXMOL; m++) {
        if (!L->sel[m]) continue;
        if (m > CLBLM_NMOL_DATA) return fail("molecule without TIPS/mass data (1..7 supported)");
        for (int iso = 1; iso <= CLBLM_MAXISO[m - 1]; iso++) {
            if (orc_tips_scor(m, iso, L->tave, &scor[m * 11 + iso])) return -1;
            rhoslf[m] = RHORAT * L->w[m] / L->wtot;
            alfd1[m * 11 + iso] = sqrt(FAD * L->tave / CLBLM_ISO_MASS[CLBLM_ISO_OFFSET[m - 1] + iso - 1]);
        }
    }
    return 0;
}

/* RADFN, clblm_ODLAY_CommonSub.f90:293-328. */
double orc_radfn(double vi, double xkt) {
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

/* RADFNI, clblm_ODLAY_CommonSub.f90:379-525. */
double orc_radfni(double vi, double dvi, double xkt, double *vinew, double *rdel, double *rdlast) {
    const double FACT1 = 3.0E-03;
    double xvi = vi;
    if (*rdlast < 0.) *rdlast = orc_radfn(vi, xkt);
    double radfni = *rdlast;
    int intvls = 1;
    double rdnext;
    if (xkt > 0.0) {
        double xviokt = xvi / xkt;
        if (xviokt <= 0.01) {
            if (*vinew >= 0.0) {
                *vinew = vi + FACT1 * 0.5 * xvi;
                intvls = ftrunc((*vinew - vi) / dvi);
                if (intvls < 1) intvls = 1;
                *vinew = vi + dvi * (double)intvls;
            } else {
                *vinew = fabs(*vinew);
                intvls = ftrunc((*vinew - vi) / dvi);
                if (intvls < 1) intvls = 1;
            }
            double xvinew = *vinew;
            double xvinewokt = xvinew / xkt;
            rdnext = 0.5 * xvinewokt * xvinew;
        } else if (xviokt <= 10.0) {
            double expvkt = exp(-xviokt);
            double xminus = 1. - expvkt;
            double xplus = 1. + expvkt;
            if (*vinew >= 0.0) {
                double cvikt = xviokt * expvkt;
                *vinew = vi + FACT1 * xvi / (1. + (cvikt / xminus + cvikt / xplus));
                intvls = ftrunc((*vinew - vi) / dvi);
                if (intvls < 1) intvls = 1;
                *vinew = vi + dvi * (double)intvls;
            } else {
                *vinew = fabs(*vinew);
                intvls = ftrunc((*vinew - vi) / dvi);
                if (intvls < 1) intvls = 1;
            }
            double xvinew = *vinew;
            double xvinewokt = xvinew / xkt;
            double e = exp(-xvinewokt);
            rdnext = xvinew * (1. - e) / (1. + e);
        } else {
            if (*vinew >= 0.0) {
                *vinew = vi + (FACT1 * xvi);
                intvls = ftrunc((*vinew - vi) / dvi);
                if (intvls < 1) intvls = 1;
                *vinew = vi + dvi * (double)intvls;
            } else {
                *vinew = fabs(*vinew);
                intvls = ftrunc((*vinew - vi) / dvi);
                if (intvls < 1) intvls = 1;
            }
            rdnext = *vinew;
        }
    } else {
        if (*vinew >= 0.0) {
            *vinew = vi + (FACT1 * xvi);
            intvls = ftrunc((*vinew - vi) / dvi);
            if (intvls < 1) intvls = 1;
            *vinew = vi + dvi * (double)intvls;
        } else {
            *vinew = fabs(*vinew);
            intvls = ftrunc((*vinew - vi) / dvi);
            if (intvls < 1) intvls = 1;
        }
        rdnext = *vinew;
    }
    *rdel = (rdnext - radfni) / (double)intvls;
    *rdlast = rdnext;
    return radfni;
}

/* XINT, clblm_Spectrum.f90:874-914.  a1 and r3_1 are 1-based views. */
void orc_xint(double v1a, double v2a, double dva, const double *A, double afact,
              double vft, double dvr3, double *R3, int n1r3, int n2r3) {
    double recdva = 1. / dva;
    int ilo = ftrunc((v1a + dva - vft) / dvr3 + 1. + ONEMI);
    if (ilo < n1r3) ilo = n1r3;
    int ihi = ftrunc((v2a - dva - vft) / dvr3 + ONEMI);
    if (ihi > n2r3) ihi = n2r3;
    for (int i = ilo; i <= ihi; i++) {
        double vi = vft + dvr3 * (double)(i - 1);
        int j = ftrunc((vi - v1a) * recdva + ONEPL);
        double vj = v1a + dva * (double)(j - 1);
        double p = recdva * (vi - vj);
        double c = (3. - 2. * p) * p * p;
        double b = 0.5 * p * (1. - p);
        double b1 = b * (1. - p);
        double b2 = b * p;
        double conti = -A[j - 1] * b1 + A[j] * (1. - c + b2) + A[j + 1] * (c + b1) - A[j + 2] * b2;
        R3[i] = R3[i] + conti * afact;
    }
}

/* calculateDV, clblm_setDV.f90:221-271 (one layer). */
double orc_calculate_dv(double vbar, double pave, double tave, double frh2o, double alfal0, double sample) {
    double alfcor = (pave / P0) * sqrt(TEMP0 / tave);
    double h2oslf = (1. - frh2o + 5. * frh2o);
    double albar = alfal0 * alfcor * h2oslf;
    double adbar = 3.58115E-07 * vbar * sqrt(tave / 36.0);
    double avbar = 0.5 * (albar + sqrt(albar * albar + 4. * adbar * adbar));
    return avbar / sample;
}

/* interpSpectrum_array2array, clblm_Spectrum.f90:326-590. */
void orc_interp(const double *spIn, double DVin, double V1in, int NLIMin,
                double *spOut, double DVout, double V1out, int NLIMout) {
    const double TOL = 1.e-8;
    const int NUMCOF = 201;
    const double EPS = 2.220446049250313e-16;
    if (fabs(V1in - V1out) < EPS && fabs(DVin - DVout) < EPS && NLIMin == NLIMout) {
        memcpy(spOut, spIn, sizeof(double) * NLIMin);
        return;
    }
    static double A1[203], A2[203], A3[203], A4[203];
    double XNUMCF = (double)NUMCOF;
    for (int ip = 0; ip <= NUMCOF + 1; ip++) {
        double p = ((double)ip - 1.0) / (XNUMCF - 1.0);
        double pp = p * p;
        A1[ip] = -p / 2.0 * ((1 - p) * (1 - p));
        A2[ip] = 1.0 - pp * (3.0 - 2.0 * p) + pp / 2.0 * (1.0 - p);
        A3[ip] = pp * (3.0 - 2.0 * p) + p / 2.0 * ((1.0 - p) * (1.0 - p));
        A4[ip] = -pp / 2 * (1.0 - p);
    }
    double V2out = V1out + (double)(NLIMout - 1) * DVout;
    int I1out = 1, I2out = NLIMout;
    int sameV1 = fabs(V1in - V1out) < EPS;
    double dvRatio = DVout / DVin;
    if (fabs(DVout - DVin) < TOL) dvRatio = 1.;
    int lenIn, lenOut, ITYPE;
    static const double hi2lo[5] = {2. / 1., 3. / 2., 4. / 3., 5. / 4., 6. / 5.};
    static const double lo2hi[5] = {1. / 2., 2. / 3., 3. / 4., 4. / 5., 5. / 6.};
    int anyhi = 0, anylo = 0;
    for (int k = 0; k < 5; k++) {
        if (fabs(dvRatio - hi2lo[k]) < TOL) anyhi = 1;
        if (fabs(dvRatio - lo2hi[k]) < TOL) anylo = 1;
    }
    if (sameV1 && fabs(dvRatio - 1.) < EPS) {
        lenIn = 1; lenOut = 1; ITYPE = 0;
    } else if (sameV1 && anyhi) {
        lenIn = ftrunc(DVout / (DVout - DVin) + 0.5);
        lenOut = lenIn - 1;
        ITYPE = lenOut - 1;
    } else if (sameV1 && anylo) {
        lenIn = ftrunc(DVout / (DVin - DVout) + 0.5);
        lenOut = lenIn + 1;
        ITYPE = lenOut - 1;
    } else {
        lenIn = 0; lenOut = 0; ITYPE = -1;
    }
    /* spInExt(1-4 : NLIMin+4); an output grid reaching further past the input reads
     * beyond it in the reference (undefined there): zeros here, as on the GPU */
    int64_t reach = (int64_t)((V2out - V1in) / DVin) + 8;
    if (reach < NLIMin) reach = NLIMin;
    double *ext = (double *)calloc((size_t)reach + 9, sizeof(double));
    double *X = ext + 4; /* X[i] == spInExt(i) for i in -3..NLIMin+4 */
    for (int i = 1; i <= NLIMin; i++) X[i] = spIn[i - 1];
    X[0] = 2.0 * X[1] - X[2];
    X[-1] = X[0];
    X[NLIMin + 1] = 2.0 * X[NLIMin] - X[NLIMin - 1];
    X[NLIMin + 2] = X[NLIMin + 1];
    double *O = spOut - 1;
    if (ITYPE == 0) {
        int step = (int)lround(dvRatio);
        int iIn = 1 - step;
        for (int i = I1out; i <= I2out; i++) {
            iIn += step;
            O[i] = X[iIn];
        }
    } else if (ITYPE > 0) {
        double B1[12], B2[12], B3[12], B4[12];
        double AP = 1.0 / ((double)ITYPE + 1.0);
        for (int jp = 1; jp <= ITYPE; jp++) {
            int ip = jp + 1;
            double p = AP * (double)jp;
            if (DVin > DVout) p = 1.0 - p;
            B1[ip] = -p * (p - 1.0) * (p - 2.0) / 6.0;
            B2[ip] = (p * p - 1.0) * (p - 2.0) * 0.5;
            B3[ip] = -p * (p + 1.0) * (p - 2.0) * 0.5;
            B4[ip] = p * (p * p - 1.0) / 6.0;
        }
        for (int nl = 1; nl <= lenOut; nl++) {
            int iIn = nl - lenIn;
            if (nl > 1 && DVin > DVout) iIn = iIn - 1;
            if (nl == 1) {
                for (int i = I1out; i <= I2out; i += lenOut) {
                    iIn += lenIn;
                    O[i] = X[iIn];
                }
            } else {
                double a1n = B1[nl], a2n = B2[nl], a3n = B3[nl], a4n = B4[nl];
                for (int i = I1out + nl - 1; i <= I2out; i += lenOut) {
                    iIn += lenIn;
                    O[i] = a1n * X[iIn - 1] + a2n * X[iIn] + a3n * X[iIn + 1] + a4n * X[iIn + 2];
                }
            }
        }
    } else {
        double Vo = V1out;
        while (Vo < V2out + TOL) {
            int iIn = ftrunc((Vo - V1in) / DVin + 1.00001);
            double Vi = V1in + (double)(iIn - 1) * DVin;
            double P = (Vo - Vi) / DVin;
            int J1 = ftrunc((Vo - V1out) / DVout + 1.001) + I1out - 1;
            int J2 = (J1 + 2400 - 1 < I2out ? J1 + 2400 - 1 : I2out) + I1out - 1;
            for (int j = J1; j <= J2; j++) {
                int ip = ftrunc(P * XNUMCF + 1.00001);
                O[j] = A1[ip] * X[iIn - 1] + A2[ip] * X[iIn] + A3[ip] * X[iIn + 1] + A4[ip] * X[iIn + 2];
                P = P + dvRatio;
                if (P >= 1.0) {
                    iIn = iIn + ftrunc(P);
                    P = P - (double)ftrunc(P);
                }
            }
            Vo = V1out + (double)J2 * DVout;
        }
    }
    free(ext);
}

/* ---------------------------------------------------------------------------
 * F4 far-wing path.
 * ------------------------------------------------------------------------- */
typedef struct {
    double *vnu, *sp, *alfa, *epp, *spp;
    int *mol;
    int64_t n, cap;
    int64_t *blk_end; /* TAPE9 block boundaries: block b holds [blk_end[b-1], blk_end[b]) */
    int64_t nblk, blkcap;
} shrunk_t;

static void shr_push(shrunk_t *s, double vnu, double sp, double alfa, double epp, double spp, int mol) {
    if (s->n == s->cap) {
        s->cap = s->cap ? 2 * s->cap : 4096;
        s->vnu = realloc(s->vnu, s->cap * sizeof(double));
        s->sp = realloc(s->sp, s->cap * sizeof(double));
        s->alfa = realloc(s->alfa, s->cap * sizeof(double));
        s->epp = realloc(s->epp, s->cap * sizeof(double));
        s->spp = realloc(s->spp, s->cap * sizeof(double));
        s->mol = realloc(s->mol, s->cap * sizeof(int));
    }
    s->vnu[s->n] = vnu; s->sp[s->n] = sp; s->alfa[s->n] = alfa;
    s->epp[s->n] = epp; s->spp[s->n] = spp; s->mol[s->n] = mol;
    s->n++;
}

static void shr_end_block(shrunk_t *s) {
    if (s->nblk == s->blkcap) {
        s->blkcap = s->blkcap ? 2 * s->blkcap : 256;
        s->blk_end = realloc(s->blk_end, s->blkcap * sizeof(int64_t));
    }
    s->blk_end[s->nblk++] = s->n;
}

static void shr_free(shrunk_t *s) {
    free(s->vnu); free(s->sp); free(s->alfa); free(s->epp); free(s->spp); free(s->mol);
    free(s->blk_end);
}

/* SHRINQ, clblm_ODLAY_LineF4.f90:1516-1647, on buf[ilo..ihi] (1-based);
 * compacts the result into buf[ilo..] and returns the new count of the buffer. */
typedef struct { double vnu, sp, alfa, epp, spp; int mol; } sline;

static int shrinq(sline *B, int ilo, int ihi, double dvr4) {
    int j = ilo - 1;
    double dv = dvr4 / 2.;
    double vlmt = B[ilo].vnu + dv;
    double sumal = 0, sumad = 0, sums = 0, sumv = 0, sumc = 0;
    double sumal2 = 0, sumad2 = 0, sums2 = 0, sumv2 = 0, sumc2 = 0;
    for (int i = ilo; i <= ihi; i++) {
        if (B[i].sp < 1.0e-35) { B[i].sp = 0.0; B[i].spp = 0.0; }
        if (B[i].spp != 0.0) {
            j++;
            B[j] = B[i];
            if (B[j].mol != 2) B[j].mol = 0;
        } else if (B[i].mol != 2) {
            sumv = sumv + B[i].vnu * B[i].sp;
            sums = sums + B[i].sp;
            sumal = sumal + B[i].sp * B[i].alfa;
            sumad = sumad + B[i].sp * B[i].epp;
            sumc = sumc + B[i].spp;
        } else {
            sumv2 = sumv2 + B[i].vnu * B[i].sp;
            sums2 = sums2 + B[i].sp;
            sumal2 = sumal2 + B[i].sp * B[i].alfa;
            sumad2 = sumad2 + B[i].sp * B[i].epp;
            sumc2 = sumc2 + B[i].spp;
        }
        if (i < ihi && B[i + 1].vnu <= vlmt) continue;
        vlmt = B[i].vnu + dv;
        if (sums > 0.) {
            j++;
            B[j].sp = sums; B[j].alfa = sumal / sums; B[j].epp = sumad / sums;
            B[j].vnu = sumv / sums; B[j].spp = sumc; B[j].mol = 0;
            sumal = sumad = sums = sumv = sumc = 0.;
        }
        if (sums2 > 0.) {
            j++;
            B[j].sp = sums2; B[j].alfa = sumal2 / sums2; B[j].epp = sumad2 / sums2;
            B[j].vnu = sumv2 / sums2; B[j].mol = 2; B[j].spp = sumc2;
            sumal2 = sumad2 = sums2 = sumv2 = sumc2 = 0.;
        }
    }
    return j;
}

/* Line coupling.  The record after an IFLG 1 / 3 line (IFLG = -1) holds its
 * coefficients at TEMPLC = 200, 250, 296, 340 K: A(1..4) in its VNU, ALFA, MOL (the bits
 * of a REAL*4) and TMPALF fields, B(1..4) in its SP, EPP, HWHM and PSHIFT fields; A is
 * Y (IFLAG 1) or GAMMA1 (IFLAG 3), B is G or GAMMA2 (clblm_ODLAY.f90:1821-1829,
 * clblm_ODLAY_LineF4.f90:971-978). */
static const double TEMPLC[4] = {200.0, 250.0, 296.0, 340.0};

static void lc_nodes(const orc_lines *Ls, int64_t i, double A[5], double B[5]) {
    const int64_t k = i + 1;
    float am;
    int32_t mb = Ls->mol[k];
    memcpy(&am, &mb, sizeof am);
    A[1] = Ls->nu[k]; A[2] = Ls->a_air[k]; A[3] = am; A[4] = Ls->tmpalf[k];
    B[1] = Ls->s[k]; B[2] = Ls->e_lo[k]; B[3] = Ls->a_self[k]; B[4] = Ls->shift[k];
}

/* LNCORQ, clblm_ODLAY.f90:1660-1667 and 1833-1845 (1-based A, B). */
static void lc_lncorq(const orc_lines *Ls, int64_t i, double tave, double *a, double *b) {
    double A[5], B[5];
    lc_nodes(Ls, i, A, B);
    int ILC = 1;
    for (int IL = 1; IL <= 3; IL++) {
        ILC = IL;
        if (tave < TEMPLC[ILC + 1 - 1]) break;
    }
    if (ILC == 4) ILC = 3;
    double RECTLC = 1.0 / (TEMPLC[ILC + 1 - 1] - TEMPLC[ILC - 1]);
    double TMPDIF = tave - TEMPLC[ILC - 1];
    double SLOPEA = (A[ILC + 1] - A[ILC]) * RECTLC;
    double SLOPEB = (B[ILC + 1] - B[ILC]) * RECTLC;
    *a = A[ILC] + SLOPEA * TMPDIF;
    *b = B[ILC] + SLOPEB * TMPDIF;
}

/* LINF4Q, clblm_ODLAY_LineF4.f90:860-866 and 971-988. */
static void lc_linf4q(const orc_lines *Ls, int64_t i, double tave, double *a, double *b) {
    double A[5], B[5];
    lc_nodes(Ls, i, A, B);
    int ILC;
    for (ILC = 1; ILC <= 4; ILC++)
        if (tave <= TEMPLC[ILC - 1]) break;
    if (ILC == 1) ILC = 2;
    if (ILC == 5) ILC = 4;
    double RECTLC = 1.0 / (TEMPLC[ILC - 1] - TEMPLC[ILC - 1 - 1]);
    double TMPDIF = tave - TEMPLC[ILC - 1];
    double SLOPEY = (A[ILC] - A[ILC - 1]) * RECTLC;
    double SLOPEG = (B[ILC] - B[ILC - 1]) * RECTLC;
    *a = A[ILC] + SLOPEY * TMPDIF;
    *b = B[ILC] + SLOPEG * TMPDIF;
}

/* The record bounds VMIN / VMAX of the TAPE3 record header: the centres of its first
 * and last lines (records with IFLG < 0 are not lines). */
static double rec_vmin(const orc_lines *Ls, int64_t a, int64_t b) {
    for (int64_t i = a; i < b; i++)
        if (Ls->iflg[i] >= 0) return Ls->nu[i];
    return -1e300;
}
static double rec_vmax(const orc_lines *Ls, int64_t a, int64_t b) {
    for (int64_t i = b - 1; i >= a; i--)
        if (Ls->iflg[i] >= 0) return Ls->nu[i];
    return -1e300;
}

/* Layer-wide constants shared by LINF4Q and LNCORQ. */
typedef struct {
    double scor[48 * 11], rhoslf[48], alfd1[48 * 11];
} molec_t;

/* PSHIFT as RDLIN / RDLNFL read it (clblm_ODLAY.f90:2730-2738,
 * clblm_ODLAY_LineF4.f90:1463-1468): an O2 line with its own species data. */
static double read_pshift(const orc_lines *Ls, int64_t i) {
    double pshift = Ls->shift[i];
    int m = Ls->mol[i] % 100;
    if (m == 7 && Ls->iflg[i] >= 0 && Ls->brd_flg && Ls->brd_flg[7 * i + 7 - 1] > 0) {
        double rvmr = 0.21;
        pshift = (pshift - rvmr * (double)Ls->brd_shft[7 * i + 7 - 1]) / (1.0 - rvmr);
    }
    return pshift;
}

/* sum(brd_mol_flg(:,i)) > 0 .and. speciesBroad */
static int brd_active(const orc_lines *Ls, int64_t i, int species_broad) {
    if (!species_broad || !Ls->brd_flg) return 0;
    int t = 0;
    for (int j = 1; j <= 7; j++) t += Ls->brd_flg[7 * i + j - 1];
    return t > 0;
}

/* vnu(i) + sum(rhoslf(1:7) * brd_mol_flg(:,i) * (brd_mol_shft(:,i) - pshift(i)))
 * (clblm_ODLAY.f90:1853-1856, clblm_ODLAY_LineF4.f90:996-999) */
static double brd_shift(const orc_lines *Ls, int64_t i, const molec_t *M, double vnu, double pshift) {
    double s = 0.;
    for (int j = 1; j <= 7; j++)
        s = s + (M->rhoslf[j] * (double)Ls->brd_flg[7 * i + j - 1]) * ((double)Ls->brd_shft[7 * i + j - 1] - pshift);
    return vnu + s;
}

/* ALFL with species broadening, clblm_ODLAY.f90:1866-1883 and
 * clblm_ODLAY_LineF4.f90:1038-1054. */
static double brd_width(const orc_lines *Ls, int64_t i, const molec_t *M, int m, double RHORAT, double TRATIO,
                        double alfa0i, double hwhmsi, double alfl, int species_broad) {
    if (brd_active(Ls, i, species_broad)) {
        double tmpcor_arr[8], alfa_tmp[8];
        for (int j = 1; j <= 7; j++) {
            tmpcor_arr[j] = pow(TRATIO, (double)Ls->brd_tmp[7 * i + j - 1]);
            alfa_tmp[j] = (double)Ls->brd_hw[7 * i + j - 1] * tmpcor_arr[j];
        }
        double alfsum = 0., rsum = 0.;
        for (int j = 1; j <= 7; j++) alfsum = alfsum + (M->rhoslf[j] * (double)Ls->brd_flg[7 * i + j - 1]) * alfa_tmp[j];
        for (int j = 1; j <= 7; j++) rsum = rsum + M->rhoslf[j] * (double)Ls->brd_flg[7 * i + j - 1];
        alfl = (RHORAT - rsum) * alfa0i + alfsum;
        if (Ls->brd_flg[7 * i + m - 1] == 0) alfl = alfl + M->rhoslf[m] * (hwhmsi - alfa0i);
    }
    if (species_broad && Ls->brd_flg && m <= 7 && Ls->brd_flg[7 * i + m - 1] == -1) alfl = hwhmsi;
    return alfl;
}

static int sel_ok(const orc_layer *L, int m, int iso) {
    if (m < 1 || m > MXMOL || !L->sel[m]) return 0;
    if (m > CLBLM_NMOL_DATA) return 0;
    if (iso > CLBLM_MAXISO[m - 1]) return 0;
    return 1;
}

/* O2 / N2 air-width rescale applied on read: RDLIN clblm_ODLAY.f90:2728-2747,
 * RDLNFL clblm_ODLAY_LineF4.f90:1462-1469. */
static double read_alfa(const orc_lines *Ls, int64_t i) {
    double alfa = Ls->a_air[i];
    int m = Ls->mol[i] % 100;
    if (m == 7 && Ls->iflg[i] >= 0) {
        double rvmr = 0.21;
        alfa = (alfa - rvmr * (double)Ls->a_self[i]) / (1.0 - rvmr);
    }
    if (m == 22 && Ls->iflg[i] >= 0) {
        double rvmr = 0.79;
        alfa = (alfa - rvmr * (double)Ls->a_self[i]) / (1.0 - rvmr);
    }
    return alfa;
}

/* LINF4Q + RDLNFL + SHRINQ, clblm_ODLAY_LineF4.f90:647-1263,1355-1498,1516-1647.
 * Records of 250 lines; the shrink buffer holds LIMIN=1000 lines (+250 slack). */
static int linf4q(const orc_lines *Ls, const orc_layer *L, const molec_t *M,
                  double v1l4, double v2l4, double dvr4, int species_broad, shrunk_t *out) {
    const int LIMIN = 1000;
    double PAVP0 = L->pave / P0;
    double PAVP2 = PAVP0 * PAVP0;
    double TRATIO = L->tave / TEMP0;
    double RHORAT = (L->pave / P0) * (TEMP0 / L->tave);
    double BETA = RADCN2 / L->tave;
    double BETA0 = RADCN2 / TEMP0;
    double BETACR = BETA - BETA0;
    double VLO = v1l4, VHI = v2l4;
    sline *B = (sline *)malloc(sizeof(sline) * 1252);
    int IJ = 0, ILO = 1, IHI = 0;
    int ieof = 0;
    int64_t nrec = (Ls->n + 249) / 250;
    int64_t rec = 0;
    for (;;) {
        /* 30: read one record (RDLNFL) */
        int got = 0;
        while (rec < nrec) {
            int64_t a = rec * 250, b = a + 250 < Ls->n ? a + 250 : Ls->n;
            rec++;
            if (rec_vmax(Ls, a, b) < VLO) continue; /* VMAX < VLO: skip record */
            got = 1;
            for (int64_t i = a; i < b; i++) {
                if (Ls->iflg[i] < 0) continue;
                const int IFLAG = Ls->iflg[i];
                int m = Ls->mol[i] % 100;
                int iso = (Ls->mol[i] % 1000) / 100;
                if (iso == 0) iso = 10;
                if (!sel_ok(L, m, iso)) continue;
                double sui = (double)Ls->s[i] * L->w[m];
                if (sui == 0.) continue;
                if (Ls->nu[i] < VLO) continue;
                IJ++;
                double YI = 0., GI = 0., GAMMA1 = 0., GAMMA2 = 0.;
                if (IFLAG == 1) lc_linf4q(Ls, i, L->tave, &YI, &GI);
                if (IFLAG == 3) lc_linf4q(Ls, i, L->tave, &GAMMA1, &GAMMA2);
                sline *S = &B[IJ];
                double pshift = read_pshift(Ls, i);
                S->vnu = Ls->nu[i] + RHORAT * pshift;
                if (brd_active(Ls, i, species_broad)) S->vnu = brd_shift(Ls, i, M, S->vnu, pshift);
                S->alfa = read_alfa(Ls, i);
                S->epp = Ls->e_lo[i];
                S->mol = m;
                if (S->vnu == 0.) sui = 2. * sui;
                if (S->epp <= -1.001) S->epp = fabs(S->epp);
                if (S->epp > -0.999)
                    sui = sui * M->scor[m * 11 + iso] * exp(-(S->epp * BETACR)) * (1. + exp(-(S->vnu * BETA)));
                double tmpcor = pow(TRATIO, (double)Ls->tmpalf[i]);
                double alfa0i = S->alfa * tmpcor;
                double hwhmsi = (double)Ls->a_self[i] * tmpcor;
                S->alfa = alfa0i * (RHORAT - M->rhoslf[m]) + hwhmsi * M->rhoslf[m];
                S->alfa = brd_width(Ls, i, M, m, RHORAT, TRATIO, alfa0i, hwhmsi, S->alfa, species_broad);
                if (IFLAG == 3) S->alfa = S->alfa * (1.0 - GAMMA1 * PAVP0 - GAMMA2 * PAVP2);
                S->epp = S->vnu * M->alfd1[m * 11 + iso];
                S->sp = sui * (1. + GI * PAVP2);
                S->spp = sui * YI * PAVP0;
                if (S->vnu > VHI) { ieof = 1; break; }
            }
            break;
        }
        if (!got && !ieof) ieof = 2; /* end of file */
        if (ieof == 0 && IJ < LIMIN) continue;
        /* 60 */
        IHI = IJ;
        if (IHI >= ILO) IHI = shrinq(B, ILO, IHI, dvr4);
        else IHI = ILO - 1;
        ILO = IHI + 1;
        IJ = ILO - 1;
        if (IHI < LIMIN && ieof == 0) continue;
        if (IHI > 0) {
            for (int k = 1; k <= IHI; k++)
                shr_push(out, B[k].vnu, B[k].sp, B[k].alfa, B[k].epp, B[k].spp, B[k].mol);
            shr_end_block(out);
        }
        if (ieof) break;
        IJ = 0;
        ILO = 1;
    }
    free(B);
    return 0;
}

static const double ZBND = 64.;

/* LBLF4Q + CNVF4Q, clblm_ODLAY_LineF4.f90:25-184,187-494.  R4 is 1-based with
 * 2502 slots; returns NPTR4 and updates *v2r4. */
static int lblf4q(int jrad, double v1r4, double *v2r4, double *R4, double dvr4, double bound4,
                  double tave, shrunk_t *sh) {
    int LIMOUT = 2500;
    int nptr4 = ftrunc((*v2r4 - v1r4) / dvr4 + ONEPL);
    if (nptr4 > LIMOUT) nptr4 = LIMOUT;
    *v2r4 = v1r4 + dvr4 * (double)(nptr4 - 1);
    for (int i = 1; i <= LIMOUT + 2; i++) R4[i] = 0.;
    double BETA = RADCN2 / tave;
    double vnulst = *v2r4 + bound4;
    double RECPI = 1. / (2. * asin(1.));
    double ZSQBND = ZBND * ZBND;
    double A3 = (1. + 2. * ZSQBND) / ((1. + ZSQBND) * (1. + ZSQBND));
    double B3 = -1. / ((1. + ZSQBND) * (1. + ZSQBND));
    double BNDSQ = bound4 * bound4;
    double r_bndsq = 1. / BNDSQ;
    double VLO = v1r4 - bound4;
    int stop = 0;
    for (int64_t blk = 0; blk < sh->nblk && !stop; blk++) {
      int64_t b0 = blk ? sh->blk_end[blk - 1] : 0, b1 = sh->blk_end[blk];
      if (sh->vnu[b1 - 1] < VLO) continue; /* RDLN4Q: VMAX < VLO, skip block */
      for (int64_t i = b0; i < b1; i++) {
        if (sh->vnu[i] >= vnulst) { stop = 1; break; }
        double sp = sh->sp[i], spp = sh->spp[i];
        if (sp == 0. && spp == 0.) continue;
        double alfadi = sh->epp[i], alfali = sh->alfa[i];
        double zetai = alfali / (alfali + alfadi);
        int iz = ftrunc(100. * zetai + ONEPL);
        double zetdif = 100. * zetai - (double)(iz - 1);
        double alfavi = (AVRAT(iz) + zetdif * (AVRAT(iz + 1) - AVRAT(iz))) * (alfali + alfadi);
        double ralfvi = 1. / alfavi;
        double sil = sp * RECPI * alfali;
        double siv = (alfali * ralfvi) * sp * RECPI * ralfvi;
        double speak = A3 * fabs(siv);
        if (speak <= 0. && spp == 0.) continue; /* DPTMIN = DPTFAC = 0 */
        double vnui = sh->vnu[i];
        for (;;) {
            double xnui = vnui - v1r4;
            int jmin = ftrunc((xnui - bound4) / dvr4 + 2.);
            if (jmin > nptr4) break; /* GO TO 60 */
            if (jmin < 1) jmin = 1;
            int jmax = ftrunc((xnui + bound4) / dvr4 + 1.);
            if (jmax >= jmin) {
                if (jmax > nptr4) jmax = nptr4;
                double alfli2 = alfali * alfali;
                double alfvi2 = alfavi * alfavi;
                double xjj = (double)(jmin - 1) * dvr4;
                double f4bnd = sil / (alfli2 + BNDSQ);
                double dptrat = spp / (sp * alfavi);
                double rec_alfvi2 = 1. / alfvi2;
                double siv_a3 = siv * A3, siv_b3 = siv * B3;
                int sppz = (spp == 0.);
                int co2 = (sh->mol[i] == 2);
                for (int jj = jmin; jj <= jmax; jj++) {
                    double xm = xjj - xnui;
                    double xmsq = xm * xm;
                    double zvsq = xmsq * rec_alfvi2;
                    double f4fn;
                    if (co2) {
                        double fcnt = (2. - (xmsq * r_bndsq)) * f4bnd;
                        if (zvsq <= ZSQBND) f4fn = (siv_a3 + zvsq * siv_b3) - fcnt;
                        else f4fn = sil / (alfli2 + xmsq) - fcnt;
                    } else {
                        if (zvsq <= ZSQBND) f4fn = (siv_a3 + zvsq * siv_b3) - f4bnd;
                        else f4fn = sil / (alfli2 + xmsq) - f4bnd;
                    }
                    if (sppz) R4[jj] = R4[jj] + f4fn;
                    else {
                        f4fn = f4fn + xm * dptrat * f4fn;
                        R4[jj] = R4[jj] + f4fn;
                    }
                    xjj = xjj + dvr4;
                }
            }
            /* 50: Van Vleck-Weisskopf mirror */
            if (vnui > 0. && vnui <= 25.) {
                vnui = -sh->vnu[i];
                spp = -spp;
                continue;
            }
            break;
        }
      }
    }
    if (jrad == 1) {
        double xkt = 1. / BETA;
        double vitst = v1r4 - dvr4, rdlast = -1., rdel;
        int n1 = 0, n2 = 0;
        do {
            n1 = n2 + 1;
            double vi = v1r4 + dvr4 * (double)(n1 - 1);
            double radvi = orc_radfni(vi, dvr4, xkt, &vitst, &rdel, &rdlast);
            n2 = ftrunc((vitst - v1r4) / dvr4 + 0.001);
            if (n2 > nptr4) n2 = nptr4;
            double t = radvi;
            for (int i = n1; i <= n2; i++) { R4[i] = R4[i] * t; t = t + rdel; }
        } while (n2 < nptr4);
    }
    return nptr4;
}

/* ---------------------------------------------------------------------------
 * HIRACQ line path.
 * ------------------------------------------------------------------------- */
typedef struct {
    double vnu, sp, sppsp, recalf, zetai, srad;
    int izeta;
} hline;

int orc_odlay(const orc_lines *Ls, const orc_layer *L, orc_params *P,
              double *out, int64_t cap, int64_t *nsamp, double *dv_used) {
    const double V1 = P->v1, V2 = P->v2;
    /* ODLAY_nonStruct clblm_ODLAY.f90:129-611 */
    double DV = P->dvnormal;
    double ALFAV = P->sample * (P->dvnormal > P->dvnarrow ? P->dvnormal : P->dvnarrow);
    if (!(fabs(P->dvnormal - P->dvnarrow) / P->dvnormal > 0.2)) DV = P->dvnarrow;
    const double TAVE = L->tave, PAVE = L->pave;
    /* OPDPTH clblm_ODLAY.f90:757-967 */
    double XKT = TAVE / RADCN2;
    if ((V1 / XKT) < 5. && P->jrad != -1) P->jrad = 0;
    int JRAD = P->jrad;
    int V1ABS_i = 0;
    double V1ABS = 0., V2ABS = 0., DVABS = 0.;
    int NPTABS = 0;
    if (P->contnm) {
        DVABS = 1.;
        V1ABS = (double)ftrunc(V1);
        if (V1 < 0.) V1ABS = V1ABS - 1.;
        V1ABS = V1ABS - 3. * DVABS;
        V2ABS = (double)ftrunc(V2 + 3. * DVABS + 0.5);
        NPTABS = ftrunc((V2ABS - V1ABS) / DVABS + 1.5);
        if (P->nptabs != NPTABS) {
            snprintf(g_err, sizeof g_err, "continuum array must have %d points from %g", NPTABS, V1ABS);
            return -1;
        }
        (void)V1ABS_i;
    }
    double SAMPLE = P->sample;
    double ALFAV4 = 64. * SAMPLE * DV;
    double DVR4 = ALFAV4 / SAMPLE;
    double BOUND4 = 25.;
    double V1R4 = V1 - 2. * DVR4;
    double V2R4 = V2 + 2. * DVR4;
    double V1L4 = V1R4 - BOUND4 - DVR4;
    double V2L4 = V2R4 + BOUND4 + 2 * DVR4;

    molec_t M;
    memset(&M, 0, sizeof M);
    if (orc_molec(L, M.scor, M.rhoslf, M.alfd1)) return -1;

    shrunk_t sh = {0};
    if (linf4q(Ls, L, &M, V1L4, V2L4, DVR4, P->species_broad, &sh)) { shr_free(&sh); return -1; }

    /* HIRACQ clblm_ODLAY.f90:1002-1562 */
    static double F1[NFMX + 1], F2[NFMX + 1], F3[NFMX + 1], FG[NFMX + 1], XVER[NFMX + 1];
    static int tables_done = 0;
    if (!tables_done) { orc_tables(F1, F2, F3, FG, XVER); tables_done = 1; }
    const double HWF1 = 4., HWF3 = 64., DXF1 = 0.002, DXF2 = 0.008, DXF3 = 0.032;
    const int NX1 = NFPTS, NX2 = NFPTS, NX3 = NFPTS;
    int NSHIFT = 32;
    double ALFMAX = 4. * ALFAV;
    int nALFMAX = ftrunc(ALFMAX / DV);
    int NBOUND = ftrunc((2. * HWF3) * (double)nALFMAX + 0.01);
    int NLIM1 = 2401, NLIM2 = (NLIM1 / 4) + 1, NLIM3 = (NLIM2 / 4) + 1;
    double DVR2 = (DXF2 / DXF1) * DV, DVR3 = (DXF3 / DXF1) * DV;
    int MAX1 = NSHIFT + NLIM1 + NSHIFT + NBOUND / 2;
    int MAX2 = MAX1 / 4, MAX3 = MAX1 / 16;
    MAX1 = MAX1 + 4 * nALFMAX;
    MAX2 = MAX2 + 16 * nALFMAX / 4 + 1;
    MAX3 = MAX3 + 64 * nALFMAX / 16 + 1;
    double BOUND = (double)NBOUND * DV / 2.;
    int NLO = NSHIFT + 1, NHI = NLIM1 + NSHIFT - 1;
    /* generous bounds; negative indices are write-only scratch as in the reference */
    int lb = 8 * MAX1 + 64;
    double *R1b = calloc((size_t)(MAX1 + 2 * lb), sizeof(double)), *R1 = R1b + lb;
    double *R2b = calloc((size_t)(MAX2 + 2 * lb), sizeof(double)), *R2 = R2b + lb;
    double *R3b = calloc((size_t)(MAX3 + 2 * lb), sizeof(double)), *R3 = R3b + lb;
    double R4[2503];
    hline *H = malloc(sizeof(hline) * 251);
    int ret = 0;

    double VFT = V1 - (double)NSHIFT * DV;
    double VBOT = V1 - BOUND, VTOP = V2 + BOUND;
    double V1R4ST = V1R4, V2R4ST = V2R4;
    (void)V1R4ST;
    int NPTR4 = lblf4q(JRAD, V1R4, &V2R4, R4, DVR4, BOUND4, TAVE, &sh);

    /* LNCORQ layer constants (NLNCR == 1) clblm_ODLAY.f90:1649-1680 */
    double XKT0 = TEMP0 / RADCN2;
    double BETACR = (1. / XKT) - (1. / XKT0);
    double TRATIO = TAVE / TEMP0;
    double RHORAT = (PAVE / P0) * (TEMP0 / TAVE);
    double PAVP0 = PAVE / P0;
    double PAVP2 = PAVP0 * PAVP0;
    double CLC1 = 4. / ((double)(NX1 - 1)), CLC2 = 16. / ((double)(NX2 - 1)), CLC3 = 64. / ((double)(NX3 - 1));
    double WAVDXF = DV / DXF1, HWDXF = HWF1 / DXF1;
    double CONF2 = DV / DVR2, CONF3 = DV / DVR3;

    int64_t numTot = 0;
    int IEOF = 0, IDATA = 0, ILO = 0, IHI = -999, IPANEL = 0, ISTOP = 0, IFPAN = 1;
    int64_t nrec = (Ls->n + 249) / 250, rec = 0;
    int nrec_lines = 0;
    int N1R2 = 1, N1R3 = 1, N2R3 = 155;

    for (;;) {
        /* 60: read a record if needed */
        int need_read = 1;
        while (need_read) {
            need_read = 0;
            if (IEOF != 0) break;
            /* RDLIN */
            int found = 0;
            int64_t a = 0, b = 0;
            while (rec < nrec) {
                a = rec * 250; b = a + 250 < Ls->n ? a + 250 : Ls->n;
                rec++;
                if (rec_vmax(Ls, a, b) < VBOT) continue;
                found = 1;
                break;
            }
            if (!found) { IEOF = 1; break; }
            IDATA = 0;
            int IJ = 0;
            int iout[251];
            nrec_lines = (int)(b - a);
            for (int64_t i = a; i < b; i++) {
                int k = (int)(i - a) + 1;
                if (Ls->iflg[i] >= 0) iout[++IJ] = k;
            }
            ILO = 1;
            if (rec_vmin(Ls, a, b) < VBOT) {
                for (int j = 1; j <= IJ; j++) {
                    ILO = j;
                    if (Ls->nu[a + iout[j] - 1] >= VBOT) break;
                }
            }
            for (int j = ILO; j <= IJ; j++) {
                int64_t i = a + iout[j] - 1;
                if (Ls->mol[i] > 0) {
                    IHI = j;
                    if (Ls->nu[i] > VTOP) break;
                }
            }
            if (IHI < IJ) IDATA = 1;
            /* LNCORQ on ILO..IHI */
            for (int j = ILO; j <= IHI; j++) {
                int64_t i = a + iout[j] - 1;
                hline *h = &H[j];
                h->sp = 0.; h->srad = 0.; h->sppsp = 0.;
                h->vnu = Ls->nu[i];
                int m = Ls->mol[i] % 100;
                int iso = (Ls->mol[i] % 1000) / 100;
                if (iso == 0) iso = 10;
                if (!sel_ok(L, m, iso)) continue;
                double sui = (double)Ls->s[i] * L->w[m];
                if (JRAD == 1) sui = sui * orc_radfn(h->vnu, XKT);
                if (sui == 0.) continue;
                const int IFLAG = Ls->iflg[i];
                double YI = 0., GI = 0., GAMMA1 = 0., GAMMA2 = 0.;
                if (IFLAG == 1) lc_lncorq(Ls, i, TAVE, &YI, &GI);
                if (IFLAG == 3) lc_lncorq(Ls, i, TAVE, &GAMMA1, &GAMMA2);
                double pshift = read_pshift(Ls, i);
                h->vnu = h->vnu + RHORAT * pshift;
                if (brd_active(Ls, i, P->species_broad)) h->vnu = brd_shift(Ls, i, &M, h->vnu, pshift);
                double tmpcor = pow(TRATIO, (double)Ls->tmpalf[i]);
                double alfa0i = read_alfa(Ls, i) * tmpcor;
                double hwhmsi = (double)Ls->a_self[i] * tmpcor;
                double alfl = alfa0i * (RHORAT - M.rhoslf[m]) + hwhmsi * M.rhoslf[m];
                alfl = brd_width(Ls, i, &M, m, RHORAT, TRATIO, alfa0i, hwhmsi, alfl, P->species_broad);
                if (IFLAG == 3) alfl = alfl * (1.0 - GAMMA1 * PAVP0 - GAMMA2 * PAVP2);
                double alfad = h->vnu * M.alfd1[m * 11 + iso];
                double zeta = alfl / (alfl + alfad);
                h->zetai = zeta;
                double fzeta = 100. * zeta;
                int iz = ftrunc(fzeta + ONEPL);
                h->izeta = iz;
                double zetdif = fzeta - (double)(iz - 1);
                double alfv = (AVRAT(iz) + zetdif * (AVRAT(iz + 1) - AVRAT(iz))) * (alfl + alfad);
                if (alfv < DV) alfv = DV;
                if (alfv > ALFMAX) alfv = ALFMAX;
                if (HWF3 * alfv + h->vnu < VFT) continue;
                h->recalf = 1. / alfv;
                double srad = Ls->e_lo[i];
                if (srad <= -1.001) srad = fabs(srad);
                if (srad > -0.999)
                    sui = sui * M.scor[m * 11 + iso] * exp(-(srad * BETACR)) * (1. + exp(-(h->vnu / XKT)));
                h->sp = sui * (1. + GI * PAVP2);
                double sppi = sui * YI * PAVP0;
                h->sppsp = sppi / h->sp;
                h->srad = 0.0;
                /* rejection of uncoupled lines with DPTMIN = DPTFAC = 0 (:1992-2013) */
                if (IFLAG == 0 && sui * h->recalf <= 0. && h->sppsp == 0.) {
                    h->sp = 0.; h->sppsp = 0.; continue;
                }
            }
            /* 70..: convolve, panel out, loop */
            for (;;) {
                /* CNVFNQ clblm_ODLAY.f90:2067-2258 */
                int ILAST = ILO - 1;
                if (ILO <= IHI) {
                    IPANEL = -1;
                    for (int j = ILO; j <= IHI; j++) {
                        hline *h = &H[j];
                        if (h->sp == 0.) continue;
                        double depth = h->sp * h->recalf;
                        int izm = h->izeta;
                        double zetdif = 100. * h->zetai - (double)(izm - 1);
                        double s1 = depth * (CF1(izm) + zetdif * (CF1(izm + 1) - CF1(izm)));
                        double s2 = depth * (CF2(izm) + zetdif * (CF2(izm + 1) - CF2(izm)));
                        double s3 = depth * (CF3(izm) + zetdif * (CF3(izm + 1) - CF3(izm)));
                        double sd = depth * (CGAUSS(izm) + zetdif * (CGAUSS(izm + 1) - CGAUSS(izm)));
                        double sv = depth * (CER(izm) + zetdif * (CER(izm + 1) - CER(izm)));
                        double zslope = h->recalf * WAVDXF;
                        double zint = (h->vnu - VFT) / DV;
                        double bhwdxf = HWDXF / zslope;
                        int jmax1 = ftrunc(zint + bhwdxf + 1.5);
                        if (jmax1 > MAX1) { ILAST = j - 1; IPANEL = 1; break; }
                        int jmin1 = ftrunc(zint - bhwdxf + 1.5);
                        double rshft = 0.5;
                        if (zint < 0.0) rshft = -rshft;
                        int j2shft = ftrunc(zint * (1. - CONF2) + rshft);
                        int j3shft = ftrunc(zint * (1. - CONF3) + rshft);
                        int jmin2 = jmin1 - j2shft, jmin3 = jmin1 - j3shft;
                        double zf1l = ((double)(jmin1 - 2) - zint) * zslope;
                        double zf2l = ((double)(jmin2 - 2) - zint * CONF2) * zslope;
                        double zf3l = ((double)(jmin3 - 2) - zint * CONF3) * zslope;
                        double zf1 = zf1l, zf2 = zf2l, zf3 = zf3l;
                        for (int j1 = jmin1; j1 <= jmax1; j1++) {
                            int j2 = j1 - j2shft, j3 = j1 - j3shft;
                            zf3 = zf3 + zslope;
                            zf2 = zf2 + zslope;
                            zf1 = zf1 + zslope;
                            int iz3 = ftrunc(fabs(zf3) + 1.5);
                            int iz2 = ftrunc(fabs(zf2) + 1.5);
                            int iz1 = ftrunc(fabs(zf1) + 1.5);
                            if (iz1 > NFMX || iz2 > NFMX || iz3 > NFMX) { ret = fail("table index out of range"); goto done; }
                            if (j1 < -lb || j2 < -lb || j3 < -lb) { ret = fail("panel index out of range"); goto done; }
                            R3[j3] = R3[j3] + s3 * F3[iz3];
                            R2[j2] = R2[j2] + s2 * F2[iz2];
                            R1[j1] = R1[j1] + s1 * F1[iz1] + sd * FG[iz1] + sv * XVER[iz1];
                        }
                        if (h->sppsp != 0.) {
                            /* line coupling, :2195-2235 */
                            double dptrat = h->sppsp;
                            s3 = s3 * CLC3 * dptrat;
                            s2 = s2 * CLC2 * dptrat;
                            s1 = s1 * CLC1 * dptrat;
                            sd = sd * CLC1 * dptrat;
                            sv = sv * CLC1 * dptrat;
                            for (int j1 = jmin1; j1 <= jmax1; j1++) {
                                int j2 = j1 - j2shft, j3 = j1 - j3shft;
                                zf3l = zf3l + zslope;
                                zf2l = zf2l + zslope;
                                zf1l = zf1l + zslope;
                                int iz3 = ftrunc(fabs(zf3l) + 1.5);
                                int iz2 = ftrunc(fabs(zf2l) + 1.5);
                                int iz1 = ftrunc(fabs(zf1l) + 1.5);
                                R3[j3] = R3[j3] + s3 * F3[iz3] * zf3l;
                                R2[j2] = R2[j2] + s2 * F2[iz2] * zf2l;
                                R1[j1] = R1[j1] + (s1 * F1[iz1] + sd * FG[iz1] + sv * XVER[iz1]) * zf1l;
                            }
                        }
                    }
                    if (IPANEL == -1) { ILAST = IHI; IPANEL = IDATA; }
                } else {
                    IPANEL = 1;
                }
                ILO = ILAST + 1;
                if (IPANEL == 0) break; /* need more data */
                /* 80: panel output */
                if (IFPAN == 1) { IFPAN = 0; N1R2 = 1; N1R3 = 1; }
                else { N1R2 = (NSHIFT / 4) + 1 + 4; N1R3 = (NSHIFT / 16) + 1 + 3; }
                N2R3 = NLIM3 + (NSHIFT / 16) - 1 + 3;
                if (VFT <= 0.) { ret = fail("panel start <= 0 cm-1 (RSYM) not supported"); goto done; }
                orc_xint(V1R4, V2R4, DVR4, R4, 1.0, VFT, DVR3, R3, N1R3, N2R3);
                if (P->contnm) orc_xint(V1ABS, V2ABS, DVABS, P->absrb - 1, 1., VFT, DVR3, R3, N1R3, N2R3);
                /* panelOut clblm_ODLAY.f90:2289-2545 */
                {
                    const double X00 = -7. / 128., X01 = 105. / 128., X02 = 35. / 128., X03 = -5. / 128.;
                    const double X10 = -1. / 16., X11 = 9. / 16.;
                    double DVP = DV;
                    ISTOP = 0;
                    double V2P;
                    if ((VFT + (double)(NHI - 1) * DVP) > V2) {
                        NHI = ftrunc((V2 - VFT) / DVP + 1.);
                        V2P = VFT + (double)(NHI - 1) * DVP;
                        if (V2P < V2) { V2P = V2P + DVP; NHI = NHI + 1; }
                        ISTOP = 1;
                    } else {
                        V2P = VFT + (double)(NHI - 1) * DV;
                    }
                    int NLIM = NHI - NLO + 1;
                    double V1P = VFT + (double)(NLO - 1) * DV;
                    int LIMLO = N1R2;
                    if (N1R2 == 1) LIMLO = LIMLO + 4;
                    int LIMHI = (NHI / 4) + 1;
                    int J = LIMLO, J3 = 0;
                    for (J = LIMLO; J <= LIMHI; J += 4) {
                        J3 = (J - 1) / 4 + 1;
                        R2[J] = R2[J] + R3[J3];
                        R2[J + 1] = R2[J + 1] + X00 * R3[J3 - 1] + X01 * R3[J3] + X02 * R3[J3 + 1] + X03 * R3[J3 + 2];
                        R2[J + 2] = R2[J + 2] + X10 * (R3[J3 - 1] + R3[J3 + 2]) + X11 * (R3[J3] + R3[J3 + 1]);
                        R2[J + 3] = R2[J + 3] + X03 * R3[J3 - 1] + X02 * R3[J3] + X01 * R3[J3 + 1] + X00 * R3[J3 + 2];
                    }
                    if (LIMLO <= LIMHI && ISTOP == 1) {
                        J3 = J3 + 1;
                        R2[J] = R2[J] + R3[J3];
                        R2[J + 1] = R2[J + 1] + X00 * R3[J3 - 1] + X01 * R3[J3] + X02 * R3[J3 + 1] + X03 * R3[J3 + 2];
                    }
                    for (J = NLO; J <= NHI; J += 4) {
                        int J2 = (J - 1) / 4 + 1;
                        R1[J] = R1[J] + R2[J2];
                        R1[J + 1] = R1[J + 1] + X00 * R2[J2 - 1] + X01 * R2[J2] + X02 * R2[J2 + 1] + X03 * R2[J2 + 2];
                        R1[J + 2] = R1[J + 2] + X10 * (R2[J2 - 1] + R2[J2 + 2]) + X11 * (R2[J2] + R2[J2 + 1]);
                        R1[J + 3] = R1[J + 3] + X03 * R2[J2 - 1] + X02 * R2[J2] + X01 * R2[J2 + 1] + X00 * R2[J2 + 2];
                    }
                    if (JRAD == 0) {
                        double xkt = TAVE / RADCN2;
                        double VI = V1P - DV;
                        double VITST = VI, RDLAST = -1., RDEL;
                        int NPTSI1, NPTSI2 = NLO - 1;
                        do {
                            NPTSI1 = NPTSI2 + 1;
                            VI = VFT + (double)(NPTSI1 - 1) * DV;
                            double RADVI = orc_radfni(VI, DV, xkt, &VITST, &RDEL, &RDLAST);
                            NPTSI2 = ftrunc((VITST - VFT) / DV + 0.001);
                            if (NPTSI2 > NHI) NPTSI2 = NHI;
                            double t = RADVI;
                            for (int i = NPTSI1; i <= NPTSI2; i++) { R1[i] = R1[i] * t; t = t + RDEL; }
                        } while (NPTSI2 < NHI);
                    }
                    (void)V2P;
                    if (out) {
                        if (numTot + NLIM > cap) { ret = fail("output capacity too small"); goto done; }
                        memcpy(out + numTot, R1 + NLO, sizeof(double) * (size_t)NLIM);
                    }
                    numTot += NLIM;
                    VFT = VFT + (double)(NLIM1 - 1) * DV;
                    if (ISTOP != 1) {
                        memmove(R1 + 1, R1 + NLIM1, sizeof(double) * (size_t)(MAX1 - NLIM1 + 1));
                        for (int k = MAX1 - NLIM1 + 2; k <= MAX1; k++) R1[k] = 0.;
                        memmove(R2 + 1, R2 + NLIM2, sizeof(double) * (size_t)(MAX2 - NLIM2 + 1));
                        for (int k = MAX2 - NLIM2 + 2; k <= MAX2; k++) R2[k] = 0.;
                        memmove(R3 + 1, R3 + NLIM3, sizeof(double) * (size_t)(MAX3 - NLIM3 + 1));
                        for (int k = MAX3 - NLIM3 + 2; k <= MAX3; k++) R3[k] = 0.;
                        NLO = NSHIFT + 1;
                    }
                }
                if (ISTOP == 1) goto finished;
                /* F4 block refresh, clblm_ODLAY.f90:1469-1499 */
                {
                    double VF1 = VFT - 2. * DVR4;
                    VF1 = V1 + (double)(int)floor((VF1 - V1) / DVR4) * DVR4;
                    double VF2 = VFT + 2. * DVR4 + (double)(N2R3 + 4) * DVR3;
                    if (VF2 > V2R4 && V2R4 != V2R4ST) {
                        V1R4 = VF1;
                        V2R4 = V2R4ST;
                        NPTR4 = lblf4q(JRAD, V1R4, &V2R4, R4, DVR4, BOUND4, TAVE, &sh);
                    }
                }
                /* GO TO 70 */
            }
            need_read = 1;
        }
        if (IEOF != 0) {
            /* 80 loop until ISTOP with no more lines */
            ILO = 1; IHI = 0;
            for (;;) {
                if (IFPAN == 1) { IFPAN = 0; N1R2 = 1; N1R3 = 1; }
                else { N1R2 = (NSHIFT / 4) + 1 + 4; N1R3 = (NSHIFT / 16) + 1 + 3; }
                N2R3 = NLIM3 + (NSHIFT / 16) - 1 + 3;
                if (VFT <= 0.) { ret = fail("panel start <= 0 cm-1 (RSYM) not supported"); goto done; }
                orc_xint(V1R4, V2R4, DVR4, R4, 1.0, VFT, DVR3, R3, N1R3, N2R3);
                if (P->contnm) orc_xint(V1ABS, V2ABS, DVABS, P->absrb - 1, 1., VFT, DVR3, R3, N1R3, N2R3);
                {
                    const double X00 = -7. / 128., X01 = 105. / 128., X02 = 35. / 128., X03 = -5. / 128.;
                    const double X10 = -1. / 16., X11 = 9. / 16.;
                    double DVP = DV;
                    ISTOP = 0;
                    double V2P;
                    if ((VFT + (double)(NHI - 1) * DVP) > V2) {
                        NHI = ftrunc((V2 - VFT) / DVP + 1.);
                        V2P = VFT + (double)(NHI - 1) * DVP;
                        if (V2P < V2) { V2P = V2P + DVP; NHI = NHI + 1; }
                        ISTOP = 1;
                    }
                    int NLIM = NHI - NLO + 1;
                    double V1P = VFT + (double)(NLO - 1) * DV;
                    int LIMLO = N1R2;
                    if (N1R2 == 1) LIMLO = LIMLO + 4;
                    int LIMHI = (NHI / 4) + 1;
                    int J = LIMLO, J3 = 0;
                    for (J = LIMLO; J <= LIMHI; J += 4) {
                        J3 = (J - 1) / 4 + 1;
                        R2[J] = R2[J] + R3[J3];
                        R2[J + 1] = R2[J + 1] + X00 * R3[J3 - 1] + X01 * R3[J3] + X02 * R3[J3 + 1] + X03 * R3[J3 + 2];
                        R2[J + 2] = R2[J + 2] + X10 * (R3[J3 - 1] + R3[J3 + 2]) + X11 * (R3[J3] + R3[J3 + 1]);
                        R2[J + 3] = R2[J + 3] + X03 * R3[J3 - 1] + X02 * R3[J3] + X01 * R3[J3 + 1] + X00 * R3[J3 + 2];
                    }
                    if (LIMLO <= LIMHI && ISTOP == 1) {
                        J3 = J3 + 1;
                        R2[J] = R2[J] + R3[J3];
                        R2[J + 1] = R2[J + 1] + X00 * R3[J3 - 1] + X01 * R3[J3] + X02 * R3[J3 + 1] + X03 * R3[J3 + 2];
                    }
                    for (J = NLO; J <= NHI; J += 4) {
                        int J2 = (J - 1) / 4 + 1;
                        R1[J] = R1[J] + R2[J2];
                        R1[J + 1] = R1[J + 1] + X00 * R2[J2 - 1] + X01 * R2[J2] + X02 * R2[J2 + 1] + X03 * R2[J2 + 2];
                        R1[J + 2] = R1[J + 2] + X10 * (R2[J2 - 1] + R2[J2 + 2]) + X11 * (R2[J2] + R2[J2 + 1]);
                        R1[J + 3] = R1[J + 3] + X03 * R2[J2 - 1] + X02 * R2[J2] + X01 * R2[J2 + 1] + X00 * R2[J2 + 2];
                    }
                    if (JRAD == 0) {
                        double xkt = TAVE / RADCN2;
                        double VI = V1P - DV;
                        double VITST = VI, RDLAST = -1., RDEL;
                        int NPTSI1, NPTSI2 = NLO - 1;
                        do {
                            NPTSI1 = NPTSI2 + 1;
                            VI = VFT + (double)(NPTSI1 - 1) * DV;
                            double RADVI = orc_radfni(VI, DV, xkt, &VITST, &RDEL, &RDLAST);
                            NPTSI2 = ftrunc((VITST - VFT) / DV + 0.001);
                            if (NPTSI2 > NHI) NPTSI2 = NHI;
                            double t = RADVI;
                            for (int i = NPTSI1; i <= NPTSI2; i++) { R1[i] = R1[i] * t; t = t + RDEL; }
                        } while (NPTSI2 < NHI);
                    }
                    (void)V2P;
                    if (out) {
                        if (numTot + NLIM > cap) { ret = fail("output capacity too small"); goto done; }
                        memcpy(out + numTot, R1 + NLO, sizeof(double) * (size_t)NLIM);
                    }
                    numTot += NLIM;
                    VFT = VFT + (double)(NLIM1 - 1) * DV;
                    if (ISTOP != 1) {
                        memmove(R1 + 1, R1 + NLIM1, sizeof(double) * (size_t)(MAX1 - NLIM1 + 1));
                        for (int k = MAX1 - NLIM1 + 2; k <= MAX1; k++) R1[k] = 0.;
                        memmove(R2 + 1, R2 + NLIM2, sizeof(double) * (size_t)(MAX2 - NLIM2 + 1));
                        for (int k = MAX2 - NLIM2 + 2; k <= MAX2; k++) R2[k] = 0.;
                        memmove(R3 + 1, R3 + NLIM3, sizeof(double) * (size_t)(MAX3 - NLIM3 + 1));
                        for (int k = MAX3 - NLIM3 + 2; k <= MAX3; k++) R3[k] = 0.;
                        NLO = NSHIFT + 1;
                    }
                }
                if (ISTOP == 1) goto finished;
                {
                    double VF1 = VFT - 2. * DVR4;
                    VF1 = V1 + (double)(int)floor((VF1 - V1) / DVR4) * DVR4;
                    double VF2 = VFT + 2. * DVR4 + (double)(N2R3 + 4) * DVR3;
                    if (VF2 > V2R4 && V2R4 != V2R4ST) {
                        V1R4 = VF1;
                        V2R4 = V2R4ST;
                        NPTR4 = lblf4q(JRAD, V1R4, &V2R4, R4, DVR4, BOUND4, TAVE, &sh);
                    }
                }
            }
        }
    }
finished:
    (void)NPTR4; (void)nrec_lines;
    if (dv_used) *dv_used = DV;
    /* optional re-grid to DVout, clblm_ODLAY.f90:573-599 */
    if (P->dvout > 0. && !(fabs(P->dvout - DV) < 2.220446049250313e-16)) {
        int len = (int)ceil((V2 - V1) / P->dvout + 1.);
        if (out) {
            if (len > cap) { ret = fail("output capacity too small"); goto done; }
            double *tmp = malloc(sizeof(double) * (size_t)numTot);
            memcpy(tmp, out, sizeof(double) * (size_t)numTot);
            orc_interp(tmp, DV, V1, (int)numTot, out, P->dvout, V1, len);
            free(tmp);
        }
        numTot = len;
        if (dv_used) *dv_used = P->dvout;
    }
    *nsamp = numTot;
done:
    free(R1b); free(R2b); free(R3b); free(H);
    shr_free(&sh);
    return ret;
}

/* ---------------------------------------------------------------------------
 * Source functions, clblm_EMLAY.f90:18-72,104-223.
 * ------------------------------------------------------------------------- */
static double planck_x(double v, double xkt) {
    if (xkt <= 1.e-6) return 0.;
    double cv3 = RADCN1 * v * v * v;
    double evt = exp(v / xkt);
    return cv3 / (evt - 1.);
}
double orc_planck(double v, double t) { return planck_x(v, t / RADCN2); }

void orc_planck_drv(double v, double t, double *b, double *dbdt) {
    double xkt = t / RADCN2;
    if (xkt <= 1.e-6) { *b = 0.; *dbdt = 0.; return; }
    double cv3 = RADCN1 * v * v * v;
    double evt = exp(v / xkt);
    *b = cv3 / (evt - 1.);
    *dbdt = *b * *b * (evt * v) / (RADCN2 * cv3 * xkt * xkt);
}

void orc_rlin(double od, double *hlf, double *dhlf) {
    const double dbAvg2dtau = -1. / 12;
    double tex = 0.;
    if (od < 1.e-03) *hlf = .5 + od * dbAvg2dtau;
    else if (od > 20.) *hlf = 1. / od;
    else { tex = exp(od); *hlf = 1. / (od) - 1. / (tex - 1.); }
    if (dhlf) {
        if (od < 1.e-03) *dhlf = dbAvg2dtau;
        else if (od > 20.) *dhlf = -1. / (od * od);
        else *dhlf = -1. / (od * od) + tex / ((tex - 1) * (tex - 1));
    }
}

/* layerEmis (clblm_EMLAY.f90:104-178) at one point: ta is the boundary the layer's
 * linear-in-tau source starts from, tave the layer temperature. */
static double layer_emis(double v, double ta, double tb, double tave, double tx, double od, int lin) {
    double abso = 1 - tx, hlf = 0.;
    if (lin != 0) orc_rlin(od, &hlf, NULL);
    double xkta = ta / RADCN2, xktb = tb / RADCN2, xkt = tave / RADCN2;
    if (lin == 0) return abso * planck_x(v, xkt);
    if (lin == 1) {
        double bba = planck_x(v, xkta), bb = planck_x(v, xkt);
        return abso * (bba + 2 * (bb - bba) * hlf);
    }
    double bba = planck_x(v, xkta), bbb = planck_x(v, xktb);
    return abso * (bba + (bbb - bba) * hlf);
}

/* layerEmis at every point of v1 + i dv, transmittance exp(-od): the entry point
 * the tests compare with the reference's own layerEmis. */
void orc_layer_emis(int64_t npts, const double *od, double v1, double dv, double ta, double tb, double tave,
                    int lin, double *out) {
    for (int64_t i = 0; i < npts; i++)
        out[i] = layer_emis(v1 + (double)i * dv, ta, tb, tave, exp(-od[i]), od[i], lin);
}

/* mode3_lookUp_mergeUp, clblm_noScattRT.f90:843-1108: down-welling radiance at the
 * observer, layers from the observer up (layer 0 nearest), merged upward: each
 * layer's downward emission (layerEmis from its lower level, :972-974) is attenuated
 * by the layers below it (:981-984); sol (nullable) the solar radiance at the top,
 * added through the whole path (:1021-1027). */
void orc_rt_down(int nlay, int64_t npts, double v1, double dv, const double *od, const double *t_lay,
                 const double *t_lev, int lin, const double *sol, double *rad, double *tx) {
    for (int64_t i = 0; i < npts; i++) {
        double v = v1 + (double)i * dv;
        double r = 0., tacc = 1.;
        for (int l = 0; l < nlay; l++) {
            double o = od[(int64_t)l * npts + i];
            double t = exp(-o);
            double em = layer_emis(v, t_lev[l], t_lev[l + 1], t_lay[l], t, o, lin);
            r = r + em * tacc;
            tacc = tacc * t;
        }
        if (sol) r = r + tacc * sol[i];
        rad[i] = r;
        if (tx) tx[i] = tacc;
    }
}

/* Upwelling radiance at the top of nlay layers (layer 0 at the surface), every
 * layer below the observer: mode2_lookDn_mergeUp / layerMerge_mode2
 * (clblm_noScattRT.f90:1287-1296, 1704-1794) on a common grid.  od[l*npts+i] is the
 * layer OD along the view path; the down path uses od * sec_dn. */
void orc_rt_up(int nlay, int64_t npts, double v1, double dv, const double *od, const double *t_lay,
               const double *t_lev, double emis, double tskin, int lin, double sec_dn, double *rad) {
    for (int64_t i = 0; i < npts; i++) {
        double v = v1 + (double)i * dv;
        double r = emis * orc_planck(v, tskin);
        double rho = 1. - emis;
        for (int l = 0; l < nlay; l++) {
            double oda = od[(int64_t)l * npts + i];
            double odb = oda * sec_dn;
            double txa = exp(-oda), txb = exp(-odb);
            double tbot = t_lev[l], ttop = t_lev[l + 1], tave = t_lay[l];
            double emdn = layer_emis(v, tbot, ttop, tave, txb, odb, lin);
            double emup = lin != 0 ? layer_emis(v, ttop, tbot, tave, txa, oda, lin) : emdn;
            r = emup + txa * (r + emdn * rho);
            rho = txa * txb * rho;
        }
        rad[i] = r;
    }
}

/* The flux branches of mode1_lookDn_mergeDn (clblm_noScattRT.f90:133-422, 482-507) and
 * mode2_lookDn_mergeUp (:1214-1450) with the driver's net flux and heating rate
 * (clblm_Drivers.f90:418-466), every layer on one grid (od[l*npts + i], layer 0 at the
 * surface, v1 + i dv, i < npts), with the quadrature of nang = 1..3 angles (secants and
 * weights of Clough et al. 1992, clblm_ConstParam.f90:31-38).  Per angle the ODs are
 * scaled by the secant; the down radiance is merged from the top (EmDn + R TxB), the
 * surface gives eps B(Tskin) plus (1 - eps) times the downwelling (lambertian, the
 * reference's ThmReflMode 0: the quadrature flux over pi, else the angle's own
 * radiance; :364-376, :482-507 and addThermalBoundary's flux branch :2020-2024), the
 * up radiance is merged
 * from the surface (EmUp + TxA R; the reflected term is zero in the flux branch).  Each
 * level's radiances are summed in groups of OUTINRAT = int(dv_flux / dv) points and
 * weighted: flux = (sum_a w_a S_a) dv 1e4 2 pi.  Outputs per level j = 0..nlay
 * (0: surface) and group k < nbin: up[j*nbin + k], dn, net = up - dn and
 * htr = HEATFC (net_j - net_j+1) / (p_j - p_j+1) (0 at the top).  Returns nbin (NOUT)
 * or -1 when (v2 - v1) / dv_flux is not a whole number. */
int orc_rt_flux(int nlay, int64_t npts, double v1, double v2, double dv, const double *od, const double *t_lay,
                const double *t_lev, const double *p_lev, int lin, double emis, double tskin, int nang,
                double dv_flux, int lambertian, double *up, double *dn, double *net, double *htr) {
    const double PIC = 3.1415926535898, EPSC = 1.E-4, HEATFC = 8.4391;
    double sec[3] = {0., 0., 0.}, w[3] = {0., 0., 0.};
    if (nang == 1) { sec[0] = 1. / 0.66666666667; w[0] = 0.50; }
    else if (nang == 2) { sec[0] = 1.0 / 0.84494897; sec[1] = 1.0 / 0.35505103; w[0] = 0.31804138; w[1] = 0.18195862; }
    else if (nang == 3) {
        sec[0] = 1.0 / 0.91141204; sec[1] = 1.0 / 0.59053314; sec[2] = 1.0 / 0.21234054;
        w[0] = 0.20093191; w[1] = 0.22924111; w[2] = 0.06982698;
    } else return fail("flux: nang must be 1, 2 or 3");
    const double OUT = (v2 - v1) / dv_flux;
    const int NOUT = (int)(OUT + EPSC);
    if (fabs((double)NOUT - OUT) > EPSC) return fail("flux: V1, V2, DV_flux are inconsistent");
    const int outinrat = (int)(dv_flux / dv);
    const double FACTOR = dv * 1.E04 * 2. * PIC;
    const int nlev = nlay + 1;
    double *radd = (double *)calloc((size_t)nang * nlay * npts, sizeof(double));
    double *radu = (double *)calloc((size_t)nang * nlay * npts, sizeof(double));
    double *dflux = (double *)calloc((size_t)npts, sizeof(double));
    double *srad = (double *)calloc((size_t)(NOUT + 2) * nang, sizeof(double));
    double *sfcu = (double *)calloc((size_t)NOUT + 2, sizeof(double));
    /* down, per angle and point */
    for (int a = 0; a < nang; a++)
        for (int64_t i = 0; i < npts; i++) {
            const double v = v1 + (double)i * dv;
            double r = 0.;
            for (int l = nlay - 1; l >= 0; l--) {
                const double odb = od[(int64_t)l * npts + i] * sec[a];
                const double txb = exp(-odb);
                const double emdn = layer_emis(v, t_lev[l], t_lev[l + 1], t_lay[l], txb, odb, lin);
                r = emdn + r * txb;
                radd[((int64_t)a * nlay + l) * npts + i] = r;
            }
        }
    /* surface: downwelling flux density and the upwelling surface radiance per angle */
    for (int64_t i = 0; i < npts; i++) {
        const double *r0 = radd + i;   /* layer 0, angle a at r0[a * nlay * npts] */
        double d;
        if (nang == 1) d = FACTOR * w[0] * r0[0];
        else if (nang == 2) d = FACTOR * (w[0] * r0[0] + w[1] * r0[(int64_t)nlay * npts]);
        else d = FACTOR * (w[0] * r0[0] + w[1] * r0[(int64_t)nlay * npts] + w[2] * r0[2 * (int64_t)nlay * npts]);
        dflux[i] = d;
    }
    for (int a = 0; a < nang; a++)
        for (int64_t i = 0; i < npts; i++) {
            const double v = v1 + (double)i * dv;
            const double srf = orc_planck(v, tskin) * emis;
            /* the lambertian form is set for points 1 .. nlim - 1 (:488-495); the last
             * point keeps addThermalBoundary's, the angle's own reflected radiance */
            double r = (lambertian && i < npts - 1) ? srf + ((1. - emis) * (dflux[i] / (PIC * dv * 1.E04)))
                                                    : srf + (1. - emis) * radd[((int64_t)a * nlay + 0) * npts + i];
            for (int l = 0; l < nlay; l++) {
                const double oda = od[(int64_t)l * npts + i] * sec[a];
                const double txa = exp(-oda);
                const double emdn = layer_emis(v, t_lev[l], t_lev[l + 1], t_lay[l], txa, oda, lin);
                const double emup = lin != 0 ? layer_emis(v, t_lev[l + 1], t_lev[l], t_lay[l], txa, oda, lin) : emdn;
                r = emup + txa * (r + emdn * 0.);
                radu[((int64_t)a * nlay + l) * npts + i] = r;
            }
        }
    for (int64_t q = 0; q < (int64_t)nlev * NOUT; q++) up[q] = dn[q] = 0.;
    /* groups of each level: layer l's down radiance is the flux at level l, its up
     * radiance the flux at level l + 1 */
    for (int pass = 0; pass < 2; pass++)
        for (int l = 0; l < nlay; l++) {
            const double *R = pass == 0 ? radd : radu;
            for (int q = 0; q < (NOUT + 2) * nang; q++) srad[q] = 0.;
            int iout = 1, icount = 0;
            for (int64_t k = 0; k < npts; k++) {
                for (int a = 0; a < nang; a++) srad[(iout - 1) * nang + a] += R[((int64_t)a * nlay + l) * npts + k];
                icount++;
                if (icount >= outinrat) {
                    icount = 0;
                    iout++;
                    if (iout > NOUT + 1) iout = NOUT + 1;   /* the reference's arrays end there */
                }
            }
            double *F = pass == 0 ? dn + (int64_t)l * NOUT : up + (int64_t)(l + 1) * NOUT;
            for (int b = 0; b < NOUT; b++) {
                const double *S = srad + b * nang;
                if (nang == 1) F[b] = w[0] * S[0] * FACTOR;
                else if (nang == 2) F[b] = (w[0] * S[0] + w[1] * S[1]) * FACTOR;
                else F[b] = (w[0] * S[0] + w[1] * S[1] + w[2] * S[2]) * FACTOR;
            }
        }
    /* surface upwelling flux, points 1 .. nlim - 1 */
    {
        int iout = 1, icount = 0;
        for (int64_t k = 0; k < npts - 1; k++) {
            const double v = v1 + (double)k * dv;
            const double srf = orc_planck(v, tskin) * emis;
            sfcu[iout - 1] = sfcu[iout - 1] + srf * dv * 1.E04 * PIC + (1. - emis) * dflux[k];
            icount++;
            if (icount >= outinrat) {
                icount = 0;
                iout++;
                if (iout > NOUT + 1) iout = NOUT + 1;
            }
        }
        for (int b = 0; b < NOUT; b++) up[b] = sfcu[b];
    }
    for (int j = 0; j < nlev; j++)
        for (int b = 0; b < NOUT; b++) {
            const int64_t q = (int64_t)j * NOUT + b;
            net[q] = up[q] - dn[q];
        }
    for (int j = nlev - 1; j >= 0; j--)
        for (int b = 0; b < NOUT; b++) {
            const int64_t q = (int64_t)j * NOUT + b;
            if (j == nlev - 1) htr[q] = 0.;
            else htr[q] = HEATFC * (net[q] - net[q + NOUT]) / (p_lev[j] - p_lev[j + 1]);
        }
    free(radd); free(radu); free(dflux); free(srad); free(sfcu);
    return NOUT;
}

/* spectraHaveSameGrid (clblm_Spectrum.f90) for two spectra from one V1: DV,
 * V2 = V1 + DV (NLIM - 1) (CLBLM_Spectrum_init) and NLIM equal, to epsilon(1d0). */
static int same_grid(double v1, double dva, int64_t na, double dvb, int64_t nb) {
    const double EPS = 2.220446049250313e-16;
    return fabs(dva - dvb) < EPS && na == nb &&
           fabs((v1 + dva * (double)(na - 1)) - (v1 + dvb * (double)(nb - 1))) < EPS;
}

/* interpToFinestGrid_array, clblm_Spectrum.f90:799-866, for spectra from one V1: the
 * first spectrum of smallest DV sets the grid; every other one not already on it is
 * interpolated onto it (interpSpectrum).  out[k] receives spectrum k on its final
 * grid; returns the index of the finest. */
int orc_interp_to_finest(int nspec, double v1, const double *dv, const int64_t *nlim, const double *const *in,
                         double *const *out) {
    int loc = 0;
    for (int k = 1; k < nspec; k++)
        if (dv[k] < dv[loc]) loc = k;
    for (int k = 0; k < nspec; k++) {
        if (k == loc || same_grid(v1, dv[k], nlim[k], dv[loc], nlim[loc])) {
            memmove(out[k], in[k], sizeof(double) * (size_t)nlim[k]);
            continue;
        }
        orc_interp(in[k], dv[k], v1, (int)nlim[k], out[k], dv[loc], v1, (int)nlim[loc]);
    }
    return loc;
}

/* mode2_lookDn_mergeUp / layerMerge_mode2 (clblm_noScattRT.f90:1287-1296, 1700-1794)
 * with every layer on its own grid: layer l's OD is od + off[l], nlim[l] points on
 * v1 + i dv[l]; the surface starts on surf_dv, surf_nlim.  Before each merge
 * interpToFinestGrid (mrgRadUp, mrgThmRefl, lyrTxA, lyrTxB, lyrEmUp, lyrEmDn) puts
 * all six on the finest grid.  rad (capacity cap) receives the radiance; returns its
 * length (*dv_out its DV), or -1. */
int64_t orc_rt_up_grids(int nlay, double v1, const double *dv, const int64_t *nlim, const int64_t *off,
                        const double *od, const double *t_lay, const double *t_lev, double emis, double tskin,
                        double surf_dv, int64_t surf_nlim, int lin, double sec_dn, double *rad, int64_t cap,
                        double *dv_out) {
    int64_t mx = surf_nlim;
    for (int l = 0; l < nlay; l++) mx = nlim[l] > mx ? nlim[l] : mx;
    double *R = malloc(sizeof(double) * mx), *RH = malloc(sizeof(double) * mx);
    double *Q = malloc(sizeof(double) * 4 * mx), *T = malloc(sizeof(double) * 6 * mx);
    double cdv = surf_dv;
    int64_t cn = surf_nlim;
    for (int64_t i = 0; i < cn; i++) {
        double v = v1 + (double)i * cdv;
        R[i] = orc_planck(v, tskin) * emis;
        RH[i] = 1.0 - emis;
    }
    for (int l = 0; l < nlay; l++) {
        const double ldv = dv[l];
        const int64_t ln = nlim[l];
        const double *o = od + off[l];
        double *txa = Q, *txb = Q + ln, *emup = Q + 2 * ln, *emdn = Q + 3 * ln;
        for (int64_t i = 0; i < ln; i++) {
            double v = v1 + (double)i * ldv;
            double oda = o[i], odb = oda * sec_dn;
            txa[i] = exp(-oda);
            txb[i] = exp(-odb);
            emdn[i] = layer_emis(v, t_lev[l], t_lev[l + 1], t_lay[l], txb[i], odb, lin);
            emup[i] = lin != 0 ? layer_emis(v, t_lev[l + 1], t_lev[l], t_lay[l], txa[i], oda, lin) : emdn[i];
        }
        double dvs[6] = {cdv, cdv, ldv, ldv, ldv, ldv};
        int64_t ns[6] = {cn, cn, ln, ln, ln, ln};
        const double *in[6] = {R, RH, txa, txb, emup, emdn};
        double *outp[6];
        for (int k = 0; k < 6; k++) outp[k] = T + k * mx;
        int loc = orc_interp_to_finest(6, v1, dvs, ns, in, outp);
        /* the merged arrays take the grid they end on (their own when already on it) */
        int64_t rn = (loc == 0 || same_grid(v1, cdv, cn, dvs[loc], ns[loc])) ? cn : ns[loc];
        if (!(loc == 0 || same_grid(v1, cdv, cn, dvs[loc], ns[loc]))) cdv = dvs[loc];
        cn = rn;
        for (int64_t i = 0; i < cn; i++) {
            double r = outp[0][i], rho = outp[1][i];
            r = outp[4][i] + outp[2][i] * (r + outp[5][i] * rho);
            rho = outp[2][i] * outp[3][i] * rho;
            R[i] = r;
            RH[i] = rho;
        }
    }
    int64_t n = cn < cap ? cn : cap;
    memcpy(rad, R, sizeof(double) * (size_t)n);
    *dv_out = cdv;
    free(R); free(RH); free(Q); free(T);
    return cn;
}

/* ==== Jacobians ================================================================
 * level2layer (clblm_noScattJac.f90:1208-1279): derivatives of the layer mean
 * temperature and of ln(layer amount) with respect to the bounding levels' T and
 * ln(number density).  qlev[m*(nlay+1)+k] is the level density of molecule m. */
void orc_level2layer(int nlay, int nmol, const double *plev, const double *tlev, const double *tave,
                     const double *qlev, double *dt_dtup, double *dt_dtlo, double *dw_dqup, double *dw_dqlo) {
    const double BOLTZ_HPA = 1.3806503E-19;   /* hPa cm3 / K, clblm_noScattJac.f90:1240 */
    for (int l = 0; l < nlay; l++) {
        double pup = plev[l + 1], plo = plev[l], tup = tlev[l + 1], tlo = tlev[l], tbar = tave[l];
        double rhoU = pup / (tup * BOLTZ_HPA);
        double rhoL = plo / (tlo * BOLTZ_HPA);
        double alpha = rhoU / rhoL;
        double alphaT = -(tup - tlo) / log(alpha);
        dt_dtlo[l] = ((tbar - alphaT) / tlo) * (rhoL / (rhoL - rhoU)) + (1.0 - alphaT / tlo) / log(alpha);
        dt_dtup[l] = ((tbar - alphaT) / tup) * (-rhoU / (rhoL - rhoU)) - (1.0 - alphaT / tup) / log(alpha);
        for (int k = 0; k < nmol; k++) {
            double denUp = qlev[(int64_t)k * (nlay + 1) + l + 1], denLo = qlev[(int64_t)k * (nlay + 1) + l];
            double *up = &dw_dqup[(int64_t)k * nlay + l], *lo = &dw_dqlo[(int64_t)k * nlay + l];
            if (denLo != 0.0) {
                double ratU = denUp / rhoU, ratL = denLo / rhoL;
                *lo = (ratL / (ratL - alpha * ratU)) + 1.0 / log(alpha * ratU / ratL);
                *up = ((-alpha * ratU) / (ratL - alpha * ratU)) - 1.0 / log(alpha * ratU / ratL);
            } else {
                *lo = 0.0;
                *up = 0.0;
            }
        }
    }
}

/* Jacobians (clblm_noScattJac.f90:346-716) at every point of v1 + i*dv for a
 * down-looking sensor above the top layer (obsLev = nlay + 1), thermal only: the
 * view OD of layer l is od[l*npts+i] (amfView 1), the down-path OD is sec_dn times it
 * (the air-mass-scaled path, amfDown = sec_dn, pathOD_up_dn_sun :826-838).  Layer 0 is
 * at the surface.  Outputs: rad, dr_dtau[l*npts+i] (derivative with respect to the
 * layer's vertical OD), dr_dtlev[k*npts+i] for the nlay+1 levels (level Planck terms
 * plus, for lin_in_tau 0 and 1, the layer-mean terms carried by dt_dtup/dt_dtlo),
 * dr_dtskin and dr_demis.  The summation order is the reference's. */
void orc_jacobians(int nlay, int64_t npts, double v1, double dv, const double *od, const double *t_lay,
                   const double *t_lev, double emis, double tskin, int lin, double sec_dn, const double *dt_dtup,
                   const double *dt_dtlo, double *rad_o, double *dr_dtau, double *dr_dtlev, double *dr_dtskin,
                   double *dr_demis) {
    const int L = nlay;
    double *Blev = malloc(sizeof(double) * (L + 2)), *dBlev = malloc(sizeof(double) * (L + 2));
    double *Bbar = malloc(sizeof(double) * (L + 1)), *dBbar = malloc(sizeof(double) * (L + 1));
    double *cumTxA = malloc(sizeof(double) * (L + 2)), *cumTxB2 = malloc(sizeof(double) * (L + 2));
    double *dTau = malloc(sizeof(double) * (L + 1)), *dTbar = malloc(sizeof(double) * (L + 1));
    double *dTup = malloc(sizeof(double) * (L + 2)), *dTdn = malloc(sizeof(double) * (L + 2));
    double *dTlev = malloc(sizeof(double) * (L + 2));
    double *vOD = malloc(sizeof(double) * (L + 1)), *dOD = malloc(sizeof(double) * (L + 1));
    /* 1-based arrays as in the reference: layers 1..L, levels 1..L+1 */
    for (int64_t i = 0; i < npts; i++) {
        double vn = v1 + (double)i * dv;
        for (int l = 1; l <= L; l++) {
            vOD[l] = od[(int64_t)(l - 1) * npts + i];
            dOD[l] = vOD[l] * sec_dn;
        }
        if (lin != 0)
            for (int k = 1; k <= L + 1; k++) orc_planck_drv(vn, t_lev[k - 1], &Blev[k], &dBlev[k]);
        for (int l = 1; l <= L; l++) orc_planck_drv(vn, t_lay[l - 1], &Bbar[l], &dBbar[l]);
        double Bsfc, dBsfc;
        orc_planck_drv(vn, tskin, &Bsfc, &dBsfc);
        double rad = 0.;
        for (int l = 1; l <= L; l++) { dTau[l] = 0.; dTbar[l] = 0.; }
        for (int k = 1; k <= L + 1; k++) { dTup[k] = 0.; dTdn[k] = 0.; dTlev[k] = 0.; }
        const int obsLay = L + 1;
        /* 0- transmittance along the view path from the top down to the surface */
        double cumODA = 0.;
        cumTxA[obsLay] = 1.;
        for (int l = obsLay - 1; l >= 1; l--) {
            cumODA = cumODA + vOD[l];
            cumTxA[l] = exp(-cumODA);
        }
        double totODUp = cumODA, totTxUp = cumTxA[1];
        /* 1- downwelling thermal radiance, reflected at the surface */
        if (totTxUp > 1.e-06) {
            cumTxB2[1] = totTxUp;
            double cumODB = 0.;
            for (int l = 1; l <= L; l++) {
                cumODB = cumODB + dOD[l];
                cumTxB2[l + 1] = exp(-(totODUp + cumODB));
            }
            for (int l = L; l >= 1; l--) {
                double dTran = cumTxB2[l] - cumTxB2[l + 1], Blt;
                if (lin != 0) {
                    double h, dh;
                    orc_rlin(dOD[l], &h, &dh);
                    if (lin == 2) {
                        Blt = Blev[l] + (Blev[l + 1] - Blev[l]) * h;
                        dTau[l] = (cumTxB2[l + 1] * Blt - rad + dTran * dh * (Blev[l + 1] - Blev[l])) * sec_dn;
                        dTdn[l + 1] = dTdn[l + 1] + dTran * dBlev[l + 1] * h;
                        dTdn[l] = dTran * dBlev[l] * (1. - h);
                        dTbar[l] = 0.0;
                    } else {
                        Blt = 2.0 * Bbar[l] * h + Blev[l] * (1.0 - 2.0 * h);
                        dTau[l] = (cumTxB2[l + 1] * Blt - rad + 2.0 * dTran * dh * (Bbar[l] - Blev[l])) * sec_dn;
                        dTdn[l] = dTran * dBlev[l] * (1.0 - 2.0 * h);
                        dTbar[l] = dTran * 2.0 * dBbar[l] * h;
                    }
                    rad = rad + dTran * Blt;
                } else {
                    dTau[l] = (cumTxB2[l + 1] * Bbar[l] - rad) * sec_dn;
                    rad = rad + dTran * Bbar[l];
                    dTbar[l] = dTran * dBbar[l];
                }
            }
        }
        /* 2- scale by the surface reflectance */
        double rsfc = (1. - emis);
        for (int l = 1; l <= L; l++) {
            dTau[l] = dTau[l] * rsfc;
            dTbar[l] = dTbar[l] * rsfc;
        }
        /* 3- surface emission */
        double dEmis = totTxUp * Bsfc - rad;
        double dTsfc = emis * totTxUp * dBsfc;
        double sfcEmit = emis * totTxUp * Bsfc;
        rad = rad * rsfc + sfcEmit;
        /* 5- upwelling thermal radiance */
        for (int l = 1; l <= obsLay - 1; l++) {
            double dTran = cumTxA[l + 1] - cumTxA[l], Blt;
            if (lin != 0) {
                double h, dh;
                orc_rlin(vOD[l], &h, &dh);
                if (lin == 2) {
                    Blt = Blev[l + 1] * (1 - h) + Blev[l] * h;
                    dTau[l] = dTau[l] + (cumTxA[l] * Blt - rad + dTran * dh * (Blev[l] - Blev[l + 1]));
                    dTup[l + 1] = dTran * dBlev[l + 1] * (1 - h);
                    dTup[l] = dTup[l] + dTran * dBlev[l] * h;
                } else {
                    Blt = 2.0 * Bbar[l] * h + Blev[l + 1] * (1.0 - 2.0 * h);
                    dTau[l] = dTau[l] + (cumTxA[l] * Blt - rad + dTran * dh * 2.0 * (Bbar[l] - Blev[l + 1]));
                    dTup[l + 1] = dTran * dBlev[l + 1] * (1.0 - 2.0 * h);
                    dTbar[l] = dTbar[l] + dTran * 2.0 * dBbar[l] * h;
                }
                rad = rad + dTran * Blt;
            } else {
                dTau[l] = dTau[l] + (cumTxA[l] * Bbar[l] - rad);
                rad = rad + dTran * Bbar[l];
                dTbar[l] = dTbar[l] + dTran * dBbar[l];
            }
        }
        /* level temperature derivatives */
        if (lin != 0)
            for (int k = 1; k <= L + 1; k++) dTlev[k] = dTup[k] + dTdn[k] * rsfc;
        if (lin == 0 || lin == 1)
            for (int l = L; l >= 1; l--) {
                dTlev[l + 1] = dTlev[l + 1] + dTbar[l] * dt_dtup[l - 1];
                dTlev[l] = dTlev[l] + dTbar[l] * dt_dtlo[l - 1];
            }
        rad_o[i] = rad;
        if (dr_dtau)
            for (int l = 1; l <= L; l++) dr_dtau[(int64_t)(l - 1) * npts + i] = dTau[l];
        if (dr_dtlev)
            for (int k = 1; k <= L + 1; k++) dr_dtlev[(int64_t)(k - 1) * npts + i] = dTlev[k];
        if (dr_dtskin) dr_dtskin[i] = dTsfc;
        if (dr_demis) dr_demis[i] = dEmis;
    }
    free(Blev); free(dBlev); free(Bbar); free(dBbar); free(cumTxA); free(cumTxB2);
    free(dTau); free(dTbar); free(dTup); free(dTdn); free(dTlev); free(vOD); free(dOD);
}

/* Jacobians (clblm_noScattJac.f90:346-716) with every branch: the path ODs are the
 * layer's vertical OD od[l*npts+i] times the air-mass factors amf_view[l], amf_down[l],
 * amf_sun[l] (pathOD_up_dn_sun's air-mass-scaled paths, :826-838), the observer at level
 * obs (0 = surface .. nlay = top of the atmosphere), up-looking (the down-welling
 * radiance at the observer from the layers above it, :436-494) or down-looking (the
 * reflected downwelling, the surface terms, the solar term with sol_toa[i] and
 * rsfc_sun when solar is set, and the upwelling of the layers below the observer,
 * :496-678), then the level assembly (:681-715).  Layer 0 is at the surface.  Outputs
 * rad, dr_dtau[l*npts+i] (with respect to the vertical OD), dr_dtlev[k*npts+i],
 * dr_dtskin, dr_demis, dr_drsfc (each may be NULL).  The summation order is the
 * reference's. */
void orc_jacobians_paths(int nlay, int64_t npts, double v1, double dv, const double *od, const double *amf_view,
                         const double *amf_down, const double *amf_sun, const double *t_lay, const double *t_lev,
                         double emis, double tskin, int solar, const double *sol_toa, double rsfc_sun, int obs,
                         int uplook, int lin, const double *dt_dtup, const double *dt_dtlo, double *rad_o,
                         double *dr_dtau, double *dr_dtlev, double *dr_dtskin, double *dr_demis, double *dr_drsfc) {
    const int L = nlay, N1 = 1, N2 = nlay;
    const int obsLev = obs + 1;   /* 1-based level */
    double *Blev = malloc(sizeof(double) * (L + 2)), *dBlev = malloc(sizeof(double) * (L + 2));
    double *Bbar = malloc(sizeof(double) * (L + 1)), *dBbar = malloc(sizeof(double) * (L + 1));
    double *cumTxA = malloc(sizeof(double) * (L + 2)), *cumTxB = malloc(sizeof(double) * (L + 2));
    double *cumTxB2 = malloc(sizeof(double) * (L + 2));
    double *dTau = malloc(sizeof(double) * (L + 1)), *dTbar = malloc(sizeof(double) * (L + 1));
    double *dTup = malloc(sizeof(double) * (L + 2)), *dTdn = malloc(sizeof(double) * (L + 2));
    double *dTlev = malloc(sizeof(double) * (L + 2));
    double *vOD = malloc(sizeof(double) * (L + 1)), *dOD = malloc(sizeof(double) * (L + 1));
    double *sOD = malloc(sizeof(double) * (L + 1));
    for (int64_t i = 0; i < npts; i++) {
        const double vn = v1 + (double)i * dv;
        for (int l = 1; l <= L; l++) {
            const double o = od[(int64_t)(l - 1) * npts + i];
            vOD[l] = o * amf_view[l - 1];
            dOD[l] = o * amf_down[l - 1];
            sOD[l] = o * amf_sun[l - 1];
        }
        if (lin != 0)
            for (int k = 1; k <= L + 1; k++) orc_planck_drv(vn, t_lev[k - 1], &Blev[k], &dBlev[k]);
        for (int l = 1; l <= L; l++) orc_planck_drv(vn, t_lay[l - 1], &Bbar[l], &dBbar[l]);
        double Bsfc, dBsfc;
        orc_planck_drv(vn, tskin, &Bsfc, &dBsfc);
        double rad = 0., radsun = 0., dEmis = 0., dTsfc = 0., dRsfc = 0.;
        for (int l = 1; l <= L; l++) { dTau[l] = 0.; dTbar[l] = 0.; }
        for (int k = 1; k <= L + 1; k++) { dTup[k] = 0.; dTdn[k] = 0.; dTlev[k] = 0.; }
        int obsLay = obsLev;
        double rsfc = 0.;
        if (uplook) {
            for (int k = N1; k <= obsLev; k++) cumTxB[k] = 1.0;
            double cumODB = 0.0;
            obsLay = obsLay > N1 ? obsLay : N1;
            for (int l = obsLay; l <= N2; l++) {
                cumODB = cumODB + vOD[l];
                cumTxB[l + 1] = exp(-cumODB);
            }
            for (int l = N2; l >= obsLay; l--) {
                const double dTran = cumTxB[l] - cumTxB[l + 1];
                double Blt;
                if (lin != 0) {
                    double h, dh;
                    orc_rlin(vOD[l], &h, &dh);
                    if (lin == 2) {
                        Blt = Blev[l] * (1 - h) + Blev[l + 1] * h;
                        dTau[l] = (cumTxB[l + 1] * Blt - rad + dTran * dh * (Blev[l + 1] - Blev[l])) * amf_view[l - 1];
                        dTdn[l + 1] = dTdn[l + 1] + dTran * dBlev[l + 1] * h;
                        dTdn[l] = dTran * dBlev[l] * (1.0 - h);
                        dTbar[l] = 0.0;
                    } else {
                        Blt = 2.0 * Bbar[l] * h + Blev[l] * (1.0 - 2.0 * h);
                        dTau[l] = (cumTxB[l + 1] * Blt - rad + 2.0 * dTran * dh * (Bbar[l] - Blev[l])) * amf_view[l - 1];
                        dTdn[l] = dTran * dBlev[l] * (1.0 - 2.0 * h);
                        dTbar[l] = dTran * 2.0 * dBbar[l] * h;
                    }
                    rad = rad + dTran * Blt;
                } else {
                    dTau[l] = (cumTxB[l + 1] * Bbar[l] - rad) * amf_view[l - 1];
                    rad = dTran * Bbar[l] + rad;
                    dTbar[l] = dTran * dBbar[l];
                }
            }
        } else {
            /* 0- view-path transmittance from the observer down to the surface */
            double cumODA = 0.;
            for (int k = obsLev; k <= N2 + 1; k++) cumTxA[k] = 1.;
            for (int l = obsLay - 1; l >= N1; l--) {
                cumODA = cumODA + vOD[l];
                cumTxA[l] = exp(-cumODA);
            }
            const double totODUp = cumODA, totTxUp = cumTxA[N1];
            /* 1- downwelling thermal radiance */
            if (totTxUp > 1.e-06) {
                cumTxB2[N1] = totTxUp;
                double cumODB = 0.;
                for (int l = N1; l <= N2; l++) {
                    cumODB = cumODB + dOD[l];
                    cumTxB2[l + 1] = exp(-(totODUp + cumODB));
                }
                for (int l = N2; l >= N1; l--) {
                    const double dTran = cumTxB2[l] - cumTxB2[l + 1];
                    double Blt;
                    if (lin != 0) {
                        double h, dh;
                        orc_rlin(dOD[l], &h, &dh);
                        if (lin == 2) {
                            Blt = Blev[l] + (Blev[l + 1] - Blev[l]) * h;
                            dTau[l] = (cumTxB2[l + 1] * Blt - rad + dTran * dh * (Blev[l + 1] - Blev[l])) *
                                      amf_down[l - 1];
                            dTdn[l + 1] = dTdn[l + 1] + dTran * dBlev[l + 1] * h;
                            dTdn[l] = dTran * dBlev[l] * (1. - h);
                            dTbar[l] = 0.0;
                        } else {
                            Blt = 2.0 * Bbar[l] * h + Blev[l] * (1.0 - 2.0 * h);
                            dTau[l] = (cumTxB2[l + 1] * Blt - rad + 2.0 * dTran * dh * (Bbar[l] - Blev[l])) *
                                      amf_down[l - 1];
                            dTdn[l] = dTran * dBlev[l] * (1.0 - 2.0 * h);
                            dTbar[l] = dTran * 2.0 * dBbar[l] * h;
                        }
                        rad = rad + dTran * Blt;
                    } else {
                        dTau[l] = (cumTxB2[l + 1] * Bbar[l] - rad) * amf_down[l - 1];
                        rad = rad + dTran * Bbar[l];
                        dTbar[l] = dTran * dBbar[l];
                    }
                }
            }
            /* 2- scale by the surface reflectance */
            rsfc = (1. - emis);
            for (int l = N1; l <= N2; l++) {
                dTau[l] = dTau[l] * rsfc;
                dTbar[l] = dTbar[l] * rsfc;
            }
            /* 3- surface terms */
            dEmis = totTxUp * Bsfc - rad;
            dTsfc = emis * totTxUp * dBsfc;
            const double sfcEmit = emis * totTxUp * Bsfc;
            rad = rad * rsfc + sfcEmit;
            /* 4- solar */
            if (solar) {
                double totODSun = 0.;
                for (int l = N1; l <= N2; l++) totODSun = totODSun + sOD[l];
                const double totTxSun2 = exp(-(totODUp + totODSun));
                dRsfc = totTxSun2 * sol_toa[i];
                radsun = rsfc_sun * dRsfc;
                for (int l = N1; l <= N2; l++) dTau[l] = dTau[l] - radsun * amf_sun[l - 1];
                rad = rad + radsun;
            }
            /* 5- upwelling thermal radiance of the layers below the observer */
            for (int l = N1; l <= obsLay - 1; l++) {
                const double dTran = cumTxA[l + 1] - cumTxA[l];
                double Blt;
                if (lin != 0) {
                    double h, dh;
                    orc_rlin(vOD[l], &h, &dh);
                    if (lin == 2) {
                        Blt = Blev[l + 1] * (1 - h) + Blev[l] * h;
                        dTau[l] = dTau[l] + (cumTxA[l] * Blt - rad + dTran * dh * (Blev[l] - Blev[l + 1])) *
                                                amf_view[l - 1];
                        dTup[l + 1] = dTran * dBlev[l + 1] * (1 - h);
                        dTup[l] = dTup[l] + dTran * dBlev[l] * h;
                    } else {
                        Blt = 2.0 * Bbar[l] * h + Blev[l + 1] * (1.0 - 2.0 * h);
                        dTau[l] = dTau[l] + (cumTxA[l] * Blt - rad + dTran * dh * 2.0 * (Bbar[l] - Blev[l + 1])) *
                                                amf_view[l - 1];
                        dTup[l + 1] = dTran * dBlev[l + 1] * (1.0 - 2.0 * h);
                        dTbar[l] = dTbar[l] + dTran * 2.0 * dBbar[l] * h;
                    }
                    rad = rad + dTran * Blt;
                } else {
                    dTau[l] = dTau[l] + (cumTxA[l] * Bbar[l] - rad) * amf_view[l - 1];
                    rad = rad + dTran * Bbar[l];
                    dTbar[l] = dTbar[l] + dTran * dBbar[l];
                }
            }
        }
        /* level temperature derivatives */
        if (lin != 0) {
            if (uplook)
                for (int k = obsLev; k <= N2 + 1; k++) dTlev[k] = dTdn[k];
            else
                for (int k = N1; k <= N2 + 1; k++) dTlev[k] = dTup[k] + dTdn[k] * rsfc;
        }
        if (lin == 0 || lin == 1) {
            const int L2 = uplook ? obsLay : N1;
            for (int l = N2; l >= L2; l--) {
                dTlev[l + 1] = dTlev[l + 1] + dTbar[l] * dt_dtup[l - 1];
                dTlev[l] = dTlev[l] + dTbar[l] * dt_dtlo[l - 1];
            }
        }
        rad_o[i] = rad;
        if (dr_dtau)
            for (int l = 1; l <= L; l++) dr_dtau[(int64_t)(l - 1) * npts + i] = dTau[l];
        if (dr_dtlev)
            for (int k = 1; k <= L + 1; k++) dr_dtlev[(int64_t)(k - 1) * npts + i] = dTlev[k];
        if (dr_dtskin) dr_dtskin[i] = dTsfc;
        if (dr_demis) dr_demis[i] = dEmis;
        if (dr_drsfc) dr_drsfc[i] = dRsfc;
    }
    free(Blev); free(dBlev); free(Bbar); free(dBbar); free(cumTxA); free(cumTxB); free(cumTxB2);
    free(dTau); free(dTbar); free(dTup); free(dTdn); free(dTlev); free(vOD); free(dOD); free(sOD);
}

/* drv_LevelTemp and drv_LevelConc (clblm_noScattJac.f90:899-1203): level Jacobians
 * from the per-point derivatives and the layer-OD variants.  od_var holds 3 + nmol
 * blocks of [nlay][npts]: the layer OD, the OD at Tave + dt, the OD at
 * Wtot + dw_frac*Wtot, then the OD of each molecule alone.  dtau/dT is
 * (od(T+dt) - od)/dt - dtau/dWtot * Wtot/Tave (the reference: dt = 1 K, dw_frac = 0.01).
 * Outputs jac_t[k*npts+i] (nlay+1 levels) and jac_q[(m*(nlay+1)+k)*npts+i]
 * (derivatives with respect to ln of the level density). */
void orc_level_jac(int nlay, int nmol, int64_t npts, const double *od_var, const double *dr_dtau,
                   const double *dr_dtlev, const double *wtot, const double *tave, double dt, double dw_frac,
                   const double *dt_dtup, const double *dt_dtlo, const double *dw_dqup, const double *dw_dqlo,
                   double *jac_t, double *jac_q) {
    const int64_t LN = (int64_t)nlay * npts;
    const double *od = od_var, *odp = od_var + LN, *odw = od_var + 2 * LN;
    for (int64_t i = 0; i < npts; i++) {
        if (jac_t) {
            double lo = 0., up = 0.;
            for (int l = 0; l < nlay; l++) {
                int64_t j = (int64_t)l * npts + i;
                double dTau_dT = (odp[j] - od[j]) / dt;
                double deltaW = dw_frac * wtot[l];
                double dTau_dW = (odw[j] - od[j]) / deltaW;
                double fac = wtot[l] / tave[l];
                dTau_dT = dTau_dT - dTau_dW * fac;
                double dRaddT = dr_dtau[j] * dTau_dT;
                up = dRaddT * dt_dtup[l];
                lo = lo + dRaddT * dt_dtlo[l] + dr_dtlev[j];
                jac_t[j] = lo;
                lo = up;
            }
            jac_t[LN + i] = lo + dr_dtlev[LN + i];
        }
        if (jac_q)
            for (int m = 0; m < nmol; m++) {
                const double *odm = od_var + (3 + (int64_t)m) * LN;
                double *jq = jac_q + (int64_t)m * (nlay + 1) * npts;
                double lo = 0., up = 0.;
                for (int l = 0; l < nlay; l++) {
                    int64_t j = (int64_t)l * npts + i;
                    double dRad_dW = dr_dtau[j] * odm[j];
                    up = dRad_dW * dw_dqup[(int64_t)m * nlay + l];
                    lo = lo + dRad_dW * dw_dqlo[(int64_t)m * nlay + l];
                    jq[j] = lo;
                    lo = up;
                }
                jq[LN + i] = lo;
            }
    }
}

/* ==== Instrument scan ===========================================================
 * clblm_FFTSCN_struct (clblm_FFT.f90:352-844) for the scanning functions with an
 * analytic apodization (IVX = 1), with LoadSpect (:1114-1334), Scntrn/Scnfnt
 * (:2070-2244, 1411-1958), Multrn (:1334-1411), wrtSpect (:2260-2382) and the
 * 4-point Intpdr/clblm_INTERP/panelInput (clblm_FFT.f90:2392-2443,
 * clblm_ScanFilter.f90:1975-2375).  The FFT itself is left to the caller. */

#define SCAN_LPTSMX 65536
static const double SCAN_C[14] = {0.0, 2.0, 0.849322, 2.257609, 3.314800, 2.100669, 2.195676,
                                  2.0, 2.570274, 2.367714, 2.071759, -3.3, -2.5, 3.314800};
static const double SCAN_CRATIO[14] = {0.0, 40., 10., 40., 160., 20., 20., 20., 40., 40., 20., 100., 10., 1e300};
static const double SCAN_CLIMIT[14] = {1., 2., 3., 40., 160., 20., 20., 20., 40., 40., 20., 100., 10., 160.};

typedef struct {
    double v1;      /* spectrum start */
    double dv;
    int64_t n;
} scan_grid;

/* panelInput (clblm_ScanFilter.f90:2302-2370): the next 2400-point panel of the
 * input that reaches VBOT, into T(5..NNI) (1-based T), panel frequencies accumulated
 * as LBLRTM does.  Returns 0 at the end of the input (outputs untouched). */
static int scan_panel(const double *x, const scan_grid *g, double *T, double vbot, double *vmin_o,
                      double *vmax_o, double *dvi_o, int *nni_o, int64_t *done, double *done_v2) {
    int nlim;
    double vmin, vmax, dvi;
    for (;;) {
        nlim = 2400;
        if (*done + nlim > g->n) nlim = (int)(g->n - *done);
        if (nlim == 0) return 0;
        dvi = g->dv;
        vmin = *done_v2 + dvi;
        vmax = vmin + (double)(nlim - 1) * dvi;
        *done += nlim;
        *done_v2 = vmax;
        if (vmax >= vbot) break;
    }
    const int nlow = 4 + 1;
    vmin = vmin - (double)(nlow - 1) * dvi;
    int nni = nlim + nlow - 1;
    for (int i = 0; i < nlim; i++) T[nlow + i] = x[*done - nlim + i];
    *vmin_o = vmin; *vmax_o = vmax; *dvi_o = dvi; *nni_o = nni;
    return 1;
}

/* Intpdr + clblm_INTERP with 4-point interpolation.  x[0..n) on v1 + i dv onto the
 * grid vv1 + j dvo up to vv2.  Writes min(count, cap) points; returns the count and
 * the output start in *v1res. */
int64_t orc_scan_interp(const double *x, int64_t n, double v1, double dv, double vv1, double vv2, double dvo,
                        double *out, int64_t cap, double *v1res) {
    scan_grid g = {v1, dv, n};
    const double V1C = v1, V2C = v1 + dv * (double)(n - 1);
    double VV1 = vv1, VV2 = vv2;
    if (VV1 < V1C) {
        double AN = (V1C - VV1) / dvo;
        if (AN == (double)(int)AN) VV1 = VV1 + dvo * (double)(int)AN;
        else VV1 = VV1 + dvo * (double)((int)AN + 1);
    }
    if (VV2 > V2C) {
        double AN = (VV2 - V2C) / dvo;
        if (AN == (double)(int)AN) VV2 = VV2 - dvo * (double)(int)AN;
        else VV2 = VV2 - dvo * (double)((int)AN + 1);
    }
    const double V1 = VV1, V2 = VV2, DVO = dvo;
    static double C1[203], C2[203], C3[203], C4[203];
    const double XNUMCF = 201.;
    for (int ip = 0; ip <= 202; ip++) {
        double P = ((double)ip - 1.0) / (XNUMCF - 1.0);
        double PP = P * P;
        C1[ip] = -(P / 2.0 * ((1 - P) * (1 - P)));
        C2[ip] = 1.0 - PP * (3.0 - 2.0 * P) + PP / 2.0 * (1.0 - P);
        C3[ip] = PP * (3.0 - 2.0 * P) + P / 2.0 * ((1.0 - P) * (1.0 - P));
        C4[ip] = -(PP / 2. * (1.0 - P));
    }
    const double BOUND = 4.0 * DVO;
    const double VBOT = V1 - BOUND;
    double T[2412];
    double R[2402];
    int64_t nout = 0, done = 0;
    double done_v2 = v1 - dv;
    for (int i = 1; i <= 4; i++) T[i] = 0.0;
    double V1I = 0, V2I = 0, DVI = 0;
    int NNI = 0;
    int eof = !scan_panel(x, &g, T, VBOT, &V1I, &V2I, &DVI, &NNI, &done, &done_v2);
    const double DVJ = DVO;
    double V1J = V1, VJ = V1J, V2J;
    const double VRATIO = DVJ / DVI;
    T[4] = 2.0 * T[5] - T[6];
#define EMIT_PANEL(nnj)                                                   \
    do {                                                                  \
        for (int q = 1; q <= (nnj); q++) {                                \
            if (nout < cap) out[nout] = R[q];                             \
            nout++;                                                       \
        }                                                                 \
    } while (0)
    while (V1J < V1I) {
        int J2 = (int)((V1I - V1J) / DVJ) + 1;
        if (J2 > 2400) J2 = 2400;
        for (int j = 1; j <= J2; j++) R[j] = 0.0;
        V2J = V1J + DVJ * (double)(J2 - 1);
        EMIT_PANEL(J2);
        V1J = V2J + DVJ;
        VJ = V1J;
    }
    int J2 = 0;
    for (;;) {
        int I = (int)((VJ - V1I) / DVI + 1.00001);
        if (I < 1) break;
        double VI = V1I + DVI * (double)(I - 1);
        double P = (VJ - VI) / DVI;
        int J1 = (int)((VJ - V1J) / DVJ + 1.001);
        int a = (int)((V2 - V1J) / DVJ + 1.001), b = (int)((V2I - DVI - V1J) / DVJ + 1.);
        J2 = a < b ? a : b;
        if (J2 > 2400) J2 = 2400;
        for (int J = J1; J <= J2; J++) {
            int IP = (int)(P * XNUMCF + 1.00001);
            R[J] = C1[IP] * T[I - 1] + C2[IP] * T[I] + C3[IP] * T[I + 1] + C4[IP] * T[I + 2];
            P = P + VRATIO;
            if (P >= 1.0) {
                I = (int)((double)I + P);
                P = P - (double)(int)P;
            }
        }
        VJ = V1J + DVJ * (double)J2;
        if (J2 >= 2400 || VJ >= V2) {
            V2J = V1J + DVJ * (double)(J2 - 1);
            EMIT_PANEL(J2);
            V1J = V2J + DVJ;
            J2 = 0;
        }
        if (VJ >= V2) goto done_interp;
        if (eof) break;
        if (I >= NNI - 2) {
            for (int q = 1; q <= 4; q++) T[q] = T[q + NNI - 4];
            double a1 = V1I, a2 = V2I, a3 = DVI;
            int a4 = NNI;
            if (scan_panel(x, &g, T, VBOT, &a1, &a2, &a3, &a4, &done, &done_v2)) {
                V1I = a1; V2I = a2; DVI = a3; NNI = a4;
            } else {
                eof = 1;
                T[NNI + 1] = 2.0 * T[NNI] - T[NNI - 1];
                T[NNI + 2] = 0.0;
                V2I = V2I + DVI;
            }
        }
    }
    /* label 80: zero fill to V2 */
    for (;;) {
        int J1 = J2 + 1;
        J2 = (int)((V2 - V1J) / DVJ + 1.0001);
        if (J2 > 2400) J2 = 2400;
        for (int J = J1; J <= J2; J++) R[J] = 0.0;
        VJ = V1J + DVJ * (double)J2;
        if (J2 >= 2400 || VJ >= V2) {
            V2J = V1J + DVJ * (double)(J2 - 1);
            EMIT_PANEL(J2);
            V1J = V2J + DVJ;
            J2 = 0;
        }
        if (!(VJ < V2)) break;
    }
done_interp:
#undef EMIT_PANEL
    if (v1res) *v1res = V1;
    return nout;
}

/* The bookkeeping of clblm_FFTSCN_struct for a spectrum on v1c + i dv (i < n),
 * scan function functID (2..8: triangle, Gaussian, sinc**2, sinc, Beer, Hamming,
 * Hanning) of half width hwhm, limits v1out..v2out.  The re-gridded spectrum (when
 * pre_interp) starts at the grid point returned in v1z with dv = dvn.  Returns 0,
 * or -1 with a message. */
/* clblm_FFTSCN_struct's set-up (clblm_FFT.f90:434-614): the scanning function (JFN =
 * funct_id - 1; 101..113 and -2 are the externally pre-scanned and deconvolution-only
 * modes), the boxcar pre-scan and deconvolution switches, the scan window V1Z..V2Z
 * and the optional re-gridding onto DVN.  Brault and Kaiser-Bessel (JFN 11, 12) stop
 * in the reference (they need an optical path difference, :480-489); Kiruna (JFN 13)
 * is not restated (its multi-record scanning function sets the next block's start
 * from a variable that is never assigned, Scntrn :2136). */
int orc_fftscn_plan2(double v1c, double dv, int64_t n, double v1out, double v2out, int funct_id, double hwhm,
                     double fill, double pre_box_hw, int deconv_pre_box, orc_scan_plan *p) {
    memset(p, 0, sizeof *p);
    if (fill < 0 || fill > 1) return fail("FILL value is out of range");
    if (!((funct_id >= 1 && funct_id <= 14) || (funct_id >= 102 && funct_id <= 114) || funct_id == -1))
        return fail("functID must be in [1,14], [102,114] or -1");
    int JFN = funct_id - 1;
    if (JFN == 0) {
        p->prebox = 1;
        p->deconv = 0;
        p->deconv_a = hwhm;   /* boxcar half width until the boxcar factor is known */
    } else if (JFN >= 1 && JFN <= 13) {
        if (pre_box_hw > 0.) {
            p->prebox = 1;
            p->deconv = deconv_pre_box != 0;
            p->deconv_a = pre_box_hw;
        }
    } else if (JFN >= 101 && JFN <= 113) {
        JFN -= 100;
        p->deconv = 1;
        p->deconv_a = pre_box_hw;
    } else if (JFN == -2) {
        p->deconv_only = 1;
        JFN = 0;
        p->deconv = 1;
        p->deconv_a = pre_box_hw;
    }
    if (JFN == 11 || JFN == 12) return fail("this scanning function needs the interferometer optical path difference");
    if (JFN == 13) return fail("the Kiruna scanning function is not supported");
    p->jfn = JFN;
    p->hwhm = hwhm;
    p->a = JFN > 0 ? SCAN_C[JFN] * hwhm : 0.;
    double V2C = v1c + dv * (double)(n - 1);
    if (v1out >= v2out) return fail("v1out >= v2out");
    if (v1out >= V2C || v2out <= v1c) return fail("v1out..v2out outside the data");
    double V1Z, V2Z;
    if (p->deconv_only) {
        V1Z = v1c;
        V2Z = V2C;
    } else {
        V1Z = v1out - hwhm * SCAN_CLIMIT[JFN];
        V2Z = v2out + hwhm * SCAN_CLIMIT[JFN];
        V1Z = v1c + (double)(int)((V1Z - v1c) / dv) * dv;
        V2Z = v1c + (double)(int)((V2Z - v1c) / dv + 1.01) * dv;
    }
    if (V1Z < v1c) V1Z = v1c;
    if (V2Z > V2C) V2Z = V2C;
    if ((V2Z - V1Z) / hwhm < SCAN_CLIMIT[JFN]) return fail("insufficient data for the scan");
    double DVN = 2. * hwhm / 12;
    double DV = dv;
    p->dvn = DVN;
    if (DV > DVN && JFN != 0) {
        p->pre_interp = 1;
        /* Intpdr(spect, V1Z, V2Z, DVN): the re-gridded spectrum starts at the clamped
         * V1Z with DVN; its end follows from the point count (computed by the caller) */
        double a = V1Z;
        if (a < v1c) {
            double AN = (v1c - a) / DVN;
            if (AN == (double)(int)AN) a = a + DVN * (double)(int)AN;
            else a = a + DVN * (double)((int)AN + 1);
        }
        V1Z = a;
        DV = DVN;
    }
    p->v1z = V1Z;
    p->v2z = V2Z;
    p->dv = DV;
    return 0;
}

int orc_fftscn_plan(double v1c, double dv, int64_t n, double v1out, double v2out, int funct_id, double hwhm,
                    orc_scan_plan *p) {
    if (funct_id < 2 || funct_id > 8) return fail("scan function not supported (2..8 only)");
    return orc_fftscn_plan2(v1c, dv, n, v1out, v2out, funct_id, hwhm, 0., -1., 0, p);
}

static int64_t scan_records(int64_t ntot) {  /* LoadSpect/Boxcar record count (:1189-1197, :890-910) */
    double ARDATA = (double)ntot / (double)SCAN_LPTSMX;
    int64_t LRDATA = (ARDATA - (double)(int)ARDATA == 0.0) ? (int64_t)(int)ARDATA : (int64_t)(int)(ARDATA + 1);
    double POWER = log((double)LRDATA) / log(2.0);
    return (POWER - (double)(int)POWER == 0) ? ((int64_t)1 << (int)(POWER + 2)) : ((int64_t)1 << ((int)POWER + 3));
}

/* Second half of the bookkeeping once the scan-grid spectrum is known (v1p, its
 * n points): the choice of apodization (IVX, :616-634), LoadSpect's count and
 * record rounding, the boxcar pre-scan's new grid (:689-706), the transform length,
 * and wrtSpect's window (:776-796). */
int orc_fftscn_finish(orc_scan_plan *p, double v1p, int64_t np, double v1out, double v2out) {
    double DV = p->dv;
    double V1Z = p->v1z, V2Z = p->v2z;
    if (p->pre_interp) {
        V1Z = v1p;
        V2Z = v1p + DV * (double)(np - 1);
    }
    if (p->jfn == 0) p->ivx = 0;
    else p->ivx = ((V2Z - V1Z) / p->hwhm > SCAN_CRATIO[p->jfn]) ? 1 : -1;
    if (p->ivx == -1 && p->jfn >= 8)
        return fail("Norton-Beer scanning functions have no spectral-domain form (range too short)");
    if (v1p > V1Z) return fail("LoadSpect: first data frequency above V1Z");
    int64_t LTOTAL = (int64_t)(int)((V2Z - V1Z) / DV + 1.1);
    V2Z = V1Z + DV * (double)(LTOTAL - 1);
    int64_t LREC = scan_records(LTOTAL);
    int64_t N1L = (int64_t)(int)((V1Z - v1p) / DV + 1.00001);
    if (N1L + LTOTAL - 1 > np) return fail("LoadSpect: ran out of data");
    p->n1load = N1L;
    p->ltotal0 = LTOTAL;
    p->lrec0 = LREC;
    p->v1z0 = V1Z;
    p->dv0 = DV;
    if (p->prebox) {
        const double boxhw = p->deconv_a;
        int M = (int)(2.0 * boxhw / DV);
        if (M >= 2) {
            V1Z = V1Z + DV * (double)(M - 1) / 2.0;
            int64_t KTOTAL = LTOTAL / M;
            if (KTOTAL < 1) return fail("boxcar wider than the scan range");
            DV = (double)M * DV;
            V2Z = V1Z + DV * (double)(KTOTAL - 1);
            LREC = scan_records(KTOTAL);
            LTOTAL = KTOTAL;
            p->m = M;
            p->deconv_a = DV;
        } else {
            p->prebox = 0;
            p->deconv = 0;
        }
    }
    p->fft = !(p->jfn == 0 && !p->deconv);
    p->v1z = V1Z; p->v2z = V2Z; p->dv = DV;
    p->ltotal = LTOTAL;
    p->lrec = LREC;
    p->nfft = LREC * SCAN_LPTSMX;
    int64_t N1 = (int64_t)(int)((v1out - V1Z) / DV);
    double V1S = V1Z + (double)(N1 - 1) * DV;
    if (V1S < V1Z) { N1 = 1; V1S = V1Z; }
    int64_t N2 = (int64_t)(int)((v2out - V1Z) / DV + 3.01);
    double V2S = V1Z + (double)(N2 - 1) * DV;
    if (V2S > V2Z) {
        N2 = (int64_t)(int)((V2Z - V1Z) / DV + 1.01);
        V2S = V1Z + (double)(N2 - 1) * DV;
    }
    (void)V2S;
    p->n1 = N1;
    p->ntotal = N2 - N1 + 1;
    p->v1s = V1S;
    return 0;
}

/* Boxcar (clblm_FFT.f90:867-990): means of m consecutive loaded points, sequential
 * sums; the transform buffer (nfft points) is filled with `fill` beyond them. */
void orc_scan_boxcar(const orc_scan_plan *p, const double *loaded, double fill, double *out) {
    for (int64_t k = 0; k < p->nfft; k++) {
        if (k < p->ltotal) {
            double sum = 0.0;
            for (int mm = 0; mm < p->m; mm++) sum = sum + loaded[k * p->m + mm];
            out[k] = sum / (double)p->m;
        } else {
            out[k] = fill;
        }
    }
}

/* Scnfnt's scanning functions (IAPSC = -1) for JFN 1..7 at X >= 0. */
static double scan_sfn(int jfn, double X, double A) {
    const double Pi = 2.0 * asin(1.0);
    switch (jfn) {
    case 1: return (X > A) ? 0.0 : (1.0 - X / A);
    case 2: { double t = 0.5 * ((X / A) * (X / A)); return exp(-(t < 34. ? t : 34.)); }
    case 3: if (X != 0) { double s = sin(Pi * X / A) / (Pi * X / A); return s * s; } return 1.0;
    case 4: return (X != 0) ? sin(2.0 * Pi * X / A) / (2.0 * Pi * X / A) : 1.0;
    case 5: {
        double U = 2.0 * Pi * X / A;
        if (U >= 0.5) {
            double U2 = U * U, U5 = U2 * U2 * U;
            return 15. * ((3.0 - U2) * sin(U) - 3.0 * U * cos(U)) / U5;
        }
        return (1.0 - 15.0 / 210 * (U * U));
    }
    default: {
        const double c = jfn == 6 ? 0.428752 : 0.5;
        double U = 2.0 * Pi * X / A;
        if (U != 0 && fabs(U) != Pi) return (sin(U) / U + c * (sin(U + Pi) / (U + Pi) + sin(U - Pi) / (U - Pi)));
        if (U == 0) return 1.0;
        return c;
    }
    }
}

/* Scntrn with IVX = -1 (clblm_FFT.f90:2113-2175): the scanning function on the
 * transform grid, block by block, for an even number of records: block J holds
 * positive frequencies from V(J) = V(J-1) + LPTS DV, and block LREC+1-J their mirror
 * images (first element the value at V(J+1)).  The caller transforms it and divides
 * by its first (DC) coefficient. */
void orc_scan_sfn(const orc_scan_plan *p, double *f) {
    const int LPTS = SCAN_LPTSMX;
    const double DV = p->dv;
    double *fv = malloc(sizeof(double) * (LPTS + 1));
    double V = 0.0;
    for (int64_t J = 1; J <= p->lrec / 2; J++) {
        const double VSAVE = V;
        for (int L = 1; L <= LPTS; L++) {
            double X = (L == 1) ? VSAVE : VSAVE + (double)(L - 1) * DV;
            fv[L] = scan_sfn(p->jfn, X, p->a);
        }
        V = VSAVE + (double)LPTS * DV;
        double *pos = f + (J - 1) * LPTS, *neg = f + (p->lrec - J) * LPTS;
        for (int L = 1; L <= LPTS; L++) pos[L - 1] = fv[L];
        neg[0] = scan_sfn(p->jfn, V, p->a);
        for (int q = 1; q < LPTS; q++) neg[q] = fv[LPTS + 1 - q];
    }
    free(fv);
}

/* The deconvolution factor (Scnfnt function -1, :1920-1946): Z/sin(Z), Z = 2 pi X A,
 * at the positions of the IVX = 1 apodization, for the boxcar of width parameter A. */
void orc_fft_deconv(const orc_scan_plan *p, double *h) {
    const int LPTS = SCAN_LPTSMX;
    const double PI = 2.0 * asin(1.0);
    const double A = p->deconv_a;
    const double DX = 1.0 / ((double)(LPTS * p->lrec) * p->dv);
    double X = 0.0;
    for (int64_t J = 0; J < p->lrec; J++) {
        const double X0 = X;
        double *f = h + J * (LPTS / 2);
        for (int L = 1; L <= LPTS - 1; L += 2) {
            double Z = 2.0 * PI * X * A;
            f[(L - 1) / 2] = (Z != 0) ? Z / sin(Z) : 1.0;
            X = X0 + (double)(L + 1) * DX / 2.;
        }
    }
}

/* Scntrn with IVX = 1 (clblm_FFT.f90:2217-2229) and Scnfnt's apodization branch:
 * h[m] for the m-th complex coefficient of the packed real transform, m < nfft/2.
 * The Nyquist coefficient is multiplied by the imaginary slot of h's first element,
 * which Scnfnt sets to 0. */
void orc_fft_apod(const orc_scan_plan *p, double *h) {
    const int LPTS = SCAN_LPTSMX;
    const double PI = 2.0 * asin(1.0);
    const double A = p->a;
    const double DX = 1.0 / ((double)(LPTS * p->lrec) * p->dv);
    double X = 0.0;
    for (int64_t J = 0; J < p->lrec; J++) {
        const double X0 = X, XSTART = X;
        double *f = h + J * (LPTS / 2);
        for (int L = 1; L <= LPTS - 1; L += 2) {
            double v;
            switch (p->jfn) {
            case 1: v = (X == 0) ? 1.0 : (sin(PI * X * A) / (PI * X * A)) * (sin(PI * X * A) / (PI * X * A)); break;
            case 2: { double t = 2.0 * ((PI * X * A) * (PI * X * A)); v = exp(-(t < 34. ? t : 34.)); } break;
            case 3: v = (X < 1 / A) ? 1.0 - X * A : 0.0; break;
            case 4: v = (X < 1.0 / A) ? 1.0 : 0.0; break;
            case 5: v = (X < 1 / A) ? (1.0 - (X * A) * (X * A)) * (1.0 - (X * A) * (X * A)) : 0.0; break;
            case 6: v = (X < 1 / A) ? 0.53835685 * (1.0 + 0.857504 * cos(PI * X * A)) : 0.0; break;
            case 7: v = (X < 1. / A) ? 0.5 * (1.0 + cos(PI * X * A)) : 0.0; break;
            default: {  /* Norton-Beer weak, medium, strong (:1758-1798) */
                double U = 1.0 - (X * A) * (X * A);
                if (X < 1. / A) {
                    if (p->jfn == 8) v = 0.384093 - 0.087577 * U + 0.703484 * (U * U);
                    else if (p->jfn == 9) v = 0.152442 - 0.136176 * U + 0.983734 * (U * U);
                    else v = 0.045355 + 0.554883 * (U * U) + 0.399782 * ((U * U) * (U * U));
                } else {
                    v = 0.0;
                }
            } break;
            }
            f[(L - 1) / 2] = v;
            X = X0 + (double)(L + 1) * DX / 2.;
        }
        if (p->jfn == 4) {  /* the rectangle's last point, interpolated (Scnfnt :1642-1652) */
            const double ALEN = 1.0 / A;
            int N = (int)((ALEN - XSTART) / DX) + 1;
            double R = (ALEN - XSTART) / DX + 1 - N;
            if (R <= 0.5) { R = R + 1.0; N = N - 1; }
            if (N >= 1 && N <= LPTS / 2 - 1) f[N] = R - 0.5;
        }
    }
}

/* ---------------------------------------------------------------------------
 * clblm_SCANFN_struct (clblm_ScanFilter.f90:70-544) for functID 1 (JFN 0, the
 * rectangle of adjacent boxes, CNVRCT :1478-1698), with panelInput (:2320-2404,
 * NREN = 0) and panelOutput (:2411-2506).  The spectrum x[0..n) lies on
 * v1 + i dv.  This is the scan output_JacArray applies to Jacobian arrays when a
 * boxcar pre-average is asked for (clblm_noScattJac.f90:1341-1409).
 * Returns the output count (writes min(count, cap) values) and the output grid in
 * *v1res, *dvres; -1 on error.  When IRATIO < 2 the reference returns the input
 * unchanged: so does this (count n, grid v1, dv).
 * ------------------------------------------------------------------------- */
typedef struct {   /* CNVRCT's SAVE variables and flags */
    double sumj, rnj;
    int jn, nb, ipc;
} rect_state;

static void rect_cnvrct(const double *S, double HWHM, double *R1, double V1I, double V2I, double DVI, int *NNI,
                        int NLO, int NHI, double V1, double V2, double DVO, int IDATA, int *IPANEL, int *JFLG,
                        rect_state *st) {
    if (*IPANEL == -1) {
        st->sumj = 0.; st->rnj = 0.; st->jn = NLO;
        if (*JFLG == -1) { st->nb = 1; st->ipc = 1; }
    }
    int NNIV2 = (int)((V2I - V1I) / DVI + 1.0001);
    if (*NNI > NNIV2) *NNI = NNIV2;
    while (NLO <= NHI) {
        double VCNT = V1 + (double)(st->nb - 1) * DVO;
        if (VCNT > V2) { *IPANEL = 1; return; }
        double VLFT = VCNT - HWHM, VRGT = VCNT + HWHM;
        double RL = (VLFT - V1I) / DVI + 1, RR = (VRGT - V1I) / DVI + 1;
        int IL = (int)(RL + 0.5), IH = (int)(RR + 0.5);
        double VLBLR = (V1I + (double)(IL - 1) * DVI) + DVI / 2.;
        double WGTL = (VLBLR - VLFT) / DVI;
        double VLBLL = (V1I + (double)(IH - 1) * DVI) - DVI / 2.;
        double WGTR = (VRGT - VLBLL) / DVI;
        int NEP = 2;
        if (IH > *NNI) {
            IH = *NNI;
            *JFLG = 1;
            WGTR = 1.;
            if (IL == IH) { WGTR = 0.; NEP = 1; }
        }
        if (IL <= 1) {
            IL = 1;
            WGTL = 1.;
            if (IL >= IH) { WGTL = 0.; NEP = 1; }
        }
        if (*JFLG == 1 && IDATA == 0 && IL > *NNI) { *IPANEL = 0; *JFLG = 0; return; }
        if (*JFLG == 1 && IDATA == 1) return;
        for (int I = IL + 1; I <= IH - 1; I++) st->sumj = st->sumj + S[I];
        st->sumj = st->sumj + S[IL] * WGTL;
        st->sumj = st->sumj + S[IH] * WGTR;
        st->rnj = st->rnj + (double)(IH - IL + 1 - NEP) + WGTL + WGTR;
        if (*JFLG == 1 && IDATA == 0) { *IPANEL = 0; *JFLG = 0; return; }
        R1[st->jn] = st->sumj / st->rnj;
        st->jn = st->jn + 1;
        st->sumj = 0.; st->rnj = 0.;
        if (st->nb == st->ipc * (NHI - NLO + 1)) {
            *IPANEL = 1;
            st->ipc = st->ipc + 1;
            st->nb = st->nb + 1;
            return;
        }
        st->nb = st->nb + 1;
    }
}

int64_t orc_scanfn_rect(const double *x, int64_t n, double v1, double dv, double v1out, double v2out, double dvout,
                        double hwhm, double *out, int64_t cap, double *v1res, double *dvres) {
    if (!(hwhm > 0.) || n < 1) return fail("rectangle scan: bad half width or empty input");
    const double DVINT = hwhm / 12.;
    const double SAMPLE = hwhm / fabs(dvout);
    double *spec = NULL;
    int64_t ns = n;
    double sv1 = v1, sdv = dv;
    if (dv > DVINT) {   /* clblm_INTERP onto DVINT first (:331-334) */
        double v2 = v1 + dv * (double)(n - 1), r1 = 0.;
        ns = orc_scan_interp(x, n, v1, dv, v1, v2, DVINT, NULL, 0, &r1);
        if (ns <= 0) return -1;
        spec = (double *)malloc(sizeof(double) * (size_t)ns);
        orc_scan_interp(x, n, v1, dv, v1, v2, DVINT, spec, ns, &r1);
        sv1 = r1;
        sdv = DVINT;
    }
    const double *sp = spec ? spec : x;
    const double DVI0 = sdv;
    double DVO = hwhm / SAMPLE;
    const int IRATIO = (int)(DVO / DVI0 + 0.5);
    DVO = (double)IRATIO * DVI0;
    if (IRATIO < 2) {   /* no scan: the (interpolated) input is the output */
        for (int64_t i = 0; i < ns && i < cap; i++) out[i] = sp[i];
        *v1res = sv1; *dvres = sdv;
        free(spec);
        return ns;
    }
    const double HWF = 1.;
    const int NLIMF = 2401, NSHIFT = 32;
    const int NBOUND = (int)((2. * HWF) * SAMPLE + 0.01);
    const double BOUND = (double)NBOUND * DVO / 2.;
    const int MAXF = NLIMF + 2 * NBOUND + NSHIFT;
    const double V1 = v1out, V2 = v2out;
    int NLO = NSHIFT + 1, NHI = NLIMF + NSHIFT - 1;
    double *R1 = (double *)calloc((size_t)MAXF + 2, sizeof(double));
    double *S = (double *)calloc(2402, sizeof(double));
    int IEOFSC = 1, IDATA = -1, IPANEL = -1, JFLG = -1, NNI = 0, ISTOP = 0;
    double VFT = V1 - (double)NSHIFT * DVO;
    const double VBOT = V1 - BOUND, VTOP = V2 + BOUND;
    int64_t done = 0, nout = 0;
    double done_v2 = sv1 - sdv;
    double V1I = 0., V2I = 0., DVI = sdv;
    rect_state st = {0., 0., 0, 1, 1};
    int state = 40;
    for (;;) {
        if (state == 40) {
            if (IEOFSC <= 0) { state = 60; continue; }
            /* panelInput with NREN = 0 */
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
            IDATA = 0;
            for (int i = 0; i < nlim; i++) S[1 + i] = sp[done - nlim + i];
            if (V2I > VTOP) IDATA = 1;
            state = 50;
        }
        if (state == 50) {
            rect_cnvrct(S, hwhm, R1, V1I, V2I, DVI, &NNI, NLO, NHI, V1, V2, DVO, IDATA, &IPANEL, &JFLG, &st);
            if (IPANEL == 0) { state = 40; continue; }
            state = 60;
        }
        /* 60: panelOutput */
        ISTOP = 0;
        int NNHI = (int)((V2 - VFT) / DVO + 1.5);
        if (NHI >= NNHI) ISTOP = 1;
        if (ISTOP == 1) NHI = NNHI;
        for (int j = NLO; j <= NHI; j++, nout++)
            if (nout < cap) out[nout] = R1[j];
        VFT = VFT + (double)(NLIMF - 1) * DVO;
        if (ISTOP != 1) {
            for (int J = NLIMF; J <= MAXF; J++) R1[J - NLIMF + 1] = R1[J];
            for (int J = MAXF - NLIMF + 2; J <= MAXF; J++) R1[J] = 0.;
            NLO = NSHIFT + 1;
        }
        IPANEL = -1;
        if (ISTOP != 1 && IEOFSC < 0) { state = 60; continue; }
        if (ISTOP != 1 && IEOFSC > 0) { state = 50; continue; }
        break;
    }
    *v1res = V1;
    *dvres = DVO;
    free(R1); free(S); free(spec);
    return nout;
}

/* ---------------------------------------------------------------------------
 * clblm_SCANFN_struct (clblm_ScanFilter.f90:70-544) for functIDs 2..5 (JFN 1..4:
 * triangle, Gaussian, sinc**2, sinc; functID -4 / -5 take the first zero crossing
 * in place of the half width, :272-280): the scanning function tabulated by
 * SHAPET / SHAPEG / SINCSQ / SINC (:551-723), the input compressed by SHRKSC
 * (:732-850, boxes of IRATSH points carried across input panels), convolved by
 * CONVSC (:856-955) into the 5000-point R1 buffer, read by panelInput (:2320-2404)
 * and written by panelOutput (:2411-2506).
 * CONVSC entered with ILO > IHI reads an unset ILAST (:942); this restatement takes
 * that call as processing nothing (ILO stays above IHI), which is what reaching
 * label 60 without a loop pass means for every later call.
 * Returns the output count (writes min(count, cap) values), -1 on error.
 * ------------------------------------------------------------------------- */
typedef struct { double hwf, dxf; int nf, nfmax; double smpl, xscal; } scan_fn_t;
static const scan_fn_t SCN_FN[5] = {   /* HWJ, DXJ, NJ, NJMX, SMPLJ, XSCAL (:182-207) */
    {1., 0.0, 0, 0, .5, 0.},
    {2., 0.02, 101, 251, 2., 0.},
    {4., 0.02, 201, 251, 4., 0.},
    {54.1826, 0.02, 2710, 2760, 4., 1.391557377},
    {119.332818, 0.02, 5968, 6018, 4., 1.89549425},
};

/* SHAPET, SHAPEG, SINCSQ, SINC: xf[1..nfmax] (1-based), zero past nf. */
void orc_scan_shape(int jfn, double *xf) {
    const scan_fn_t *f = &SCN_FN[jfn];
    const double PI = 2. * asin(1.);
    for (int i = 1; i <= f->nfmax; i++) xf[i] = 0.;
    if (jfn == 1) {
        xf[1] = 0.5;
        for (int i = 2; i <= f->nf; i++) {
            double X = (double)(i - 1) * f->dxf;
            xf[i] = 0.5 * (1. - 0.5 * X);
        }
    } else if (jfn == 2) {
        const double FLN2 = log(2.), RECPI = 1. / (2. * asin(1.)), FGNORM = sqrt(FLN2 * RECPI);
        xf[1] = FGNORM * exp(-FLN2 * 0.);
        for (int j = 2; j <= f->nf; j++) {
            double X = (double)(j - 1) * f->dxf, XSQ = X * X;
            xf[j] = FGNORM * exp(-FLN2 * XSQ);
        }
        xf[f->nf] = 0.;
    } else {
        const double XNORM = f->xscal / PI;
        xf[1] = XNORM;
        for (int i = 2; i <= f->nf; i++) {
            double X = (double)(i - 1) * f->dxf, T = X * f->xscal, q = sin(T) / T;
            xf[i] = (jfn == 3) ? XNORM * (q * q) : XNORM * q;
        }
    }
}

int64_t orc_scanfn_conv(const double *x, int64_t n, double v1, double dv, double v1out, double v2out, double dvout,
                        int funct_id, double hwhm, double *out, int64_t cap, double *v1res, double *dvres) {
    const double PI = 2. * asin(1.);
    int JFN = abs(funct_id) - 1;
    if (funct_id < 0) JFN = -JFN;
    if (!(hwhm > 0.) || n < 1) return fail("scan: bad half width or empty input");
    if (JFN == 0 || abs(JFN) > 4) return fail("scan: functID must be 2..5 or -4, -5");
    const scan_fn_t *F = &SCN_FN[abs(JFN)];
    double HWHM = hwhm;
    if (JFN < 0) {
        JFN = -JFN;
        if (JFN != 3 && JFN != 4) return fail("scan: negative functID only for sinc**2 and sinc");
        HWHM = HWHM / (PI / F->xscal);
    }
    const double DVINT = HWHM / 12.;
    const double SAMPL = -fabs(dvout);
    double SAMPLE = F->smpl;
    if (SAMPL < 0.) SAMPLE = HWHM / (-SAMPL);
    double XF[6019];
    orc_scan_shape(JFN, XF);
    double *spec = NULL;
    int64_t ns = n;
    double sv1 = v1, sdv = dv;
    if (dv > DVINT) {   /* clblm_INTERP onto DVINT (:331-334) */
        double v2 = v1 + dv * (double)(n - 1), r1 = 0.;
        ns = orc_scan_interp(x, n, v1, dv, v1, v2, DVINT, NULL, 0, &r1);
        if (ns <= 0) return -1;
        spec = (double *)malloc(sizeof(double) * (size_t)ns);
        orc_scan_interp(x, n, v1, dv, v1, v2, DVINT, spec, ns, &r1);
        sv1 = r1;
        sdv = DVINT;
    }
    const double *sp = spec ? spec : x;
    const double DVO = HWHM / SAMPLE;
    const double HWF = F->hwf, DXF = F->dxf;
    const int NLIMF = 2401, NSHIFT = 32;
    const int NBOUND = (int)((2. * HWF) * SAMPLE + 0.01);
    const double BOUND = (double)NBOUND * DVO / 2.;
    const int MAXF = NLIMF + 2 * NBOUND + NSHIFT;
    if (MAXF > 5000) { free(spec); return fail("scan: output spacing too fine for R1(5000)"); }
    /* CONVSC's table index stays below (ZBOUND + 1/2) ZSLOPE + 1.5; past NJMX the
     * reference reads outside its table */
    if ((HWF / DVO * HWHM + 0.5) * (DVO / DXF / HWHM) * (1. + 1e-9) + 1.5 >= (double)(F->nfmax + 1)) {
        free(spec);
        return fail("scan: output spacing too coarse for the scanning function's table");
    }
    const double V1 = v1out, V2 = v2out;
    int NLO = NSHIFT + 1, NHI = NLIMF + NSHIFT - 1;
    double *R1 = (double *)calloc((size_t)MAXF + 2, sizeof(double));
    double *S = (double *)calloc(3851, sizeof(double)), *SS = (double *)calloc(5001, sizeof(double));
    int IEOFSC = 1, IDATA = -1, IPANEL = -1, ISTOP = 0, INIT = 0, NREN = 0, ILO = 1, IHI = 0, NNI = 0;
    double VFT = V1 - (double)NSHIFT * DVO;
    const double VBOT = V1 - BOUND, VTOP = V2 + BOUND;
    int64_t done = 0, nout = 0;
    double done_v2 = sv1 - sdv;
    double V1I = 0., V2I = 0., DVI = sdv;
    /* SHRKSC's SAVE variables */
    double DVSC = 0., SRATIO = 1., V1SHFT = 0.;
    int IRATSH = 1, IRATM1 = 0;
    static const int JRATIO[24] = {1, 2, 3, 4, 5, 6, 8, 10, 12, 15, 16, 20, 24, 25, 30, 32, 40, 48, 50, 60, 75, 80, 100, 120};
    int state = 40;
    for (;;) {
        if (state == 40) {
            if (IEOFSC <= 0) { state = 60; continue; }
            /* panelInput */
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
            const int NLOW = NREN > 0 ? NREN + 1 : 1;
            V1I = V1I - (double)(NLOW - 1) * DVI;
            NNI = nlim + NLOW - 1;
            IDATA = 0;
            for (int i = 0; i < nlim; i++) S[NLOW + i] = sp[done - nlim + i];
            ILO = 1;
            IHI = NNI;
            if (V2I > VTOP) {
                IHI = (int)((VTOP - V1I) / DVI + 1.5);
                IDATA = 1;
            }
            /* SHRKSC */
            {
                const int NLIMS = NNI;
                for (int i = 1; i <= NREN; i++) S[i] = SS[i];
                NREN = 0;
                int skip = 0;
                if (INIT == 0) {
                    DVSC = HWHM / 12.;
                    IRATSH = (int)(DVSC / DVI + 0.5);
                    for (int i = 2; i <= 24; i++)
                        if (JRATIO[i - 1] > IRATSH) { IRATSH = JRATIO[i - 2]; break; }
                    if (IRATSH > JRATIO[23]) IRATSH = JRATIO[23];
                    if (IRATSH <= 1) skip = 1;
                    else {
                        DVSC = (double)IRATSH * DVI;
                        V1SHFT = (double)(IRATSH - 1) * DVI / 2.;
                        SRATIO = IRATSH;
                        IRATM1 = IRATSH - 1;
                        INIT = 1;
                    }
                }
                if (!skip && IRATSH > 1) {
                    NREN = NNI - (NNI / IRATSH) * IRATSH;
                    V1I = V1I + V1SHFT;
                    const int IMAX = NNI - IRATM1 - NREN;
                    int K = 0;
                    for (int I = 1; I <= IMAX; I += IRATSH) {
                        double SUMK = 0.;
                        K = K + 1;
                        for (int J = I; J <= I + IRATM1; J++) SUMK = SUMK + S[J];
                        S[K] = SUMK / SRATIO;
                    }
                    V2I = V1I + DVSC * (double)(K - 1);
                    NNI = K;
                    DVI = DVSC;
                    ILO = (int)((VBOT - V1I) / DVI + 1.5);
                    if (ILO < 1) ILO = 1;
                    IHI = (int)((VTOP - V1I) / DVI + 1.5);
                    if (IHI > NNI) IHI = NNI;
                    for (int i = 1; i <= NREN; i++) SS[i] = S[NLIMS - NREN + i];
                }
            }
            state = 50;
        }
        if (state == 50) {
            /* CONVSC (JVAR = 0) */
            if (ILO > IHI) {
                IPANEL = 1;
            } else {
                const double RATIO = DVI / DVO, DVODX = DVO / DXF, HWBND = HWF / DVO;
                const double ZINT = (V1I - VFT) / DVO;
                const double ZSLOPE = DVODX / HWHM, ZBOUND = HWBND * HWHM, XNORM = DVI / HWHM;
                int ILAST = IHI, full = 0;
                for (int I = ILO; I <= IHI; I++) {
                    if (S[I] == 0.) continue;
                    const double ZPEAK = (double)(I - 1) * RATIO + ZINT;
                    int JMAX = (int)(ZPEAK + ZBOUND + 1.5);
                    if (JMAX > MAXF) { ILAST = I - 1; full = 1; break; }
                    int JMIN = (int)(ZPEAK - ZBOUND + 1.5);
                    if (JMIN < 1) JMIN = 1;
                    const double SI = XNORM * S[I];
                    double ZF = ((double)(JMIN - 1) - ZPEAK) * ZSLOPE;
                    for (int JF = JMIN; JF <= JMAX; JF++) {
                        int IT = (int)(fabs(ZF) + 1.5);
                        R1[JF] = R1[JF] + SI * XF[IT];
                        ZF = ZF + ZSLOPE;
                    }
                }
                IPANEL = full ? 1 : IDATA;
                ILO = ILAST + 1;
            }
            if (IPANEL == 0) { state = 40; continue; }
            state = 60;
        }
        /* 60: panelOutput */
        ISTOP = 0;
        int NNHI = (int)((V2 - VFT) / DVO + 1.5);
        if (NHI >= NNHI) ISTOP = 1;
        if (ISTOP == 1) NHI = NNHI;
        for (int j = NLO; j <= NHI; j++, nout++)
            if (nout < cap) out[nout] = R1[j];
        if (NHI < NLO) nout += NHI - NLO + 1;   /* a negative count shortens what came before */
        VFT = VFT + (double)(NLIMF - 1) * DVO;
        if (ISTOP != 1) {
            for (int J = NLIMF; J <= MAXF; J++) R1[J - NLIMF + 1] = R1[J];
            for (int J = MAXF - NLIMF + 2; J <= MAXF; J++) R1[J] = 0.;
            NLO = NSHIFT + 1;
        }
        if (ISTOP != 1 && IEOFSC < 0) { state = 60; continue; }
        if (ISTOP != 1 && IEOFSC > 0) { state = 50; continue; }
        break;
    }
    *v1res = V1;
    *dvres = DVO;
    free(R1); free(S); free(SS); free(spec);
    return nout;
}

/* ---------------------------------------------------------------------------
 * MT_CKD water-vapour continuum: the H2O part of CONTNM
 * (clblm_ODLAY_Continuum.f90:356-404) with mt_ckd_h2o_absco, myradfn, myxint and
 * pre_xint (mt_ckd_h2o_module.f90:27-294), on OPDPTH's ABSRB grid
 * (clblm_ODLAY.f90:880-887: V1ABS = INT(V1) - 3, 1 cm-1, to INT(V2 + 3.5)).  The
 * coefficient table (wavenumber, self and foreign coefficients at ref_temp /
 * ref_press, self temperature exponent) is the caller's, read from the reference's
 * absco-ref_wv-mt-ckd.nc.  jrad is OPDPTH's JRAD after its flip; the radiation term
 * is applied unless JRAD is 0 (:365-369).  absrb[0..nptabs) is ADDED to.
 * Returns nptabs, -1 on error.
 * ------------------------------------------------------------------------- */
static void mtckd_xint(double V1A, double V2A, double DVA, const double *A /* 1-based, n */, int n, double AFACT,
                       double VFT, double DVR3, double *R3 /* 1-based */, int N1R3, int N2R3) {
    const double onepl = 1.001, onemi = 0.999;   /* the module's own (mt_ckd_h2o_module.f90:21) */
    const double RECDVA = 1. / DVA;
    int ILO = (int)((V1A + DVA - VFT) / DVR3 + 1. + onemi);
    if (ILO < N1R3) ILO = N1R3;
    int IHI = (int)((V2A - DVA - VFT) / DVR3 + onemi);
    if (IHI > N2R3) IHI = N2R3;
    for (int I = ILO; I <= IHI; I++) {
        const double VI = VFT + DVR3 * (double)(I - 1);
        const int J = (int)((VI - V1A) * RECDVA + onepl);
        const double VJ = V1A + DVA * (double)(J - 1);
        const double P = RECDVA * (VI - VJ);
        const double C = (3. - 2. * P) * P * P;
        const double B = 0.5 * P * (1. - P);
        const double B1 = B * (1. - P);
        const double B2 = B * P;
        const double a3 = (J + 2 <= n) ? A[J + 2] : 0.;   /* past the window only where B2 = 0 */
        const double CONTI = -A[J - 1] * B1 + A[J] * (1. - C + B2) + A[J + 1] * (C + B1) - a3 * B2;
        R3[I] = R3[I] + CONTI * AFACT;
    }
}

int orc_mtckd_h2o(const double *wvn, const double *self_ref, const double *for_ref, const double *texp, int ncoeffin,
                  double ref_temp, double ref_press, double pave, double tave, double wk1, double wtot, double v1,
                  double v2, int jrad, double xself, double xfrgn, double *absrb, int cap) {
    /* OPDPTH's ABSRB grid */
    double V1ABS = (double)(int)v1;
    if (v1 < 0.) V1ABS = V1ABS - 1.;
    V1ABS = V1ABS - 3.;
    const double V2ABS = (double)(int)(v2 + 3. + 0.5);
    const int NPTABS = (int)((V2ABS - V1ABS) / 1. + 1.5);
    if (NPTABS > cap || NPTABS > 5050) return fail("mtckd: ABSRB longer than 5050 points");
    if (!((v2 > -20.0) && (v1 < 20000.) && ((xself > 0.) || (xfrgn > 0.)))) return NPTABS;
    const double h2o_fac = wk1 / wtot;
    const int radflag = jrad != 0;
    /* mt_ckd_h2o_absco */
    const double dvabs = 1.;
    const double dvc = wvn[1] - wvn[0];
    int i = 1;                                    /* 1-based over wvn */
    while (i <= ncoeffin && wvn[i - 1] <= (V1ABS - 2 * dvc)) i++;
    int i1 = i - 1;
    if (i1 == 0) i1 = 1;
    while (i <= ncoeffin && wvn[i - 1] < (V2ABS + 2 * dvc)) i++;
    if (i > ncoeffin) return fail("mtckd: the ABSRB grid reaches past the coefficient table");
    const int i2 = i;
    const int ncoeff = i2 - i1 + 1;
    double *sh = (double *)calloc((size_t)ncoeff + 2, sizeof(double));
    double *fh = (double *)calloc((size_t)ncoeff + 2, sizeof(double));
    const double xkt = tave / RADCN2;
    const double rho_rat = (pave / ref_press) * (ref_temp / tave);
    for (int k = 1; k <= ncoeff; k++) {
        const int m = i1 + k - 1;
        double s = self_ref[m - 1] * pow(ref_temp / tave, texp[m - 1]);
        s = s * h2o_fac * rho_rat;
        double f = for_ref[m - 1];
        f = f * (1 - h2o_fac) * rho_rat;
        if (radflag) {   /* myradfn */
            const double xvi = wvn[m - 1];
            double rad = xvi;
            const double xviokt = xvi / xkt;
            if (xviokt <= 0.01) {
                rad = 0.5 * xviokt * xvi;
            } else if (xviokt <= 10.) {
                const double expvkt = exp(-xviokt);
                rad = xvi * (1 - expvkt) / (1. + expvkt);
            }
            s = s * rad;
            f = f * rad;
        }
        sh[k] = s;
        fh[k] = f;
    }
    const int nptabs = (int)((V2ABS - V1ABS) / dvabs + 1);
    /* pre_xint over the whole table */
    const double v1ss = wvn[0], v2ss = wvn[ncoeffin - 1];
    const int nbnd_v1c = (int)(2 + (v1ss - V1ABS) / dvabs + 1.e-5);
    const int ist = nbnd_v1c > 1 ? nbnd_v1c : 1;
    const int nbnd_v2c = (int)(1 + (v2ss - V1ABS) / dvabs + 1.e-5);
    const int lst = nptabs < nbnd_v2c ? nptabs : nbnd_v2c;
    double *self_abs = (double *)calloc((size_t)NPTABS + 2, sizeof(double));
    double *for_abs = (double *)calloc((size_t)NPTABS + 2, sizeof(double));
    mtckd_xint(wvn[i1 - 1], wvn[i2 - 1], dvc, sh, ncoeff, 1.0, V1ABS, dvabs, self_abs, ist, lst);
    mtckd_xint(wvn[i1 - 1], wvn[i2 - 1], dvc, fh, ncoeff, 1.0, V1ABS, dvabs, for_abs, ist, lst);
    /* CONTNM: cself = wk(1) sh2o xself, cfor = wk(1) fh2o xfrgn, absrb = absrb + cself + cfor */
    for (int k = 1; k <= NPTABS; k++) {
        const double cself = wk1 * self_abs[k] * xself;
        const double cfor = wk1 * for_abs[k] * xfrgn;
        absrb[k - 1] = absrb[k - 1] + cself + cfor;
    }
    free(sh); free(fh); free(self_abs); free(for_abs);
    return NPTABS;
}

/* ---------------------------------------------------------------------------
 * CONTNM's continua other than water vapour (clblm_ODLAY_Continuum.f90:408-1060), each
 * added into absrb (nptabs points on v1abs + i at 1 cm-1, 0-based) in CONTNM's order:
 *   CO2 (FRNCO2 :1333-1400 with the XFACCO2 scaling, CONTNM :415-457);
 *   O3: the Chappuis and Wulf bands (XO3CHP :3061-3205, CONTNM :464-481), the
 *     Hartley-Huggins bands with their temperature terms (O3HHT0-2 :5210-7085, CONTNM
 *     :483-525) and in the UV (O3HHUV :7090-7175, CONTNM :530-566), with CONTNM's save and restore
 *     of ABSRB around 40800 cm-1;
 *   O2: the collision-induced fundamental (o2_ver_1 :7187-7255, CONTNM :582-615), the
 *     1.27 um band (O2INF1 :7327-7488, CONTNM :633-660), the 1.06 um band (O2INF2
 *     :7494-7560, CONTNM :670-690), the A-band (O2INF3 :7573-7682, CONTNM :700-719, XINT
 *     bounded by pre_xint), the visible (O2_vis :7699-8090, CONTNM :734-758), the
 *     Herzberg continuum (O2HERZ :8095-8145 with HERTDA and HERPRS, CONTNM :762-779) and
 *     the far UV (O2FUV :8257-8650, CONTNM :784-802, bounded by pre_xint);
 *   N2: the collision-induced pure rotation (xn2_r :2533-2600, CONTNM :831-865), the
 *     fundamental (n2_ver_1 :2701-2790, CONTNM :889-931) and the first overtone
 *     (n2_overtone1 :2956-3056, CONTNM :946-987);
 *   Rayleigh extinction (:970-1060, no aerosols).
 * The tabulated bands go through XINT.  Parity unpinned: HERTDA's correction term (at
 * most 7.9e-27, below double precision against the continuum it corrects) and O2FUV's
 * 1e-5 truncation offset (the same window as 0.01 for integer V1ABS) cannot change a
 * result, so no check reaches them.  wk[m]: column of HITRAN molecule m (wk[1] H2O,
 * wk[2] CO2, wk[3] O3, wk[7] O2); scale: XCO2C, XO3CN, XO2CN, XN2CN, XRAYL.  Returns 0,
 * or -1 when ABSRB is longer than the reference's N_ABSRB (5050 points).
 * --------------------------------------------------------------------------- */

/* the window a table routine of CONTNM takes around [V1ABS, V2ABS]: returns NPTC.  eps
 * is the offset of the index truncation (0.01; O2FUV 1e-5); O2HERZ does not clip NPTC
 * at its table length (clip 0). */
static int cn_window_g(double V1ABS, double V2ABS, double V1S, double DVS, int NPTS, double eps, int clip,
                       double *V1C, double *V2C, double *DVC, int *I1) {
    *DVC = DVS;
    double v1c = V1ABS - *DVC, v2c = V2ABS + *DVC;
    int i1 = v1c < V1S ? -1 : (int)((v1c - V1S) / DVS + eps);
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
    return cn_window_g(V1ABS, V2ABS, grid[0], grid[2], (int)grid[3], 0.01, 1, V1C, V2C, DVC, I1);
}

/* pre_xint (clblm_Spectrum.f90:917-944): the ABSRB points [ist, last] that XINT may
 * touch for a table on [v1ss, v2ss] */
static void cn_pre_xint(double v1ss, double v2ss, double v1abs, double dvabs, int nptabs, int *ist, int *last) {
    const int nbnd_v1c = (int)(2. + (v1ss - v1abs) / dvabs + 1.e-5);
    *ist = nbnd_v1c > 1 ? nbnd_v1c : 1;
    const int nbnd_v2c = (int)(1. + (v2ss - v1abs) / dvabs + 1.e-5);
    *last = nptabs < nbnd_v2c ? nptabs : nbnd_v2c;
}

/* HERTDA and HERPRS (clblm_ODLAY_Continuum.f90:8164-8253): the Herzberg O2 continuum at
 * v and its pressure dependence */
static double cn_hertda(double V) {
    if (V <= 36000.00) return 0.;
    double CORR = 0.;
    if (V <= 40000.) CORR = ((40000. - V) / 4000.) * 7.917E-27;
    const double YRATIO = V / 48811.0;
    const double ly = log(YRATIO);
    return 6.884E-04 * (YRATIO)*exp(-69.738 * (ly * ly)) - CORR;
}

static double cn_herprs(double HERZ, double T, double P) {
    const double PO = 1013., TO = 273.16;
    return HERZ * (1. + .83 * (P / PO) * (TO / T));
}

int orc_contnm_ir(double pave, double tave, double wtot, const double *wk, double v1, double v2, int jrad,
                  const double *scale, double v1abs, int nptabs, double *absrb) {
    const double P0C = 1013., T0C = 296., XLOSMT = 2.68675E+19;
    const double xco2c = scale[0], xo3cn = scale[1], xo2cn = scale[2], xn2cn = scale[3], xrayl = scale[4];
    const double V1ABS = v1abs, V2ABS = v1abs + (double)(nptabs - 1), DVABS = 1.;
    const double RHOAVE = (pave / P0C) * (T0C / tave);
    const double XKT = tave / RADCN2;
    const double amagat = (pave / P0C) * (273. / tave);
    const double x_vmr_h2o = wk[1] / wtot, x_vmr_o2 = wk[7] / wtot;
    const double x_vmr_n2 = 1. - x_vmr_h2o - x_vmr_o2;
    const double wn2 = x_vmr_n2 * wtot;
    double *R = absrb - 1;   /* 1-based */
    double C[6000], V1C, V2C, DVC;
    int I1, NPTC;
    if (nptabs > 5050) return -1;

    /* CO2 */
    if ((v2 > -20.0) && (v1 < 10000.) && xco2c > 0) {
        const double WCO2 = wk[2] * RHOAVE * 1.0E-20 * xco2c;
        NPTC = cn_window(V1ABS, V2ABS, CN_CO2_GRID, &V1C, &V2C, &DVC, &I1);
        const int NPTS = (int)CN_CO2_GRID[3];
        const double trat = tave / CN_CO2_TEFF[0];
        for (int J = 1; J <= NPTC; J++) {
            const int I = I1 + (J - 1);
            double f = 0.;
            if (I >= 1 && I <= NPTS) {
                double tcor = 1.;
                if (I >= 1196 && I <= 1220) tcor = pow(trat, CN_CO2_TDEP[I - 1196]);
                f = tcor * CN_CO2_S[I - 1];
            }
            const double VJ = V1C + DVC * (double)(J - 1);
            double CFAC = 1.;
            if (VJ >= 2000. && VJ <= 2998.) CFAC = CN_XFACCO2[(int)((VJ - 1998.) / 2. + 0.00001) - 1];
            f = CFAC * f;
            C[J] = f * WCO2;
            if (jrad == 1) C[J] = C[J] * orc_radfn(VJ, XKT);
        }
        C[NPTC + 1] = 0.;
        orc_xint(V1C, V2C, DVC, C, 1.0, V1ABS, DVABS, R, 1, nptabs);
    }
    /* O3: Chappuis and Wulf bands */
    if (v2 > 8920.0 && v1 <= 24665.0 && xo3cn > 0.) {
        const double WO3 = wk[3] * 1.0E-20 * xo3cn;
        NPTC = cn_window(V1ABS, V2ABS, CN_O3C_GRID, &V1C, &V2C, &DVC, &I1);
        const int NPTS = (int)CN_O3C_GRID[3];
        const double DT = tave - 273.15;
        for (int J = 1; J <= NPTC; J++) {
            const int I = I1 + (J - 1);
            const double VJ = V1C + DVC * (double)(J - 1);
            double c0 = 0., c1 = 0., c2 = 0.;
            if (I >= 1 && I <= NPTS) {
                c0 = CN_O3C_X[I - 1] / VJ;
                c1 = CN_O3C_Y[I - 1] / VJ;
                c2 = CN_O3C_Z[I - 1] / VJ;
            }
            C[J] = (c0 + (c1 + c2 * DT) * DT) * WO3;
            if (jrad == 1) C[J] = C[J] * orc_radfn(VJ, XKT);
        }
        C[NPTC + 1] = 0.;
        orc_xint(V1C, V2C, DVC, C, 1.0, V1ABS, DVABS, R, 1, nptabs);
    }
    /* O3: Hartley-Huggins bands, 27370-40800 cm-1, with their temperature terms; the
     * points from 40800 cm-1 keep their value when the window reaches past 40815 */
    if (v2 > 27370. && v1 < 40800. && xo3cn > 0.) {
        const double WO3 = wk[3] * 1.E-20 * xo3cn;
        const double TC = tave - 273.15;
        NPTC = cn_window(V1ABS, V2ABS, CN_O3T_GRID, &V1C, &V2C, &DVC, &I1);
        const int NPTS = (int)CN_O3T_GRID[3];
        double VJ = 0.;
        for (int J = 1; J <= NPTC; J++) {
            const int I = I1 + (J - 1);
            const int in = I >= 1 && I <= NPTS;
            VJ = V1C + DVC * (double)(J - 1);
            const double c0 = in ? CN_O3T0_S[I - 1] / VJ : 0.;
            const double ct1 = in ? CN_O3T1_S[I - 1] : 0., ct2 = in ? CN_O3T2_S[I - 1] : 0.;
            C[J] = c0 * WO3;
            if (jrad == 1) C[J] = C[J] * orc_radfn(VJ, XKT);
            C[J] = C[J] * (1. + ct1 * TC + ct2 * TC * TC);
        }
        C[NPTC + 1] = 0.;
        const int fix = (VJ > 40815.) && (v2 > 40800);
        int I_FIX = 0;
        double *save = NULL;
        if (fix) {
            I_FIX = (int)((40800. - V1ABS) / DVABS + 1.001);
            save = (double *)malloc(sizeof(double) * (size_t)(nptabs + 1));
            for (int i = I_FIX; i <= nptabs; i++) save[i] = R[i];
        }
        orc_xint(V1C, V2C, DVC, C, 1.0, V1ABS, DVABS, R, 1, nptabs);
        if (fix) {
            for (int i = I_FIX; i <= nptabs; i++) R[i] = save[i];
            free(save);
        }
    }
    /* O3: Hartley-Huggins in the UV, 40800-54000 cm-1; the points below 40800 cm-1 keep
     * their value */
    if (v2 > 40800. && v1 < 54000. && xo3cn > 0.) {
        const double WO3 = wk[3] * xo3cn;
        NPTC = cn_window(V1ABS, V2ABS, CN_O3UV_GRID, &V1C, &V2C, &DVC, &I1);
        const int NPTS = (int)CN_O3UV_GRID[3];
        for (int J = 1; J <= NPTC; J++) {
            const int I = I1 + (J - 1);
            const double VJ = V1C + DVC * (double)(J - 1);
            const double c0 = (I >= 1 && I <= NPTS) ? CN_O3UV_S[I - 1] / VJ : 0.;
            C[J] = c0 * WO3;
            if (jrad == 1) C[J] = C[J] * orc_radfn(VJ, XKT);
        }
        C[NPTC + 1] = 0.;
        int I_FIX = 0;
        double *save = NULL;
        if (v1 < 40800) {
            I_FIX = (int)((40800. - V1ABS) / DVABS + 1.001);
            save = (double *)malloc(sizeof(double) * (size_t)(nptabs + 1));
            for (int i = 1; i <= I_FIX - 1; i++) save[i] = R[i];
        }
        orc_xint(V1C, V2C, DVC, C, 1.0, V1ABS, DVABS, R, 1, nptabs);
        if (v1 < 40800) {
            for (int i = 1; i <= I_FIX - 1; i++) R[i] = save[i];
            free(save);
        }
    }
    /* O2 collision-induced fundamental */
    if ((v2 > 1340.0) && (v1 < 1850.) && xo2cn > 0.) {
        const double tau_fac = xo2cn * wk[7] * 1.e-20 * amagat;
        NPTC = cn_window(V1ABS, V2ABS, CN_O2F_GRID, &V1C, &V2C, &DVC, &I1);
        const int NPTS = (int)CN_O2F_GRID[3];
        const double xktfac = (1. / 296.) - (1. / tave);
        const double factor = (1.e+20 / 2.68675e+19);
        for (int J = 1; J <= NPTC; J++) {
            const int I = I1 + (J - 1);
            const double VJ = V1C + DVC * (double)(J - 1);
            double c0 = 0.;
            if (I >= 1 && I <= NPTS) c0 = factor * CN_O2F_X[I - 1] * exp(CN_O2F_XT[I - 1] * xktfac) / VJ;
            C[J] = tau_fac * c0;
            if (jrad == 1) C[J] = C[J] * orc_radfn(VJ, XKT);
        }
        C[NPTC + 1] = 0.;
        orc_xint(V1C, V2C, DVC, C, 1.0, V1ABS, DVABS, R, 1, nptabs);
    }
    /* O2: the 1.27 um band (Mate et al.) */
    if ((v2 > 7536.0) && (v1 < 8500.) && xo2cn > 0.) {
        const double a_o2 = 1. / 0.446, a_n2 = 0.3 / 0.446, a_h2o = 1.;
        const double tau_fac =
            xo2cn * (wk[7] / XLOSMT) * amagat * (a_o2 * x_vmr_o2 + a_n2 * x_vmr_n2 + a_h2o * x_vmr_h2o);
        NPTC = cn_window(V1ABS, V2ABS, CN_O2I1_GRID, &V1C, &V2C, &DVC, &I1);
        const int NPTS = (int)CN_O2I1_GRID[3];
        for (int J = 1; J <= NPTC; J++) {
            const int I = I1 + (J - 1);
            const double VJ = V1C + DVC * (double)(J - 1);
            const double c0 = (I >= 1 && I <= NPTS) ? CN_O2I1_X[I - 1] / VJ : 0.;
            C[J] = tau_fac * c0;
            if (jrad == 1) C[J] = C[J] * orc_radfn(VJ, XKT);
        }
        C[NPTC + 1] = 0.;
        orc_xint(V1C, V2C, DVC, C, 1.0, V1ABS, DVABS, R, 1, nptabs);
    }
    /* O2: the 1.06 um band (Mlawer et al.), two damped Lorentzians on a 2 cm-1 grid */
    if ((v2 > 9100.0) && (v1 < 11000.) && xo2cn > 0.) {
        const double V1_osc = 9375., HW1 = 58.96, V2_osc = 9439., HW2 = 45.04;
        const double S1 = 1.166E-04, S2 = 3.086E-05;
        const double V1S = 9100., v2s = 11000., DVS = 2.;
        DVC = DVS;
        V1C = V1ABS - DVC;
        V2C = V2ABS + DVC;
        if (V1C < V1S) V1C = V1S - 2. * DVS;
        if (V2C > v2s) V2C = v2s + 2. * DVS;
        NPTC = (int)((V2C - V1C) / DVC + 3.01);
        V2C = V1C + DVC * (double)(NPTC - 1);
        const double WO2 = xo2cn * (wk[7] * 1.e-20) * RHOAVE;
        const double ADJWO2 = (wk[7] / wtot) * (1. / 0.209) * WO2;
        for (int J = 1; J <= NPTC; J++) {
            double c0 = 0.;
            const double VJ = V1C + DVC * (double)(J - 1);
            if ((VJ > V1S) && (VJ < v2s)) {
                const double DV1 = VJ - V1_osc, DV2 = VJ - V2_osc;
                const double DAMP1 = DV1 < 0.0 ? exp(DV1 / 176.1) : 1.0;
                const double DAMP2 = DV2 < 0.0 ? exp(DV2 / 176.1) : 1.0;
                const double q1 = DV1 / HW1, q2 = DV2 / HW2;
                const double O2INF =
                    0.31831 * (((S1 * DAMP1 / HW1) / (1. + q1 * q1)) + ((S2 * DAMP2 / HW2) / (1. + q2 * q2))) * 1.054;
                c0 = O2INF / VJ;
            }
            C[J] = c0 * ADJWO2;
            if (jrad == 1) C[J] = C[J] * orc_radfn(VJ, XKT);
        }
        C[NPTC + 1] = 0.;
        orc_xint(V1C, V2C, DVC, C, 1.0, V1ABS, DVABS, R, 1, nptabs);
    }
    /* O2: the A-band (Mlawer and Gombos), XINT bounded by pre_xint */
    if ((v2 > 12961.5) && (v1 < 13221.5) && xo2cn > 0.) {
        const double tau_fac = xo2cn * (wk[7] / XLOSMT) * amagat;
        NPTC = cn_window(V1ABS, V2ABS, CN_O2I3_GRID, &V1C, &V2C, &DVC, &I1);
        const int NPTS = (int)CN_O2I3_GRID[3];
        for (int J = 1; J <= NPTC; J++) {
            const int I = I1 + (J - 1);
            const double VJ = V1C + DVC * (double)(J - 1);
            const double c0 = (I >= 1 && I <= NPTS) ? CN_O2I3_X[I - 1] / VJ : 0.;
            C[J] = tau_fac * c0;
            if (jrad == 1) C[J] = C[J] * orc_radfn(VJ, XKT);
        }
        C[NPTC + 1] = 0.;
        int ist, last;
        cn_pre_xint(CN_O2I3_GRID[0], CN_O2I3_GRID[1], V1ABS, DVABS, nptabs, &ist, &last);
        orc_xint(V1C, V2C, DVC, C, 1.0, V1ABS, DVABS, R, ist, last);
    }
    /* O2: the visible bands (Greenblatt et al.) */
    if ((v2 > 15000.0) && (v1 < 29870.) && xo2cn > 0.) {
        const double WO2 = wk[7] * 1.e-20 * ((pave / 1013.) * (273. / tave)) * xo2cn;
        const double CHIO2 = wk[7] / wtot;
        const double ADJWO2 = CHIO2 * WO2;
        const double t = 55. * 273. / 296.;
        const double factor = 1. / ((XLOSMT * 1.e-20 * (t * t)) * 89.5);
        NPTC = cn_window(V1ABS, V2ABS, CN_O2V_GRID, &V1C, &V2C, &DVC, &I1);
        const int NPTS = (int)CN_O2V_GRID[3];
        for (int J = 1; J <= NPTC; J++) {
            const int I = I1 + (J - 1);
            const double VJ = V1C + DVC * (double)(J - 1);
            const double c0 = (I >= 1 && I <= NPTS) ? factor * CN_O2V_S[I - 1] / VJ : 0.;
            C[J] = c0 * ADJWO2;
            if (jrad == 1) C[J] = C[J] * orc_radfn(VJ, XKT);
        }
        C[NPTC + 1] = 0.;
        orc_xint(V1C, V2C, DVC, C, 1.0, V1ABS, DVABS, R, 1, nptabs);
    }
    /* O2: the Herzberg continuum from 36000 cm-1 */
    if (v2 > 36000.0 && xo2cn > 0.) {
        const double WO2 = wk[7] * 1.e-20 * xo2cn;
        NPTC = cn_window_g(V1ABS, V2ABS, 36000., 10., 0, 0.01, 0, &V1C, &V2C, &DVC, &I1);
        if (NPTC + 2 > 6000) return -1;
        for (int J = 1; J <= NPTC; J++) {
            const int I = I1 + (J - 1);
            const double VJ = V1C + DVC * (double)(J - 1);
            double c0 = 0.;
            if (I >= 1) c0 = cn_herprs(cn_hertda(VJ), tave, pave) / VJ;
            C[J] = c0 * WO2;
            if (jrad == 1) C[J] = C[J] * orc_radfn(VJ, XKT);
        }
        C[NPTC + 1] = 0.;
        orc_xint(V1C, V2C, DVC, C, 1.0, V1ABS, DVABS, R, 1, nptabs);
    }
    /* O2: the far UV (Schumann-Runge continuum) from 56740 cm-1, bounded by pre_xint */
    if (v2 > 56740.0 && xo2cn > 0.) {
        const double WO2 = wk[7] * 1.e-20 * xo2cn;
        NPTC = cn_window_g(V1ABS, V2ABS, CN_O2U_GRID[0], CN_O2U_GRID[2], (int)CN_O2U_GRID[3], 1.e-5, 1, &V1C, &V2C,
                           &DVC, &I1);
        const int NPTS = (int)CN_O2U_GRID[3];
        for (int J = 1; J <= NPTC; J++) {
            const int I = I1 + (J - 1);
            const double VJ = V1C + DVC * (double)(J - 1);
            const double c0 = (I >= 1 && I <= NPTS) ? CN_O2U_S[I - 1] / VJ : 0.;
            C[J] = c0 * WO2;
            if (jrad == 1) C[J] = C[J] * orc_radfn(VJ, XKT);
        }
        C[NPTC + 1] = 0.;
        int ist, last;
        cn_pre_xint(CN_O2U_GRID[0], CN_O2U_GRID[1], V1ABS, DVABS, nptabs, &ist, &last);
        orc_xint(V1C, V2C, DVC, C, 1.0, V1ABS, DVABS, R, ist, last);
    }
    /* N2 collision-induced pure rotation */
    if ((v2 > -10.0) && (v1 < 350.) && xn2cn > 0.) {
        const double a_h2o = 1.;
        const double tau_fac = xn2cn * (wn2 / XLOSMT) * amagat;
        NPTC = cn_window(V1ABS, V2ABS, CN_N2R_GRID, &V1C, &V2C, &DVC, &I1);
        const int NPTS = (int)CN_N2R_GRID[3];
        const double tfac = (tave - 296.) / (220. - 296.);
        for (int J = 1; J <= NPTC; J++) {
            const int I = I1 + (J - 1);
            const double VJ = V1C + DVC * (double)(J - 1);
            double c0 = 0., c1 = 0.;
            if (I >= 1 && I <= NPTS) {
                c0 = CN_N2R_C296[I - 1] * pow(CN_N2R_C220[I - 1] / CN_N2R_C296[I - 1], tfac);
                const double sf_T = CN_N2R_SF296[I - 1] * pow(CN_N2R_SF220[I - 1] / CN_N2R_SF296[I - 1], tfac);
                c1 = (sf_T - 1.) * (0.79) / (0.21);
            }
            C[J] = tau_fac * c0 * (x_vmr_n2 + c1 * x_vmr_o2 + a_h2o * x_vmr_h2o);
            if (jrad == 1) C[J] = C[J] * orc_radfn(VJ, XKT);
        }
        C[NPTC + 1] = 0.;
        orc_xint(V1C, V2C, DVC, C, 1.0, V1ABS, DVABS, R, 1, nptabs);
    }
    /* N2 collision-induced fundamental */
    if ((v2 > 2001.77) && (v1 < 2897.59) && xn2cn > 0.) {
        const double tau_fac = xn2cn * (wn2 / XLOSMT) * amagat;
        NPTC = cn_window(V1ABS, V2ABS, CN_N2F_GRID, &V1C, &V2C, &DVC, &I1);
        const int NPTS = (int)CN_N2F_GRID[3];
        const double T = tave;
        const double xtfac = ((1. / T) - (1. / 272.)) / ((1. / 228.) - (1. / 272.));
        const double xt_lin = (T - 272.) / (228. - 272.);
        const double a_o2 = 1.294 - 0.4545 * T / 296.;
        for (int J = 1; J <= NPTC; J++) {
            const int I = I1 + (J - 1);
            double c0 = 0., c1 = 0., c2 = 0.;
            if (I >= 1 && I <= NPTS) {
                const double VJ = V1C + DVC * (double)(J - 1);
                const double x272 = CN_N2F_X272[I - 1], x228 = CN_N2F_X228[I - 1];
                if (x272 > 0. && x228 > 0.) c0 = x272 * pow(x228 / x272, xtfac);
                else c0 = x272 + (x228 - x272) * xt_lin;
                c0 = c0 / VJ;
                c1 = a_o2 * c0;
                c2 = (9. / 7.) * CN_N2F_AH2O[I - 1] * c0;
            }
            const double VJ = V1C + DVC * (double)(J - 1);
            C[J] = tau_fac * (x_vmr_n2 * c0 + x_vmr_o2 * c1 + x_vmr_h2o * c2);
            if (jrad == 1) C[J] = C[J] * orc_radfn(VJ, XKT);
        }
        C[NPTC + 1] = 0.;
        orc_xint(V1C, V2C, DVC, C, 1.0, V1ABS, DVABS, R, 1, nptabs);
    }
    /* N2 first overtone (Shapiro and Gush, modified by Mlawer and Gombos) */
    if ((v2 > 4340.0) && (v1 < 4910.) && xn2cn > 0.) {
        const double a_o2 = 1., a_h2o = 1.;
        const double tau_fac = xn2cn * (wn2 / XLOSMT) * amagat * (x_vmr_n2 + a_o2 * x_vmr_o2 + a_h2o * x_vmr_h2o);
        NPTC = cn_window(V1ABS, V2ABS, CN_N2O_GRID, &V1C, &V2C, &DVC, &I1);
        const int NPTS = (int)CN_N2O_GRID[3];
        for (int J = 1; J <= NPTC; J++) {
            const int I = I1 + (J - 1);
            const double VJ = V1C + DVC * (double)(J - 1);
            const double c0 = (I >= 1 && I <= NPTS) ? CN_N2O_X[I - 1] / VJ : 0.;
            C[J] = tau_fac * c0;
            if (jrad == 1) C[J] = C[J] * orc_radfn(VJ, XKT);
        }
        C[NPTC + 1] = 0.;
        orc_xint(V1C, V2C, DVC, C, 1.0, V1ABS, DVABS, R, 1, nptabs);
    }
    /* Rayleigh extinction (IAERSL 0) */
    if (v2 >= 820. && xrayl > 0.) {
        const double conv_cm2mol = xrayl * 1.E-20 / (2.68675e-1 * 1.e5);
        for (int i = 1; i <= nptabs; i++) {
            const double vrayleigh = V1ABS + (double)(i - 1) * DVABS;
            const double x = vrayleigh / 1.e4;
            double ray_ext = (x * x * x / (9.38076E2 - 10.8426 * (x * x))) * (wtot * conv_cm2mol);
            if (jrad == 1) ray_ext = ray_ext * x;
            else if (jrad == 0) ray_ext = ray_ext / 1.e4;
            R[i] = R[i] + ray_ext;
        }
    }
    return 0;
}

/* ---------------------------------------------------------------------------
 * mode1_lookDn_mergeDn / layerMerge_mode1 (clblm_noScattRT.f90:42-528, 529-842) with
 * addThermalBoundary and addSolarBoundary (:1945-2148): the radiance at an observer
 * above the layers, merged from the top layer down, every layer on its own grid
 * (od_* + off[l], nlim[l] points on v1 + i dv[l]).  Layer l is below the observer when
 * l < obs_level (levels from 0 at the surface).  The view, down and solar path ODs of
 * layer l are od_up, od_dn, od_sun times fac_up[l], fac_dn[l], fac_sun[l] (the
 * airmass-scaled form; factor 1 for ODs computed along each path).  The merged
 * transmittance, up and down radiance and solar transmittance start on the two-point
 * grid V1, V2 (:217-245), and every interpToFinestGrid of the reference is applied in
 * its order.  Surface: constant emissivity emis (reflecting when below 1), skin
 * temperature tskin, solar reflectance refl; sol: the solar radiance at the top on
 * v1 + i sol_dv (sol_n points).  rad and tx (capacity cap) receive the radiance and
 * the total transmittance on the final grid; returns its length (*dv_out its DV), or -1.
 * ------------------------------------------------------------------------- */
typedef struct { double *x; double dv; int64_t n; } m1spec;

static void m1_new(m1spec *s, double dv, int64_t n) {
    s->x = (double *)calloc((size_t)n, sizeof(double));
    s->dv = dv;
    s->n = n;
}

/* interpToFinestGrid on the spectra s[0..ns) in place */
static void m1_finest(double v1, int ns, m1spec **s) {
    int loc = 0;
    for (int k = 1; k < ns; k++)
        if (s[k]->dv < s[loc]->dv) loc = k;
    const double dvl = s[loc]->dv;
    const int64_t nl = s[loc]->n;
    for (int k = 0; k < ns; k++) {
        if (k == loc || same_grid(v1, s[k]->dv, s[k]->n, dvl, nl)) continue;
        double *o = (double *)calloc((size_t)nl, sizeof(double));
        orc_interp(s[k]->x, s[k]->dv, v1, (int)s[k]->n, o, dvl, v1, (int)nl);
        free(s[k]->x);
        s[k]->x = o;
        s[k]->dv = dvl;
        s[k]->n = nl;
    }
}

int64_t orc_rt_mode1(int nlay, double v1, double v2, const double *dv, const int64_t *nlim, const int64_t *off,
                     const double *od_up, const double *od_dn, const double *od_sun, const double *fac_up,
                     const double *fac_dn, const double *fac_sun, const double *t_lay, const double *t_lev,
                     int obs_level, int lin, int thermal, int solar, double emis, double tskin, double refl,
                     const double *sol, int64_t sol_n, double sol_dv, double *rad, double *tx, int64_t cap,
                     double *dv_out) {
    if (!thermal && emis < 1.) return fail("mode1: a reflecting surface needs the thermal down path");
    if (!(emis < 1.) && !(tskin > 0.)) return fail("mode1: no surface term (emissivity 1 and no skin temperature)");
    m1spec TxA, RadUp, RadDn, SolTx;
    m1_new(&TxA, v2 - v1, 2);
    m1_new(&RadUp, v2 - v1, 2);
    m1_new(&RadDn, v2 - v1, 2);
    m1_new(&SolTx, v2 - v1, 2);
    for (int i = 0; i < 2; i++) { TxA.x[i] = 1.; SolTx.x[i] = 1.; }
    for (int l = nlay - 1; l >= 0; l--) {
        const int below = l < obs_level;
        const double ldv = dv[l];
        const int64_t ln = nlim[l];
        const double tbot = t_lev[l], ttop = t_lev[l + 1], tave = t_lay[l];
        m1spec lTxA, lTxB, lEmDn, lEmUp, lSol;
        m1_new(&lTxA, ldv, ln);
        m1_new(&lTxB, ldv, ln);
        m1_new(&lEmDn, ldv, ln);
        m1_new(&lEmUp, ldv, ln);
        m1_new(&lSol, ldv, ln);
        double *oda = (double *)calloc((size_t)ln, sizeof(double));
        for (int64_t i = 0; i < ln; i++) {
            const double v = v1 + (double)i * ldv;
            if (below) {
                oda[i] = od_up[off[l] + i] * fac_up[l];
                lTxA.x[i] = exp(-oda[i]);
            }
            if (thermal) {
                const double odb = od_dn[off[l] + i] * fac_dn[l];
                lTxB.x[i] = exp(-odb);
                lEmDn.x[i] = layer_emis(v, tbot, ttop, tave, lTxB.x[i], odb, lin);
            }
            if (solar) lSol.x[i] = exp(-(od_sun[off[l] + i] * fac_sun[l]));
        }
        if (thermal) {
            m1spec *a[3] = {&lEmDn, &lTxB, &RadDn};
            m1_finest(v1, 3, a);
            for (int64_t i = 0; i < RadDn.n; i++) RadDn.x[i] = lEmDn.x[i] + RadDn.x[i] * lTxB.x[i];
            if (below) {
                if (lin != 0) {
                    for (int64_t i = 0; i < ln; i++)
                        lEmUp.x[i] = layer_emis(v1 + (double)i * ldv, ttop, tbot, tave, lTxA.x[i], oda[i], lin);
                } else {            /* lyrEmUp = lyrEmDn, on its grid after the merge above */
                    free(lEmUp.x);
                    m1_new(&lEmUp, lEmDn.dv, lEmDn.n);
                    memcpy(lEmUp.x, lEmDn.x, sizeof(double) * (size_t)lEmDn.n);
                }
                m1spec *b[3] = {&lEmUp, &TxA, &RadUp};
                m1_finest(v1, 3, b);
                for (int64_t i = 0; i < RadUp.n; i++) RadUp.x[i] = RadUp.x[i] + TxA.x[i] * lEmUp.x[i];
            }
        }
        if (below) {
            m1spec *c[2] = {&lTxA, &TxA};
            m1_finest(v1, 2, c);
            for (int64_t i = 0; i < TxA.n; i++) TxA.x[i] = TxA.x[i] * lTxA.x[i];
        }
        if (solar) {
            m1spec *d[2] = {&lSol, &SolTx};
            m1_finest(v1, 2, d);
            for (int64_t i = 0; i < SolTx.n; i++) SolTx.x[i] = SolTx.x[i] * lSol.x[i];
        }
        free(oda); free(lTxA.x); free(lTxB.x); free(lEmDn.x); free(lEmUp.x); free(lSol.x);
    }
    /* addThermalBoundary (flux off) */
    m1spec Tot;
    if (emis < 1.) {
        m1spec *a[3] = {&TxA, &RadUp, &RadDn};
        m1_finest(v1, 3, a);
        m1_new(&Tot, RadUp.dv, RadUp.n);
        for (int64_t i = 0; i < Tot.n; i++) {
            const double v = v1 + (double)i * RadUp.dv;
            const double srad = planck_x(v, tskin / RADCN2) * emis, srefl = 1.0 - emis;
            Tot.x[i] = RadUp.x[i] + TxA.x[i] * (srad + srefl * RadDn.x[i]);
        }
    } else {
        m1spec *a[2] = {&RadUp, &TxA};
        m1_finest(v1, 2, a);
        m1_new(&Tot, RadUp.dv, RadUp.n);
        for (int64_t i = 0; i < Tot.n; i++) {
            const double v = v1 + (double)i * RadUp.dv;
            Tot.x[i] = RadUp.x[i] + TxA.x[i] * (planck_x(v, tskin / RADCN2) * emis);
        }
    }
    /* addSolarBoundary */
    if (solar) {
        m1spec S0;
        m1_new(&S0, sol_dv, sol_n);
        memcpy(S0.x, sol, sizeof(double) * (size_t)sol_n);
        m1spec *a[4] = {&TxA, &Tot, &SolTx, &S0};
        m1_finest(v1, 4, a);
        for (int64_t i = 0; i < TxA.n; i++) Tot.x[i] = Tot.x[i] + TxA.x[i] * refl * SolTx.x[i] * S0.x[i];
        free(S0.x);
    }
    const int64_t n = Tot.n;
    for (int64_t i = 0; i < n && i < cap; i++) {
        rad[i] = Tot.x[i];
        if (tx) tx[i] = TxA.n == n ? TxA.x[i] : 0.;
    }
    *dv_out = Tot.dv;
    free(TxA.x); free(RadUp.x); free(RadDn.x); free(SolTx.x); free(Tot.x);
    return n;
}

/* ---------------------------------------------------------------------------
 * clblm_FLTRFN with CNVFLT (clblm_ScanFilter.f90:1703-1900), the spectral response
 * function filter of postProc_filter_singleStruct (clblm_PostProc.f90:289-336): the
 * spectrum x[0..n) on v1 + i dv, read in 2400-point panels (panelInput :2320-2404,
 * VBOT = V1F, VTOP = V2F), summed against the filter xf[0..nptf) on v1f + i dvf (its end
 * values halved, each spectral point weighted by the filter interpolated linearly at
 * the point's interval), times DV, over the filter's own sum times DVF.  dvf < 0:
 * v1f is the filter's centre (odd nptf).  Returns 0 (result in *out), -1 on error.
 * ------------------------------------------------------------------------- */
int orc_fltrfn(const double *x, int64_t n, double v1, double dv, double v1f, double dvf, int nptf, const double *xf,
               double *out) {
    if (nptf < 2 || nptf > 3001) return fail("filter: 2..3001 filter points (NFLTPT)");
    double DVF = dvf, V1F = v1f;
    *out = 0.;
    if (V1F < 0) return 0;   /* the reference returns without setting filtOut */
    if (DVF < 0.) {
        DVF = fabs(DVF);
        if ((nptf - 1) % 2 != 0) return fail("filter: a centre frequency needs an odd number of points");
        const int half = (nptf - 1) / 2;
        V1F = V1F - DVF * (double)half;
    }
    double *XF = (double *)malloc(sizeof(double) * (size_t)(nptf + 1));   /* 1-based */
    for (int i = 0; i < nptf; i++) XF[i + 1] = xf[i];
    const int NPTS = nptf;
    const double V2F = V1F + DVF * (double)(NPTS - 1);
    XF[1] = 0.5 * XF[1];
    XF[NPTS] = 0.5 * XF[NPTS];
    double SUMFLT = 0.0;
    for (int i = 1; i <= NPTS; i++) SUMFLT = SUMFLT + XF[i];
    SUMFLT = SUMFLT * DVF;
    double RFILTR = 0.0;
    const double VBOT = V1F, VTOP = V2F;
    int64_t done = 0;
    double done_v2 = v1 - dv;
    for (;;) {
        /* panelInput, NREN = 0 */
        int nlim;
        double VMIN = 0., VMAX = 0.;
        for (;;) {
            nlim = 2400;
            if (done + nlim > n) nlim = (int)(n - done);
            if (nlim == 0) break;
            VMIN = done_v2 + dv;
            VMAX = VMIN + (double)(nlim - 1) * dv;
            done += nlim;
            done_v2 = VMAX;
            if (VMAX >= VBOT) break;
        }
        if (nlim == 0) break;                  /* IEOFSC = 0 */
        const double *S = x + (done - nlim) - 1;   /* S[1..nlim] */
        int ILO = 1, IHI = nlim, IDATA = 0;
        if (VMAX > VTOP) {
            IHI = (int)((VTOP - VMIN) / dv + 1.5);
            IDATA = 1;
        }
        /* CNVFLT */
        const double DVI = dv, V1I = VMIN;
        int IMIN = (int)((V1F - V1I) / DVI + 1.0001);
        if (IMIN < ILO) IMIN = ILO;
        int IMAX = (int)((V2F + V1F - V1I) / DVI + 1.0001);
        if (IMAX > IHI) IMAX = IHI;
        const double XIF0 = (V1I - V1F) / DVF + 1.0001;
        const double XDVIF = DVI / DVF;
        if ((int)(XIF0 + XDVIF * (double)IMAX) + 1 > NPTS) {
            for (int i = IMAX - 1; i >= IMIN; i--)
                if ((int)(XIF0 + XDVIF * (double)i) + 1 <= NPTS) {
                    IMAX = i;
                    break;
                }
        }
        if ((int)(XIF0 + XDVIF * (double)IMAX) + 1 > NPTS) {
            free(XF);
            return fail("filter: subscript beyond the filter (CNVFLT)");
        }
        for (int I = IMIN; I <= IMAX; I++) {
            const int IFL = (int)(XIF0 + XDVIF * (double)I);
            const double v1s = V1I + (double)(I - 1) * DVI;
            const double v2s = V1I + (double)I * DVI;
            const double vxf1 = V1F + (double)(IFL - 1) * DVF;
            const double vxf2 = V1F + (double)IFL * DVF;
            const double p = (vxf2 - v2s) / (vxf2 - vxf1);
            (void)v1s;
            RFILTR = RFILTR + S[I] * (p * XF[IFL] + (1. - p) * XF[IFL + 1]);
        }
        if (IDATA == 1) break;
    }
    *out = (RFILTR * dv) / SUMFLT;
    free(XF);
    return 0;
}

/* Layer OD with its directional derivatives (the adjoint leg's checker). */
#include "clblm_oracle_tan.c"
