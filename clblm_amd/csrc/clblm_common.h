// Shared definitions of the MI355X layer-OD library (host and device).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace clblm {

// clblm_ConstParam.f90:49-50,59-71,110,112
constexpr double ONEPL = 1.01;
constexpr double ONEMI = 0.999;
constexpr double RADCN1 = 1.191042722E-12;
constexpr double RADCN2 = 1.4387752;
constexpr double BOLTZ = 1.3806503E-16;
constexpr double CLIGHT = 2.99792458E+10;
constexpr double AVOGAD = 6.02214199E+23;
constexpr double PI_REF = 3.1415926535898;
constexpr double P0 = 1013.25;
constexpr double TEMP0 = 296.0;
constexpr int NFPTS = 2001;
constexpr int NFMX = 3001;            // NFMX = 1.5*NFPTS truncated
constexpr int NMOL = 7;               // molecules with TIPS/mass data here
constexpr int NISO = 11;              // isotopologue slots (1-based, <= 10)
constexpr int NSHIFT = 32;
constexpr int NLIM1 = 2401, NLIM2 = 601, NLIM3 = 151;
constexpr int LIMOUT4 = 2500;         // R4 block length (LBLF4Q)
constexpr int LIMIN4 = 1000;          // LINF4Q shrink buffer
constexpr int REC = 250;              // TAPE3 record length
constexpr int MAXCHUNK_REC = 6;       // records a shrink chunk can span (1000/250 + slack)

// Per-layer job descriptor (device resident).
struct Job {
    double v1, v2, dv, dvr2, dvr3, dvr4, alfmax, vft0, bound4;
    double xkt, betacr, tratio, rhorat, pavp0, pavp2;   // LNCORQ constants
    // line coupling at TAVE: coefficient = X[anchor] + ((X[lo + 1] - X[lo]) rect) dif over the
    // nodes X[0..3] at 200/250/296/340 K, with LNCORQ's bracket (lcq) and LINF4Q's (lc4)
    int lcq_lo, lcq_anchor, lc4_lo, lc4_anchor;
    double lcq_rect, lcq_dif, lc4_rect, lc4_dif;
    double beta4, betacr4;                       // LINF4Q constants
    double v1l4, v2l4, tave;
    double w[NMOL + 1], rhoslf[NMOL + 1];
    double scor[NMOL + 1][NISO], alfd1[NMOL + 1][NISO];
    int sel[NMOL + 1];
    int jrad, max1, max2, max3, stride, nalfmax;
    int npanel, nblk;
    int64_t panel_off, blk_off, out_off, npts;
    int64_t line_lo, line_n, lp_off;             // RDLIN range and scratch offset
    int64_t f4_lo, f4_n, f4_off;                 // LINF4Q range and scratch offset
    int64_t rec0, nrec, rec_off;                 // records covering the LINF4Q range
    int64_t chunk_off;                           // scratch for chunk table (nrec entries)
    int contnm, nptabs;
    int species_broad;                           // odCtrl%speciesBroad
    double v1abs;
    int64_t absrb_off;
    int64_t r4_off;                              // start of this layer's R4 blocks
    int64_t f4rec_off, f4rec_cap;                // per-(line, block) F4 records of this layer
    // grid of the returned spectrum (DVout re-grid)
    double dvout;
    int64_t nat_off;                             // native OD offset when re-gridding
    int64_t nnat;
};

struct Panel {
    double vft;
    int nhi, istop, blk, layer;
    int64_t out_off;        // absolute output offset of point NLO
    int64_t a_off;          // offset of this panel's A arrays (doubles or floats)
    int64_t lstart, lend;   // line range (absolute scratch index) homed here
};

struct R4Blk {
    double v1r4, v2r4;
    int nptr4, layer;
    int64_t r4_off;         // offset of R4(1) in the R4 buffer (2502 slots per block)
    int64_t lo, hi;         // shrunk-line range read by RDLN4Q/CNVF4Q (absolute)
    int64_t rec;            // first per-(line, block) F4 record of this block
};

}  // namespace clblm
