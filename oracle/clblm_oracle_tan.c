/*
 * CLBLM oracle -- layer optical depth with its directional derivatives.
 *
 * TEST INFRASTRUCTURE ONLY (it checks results).  The product never links or calls it.
 * This file is compiled as part of clblm_oracle.c (it uses that file's statics).
 *
 * The model differentiated is the reference's ODLAY (clblm_ODLAY.f90:757-2545,
 * clblm_ODLAY_LineF4.f90:25-1647) with one change: CNVFNQ's table lookups F(IZ),
 * IZ = INT(|ZF| + 1.5), are replaced by the closed forms the tables sample
 * (SHAPEL clblm_ODLAY.f90:2804-2909, SHAPEG :2914-2953, VERFN :2959-3016) evaluated
 * at the exact argument x = |ZF| DXF.  Everything else -- the panels, the walk bounds
 * JMIN1/JMAX1, the R2/R3 shifts, the zeta-table interpolation, LINF4Q's buffer, SHRINQ's
 * groups, CNVF4Q's far-wing function, XINT and the Lagrange steps of panelOut, the
 * radiation terms -- is the reference's.  Discrete choices (table rows IZETA, walk
 * bounds, groups, the JRAD flip) are taken from the values and held fixed, so the
 * derivative is that of the piecewise-smooth function the reference evaluates with
 * its tables refined to their closed forms.
 *
 * A direction d = (dT, dlnWtot, dlnW(1..47)) moves the layer temperature by dT, the
 * broadening column by the factor (1 + dlnWtot) and each column W(m) by (1 + dlnW(m)),
 * to first order.  The layer constants of MOLEC and LNCORQ follow: RHORAT = (P/P0)
 * (T0/T), RHOSLF(m) = RHORAT W(m)/Wtot, ALFD1 ~ sqrt(T), SCOR = Q(296)/Q(T) through
 * TIPS's Lagrange interpolation, BETACR, XKT.  The continuum (ABSRB) is the caller's:
 * its derivative along each direction comes in dabsrb.
 *
 * Output: channel 0 is the OD of this smooth model, channel 1 + k its derivative
 * along dirs[k]; each channel holds the layer's points as orc_odlay returns them.
 * Not differentiated (an error is returned): species broadening.
 */

#define TAN_MAXDIR 16

/* derivative of a layer constant along one direction */
typedef struct {
    double dT, dxkt, dbetacr, drhorat, dbeta4, dtmp;   /* dtmp = dT / T */
    double dw[48], drhoslf[48], dscor[48 * 11], dalfd1[48 * 11];
} dlay_t;

/* d/dt of AtoB's Lagrange interpolant (clblm_ODLAY_TIPS.f90:5451-5547): the same
 * three- or four-point stencil, differentiated in aa. */
static double atob_d(double aa, const double *A, const double *B, int npt) {
    for (int i = 2; i <= npt; i++) {
        if (A[i] >= aa) {
            int j = i;
            int three = (i < 3 || i == npt);
            if (i < 3) j = 3;
            if (i == npt) j = npt;
            int lo = j - 2, hi = three ? j : j + 1;
            double s = 0.;
            for (int k = lo; k <= hi; k++) {
                double den = 1.;
                for (int q = lo; q <= hi; q++) {
                    if (q == k) continue;
                    double d = A[k] - A[q];
                    if (d == 0.) d = 0.0001;
                    den = den * d;
                }
                /* d/daa of prod_{q != k} (aa - A[q]) */
                double num = 0.;
                for (int r = lo; r <= hi; r++) {
                    if (r == k) continue;
                    double p = 1.;
                    for (int q = lo; q <= hi; q++)
                        if (q != k && q != r) p = p * (aa - A[q]);
                    num = num + p;
                }
                s = s + num / den * B[k];
            }
            return s;
        }
    }
    return 0.;
}

/* d ln Q / dT of TIPS at t */
static int qt_dlog(int mol, int iso, double t, double *dlq) {
    const double *row = CLBLM_TIPS_Q + (size_t)(CLBLM_TIPS_ROW[mol - 1] + iso - 1) * CLBLM_NT;
    double A[CLBLM_NT + 2], B[CLBLM_NT + 2];
    A[0] = B[0] = A[CLBLM_NT + 1] = B[CLBLM_NT + 1] = 0.;
    for (int k = 0; k < CLBLM_NT; k++) { A[k + 1] = CLBLM_TIPS_TDAT[k]; B[k + 1] = row[k]; }
    double q = atob(t, A, B, CLBLM_NT), dq = atob_d(t, A, B, CLBLM_NT);
    if (!(q > 0.)) return fail("TIPS: partition sum less than 0");
    *dlq = dq / q;
    return 0;
}

double orc_tips_dlogq(int mol, int iso, double t) {
    double d = 0.;
    if (qt_dlog(mol, iso, t, &d)) return 0.;
    return d;
}

/* d/dT of RADFN(v, T / RADCN2) (clblm_ODLAY_CommonSub.f90:293-328), branch by branch */
double orc_dradfn_dt(double v, double t) {
    double xkt = t / RADCN2;
    if (!(xkt > 0.)) return 0.;
    double x = v / xkt;
    if (x <= 0.01) return -0.5 * x * v / t;
    if (x <= 10.0) {
        double e = exp(-x);
        return -2. * v * x * e / ((1. + e) * (1. + e) * t);
    }
    return 0.;
}

static int make_dlay(const orc_layer *L, const molec_t *M, const orc_dir *d, dlay_t *o) {
    memset(o, 0, sizeof *o);
    const double T = L->tave;
    double RHORAT = (L->pave / P0) * (TEMP0 / T);
    o->dT = d->dt;
    o->dtmp = d->dt / T;
    o->dxkt = d->dt / RADCN2;
    o->dbetacr = -RADCN2 / (T * T) * d->dt;
    o->dbeta4 = o->dbetacr;
    o->drhorat = -RHORAT * o->dtmp;
    for (int m = 1; m <= MXMOL; m++) {
        o->dw[m] = L->w[m] * d->dlnw[m];
        if (!L->sel[m]) continue;
        o->drhoslf[m] = M->rhoslf[m] * (-o->dtmp - d->dlnwtot + d->dlnw[m]);
        for (int iso = 1; iso <= CLBLM_MAXISO[m - 1]; iso++) {
            double dlq;
            if (qt_dlog(m, iso, T, &dlq)) return -1;
            o->dscor[m * 11 + iso] = -M->scor[m * 11 + iso] * dlq * d->dt;
            o->dalfd1[m * 11 + iso] = M->alfd1[m * 11 + iso] * 0.5 * o->dtmp;
        }
    }
    return 0;
}

/* d/dT of the coupling coefficients: the slopes of lc_lncorq / lc_linf4q's brackets */
static void lc_lncorq_slope(const orc_lines *Ls, int64_t i, double tave, double *sa, double *sb) {
    double A[5], B[5];
    lc_nodes(Ls, i, A, B);
    int ILC = 1;
    for (int IL = 1; IL <= 3; IL++) {
        ILC = IL;
        if (tave < TEMPLC[ILC + 1 - 1]) break;
    }
    if (ILC == 4) ILC = 3;
    double RECTLC = 1.0 / (TEMPLC[ILC + 1 - 1] - TEMPLC[ILC - 1]);
    *sa = (A[ILC + 1] - A[ILC]) * RECTLC;
    *sb = (B[ILC + 1] - B[ILC]) * RECTLC;
}
static void lc_linf4q_slope(const orc_lines *Ls, int64_t i, double tave, double *sa, double *sb) {
    double A[5], B[5];
    lc_nodes(Ls, i, A, B);
    int ILC;
    for (ILC = 1; ILC <= 4; ILC++)
        if (tave <= TEMPLC[ILC - 1]) break;
    if (ILC == 1) ILC = 2;
    if (ILC == 5) ILC = 4;
    double RECTLC = 1.0 / (TEMPLC[ILC - 1] - TEMPLC[ILC - 1 - 1]);
    *sa = (A[ILC] - A[ILC - 1]) * RECTLC;
    *sb = (B[ILC] - B[ILC - 1]) * RECTLC;
}

/* ---- closed forms of the shape tables and their derivatives in x ---------------- */
typedef struct { double f, d; } fd_t;
typedef struct { double recpi, a1, b1, a2, b2, a3, b3, fgnorm, fln2, se0, ae2, ae4, cn; } shc_t;
static shc_t shape_consts(void) {
    shc_t c;
    c.recpi = 1. / (2. * asin(1.));
    c.a1 = Afn(4.); c.b1 = Bfn(4.); c.a2 = Afn(16.); c.b2 = Bfn(16.); c.a3 = Afn(64.); c.b3 = Bfn(64.);
    c.fln2 = log(2.);
    c.fgnorm = sqrt(c.fln2 * c.recpi);
    const double CEXP = 0.45, CE0 = 1., CE2 = -.20737285249, CE4 = -.00872684335747;
    c.se0 = sqrt(CEXP / PI_REF);
    c.ae2 = 2. * CEXP;
    c.ae4 = c.ae2 * c.ae2 / 3.;
    c.cn = 1. / (CE0 + CE2 + CE4);
    return c;
}
/* F1 (|x| < 4), F2 (< 16), F3 (< 64): Lorentz or inner quadratic minus the outer one */
static fd_t sh_f(const shc_t *c, int k, double x) {
    fd_t r = {0., 0.};
    const double hw_in = k == 1 ? 0. : (k == 2 ? 4. : 16.), hw = k == 1 ? 4. : (k == 2 ? 16. : 64.);
    const double ai = k == 2 ? c->a1 : c->a2, bi = k == 2 ? c->b1 : c->b2;
    const double ao = k == 1 ? c->a1 : (k == 2 ? c->a2 : c->a3), bo = k == 1 ? c->b1 : (k == 2 ? c->b2 : c->b3);
    if (x >= hw) return r;
    double xs = x * x;
    if (k > 1 && x <= hw_in) {
        r.f = c->recpi * ((ai + bi * xs) - (ao + bo * xs));
        r.d = c->recpi * 2. * x * (bi - bo);
    } else {
        double l = 1. / (1. + xs);
        r.f = c->recpi * (l - (ao + bo * xs));
        r.d = c->recpi * (-2. * x * l * l - 2. * bo * x);
    }
    return r;
}
static fd_t sh_fg(const shc_t *c, double x) {
    fd_t r = {0., 0.};
    if (x >= 4.) return r;
    r.f = c->fgnorm * exp(-(c->fln2 * x * x));
    r.d = -2. * c->fln2 * x * r.f;
    return r;
}
static fd_t sh_xver(const shc_t *c, double x) {
    const double CEXP = 0.45, CE0 = 1., CE2 = -.20737285249, CE4 = -.00872684335747;
    double z2 = x * x;
    double xe0 = c->se0 * exp(-(CEXP * z2));
    double poly = c->cn * (CE0 + CE2 * c->ae2 * z2 + CE4 * c->ae4 * z2 * z2);
    double dpoly = c->cn * (2. * CE2 * c->ae2 * x + 4. * CE4 * c->ae4 * z2 * x);
    fd_t r;
    r.f = poly * xe0;
    r.d = dpoly * xe0 + poly * xe0 * (-2. * CEXP * x);
    return r;
}

/* ---- F4: LINF4Q buffer lines and shrunk lines with their derivatives ------------- */
typedef struct { double v, s, al, ad, p; } dsl_t;   /* d(vnu, sp, alfa, epp, spp) */
typedef struct { dsl_t *d; int64_t n, cap; int nd; } shtan_t;

static void sht_push(shtan_t *t, const dsl_t *d) {
    if (t->n == t->cap) {
        t->cap = t->cap ? 2 * t->cap : 4096;
        t->d = realloc(t->d, (size_t)t->cap * (size_t)t->nd * sizeof(dsl_t));
    }
    memcpy(t->d + t->n * t->nd, d, sizeof(dsl_t) * (size_t)t->nd);
    t->n++;
}

/* SHRINQ (clblm_ODLAY_LineF4.f90:1516-1647) on B[ilo..ihi] with the derivatives DB
 * (nd per line): the same groups; a group's line carries the derivatives of its
 * strength-weighted sums. */
static int shrinq_tan(sline *B, dsl_t *DB, int nd, int ilo, int ihi, double dvr4) {
    int j = ilo - 1;
    double dv = dvr4 / 2.;
    double vlmt = B[ilo].vnu + dv;
    double sumal = 0, sumad = 0, sums = 0, sumv = 0, sumc = 0;
    double sumal2 = 0, sumad2 = 0, sums2 = 0, sumv2 = 0, sumc2 = 0;
    dsl_t ga[TAN_MAXDIR], gb[TAN_MAXDIR];   /* d(sumv, sums, sumal, sumad, sumc) */
    memset(ga, 0, sizeof ga);
    memset(gb, 0, sizeof gb);
    for (int i = ilo; i <= ihi; i++) {
        dsl_t *di = DB + (size_t)i * nd;
        if (B[i].sp < 1.0e-35) {
            B[i].sp = 0.0; B[i].spp = 0.0;
            for (int c = 0; c < nd; c++) { di[c].s = 0.; di[c].p = 0.; }
        }
        if (B[i].spp != 0.0) {
            j++;
            B[j] = B[i];
            memmove(DB + (size_t)j * nd, di, sizeof(dsl_t) * (size_t)nd);
            if (B[j].mol != 2) B[j].mol = 0;
        } else {
            int co2 = B[i].mol == 2;
            dsl_t *g = co2 ? gb : ga;
            if (!co2) {
                sumv = sumv + B[i].vnu * B[i].sp; sums = sums + B[i].sp;
                sumal = sumal + B[i].sp * B[i].alfa; sumad = sumad + B[i].sp * B[i].epp; sumc = sumc + B[i].spp;
            } else {
                sumv2 = sumv2 + B[i].vnu * B[i].sp; sums2 = sums2 + B[i].sp;
                sumal2 = sumal2 + B[i].sp * B[i].alfa; sumad2 = sumad2 + B[i].sp * B[i].epp; sumc2 = sumc2 + B[i].spp;
            }
            for (int c = 0; c < nd; c++) {
                g[c].v += di[c].v * B[i].sp + B[i].vnu * di[c].s;
                g[c].s += di[c].s;
                g[c].al += di[c].s * B[i].alfa + B[i].sp * di[c].al;
                g[c].ad += di[c].s * B[i].epp + B[i].sp * di[c].ad;
                g[c].p += di[c].p;
            }
        }
        if (i < ihi && B[i + 1].vnu <= vlmt) continue;
        vlmt = B[i].vnu + dv;
        if (sums > 0.) {
            j++;
            B[j].sp = sums; B[j].alfa = sumal / sums; B[j].epp = sumad / sums;
            B[j].vnu = sumv / sums; B[j].spp = sumc; B[j].mol = 0;
            dsl_t *dj = DB + (size_t)j * nd;
            for (int c = 0; c < nd; c++) {
                dj[c].s = ga[c].s;
                dj[c].al = (ga[c].al - B[j].alfa * ga[c].s) / sums;
                dj[c].ad = (ga[c].ad - B[j].epp * ga[c].s) / sums;
                dj[c].v = (ga[c].v - B[j].vnu * ga[c].s) / sums;
                dj[c].p = ga[c].p;
            }
            sumal = sumad = sums = sumv = sumc = 0.;
        }
        memset(ga, 0, sizeof ga);
        if (sums2 > 0.) {
            j++;
            B[j].sp = sums2; B[j].alfa = sumal2 / sums2; B[j].epp = sumad2 / sums2;
            B[j].vnu = sumv2 / sums2; B[j].mol = 2; B[j].spp = sumc2;
            dsl_t *dj = DB + (size_t)j * nd;
            for (int c = 0; c < nd; c++) {
                dj[c].s = gb[c].s;
                dj[c].al = (gb[c].al - B[j].alfa * gb[c].s) / sums2;
                dj[c].ad = (gb[c].ad - B[j].epp * gb[c].s) / sums2;
                dj[c].v = (gb[c].v - B[j].vnu * gb[c].s) / sums2;
                dj[c].p = gb[c].p;
            }
            sumal2 = sumad2 = sums2 = sumv2 = sumc2 = 0.;
        }
        memset(gb, 0, sizeof gb);
    }
    return j;
}

/* LINF4Q + RDLNFL + SHRINQ (clblm_ODLAY_LineF4.f90:647-1263) with derivatives. */
static int linf4q_tan(const orc_lines *Ls, const orc_layer *L, const molec_t *M, int nd, const dlay_t *dl,
                      double v1l4, double v2l4, double dvr4, shrunk_t *out, shtan_t *tout) {
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
    dsl_t *DB = (dsl_t *)calloc((size_t)1252 * (size_t)(nd > 0 ? nd : 1), sizeof(dsl_t));
    int IJ = 0, ILO = 1, IHI = 0;
    int ieof = 0;
    int64_t nrec = (Ls->n + 249) / 250;
    int64_t rec = 0;
    for (;;) {
        int got = 0;
        while (rec < nrec) {
            int64_t a = rec * 250, b = a + 250 < Ls->n ? a + 250 : Ls->n;
            rec++;
            if (rec_vmax(Ls, a, b) < VLO) continue;
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
                double dsui[TAN_MAXDIR];
                for (int c = 0; c < nd; c++) dsui[c] = (double)Ls->s[i] * dl[c].dw[m];
                double YI = 0., GI = 0., GAMMA1 = 0., GAMMA2 = 0., sa = 0., sb = 0.;
                if (IFLAG == 1) lc_linf4q(Ls, i, L->tave, &YI, &GI);
                if (IFLAG == 3) lc_linf4q(Ls, i, L->tave, &GAMMA1, &GAMMA2);
                if (IFLAG == 1 || IFLAG == 3) lc_linf4q_slope(Ls, i, L->tave, &sa, &sb);
                sline *S = &B[IJ];
                dsl_t *D = DB + (size_t)IJ * nd;
                double pshift = read_pshift(Ls, i);
                S->vnu = Ls->nu[i] + RHORAT * pshift;
                for (int c = 0; c < nd; c++) D[c].v = dl[c].drhorat * pshift;
                S->alfa = read_alfa(Ls, i);
                S->epp = Ls->e_lo[i];
                S->mol = m;
                if (S->vnu == 0.) { sui = 2. * sui; for (int c = 0; c < nd; c++) dsui[c] *= 2.; }
                if (S->epp <= -1.001) S->epp = fabs(S->epp);
                if (S->epp > -0.999) {
                    double sc = M->scor[m * 11 + iso], E = exp(-(S->epp * BETACR)), e = exp(-(S->vnu * BETA));
                    double f = sc * E * (1. + e);
                    for (int c = 0; c < nd; c++) {
                        double de = e * (-(D[c].v * BETA + S->vnu * dl[c].dbeta4));
                        double df = dl[c].dscor[m * 11 + iso] * E * (1. + e) + sc * E * (-(S->epp * dl[c].dbetacr)) * (1. + e) +
                                    sc * E * de;
                        dsui[c] = dsui[c] * f + sui * df;
                    }
                    sui = sui * f;
                }
                double tmpcor = pow(TRATIO, (double)Ls->tmpalf[i]);
                double alfa0i = S->alfa * tmpcor;
                double hwhmsi = (double)Ls->a_self[i] * tmpcor;
                double rs = M->rhoslf[m];
                double alfl = alfa0i * (RHORAT - rs) + hwhmsi * rs;
                double f3 = 1.0 - GAMMA1 * PAVP0 - GAMMA2 * PAVP2;
                for (int c = 0; c < nd; c++) {
                    double dtc = (double)Ls->tmpalf[i] * dl[c].dtmp;   /* d tmpcor / tmpcor */
                    double da = alfa0i * dtc * (RHORAT - rs) + alfa0i * (dl[c].drhorat - dl[c].drhoslf[m]) +
                                hwhmsi * dtc * rs + hwhmsi * dl[c].drhoslf[m];
                    if (IFLAG == 3) da = da * f3 + alfl * (-(sa * dl[c].dT) * PAVP0 - (sb * dl[c].dT) * PAVP2);
                    D[c].al = da;
                }
                if (IFLAG == 3) alfl = alfl * f3;
                S->alfa = alfl;
                double ad1 = M->alfd1[m * 11 + iso];
                S->epp = S->vnu * ad1;
                for (int c = 0; c < nd; c++) D[c].ad = D[c].v * ad1 + S->vnu * dl[c].dalfd1[m * 11 + iso];
                S->sp = sui * (1. + GI * PAVP2);
                S->spp = sui * YI * PAVP0;
                for (int c = 0; c < nd; c++) {
                    D[c].s = dsui[c] * (1. + GI * PAVP2) + sui * (sb * dl[c].dT) * PAVP2;
                    D[c].p = dsui[c] * YI * PAVP0 + sui * (sa * dl[c].dT) * PAVP0;
                    if (IFLAG != 1) { D[c].s = dsui[c]; D[c].p = 0.; }
                }
                if (S->vnu > VHI) { ieof = 1; break; }
            }
            break;
        }
        if (!got && !ieof) ieof = 2;
        if (ieof == 0 && IJ < LIMIN) continue;
        IHI = IJ;
        if (IHI >= ILO) IHI = shrinq_tan(B, DB, nd, ILO, IHI, dvr4);
        else IHI = ILO - 1;
        ILO = IHI + 1;
        IJ = ILO - 1;
        if (IHI < LIMIN && ieof == 0) continue;
        if (IHI > 0) {
            for (int k = 1; k <= IHI; k++) {
                shr_push(out, B[k].vnu, B[k].sp, B[k].alfa, B[k].epp, B[k].spp, B[k].mol);
                sht_push(tout, DB + (size_t)k * nd);
            }
            shr_end_block(out);
        }
        if (ieof) break;
        IJ = 0;
        ILO = 1;
    }
    free(B);
    free(DB);
    return 0;
}

/* RADFNI (clblm_ODLAY_CommonSub.f90:379-525) with the T-derivative of its ramp: the
 * same nodes, node values d RADFN / dT.  *drlast carries the derivative of RDLAST. */
static double radfni_d(double vi, double dvi, double t, double *vinew, double *rdel, double *rdlast, double *dr,
                       double *drdel, double *drlast) {
    double xkt = t / RADCN2;
    double v0 = vi;
    int first = *rdlast < 0.;
    double r = orc_radfni(vi, dvi, xkt, vinew, rdel, rdlast);
    double d0 = first ? orc_dradfn_dt(v0, t) : *drlast;
    int intvls = (int)((*vinew - v0) / dvi + 0.5);
    if (intvls < 1) intvls = 1;
    double dn = orc_dradfn_dt(*vinew, t);
    *dr = d0;
    *drdel = (dn - d0) / (double)intvls;
    *drlast = dn;
    return r;
}

/* LBLF4Q + CNVF4Q (clblm_ODLAY_LineF4.f90:25-494) with derivatives: R4[c] for the
 * value (c = 0) and each direction.  Returns NPTR4. */
static int lblf4q_tan(int jrad, double v1r4, double *v2r4, double **R4, int nd, const dlay_t *dl, double dvr4,
                      double bound4, double tave, shrunk_t *sh, shtan_t *st) {
    int LIMOUT = 2500;
    int nptr4 = ftrunc((*v2r4 - v1r4) / dvr4 + ONEPL);
    if (nptr4 > LIMOUT) nptr4 = LIMOUT;
    *v2r4 = v1r4 + dvr4 * (double)(nptr4 - 1);
    for (int c = 0; c <= nd; c++)
        for (int i = 1; i <= LIMOUT + 2; i++) R4[c][i] = 0.;
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
        if (sh->vnu[b1 - 1] < VLO) continue;
        for (int64_t i = b0; i < b1; i++) {
            if (sh->vnu[i] >= vnulst) { stop = 1; break; }
            double sp = sh->sp[i], spp = sh->spp[i];
            if (sp == 0. && spp == 0.) continue;
            const dsl_t *D = st->d + i * nd;
            double alfadi = sh->epp[i], alfali = sh->alfa[i];
            double zetai = alfali / (alfali + alfadi);
            int iz = ftrunc(100. * zetai + ONEPL);
            double zetdif = 100. * zetai - (double)(iz - 1);
            double avr = AVRAT(iz) + zetdif * (AVRAT(iz + 1) - AVRAT(iz));
            double alfavi = avr * (alfali + alfadi);
            double ralfvi = 1. / alfavi;
            double sil = sp * RECPI * alfali;
            double siv = (alfali * ralfvi) * sp * RECPI * ralfvi;
            double speak = A3 * fabs(siv);
            if (speak <= 0. && spp == 0.) continue;
            /* per-direction line constants */
            double dal[TAN_MAXDIR], dav[TAN_MAXDIR], dsil[TAN_MAXDIR], dsiv[TAN_MAXDIR], dbnd[TAN_MAXDIR];
            double drec2[TAN_MAXDIR];
            double sw = alfali + alfadi;
            double f4bnd0 = sil / (alfali * alfali + BNDSQ);
            for (int c = 0; c < nd; c++) {
                double dz = (D[c].al * alfadi - alfali * D[c].ad) / (sw * sw);
                double davr = 100. * dz * (AVRAT(iz + 1) - AVRAT(iz));
                dal[c] = D[c].al;
                dav[c] = davr * sw + avr * (D[c].al + D[c].ad);
                double dr = -dav[c] * ralfvi * ralfvi;
                dsil[c] = RECPI * (D[c].s * alfali + sp * D[c].al);
                dsiv[c] = RECPI * (D[c].s * alfali * ralfvi * ralfvi + sp * D[c].al * ralfvi * ralfvi +
                                   2. * sp * alfali * ralfvi * dr);
                double den = alfali * alfali + BNDSQ;
                dbnd[c] = dsil[c] / den - sil * 2. * alfali * D[c].al / (den * den);
                drec2[c] = -2. * dav[c] * ralfvi * ralfvi * ralfvi;
            }
            double vnui = sh->vnu[i];
            double sgn = 1.;   /* +1, then -1 for the Van Vleck-Weisskopf mirror */
            for (;;) {
                double xnui = vnui - v1r4;
                int jmin = ftrunc((xnui - bound4) / dvr4 + 2.);
                if (jmin > nptr4) break;
                if (jmin < 1) jmin = 1;
                int jmax = ftrunc((xnui + bound4) / dvr4 + 1.);
                if (jmax >= jmin) {
                    if (jmax > nptr4) jmax = nptr4;
                    double alfli2 = alfali * alfali;
                    double rec_alfvi2 = 1. / (alfavi * alfavi);
                    double dptrat = spp / (sp * alfavi);
                    int co2 = (sh->mol[i] == 2);
                    for (int jj = jmin; jj <= jmax; jj++) {
                        double xjj = (double)(jj - 1) * dvr4;
                        double xm = xjj - xnui;
                        double xmsq = xm * xm;
                        double zvsq = xmsq * rec_alfvi2;
                        int zone = zvsq <= ZSQBND;
                        double base = co2 ? (2. - xmsq * r_bndsq) * f4bnd0 : f4bnd0;
                        double f = zone ? (siv * A3 + zvsq * (siv * B3)) - base : sil / (alfli2 + xmsq) - base;
                        R4[0][jj] += spp == 0. ? f : f + xm * dptrat * f;
                        for (int c = 0; c < nd; c++) {
                            double dxm = -sgn * D[c].v;
                            double dxmsq = 2. * xm * dxm;
                            double dzv = dxmsq * rec_alfvi2 + xmsq * drec2[c];
                            double dbase = co2 ? -dxmsq * r_bndsq * f4bnd0 + (2. - xmsq * r_bndsq) * dbnd[c] : dbnd[c];
                            double df;
                            if (zone) df = dsiv[c] * A3 + dzv * siv * B3 + zvsq * dsiv[c] * B3 - dbase;
                            else {
                                double den = alfli2 + xmsq;
                                df = dsil[c] / den - sil * (2. * alfali * dal[c] + dxmsq) / (den * den) - dbase;
                            }
                            if (spp != 0.) {
                                double ddp = sgn * D[c].p / (sp * alfavi) - dptrat * (D[c].s / sp + dav[c] / alfavi);
                                df = df + (dxm * dptrat + xm * ddp) * f + xm * dptrat * df;
                            }
                            R4[1 + c][jj] += df;
                        }
                    }
                }
                if (vnui > 0. && vnui <= 25.) {
                    vnui = -sh->vnu[i];
                    spp = -spp;
                    sgn = -1.;
                    continue;
                }
                break;
            }
        }
    }
    if (jrad == 1) {
        double xkt = 1. / BETA;
        double vitst = v1r4 - dvr4, rdlast = -1., rdel, dr, drdel, drlast = 0.;
        int n1 = 0, n2 = 0;
        do {
            n1 = n2 + 1;
            double vi = v1r4 + dvr4 * (double)(n1 - 1);
            double radvi = radfni_d(vi, dvr4, tave, &vitst, &rdel, &rdlast, &dr, &drdel, &drlast);
            (void)xkt;
            n2 = ftrunc((vitst - v1r4) / dvr4 + 0.001);
            if (n2 > nptr4) n2 = nptr4;
            double t = radvi;
            for (int i = n1; i <= n2; i++) {
                double drad = dr + (double)(i - n1) * drdel;
                for (int c = 0; c < nd; c++) R4[1 + c][i] = R4[1 + c][i] * t + R4[0][i] * drad * dl[c].dT;
                R4[0][i] = R4[0][i] * t;
                t = t + rdel;
            }
        } while (n2 < nptr4);
    }
    return nptr4;
}

/* per-line LNCORQ values and derivatives held for CNVFNQ */
typedef struct {
    double vnu, sp, sppsp, recalf, zetai;
    int izeta;
    double dvnu[TAN_MAXDIR], dsp[TAN_MAXDIR], dsppsp[TAN_MAXDIR], drecalf[TAN_MAXDIR], dzeta[TAN_MAXDIR];
} htan;

int orc_odlay_tan(const orc_lines *Ls, const orc_layer *L, orc_params *P, int ndir, const orc_dir *dirs,
                  const double *dabsrb, double *out, int64_t cap, int64_t *nsamp, double *dv_used) {
    const int nd = ndir, nch = ndir + 1;
    if (nd < 0 || nd > TAN_MAXDIR) return fail("ndir out of range (0..16)");
    if (P->species_broad && Ls->brd_flg) return fail("species broadening is not differentiated");
    const double V1 = P->v1, V2 = P->v2;
    double DV = P->dvnormal;
    double ALFAV = P->sample * (P->dvnormal > P->dvnarrow ? P->dvnormal : P->dvnarrow);
    if (!(fabs(P->dvnormal - P->dvnarrow) / P->dvnormal > 0.2)) DV = P->dvnarrow;
    const double TAVE = L->tave, PAVE = L->pave;
    double XKT = TAVE / RADCN2;
    if ((V1 / XKT) < 5. && P->jrad != -1) P->jrad = 0;
    int JRAD = P->jrad;
    double V1ABS = 0., V2ABS = 0., DVABS = 0.;
    int NPTABS = 0;
    if (P->contnm) {
        DVABS = 1.;
        V1ABS = (double)ftrunc(V1);
        if (V1 < 0.) V1ABS = V1ABS - 1.;
        V1ABS = V1ABS - 3. * DVABS;
        V2ABS = (double)ftrunc(V2 + 3. * DVABS + 0.5);
        NPTABS = ftrunc((V2ABS - V1ABS) / DVABS + 1.5);
        if (P->nptabs != NPTABS) return fail("continuum array has the wrong length");
    }
    double SAMPLE = P->sample;
    double DVR4 = 64. * SAMPLE * DV / SAMPLE;
    double BOUND4 = 25.;
    double V1R4 = V1 - 2. * DVR4;
    double V2R4 = V2 + 2. * DVR4;
    double V1L4 = V1R4 - BOUND4 - DVR4;
    double V2L4 = V2R4 + BOUND4 + 2 * DVR4;

    molec_t M;
    memset(&M, 0, sizeof M);
    if (orc_molec(L, M.scor, M.rhoslf, M.alfd1)) return -1;
    dlay_t dl[TAN_MAXDIR];
    for (int c = 0; c < nd; c++)
        if (make_dlay(L, &M, &dirs[c], &dl[c])) return -1;

    shrunk_t sh = {0};
    shtan_t st = {0};
    st.nd = nd > 0 ? nd : 1;
    if (linf4q_tan(Ls, L, &M, nd, dl, V1L4, V2L4, DVR4, &sh, &st)) { shr_free(&sh); free(st.d); return -1; }

    const shc_t SC = shape_consts();
    const double HWF3 = 64., DXF1 = 0.002, DXF2 = 0.008, DXF3 = 0.032, HWF1 = 4.;
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
    int lb = 8 * MAX1 + 64;
    double *R1b[TAN_MAXDIR + 1], *R2b[TAN_MAXDIR + 1], *R3b[TAN_MAXDIR + 1], *R4b[TAN_MAXDIR + 1];
    double *R1[TAN_MAXDIR + 1], *R2[TAN_MAXDIR + 1], *R3[TAN_MAXDIR + 1], *R4[TAN_MAXDIR + 1];
    for (int c = 0; c < nch; c++) {
        R1b[c] = calloc((size_t)(MAX1 + 2 * lb), sizeof(double)); R1[c] = R1b[c] + lb;
        R2b[c] = calloc((size_t)(MAX2 + 2 * lb), sizeof(double)); R2[c] = R2b[c] + lb;
        R3b[c] = calloc((size_t)(MAX3 + 2 * lb), sizeof(double)); R3[c] = R3b[c] + lb;
        R4b[c] = calloc(2503, sizeof(double)); R4[c] = R4b[c];
    }
    htan *H = malloc(sizeof(htan) * 251);
    int ret = 0;

    double VFT = V1 - (double)NSHIFT * DV;
    double VBOT = V1 - BOUND, VTOP = V2 + BOUND;
    double V2R4ST = V2R4;
    lblf4q_tan(JRAD, V1R4, &V2R4, R4, nd, dl, DVR4, BOUND4, TAVE, &sh, &st);

    double XKT0 = TEMP0 / RADCN2;
    double BETACR = (1. / XKT) - (1. / XKT0);
    double TRATIO = TAVE / TEMP0;
    double RHORAT = (PAVE / P0) * (TEMP0 / TAVE);
    double PAVP0 = PAVE / P0;
    double PAVP2 = PAVP0 * PAVP0;
    const double CLC1 = 4. / 2000., CLC2 = 16. / 2000., CLC3 = 64. / 2000.;
    double WAVDXF = DV / DXF1, HWDXF = HWF1 / DXF1;
    double CONF2 = DV / DVR2, CONF3 = DV / DVR3;

    int64_t numTot = 0;
    int IEOF = 0, IDATA = 0, ILO = 0, IHI = -999, IPANEL = 0, ISTOP = 0, IFPAN = 1;
    int64_t nrec = (Ls->n + 249) / 250, rec = 0;
    int N1R2 = 1, N1R3 = 1, N2R3 = 155;
    int more_lines = 1;

    for (;;) {
        /* RDLIN + LNCORQ for the next record */
        if (more_lines && IEOF == 0) {
            int found = 0;
            int64_t a = 0, b = 0;
            while (rec < nrec) {
                a = rec * 250; b = a + 250 < Ls->n ? a + 250 : Ls->n;
                rec++;
                if (rec_vmax(Ls, a, b) < VBOT) continue;
                found = 1;
                break;
            }
            if (!found) IEOF = 1;
            else {
                IDATA = 0;
                int IJ = 0;
                int iout[251];
                for (int64_t i = a; i < b; i++)
                    if (Ls->iflg[i] >= 0) iout[++IJ] = (int)(i - a) + 1;
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
                for (int j = ILO; j <= IHI; j++) {
                    int64_t i = a + iout[j] - 1;
                    htan *h = &H[j];
                    h->sp = 0.; h->sppsp = 0.;
                    h->vnu = Ls->nu[i];
                    int m = Ls->mol[i] % 100;
                    int iso = (Ls->mol[i] % 1000) / 100;
                    if (iso == 0) iso = 10;
                    if (!sel_ok(L, m, iso)) continue;
                    double sui = (double)Ls->s[i] * L->w[m];
                    double dsui[TAN_MAXDIR];
                    for (int c = 0; c < nd; c++) dsui[c] = (double)Ls->s[i] * dl[c].dw[m];
                    if (JRAD == 1) {
                        double r = orc_radfn(h->vnu, XKT), dr = orc_dradfn_dt(h->vnu, TAVE);
                        for (int c = 0; c < nd; c++) dsui[c] = dsui[c] * r + sui * dr * dl[c].dT;
                        sui = sui * r;
                    }
                    if (sui == 0.) continue;
                    const int IFLAG = Ls->iflg[i];
                    double YI = 0., GI = 0., GAMMA1 = 0., GAMMA2 = 0., sa = 0., sb = 0.;
                    if (IFLAG == 1) lc_lncorq(Ls, i, TAVE, &YI, &GI);
                    if (IFLAG == 3) lc_lncorq(Ls, i, TAVE, &GAMMA1, &GAMMA2);
                    if (IFLAG == 1 || IFLAG == 3) lc_lncorq_slope(Ls, i, TAVE, &sa, &sb);
                    double pshift = read_pshift(Ls, i);
                    h->vnu = h->vnu + RHORAT * pshift;
                    for (int c = 0; c < nd; c++) h->dvnu[c] = dl[c].drhorat * pshift;
                    double tmpcor = pow(TRATIO, (double)Ls->tmpalf[i]);
                    double alfa0i = read_alfa(Ls, i) * tmpcor;
                    double hwhmsi = (double)Ls->a_self[i] * tmpcor;
                    double rs = M.rhoslf[m];
                    double alfl = alfa0i * (RHORAT - rs) + hwhmsi * rs;
                    double f3 = 1.0 - GAMMA1 * PAVP0 - GAMMA2 * PAVP2;
                    double dalfl[TAN_MAXDIR];
                    for (int c = 0; c < nd; c++) {
                        double dtc = (double)Ls->tmpalf[i] * dl[c].dtmp;
                        dalfl[c] = alfa0i * dtc * (RHORAT - rs) + alfa0i * (dl[c].drhorat - dl[c].drhoslf[m]) +
                                   hwhmsi * dtc * rs + hwhmsi * dl[c].drhoslf[m];
                        if (IFLAG == 3)
                            dalfl[c] = dalfl[c] * f3 + alfl * (-(sa * dl[c].dT) * PAVP0 - (sb * dl[c].dT) * PAVP2);
                    }
                    if (IFLAG == 3) alfl = alfl * f3;
                    double ad1 = M.alfd1[m * 11 + iso];
                    double alfad = h->vnu * ad1;
                    double zeta = alfl / (alfl + alfad);
                    h->zetai = zeta;
                    double fzeta = 100. * zeta;
                    int iz = ftrunc(fzeta + ONEPL);
                    h->izeta = iz;
                    double zetdif = fzeta - (double)(iz - 1);
                    double avr = AVRAT(iz) + zetdif * (AVRAT(iz + 1) - AVRAT(iz));
                    double alfv = avr * (alfl + alfad);
                    int clamped = 0;
                    if (alfv < DV) { alfv = DV; clamped = 1; }
                    if (alfv > ALFMAX) { alfv = ALFMAX; clamped = 1; }
                    if (HWF3 * alfv + h->vnu < VFT) continue;
                    h->recalf = 1. / alfv;
                    double sw = alfl + alfad;
                    for (int c = 0; c < nd; c++) {
                        double dad = h->dvnu[c] * ad1 + h->vnu * dl[c].dalfd1[m * 11 + iso];
                        h->dzeta[c] = (dalfl[c] * alfad - alfl * dad) / (sw * sw);
                        double dalfv = 100. * h->dzeta[c] * (AVRAT(iz + 1) - AVRAT(iz)) * sw + avr * (dalfl[c] + dad);
                        if (clamped) dalfv = 0.;
                        h->drecalf[c] = -dalfv / (alfv * alfv);
                    }
                    double srad = Ls->e_lo[i];
                    if (srad <= -1.001) srad = fabs(srad);
                    if (srad > -0.999) {
                        double sc = M.scor[m * 11 + iso], E = exp(-(srad * BETACR)), e = exp(-(h->vnu / XKT));
                        double f = sc * E * (1. + e);
                        for (int c = 0; c < nd; c++) {
                            double de = e * (-(h->dvnu[c] * XKT - h->vnu * dl[c].dxkt) / (XKT * XKT));
                            double df = dl[c].dscor[m * 11 + iso] * E * (1. + e) +
                                        sc * E * (-(srad * dl[c].dbetacr)) * (1. + e) + sc * E * de;
                            dsui[c] = dsui[c] * f + sui * df;
                        }
                        sui = sui * f;
                    }
                    h->sp = sui * (1. + GI * PAVP2);
                    double sppi = sui * YI * PAVP0;
                    h->sppsp = sppi / h->sp;
                    for (int c = 0; c < nd; c++) {
                        double dsp = IFLAG == 1 ? dsui[c] * (1. + GI * PAVP2) + sui * (sb * dl[c].dT) * PAVP2 : dsui[c];
                        double dspp = IFLAG == 1 ? dsui[c] * YI * PAVP0 + sui * (sa * dl[c].dT) * PAVP0 : 0.;
                        h->dsp[c] = dsp;
                        h->dsppsp[c] = (dspp - h->sppsp * dsp) / h->sp;
                    }
                    if (IFLAG == 0 && sui * h->recalf <= 0. && h->sppsp == 0.) {
                        h->sp = 0.; h->sppsp = 0.; continue;
                    }
                }
            }
            more_lines = 0;
        }
        if (IEOF != 0) { ILO = 1; IHI = 0; }
        /* CNVFNQ (clblm_ODLAY.f90:2067-2258) with the closed-form shapes */
        int ILAST = ILO - 1;
        if (ILO <= IHI) {
            IPANEL = -1;
            for (int j = ILO; j <= IHI; j++) {
                htan *h = &H[j];
                if (h->sp == 0.) continue;
                double zslope = h->recalf * WAVDXF;
                double zint = (h->vnu - VFT) / DV;
                double bhwdxf = HWDXF / zslope;
                int jmax1 = ftrunc(zint + bhwdxf + 1.5);
                if (jmax1 > MAX1) { ILAST = j - 1; IPANEL = 1; break; }
                int jmin1 = ftrunc(zint - bhwdxf + 1.5);
                double rshft = zint < 0.0 ? -0.5 : 0.5;
                int j2shft = ftrunc(zint * (1. - CONF2) + rshft);
                int j3shft = ftrunc(zint * (1. - CONF3) + rshft);
                double depth = h->sp * h->recalf;
                int izm = h->izeta;
                double zetdif = 100. * h->zetai - (double)(izm - 1);
                const double *tabs[5] = {CLBLM_CF1, CLBLM_CF2, CLBLM_CF3, CLBLM_CGAUSS, CLBLM_CER};
                double cf[5], ds[5][TAN_MAXDIR], s[5];
                for (int q = 0; q < 5; q++) {
                    const double *t = tabs[q];
                    cf[q] = t[izm - 1] + zetdif * (t[izm] - t[izm - 1]);
                    s[q] = depth * cf[q];
                    for (int c = 0; c < nd; c++) {
                        double dd = h->dsp[c] * h->recalf + h->sp * h->drecalf[c];
                        ds[q][c] = dd * cf[q] + depth * 100. * h->dzeta[c] * (t[izm] - t[izm - 1]);
                    }
                }
                double dptrat = h->sppsp;
                double dzsl[TAN_MAXDIR], dzint[TAN_MAXDIR];
                for (int c = 0; c < nd; c++) { dzsl[c] = h->drecalf[c] * WAVDXF; dzint[c] = h->dvnu[c] / DV; }
                const double conf[3] = {1., CONF2, CONF3}, dxf[3] = {DXF1, DXF2, DXF3};
                for (int j1 = jmin1; j1 <= jmax1; j1++) {
                    int jk[3] = {j1, j1 - j2shft, j1 - j3shft};
                    for (int k = 0; k < 3; k++) {
                        if (jk[k] < -lb) { ret = fail("panel index out of range"); goto done; }
                        double u = (double)(jk[k] - 1) - zint * conf[k];
                        double zf = u * zslope;
                        double x = fabs(zf) * dxf[k];
                        double sg = zf < 0. ? -1. : 1.;
                        fd_t F[3];
                        int nf = k == 0 ? 3 : 1;
                        double sk[3];
                        double dsk[3][TAN_MAXDIR];
                        if (k == 0) {
                            F[0] = sh_f(&SC, 1, x); F[1] = sh_fg(&SC, x); F[2] = sh_xver(&SC, x);
                            sk[0] = s[0]; sk[1] = s[3]; sk[2] = s[4];
                            for (int c = 0; c < nd; c++) { dsk[0][c] = ds[0][c]; dsk[1][c] = ds[3][c]; dsk[2][c] = ds[4][c]; }
                        } else {
                            F[0] = sh_f(&SC, k + 1, x);
                            sk[0] = s[k];
                            for (int c = 0; c < nd; c++) dsk[0][c] = ds[k][c];
                        }
                        double clc = k == 0 ? CLC1 : (k == 1 ? CLC2 : CLC3);
                        double *Rv = k == 0 ? R1[0] : (k == 1 ? R2[0] : R3[0]);
                        double val = 0.;
                        for (int q = 0; q < nf; q++) val += sk[q] * F[q].f;
                        double cpl = dptrat != 0. ? clc * dptrat * zf : 0.;
                        Rv[jk[k]] += val + val * cpl;
                        for (int c = 0; c < nd; c++) {
                            double dzf = u * dzsl[c] - conf[k] * dzint[c] * zslope;
                            double dx = sg * dxf[k] * dzf;
                            double dv = 0.;
                            for (int q = 0; q < nf; q++) dv += dsk[q][c] * F[q].f + sk[q] * F[q].d * dx;
                            double dcpl = dptrat != 0. ? clc * (h->dsppsp[c] * zf + dptrat * dzf) : 0.;
                            double *Rc = k == 0 ? R1[1 + c] : (k == 1 ? R2[1 + c] : R3[1 + c]);
                            Rc[jk[k]] += dv + dv * cpl + val * dcpl;
                        }
                    }
                }
            }
            if (IPANEL == -1) { ILAST = IHI; IPANEL = IDATA; }
        } else {
            IPANEL = 1;
        }
        ILO = ILAST + 1;
        if (IPANEL == 0 && IEOF == 0) { more_lines = 1; continue; }
        /* panelOut (clblm_ODLAY.f90:2289-2545) for every channel */
        if (IFPAN == 1) { IFPAN = 0; N1R2 = 1; N1R3 = 1; }
        else { N1R2 = (NSHIFT / 4) + 1 + 4; N1R3 = (NSHIFT / 16) + 1 + 3; }
        N2R3 = NLIM3 + (NSHIFT / 16) - 1 + 3;
        if (VFT <= 0.) { ret = fail("panel start <= 0 cm-1 (RSYM) not supported"); goto done; }
        for (int c = 0; c < nch; c++) {
            orc_xint(V1R4, V2R4, DVR4, R4[c], 1.0, VFT, DVR3, R3[c], N1R3, N2R3);
            if (P->contnm) {
                const double *ab = c == 0 ? P->absrb : (dabsrb ? dabsrb + (size_t)(c - 1) * (size_t)NPTABS : NULL);
                if (ab) orc_xint(V1ABS, V2ABS, DVABS, ab - 1, 1., VFT, DVR3, R3[c], N1R3, N2R3);
            }
        }
        {
            const double X00 = -7. / 128., X01 = 105. / 128., X02 = 35. / 128., X03 = -5. / 128.;
            const double X10 = -1. / 16., X11 = 9. / 16.;
            ISTOP = 0;
            if ((VFT + (double)(NHI - 1) * DV) > V2) {
                NHI = ftrunc((V2 - VFT) / DV + 1.);
                double V2P = VFT + (double)(NHI - 1) * DV;
                if (V2P < V2) NHI = NHI + 1;
                ISTOP = 1;
            }
            int NLIM = NHI - NLO + 1;
            double V1P = VFT + (double)(NLO - 1) * DV;
            int LIMLO = N1R2;
            if (N1R2 == 1) LIMLO = LIMLO + 4;
            int LIMHI = (NHI / 4) + 1;
            for (int c = 0; c < nch; c++) {
                double *r1 = R1[c], *r2 = R2[c], *r3 = R3[c];
                int J = LIMLO, J3 = 0;
                for (J = LIMLO; J <= LIMHI; J += 4) {
                    J3 = (J - 1) / 4 + 1;
                    r2[J] = r2[J] + r3[J3];
                    r2[J + 1] = r2[J + 1] + X00 * r3[J3 - 1] + X01 * r3[J3] + X02 * r3[J3 + 1] + X03 * r3[J3 + 2];
                    r2[J + 2] = r2[J + 2] + X10 * (r3[J3 - 1] + r3[J3 + 2]) + X11 * (r3[J3] + r3[J3 + 1]);
                    r2[J + 3] = r2[J + 3] + X03 * r3[J3 - 1] + X02 * r3[J3] + X01 * r3[J3 + 1] + X00 * r3[J3 + 2];
                }
                if (LIMLO <= LIMHI && ISTOP == 1) {
                    J3 = J3 + 1;
                    r2[J] = r2[J] + r3[J3];
                    r2[J + 1] = r2[J + 1] + X00 * r3[J3 - 1] + X01 * r3[J3] + X02 * r3[J3 + 1] + X03 * r3[J3 + 2];
                }
                for (J = NLO; J <= NHI; J += 4) {
                    int J2 = (J - 1) / 4 + 1;
                    r1[J] = r1[J] + r2[J2];
                    r1[J + 1] = r1[J + 1] + X00 * r2[J2 - 1] + X01 * r2[J2] + X02 * r2[J2 + 1] + X03 * r2[J2 + 2];
                    r1[J + 2] = r1[J + 2] + X10 * (r2[J2 - 1] + r2[J2 + 2]) + X11 * (r2[J2] + r2[J2 + 1]);
                    r1[J + 3] = r1[J + 3] + X03 * r2[J2 - 1] + X02 * r2[J2] + X01 * r2[J2 + 1] + X00 * r2[J2 + 2];
                }
            }
            if (JRAD == 0) {
                double VI = V1P - DV;
                double VITST = VI, RDLAST = -1., RDEL, DR, DRDEL, DRLAST = 0.;
                int NPTSI1, NPTSI2 = NLO - 1;
                do {
                    NPTSI1 = NPTSI2 + 1;
                    VI = VFT + (double)(NPTSI1 - 1) * DV;
                    double RADVI = radfni_d(VI, DV, TAVE, &VITST, &RDEL, &RDLAST, &DR, &DRDEL, &DRLAST);
                    NPTSI2 = ftrunc((VITST - VFT) / DV + 0.001);
                    if (NPTSI2 > NHI) NPTSI2 = NHI;
                    double t = RADVI;
                    for (int i = NPTSI1; i <= NPTSI2; i++) {
                        double drad = DR + (double)(i - NPTSI1) * DRDEL;
                        for (int c = 0; c < nd; c++) R1[1 + c][i] = R1[1 + c][i] * t + R1[0][i] * drad * dl[c].dT;
                        R1[0][i] = R1[0][i] * t;
                        t = t + RDEL;
                    }
                } while (NPTSI2 < NHI);
            }
            if (out) {
                if (numTot + NLIM > cap) { ret = fail("output capacity too small"); goto done; }
                for (int c = 0; c < nch; c++) memcpy(out + (size_t)c * (size_t)cap + numTot, R1[c] + NLO, sizeof(double) * (size_t)NLIM);
            }
            numTot += NLIM;
            VFT = VFT + (double)(NLIM1 - 1) * DV;
            if (ISTOP != 1) {
                for (int c = 0; c < nch; c++) {
                    memmove(R1[c] + 1, R1[c] + NLIM1, sizeof(double) * (size_t)(MAX1 - NLIM1 + 1));
                    for (int k = MAX1 - NLIM1 + 2; k <= MAX1; k++) R1[c][k] = 0.;
                    memmove(R2[c] + 1, R2[c] + NLIM2, sizeof(double) * (size_t)(MAX2 - NLIM2 + 1));
                    for (int k = MAX2 - NLIM2 + 2; k <= MAX2; k++) R2[c][k] = 0.;
                    memmove(R3[c] + 1, R3[c] + NLIM3, sizeof(double) * (size_t)(MAX3 - NLIM3 + 1));
                    for (int k = MAX3 - NLIM3 + 2; k <= MAX3; k++) R3[c][k] = 0.;
                }
                NLO = NSHIFT + 1;
            }
        }
        if (ISTOP == 1) break;
        {   /* F4 block refresh, clblm_ODLAY.f90:1469-1499 */
            double VF1 = VFT - 2. * DVR4;
            VF1 = V1 + (double)(int)floor((VF1 - V1) / DVR4) * DVR4;
            double VF2 = VFT + 2. * DVR4 + (double)(N2R3 + 4) * DVR3;
            if (VF2 > V2R4 && V2R4 != V2R4ST) {
                V1R4 = VF1;
                V2R4 = V2R4ST;
                lblf4q_tan(JRAD, V1R4, &V2R4, R4, nd, dl, DVR4, BOUND4, TAVE, &sh, &st);
            }
        }
    }
    if (dv_used) *dv_used = DV;
    if (P->dvout > 0. && !(fabs(P->dvout - DV) < 2.220446049250313e-16)) {
        int len = (int)ceil((V2 - V1) / P->dvout + 1.);
        if (out) {
            if (len > cap) { ret = fail("output capacity too small"); goto done; }
            double *tmp = malloc(sizeof(double) * (size_t)numTot);
            for (int c = 0; c < nch; c++) {
                memcpy(tmp, out + (size_t)c * (size_t)cap, sizeof(double) * (size_t)numTot);
                orc_interp(tmp, DV, V1, (int)numTot, out + (size_t)c * (size_t)cap, P->dvout, V1, len);
            }
            free(tmp);
        }
        numTot = len;
        if (dv_used) *dv_used = P->dvout;
    }
    *nsamp = numTot;
done:
    for (int c = 0; c < nch; c++) { free(R1b[c]); free(R2b[c]); free(R3b[c]); free(R4b[c]); }
    free(H);
    shr_free(&sh);
    free(st.d);
    return ret;
}
