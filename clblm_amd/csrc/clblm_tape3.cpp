// TAPE3 line file reader: the binary line file LNFL writes and the reference reads
// with CLBLM_LineData_init (src/clblm_src/clblm_LineData.f90:99-256) and RDLIN
// (src/clblm_src/clblm_ODLAY.f90:2571-2798), RDLNFL (clblm_ODLAY_LineF4.f90:1355-1498).
//
// The file is Fortran unformatted sequential: every record is framed by its length
// in bytes (a 4-byte little-endian integer before and after it).
//   record 1     the file header; HLINID(10) (CHARACTER*8) leads it.  An '^' in the
//                eighth character of HLINID(7) announces a second header record
//                (the negative-E'' counts, clblm_LineData.f90:150-155), skipped here.
//   then, per panel of at most 250 lines, two records:
//     the panel header: VMIN, VMAX (REAL*8), NREC, NWDS (INTEGER*4)
//       (INPUT_HEADER, clblm_LineData.f90:52-58);
//     the panel: INPUT_BLOCK (clblm_LineData.f90:60-70), arrays of 250 whatever NREC:
//       VNU(250) REAL*8; SP, ALFA, EPP(250) REAL*4; MOL(250) INTEGER*4;
//       HWHM, TMPALF, PSHIFT(250) REAL*4; IFLG(250) INTEGER*4;
//       BRD_MOL_FLG_IN(7,250) INTEGER*4; BRD_MOL_DAT(21,250) REAL*4 (width, exponent
//       and shift of each of the 7 broadeners); SPEED_DEP(250) REAL*4.
//     A panel record of 10000 bytes (no broadening block) is read with zero flags.
// Reading stops at the end of the file or at a panel header whose fourth word is
// -99 (BUFIN_sgl's end mark, src/common/util_linux_intel.f90:121-148).
//
// The panels are the reference's records: RDLIN and LINF4Q take a panel at a time.
// The library's line lists hold records of 250 entries, so a panel of NREC < 250
// lines is padded with entries that are not lines (IFLG -1, the centre of the
// panel's last line), which every reader skips.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/clblm_hip.h"

extern "C" __attribute__((visibility("hidden"))) int clblm_set_error(int code, const char *msg);

namespace {

constexpr int NREC = 250;

struct Reader {
    FILE *f = nullptr;
    ~Reader() {
        if (f) fclose(f);
    }
    // one record; false at the end of the file
    bool next(std::vector<unsigned char> &buf, bool *bad) {
        int32_t len = 0, len2 = 0;
        *bad = false;
        if (fread(&len, 4, 1, f) != 1) return false;
        if (len < 0) { *bad = true; return false; }
        buf.resize((size_t)len);
        if (len && fread(buf.data(), 1, (size_t)len, f) != (size_t)len) { *bad = true; return false; }
        if (fread(&len2, 4, 1, f) != 1 || len2 != len) { *bad = true; return false; }
        return true;
    }
};

template <typename T>
T get(const std::vector<unsigned char> &b, size_t off) {
    T v;
    std::memcpy(&v, b.data() + off, sizeof v);
    return v;
}

struct Lists {
    std::vector<double> nu;
    std::vector<float> s, a_air, a_self, e_lo, tmpalf, shift, hw, tmp, shft;
    std::vector<int32_t> mol, iflg, flg;
};

// Reads the panels overlapping [vlo, vhi] (all of them when vlo > vhi), plus the
// first panel beyond vhi, into records of 250 entries.
int read_tape3(const char *path, double vlo, double vhi, Lists &L) {
    Reader r;
    r.f = fopen(path, "rb");
    if (!r.f) return clblm_set_error(CLBLM_ERR_ARG, (std::string("cannot open the line file ") + path).c_str());
    std::vector<unsigned char> hdr, ph, pb;
    bool bad = false;
    if (!r.next(hdr, &bad) || hdr.size() < 80)
        return clblm_set_error(CLBLM_ERR_ARG, "line file: no header record");
    if (hdr[6 * 8 + 7] == '^' && !r.next(pb, &bad))
        return clblm_set_error(CLBLM_ERR_ARG, "line file: the second header record is missing");
    const bool all = vlo > vhi;
    for (;;) {
        if (!r.next(ph, &bad)) {
            if (bad) return clblm_set_error(CLBLM_ERR_ARG, "line file: truncated or malformed record");
            break;
        }
        if (ph.size() < 24) return clblm_set_error(CLBLM_ERR_ARG, "line file: short panel header");
        if (get<int32_t>(ph, 12) == -99) break;
        const double vmin = get<double>(ph, 0), vmax = get<double>(ph, 8);
        const int nrec = get<int32_t>(ph, 16);
        if (!r.next(pb, &bad)) return clblm_set_error(CLBLM_ERR_ARG, "line file: panel header without its panel");
        if (nrec <= 0) continue;
        if (nrec > NREC) return clblm_set_error(CLBLM_ERR_ARG, "line file: panel of more than 250 lines");
        if (!all && vmax < vlo) continue;
        if (pb.size() < 10000) return clblm_set_error(CLBLM_ERR_ARG, "line file: short panel record");
        const bool brd = pb.size() >= 38000;
        double last = 0.;
        for (int i = 0; i < NREC; i++) {
            if (i < nrec) {
                const int32_t fl = get<int32_t>(pb, 9000 + 4 * i);
                const double v = get<double>(pb, 8 * i);
                L.nu.push_back(v);
                L.s.push_back(get<float>(pb, 2000 + 4 * i));
                L.a_air.push_back(get<float>(pb, 3000 + 4 * i));
                L.e_lo.push_back(get<float>(pb, 4000 + 4 * i));
                L.mol.push_back(get<int32_t>(pb, 5000 + 4 * i));
                L.a_self.push_back(get<float>(pb, 6000 + 4 * i));
                L.tmpalf.push_back(get<float>(pb, 7000 + 4 * i));
                L.shift.push_back(get<float>(pb, 8000 + 4 * i));
                L.iflg.push_back(fl);
                for (int j = 0; j < 7; j++) {
                    L.flg.push_back(brd ? get<int32_t>(pb, 10000 + 4 * (7 * i + j)) : 0);
                    L.hw.push_back(brd ? get<float>(pb, 17000 + 4 * (21 * i + 3 * j)) : 0.f);
                    L.tmp.push_back(brd ? get<float>(pb, 17000 + 4 * (21 * i + 3 * j + 1)) : 0.f);
                    L.shft.push_back(brd ? get<float>(pb, 17000 + 4 * (21 * i + 3 * j + 2)) : 0.f);
                }
                if (fl >= 0) last = v;
            } else {   // padding to the 250-entry record: not a line
                L.nu.push_back(last);
                L.s.push_back(0.f); L.a_air.push_back(0.f); L.e_lo.push_back(0.f); L.mol.push_back(0);
                L.a_self.push_back(0.f); L.tmpalf.push_back(0.f); L.shift.push_back(0.f);
                L.iflg.push_back(-1);
                for (int j = 0; j < 7; j++) {
                    L.flg.push_back(0); L.hw.push_back(0.f); L.tmp.push_back(0.f); L.shft.push_back(0.f);
                }
            }
        }
        if (!all && vmin > vhi) break;   // the first panel beyond the range is the last read
    }
    if (L.nu.empty()) return clblm_set_error(CLBLM_ERR_ARG, "line file: no panel in the requested range");
    return 0;
}

template <typename T>
void copy_out(T *dst, const std::vector<T> &src) {
    if (dst) std::memcpy(dst, src.data(), sizeof(T) * src.size());
}

}  // namespace

extern "C" {

int clblm_tape3_read(const char *path, double vlo, double vhi, int64_t *n, double *nu, float *s, float *a_air,
                     float *a_self, float *e_lo, float *tmpalf, float *shift, int32_t *mol, int32_t *iflg,
                     int32_t *brd_flg, float *brd_hw, float *brd_tmp, float *brd_shft) {
    if (!path || !n) return clblm_set_error(CLBLM_ERR_ARG, "path or n is NULL");
    Lists L;
    int rc = read_tape3(path, vlo, vhi, L);
    if (rc) return rc;
    const int64_t cnt = (int64_t)L.nu.size();
    if (!nu) {
        *n = cnt;
        return 0;
    }
    if (*n < cnt) return clblm_set_error(CLBLM_ERR_ARG, "arrays shorter than the record count");
    *n = cnt;
    copy_out(nu, L.nu); copy_out(s, L.s); copy_out(a_air, L.a_air); copy_out(a_self, L.a_self);
    copy_out(e_lo, L.e_lo); copy_out(tmpalf, L.tmpalf); copy_out(shift, L.shift); copy_out(mol, L.mol);
    copy_out(iflg, L.iflg); copy_out(brd_flg, L.flg); copy_out(brd_hw, L.hw); copy_out(brd_tmp, L.tmp);
    copy_out(brd_shft, L.shft);
    return 0;
}

int clblm_linelist_load_tape3(clblm_ctx *ctx, const char *path, double vlo, double vhi, int *handle, int64_t *n) {
    if (!ctx || !path || !handle) return clblm_set_error(CLBLM_ERR_ARG, "null argument");
    Lists L;
    int rc = read_tape3(path, vlo, vhi, L);
    if (rc) return rc;
    const int64_t cnt = (int64_t)L.nu.size();
    rc = clblm_linelist_upload(ctx, cnt, L.nu.data(), L.s.data(), L.a_air.data(), L.a_self.data(), L.e_lo.data(),
                               L.tmpalf.data(), L.shift.data(), L.mol.data(), L.iflg.data(), handle);
    if (rc) return rc;
    bool any = false;
    for (int32_t v : L.flg) any = any || v != 0;
    if (any) {
        rc = clblm_linelist_broadening(ctx, *handle, L.flg.data(), L.hw.data(), L.tmp.data(), L.shft.data());
        if (rc) {
            clblm_linelist_free(ctx, *handle);
            return rc;
        }
    }
    if (n) *n = cnt;
    return 0;
}

}  // extern "C"
