// The MT_CKD water-vapour continuum table from absco-ref_wv-mt-ckd.nc, the file
// mt_ckd_h2o_absco reads through netCDF-Fortran (src/clblm_src/mt_ckd_h2o_module.f90:16,
// read_module.f90:50-101): a netCDF classic file (CDF-1, or CDF-2 with 64-bit
// offsets), of which this reads the header and the six variables the continuum uses.
// Host code; no netCDF library is needed.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/clblm_hip.h"

extern "C" __attribute__((visibility("hidden"))) int clblm_set_error(int rc, const char *msg);

namespace {

struct NcVar {
    std::string name;
    std::vector<int> dims;
    int type = 0;
    uint64_t begin = 0;
};

struct Reader {
    std::vector<unsigned char> b;
    size_t p = 0;
    bool bad = false;
    uint32_t u32() {
        if (p + 4 > b.size()) { bad = true; return 0; }
        uint32_t v = (uint32_t)b[p] << 24 | (uint32_t)b[p + 1] << 16 | (uint32_t)b[p + 2] << 8 | b[p + 3];
        p += 4;
        return v;
    }
    uint64_t u64() {
        uint64_t hi = u32();
        return hi << 32 | u32();
    }
    std::string name() {
        const uint32_t n = u32();
        if (bad || p + n > b.size()) { bad = true; return ""; }
        std::string s((const char *)&b[p], n);
        p += (n + 3) & ~3u;
        return s;
    }
    void skip_values(int type, uint32_t n) {
        static const int size[7] = {0, 1, 1, 2, 4, 4, 8};
        if (type < 1 || type > 6) { bad = true; return; }
        p += ((size_t)n * size[type] + 3) & ~(size_t)3;
        if (p > b.size()) bad = true;
    }
    void skip_atts() {
        const uint32_t tag = u32(), n = u32();
        if (tag == 0) return;
        for (uint32_t i = 0; i < n && !bad; i++) {
            name();
            const int t = (int)u32();
            skip_values(t, u32());
        }
    }
    double at(const NcVar &v, size_t i) {   // element i of a DOUBLE or FLOAT variable
        if (v.type == 6) {
            const size_t o = v.begin + 8 * i;
            if (o + 8 > b.size()) { bad = true; return 0.; }
            uint64_t u = 0;
            for (int k = 0; k < 8; k++) u = u << 8 | b[o + k];
            double d;
            std::memcpy(&d, &u, 8);
            return d;
        }
        if (v.type == 5) {
            const size_t o = v.begin + 4 * i;
            if (o + 4 > b.size()) { bad = true; return 0.; }
            uint32_t u = 0;
            for (int k = 0; k < 4; k++) u = u << 8 | b[o + k];
            float f;
            std::memcpy(&f, &u, 4);
            return f;
        }
        bad = true;
        return 0.;
    }
};

}  // namespace

extern "C" int clblm_mtckd_read_nc(const char *path, int *n, double *wavenumber, double *self_absco_ref,
                                   double *for_absco_ref, double *self_texp, double *ref_temp, double *ref_press) {
    if (!path || !n) return clblm_set_error(CLBLM_ERR_ARG, "null argument");
    FILE *f = std::fopen(path, "rb");
    if (!f) return clblm_set_error(CLBLM_ERR_ARG, (std::string("cannot open ") + path).c_str());
    Reader R;
    unsigned char buf[1 << 16];
    size_t k;
    while ((k = std::fread(buf, 1, sizeof buf, f)) > 0) R.b.insert(R.b.end(), buf, buf + k);
    std::fclose(f);
    if (R.b.size() < 8 || R.b[0] != 'C' || R.b[1] != 'D' || R.b[2] != 'F' || (R.b[3] != 1 && R.b[3] != 2))
        return clblm_set_error(CLBLM_ERR_ARG, "not a netCDF classic file (CDF-1 or CDF-2)");
    const bool off64 = R.b[3] == 2;
    R.p = 4;
    R.u32();   // numrecs
    std::vector<uint32_t> dimlen;
    {
        const uint32_t tag = R.u32(), nd = R.u32();
        if (tag != 0)
            for (uint32_t i = 0; i < nd && !R.bad; i++) {
                R.name();
                dimlen.push_back(R.u32());
            }
    }
    R.skip_atts();   // global attributes
    std::vector<NcVar> vars;
    {
        const uint32_t tag = R.u32(), nv = R.u32();
        if (tag != 0)
            for (uint32_t i = 0; i < nv && !R.bad; i++) {
                NcVar v;
                v.name = R.name();
                const uint32_t nd = R.u32();
                for (uint32_t d = 0; d < nd && !R.bad; d++) v.dims.push_back((int)R.u32());
                R.skip_atts();
                v.type = (int)R.u32();
                R.u32();   // vsize
                v.begin = off64 ? R.u64() : R.u32();
                vars.push_back(v);
            }
    }
    if (R.bad) return clblm_set_error(CLBLM_ERR_ARG, "truncated netCDF header");
    auto find = [&](const char *name) -> const NcVar * {
        for (const NcVar &v : vars)
            if (v.name == name) return &v;
        return nullptr;
    };
    const char *arrays[4] = {"wavenumbers", "self_absco_ref", "for_absco_ref", "self_texp"};
    const NcVar *va[4];
    for (int i = 0; i < 4; i++) {
        va[i] = find(arrays[i]);
        if (!va[i] || va[i]->dims.size() != 1 || va[i]->dims[0] < 0 || va[i]->dims[0] >= (int)dimlen.size())
            return clblm_set_error(CLBLM_ERR_ARG, (std::string("no 1-D variable ") + arrays[i]).c_str());
    }
    const NcVar *vt = find("ref_temp"), *vp = find("ref_press");
    if (!vt || !vp || !vt->dims.empty() || !vp->dims.empty())
        return clblm_set_error(CLBLM_ERR_ARG, "no scalar ref_temp / ref_press");
    const int len = (int)dimlen[(size_t)va[0]->dims[0]];
    for (int i = 1; i < 4; i++)
        if ((int)dimlen[(size_t)va[i]->dims[0]] != len) return clblm_set_error(CLBLM_ERR_ARG, "array lengths differ");
    *n = len;
    if (!wavenumber) return 0;
    double *out[4] = {wavenumber, self_absco_ref, for_absco_ref, self_texp};
    for (int i = 0; i < 4; i++)
        if (out[i])
            for (int j = 0; j < len; j++) out[i][j] = R.at(*va[i], (size_t)j);
    if (ref_temp) *ref_temp = R.at(*vt, 0);
    if (ref_press) *ref_press = R.at(*vp, 0);
    if (R.bad) return clblm_set_error(CLBLM_ERR_ARG, "netCDF data beyond the end of the file, or not FLOAT/DOUBLE");
    return 0;
}

extern "C" int clblm_mtckd_load_nc(clblm_ctx *ctx, const char *path, double xself, double xfrgn) {
    int n = 0;
    int rc = clblm_mtckd_read_nc(path, &n, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
    if (rc) return rc;
    std::vector<double> w((size_t)n), s((size_t)n), fo((size_t)n), t((size_t)n);
    double rt = 0., rp = 0.;
    rc = clblm_mtckd_read_nc(path, &n, w.data(), s.data(), fo.data(), t.data(), &rt, &rp);
    if (rc) return rc;
    return clblm_mtckd_load(ctx, n, w.data(), s.data(), fo.data(), t.data(), rt, rp, xself, xfrgn);
}
