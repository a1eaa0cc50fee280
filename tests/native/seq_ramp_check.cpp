// CPU check of seqramp::ramp_pieces against the sequential loop it replaces.
// Prints "cases mismatches naive_mismatches" and exits 0; the Python test asserts on them.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <random>
#include <vector>

#include "../../clblm_amd/csrc/seq_ramp.h"

struct Piece { int k; double base, step; };

static bool check(double t0, double d, int n, long *naive_bad) {
    std::vector<Piece> ps;
    int np = seqramp::ramp_pieces(t0, d, n, 1 << 20, [&](int k, double b, double s) { ps.push_back({k, b, s}); });
    if (np != (int)ps.size() || np < 1 || ps[0].k != 0) return false;
    double t = t0;
    size_t p = 0;
    bool naive_ok = true;
    for (int j = 0; j < n; j++) {
        while (p + 1 < ps.size() && ps[p + 1].k <= j) p++;
        volatile double prod = (double)(j - ps[p].k) * ps[p].step;
        double v = ps[p].base + prod;
        if (!(v == t)) return false;
        volatile double nv = t0 + (double)j * d;
        if (!(nv == t)) naive_ok = false;
        volatile double tn = t + d;
        t = tn;
    }
    if (!naive_ok) (*naive_bad)++;
    return true;
}

int main() {
    std::mt19937_64 g(12345);
    std::uniform_real_distribution<double> U(0., 1.);
    long cases = 0, bad = 0, naive_bad = 0;
    auto run = [&](double t0, double d, int n) {
        cases++;
        if (!check(t0, d, n, &naive_bad)) {
            bad++;
            if (bad < 10) std::printf("MISMATCH t0=%.17g d=%.17g n=%d\n", t0, d, n);
        }
    };
    // 1. generic: log-uniform magnitudes and relative steps, both signs
    for (int i = 0; i < 20000; i++) {
        double t0 = std::pow(10., -30. + 60. * U(g));
        double r = std::pow(10., -14. + 11. * U(g)) * (U(g) < 0.5 ? -1. : 1.);
        run(t0, t0 * r, 1 + (int)(3000 * U(g)));
    }
    // 2. crossings upwards and downwards close to powers of two
    for (int i = 0; i < 20000; i++) {
        int e = -60 + (int)(120 * U(g));
        double p2 = std::ldexp(1., e);
        double r = std::pow(10., -9. + 5. * U(g));
        int n = 1 + (int)(2500 * U(g));
        double frac = r * n * U(g);
        if (U(g) < 0.5) run(p2 * (1. - frac), p2 * r, n);
        else run(p2 * (1. + frac), -p2 * r, n);
    }
    // 3. exact ties: d = (m + 1/2) u of the start binade
    for (int i = 0; i < 5000; i++) {
        int e = -20 + (int)(40 * U(g));
        double p2 = std::ldexp(1., e), u = std::ldexp(1., e - 52);
        double t0 = p2 * (1. + 0.999 * U(g));
        t0 = std::floor(t0 / u) * u;
        double m = std::floor(1. + 1000. * U(g));
        run(t0, (m + 0.5) * u * (U(g) < 0.5 ? -1. : 1.), 1 + (int)(2500 * U(g)));
    }
    // 4. degenerate steps: zero, below half an ulp, large
    for (int i = 0; i < 2000; i++) {
        double t0 = std::pow(10., -5. + 10. * U(g));
        run(t0, 0., 100);
        run(t0, t0 * 1e-17, 500);
        run(t0, t0 * (0.01 + U(g)), 50);
        run(t0, -t0 * 0.3 * U(g), 20);
    }
    // 5. radiation-term-like ramps: RADFN(v, kT) magnitude, step over 1..10000 intervals
    for (int i = 0; i < 10000; i++) {
        double v = 10. + 3000. * U(g), xkt = 200. * (0.5 + U(g));
        double x = v / xkt, r = v * std::tanh(0.5 * x);
        double rn = (v * 1.003) * std::tanh(0.5 * x * 1.003);
        int iv = 1 + (int)(10000 * U(g));
        run(r, (rn - r) / iv, std::min(iv + 1, 2400));
    }
    std::printf("%ld %ld %ld\n", cases, bad, naive_bad);
    return 0;
}
