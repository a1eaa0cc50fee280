// Exact closed form of a sequentially accumulated ramp.
//
// The reference applies its piecewise-linear radiation term by accumulating
//     t_0 = r,   t_{j+1} = fl(t_j + d)          (IEEE double, round to nearest even)
// one point after another (RADFNI callers: clblm_ODLAY.f90:2458-2481,
// clblm_ODLAY_LineF4.f90:152-177).  Inside one binade [2^E, 2^(E+1)) every t_j is a
// multiple of u = 2^(E-52), so fl(t_j + d) = t_j + delta with delta = d rounded to a
// multiple of u, as long as the result stays inside the binade: the ramp is exactly
// linear there.  ramp_pieces() splits j = 0..n-1 into pieces on which
//     t_j = base + (j - k) * step
// holds exactly (every product and sum in it is exact), taking the single steps that
// cross a binade edge as the sequential loop does.  Rounding ties (d exactly half-way
// between multiples of u, where the result depends on the parity of t_j) and large
// steps fall back to one piece per point.  Host and device code share this header.
#pragma once
#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#define SEQRAMP_HD __host__ __device__
#else
#define SEQRAMP_HD
#endif

namespace seqramp {

SEQRAMP_HD inline double pow2_of(double t) {  // 2^E for a positive normal t in [2^E, 2^(E+1))
    uint64_t b = __builtin_bit_cast(uint64_t, t);
    return __builtin_bit_cast(double, b & 0x7ff0000000000000ull);
}

// Calls emit(k, base, step) for each piece, in order of k (k = 0 first).  Returns the
// number of pieces, or -1 once more than cap would be needed.
template <class Emit>
SEQRAMP_HD inline int ramp_pieces(double t0, double d, int n, int cap, Emit &&emit) {
    const double TWO_M52 = 2.220446049250313e-16;  // 2^-52
    const double TWO_M12 = 2.44140625e-4;          // 2^-12
    double t = t0;
    int k = 0, np = 0;
    while (k < n) {
        if (np >= cap) return -1;
        const bool normal = t > 0. && t < 8.98846567431158e307;  // positive, below 2^1023
        const double p2 = normal ? pow2_of(t) : 0.;
        const double u = p2 * TWO_M52;
        if (d == 0. && normal) {
            emit(k, t, 0.);
            return np + 1;
        }
        bool single = !normal || !(p2 >= 2.2250738585072014e-308) || !(fabs(d) < p2 * TWO_M12);
        double delta = 0.;
        if (!single) {
            const double T = 1.5 * p2;
            delta = (T + d) - T;             // d rounded to a multiple of u
            const double rem = d - delta;    // exact
            single = fabs(rem) == 0.5 * u;   // tie: depends on the parity of t_j
        }
        if (single) {
            emit(k, t, 0.);
            np++;
            t = t + d;
            k++;
            continue;
        }
        if (delta == 0.) {  // |d| < u/2: fl(t + d) == t from here on
            emit(k, t, 0.);
            return np + 1;
        }
        // steps that stay at least one u inside the binade
        const double room = delta > 0. ? (2. * p2 - u) - t : t - (p2 + u);  // exact
        const int64_t D = (int64_t)(fabs(delta) / u);
        const int64_t nin = room < 0. ? 0 : (int64_t)(room / u) / D;
        emit(k, t, delta);
        np++;
        if ((int64_t)k + nin >= (int64_t)n - 1) return np;
        t = (t + (double)nin * delta) + d;  // the step that reaches the binade edge
        k += (int)nin + 1;
    }
    return np;
}

}  // namespace seqramp
