// Correctly rounded FP64 division and square root for gfx950, cheaper than
// the general `/` and `sqrt` inside a guarded exponent range.
//
// The compiler lowers an FP64 division a / b to: v_div_scale on both
// operands, v_rcp_f64, two Newton-Raphson steps on the reciprocal
// (e = 1 - b y; y = y + y e), q = a y, r = a - b q (fma), a final
// fma(r, y, q) (v_div_fmas) and v_div_fixup for special operands; and sqrt(x)
// to: a scaling of tiny x, v_rsq_f64, a Goldschmidt/Newton refinement of
// s ~ sqrt(x) and h ~ 1/(2 sqrt(x)), the inverse scaling and a select for
// +-0 / +inf.  When b is in [2^-500, 2^500] and a is 0 or has |a| in
// [2^-500, 2 b], v_div_scale returns its operand unchanged and sets no
// scale flag, v_div_fmas is a plain fma and v_div_fixup returns its first
// operand with the sign of a / b; when x is in [2^-767, 2^1000] the sqrt
// scaling is by 2^0 and the special-case select keeps s.  In those ranges the
// operations below are the same operations on the same values, so the results
// are the correctly rounded IEEE ones, bit for bit, except that a zero
// numerator gives +0 where v_div_fixup gives the numerator's sign (a zero's
// sign never changes a nonzero result downstream).  The reciprocal
// refinement, which depends only on b, is computed once and shared by the
// three components of a unit vector.
//
// Callers test fast_range() and fall back to `/` and `sqrt` outside it.
#pragma once

#include <hip/hip_runtime.h>

namespace titan_mi355x {

// sqrt(x) for x in [2^-767, 2^1000]
__device__ __forceinline__ double fast_sqrt(double x) {
    const double y = __builtin_amdgcn_rsq(x);
    double s = x * y;
    double h = y * 0.5;
    const double e = fma(-h, s, 0.5);
    s = fma(s, e, s);
    h = fma(h, e, h);
    double d = fma(-s, s, x);
    s = fma(d, h, s);
    d = fma(-s, s, x);
    s = fma(d, h, s);
    return s;
}

// The refined reciprocal of b in [2^-500, 2^500].
__device__ __forceinline__ double fast_recip(double b) {
    double y = __builtin_amdgcn_rcp(b);
    double e = fma(-b, y, 1.0);
    y = fma(y, e, y);
    e = fma(-b, y, 1.0);
    y = fma(y, e, y);
    return y;
}

// a / b given y = fast_recip(b), for b > 0 and a == 0 or |a| in [2^-500, 2 b]
// (a == -0 gives +0).
__device__ __forceinline__ double fast_div(double a, double b, double y) {
    const double q = a * y;
    const double r = fma(-b, q, a);
    return fma(r, y, q);
}

// (bitwise & and | instead of && and ||: every comparison is evaluated, so
// the guards compile to compares and mask operations without branches)
__device__ __forceinline__ int num_ok(double a) { return (int)(a == 0.0) | (int)(fabs(a) >= 0x1p-500); }

// The guard for a spring axis t with squared length n2 = |t|^2: sqrt(n2) and
// t / sqrt(n2) may take the fast path.  (n2 in [2^-767, 2^1000] puts
// sqrt(n2) in [2^-384, 2^500], and every |t_c| <= sqrt(n2) (1 + 2^-52).)
__device__ __forceinline__ bool fast_range(double n2, double tx, double ty, double tz) {
    return ((int)(n2 >= 0x1p-767) & (int)(n2 <= 0x1p1000) & num_ok(tx) & num_ok(ty) & num_ok(tz)) != 0;
}

// A cheaper, stricter form of fast_range for straight-line code, as a wave
// mask: the lanes whose axis has every |t_c| >= 2^-383 and n2 <= 2^1000.  A
// zero component fails (such axes take the general path, which admits
// zeros).  Every t_c^2 is then at least 2^-766 (a correctly rounded square
// is monotone in |t_c|, and so is a sum of non-negative terms), so n2 >=
// 2^-766 needs no test of its own.  Each condition is one compare whose
// lane mask the ballot takes as it is, so the guard costs two min and two
// compares per axis and no VALU work to combine them.
__device__ __forceinline__ uint64_t fast_lanes_nz(double n2, double tx, double ty, double tz) {
    const double m = fmin(fmin(fabs(tx), fabs(ty)), fabs(tz));
    return __builtin_amdgcn_ballot_w64(n2 <= 0x1p1000) & __builtin_amdgcn_ballot_w64(m >= 0x1p-383);
}

// nrm = sqrt(tx tx + ty ty + tz tz) and u = t / nrm, correctly rounded, as
// the reference computes a spring's length and direction (a zero component
// of t gives a zero component of u, of either sign).
__device__ __forceinline__ void unit_axis(double tx, double ty, double tz, double &nrm, double &ux,
                                          double &uy, double &uz) {
    const double n2 = tx * tx + ty * ty + tz * tz;
    if (fast_range(n2, tx, ty, tz)) {
        nrm = fast_sqrt(n2);
        const double y = fast_recip(nrm);
        ux = fast_div(tx, nrm, y);
        uy = fast_div(ty, nrm, y);
        uz = fast_div(tz, nrm, y);
    } else {
        nrm = sqrt(n2);
        ux = tx / nrm;
        uy = ty / nrm;
        uz = tz / nrm;
    }
}

}  // namespace titan_mi355x
