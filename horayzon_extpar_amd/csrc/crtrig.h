// Correctly rounded sin and cos of an FP64 argument, for the device prologue.
//
// The argument is reduced by pi/2 in double-double arithmetic (pi/2 split into
// 33 + 33 + 53 + 53 bits, so that k * C1 and k * C2 are exact for |k| < 2^20
// and x - k * C1 is exact by Sterbenz's lemma), sin and cos of the reduced
// argument r (|r| <= pi/4 + 1e-9) are summed as their Taylor series in
// double-double (Horner in r^2, terms up to r^29 / 29! and r^30 / 30!, below
// 2^-110 for |r| < 0.8), and the double-double result is rounded once to the
// nearest double.  The error before that rounding is near 2^-100 relative, so
// the result is the correctly rounded value unless the exact one lies within
// that distance of a rounding midpoint.  Arithmetic: IEEE add, multiply and
// fused multiply-add only (error-free transformations), so the host and the
// device produce the same bits.  Valid for |x| < 2^20 (callers check).

#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace hz {
namespace crt {

struct DD {
    double h, l;
};

__host__ __device__ inline DD two_sum(double a, double b) {
    const double s = a + b;
    const double v = s - a;
    return {s, (a - (s - v)) + (b - v)};
}

__host__ __device__ inline DD fast_two_sum(double a, double b) {
    const double s = a + b;
    return {s, b - (s - a)};
}

__host__ __device__ inline DD two_prod(double a, double b) {
    const double p = a * b;
    return {p, std::fma(a, b, -p)};
}

__host__ __device__ inline DD dd_add(DD x, DD y) {
    DD s = two_sum(x.h, y.h);
    const DD t = two_sum(x.l, y.l);
    s.l += t.h;
    s = fast_two_sum(s.h, s.l);
    s.l += t.l;
    return fast_two_sum(s.h, s.l);
}

__host__ __device__ inline DD dd_mul(DD x, DD y) {
    DD p = two_prod(x.h, y.h);
    p.l += x.h * y.l + x.l * y.h;
    return fast_two_sum(p.h, p.l);
}

// (-1)^n / (2n + 1)!, n = 0 .. 14, and (-1)^n / (2n)!, n = 0 .. 15, as
// double-double (high part rounded to nearest, low part the rounded rest)
__host__ __device__ inline DD sin_coef(int n) {
    switch (n) {
        case 0: return {1.0, 0.0};
        case 1: return {-0x1.5555555555555p-3, -0x1.5555555555555p-57};
        case 2: return {0x1.1111111111111p-7, 0x1.1111111111111p-63};
        case 3: return {-0x1.a01a01a01a01ap-13, -0x1.a01a01a01a01ap-73};
        case 4: return {0x1.71de3a556c734p-19, -0x1.c154f8ddc6c00p-73};
        case 5: return {-0x1.ae64567f544e4p-26, 0x1.c062e06d1f209p-80};
        case 6: return {0x1.6124613a86d09p-33, 0x1.f28e0cc748ebep-87};
        case 7: return {-0x1.ae7f3e733b81fp-41, -0x1.1d8656b0ee8cbp-97};
        case 8: return {0x1.952c77030ad4ap-49, 0x1.ac981465ddc6cp-103};
        case 9: return {-0x1.2f49b46814157p-57, -0x1.2650f61dbdcb4p-112};
        case 10: return {0x1.71b8ef6dcf572p-66, -0x1.d043ae40c4647p-120};
        case 11: return {-0x1.761b41316381ap-75, 0x1.3423c7d91404fp-130};
        case 12: return {0x1.3f3ccdd165fa9p-84, -0x1.58ddadf344487p-139};
        case 13: return {-0x1.d1ab1c2dccea3p-94, -0x1.054d0c78aea14p-149};
        default: return {0x1.259f98b4358adp-103, 0x1.eaf8c39dd9bc5p-157};
    }
}

__host__ __device__ inline DD cos_coef(int n) {
    switch (n) {
        case 0: return {1.0, 0.0};
        case 1: return {-0x1.0000000000000p-1, 0.0};
        case 2: return {0x1.5555555555555p-5, 0x1.5555555555555p-59};
        case 3: return {-0x1.6c16c16c16c17p-10, 0x1.f49f49f49f49fp-65};
        case 4: return {0x1.a01a01a01a01ap-16, 0x1.a01a01a01a01ap-76};
        case 5: return {-0x1.27e4fb7789f5cp-22, -0x1.cbbc05b4fa99ap-76};
        case 6: return {0x1.1eed8eff8d898p-29, -0x1.2aec959e14c06p-83};
        case 7: return {-0x1.93974a8c07c9dp-37, -0x1.05d6f8a2efd1fp-92};
        case 8: return {0x1.ae7f3e733b81fp-45, 0x1.1d8656b0ee8cbp-101};
        case 9: return {-0x1.6827863b97d97p-53, -0x1.eec01221a8b0bp-107};
        case 10: return {0x1.e542ba4020225p-62, 0x1.ea72b4afe3c2fp-120};
        case 11: return {-0x1.0ce396db7f853p-70, 0x1.aebcdbd20331cp-124};
        case 12: return {0x1.f2cf01972f578p-80, -0x1.9ada5fcc1ab14p-135};
        case 13: return {-0x1.88e85fc6a4e5ap-89, 0x1.71c37ebd16540p-143};
        case 14: return {0x1.0a18a2635085dp-98, 0x1.b9e2e28e1aa54p-153};
        default: return {-0x1.3932c5047d60ep-108, -0x1.832b7b530a627p-162};
    }
}

__host__ __device__ inline void sincos_cr(double x, double* sn, double* cs) {
    if (x == 0.0) {  // keeps the sign of zero
        *sn = x;
        *cs = 1.0;
        return;
    }
    // pi/2 = C1 + C2 + C3 + C4 (+ < 2^-176)
    const double C1 = 0x1.921fb54400000p+0;
    const double C2 = 0x1.0b4611a600000p-34;
    const double C3 = 0x1.3198a2e037073p-69;
    const double C4 = 0x1.129024e088a68p-123;
    const double k = std::rint(x * 0x1.45f306dc9c883p-1);
    const double t = x - k * C1;  // exact
    DD r = two_sum(t, -(k * C2));  // k * C2 exact
    r = dd_add(r, two_prod(-k, C3));
    r = dd_add(r, two_prod(-k, C4));
    const DD r2 = dd_mul(r, r);
    DD ps = sin_coef(14);
    for (int n = 13; n >= 0; n--) ps = dd_add(dd_mul(ps, r2), sin_coef(n));
    const DD s = dd_mul(r, ps);
    DD pc = cos_coef(15);
    for (int n = 14; n >= 0; n--) pc = dd_add(dd_mul(pc, r2), cos_coef(n));
    // the normalised pair's high part is the nearest double to the pair
    const double S = s.h, C = pc.h;
    const int q = (int)((long long)k & 3);
    if (q == 0) {
        *sn = S;
        *cs = C;
    } else if (q == 1) {
        *sn = C;
        *cs = -S;
    } else if (q == 2) {
        *sn = -S;
        *cs = -C;
    } else {
        *sn = -C;
        *cs = S;
    }
}

}  // namespace crt
}  // namespace hz
