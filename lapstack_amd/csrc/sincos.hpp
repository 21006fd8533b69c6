// sincos.hpp -- sin and cos of a Sim's rotation, correctly rounded, the same
// bits on the host and on the device.
//
// The reference takes f64::sin / f64::cos (align.rs:243), which are the
// platform's libm: glibc 2.35 returns a value one ulp from the correctly
// rounded one for about 0.3 % of arguments in [-0.3, 0.3], and the device
// library differs from both in other cases.  The transform maps built from
// them (the warp's, the search's) must be the same bits wherever they are
// built, so every map here takes these instead: for |x| <= pi/4 the Taylor
// series summed by Horner's rule in double-double arithmetic (Dekker's exact
// product by a fused multiply-add and Knuth's two-sum; relative error about
// 2^-100, so the rounded result is the correctly rounded sin and cos except
// when the exact value lies within about 2^-100 of a rounding midpoint), and
// the library's functions beyond (no rotation the search reaches).  The
// coefficients are (-1)^k / (2k+1)! and (-1)^k / (2k)! split into a double and
// the double nearest the remainder.
#pragma once

#ifndef LSH_HD
#define LSH_HD __host__ __device__
#endif

namespace lsh {
namespace cr {

struct dd {
    double h, l;
};

LSH_HD inline dd two_prod(double a, double b) {
    dd r;
    r.h = a * b;
    r.l = __builtin_fma(a, b, -r.h);
    return r;
}
LSH_HD inline dd two_sum(double a, double b) {
    dd r;
    r.h = a + b;
    const double bb = r.h - a;
    r.l = (a - (r.h - bb)) + (b - bb);
    return r;
}
LSH_HD inline dd fast_two_sum(double a, double b) {
    dd r;
    r.h = a + b;
    r.l = b - (r.h - a);
    return r;
}
LSH_HD inline dd mul(dd a, dd b) {
    dd p = two_prod(a.h, b.h);
    p.l += a.h * b.l + a.l * b.h;
    return fast_two_sum(p.h, p.l);
}
LSH_HD inline dd add(dd a, dd b) {
    dd s = two_sum(a.h, b.h);
    s.l += a.l + b.l;
    return fast_two_sum(s.h, s.l);
}

}  // namespace cr

// sin(x) and cos(x), correctly rounded for |x| <= pi/4 (above)
LSH_HD inline void sincos_cr(double x, double *s, double *c) {
    if (!(x >= -0.78539816339744828 && x <= 0.78539816339744828)) {
        *s = sin(x);
        *c = cos(x);
        return;
    }
    // x^(2k) coefficients, k = 1..15, of the sine (over x) and of the cosine
    const double sc[30] = {-0.16666666666666666,    -9.25185853854297e-18,   0.008333333333333333,
                           1.1564823173178714e-19,  -0.0001984126984126984,  -1.7209558293420705e-22,
                           2.7557319223985893e-06,  -1.858393274046472e-22,  -2.505210838544172e-08,
                           1.448814070935912e-24,   1.6059043836821613e-10,  1.2585294588752098e-26,
                           -7.647163731819816e-13,  -7.03872877733453e-30,   2.8114572543455206e-15,
                           1.6508842730861433e-31,  -8.22063524662433e-18,   -2.2141894119604265e-34,
                           1.9572941063391263e-20,  -1.3643503830087908e-36, -3.868170170630684e-23,
                           8.843177655482344e-40,   6.446950284384474e-26,   -1.9330404233703465e-42,
                           -9.183689863795546e-29,  -1.4303150396787322e-45, 1.1309962886447716e-31,
                           1.0498015412959506e-47,  -1.216125041553518e-34,  -5.586290567888806e-51};
    const double cc[30] = {-0.5,                    0.0,                     0.041666666666666664,
                           2.3129646346357427e-18,  -0.001388888888888889,   5.300543954373577e-20,
                           2.48015873015873e-05,    2.1511947866775882e-23,  -2.755731922398589e-07,
                           -2.3767714622250297e-23, 2.08767569878681e-09,    -1.20734505911326e-25,
                           -1.1470745597729725e-11, -2.0655512752830745e-28, 4.779477332387385e-14,
                           4.399205485834081e-31,   -1.5619206968586225e-16, -1.1910679660273754e-32,
                           4.110317623312165e-19,   1.4412973378659527e-36,  -8.896791392450574e-22,
                           7.911402614872376e-38,   1.6117375710961184e-24,  -3.6846573564509766e-41,
                           -2.4795962632247976e-27, 1.2953730964765229e-43,  3.279889237069838e-30,
                           1.5117542744029879e-46,  -3.7699876288159054e-33, -2.5870347832750324e-49};
    const cr::dd x2 = cr::two_prod(x, x);
    cr::dd ps = {sc[28], sc[29]}, pc = {cc[28], cc[29]};
#pragma unroll
    for (int k = 13; k >= 0; k--) {
        ps = cr::add(cr::mul(ps, x2), cr::dd{sc[2 * k], sc[2 * k + 1]});
        pc = cr::add(cr::mul(pc, x2), cr::dd{cc[2 * k], cc[2 * k + 1]});
    }
    const cr::dd xs = {x, 0.0};
    *s = cr::add(xs, cr::mul(cr::mul(x2, ps), xs)).h;
    *c = cr::add(cr::dd{1.0, 0.0}, cr::mul(x2, pc)).h;
}

}  // namespace lsh
