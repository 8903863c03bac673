// Operation count of one near-field (point, triangle) pair: the library's
// pair quadrature (elfmm_math.h, the code k_near runs) instantiated with a
// counting scalar. Additions/subtractions, multiplications and divisions
// are counted separately; sqrt, log and atan2 count as one transcendental
// each. The record of a triangle (built once per triangle) is not counted.
#include <math.h>

#include <cmath>
#include <cstring>

struct CD {
    double v;
    CD() = default;
    CD(double x) : v(x) {}
    explicit operator int() const { return (int)v; }
};

static long n_add, n_mul, n_div, n_tr;

inline CD operator+(CD a, CD b) { n_add++; return CD(a.v + b.v); }
inline CD operator-(CD a, CD b) { n_add++; return CD(a.v - b.v); }
inline CD operator*(CD a, CD b) { n_mul++; return CD(a.v * b.v); }
inline CD operator/(CD a, CD b) { n_div++; return CD(a.v / b.v); }
inline CD operator-(CD a) { return CD(-a.v); }
inline CD operator+(CD a) { return a; }
inline CD &operator+=(CD &a, CD b) { return a = a + b; }
inline CD &operator-=(CD &a, CD b) { return a = a - b; }
inline CD &operator*=(CD &a, CD b) { return a = a * b; }
inline CD &operator/=(CD &a, CD b) { return a = a / b; }
inline bool operator<(CD a, CD b) { return a.v < b.v; }
inline bool operator>(CD a, CD b) { return a.v > b.v; }
inline bool operator<=(CD a, CD b) { return a.v <= b.v; }
inline bool operator>=(CD a, CD b) { return a.v >= b.v; }
inline bool operator==(CD a, CD b) { return a.v == b.v; }
inline bool operator!=(CD a, CD b) { return a.v != b.v; }
inline CD sqrt(CD a) { n_tr++; return CD(std::sqrt(a.v)); }
inline CD log(CD a) { n_tr++; return CD(std::log(a.v)); }
inline CD atan2(CD a, CD b) { n_tr++; return CD(std::atan2(a.v, b.v)); }

#define double CD
#include "../../elfmmlib3d_amd/csrc/elfmm_math.h"
#undef double

// counts[4] = adds, muls, divs, transcendentals of nf_pair; out[9] its result
extern "C" void nf_opcount(const double *tri, const double *ssl, const double *sdl, const double *P,
                           int iself, double rlam, double rmu, int ifsl, int ifdl, long *counts,
                           double *out)
{
    CD t[9], s1[3], s2[3], p[3], rec[elfmm::NF_REC], acc[9];
    for (int i = 0; i < 9; i++) t[i] = tri[i];
    for (int i = 0; i < 3; i++) {
        s1[i] = ssl[i];
        s2[i] = sdl[i];
        p[i] = P[i];
    }
    for (int i = 0; i < 9; i++) acc[i] = 0.0;
    elfmm::nf_make_record(t, s1, s2, rec);
    CD c1 = CD((rlam + rmu) / (rlam + 2 * rmu));
    n_add = n_mul = n_div = n_tr = 0;
    elfmm::nf_pair(rec, p, iself, CD(rlam), CD(rmu), c1, ifsl, ifdl, acc);
    counts[0] = n_add;
    counts[1] = n_mul;
    counts[2] = n_div;
    counts[3] = n_tr;
    for (int i = 0; i < 9; i++) out[i] = acc[i].v;
}
