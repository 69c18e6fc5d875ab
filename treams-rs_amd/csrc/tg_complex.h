// Small complex arithmetic usable on host and device (gfx950), FP64 and FP32.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

#define TG_HD __host__ __device__ __forceinline__

template <class T>
struct cx {
    T re, im;
    TG_HD cx() : re(0), im(0) {}
    TG_HD cx(T r) : re(r), im(0) {}
    TG_HD cx(T r, T i) : re(r), im(i) {}
};

using cd = cx<double>;
using cf = cx<float>;

template <class T> TG_HD cx<T> operator+(cx<T> a, cx<T> b) { return {a.re + b.re, a.im + b.im}; }
template <class T> TG_HD cx<T> operator-(cx<T> a, cx<T> b) { return {a.re - b.re, a.im - b.im}; }
template <class T> TG_HD cx<T> operator-(cx<T> a) { return {-a.re, -a.im}; }
template <class T> TG_HD cx<T> operator*(cx<T> a, cx<T> b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <class T> TG_HD cx<T> operator*(T s, cx<T> a) { return {s * a.re, s * a.im}; }
template <class T> TG_HD cx<T> operator*(cx<T> a, T s) { return {s * a.re, s * a.im}; }
template <class T> TG_HD cx<T> &operator+=(cx<T> &a, cx<T> b) { a.re += b.re; a.im += b.im; return a; }
template <class T> TG_HD cx<T> &operator-=(cx<T> &a, cx<T> b) { a.re -= b.re; a.im -= b.im; return a; }
template <class T> TG_HD cx<T> conj(cx<T> a) { return {a.re, -a.im}; }
template <class T> TG_HD T abs2(cx<T> a) { return a.re * a.re + a.im * a.im; }

// a / b with the denominator scaled by its largest component (no overflow of |b|^2).
template <class T> TG_HD cx<T> operator/(cx<T> a, cx<T> b) {
    T s = fabs(b.re) > fabs(b.im) ? fabs(b.re) : fabs(b.im);
    T br = b.re / s, bi = b.im / s, ar = a.re / s, ai = a.im / s;
    T d = br * br + bi * bi;
    return {(ar * br + ai * bi) / d, (ai * br - ar * bi) / d};
}
template <class T> TG_HD cx<T> operator/(cx<T> a, T s) { return {a.re / s, a.im / s}; }
template <class T> TG_HD cx<T> operator/(T s, cx<T> b) { return cx<T>(s) / b; }

TG_HD double cabs_(cd a) { return hypot(a.re, a.im); }

// Principal square root.
TG_HD cd csqrt_(cd z) {
    double r = hypot(z.re, z.im);
    if (r == 0.0) return cd(0.0, 0.0);
    if (z.re >= 0.0) {
        double t = sqrt(0.5 * (r + z.re));
        return cd(t, z.im / (2.0 * t));
    }
    double t = sqrt(0.5 * (r - z.re));
    return cd(fabs(z.im) / (2.0 * t), z.im >= 0.0 ? t : -t);
}

TG_HD cd cexp_(cd z) {
    double e = exp(z.re);
    return cd(e * cos(z.im), e * sin(z.im));
}

// sin(z) = sin x cosh y + i cos x sinh y
TG_HD cd csin_(cd z) { return cd(sin(z.re) * cosh(z.im), cos(z.re) * sinh(z.im)); }
TG_HD cd ccos_(cd z) { return cd(cos(z.re) * cosh(z.im), -sin(z.re) * sinh(z.im)); }

// i^n for integer n.
TG_HD cd ipow_(int n) {
    int r = ((n % 4) + 4) % 4;
    return r == 0 ? cd(1, 0) : r == 1 ? cd(0, 1) : r == 2 ? cd(-1, 0) : cd(0, -1);
}
