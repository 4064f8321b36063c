// Deterministic pow for x > 0, identical bit for bit on the host and on the
// GPU.
//
// The step controllers raise error norms to fractional powers. Library pow
// implementations differ in their last bit between the host libm and the
// device math library, and in a chaotic system that one-ulp difference in
// the step size grows until two otherwise identical integrations part ways.
// This routine uses only IEEE-754 operations that every conforming platform
// rounds identically (+, -, *, /, fused multiply-add, exact frexp/ldexp and
// floor), in a fixed order, so the HIP kernel and the CPU oracle produce the
// same bits. It is accurate to a few ulp (pow(x, y) = exp(y log x) with a
// two-term ln 2 and series evaluated in Horner form).
//
// Header-only and free of HIP-specific code: compiled by hiprtc for the
// device and by g++ for the oracle.
#pragma once

#if defined(__HIPCC_RTC__) || defined(__HIP__)
#define CUBIE_HD __host__ __device__
#else
#define CUBIE_HD
#endif

namespace cubie_det {

template <typename T> struct Consts;
template <> struct Consts<double> {
    static constexpr double ln2_hi = 0x1.62e42fefa3800p-1;
    static constexpr double ln2_lo = 0x1.ef35793c76730p-45;
    static constexpr double inv_ln2 = 0x1.71547652b82fep+0;
    static constexpr double sqrt_half = 0x1.6a09e667f3bcdp-1;
    static constexpr int log_terms = 12;
    static constexpr int exp_terms = 16;
};
template <> struct Consts<float> {
    static constexpr float ln2_hi = 0x1.62e4p-1f;
    static constexpr float ln2_lo = 0x1.7f7d1cp-20f;
    static constexpr float inv_ln2 = 0x1.715476p+0f;
    static constexpr float sqrt_half = 0x1.6a09e6p-1f;
    static constexpr int log_terms = 6;
    static constexpr int exp_terms = 9;
};

CUBIE_HD inline double frexp_(double x, int* e) { return __builtin_frexp(x, e); }
CUBIE_HD inline float frexp_(float x, int* e) { return __builtin_frexpf(x, e); }
CUBIE_HD inline double ldexp_(double x, int e) { return __builtin_ldexp(x, e); }
CUBIE_HD inline float ldexp_(float x, int e) { return __builtin_ldexpf(x, e); }
CUBIE_HD inline double floor_(double x) { return __builtin_floor(x); }
CUBIE_HD inline float floor_(float x) { return __builtin_floorf(x); }
CUBIE_HD inline double fma_(double a, double b, double c) { return __builtin_fma(a, b, c); }
CUBIE_HD inline float fma_(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

// Natural log of x > 0: x = m 2^e, m in [sqrt(1/2), sqrt(2)),
// log m = 2 atanh(f), f = (m - 1) / (m + 1).
template <typename T>
CUBIE_HD inline T log_pos(T x) {
    int e = 0;
    T m = frexp_(x, &e);            // m in [0.5, 1)
    if (m < Consts<T>::sqrt_half) {
        m = m * T(2);
        e -= 1;
    }
    const T f = (m - T(1)) / (m + T(1));
    const T s = f * f;
    T poly = T(1) / T(2 * Consts<T>::log_terms + 1);
    for (int k = Consts<T>::log_terms - 1; k >= 0; --k) poly = fma_(poly, s, T(1) / T(2 * k + 1));
    const T logm = T(2) * f * poly;
    const T ef = T(e);
    return fma_(ef, Consts<T>::ln2_hi, fma_(ef, Consts<T>::ln2_lo, logm));
}

template <typename T>
CUBIE_HD inline T exp_(T z) {
    const T k = floor_(fma_(z, Consts<T>::inv_ln2, T(0.5)));
    T r = fma_(-k, Consts<T>::ln2_hi, z);
    r = fma_(-k, Consts<T>::ln2_lo, r);
    T poly = T(1);
    for (int i = Consts<T>::exp_terms; i >= 1; --i) poly = fma_(poly, r / T(i), T(1));
    return ldexp_(poly, int(k));
}

// pow(x, y) for x > 0 and moderate |y log x| (the controllers' range).
template <typename T>
CUBIE_HD inline T pow_pos(T x, T y) {
    if (x == T(1) || y == T(0)) return T(1);
    return exp_(y * log_pos(x));
}

}  // namespace cubie_det
