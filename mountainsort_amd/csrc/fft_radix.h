// In-register complex DFTs of small compile-time lengths for the bandpass
// FFT.  dft<R, S>(x) transforms x[0..R) in place into natural order:
//   X[k] = sum_j x[j] exp(S * 2*pi*i * j*k / R),  S = -1 forward, +1 inverse
// (unnormalized).  Lengths 2, 3, 4, 5 are written out; composite lengths
// R = A*B are built by one in-register Cooley-Tukey step whose twiddles are
// compile-time constants.
#pragma once

#include <hip/hip_runtime.h>

namespace msamd {
namespace fft {

struct cf {
    float x, y;
};

__device__ __forceinline__ cf cadd(cf a, cf b) { return {a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ cf csub(cf a, cf b) { return {a.x - b.x, a.y - b.y}; }
__device__ __forceinline__ cf cmul(cf a, cf b) {
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
__device__ __forceinline__ cf cconjmul(cf a, cf b) {  // a * conj(b)
    return {a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y};
}
// multiply by S*i  (S = +1 -> i, S = -1 -> -i)
template <int S>
__device__ __forceinline__ cf mul_si(cf a) {
    return S > 0 ? cf{-a.y, a.x} : cf{a.y, -a.x};
}

// ---- compile-time sin/cos (Taylor series after reduction to [-pi, pi]) ----
constexpr double kPi = 3.14159265358979323846264338327950288;

constexpr double c_sin_red(double x) {
    double term = x, sum = x;
    for (int k = 1; k < 30; ++k) {
        term *= -x * x / ((2.0 * k) * (2.0 * k + 1.0));
        sum += term;
    }
    return sum;
}
constexpr double c_cos_red(double x) {
    double term = 1.0, sum = 1.0;
    for (int k = 1; k < 30; ++k) {
        term *= -x * x / ((2.0 * k - 1.0) * (2.0 * k));
        sum += term;
    }
    return sum;
}

// exp(S * 2*pi*i * m / R) as a table over m in [0, R)
template <int R, int S>
struct Twiddles {
    float re[R > 0 ? R : 1];
    float im[R > 0 ? R : 1];
    constexpr Twiddles() : re(), im() {
        for (int m = 0; m < R; ++m) {
            int mm = m;
            if (2 * mm > R) mm -= R;  // angle in [-pi, pi]
            const double th = 2.0 * kPi * mm / R;
            re[m] = static_cast<float>(c_cos_red(th));
            im[m] = static_cast<float>(S * c_sin_red(th));
        }
    }
};

template <int R, int S>
inline constexpr Twiddles<R, S> kTw{};

template <int R, int S>
__device__ __forceinline__ cf tw(int m) {
    return {kTw<R, S>.re[m % R], kTw<R, S>.im[m % R]};
}

// ---- smallest-factor split used for composite lengths ----
constexpr int first_factor(int R) {
    if (R % 4 == 0 && R != 4) return 4;
    if (R % 2 == 0 && R != 2) return 2;
    if (R % 3 == 0 && R != 3) return 3;
    if (R % 5 == 0 && R != 5) return 5;
    if (R % 7 == 0 && R != 7) return 7;
    return R;
}

template <int R, int S>
struct Dft;

template <int S>
struct Dft<1, S> {
    __device__ static __forceinline__ void run(cf*) {}
};

template <int S>
struct Dft<2, S> {
    __device__ static __forceinline__ void run(cf* x) {
        const cf a = x[0], b = x[1];
        x[0] = cadd(a, b);
        x[1] = csub(a, b);
    }
};

template <int S>
struct Dft<3, S> {
    __device__ static __forceinline__ void run(cf* x) {
        constexpr float c1 = -0.5f;
        constexpr float s1 = 0.866025403784438646763723170752936183f * S;
        const cf a = x[0], b = x[1], c = x[2];
        const cf t = cadd(b, c);
        const cf d = csub(b, c);
        x[0] = cadd(a, t);
        const cf m = {a.x + c1 * t.x, a.y + c1 * t.y};
        const cf r = {-s1 * d.y, s1 * d.x};  // s1 * i * d
        x[1] = cadd(m, r);
        x[2] = csub(m, r);
    }
};

template <int S>
struct Dft<4, S> {
    __device__ static __forceinline__ void run(cf* x) {
        const cf a0 = cadd(x[0], x[2]);
        const cf a1 = csub(x[0], x[2]);
        const cf b0 = cadd(x[1], x[3]);
        const cf b1 = mul_si<S>(csub(x[1], x[3]));
        x[0] = cadd(a0, b0);
        x[2] = csub(a0, b0);
        x[1] = cadd(a1, b1);
        x[3] = csub(a1, b1);
    }
};

template <int S>
struct Dft<5, S> {
    __device__ static __forceinline__ void run(cf* x) {
        constexpr float c1 = 0.309016994374947424102293417182819059f;   // cos(2pi/5)
        constexpr float c2 = -0.809016994374947424102293417182819059f;  // cos(4pi/5)
        constexpr float s1 = 0.951056516295153572116439333379382143f * S;  // sin(2pi/5)
        constexpr float s2 = 0.587785252292473129168705954639072769f * S;  // sin(4pi/5)
        const cf a = x[0];
        const cf t1 = cadd(x[1], x[4]), d1 = csub(x[1], x[4]);
        const cf t2 = cadd(x[2], x[3]), d2 = csub(x[2], x[3]);
        x[0] = {a.x + t1.x + t2.x, a.y + t1.y + t2.y};
        const cf m1 = {a.x + c1 * t1.x + c2 * t2.x, a.y + c1 * t1.y + c2 * t2.y};
        const cf m2 = {a.x + c2 * t1.x + c1 * t2.x, a.y + c2 * t1.y + c1 * t2.y};
        // i * (s1 d1 + s2 d2) and i * (s2 d1 - s1 d2)
        const cf n1 = {-(s1 * d1.y + s2 * d2.y), s1 * d1.x + s2 * d2.x};
        const cf n2 = {-(s2 * d1.y - s1 * d2.y), s2 * d1.x - s1 * d2.x};
        x[1] = cadd(m1, n1);
        x[4] = csub(m1, n1);
        x[2] = cadd(m2, n2);
        x[3] = csub(m2, n2);
    }
};

// Composite length R = A*B:
//   X[k1 + A*k2] = sum_b W_B^{b k2} W_R^{b k1} sum_a x[a*B + b] W_A^{a k1}
template <int R, int S>
struct Dft {
    static constexpr int A = first_factor(R);
    static constexpr int B = R / A;
    static_assert(A > 1 && A < R, "unsupported DFT length");
    __device__ static __forceinline__ void run(cf* x) {
        cf y[R];
#pragma unroll
        for (int b = 0; b < B; ++b) {
            cf t[A];
#pragma unroll
            for (int a = 0; a < A; ++a) t[a] = x[a * B + b];
            Dft<A, S>::run(t);
#pragma unroll
            for (int k1 = 0; k1 < A; ++k1)
                y[k1 * B + b] = (b * k1 == 0) ? t[k1] : cmul(t[k1], tw<R, S>(b * k1));
        }
#pragma unroll
        for (int k1 = 0; k1 < A; ++k1) {
            cf t[B];
#pragma unroll
            for (int b = 0; b < B; ++b) t[b] = y[k1 * B + b];
            Dft<B, S>::run(t);
#pragma unroll
            for (int k2 = 0; k2 < B; ++k2) x[k1 + A * k2] = t[k2];
        }
    }
};

}  // namespace fft
}  // namespace msamd
