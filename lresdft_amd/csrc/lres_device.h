// Device helpers shared by the kernel translation units: complex arithmetic and
// the pole set of the mu-centred circle.
#pragma once
#include <hip/hip_runtime.h>

namespace lres {

template <typename R>
struct cx {
    R re, im;
};

template <typename R>
__device__ __forceinline__ cx<R> cmul(cx<R> a, cx<R> b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}

template <typename R>
__device__ __forceinline__ cx<R> crecip(cx<R> a) {
    const R d = a.re * a.re + a.im * a.im;
    const R s = R(1) / d;
    return {a.re * s, -a.im * s};
}

// z_j = c + R e^{i theta_j}, w_j = R e^{i theta_j} / M, theta_j = pi (2j+1)/M (lrft/core.py:130-139)
__device__ __forceinline__ void pole(int j, int M, double c, double Rr, cx<double>& z, cx<double>& w,
                                     cx<double>& wz) {
    double s, co;
    sincospi((double)(2 * j + 1) / (double)M, &s, &co);
    z = {c + Rr * co, Rr * s};
    w = {Rr * co / M, Rr * s / M};
    wz = cmul(w, z);
}

}  // namespace lres
