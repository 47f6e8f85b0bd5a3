#pragma once
// Complex arithmetic shared by host and device code.  Semantics follow the
// reference's complex_type (include/compas/core/complex_type.h:24-325):
// interleaved (re, im), 8-byte aligned for float, products written as
// re*re - im*im / re*im + im*re.

#include <hip/hip_runtime.h>

namespace compas {

template<typename T>
struct alignas(2 * sizeof(T)) complex_type {
    T re, im;

    __host__ __device__ constexpr complex_type(T r = T(0), T i = T(0)) : re(r), im(i) {}

    template<typename R>
    __host__ __device__ explicit constexpr complex_type(const complex_type<R>& o) : re(T(o.re)), im(T(o.im)) {}

    __host__ __device__ T real() const { return re; }
    __host__ __device__ T imag() const { return im; }
    __host__ __device__ complex_type conj() const { return {re, -im}; }
    __host__ __device__ T norm() const { return re * re + im * im; }
};

using cfloat = complex_type<float>;
using cdouble = complex_type<double>;

template<typename T>
__host__ __device__ inline complex_type<T> conj(complex_type<T> a) {
    return {a.re, -a.im};
}
template<typename T>
__host__ __device__ inline complex_type<T> operator+(complex_type<T> a, complex_type<T> b) {
    return {a.re + b.re, a.im + b.im};
}
template<typename T>
__host__ __device__ inline complex_type<T> operator-(complex_type<T> a, complex_type<T> b) {
    return {a.re - b.re, a.im - b.im};
}
template<typename T>
__host__ __device__ inline complex_type<T> operator-(complex_type<T> a) {
    return {-a.re, -a.im};
}
template<typename T>
__host__ __device__ inline complex_type<T> operator*(complex_type<T> a, complex_type<T> b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template<typename T>
__host__ __device__ inline complex_type<T> operator*(complex_type<T> a, T s) {
    return {a.re * s, a.im * s};
}
template<typename T>
__host__ __device__ inline complex_type<T> operator*(T s, complex_type<T> a) {
    return {a.re * s, a.im * s};
}
template<typename T>
__host__ __device__ inline complex_type<T>& operator+=(complex_type<T>& a, complex_type<T> b) {
    a.re += b.re;
    a.im += b.im;
    return a;
}
template<typename T>
__host__ __device__ inline complex_type<T>& operator*=(complex_type<T>& a, complex_type<T> b) {
    return a = a * b;
}
template<typename T>
__host__ __device__ inline complex_type<T>& operator*=(complex_type<T>& a, T s) {
    a.re *= s;
    a.im *= s;
    return a;
}
template<typename T>
__host__ __device__ inline bool operator==(complex_type<T> a, complex_type<T> b) {
    return a.re == b.re && a.im == b.im;
}

}  // namespace compas
