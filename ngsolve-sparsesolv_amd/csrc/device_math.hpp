// device_math.hpp -- scalar types and arithmetic used by the gfx950 kernels.
//
// Scalars: double, float, and the complex pairs c128 (two doubles, 16 B,
// loaded as one dwordx4) and c64 (two floats).  Complex products are never
// conjugated implicitly: the complex-symmetric solvers use the bilinear
// form sum a_i b_i, and conjugation is requested explicitly (conj_if).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace ssa {

using index_t = std::int32_t;

struct __align__(16) c128 { double x, y; };
struct __align__(8) c64 { float x, y; };

template <typename T> struct real_of { using type = T; };
template <> struct real_of<c128> { using type = double; };
template <> struct real_of<c64> { using type = float; };

template <typename T> struct is_cplx { static constexpr bool value = false; };
template <> struct is_cplx<c128> { static constexpr bool value = true; };
template <> struct is_cplx<c64> { static constexpr bool value = true; };

template <typename T> __host__ __device__ inline T zero() { return T{}; }

// a vector store with the non-temporal hint (streams not read again before
// they have left the caches)
template <typename T>
__device__ inline void store_nt(T* p, const T& v)
{
    if constexpr (sizeof(T) == 16) {
        typedef double v2d __attribute__((ext_vector_type(2)));
        v2d w;
        __builtin_memcpy(&w, &v, 16);
        __builtin_nontemporal_store(w, reinterpret_cast<v2d*>(p));
    } else if constexpr (sizeof(T) == 8) {
        double w;
        __builtin_memcpy(&w, &v, 8);
        __builtin_nontemporal_store(w, reinterpret_cast<double*>(p));
    } else {
        static_assert(sizeof(T) == 4, "16-, 8- or 4-byte values");
        int w;
        __builtin_memcpy(&w, &v, 4);
        __builtin_nontemporal_store(w, reinterpret_cast<int*>(p));
    }
}

// a vector load with the non-temporal hint
template <typename T>
__device__ inline T load_nt(const T* p)
{
    T v;
    if constexpr (sizeof(T) == 16) {
        typedef double v2d __attribute__((ext_vector_type(2)));
        const v2d w = __builtin_nontemporal_load(reinterpret_cast<const v2d*>(p));
        __builtin_memcpy(&v, &w, 16);
    } else if constexpr (sizeof(T) == 8) {
        const double w = __builtin_nontemporal_load(reinterpret_cast<const double*>(p));
        __builtin_memcpy(&v, &w, 8);
    } else {
        static_assert(sizeof(T) == 4, "16-, 8- or 4-byte values");
        const int w = __builtin_nontemporal_load(reinterpret_cast<const int*>(p));
        __builtin_memcpy(&v, &w, 4);
    }
    return v;
}


// ---- conversions ------------------------------------------------------
template <typename To, typename From> __device__ inline To cvt(const From& a) { return To(a); }
template <> __device__ inline c128 cvt<c128, c128>(const c128& a) { return a; }
template <> __device__ inline c64 cvt<c64, c64>(const c64& a) { return a; }
template <> __device__ inline c128 cvt<c128, c64>(const c64& a) { return c128{double(a.x), double(a.y)}; }
template <> __device__ inline c64 cvt<c64, c128>(const c128& a) { return c64{float(a.x), float(a.y)}; }

// ---- arithmetic --------------------------------------------------------
__device__ inline double add(double a, double b) { return a + b; }
__device__ inline float add(float a, float b) { return a + b; }
__device__ inline c128 add(c128 a, c128 b) { return c128{a.x + b.x, a.y + b.y}; }
__device__ inline c64 add(c64 a, c64 b) { return c64{a.x + b.x, a.y + b.y}; }

__device__ inline double sub(double a, double b) { return a - b; }
__device__ inline float sub(float a, float b) { return a - b; }
__device__ inline c128 sub(c128 a, c128 b) { return c128{a.x - b.x, a.y - b.y}; }
__device__ inline c64 sub(c64 a, c64 b) { return c64{a.x - b.x, a.y - b.y}; }

__device__ inline double mul(double a, double b) { return a * b; }
__device__ inline float mul(float a, float b) { return a * b; }
__device__ inline c128 mul(c128 a, c128 b) { return c128{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
__device__ inline c64 mul(c64 a, c64 b) { return c64{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }

// acc + a*b (fused)
__device__ inline double fma_(double a, double b, double acc) { return __fma_rn(a, b, acc); }
__device__ inline float fma_(float a, float b, float acc) { return __fmaf_rn(a, b, acc); }
__device__ inline c128 fma_(c128 a, c128 b, c128 acc)
{
    c128 r;
    r.x = __fma_rn(a.x, b.x, acc.x);
    r.x = __fma_rn(-a.y, b.y, r.x);
    r.y = __fma_rn(a.x, b.y, acc.y);
    r.y = __fma_rn(a.y, b.x, r.y);
    return r;
}
__device__ inline c64 fma_(c64 a, c64 b, c64 acc)
{
    c64 r;
    r.x = __fmaf_rn(a.x, b.x, acc.x);
    r.x = __fmaf_rn(-a.y, b.y, r.x);
    r.y = __fmaf_rn(a.x, b.y, acc.y);
    r.y = __fmaf_rn(a.y, b.x, r.y);
    return r;
}
// acc - a*b (fused)
__device__ inline double fms_(double a, double b, double acc) { return __fma_rn(-a, b, acc); }
__device__ inline float fms_(float a, float b, float acc) { return __fmaf_rn(-a, b, acc); }
__device__ inline c128 fms_(c128 a, c128 b, c128 acc)
{
    c128 r;
    r.x = __fma_rn(-a.x, b.x, acc.x);
    r.x = __fma_rn(a.y, b.y, r.x);
    r.y = __fma_rn(-a.x, b.y, acc.y);
    r.y = __fma_rn(-a.y, b.x, r.y);
    return r;
}
__device__ inline c64 fms_(c64 a, c64 b, c64 acc)
{
    c64 r;
    r.x = __fmaf_rn(-a.x, b.x, acc.x);
    r.x = __fmaf_rn(a.y, b.y, r.x);
    r.y = __fmaf_rn(-a.x, b.y, acc.y);
    r.y = __fmaf_rn(-a.y, b.x, r.y);
    return r;
}

// c ? a : b per component (a ternary on a struct can become a select of
// addresses, which puts arrays of such values in scratch memory)
__device__ inline double select_(bool c, double a, double b) { return c ? a : b; }
__device__ inline float select_(bool c, float a, float b) { return c ? a : b; }
__device__ inline c128 select_(bool c, c128 a, c128 b) { return c128{c ? a.x : b.x, c ? a.y : b.y}; }
__device__ inline c64 select_(bool c, c64 a, c64 b) { return c64{c ? a.x : b.x, c ? a.y : b.y}; }

__device__ inline double conj_if(double a, bool) { return a; }
__device__ inline float conj_if(float a, bool) { return a; }
__device__ inline c128 conj_if(c128 a, bool c) { return c ? c128{a.x, -a.y} : a; }
__device__ inline c64 conj_if(c64 a, bool c) { return c ? c64{a.x, -a.y} : a; }

// |a|^2 in double (std::norm)
__device__ inline double norm2(double a) { return a * a; }
__device__ inline double norm2(c128 a) { return a.x * a.x + a.y * a.y; }

// scalar (re, im) pair <-> T
template <typename T> __device__ inline T from_pair(const double* p);
template <> __device__ inline double from_pair<double>(const double* p) { return p[0]; }
template <> __device__ inline c128 from_pair<c128>(const double* p) { return c128{p[0], p[1]}; }
template <> __device__ inline float from_pair<float>(const double* p) { return float(p[0]); }
template <> __device__ inline c64 from_pair<c64>(const double* p) { return c64{float(p[0]), float(p[1])}; }

__device__ inline void to_pair(double a, double* p) { p[0] = a; p[1] = 0.0; }
__device__ inline void to_pair(c128 a, double* p) { p[0] = a.x; p[1] = a.y; }

// ---- wavefront / workgroup reductions (fixed order: deterministic) -----
__device__ inline double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ inline c128 wave_sum(c128 v)
{
    v.x = wave_sum(v.x);
    v.y = wave_sum(v.y);
    return v;
}

// Sum over the workgroup; the result is valid in thread 0.  `scratch`
// must hold blockDim/64 values.
template <typename T>
__device__ inline T block_sum(T v, T* scratch)
{
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) scratch[wid] = v;
    __syncthreads();
    T s = zero<T>();
    if (threadIdx.x == 0) {
        const int nw = blockDim.x >> 6;
        for (int w = 0; w < nw; ++w) s = add(s, scratch[w]);
    }
    return s;
}

}  // namespace ssa
