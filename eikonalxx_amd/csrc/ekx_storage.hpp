// ekx_storage.hpp -- element types the 3D sweep kernels keep travel times
// and slowness in, and the arithmetic each one computes in.
//
//   float     FP32 values, FP32 arithmetic (FP64 inside the factored box)
//   double    FP64 values, FP64 arithmetic: the exact restatement
//   uint32_t  scaled integers, FP64 arithmetic.  A travel time is stored as
//             code * tStep, a slowness as code * sStep (steps per problem
//             and per model); the all-ones code is HUGE, an unreached node.
//             With tStep = Tmax / (2^32 - 2) a stored time is off by at
//             most tStep / 2 (about 1e-9 of Tmax), twenty times finer than
//             FP32 at the same 4 bytes per value.
#pragma once
#include "ekx_device.hpp"

namespace ekx
{

// Largest finite travel-time code; kTimeHugeCode marks an unreached node.
constexpr uint32_t kTimeMaxCode = 0xFFFFFFFEu;
constexpr uint32_t kTimeHugeCode = 0xFFFFFFFFu;

template <typename TS> struct Store;

template <> struct Store<float>
{
    using R = float;
    static constexpr float kHugeCode = (float) kHuge;
    static __device__ __forceinline__ float time(float v, const Problem3D &)
    { return v; }
    static __device__ __forceinline__ float slowness(float v,
                                                     const Problem3D &)
    { return v; }
    // code of a computed (finite) travel time
    static __device__ __forceinline__ float code(float v, const Problem3D &)
    { return v; }
    // value of a code that replaced a node's value (never the HUGE code)
    static __device__ __forceinline__ float finite_time(float v,
                                                        const Problem3D &)
    { return v; }
};

template <> struct Store<double>
{
    using R = double;
    static constexpr double kHugeCode = kHuge;
    static __device__ __forceinline__ double time(double v, const Problem3D &)
    { return v; }
    static __device__ __forceinline__ double slowness(double v,
                                                      const Problem3D &)
    { return v; }
    static __device__ __forceinline__ double code(double v, const Problem3D &)
    { return v; }
    static __device__ __forceinline__ double finite_time(double v,
                                                         const Problem3D &)
    { return v; }
};

template <> struct Store<uint32_t>
{
    using R = double;
    static constexpr uint32_t kHugeCode = kTimeHugeCode;
    static __device__ __forceinline__ double time(uint32_t v,
                                                  const Problem3D &p)
    {
        return v == kTimeHugeCode ? kHuge : (double) v*p.tStep;
    }
    static __device__ __forceinline__ double slowness(uint32_t v,
                                                      const Problem3D &p)
    {
        return (double) v*p.sStep;
    }
    // Round to the nearest code.  tStep comes from a bound on every finite
    // travel time of the problem, so a finite time never exceeds the
    // largest finite code; a candidate beyond it (HUGE, from unreached
    // neighbours) codes as HUGE and so never replaces a node's value.
    static __device__ __forceinline__ uint32_t code(double v,
                                                    const Problem3D &p)
    {
        double q = rint(v*p.tScale);
        return q <= (double) kTimeMaxCode ? (uint32_t) fmax(q, 0.0)
                                          : kTimeHugeCode;
    }
    static __device__ __forceinline__ double finite_time(uint32_t v,
                                                         const Problem3D &p)
    {
        return (double) v*p.tStep;
    }
};

} // namespace ekx
