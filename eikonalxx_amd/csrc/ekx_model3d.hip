// ekx_model3d.hip -- velocity models generated on the device: equal-
// thickness layers in z times (1 + a G), G a unit-variance Gaussian random
// field with covariance ~ exp(-r^2 / l^2) (white noise filtered by a
// separable Gaussian of sigma = l / 2 per axis, truncated at 4 sigma, with
// half-sample-symmetric boundaries).  The batched benchmark workloads are
// of this kind; generating them in HBM removes the host preparation and the
// upload from large runs.
//
// Every node's noise is a function of (seed, node index) only (a
// splitmix64 counter hash and the Box-Muller transform), so the field does
// not depend on the launch shape.  Everything is FP64 and HBM-bound.
#include "ekx_kernels.hpp"

namespace ekx
{

namespace
{

__device__ __forceinline__ uint64_t mix64(uint64_t x)
{
    x = (x ^ (x >> 30))*0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27))*0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// Uniform in (0, 1) from the 53 high bits of the k-th hash of seed.
__device__ __forceinline__ double uniform01(uint64_t seed, uint64_t k)
{
    const uint64_t z = mix64(seed + (k + 1)*0x9E3779B97F4A7C15ull);
    return ((double) (z >> 11) + 0.5)*0x1p-53;
}

// Half-sample-symmetric reflection of index i into [0, n).
__device__ __forceinline__ int reflect_index(int i, int n)
{
    const int period = 2*n;
    i %= period;
    if (i < 0) { i += period; }
    return i < n ? i : period - 1 - i;
}

} // namespace

__global__ void __launch_bounds__(256)
gaussian_noise_kernel(int64_t n, uint64_t seed, double *__restrict__ out)
{
    for (int64_t i = (int64_t) blockIdx.x*blockDim.x + threadIdx.x; i < n;
         i += (int64_t) gridDim.x*blockDim.x)
    {
        const double u1 = uniform01(seed, 2*(uint64_t) i);
        const double u2 = uniform01(seed, 2*(uint64_t) i + 1);
        out[i] = sqrt(-2.0*log(u1))*cos(6.283185307179586476925*u2);
    }
}

// One axis of the separable filter: out = sum_k w[k] in[reflect(i + k -
// radius)] along `axis` (0: x, 1: y, 2: z), taps in increasing k.
__global__ void __launch_bounds__(256)
filter_axis_kernel(int nx, int ny, int nz, int axis, int radius,
                   const double *__restrict__ w,
                   const double *__restrict__ in, double *__restrict__ out)
{
    const int64_t nxy = (int64_t) nx*ny;
    const int64_t n = nxy*nz;
    const int len = axis == 0 ? nx : (axis == 1 ? ny : nz);
    const int64_t stride = axis == 0 ? 1 : (axis == 1 ? nx : nxy);
    for (int64_t i = (int64_t) blockIdx.x*blockDim.x + threadIdx.x; i < n;
         i += (int64_t) gridDim.x*blockDim.x)
    {
        const int iz = (int) (i/nxy);
        const int rem = (int) (i - (int64_t) iz*nxy);
        const int iy = rem/nx;
        const int ix = rem - iy*nx;
        const int pos = axis == 0 ? ix : (axis == 1 ? iy : iz);
        const int64_t base = i - (int64_t) pos*stride;
        double acc = 0.0;
        for (int k = 0; k <= 2*radius; ++k)
        {
            const int j = reflect_index(pos + k - radius, len);
            acc += w[k]*in[base + (int64_t) j*stride];
        }
        out[i] = acc;
    }
}

// Per-block sums of v (power 1) or of (v - mean)^2 (power 2), or minima
// of v (power 0).
__global__ void __launch_bounds__(256)
moment_kernel(int64_t n, const double *__restrict__ v, double mean,
              int power, double *__restrict__ partial)
{
    __shared__ double red[256];
    double acc = power == 0 ? 1.0e308 : 0.0;
    for (int64_t i = (int64_t) blockIdx.x*blockDim.x + threadIdx.x; i < n;
         i += (int64_t) gridDim.x*blockDim.x)
    {
        const double d = v[i] - mean;
        if (power == 0) { acc = fmin(acc, v[i]); }
        else { acc += power == 1 ? v[i] : d*d; }
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1)
    {
        if ((int) threadIdx.x < k)
        {
            const double o = red[threadIdx.x + k];
            red[threadIdx.x] = power == 0 ? fmin(red[threadIdx.x], o)
                                          : red[threadIdx.x] + o;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) { partial[blockIdx.x] = red[0]; }
}

// v = layer velocity of the node's depth x (1 + amplitude (g - mean) / sd).
__global__ void __launch_bounds__(256)
layered_velocity_kernel(int nx, int ny, int nz, int nLayers,
                        const double *__restrict__ layerV, double amplitude,
                        double mean, double sd,
                        const double *__restrict__ g,
                        double *__restrict__ v)
{
    const int64_t nxy = (int64_t) nx*ny;
    const int64_t n = nxy*nz;
    for (int64_t i = (int64_t) blockIdx.x*blockDim.x + threadIdx.x; i < n;
         i += (int64_t) gridDim.x*blockDim.x)
    {
        const int iz = (int) (i/nxy);
        const int layer = min((int) (((int64_t) iz*nLayers)/nz), nLayers - 1);
        v[i] = layerV[layer]*(1.0 + amplitude*((g[i] - mean)/sd));
    }
}

static unsigned grid_of(int64_t n, unsigned cap)
{
    int64_t g = (n + 255)/256;
    if (g < 1) { g = 1; }
    return (unsigned) (g < cap ? g : cap);
}

hipError_t launch_gaussian_noise(int64_t n, uint64_t seed, double *out,
                                 hipStream_t stream)
{
    hipLaunchKernelGGL(gaussian_noise_kernel, dim3(grid_of(n, 16384)),
                       dim3(256), 0, stream, n, seed, out);
    return hipGetLastError();
}

hipError_t launch_filter_axis(int nx, int ny, int nz, int axis, int radius,
                              const double *w, const double *in, double *out,
                              hipStream_t stream)
{
    const int64_t n = (int64_t) nx*ny*nz;
    hipLaunchKernelGGL(filter_axis_kernel, dim3(grid_of(n, 16384)),
                       dim3(256), 0, stream, nx, ny, nz, axis, radius, w, in,
                       out);
    return hipGetLastError();
}

hipError_t launch_moment(int64_t n, const double *v, double mean, int power,
                         double *partial, int nPartial, hipStream_t stream)
{
    hipLaunchKernelGGL(moment_kernel, dim3((unsigned) nPartial), dim3(256), 0,
                       stream, n, v, mean, power, partial);
    return hipGetLastError();
}

hipError_t launch_layered_velocity(int nx, int ny, int nz, int nLayers,
                                   const double *layerV, double amplitude,
                                   double mean, double sd, const double *g,
                                   double *v, hipStream_t stream)
{
    const int64_t n = (int64_t) nx*ny*nz;
    hipLaunchKernelGGL(layered_velocity_kernel, dim3(grid_of(n, 16384)),
                       dim3(256), 0, stream, nx, ny, nz, nLayers, layerV,
                       amplitude, mean, sd, g, v);
    return hipGetLastError();
}

} // namespace ekx
