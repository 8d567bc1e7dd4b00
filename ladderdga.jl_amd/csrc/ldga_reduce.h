// Workgroup and device-wide sums shared by the reduction kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace ldga_red {

constexpr int kMaxWaves = 16;  // workgroups of up to 1024 threads

// sum over the workgroup; every thread gets the sums (all threads must call)
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double (*red)[kMaxWaves]) {
#pragma unroll
  for (int i = 0; i < NV; ++i)
    for (int off = 32; off > 0; off >>= 1) v[i] += __shfl_xor(v[i], off);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
  __syncthreads();
  if (lane == 0)
#pragma unroll
    for (int i = 0; i < NV; ++i) red[i][wv] = v[i];
  __syncthreads();
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    double s = 0.0;
    for (int j = 0; j < nw; ++j) s += red[i][j];
    v[i] = s;
  }
}

// Device-wide sum of v over the workgroups of the launch.  Every workgroup
// stores its partial sums (part: NV x gridDim.x doubles), then takes a ticket;
// the workgroup holding the last ticket returns true with tot = the sums of all
// partials, added in workgroup order (so the result does not depend on which
// workgroup finishes last), and resets the ticket to 0 for the next launch.
// No workgroup waits on another.  The ticket must be 0 at the first launch.
template <int NV>
__device__ bool device_sum(double (&v)[NV], double *part, unsigned *ticket, double (&tot)[NV]) {
  __shared__ double red[NV][kMaxWaves];
  __shared__ int last;
  block_sum<NV>(v, red);
  const int nb = gridDim.x;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < NV; ++i)
      __hip_atomic_store(part + (size_t)i * nb + blockIdx.x, v[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __threadfence();  // the partials are visible device-wide before the ticket
    last = atomicAdd(ticket, 1u) == (unsigned)(nb - 1);
  }
  __syncthreads();
  if (!last) return false;
  __threadfence();
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    double s = 0.0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x)
      s += __hip_atomic_load(part + (size_t)i * nb + b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    tot[i] = s;
  }
  block_sum<NV>(tot, red);
  if (threadIdx.x == 0) *ticket = 0u;
  return true;
}

}  // namespace ldga_red
