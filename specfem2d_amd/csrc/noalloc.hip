// An entry point that allocates no device memory: the max-norm check of get_norm_elastic_from_device_ with its
// reduction scratch in a file-scope device array instead of the handle's buffer.  For callers that cannot call
// hipMalloc (code under HIP graph capture, tight serving loops, harnesses that forbid allocation).
//
// The scratch holds one partial maximum per workgroup of the first pass (at most MAX_PARTIALS) and the final word.
// Each GPU holds its own copy of the array, at its own address, until the process ends; hipGetSymbolAddress gives
// that address once per device (at initialize_cuda_device_, or at the entry point's first call on a device) and
// only that address is ever handed to kernels or copies.  Calls on one GPU share the array: two calls that overlap
// on one GPU (two host threads, or two streams) race on it.  Nothing in it needs clearing between calls: the first
// pass writes every partial slot the second pass reads, and the second pass writes the final word.
#include <cmath>
#include <cstring>
#include <mutex>

#include "sf2d_mesh.h"
#include "specfem2d_gpu.h"

namespace sf2d {

namespace {

constexpr int NORM_THREADS = 256;
constexpr int MAX_PARTIALS = 2048;    // first-pass workgroups (grid-stride loop beyond that)
constexpr int MAX_DEVICES = 64;
// 2048 partial maxima + the result word, 8196 bytes per GPU
__device__ __align__(256) unsigned int scratch[MAX_PARTIALS + 1];

void* g_scratch[MAX_DEVICES] = {};
std::mutex g_scratch_mutex;

// pass 1: max |u|^2 over a grid-stride share of the points, one partial per workgroup (non-negative floats order
// as their bit patterns; a NaN wins, as in k_max_norm)
__global__ __launch_bounds__(NORM_THREADS) void k_norm_partial(const float2* __restrict__ f, int n,
                                                               unsigned int* __restrict__ partial) {
  __shared__ float s_m[NORM_THREADS / 64];
  float m = 0.f;
  for (int k = blockIdx.x * NORM_THREADS + threadIdx.x; k < n; k += gridDim.x * NORM_THREADS) {
    const float2 u = f[k];
    const float v = u.x * u.x + u.y * u.y;
    m = (v > m || v != v) ? v : m;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const float o = __shfl_xor(m, off, 64);
    m = (o > m || o != o) ? o : m;
  }
  if ((threadIdx.x & 63) == 0) s_m[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    float r = s_m[0];
#pragma unroll
    for (int w = 1; w < NORM_THREADS / 64; w++) r = (s_m[w] > r || s_m[w] != s_m[w]) ? s_m[w] : r;
    partial[blockIdx.x] = __float_as_uint(r);
  }
}

// pass 2 (one workgroup): the maximum of the np partials into *out
__global__ __launch_bounds__(NORM_THREADS) void k_norm_final(const unsigned int* __restrict__ partial, int np,
                                                             unsigned int* __restrict__ out) {
  __shared__ float s_m[NORM_THREADS / 64];
  float m = 0.f;
  for (int k = threadIdx.x; k < np; k += NORM_THREADS) {
    const float v = __uint_as_float(partial[k]);
    m = (v > m || v != v) ? v : m;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const float o = __shfl_xor(m, off, 64);
    m = (o > m || o != o) ? o : m;
  }
  if ((threadIdx.x & 63) == 0) s_m[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    float r = s_m[0];
#pragma unroll
    for (int w = 1; w < NORM_THREADS / 64; w++) r = (s_m[w] > r || s_m[w] != s_m[w]) ? s_m[w] : r;
    out[0] = __float_as_uint(r);
  }
}

}  // namespace

// device address of the scratch array on the current device, looked up once per device
unsigned int* noalloc_scratch() {
  int dev = -1;
  check_hip(hipGetDevice(&dev), "hipGetDevice");
  if (dev < 0 || dev >= MAX_DEVICES) fail("no-allocation scratch: device %d out of range", dev);
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch[dev]) {
    void* ptr = nullptr;
    check_hip(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    if (!ptr) fail("no-allocation scratch: no device address on device %d", dev);
    g_scratch[dev] = ptr;
  }
  return static_cast<unsigned int*>(g_scratch[dev]);
}

}  // namespace sf2d

using namespace sf2d;

extern "C" void get_norm_elastic_from_device_noalloc_(realw* norm, long* Mesh_pointer,
                                                      const int* FORWARD_OR_ADJOINT) {
  Mesh* mp = handle(Mesh_pointer);
  const float* f = (*FORWARD_OR_ADJOINT == 3) ? mp->d_b_displ : mp->d_displ;
  if (!f) fail("get_norm_elastic_from_device_noalloc: field not allocated");
  // the scratch serves any field size (grid-stride first pass); the point count itself must fit an int
  if (mp->nglob <= 0) fail("get_norm_elastic_from_device_noalloc: empty field (nglob %d)", mp->nglob);
  unsigned int* s = noalloc_scratch();
  long nb = ((long)mp->nglob + NORM_THREADS - 1) / NORM_THREADS;
  if (nb > MAX_PARTIALS) nb = MAX_PARTIALS;
  hipLaunchKernelGGL(k_norm_partial, dim3((unsigned)nb), dim3(NORM_THREADS), 0, mp->stream,
                     reinterpret_cast<const float2*>(f), mp->nglob, s);
  hipLaunchKernelGGL(k_norm_final, dim3(1), dim3(NORM_THREADS), 0, mp->stream, s, (int)nb, s + MAX_PARTIALS);
  check_hip(hipGetLastError(), "get_norm_elastic_from_device_noalloc launch");
  unsigned int bits = 0;
  check_hip(hipMemcpyAsync(&bits, s + MAX_PARTIALS, sizeof(bits), hipMemcpyDeviceToHost, mp->stream), "norm D2H");
  check_hip(hipStreamSynchronize(mp->stream), "get_norm_elastic_from_device_noalloc");
  float m;
  memcpy(&m, &bits, sizeof(m));
  *norm = std::sqrt(m);
}
