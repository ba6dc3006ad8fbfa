// Exact discrete gradient of the L2 waveform misfit on the GPU: reverse mode of the elastic Newmark time loop
// with checkpointed forward recomputation.  No counterpart in the reference, whose SIMULATION_TYPE=3 path
// reconstructs the forward field backwards in time instead (src/specfem2D/compute_kernels.f90).
//
// One forward step n, as the reference's GPU sequence runs it (iterate_time.F90, compute_gpu_elastic.f90):
//     d_n = d + dt v + dt^2/2 a          v' = v + dt/2 a                      (predictor)
//     r_n = -K d_n - B v' + f_n                                               (forces, Stacey, sources)
//     a_n = W r_n                        v_n = v' + dt/2 a_n                  (W = rmass_inverse; corrector)
//     s_n = R d_n                                                             (seismograms)
// With the corrector deferred to the next step, the device fields after step n hold exactly (d_n, v', r_n): a
// checkpoint is a copy of them, and restoring one re-arms the pending corrector.
//
// Reverse step n, with (D, V, A) the cotangents of (d_n, v_n, a_n) and chi = 1/2 dt sum |s - o|^2:
//     D += R^T dt (s_n - o_n)
//     V' = V;  A += dt/2 V;  R~ = W A;   d chi/d M += -A r_n W^2           (k_adj_pointwise)
//     D += -K R~;  V' += -B R~                                              (the forward force and Stacey kernels)
//     d chi/d(kappa, mu) += -d a(d_n, R~)/d(kappa, mu);  d chi/d(rho vp, rho vs) += -R~ . dB v'   (k_grad_*)
//     V = V' + dt D;  A = dt/2 V' + dt^2/2 D                                 (folded into the next k_adj_pointwise)
// K and B are symmetric, so their transposes are the forward kernels themselves.  Gradients accumulate in FP32
// over one checkpoint segment and are added into FP64 totals at its end.
//
// A shot batch (batch.hip) runs the same sweep for all its shots at once: states, adjoint fields and residuals
// carry a shot dimension, while the gradient accumulators hold the sum over shots; the kernels that touch the
// accumulators loop over the shots, so each accumulator is read and written once per step for the batch.
#include <algorithm>

#include "sf2d_mesh.h"
#include "specfem2d_gpu_ext.h"

namespace sf2d {

struct GradState {
  int nstep = 0, chunk = 0, nckpt = 0, nrec = 0, nshot = 1, batch = 0;
  size_t n2 = 0, ne = 0, nf = 0;
  long stride = 0;        // float2 between shots (shot-major: >= nglob; shot-minor: 1)
  int ps = 1;             // float2 between points (shot-minor: nshot)
  size_t nfield = 0;      // floats of one field over all shots
  float *fd = nullptr, *fv = nullptr, *fa = nullptr;  // the forward fields that states are copied from
  float* ckpt = nullptr;  // [nckpt][3][nfield]: (d, v', r) after steps chunk, 2 chunk, ...
  float* seg = nullptr;   // [chunk][3][nfield]: states of the segment being reversed
  float *D = nullptr, *Vb = nullptr, *Rb = nullptr;  // [nfield]
  float *gk32 = nullptr, *gm32 = nullptr, *pm32 = nullptr, *gvp32 = nullptr, *gvs32 = nullptr;
  double *gk = nullptr, *gm = nullptr, *pm = nullptr, *gvp = nullptr, *gvs = nullptr;
  float* res = nullptr;   // [nstep][nshot][nrec][2] = dt (s - o)
};

namespace {

constexpr int GT = 256;   // threads per workgroup

template <typename T>
__device__ __forceinline__ T ld_off(const T* base, unsigned byte_off) {  // 32-bit byte offset from a base
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + byte_off);
}
constexpr int GEPB = 10;  // elements per workgroup in k_grad_elastic

// fields of shot s hold point g at float2 g + s ss (shot-major batches and single shots); one thread per point,
// both components at once, the shots four at a time so that twelve loads are in flight per thread
__device__ __forceinline__ void adj_point(int first, float dt, float dto2, float h, float2 w, float2 vb, float2 db,
                                          float2 rr, float2& v_out, float2& r_out, float2& dm) {
  float2 v = make_float2(0.f, 0.f), a = make_float2(0.f, 0.f);
  if (!first) {  // state cotangents of the step after this one
    v.x = vb.x + dt * db.x;
    v.y = vb.y + dt * db.y;
    a.x = dto2 * vb.x + h * db.x;
    a.y = dto2 * vb.y + h * db.y;
  }
  a.x = a.x + dto2 * v.x;
  a.y = a.y + dto2 * v.y;
  v_out = v;
  r_out = make_float2(w.x * a.x, w.y * a.y);
  dm.x += a.x * rr.x * w.x * w.x;
  dm.y += a.y * rr.y * w.y * w.y;
}

__global__ __launch_bounds__(GT) void k_adj_pointwise(int first, int npts, int nshot, long ss, float dt, float dto2,
                                                      float h, const float2* __restrict__ D, float2* __restrict__ Vb,
                                                      float2* __restrict__ Rb, const float* __restrict__ rmx,
                                                      const float* __restrict__ rmz, const float2* __restrict__ r,
                                                      float2* __restrict__ pm) {
  constexpr int U = 4;
  for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < npts; g += gridDim.x * blockDim.x) {
    const float2 w = make_float2(rmx[g], rmz[g]);
    float2 dm = make_float2(0.f, 0.f);
    int s = 0;
    for (; s + U <= nshot; s += U) {
      float2 vb[U], db[U], rr[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
        const size_t q = (size_t)g + (size_t)(s + u) * ss;
        vb[u] = first ? make_float2(0.f, 0.f) : Vb[q];
        db[u] = first ? make_float2(0.f, 0.f) : D[q];
        rr[u] = r[q];
      }
#pragma unroll
      for (int u = 0; u < U; u++) {
        const size_t q = (size_t)g + (size_t)(s + u) * ss;
        float2 vo, ro;
        adj_point(first, dt, dto2, h, w, vb[u], db[u], rr[u], vo, ro, dm);
        Vb[q] = vo;
        Rb[q] = ro;
      }
    }
    for (; s < nshot; s++) {
      const size_t q = (size_t)g + (size_t)s * ss;
      float2 vo, ro;
      adj_point(first, dt, dto2, h, w, first ? make_float2(0.f, 0.f) : Vb[q], first ? make_float2(0.f, 0.f) : D[q],
                r[q], vo, ro, dm);
      Vb[q] = vo;
      Rb[q] = ro;
    }
    float2 p = pm[g];
    p.x -= dm.x;
    p.y -= dm.y;
    pm[g] = p;
  }
}

// residual injection; shot blockIdx.y (residuals [nshot][nrec][2] of this step; D point g of shot s at float2
// g ps + s ss)
__global__ __launch_bounds__(64) void k_adj_receivers(int nrec, const float* __restrict__ res, const int* __restrict__ ispec_rec,
                                                      const int* __restrict__ ibool, const float* __restrict__ xir,
                                                      const float* __restrict__ gammar, const float* __restrict__ cosrot,
                                                      const float* __restrict__ sinrot, float* __restrict__ D,
                                                      int ps, long ss) {
  const int r = blockIdx.x;
  const int p = threadIdx.x;
  if (r >= nrec || p >= NGLL2) return;
  res += (size_t)blockIdx.y * 2 * nrec;
  D += (size_t)2 * blockIdx.y * ss;
  const int i = p % NGLL, j = p / NGLL;
  const float px = res[2 * r], pz = res[2 * r + 1];
  const float c = cosrot[r], sn = sinrot[r];
  const float cx = c * px - sn * pz, cz = sn * px + c * pz;
  const float hl = xir[i * nrec + r] * gammar[j * nrec + r];
  const size_t g = (size_t)ibool[(size_t)ispec_rec[r] * NGLL2 + p] * ps;
  atomicAdd(&D[2 * g], hl * cx);
  atomicAdd(&D[2 * g + 1], hl * cz);
}

// strain products of the forward displacement d_n and the adjoint force R~ at every element point, summed over
// the shots (point g of shot s at float2 g ps + s ss)
__global__ __launch_bounds__(GT) void k_grad_elastic(int nelem, int nshot, int ps, long ss, const int* __restrict__ ibool,
                                                     const float* __restrict__ xix, const float* __restrict__ xiz,
                                                     const float* __restrict__ gammax, const float* __restrict__ gammaz,
                                                     const float2* __restrict__ dn, const float2* __restrict__ rb,
                                                     float* __restrict__ gk, float* __restrict__ gm, const Basis basis) {
  __shared__ float s_hp[NGLL2];
  __shared__ float2 s_d[GT], s_b[GT];
  const int tid = threadIdx.x;
  if (tid < NGLL2) s_hp[tid] = basis.hp[tid];
  const int le = tid / NGLL2, p = tid - le * NGLL2, i = p % NGLL, j = p / NGLL;
  const int e = blockIdx.x * GEPB + le;
  const bool active = le < GEPB && e < nelem;
  const size_t l = (size_t)e * NGLL2 + p;
  const int g = active ? ibool[l] : 0;
  float X = 0.f, Z = 0.f, G = 0.f, H = 0.f;
  if (active) {
    X = xix[l];
    Z = xiz[l];
    G = gammax[l];
    H = gammaz[l];
  }
  const int base = le * NGLL2;
  const float wij = basis.w[i] * basis.w[j];  // read once, not per shot
  float sk = 0.f, sm = 0.f;
  // the next shot's two values are loaded while the current one is computed
  float2 dnext = make_float2(0.f, 0.f), bnext = dnext;
  if (active) {
    dnext = dn[(size_t)g * ps];
    bnext = rb[(size_t)g * ps];
  }
  for (int s = 0; s < nshot; s++) {
    if (s > 0) __syncthreads();  // the previous shot's LDS reads are done
    if (active) {
      s_d[tid] = dnext;
      s_b[tid] = bnext;
    }
    __syncthreads();
    if (active && s + 1 < nshot) {
      const size_t q = (size_t)g * ps + (size_t)(s + 1) * ss;
      dnext = dn[q];
      bnext = rb[q];
    }
    if (active) {
      float a[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int k = 0; k < NGLL; k++) {
        const float hi = s_hp[i * NGLL + k], hj = s_hp[j * NGLL + k];
        const float2 dr = s_d[base + j * NGLL + k], dc = s_d[base + k * NGLL + i];
        const float2 br = s_b[base + j * NGLL + k], bc = s_b[base + k * NGLL + i];
        a[0] += dr.x * hi; a[1] += dr.y * hi; a[2] += dc.x * hj; a[3] += dc.y * hj;
        a[4] += br.x * hi; a[5] += br.y * hi; a[6] += bc.x * hj; a[7] += bc.y * hj;
      }
      const float ux = a[0] * X + a[2] * G, uxz = a[0] * Z + a[2] * H, uzx = a[1] * X + a[3] * G, uz = a[1] * Z + a[3] * H;
      const float wx = a[4] * X + a[6] * G, wxz = a[4] * Z + a[6] * H, wzx = a[5] * X + a[7] * G, wz = a[5] * Z + a[7] * H;
      const float wj = wij / (X * H - G * Z);
      sk += wj * (ux + uz) * (wx + wz);
      sm += wj * (ux * wx + uz * wz - ux * wz - uz * wx + (uxz + uzx) * (wxz + wzx));
    }
  }
  if (!active) return;
  gk[l] -= sk;
  gm[l] -= sm;
}

// Stacey traction's dependence on rho vp and rho vs, per absorbing face point (same order as abs_ij), summed over
// the shots (point g of shot s at float2 g ps + s ss)
__global__ __launch_bounds__(64) void k_grad_stacey(int nabs, int nshot, int ps, long ss, const int* __restrict__ abs_ispec,
                                                    const int* __restrict__ abs_ij,
                                                    const float* __restrict__ normal, const float* __restrict__ jacw,
                                                    const int* __restrict__ ibool, const float2* __restrict__ vpn,
                                                    const float2* __restrict__ rb, float* __restrict__ gvp,
                                                    float* __restrict__ gvs) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nabs * NGLL) return;
  const int f = q / NGLL;
  const int ii = abs_ij[2 * q], jj = abs_ij[2 * q + 1];
  if (ii < 0 || jj < 0) return;  // corner point owned by the other edge
  const int g = ibool[(size_t)abs_ispec[f] * NGLL2 + jj * NGLL + ii];
  const float nx = normal[2 * q], nz = normal[2 * q + 1], w = jacw[q];
  float sp = 0.f, sv = 0.f;
  for (int s = 0; s < nshot; s++) {
    const size_t o = (size_t)g * ps + (size_t)s * ss;
    const float2 v = vpn[o], b = rb[o];
    const float vn = nx * v.x + nz * v.y, bn = nx * b.x + nz * b.y;
    sp += w * vn * bn;
    sv += w * (v.x * b.x + v.y * b.y - vn * bn);
  }
  gvp[q] -= sp;
  gvs[q] -= sv;
}

// Shot-minor batches (nshot = 2^lshift values per point, at most 64): the same two reverse-sweep kernels with the
// shots in the lanes, so that every access is to consecutive values and the sum over the shots is a reduction
// across lanes.
// k_adj_pointwise over the flat points x shots x components entries; lanes of one (point, component) are 2 apart
__global__ __launch_bounds__(GT) void k_adj_pointwise_minor(int first, long n, int lshift, float dt, float dto2,
                                                            float h, const float* __restrict__ D,
                                                            float* __restrict__ Vb, float* __restrict__ Rb,
                                                            const float* __restrict__ rmx,
                                                            const float* __restrict__ rmz,
                                                            const float* __restrict__ r, float* __restrict__ pm) {
  const long stride = (long)gridDim.x * blockDim.x;  // a multiple of 64: groups of 2 nshot lanes stay together
  for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) {
    const long g = k >> (lshift + 1);
    const int c = (int)(k & 1);
    const float w = c ? rmz[g] : rmx[g];
    float v = 0.f, a = 0.f;
    if (!first) {
      const float vb = Vb[k], db = D[k];
      v = vb + dt * db;
      a = dto2 * vb + h * db;
    }
    a = a + dto2 * v;
    Vb[k] = v;
    Rb[k] = w * a;
    float dm = a * r[k] * w * w;
    for (int o = 2; o < (2 << lshift); o <<= 1) dm += __shfl_xor(dm, o, 64);
    if (((k >> 1) & ((1 << lshift) - 1)) == 0) pm[2 * g + c] -= dm;
  }
}

// k_grad_elastic with lane = shot: a wave takes 64 / 2^lshift elements; each lane stages its element's two fields in
// lane-private LDS and the strain products of the 2^lshift lanes of an element are summed by shuffles
__global__ __launch_bounds__(64) void k_grad_elastic_minor(int nelem, int nshot, int lshift,
                                                           const int* __restrict__ ibool,
                                                           const float4* __restrict__ geom,
                                                           const float2* __restrict__ dn,
                                                           const float2* __restrict__ rb, float* __restrict__ gk,
                                                           float* __restrict__ gm, const Basis basis) {
  __shared__ float2 s_d[NGLL2][64], s_b[NGLL2][64];
  const int lane = threadIdx.x;
  const int per_wave = 64 >> lshift;
  const int wave = blockIdx.x;
  if (wave * per_wave >= nelem) return;  // uniform
  const int e = wave * per_wave + (lane >> lshift);
  const int s = lane & ((1 << lshift) - 1);
  const bool live = e < nelem && s < nshot;
  const unsigned ec = (unsigned)min(e, nelem - 1);
  const unsigned sh = (unsigned)min(s, nshot - 1);
  const unsigned ps = 1u << lshift;
#pragma unroll
  for (int p = 0; p < NGLL2; p++) {
    const unsigned g = (unsigned)ld_off(ibool, 4u * (ec * NGLL2 + p));
    s_d[p][lane] = ld_off(dn, 8u * (g * ps + sh));
    s_b[p][lane] = ld_off(rb, 8u * (g * ps + sh));
  }
#pragma unroll
  for (int j = 0; j < NGLL; j++) {
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int i = 0; i < NGLL; i++) {
      float a[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int k = 0; k < NGLL; k++) {
        const float hi = basis.hp[i * NGLL + k], hj = basis.hp[j * NGLL + k];
        const float2 dr = s_d[j * NGLL + k][lane], dc = s_d[k * NGLL + i][lane];
        const float2 br = s_b[j * NGLL + k][lane], bc = s_b[k * NGLL + i][lane];
        a[0] += dr.x * hi; a[1] += dr.y * hi; a[2] += dc.x * hj; a[3] += dc.y * hj;
        a[4] += br.x * hi; a[5] += br.y * hi; a[6] += bc.x * hj; a[7] += bc.y * hj;
      }
      const unsigned l = ec * NGLL2 + j * NGLL + i;
      const float4 gp = ld_off(geom, 16u * l);
      const float X = gp.x, Z = gp.y, G = gp.z, H = gp.w;
      const float ux = a[0] * X + a[2] * G, uxz = a[0] * Z + a[2] * H, uzx = a[1] * X + a[3] * G, uz = a[1] * Z + a[3] * H;
      const float wx = a[4] * X + a[6] * G, wxz = a[4] * Z + a[6] * H, wzx = a[5] * X + a[7] * G, wz = a[5] * Z + a[7] * H;
      const float wj = basis.w[i] * basis.w[j] / (X * H - G * Z);
      float sk = live ? wj * (ux + uz) * (wx + wz) : 0.f;
      float sm = live ? wj * (ux * wx + uz * wz - ux * wz - uz * wx + (uxz + uzx) * (wxz + wzx)) : 0.f;
      for (int o = 1; o < (1 << lshift); o <<= 1) {
        sk += __shfl_xor(sk, o, 64);
        sm += __shfl_xor(sm, o, 64);
      }
      if (s == 0 && e < nelem) {
        gk[l] -= sk;
        gm[l] -= sm;
      }
    }
  }
}

__global__ __launch_bounds__(GT) void k_flush64(size_t n, float* __restrict__ a32, double* __restrict__ a64) {
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x) {
    a64[k] += (double)a32[k];
    a32[k] = 0.f;
  }
}

int grid_for(size_t n) { return (int)std::min<size_t>((n + GT - 1) / GT, 65536); }

GradState* need(Mesh* mp) {
  if (!mp->grad) fail("gradient buffers not prepared (sf2d_grad_prepare)");
  return mp->grad;
}

void copy_d2d(float* dst, const float* src, size_t n, hipStream_t s) {
  check_hip(hipMemcpyAsync(dst, src, n * sizeof(float), hipMemcpyDeviceToDevice, s), "gradient state copy");
}

// buffers for checkpoints every `chunk` steps, one segment of states and the accumulators, over the handle's
// own forward fields (batch = 0) or over its shot batch (batch = 1); returns the number of checkpoints
int grad_prepare(Mesh* mp, long* Mesh_pointer, int chunk, int nrec, int batch) {
  if (chunk < 1) fail("gradient: chunk %d < 1", chunk);
  if (nrec != mp->nrec_local) fail("gradient: %d receivers, the handle has %d", nrec, mp->nrec_local);
  if (GradState* old = mp->grad) {  // same layout as the buffers in place: keep them (repeated gradients)
    bool same = old->batch == batch && old->chunk == std::min(chunk, mp->nstep) && old->nstep == mp->nstep &&
                old->nrec == nrec;
    if (same && batch) {
      float *d, *v, *a;
      int ns, ps;
      long st;
      batch_fields(mp, &d, &v, &a, &ns, &st, &ps);
      same = old->fd == d && old->nshot == ns && old->stride == st && old->ps == ps;
    }
    if (same) return old->nckpt;
    sf2d_grad_release(Mesh_pointer);
  }
  GradState* G = new GradState();
  G->nstep = mp->nstep;
  G->chunk = std::min(chunk, mp->nstep);
  G->nckpt = (mp->nstep - 1) / G->chunk;
  G->nrec = nrec;
  G->n2 = (size_t)2 * mp->nglob;
  G->ne = (size_t)mp->nspec * NGLL2;
  G->nf = (size_t)mp->num_abs * NGLL;
  G->batch = batch;
  if (batch) {
    batch_fields(mp, &G->fd, &G->fv, &G->fa, &G->nshot, &G->stride, &G->ps);
  } else {
    G->fd = mp->d_displ;
    G->fv = mp->d_veloc;
    G->fa = mp->d_accel;
    G->nshot = 1;
    G->stride = mp->nglob;
    G->ps = 1;
  }
  G->nfield = G->ps > 1 ? (size_t)2 * mp->nglob * G->nshot : (size_t)2 * G->stride * G->nshot;
  if (G->ps > 1 && (G->ps > 64 || (G->ps & (G->ps - 1)) || !mp->d_geom))
    fail("gradient: shot-minor batches need a power of two of at most 64 shots and the patch assembly");
  const size_t states = (size_t)std::max(1, G->nckpt) + G->chunk;
  if (states * 3 * G->nfield * sizeof(float) > ((size_t)240 << 30))
    fail("gradient: %zu states of %d shots exceed 240 GiB; use a longer chunk or fewer shots", states, G->nshot);
  G->ckpt = device_alloc<float>(3 * G->nfield * std::max(1, G->nckpt));  // written before they are read
  G->seg = device_alloc<float>(3 * G->nfield * G->chunk);
  G->D = device_zeros<float>(G->nfield);
  G->Vb = device_zeros<float>(G->nfield);
  G->Rb = device_zeros<float>(G->nfield);
  G->gk32 = device_zeros<float>(G->ne);
  G->gm32 = device_zeros<float>(G->ne);
  G->pm32 = device_zeros<float>(G->n2);
  G->gvp32 = device_zeros<float>(std::max<size_t>(1, G->nf));
  G->gvs32 = device_zeros<float>(std::max<size_t>(1, G->nf));
  G->gk = device_zeros<double>(G->ne);
  G->gm = device_zeros<double>(G->ne);
  G->pm = device_zeros<double>(G->n2);
  G->gvp = device_zeros<double>(std::max<size_t>(1, G->nf));
  G->gvs = device_zeros<double>(std::max<size_t>(1, G->nf));
  G->res = device_zeros<float>((size_t)std::max(1, nrec) * 2 * G->nshot * mp->nstep);
  mp->grad = G;
  mp->grad_batch = batch;
  return G->nckpt;
}

}  // namespace
}  // namespace sf2d

using namespace sf2d;

extern "C" {

int sf2d_grad_prepare(long* Mesh_pointer, int chunk, int nrec) {
  Mesh* mp = handle(Mesh_pointer);
  if (!mp->sorted) fail("sf2d_grad_prepare: prepare_fields_elastic_device was not called");
  if (mp->nspec_acoustic > 0) fail("sf2d_grad_prepare: gradients of meshes with fluid elements are not supported");
  return grad_prepare(mp, Mesh_pointer, chunk, nrec, 0);
}

int sf2d_batch_grad_prepare(long* Mesh_pointer, int chunk) {
  Mesh* mp = handle(Mesh_pointer);
  if (!mp->batch) fail("sf2d_batch_grad_prepare: shot batch not prepared (sf2d_batch_prepare)");
  return grad_prepare(mp, Mesh_pointer, chunk, mp->nrec_local, 1);
}

void sf2d_grad_release(long* Mesh_pointer) {
  Mesh* mp = handle(Mesh_pointer);
  GradState* G = mp->grad;
  if (!G) return;
  check_hip(hipStreamSynchronize(mp->stream), "sf2d_grad_release");
  for (float** p : {&G->ckpt, &G->seg, &G->D, &G->Vb, &G->Rb, &G->gk32, &G->gm32, &G->pm32, &G->gvp32, &G->gvs32,
                    &G->res})
    device_free(*p);
  for (double** p : {&G->gk, &G->gm, &G->pm, &G->gvp, &G->gvs}) device_free(*p);
  delete G;
  mp->grad = nullptr;
  mp->grad_batch = 0;
}

// copies the forward fields (d_n, v', r_n) after a step into checkpoint (kind 0) or segment (kind 1) slot
void sf2d_grad_save_state(long* Mesh_pointer, int kind, int slot) {
  Mesh* mp = handle(Mesh_pointer);
  GradState* G = need(mp);
  const int nslot = kind == 0 ? G->nckpt : G->chunk;
  if (slot < 0 || slot >= nslot) fail("sf2d_grad_save_state: slot %d outside 0..%d", slot, nslot - 1);
  float* dst = (kind == 0 ? G->ckpt : G->seg) + (size_t)slot * 3 * G->nfield;
  copy_d2d(dst, G->fd, G->nfield, mp->stream);
  copy_d2d(dst + G->nfield, G->fv, G->nfield, mp->stream);
  copy_d2d(dst + 2 * G->nfield, G->fa, G->nfield, mp->stream);
}

// restores checkpoint slot into the forward fields, with the corrector of that step pending (dto2 = dt/2)
void sf2d_grad_restore_checkpoint(long* Mesh_pointer, int slot, float dto2) {
  Mesh* mp = handle(Mesh_pointer);
  GradState* G = need(mp);
  if (slot < 0 || slot >= G->nckpt) fail("sf2d_grad_restore_checkpoint: slot %d outside 0..%d", slot, G->nckpt - 1);
  const float* src = G->ckpt + (size_t)slot * 3 * G->nfield;
  copy_d2d(G->fd, src, G->nfield, mp->stream);
  copy_d2d(G->fv, src + G->nfield, G->nfield, mp->stream);
  copy_d2d(G->fa, src + 2 * G->nfield, G->nfield, mp->stream);
  if (G->batch) {
    batch_set_pending(mp, dto2);
  } else {
    mp->pending_corrector[0] = 1;
    mp->pending_dto2[0] = dto2;
    mp->accel_partial = 0;
  }
}

// residuals dt (s - o) for all steps, [nstep][nshot][nrec][2]; zeroes the adjoint state and the accumulators
void sf2d_grad_begin_reverse(long* Mesh_pointer, const float* res) {
  Mesh* mp = handle(Mesh_pointer);
  GradState* G = need(mp);
  if (G->nrec > 0)
    check_hip(hipMemcpyAsync(G->res, res, sizeof(float) * 2 * G->nrec * G->nshot * G->nstep, hipMemcpyHostToDevice,
                             mp->stream),
              "residual H2D");
  for (float* p : {G->D, G->Vb, G->Rb}) check_hip(hipMemsetAsync(p, 0, G->nfield * sizeof(float), mp->stream), "zero");
  check_hip(hipMemsetAsync(G->pm32, 0, G->n2 * sizeof(float), mp->stream), "zero");
  for (float* p : {G->gk32, G->gm32}) check_hip(hipMemsetAsync(p, 0, G->ne * sizeof(float), mp->stream), "zero");
  for (double* p : {G->gk, G->gm}) check_hip(hipMemsetAsync(p, 0, G->ne * sizeof(double), mp->stream), "zero");
  check_hip(hipMemsetAsync(G->pm, 0, G->n2 * sizeof(double), mp->stream), "zero");
  const size_t nf = std::max<size_t>(1, G->nf);
  for (float* p : {G->gvp32, G->gvs32}) check_hip(hipMemsetAsync(p, 0, nf * sizeof(float), mp->stream), "zero");
  for (double* p : {G->gvp, G->gvs}) check_hip(hipMemsetAsync(p, 0, nf * sizeof(double), mp->stream), "zero");
  check_hip(hipStreamSynchronize(mp->stream), "sf2d_grad_begin_reverse");
}

// reverse step n (1-based) from the states in segment slot (d_n, v', r_n); first = 1 for n = NSTEP
void sf2d_grad_reverse_step(long* Mesh_pointer, int n, int slot, int first, float dt) {
  Mesh* mp = handle(Mesh_pointer);
  GradState* G = need(mp);
  if (n < 1 || n > G->nstep) fail("sf2d_grad_reverse_step: step %d outside 1..%d", n, G->nstep);
  if (slot < 0 || slot >= G->chunk) fail("sf2d_grad_reverse_step: slot %d outside 0..%d", slot, G->chunk - 1);
  const float* dn = G->seg + (size_t)slot * 3 * G->nfield;
  const float* vpn = dn + G->nfield;
  const float* rn = dn + 2 * G->nfield;
  const float dto2 = (float)(0.5 * (double)dt), h = (float)(0.5 * (double)dt * (double)dt);
  int lshift = 0;
  while ((1 << lshift) < G->ps) lshift++;
  if (G->ps > 1)
    hipLaunchKernelGGL(k_adj_pointwise_minor, dim3(grid_for(G->nfield)), dim3(GT), 0, mp->stream, first,
                       (long)G->nfield, lshift, dt, dto2, h, G->D, G->Vb, G->Rb, mp->d_rmassx, mp->d_rmassz, rn,
                       G->pm32);
  else
    hipLaunchKernelGGL(k_adj_pointwise, dim3(grid_for(G->n2 / 2)), dim3(GT), 0, mp->stream, first, (int)(G->n2 / 2),
                       G->nshot, G->stride, dt, dto2, h, (const float2*)G->D, (float2*)G->Vb, (float2*)G->Rb,
                       mp->d_rmassx, mp->d_rmassz, (const float2*)rn, (float2*)G->pm32);
  if (G->nrec > 0)
    hipLaunchKernelGGL(k_adj_receivers, dim3(G->nrec, G->nshot), dim3(64), 0, mp->stream, G->nrec,
                       G->res + (size_t)(n - 1) * 2 * G->nrec * G->nshot, mp->d_ispec_selected_rec, mp->d_ibool,
                       mp->d_xir, mp->d_gammar, mp->d_cosrot, mp->d_sinrot, G->D, G->ps, G->stride);
  apply_element_forces(mp, 0, G->Rb, G->D, G->nshot, G->stride, 0, G->ps);
  apply_element_forces(mp, 1, G->Rb, G->D, G->nshot, G->stride, 0, G->ps);
  if (mp->stacey && mp->num_abs > 0) {
    launch_stacey(mp, G->Rb, G->Vb, 0, G->nshot, G->stride, G->ps);
    hipLaunchKernelGGL(k_grad_stacey, dim3((mp->num_abs * NGLL + 63) / 64), dim3(64), 0, mp->stream, mp->num_abs,
                       G->nshot, G->ps, G->stride, mp->d_abs_ispec, mp->d_abs_ij, mp->d_abs_normal, mp->d_abs_jacw,
                       mp->d_ibool, (const float2*)vpn, (const float2*)G->Rb, G->gvp32, G->gvs32);
  }
  if (G->ps > 1)
    hipLaunchKernelGGL(k_grad_elastic_minor, dim3((mp->nspec + (64 >> lshift) - 1) / (64 >> lshift)), dim3(64), 0,
                       mp->stream, mp->nspec, G->nshot, lshift, mp->d_ibool, mp->d_geom, (const float2*)dn,
                       (const float2*)G->Rb, G->gk32, G->gm32, mp->basis);
  else
    hipLaunchKernelGGL(k_grad_elastic, dim3((mp->nspec + GEPB - 1) / GEPB), dim3(GT), 0, mp->stream, mp->nspec,
                       G->nshot, G->ps, G->stride, mp->d_ibool, mp->d_xix, mp->d_xiz, mp->d_gammax, mp->d_gammaz,
                       (const float2*)dn, (const float2*)G->Rb, G->gk32, G->gm32, mp->basis);
  check_hip(hipGetLastError(), "sf2d_grad_reverse_step launch");
}

// shot batch: recomputes steps n0+1 .. n1 straight into the segment slots, from checkpoint slot ckpt (the state
// after step n0) or from rest (ckpt = -1, n0 = 0); no state is copied
void sf2d_grad_recompute_segment(long* Mesh_pointer, int ckpt, int n0, int n1, float dt) {
  Mesh* mp = handle(Mesh_pointer);
  GradState* G = need(mp);
  if (!G->batch) fail("sf2d_grad_recompute_segment: the gradient buffers do not follow a shot batch");
  if (n0 < 0 || n1 <= n0 || n1 > G->nstep || n1 - n0 > G->chunk)
    fail("sf2d_grad_recompute_segment: steps %d..%d (segments of at most %d)", n0 + 1, n1, G->chunk);
  if (ckpt < -1 || ckpt >= G->nckpt || (ckpt == -1) != (n0 == 0))
    fail("sf2d_grad_recompute_segment: checkpoint %d for a segment after step %d", ckpt, n0);
  const float dto2 = (float)(0.5 * (double)dt), h = (float)(0.5 * (double)dt * (double)dt);
  const float* src = ckpt < 0 ? nullptr : G->ckpt + (size_t)ckpt * 3 * G->nfield;
  for (int n = n0 + 1; n <= n1; n++) {
    float* dst = G->seg + (size_t)(n - n0 - 1) * 3 * G->nfield;
    batch_step_into(mp, n, src, src ? src + G->nfield : nullptr, src ? src + 2 * G->nfield : nullptr, dto2, dst,
                    dst + G->nfield, dst + 2 * G->nfield, dt, h, dto2);
    src = dst;
  }
}

// adds the FP32 segment accumulators into the FP64 totals
void sf2d_grad_flush(long* Mesh_pointer) {
  Mesh* mp = handle(Mesh_pointer);
  GradState* G = need(mp);
  hipLaunchKernelGGL(k_flush64, dim3(grid_for(G->ne)), dim3(GT), 0, mp->stream, G->ne, G->gk32, G->gk);
  hipLaunchKernelGGL(k_flush64, dim3(grid_for(G->ne)), dim3(GT), 0, mp->stream, G->ne, G->gm32, G->gm);
  hipLaunchKernelGGL(k_flush64, dim3(grid_for(G->n2)), dim3(GT), 0, mp->stream, G->n2, G->pm32, G->pm);
  if (G->nf) {
    hipLaunchKernelGGL(k_flush64, dim3(grid_for(G->nf)), dim3(GT), 0, mp->stream, G->nf, G->gvp32, G->gvp);
    hipLaunchKernelGGL(k_flush64, dim3(grid_for(G->nf)), dim3(GT), 0, mp->stream, G->nf, G->gvs32, G->gvs);
  }
  check_hip(hipGetLastError(), "sf2d_grad_flush launch");
}

// FP64 totals in the caller's numbering: g_kappa, g_mu [nspec][25] (ispec order), d chi/d M [nglob][2],
// Stacey traction parts [nabs][5] (rho vp, rho vs)
void sf2d_grad_result(long* Mesh_pointer, double* gk, double* gm, double* pm, double* gvp, double* gvs) {
  Mesh* mp = handle(Mesh_pointer);
  GradState* G = need(mp);
  check_hip(hipStreamSynchronize(mp->stream), "sf2d_grad_result");
  std::vector<double> t(std::max(G->ne, G->n2));
  auto elem = [&](const double* dev, double* out) {
    check_hip(hipMemcpy(t.data(), dev, G->ne * sizeof(double), hipMemcpyDeviceToHost), "gradient D2H");
    for (int e = 0; e < mp->nspec; e++) {
      const int pos = mp->pos_of_ispec[e];
      for (int p = 0; p < NGLL2; p++) out[(size_t)e * NGLL2 + p] = t[(size_t)pos * NGLL2 + p];
    }
  };
  elem(G->gk, gk);
  elem(G->gm, gm);
  check_hip(hipMemcpy(t.data(), G->pm, G->n2 * sizeof(double), hipMemcpyDeviceToHost), "gradient D2H");
  for (int g = 0; g < mp->nglob; g++) {
    const size_t id = mp->point_id.empty() ? (size_t)g : (size_t)mp->point_id[g];
    pm[2 * g] = t[2 * id];
    pm[2 * g + 1] = t[2 * id + 1];
  }
  if (G->nf) {
    check_hip(hipMemcpy(gvp, G->gvp, G->nf * sizeof(double), hipMemcpyDeviceToHost), "gradient D2H");
    check_hip(hipMemcpy(gvs, G->gvs, G->nf * sizeof(double), hipMemcpyDeviceToHost), "gradient D2H");
  }
}

// replaces the materials of a prepared handle: kappa, mu, rho vp, rho vs [nspec][25] in ispec order and the
// inverse mass rmassx, rmassz [nglob] in the caller's numbering
void sf2d_mesh_set_materials(long* Mesh_pointer, const float* kappav, const float* muv, const float* rho_vp,
                             const float* rho_vs, const float* rmassx, const float* rmassz) {
  Mesh* mp = handle(Mesh_pointer);
  if (!mp->sorted) fail("sf2d_mesh_set_materials: prepare_fields_elastic_device was not called");
  flush_correctors(mp);
  check_hip(hipStreamSynchronize(mp->stream), "sf2d_mesh_set_materials");
  const size_t ne = (size_t)mp->nspec * NGLL2;
  auto put = [&](float* dev, const float* host) {
    if (!dev || !host) return;
    std::vector<float> v = to_device_order(mp, host);
    check_hip(hipMemcpy(dev, v.data(), ne * sizeof(float), hipMemcpyHostToDevice), "materials H2D");
  };
  put(mp->d_kappav, kappav);
  put(mp->d_muv, muv);
  if (mp->d_moduli && (kappav || muv)) {
    launch_pack_elements(mp);
    check_hip(hipStreamSynchronize(mp->stream), "pack elements");
  }
  put(mp->d_rho_vp, rho_vp);
  put(mp->d_rho_vs, rho_vs);
  if (rmassx) points_to_device(mp, mp->d_rmassx, rmassx, 1, "rmassx H2D");
  if (rmassz) points_to_device(mp, mp->d_rmassz, rmassz, 1, "rmassz H2D");
  if (rmassx || rmassz) mp->same_mass = 0;  // the Newmark kernels read both arrays from here on
}

}  // extern "C"
