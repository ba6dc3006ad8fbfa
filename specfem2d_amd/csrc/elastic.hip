// Elastic time-stepping kernels for gfx950 (MI355X).
//
// The element force kernel is the hot spot (reference: src/gpu/kernels/Kernel_2_viscoelastic_impl.cu:294-611,
// CPU: src/specfem2D/compute_forces_viscoelastic.F90:34-885).  It is HBM-bound (~2.9 flop/B), so the design is
// about bytes and latency hiding, not math:
//   * flat GLL-point mapping: a 256-thread workgroup holds 10 elements (250 active lanes), so every per-element
//     array is read with contiguous, fully coalesced loads (no 25->32 padding traffic);
//   * default assembly by coloured patches: a workgroup gathers a patch's points into LDS once, sums the element
//     contributions there and writes each point back once; patches of one launch share no point, so no global
//     atomics.  Points are renumbered so that a patch's own points form one contiguous range;
//   * the patch kernel is persistent and software-pipelined: the loads of the next patch are in flight while
//     the current one is computed;
//   * workgroups are remapped so that each XCD walks a contiguous range: neighbouring patches share an L2.
// Coloured-element (plain read-modify-write per colour) and FP32-atomic variants remain selectable.
// All kernels run on the handle's stream.  Every launch bounds its indices against the array sizes it was given.
#include "sf2d_mesh.h"

namespace sf2d {

namespace {

constexpr int FORCE_THREADS = 256;
constexpr int FORCE_EPB = 10;  // elements per block: 250 of 256 lanes active
constexpr int NXCD = 8;        // gfx950: 8 accelerator complex dies, workgroups dispatched round-robin

static_assert(Mesh::PATCH_ELEMS == FORCE_EPB, "one workgroup holds one patch");
static_assert(FORCE_EPB * NGLL2 <= FORCE_THREADS && FORCE_THREADS <= 256, "patch points indexed by 8-bit slots");
static_assert(Mesh::PATCH_SHARE == 4, "force cells are summed as two float4 per point");
static_assert(Mesh::PATCH_ELEMS <= 15 && Mesh::CHAIN_ELEMS <= 15, "the element count of a patch descriptor has 4 bits");
static_assert(Mesh::CHAIN_ELEMS * NGLL <= 64 && (4 * Mesh::CHAIN_ELEMS + 1) * NGLL <= 256,
              "a chain patch fills one wave by rows and has 8-bit point slots");

__device__ __forceinline__ void atomic_add_f32(float* p, float v) { unsafeAtomicAdd(p, v); }

// ---------------------------------------------------------------------------
// Newmark predictor  (src/gpu/kernels/UpdateDispVeloc_kernel.cu:37-66; update_displacement_Newmark.F90:408-454)
__global__ __launch_bounds__(256) void k_predictor(float2* __restrict__ d, float2* __restrict__ v,
                                                   float2* __restrict__ a, int n, float dt, float dt2o2, float dto2,
                                                   long shot_stride, int zero_accel) {
  d += blockIdx.y * shot_stride;
  v += blockIdx.y * shot_stride;
  a += blockIdx.y * shot_stride;
  int stride = gridDim.x * blockDim.x;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) {
    float2 dd = d[k], vv = v[k], aa = a[k];
    dd.x = dd.x + dt * vv.x + dt2o2 * aa.x;
    dd.y = dd.y + dt * vv.y + dt2o2 * aa.y;
    vv.x = vv.x + dto2 * aa.x;
    vv.y = vv.y + dto2 * aa.y;
    d[k] = dd;
    v[k] = vv;
    if (zero_accel) a[k] = make_float2(0.f, 0.f);
  }
}

// Newmark corrector with the inverse mass matrix (src/gpu/kernels/kernel_3_cuda_device.cu:37-65)
__global__ __launch_bounds__(256) void k_corrector(float2* __restrict__ v, float2* __restrict__ a,
                                                   const float* __restrict__ rmx, const float* __restrict__ rmz,
                                                   int n, float dto2, int apply_mass, long shot_stride, int ps) {
  v += blockIdx.y * shot_stride;
  a += blockIdx.y * shot_stride;
  int stride = gridDim.x * blockDim.x;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) {
    float2 aa = a[k], vv = v[k];
    if (apply_mass) {
      const int g = ps == 1 ? k : k / ps;
      const float wx = rmx[g];
      aa.x = aa.x * wx;
      aa.y = aa.y * (rmz ? rmz[g] : wx);  // rmz null: one inverse mass for both components
      a[k] = aa;
    }
    vv.x = vv.x + dto2 * aa.x;
    vv.y = vv.y + dto2 * aa.y;
    v[k] = vv;
  }
}

// Corrector of the previous step fused with the predictor of this one: the same per-point operations as
// k_corrector followed by k_predictor, in one pass (56 B per point instead of 88).
__global__ __launch_bounds__(256) void k_corrector_predictor(float2* __restrict__ d, float2* __restrict__ v,
                                                             float2* __restrict__ a, const float* __restrict__ rmx,
                                                             const float* __restrict__ rmz, int n, float dto2_prev,
                                                             float dt, float dt2o2, float dto2, long shot_stride,
                                                             int zero_accel, int ps) {
  d += blockIdx.y * shot_stride;
  v += blockIdx.y * shot_stride;
  a += blockIdx.y * shot_stride;
  int stride = gridDim.x * blockDim.x;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) {
    float2 dd = d[k], vv = v[k], aa = a[k];
    const int g = ps == 1 ? k : k / ps;
    const float wx = rmx[g];
    aa.x = aa.x * wx;
    aa.y = aa.y * (rmz ? rmz[g] : wx);  // rmz null: one inverse mass for both components
    vv.x = vv.x + dto2_prev * aa.x;
    vv.y = vv.y + dto2_prev * aa.y;
    dd.x = dd.x + dt * vv.x + dt2o2 * aa.x;
    dd.y = dd.y + dt * vv.y + dt2o2 * aa.y;
    vv.x = vv.x + dto2 * aa.x;
    vv.y = vv.y + dto2 * aa.y;
    d[k] = dd;
    v[k] = vv;
    if (zero_accel) a[k] = make_float2(0.f, 0.f);
  }
}

// The same update out of place: state (d, v, a) with the corrector pending -> predicted state (dd, dv, 0), for
// recomputations that keep every step's state.
__global__ __launch_bounds__(256) void k_corrector_predictor_to(
    const float2* __restrict__ d, const float2* __restrict__ v, const float2* __restrict__ a,
    float2* __restrict__ dd_out, float2* __restrict__ dv_out, float2* __restrict__ da_out,
    const float* __restrict__ rmx, const float* __restrict__ rmz, int n, float dto2_prev, float dt, float dt2o2,
    float dto2, long shot_stride, int zero_accel, int ps) {
  const long off = blockIdx.y * shot_stride;
  d += off;
  v += off;
  a += off;
  dd_out += off;
  dv_out += off;
  da_out += off;
  int stride = gridDim.x * blockDim.x;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) {
    float2 dd = d[k], vv = v[k], aa = a[k];
    const int g = ps == 1 ? k : k / ps;
    const float wx = rmx[g];
    aa.x = aa.x * wx;
    aa.y = aa.y * (rmz ? rmz[g] : wx);  // rmz null: one inverse mass for both components
    vv.x = vv.x + dto2_prev * aa.x;
    vv.y = vv.y + dto2_prev * aa.y;
    dd.x = dd.x + dt * vv.x + dt2o2 * aa.x;
    dd.y = dd.y + dt * vv.y + dt2o2 * aa.y;
    vv.x = vv.x + dto2 * aa.x;
    vv.y = vv.y + dto2 * aa.y;
    dd_out[k] = dd;
    dv_out[k] = vv;
    if (zero_accel) da_out[k] = make_float2(0.f, 0.f);
  }
}

// accel = 0 at a list of points (first-touch assembly: the points that boundary terms and sources also update)
// (shot-minor batches: n = points x ps entries, point k / ps, shot k % ps)
__global__ __launch_bounds__(256) void k_zero_points(float2* __restrict__ a, const int* __restrict__ pts, int n,
                                                     long shot_stride, int ps) {
  a += blockIdx.y * shot_stride;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x)
    a[ps == 1 ? (long)pts[k] : (long)pts[k / ps] * ps + k % ps] = make_float2(0.f, 0.f);
}

// ---------------------------------------------------------------------------
// isotropic elastic element forces, P-SV, for the elements at positions [first, first + nelem).
// ATOMIC = false: the range is one colour (no two elements share a point), so the assembly is a plain
// read-modify-write; ATOMIC = true: FP32 atomics, valid for any range.
// ATT: viscoelastic attenuation with 3 standard linear solids -- the memory variables advance with the
// recursive convolution of compute_attenuation_viscoelastic.f90:93-121 (Kernel_2_viscoelastic_impl.cu:1019-1065
// on the reference's GPU) and correct the stress with the unrelaxed moduli; each point reads its coefficient set
// {A_kappa, B_kappa, A_mu, B_mu} x 3 solids from a small table through a per-point index; the state (memory
// variables e1, e11, e13 per solid, previous dux_dxl, duz_dzl, dux_dzl + duz_dxl) is three float4 planes of
// [3][ne] in device element order (att_load)
// One point's viscoelastic update (3 standard linear solids): the memory variables ev (e1, e11, e13 per solid)
// advance by the recursive convolution of compute_attenuation_viscoelastic.f90:93-121 with the coefficient set
// cf[k] = {A_kappa, B_kappa, A_mu, B_mu}, the previous strains ov (dux_dxl, duz_dzl, dux_dzl + duz_dxl) become
// this step's, and the stress takes the unrelaxed-moduli correction.
__device__ __forceinline__ void attenuate_point(const float4 (&cf)[3], float (&ev)[9], float (&ov)[3], float kap,
                                                float mu, float dux_dxl, float dux_dzl, float duz_dxl, float duz_dzl,
                                                float& sxx, float& sxz, float& szz) {
  const float th = dux_dxl + duz_dzl;
  const float oxx = ov[0], ozz = ov[1], oxz = ov[2];
  const float th_old = oxx + ozz;
  float e1s = 0.f, e11s = 0.f, e13s = 0.f;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const float a1 = cf[k].x, b1 = cf[k].y, a2 = cf[k].z, b2 = cf[k].w;
    const float e1 = a1 * a1 * ev[3 * k] + b1 * (th + a1 * th_old);
    const float e11 = a2 * a2 * ev[3 * k + 1] + b2 * (dux_dxl - 0.5f * th + a2 * (oxx - 0.5f * th_old));
    const float e13 = a2 * a2 * ev[3 * k + 2] + b2 * (dux_dzl + duz_dxl + a2 * oxz);
    ev[3 * k] = e1;
    ev[3 * k + 1] = e11;
    ev[3 * k + 2] = e13;
    e1s += e1;
    e11s += e11;
    e13s += e13;
  }
  ov[0] = dux_dxl;
  ov[1] = duz_dzl;
  ov[2] = dux_dzl + duz_dxl;
  sxx += kap * e1s + 2.f * mu * e11s;
  sxz += mu * e13s;
  szz += kap * e1s - 2.f * mu * e11s;
}

// A point's attenuation state as three float4 planes of [3][ne] (ne = element points): the 9 memory variables then
// the 3 previous strains, {e1, e11, e13} of solid 1, 2, 3 and {dux_dxl, duz_dzl, dux_dzl + duz_dxl}
__device__ __forceinline__ void att_load(const float* st, long ne, size_t l, float (&ev)[9], float (&ov)[3]) {
  const float4* s4 = reinterpret_cast<const float4*>(st);
  const float4 a = s4[l], b = s4[(size_t)ne + l], c = s4[2 * (size_t)ne + l];
  ev[0] = a.x; ev[1] = a.y; ev[2] = a.z; ev[3] = a.w;
  ev[4] = b.x; ev[5] = b.y; ev[6] = b.z; ev[7] = b.w;
  ev[8] = c.x; ov[0] = c.y; ov[1] = c.z; ov[2] = c.w;
}
__device__ __forceinline__ void att_store(float* st, long ne, size_t l, const float (&ev)[9], const float (&ov)[3]) {
  float4* s4 = reinterpret_cast<float4*>(st);
  s4[l] = make_float4(ev[0], ev[1], ev[2], ev[3]);
  s4[(size_t)ne + l] = make_float4(ev[4], ev[5], ev[6], ev[7]);
  s4[2 * (size_t)ne + l] = make_float4(ev[8], ov[0], ov[1], ov[2]);
}

// RECT: every element an axis-aligned rectangle (xi_z = gamma_x = 0, xi_x and gamma_z constant over the element):
// one float2 {xi_x, gamma_z} per element (geom_rect, device element order) instead of four per-point arrays
template <bool ATOMIC, bool ATT, bool RECT>
__global__ __launch_bounds__(FORCE_THREADS) void k_elastic_forces(
    int first, int nelem, int nbatch, int per_xcd, const int* __restrict__ ibool,
    const float* __restrict__ xix, const float* __restrict__ xiz, const float* __restrict__ gammax,
    const float* __restrict__ gammaz, const float* __restrict__ kappav, const float* __restrict__ muv,
    const float2* __restrict__ displ, float* __restrict__ accel, const Basis basis,
    const unsigned int* __restrict__ att_idx, const float4* __restrict__ att_table, float* __restrict__ att_e,
    float* __restrict__ att_old, long ne, const float2* __restrict__ geom_rect, const float4* __restrict__ aniso) {
  // XCD-aware remap: blocks b, b+8, b+16, ... (one XCD) take consecutive batches
  const int bid = blockIdx.x;
  const int batch = (bid % NXCD) * per_xcd + bid / NXCD;
  if (batch >= nbatch) return;

  __shared__ float s_hp[NGLL2], s_hpw[NGLL2], s_w[NGLL];
  __shared__ float s_ux[FORCE_THREADS], s_uz[FORCE_THREADS];
  __shared__ float s_t1x[FORCE_THREADS], s_t1z[FORCE_THREADS], s_t2x[FORCE_THREADS], s_t2z[FORCE_THREADS];

  const int tid = threadIdx.x;
  if (tid < NGLL2) {
    s_hp[tid] = basis.hp[tid];
    s_hpw[tid] = basis.hpw[tid];
  }
  if (tid < NGLL) s_w[tid] = basis.w[tid];

  const int le = tid / NGLL2;
  const int p = tid - le * NGLL2;
  const int i = p % NGLL;
  const int j = p / NGLL;
  const int e = batch * FORCE_EPB + le;
  const bool active = (tid < FORCE_EPB * NGLL2) && (e < nelem);

  int g = 0;
  float xixl = 0.f, xizl = 0.f, gaxl = 0.f, gazl = 0.f, kap = 0.f, mu = 0.f;
  // attenuation state of this point, loaded with the element's other inputs so the loads overlap
  float4 cf[3];
  float ev[9], ov[3];
  float2 a_old = make_float2(0.f, 0.f);  // coloured launches: no other element of the launch has the point
  float4 ca = make_float4(0.f, 0.f, 0.f, 0.f), cb = ca;  // anisotropic stiffness of the point
  if (active) {
    if (aniso) {
      ca = aniso[2 * ((size_t)(first + e) * NGLL2 + p)];
      cb = aniso[2 * ((size_t)(first + e) * NGLL2 + p) + 1];
    }
    if (!ATOMIC) a_old = reinterpret_cast<const float2*>(accel)[ibool[(size_t)(first + e) * NGLL2 + p]];
    const size_t l = (size_t)(first + e) * NGLL2 + p;
    if (ATT) {
      const float4* coef = att_table + 3 * (size_t)att_idx[l];
#pragma unroll
      for (int k = 0; k < 3; k++) cf[k] = coef[k];
      att_load(att_e, ne, l, ev, ov);
    }
    g = ibool[l];
    const float2 u = displ[g];
    s_ux[tid] = u.x;
    s_uz[tid] = u.y;
    if (RECT) {
      const float2 r = geom_rect[first + e];
      xixl = r.x;
      gazl = r.y;
    } else {
      xixl = xix[l];
      xizl = xiz[l];
      gaxl = gammax[l];
      gazl = gammaz[l];
    }
    kap = kappav[l];
    mu = muv[l];
  }
  __syncthreads();

  const int base = le * NGLL2;
  if (active) {
    float dux_dxi = 0.f, duz_dxi = 0.f, dux_dga = 0.f, duz_dga = 0.f;
#pragma unroll
    for (int k = 0; k < NGLL; k++) {
      const float hi = s_hp[i * NGLL + k];
      const float hj = s_hp[j * NGLL + k];
      dux_dxi += s_ux[base + j * NGLL + k] * hi;
      duz_dxi += s_uz[base + j * NGLL + k] * hi;
      dux_dga += s_ux[base + k * NGLL + i] * hj;
      duz_dga += s_uz[base + k * NGLL + i] * hj;
    }
    const float dux_dxl = dux_dxi * xixl + dux_dga * gaxl;
    const float dux_dzl = dux_dxi * xizl + dux_dga * gazl;
    const float duz_dxl = duz_dxi * xixl + duz_dga * gaxl;
    const float duz_dzl = duz_dxi * xizl + duz_dga * gazl;
    const float jac = 1.f / (xixl * gazl - gaxl * xizl);
    const float lambdal = kap - mu;
    const float lp2m = kap + mu;
    float sxx = lp2m * dux_dxl + lambdal * duz_dzl;
    float szz = lp2m * duz_dzl + lambdal * dux_dxl;
    float sxz = mu * (duz_dxl + dux_dzl);
    if (aniso) {  // compute_forces_viscoelastic.F90:372-447 (P-SV)
      const float sh = duz_dxl + dux_dzl;
      sxx = ca.x * dux_dxl + ca.y * duz_dzl + ca.z * sh;
      szz = ca.y * dux_dxl + ca.w * duz_dzl + cb.x * sh;
      sxz = ca.z * dux_dxl + cb.x * duz_dzl + cb.y * sh;
    }
    if (ATT) {
      const size_t l = (size_t)(first + e) * NGLL2 + p;
      attenuate_point(cf, ev, ov, kap, mu, dux_dxl, dux_dzl, duz_dxl, duz_dzl, sxx, sxz, szz);
      att_store(att_e, ne, l, ev, ov);
    }
    s_t1x[tid] = jac * (sxx * xixl + sxz * xizl);
    s_t1z[tid] = jac * (sxz * xixl + szz * xizl);
    s_t2x[tid] = jac * (sxx * gaxl + sxz * gazl);
    s_t2z[tid] = jac * (sxz * gaxl + szz * gazl);
  }
  __syncthreads();

  if (active) {
    float t1x = 0.f, t1z = 0.f, t2x = 0.f, t2z = 0.f;
#pragma unroll
    for (int k = 0; k < NGLL; k++) {
      const float wi = s_hpw[k * NGLL + i];
      const float wj = s_hpw[k * NGLL + j];
      t1x += s_t1x[base + j * NGLL + k] * wi;
      t1z += s_t1z[base + j * NGLL + k] * wi;
      t2x += s_t2x[base + k * NGLL + i] * wj;
      t2z += s_t2z[base + k * NGLL + i] * wj;
    }
    const float fx = -(s_w[j] * t1x + s_w[i] * t2x);
    const float fz = -(s_w[j] * t1z + s_w[i] * t2z);
    if (ATOMIC) {
      atomic_add_f32(&accel[2 * (size_t)g], fx);
      atomic_add_f32(&accel[2 * (size_t)g + 1], fz);
    } else {
      float2 a = a_old;
      a.x += fx;
      a.y += fz;
      reinterpret_cast<float2*>(accel)[g] = a;
    }
  }
}

// Coloured-patch variant, persistent and software-pipelined.  A patch is <= FORCE_EPB connected elements
// whose points (nown owned, a contiguous id range, then nhalo halo ids) are gathered once into LDS.  Every
// element point writes its force into its own LDS cell (point slot, rank among the element points sharing
// that point), and each patch point sums its cells and is written back with one read-modify-write (patches of
// one launch share no point): no atomics, and the sum is deterministic.  Each workgroup walks every W-th
// patch; while it computes one, the loads of the next two patches are in flight (three register sets rotate,
// so no copy forces an early wait).
struct PatchRegs {
  float4 geo;   // xix, xiz, gammax, gammaz of this lane's element point
  float2 mod;   // kappa, mu
  int slot;     // point slot | rank << 8
  int g;        // point id loaded by this lane (lanes < nown + nhalo)
  float2 u, a;
};

// load through a 32-bit byte offset from a uniform base (global_load with saddr + voffset); every array indexed
// this way is smaller than 4 GiB (checked when the handle is prepared)
template <typename T>
__device__ __forceinline__ T ld_off(const T* base, unsigned byte_off) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + byte_off);
}

// counts word of a patch descriptor: ne (bits 0-3), affine flag (bit 7), nown (bits 8-19), nhalo (bits 20-31)
__device__ __forceinline__ int patch_ne(int c) { return c & 0xf; }
__device__ __forceinline__ bool patch_affine(int c) { return (c & 0x80) != 0; }
__device__ __forceinline__ int patch_nown(int c) { return (c >> 8) & 0xfff; }
__device__ __forceinline__ int patch_nhalo(int c) { return (c >> 20) & 0xfff; }

// element-local LDS rows: 5 float2 padded to 6 (48 B), so that a row or column is read with two 16-byte and
// one 8-byte access and consecutive rows fall on different banks
constexpr int ROWF2 = 6;

__global__ __launch_bounds__(FORCE_THREADS) void k_elastic_forces_patch(
    int first_patch, int npatch, int nshot, long shot_stride, const int4* __restrict__ desc,
    const int* __restrict__ halo, const uint16_t* __restrict__ elem_slot, const float4* __restrict__ geom,
    const float4* __restrict__ geom_elem, const float2* __restrict__ moduli, const float2* __restrict__ displ,
    float2* __restrict__ accel, const Basis basis) {
  // Work items are (patch, shot) pairs, item = patch * nshot + shot: the shots of one patch are consecutive
  // items, so their element data is fetched from HBM once and then hit in L2.  Shot s's fields start
  // s * shot_stride points into displ and accel.  Workgroup w takes items w, w + W, w + 2W, ... (W = gridDim.x):
  // at any time the grid works on one window of consecutive items, so the element arrays stream through DRAM
  // in order.  Consecutive w run on one XCD (blocks are dispatched round-robin over the XCDs; gridDim.x is a
  // multiple of NXCD).
  const int bid = blockIdx.x;
  const int nw = gridDim.x;
  const int w = (bid % NXCD) * (nw / NXCD) + bid / NXCD;
  const int nitem = npatch * nshot;
  if (w >= nitem) return;
  const int qlast = w + (nitem - 1 - w) / nw * nw;  // last item of this workgroup

  __shared__ float s_hp[NGLL2], s_hpw[NGLL2], s_w[NGLL];
  __shared__ float2 s_u[FORCE_THREADS];                          // displacement of the patch's points
  __shared__ float4 s_f[FORCE_THREADS * Mesh::PATCH_SHARE / 2];  // per point: PATCH_SHARE force cells (float2)
  __shared__ float4 s_ur[FORCE_EPB * NGLL * ROWF2 / 2], s_uc[FORCE_EPB * NGLL * ROWF2 / 2];   // u by rows, columns
  __shared__ float4 s_t1r[FORCE_EPB * NGLL * ROWF2 / 2], s_t2c[FORCE_EPB * NGLL * ROWF2 / 2]; // fluxes

  const int tid = threadIdx.x;
  if (tid < NGLL2) {
    s_hp[tid] = basis.hp[tid];
    s_hpw[tid] = basis.hpw[tid];
  }
  if (tid < NGLL) s_w[tid] = basis.w[tid];
  const int le = tid / NGLL2;
  const int p = tid - le * NGLL2;
  const int i = p % NGLL;
  const int j = p / NGLL;
  const int erow = min(le, FORCE_EPB - 1) * NGLL;  // first row (and column) of this lane's element

  // Loads are unconditional with clamped indices (lanes past the patch re-read a valid entry, the patch after
  // the last one is the last one again), so nothing waits on a branch; only stores are predicated.
  // affine patches (every element's geometry constant over its 25 points) read one float4 per element
  auto load_elements = [&](const int4 d, PatchRegs& r) {
    const unsigned e = (unsigned)(d.x + min(le, patch_ne(d.w) - 1));
    const unsigned l = e * NGLL2 + p;
    r.slot = ld_off(elem_slot, 2u * l);
    r.geo = patch_affine(d.w) ? ld_off(geom_elem, 16u * e) : ld_off(geom, 16u * l);
    r.mod = ld_off(moduli, 8u * l);
  };
  auto load_halo_id = [&](const int4 d) -> int {
    const int h = min(max(tid - patch_nown(d.w), 0), max(patch_nhalo(d.w) - 1, 0));
    return ld_off(halo, 4u * (unsigned)(d.z + h));
  };
  auto item_desc = [&](int item) { return desc[first_patch + item / nshot]; };
  auto load_points = [&](const int4 d, int halo_id, PatchRegs& r, int item) {
    const long base = (long)(item % nshot) * shot_stride;  // uniform
    r.g = tid < patch_nown(d.w) ? d.y + tid : halo_id;
    r.u = ld_off(displ + base, 8u * (unsigned)r.g);
    r.a = ld_off((const float2*)accel + base, 8u * (unsigned)r.g);
  };

  int q = w;
  // three register sets rotate through (compute, loaded, loading): no copies, so each compute step waits only
  // for its own loads, issued two steps earlier
  PatchRegs ra, rb, rc;
  int4 da = item_desc(q);
  load_elements(da, ra);
  load_points(da, load_halo_id(da), ra, q);
  int4 db = item_desc(min(q + nw, qlast));
  load_elements(db, rb);
  load_points(db, load_halo_id(db), rb, min(q + nw, qlast));
  int4 dc = item_desc(min(q + 2 * nw, qlast));
  int hc = load_halo_id(dc);
  int ha = 0, hb = 0;
  __syncthreads();
  // the basis stays in LDS (broadcast reads): in registers it would cost 22 VGPRs and a wave per SIMD
#define hpi(k) s_hp[i * NGLL + (k)]
#define hpj(k) s_hp[j * NGLL + (k)]
#define hwi(k) s_hpw[(k) * NGLL + i]
#define hwj(k) s_hpw[(k) * NGLL + j]
  float2* const ur2 = reinterpret_cast<float2*>(s_ur);
  float2* const uc2 = reinterpret_cast<float2*>(s_uc);
  float2* const t1r2 = reinterpret_cast<float2*>(s_t1r);
  float2* const t2c2 = reinterpret_cast<float2*>(s_t2c);
  float2* const f2 = reinterpret_cast<float2*>(s_f);
  const int rrow = (erow + j) * ROWF2, rcol = (erow + i) * ROWF2;  // float2 index of this lane's row, column

  auto compute = [&](const PatchRegs& C, const int4 DC) {
    const int npts = patch_nown(DC.w) + patch_nhalo(DC.w);
    const bool active = le < patch_ne(DC.w);
    if (tid < npts) {
      s_u[tid] = C.u;
      s_f[2 * tid] = make_float4(0.f, 0.f, 0.f, 0.f);
      s_f[2 * tid + 1] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    __syncthreads();
    if (active) {
      const float2 u = s_u[C.slot & 0xff];
      ur2[rrow + i] = u;
      uc2[rcol + j] = u;
    }
    __syncthreads();
    if (active) {
      const float4 r01 = s_ur[rrow / 2], r23 = s_ur[rrow / 2 + 1];
      const float2 r4 = ur2[rrow + 4];
      const float4 c01 = s_uc[rcol / 2], c23 = s_uc[rcol / 2 + 1];
      const float2 c4 = uc2[rcol + 4];
      const float dux_dxi = r01.x * hpi(0) + r01.z * hpi(1) + r23.x * hpi(2) + r23.z * hpi(3) + r4.x * hpi(4);
      const float duz_dxi = r01.y * hpi(0) + r01.w * hpi(1) + r23.y * hpi(2) + r23.w * hpi(3) + r4.y * hpi(4);
      const float dux_dga = c01.x * hpj(0) + c01.z * hpj(1) + c23.x * hpj(2) + c23.z * hpj(3) + c4.x * hpj(4);
      const float duz_dga = c01.y * hpj(0) + c01.w * hpj(1) + c23.y * hpj(2) + c23.w * hpj(3) + c4.y * hpj(4);
      const float xixl = C.geo.x, xizl = C.geo.y, gaxl = C.geo.z, gazl = C.geo.w;
      const float dux_dxl = dux_dxi * xixl + dux_dga * gaxl;
      const float dux_dzl = dux_dxi * xizl + dux_dga * gazl;
      const float duz_dxl = duz_dxi * xixl + duz_dga * gaxl;
      const float duz_dzl = duz_dxi * xizl + duz_dga * gazl;
      const float jac = 1.f / (xixl * gazl - gaxl * xizl);
      const float lambdal = C.mod.x - C.mod.y;
      const float lp2m = C.mod.x + C.mod.y;
      const float sxx = lp2m * dux_dxl + lambdal * duz_dzl;
      const float szz = lp2m * duz_dzl + lambdal * dux_dxl;
      const float sxz = C.mod.y * (duz_dxl + dux_dzl);
      t1r2[rrow + i] = make_float2(jac * (sxx * xixl + sxz * xizl), jac * (sxz * xixl + szz * xizl));
      t2c2[rcol + j] = make_float2(jac * (sxx * gaxl + sxz * gazl), jac * (sxz * gaxl + szz * gazl));
    }
    __syncthreads();
    if (active) {
      const float4 a01 = s_t1r[rrow / 2], a23 = s_t1r[rrow / 2 + 1];
      const float2 a4 = t1r2[rrow + 4];
      const float4 b01 = s_t2c[rcol / 2], b23 = s_t2c[rcol / 2 + 1];
      const float2 b4 = t2c2[rcol + 4];
      const float t1x = a01.x * hwi(0) + a01.z * hwi(1) + a23.x * hwi(2) + a23.z * hwi(3) + a4.x * hwi(4);
      const float t1z = a01.y * hwi(0) + a01.w * hwi(1) + a23.y * hwi(2) + a23.w * hwi(3) + a4.y * hwi(4);
      const float t2x = b01.x * hwj(0) + b01.z * hwj(1) + b23.x * hwj(2) + b23.z * hwj(3) + b4.x * hwj(4);
      const float t2z = b01.y * hwj(0) + b01.w * hwj(1) + b23.y * hwj(2) + b23.w * hwj(3) + b4.y * hwj(4);
      f2[(C.slot & 0xff) * Mesh::PATCH_SHARE + (C.slot >> 8)] =
          make_float2(-(s_w[j] * t1x + s_w[i] * t2x), -(s_w[j] * t1z + s_w[i] * t2z));
    }
    __syncthreads();
    if (tid < npts) {
      const float4 f01 = s_f[2 * tid], f23 = s_f[2 * tid + 1];
      float2 a = C.a;
      a.x += ((f01.x + f01.z) + f23.x) + f23.z;
      a.y += ((f01.y + f01.w) + f23.y) + f23.w;
      *reinterpret_cast<float2*>(reinterpret_cast<char*>(accel + (long)(q % nshot) * shot_stride) +
                                 8u * (unsigned)C.g) = a;
    }
  };
#define SF2D_PATCH_STEP(RC, DCUR, RL, DL, HL, DNEXT, HNEXT)                 \
  {                                                                         \
    const int4 dn_ = item_desc(min(q + 3 * nw, qlast));                     \
    const int hn_ = load_halo_id(dn_);                                      \
    load_elements(DL, RL);                                                  \
    load_points(DL, HL, RL, min(q + 2 * nw, qlast));                        \
    compute(RC, DCUR);                                                      \
    DNEXT = dn_;                                                            \
    HNEXT = hn_;                                                            \
    q += nw;                                                                \
    if (q > qlast) break;                                                   \
  }
  for (;;) {
    SF2D_PATCH_STEP(ra, da, rc, dc, hc, da, ha)
    SF2D_PATCH_STEP(rb, db, ra, da, ha, db, hb)
    SF2D_PATCH_STEP(rc, dc, rb, db, hb, dc, hc)
  }
#undef SF2D_PATCH_STEP
#undef hpi
#undef hpj
#undef hwi
#undef hwj
}

// Chain patches: the patch's elements form a row in which element k's right edge (i = 4) is element k+1's left
// edge (i = 0) and no other point is shared inside the patch (the structured meshes of the internal mesher,
// checked on the host when the plan is built).  One wave per (patch, shot) item; lane 5 k + j holds row j of
// element k (its five points (i, j), i = 0..4) in registers:
//   * xi-derivatives and the xi-flux contraction run along the lane's own row, in registers;
//   * the gamma direction goes through two LDS transposes per operand: the lane writes its row, reads column j,
//     contracts along it and writes the result back by columns, then reads its row again;
//   * the force at the shared edge point (4, j) moves one element to the right with a cross-lane shuffle and is
//     added at (0, j), so every point of the patch is updated by exactly one lane: no force cells, no atomics.
// The arithmetic per point is the element operator of the other kernels (Kernel_2_noatt_iso_impl:
// compute_forces_viscoelastic_cuda.cu:163-215, :501-565).
//
// FIRST (first-touch assembly): a point id carrying Mesh::FIRST_TOUCH belongs to the first colour that updates the
// point in the step, and that colour writes the force instead of adding it to accel, so the predictor does not
// have to clear accel and the first update reads nothing.
constexpr int CHAIN_LANES = 64;
constexpr int CHAIN_SHOTS_PER_WAVE = 8;  // shots per wave in batched launches unless the handle is tuned
//
// Shots: an item is a patch and a group of up to `group` consecutive shots; the wave loads the patch's element
// data once and runs the shots of its group one after the other.
template <bool FIRST, bool MULTI, bool RECT>
__global__ __launch_bounds__(CHAIN_LANES) void k_elastic_forces_chain(
    int first_patch, int npatch, int nshot, int group, long shot_stride, const int4* __restrict__ desc,
    const int* __restrict__ ids, const float4* __restrict__ geom, const float4* __restrict__ geom_elem,
    const float2* __restrict__ moduli, const float2* __restrict__ displ, float2* __restrict__ accel,
    const Basis basis) {
  // [element][row][column], rows padded to 6; element CHAIN_ELEMS is scratch for lanes past the patch
  __shared__ float2 s_t[Mesh::CHAIN_ELEMS + 1][NGLL][NGLL + 1];
  __shared__ float2 s_v[Mesh::CHAIN_ELEMS + 1][NGLL][NGLL + 1];
  __shared__ float s_hp[NGLL2], s_hpw[NGLL2];
  const int lane = threadIdx.x;
  if (lane < NGLL2) {
    s_hp[lane] = basis.hp[lane];
    s_hpw[lane] = basis.hpw[lane];
  }
  // items in XCD-contiguous blocks: the shots of a patch run on one XCD and share its L2
  const int nw = gridDim.x;
  const int item = (blockIdx.x % NXCD) * (nw / NXCD) + blockIdx.x / NXCD;
  if (!MULTI) group = 1;  // one shot per wave: the element data need not stay live across shots
  const int ngroup = (nshot + group - 1) / group;
  const int nitem = npatch * ngroup;
  if (item >= nitem) return;  // uniform over the wave
  const int4 d = desc[first_patch + item / ngroup];
  const int s0 = (item % ngroup) * group, s1 = min(nshot, s0 + group);
  const int ne = patch_ne(d.w);
  const int k = min(lane / NGLL, Mesh::CHAIN_ELEMS);  // element of the patch (CHAIN_ELEMS: lanes past it)
  const int j = lane % NGLL;
  const bool active = lane / NGLL < ne;
  const int e = d.x + min(k, ne - 1);
  const unsigned l0 = (unsigned)e * NGLL2 + j * NGLL;  // first point of this lane's row
  int g[NGLL];
  bool first[NGLL];
  float2 u[NGLL], mod[NGLL];
  float4 geo[NGLL];
#pragma unroll
  for (int i = 0; i < NGLL; i++) {
    const int id = ld_off(ids, 4u * (l0 + i));
    g[i] = id & ~Mesh::FIRST_TOUCH;
    first[i] = FIRST && (id & Mesh::FIRST_TOUCH);
    mod[i] = ld_off(moduli, 8u * (l0 + i));
  }
  if (RECT || patch_affine(d.w)) {
    const float4 ge = ld_off(geom_elem, 16u * (unsigned)e);
#pragma unroll
    for (int i = 0; i < NGLL; i++) geo[i] = ge;
  } else {
#pragma unroll
    for (int i = 0; i < NGLL; i++) geo[i] = ld_off(geom, 16u * (l0 + i));
  }
  const bool last = k == ne - 1;
  float2 un[NGLL];  // the displacement of the group's next shot, loaded while the current one is computed
#pragma unroll
  for (int i = 0; i < NGLL; i++) un[i] = ld_off(displ + (long)s0 * shot_stride, 8u * (unsigned)g[i]);
  for (int shot = s0; shot < s1; shot++) {
  const long base = (long)shot * shot_stride;
#pragma unroll
  for (int i = 0; i < NGLL; i++) u[i] = un[i];
  if (MULTI && shot + 1 < s1) {
    const float2* dn = displ + base + shot_stride;
#pragma unroll
    for (int i = 0; i < NGLL; i++) un[i] = ld_off(dn, 8u * (unsigned)g[i]);
  }
  float2* as = accel + base;
  float2 a[NGLL];
#pragma unroll
  for (int i = 0; i < NGLL - 1; i++)
    a[i] = first[i] ? make_float2(0.f, 0.f) : ld_off((const float2*)as, 8u * (unsigned)g[i]);
  a[NGLL - 1] = (last && !first[NGLL - 1]) ? ld_off((const float2*)as, 8u * (unsigned)g[NGLL - 1])
                                          : make_float2(0.f, 0.f);

  // row j of u into the tile, gamma-derivatives by columns
#pragma unroll
  for (int i = 0; i < NGLL; i++) s_t[k][j][i] = u[i];
  __syncthreads();
  {
    const int c = j;  // this lane's column
    float2 col[NGLL];
#pragma unroll
    for (int m = 0; m < NGLL; m++) col[m] = s_t[k][m][c];
#pragma unroll
    for (int r = 0; r < NGLL; r++) {
      float2 dg = make_float2(0.f, 0.f);
#pragma unroll
      for (int m = 0; m < NGLL; m++) {
        const float h = s_hp[r * NGLL + m];
        dg.x += col[m].x * h;
        dg.y += col[m].y * h;
      }
      s_v[k][r][c] = dg;  // d u / d gamma at point (c, r)
    }
  }
  __syncthreads();
  // stress and fluxes at the row's points
  float2 t1[NGLL], t2[NGLL];
#pragma unroll
  for (int i = 0; i < NGLL; i++) {
    float dux_dxi = 0.f, duz_dxi = 0.f;
#pragma unroll
    for (int m = 0; m < NGLL; m++) {
      const float h = s_hp[i * NGLL + m];
      dux_dxi += u[m].x * h;
      duz_dxi += u[m].y * h;
    }
    const float2 dg = s_v[k][j][i];
    const float kap = mod[i].x, mu = mod[i].y;
    const float lambdal = kap - mu;
    const float lp2m = kap + mu;
    if (RECT) {  // xi_z = gamma_x = 0: J xi_x = 1 / gamma_z, J gamma_z = 1 / xi_x
      const float xixl = geo[i].x, gazl = geo[i].w;
      const float dux_dxl = dux_dxi * xixl, dux_dzl = dg.x * gazl;
      const float duz_dxl = duz_dxi * xixl, duz_dzl = dg.y * gazl;
      const float sxx = lp2m * dux_dxl + lambdal * duz_dzl;
      const float szz = lp2m * duz_dzl + lambdal * dux_dxl;
      const float sxz = mu * (duz_dxl + dux_dzl);
      const float cx = 1.f / gazl, cz = 1.f / xixl;
      t1[i] = make_float2(cx * sxx, cx * sxz);
      t2[i] = make_float2(cz * sxz, cz * szz);
    } else {
      const float xixl = geo[i].x, xizl = geo[i].y, gaxl = geo[i].z, gazl = geo[i].w;
      const float dux_dxl = dux_dxi * xixl + dg.x * gaxl;
      const float dux_dzl = dux_dxi * xizl + dg.x * gazl;
      const float duz_dxl = duz_dxi * xixl + dg.y * gaxl;
      const float duz_dzl = duz_dxi * xizl + dg.y * gazl;
      const float jac = 1.f / (xixl * gazl - gaxl * xizl);
      const float sxx = lp2m * dux_dxl + lambdal * duz_dzl;
      const float szz = lp2m * duz_dzl + lambdal * dux_dxl;
      const float sxz = mu * (duz_dxl + dux_dzl);
      t1[i] = make_float2(jac * (sxx * xixl + sxz * xizl), jac * (sxz * xixl + szz * xizl));
      t2[i] = make_float2(jac * (sxx * gaxl + sxz * gazl), jac * (sxz * gaxl + szz * gazl));
    }
  }
  __syncthreads();  // every lane has read its derivatives and the tile's u
#pragma unroll
  for (int i = 0; i < NGLL; i++) s_t[k][j][i] = t2[i];
  __syncthreads();
  {
    const int c = j;
    float2 col[NGLL];
#pragma unroll
    for (int m = 0; m < NGLL; m++) col[m] = s_t[k][m][c];
#pragma unroll
    for (int r = 0; r < NGLL; r++) {
      float2 sg = make_float2(0.f, 0.f);
#pragma unroll
      for (int m = 0; m < NGLL; m++) {
        const float h = s_hpw[m * NGLL + r];
        sg.x += col[m].x * h;
        sg.y += col[m].y * h;
      }
      s_v[k][r][c] = sg;  // gamma-flux contraction at point (c, r)
    }
  }
  __syncthreads();
  float2 f[NGLL];
  const float wj = basis.w[j];
#pragma unroll
  for (int i = 0; i < NGLL; i++) {
    float sx = 0.f, sz = 0.f;
#pragma unroll
    for (int m = 0; m < NGLL; m++) {
      const float h = s_hpw[m * NGLL + i];
      sx += t1[m].x * h;
      sz += t1[m].y * h;
    }
    const float2 sg = s_v[k][j][i];
    const float wi = basis.w[i];
    f[i] = make_float2(-(wj * sx + wi * sg.x), -(wj * sz + wi * sg.y));
  }
  // the right edge's force moves to the next element's left edge
  const float fx4 = __shfl_up(f[NGLL - 1].x, NGLL, CHAIN_LANES);
  const float fz4 = __shfl_up(f[NGLL - 1].y, NGLL, CHAIN_LANES);
  if (k > 0) {
    f[0].x += fx4;
    f[0].y += fz4;
  }
  if (active) {
#pragma unroll
    for (int i = 0; i < NGLL - 1; i++) {
      a[i].x += f[i].x;
      a[i].y += f[i].y;
      *reinterpret_cast<float2*>(reinterpret_cast<char*>(as) + 8u * (unsigned)g[i]) = a[i];
    }
    if (last) {
      a[NGLL - 1].x += f[NGLL - 1].x;
      a[NGLL - 1].y += f[NGLL - 1].y;
      *reinterpret_cast<float2*>(reinterpret_cast<char*>(as) + 8u * (unsigned)g[NGLL - 1]) = a[NGLL - 1];
    }
  }
  }  // shots
}

// The chain kernel with lanes along columns: lane 5 k + i holds column i of element k (its five points (i, j),
// j = 0..4), so that the wave's access to GLL row j of the patch touches consecutive points (the patch's points are
// numbered row by row) and its element data, stored in lane order (Mesh::d_col_ids, d_col_mod), consecutive words.
// The roles of the directions swap against the row-lane kernel: gamma-derivatives and the gamma-flux contraction run
// along the lane's own column in registers, the xi direction goes through the two LDS transposes; the force of the
// shared column (4, j) of element k moves to lane 5 (k + 1) by a one-lane shuffle.  Every sum has the row-lane
// kernel's terms in its order, and the same (element, point) pairs write accel, so the first-touch flags carry over.
// rows: when not null, the point ids and first-touch flags come from one int4 per GLL row of the patch
// (Mesh::d_col_rows: id = base + x for column x > 0 of the chain, the id of x = 0, the flags as a 64-bit mask)
// instead of col_ids
// ATT (one wavefield): viscoelastic attenuation, the memory variables and previous strains as three float4 planes
// of [3][ne] in lane order (att_e; see att_load) and a coefficient-set index per point in lane order (att_idx)
struct ColAtt {
  const unsigned int* idx;
  const float4* table;
  float* e;
  float* old;
  long ne;
};

// the attenuation update of the point at lane-order position la (lanes past the patch compute and store nothing)
__device__ __forceinline__ void col_attenuate(const ColAtt& att, bool active, unsigned la, float kap, float mu,
                                              float dux_dxl, float dux_dzl, float duz_dxl, float duz_dzl, float& sxx,
                                              float& sxz, float& szz) {
  if (!active) return;
  float4 cf[3];
  float ev[9], ov[3];
  const float4* coef = att.table + 3 * (size_t)att.idx[la];
#pragma unroll
  for (int k = 0; k < 3; k++) cf[k] = coef[k];
  att_load(att.e, att.ne, la, ev, ov);
  attenuate_point(cf, ev, ov, kap, mu, dux_dxl, dux_dzl, duz_dxl, duz_dzl, sxx, sxz, szz);
  att_store(att.e, att.ne, la, ev, ov);
}

template <bool FIRST, bool MULTI, bool RECT, bool ATT = false>
__global__ __launch_bounds__(CHAIN_LANES) void k_elastic_forces_col(
    int first_patch, int npatch, int nshot, int group, long shot_stride, const int4* __restrict__ desc,
    const int* __restrict__ col_ids, const float4* __restrict__ geom, const float4* __restrict__ geom_elem,
    const float2* __restrict__ col_mod, const float2* __restrict__ displ, float2* __restrict__ accel,
    const Basis basis, const int4* __restrict__ rows, const ColAtt att) {
  __shared__ float2 s_t[Mesh::CHAIN_ELEMS + 1][NGLL][NGLL + 1];
  __shared__ float2 s_v[Mesh::CHAIN_ELEMS + 1][NGLL][NGLL + 1];
  __shared__ float s_hp[NGLL2], s_hpw[NGLL2];
  const int lane = threadIdx.x;
  if (lane < NGLL2) {
    s_hp[lane] = basis.hp[lane];
    s_hpw[lane] = basis.hpw[lane];
  }
  const int nw = gridDim.x;
  const int item = (blockIdx.x % NXCD) * (nw / NXCD) + blockIdx.x / NXCD;
  if (!MULTI) group = 1;
  const int ngroup = (nshot + group - 1) / group;
  const int nitem = npatch * ngroup;
  if (item >= nitem) return;  // uniform over the wave
  const int4 d = desc[first_patch + item / ngroup];
  const int s0 = (item % ngroup) * group, s1 = min(nshot, s0 + group);
  const int ne = patch_ne(d.w);
  const int k = min(lane / NGLL, Mesh::CHAIN_ELEMS);  // element of the patch (CHAIN_ELEMS: lanes past it)
  const int i = lane % NGLL;                          // this lane's column
  const bool active = lane / NGLL < ne;
  const int e = d.x + min(k, ne - 1);
  const int row = NGLL * ne;                                   // words per GLL row of the patch in lane order
  const unsigned c0 = (unsigned)d.x * NGLL2 + (unsigned)min(lane, row - 1);  // this lane's word of row 0
  int g[NGLL];
  bool first[NGLL];
  float2 u[NGLL], mod[NGLL];
  float4 geo[NGLL];
  if (rows) {
    const int x = 4 * min(k, ne - 1) + i;  // column of the chain (lanes past it: the last element's)
    const int4* rp = rows + (size_t)(first_patch + item / ngroup) * NGLL;
#pragma unroll
    for (int j = 0; j < NGLL; j++) {
      const int4 r = rp[j];  // the same word for every lane
      g[j] = x == 0 ? r.y : r.x + x;
      const unsigned bits = x < 32 ? ((unsigned)r.z >> x) : ((unsigned)r.w >> (x - 32));
      first[j] = FIRST && (bits & 1u);
      mod[j] = ld_off(col_mod, 8u * (c0 + (unsigned)(j * row)));
    }
  } else {
#pragma unroll
    for (int j = 0; j < NGLL; j++) {
      const int id = ld_off(col_ids, 4u * (c0 + (unsigned)(j * row)));
      g[j] = id & ~Mesh::FIRST_TOUCH;
      first[j] = FIRST && (id & Mesh::FIRST_TOUCH);
      mod[j] = ld_off(col_mod, 8u * (c0 + (unsigned)(j * row)));
    }
  }
  if (RECT || patch_affine(d.w)) {
    const float4 ge = ld_off(geom_elem, 16u * (unsigned)e);
#pragma unroll
    for (int j = 0; j < NGLL; j++) geo[j] = ge;
  } else {
#pragma unroll
    for (int j = 0; j < NGLL; j++) geo[j] = ld_off(geom, 16u * ((unsigned)e * NGLL2 + j * NGLL + i));
  }
  const bool last = k == ne - 1;
  const bool writer = active && (i < NGLL - 1 || last);  // the shared column is written by the next element's lane
  float2 un[NGLL];
#pragma unroll
  for (int j = 0; j < NGLL; j++) un[j] = ld_off(displ + (long)s0 * shot_stride, 8u * (unsigned)g[j]);
  for (int shot = s0; shot < s1; shot++) {
  const long base = (long)shot * shot_stride;
#pragma unroll
  for (int j = 0; j < NGLL; j++) u[j] = un[j];
  if (MULTI && shot + 1 < s1) {
    const float2* dn = displ + base + shot_stride;
#pragma unroll
    for (int j = 0; j < NGLL; j++) un[j] = ld_off(dn, 8u * (unsigned)g[j]);
  }
  float2* as = accel + base;
  float2 a[NGLL];
#pragma unroll
  for (int j = 0; j < NGLL; j++)
    a[j] = (writer && !first[j]) ? ld_off((const float2*)as, 8u * (unsigned)g[j]) : make_float2(0.f, 0.f);

  // column i of u into the tile; xi-derivatives by rows: this lane takes row r = i
#pragma unroll
  for (int j = 0; j < NGLL; j++) s_t[k][j][i] = u[j];
  __syncthreads();
  {
    const int r = i;
    float2 rw[NGLL];
#pragma unroll
    for (int m = 0; m < NGLL; m++) rw[m] = s_t[k][r][m];
#pragma unroll
    for (int c = 0; c < NGLL; c++) {
      float2 dx = make_float2(0.f, 0.f);
#pragma unroll
      for (int m = 0; m < NGLL; m++) {
        const float h = s_hp[c * NGLL + m];
        dx.x += rw[m].x * h;
        dx.y += rw[m].y * h;
      }
      s_v[k][r][c] = dx;  // d u / d xi at point (c, r)
    }
  }
  __syncthreads();
  // stress and fluxes at the column's points
  float2 t1[NGLL], t2[NGLL];
#pragma unroll
  for (int j = 0; j < NGLL; j++) {
    float dux_dga = 0.f, duz_dga = 0.f;
#pragma unroll
    for (int m = 0; m < NGLL; m++) {
      const float h = s_hp[j * NGLL + m];
      dux_dga += u[m].x * h;
      duz_dga += u[m].y * h;
    }
    const float2 dxi = s_v[k][j][i];
    const float kap = mod[j].x, mu = mod[j].y;
    const float lambdal = kap - mu;
    const float lp2m = kap + mu;
    if (RECT) {  // xi_z = gamma_x = 0: J xi_x = 1 / gamma_z, J gamma_z = 1 / xi_x
      const float xixl = geo[j].x, gazl = geo[j].w;
      const float dux_dxl = dxi.x * xixl, dux_dzl = dux_dga * gazl;
      const float duz_dxl = dxi.y * xixl, duz_dzl = duz_dga * gazl;
      float sxx = lp2m * dux_dxl + lambdal * duz_dzl;
      float szz = lp2m * duz_dzl + lambdal * dux_dxl;
      float sxz = mu * (duz_dxl + dux_dzl);
      if (ATT) col_attenuate(att, active, c0 + (unsigned)(j * row), kap, mu, dux_dxl, dux_dzl, duz_dxl, duz_dzl, sxx,
                             sxz, szz);
      const float cx = 1.f / gazl, cz = 1.f / xixl;
      t1[j] = make_float2(cx * sxx, cx * sxz);
      t2[j] = make_float2(cz * sxz, cz * szz);
    } else {
      const float xixl = geo[j].x, xizl = geo[j].y, gaxl = geo[j].z, gazl = geo[j].w;
      const float dux_dxl = dxi.x * xixl + dux_dga * gaxl;
      const float dux_dzl = dxi.x * xizl + dux_dga * gazl;
      const float duz_dxl = dxi.y * xixl + duz_dga * gaxl;
      const float duz_dzl = dxi.y * xizl + duz_dga * gazl;
      const float jac = 1.f / (xixl * gazl - gaxl * xizl);
      float sxx = lp2m * dux_dxl + lambdal * duz_dzl;
      float szz = lp2m * duz_dzl + lambdal * dux_dxl;
      float sxz = mu * (duz_dxl + dux_dzl);
      if (ATT) col_attenuate(att, active, c0 + (unsigned)(j * row), kap, mu, dux_dxl, dux_dzl, duz_dxl, duz_dzl, sxx,
                             sxz, szz);
      t1[j] = make_float2(jac * (sxx * xixl + sxz * xizl), jac * (sxz * xixl + szz * xizl));
      t2[j] = make_float2(jac * (sxx * gaxl + sxz * gazl), jac * (sxz * gaxl + szz * gazl));
    }
  }
  __syncthreads();  // every lane has read its derivatives and the tile's u
#pragma unroll
  for (int j = 0; j < NGLL; j++) s_t[k][j][i] = t1[j];
  __syncthreads();
  {
    const int r = i;
    float2 rw[NGLL];
#pragma unroll
    for (int m = 0; m < NGLL; m++) rw[m] = s_t[k][r][m];
#pragma unroll
    for (int c = 0; c < NGLL; c++) {
      float2 sx = make_float2(0.f, 0.f);
#pragma unroll
      for (int m = 0; m < NGLL; m++) {
        const float h = s_hpw[m * NGLL + c];
        sx.x += rw[m].x * h;
        sx.y += rw[m].y * h;
      }
      s_v[k][r][c] = sx;  // xi-flux contraction at point (c, r)
    }
  }
  __syncthreads();
  float2 f[NGLL];
  const float wi = basis.w[i];
#pragma unroll
  for (int j = 0; j < NGLL; j++) {
    float sgx = 0.f, sgz = 0.f;
#pragma unroll
    for (int m = 0; m < NGLL; m++) {
      const float h = s_hpw[m * NGLL + j];
      sgx += t2[m].x * h;
      sgz += t2[m].y * h;
    }
    const float2 sx = s_v[k][j][i];
    const float wj = basis.w[j];
    f[j] = make_float2(-(wj * sx.x + wi * sgx), -(wj * sx.y + wi * sgz));
  }
  // the shared column's force moves from element k - 1 (lane 5 k - 1) to element k (lane 5 k)
#pragma unroll
  for (int j = 0; j < NGLL; j++) {
    const float fx = __shfl_up(f[j].x, 1, CHAIN_LANES);
    const float fz = __shfl_up(f[j].y, 1, CHAIN_LANES);
    if (i == 0 && k > 0) {
      f[j].x += fx;
      f[j].y += fz;
    }
  }
  if (writer) {
#pragma unroll
    for (int j = 0; j < NGLL; j++) {
      a[j].x += f[j].x;
      a[j].y += f[j].y;
      *reinterpret_cast<float2*>(reinterpret_cast<char*>(as) + 8u * (unsigned)g[j]) = a[j];
    }
  }
  }  // shots
}

// Shot-minor batches (fields [point][shot]: ps = nshot shots per point): the lanes of a wave are shots.  A wave holds
// 64 / L (patch, shot group) items, L = nshot rounded up to a power of two and at most 64 (lanes past the batch
// idle), and each lane walks its item's chain element by element with the element's 25 points of its shot in
// registers: the derivatives along both directions need no exchange between lanes (no LDS, no barriers), the
// element data is the same for the L lanes of an item (one cache line serves them all), and each field access
// of a point is L consecutive float2.  The left column of element k+1 is the right column of element k, so its
// displacement and force stay in registers from one element to the next.  The weights are folded into the fluxes:
// f(i, j) = -(sum_m [w_j t1(m, j)] hpw(m, i) + sum_m [w_i t2(i, m)] hpw(m, j)).
template <bool FIRST, bool AFFINE>
__device__ __forceinline__ void walk_body(
    int first_patch, int npatch, int nshot, int lshift, const int4* __restrict__ desc, const int* __restrict__ ids,
    const float4* __restrict__ geom, const float4* __restrict__ geom_elem, const float2* __restrict__ moduli,
    const float2* __restrict__ displ, float2* __restrict__ accel, const Basis& basis, float2 (*s_u)[64]) {
  const int lane = threadIdx.x;
  const int ngroup = (nshot + (1 << lshift) - 1) >> lshift;
  const int nitem = npatch * ngroup;
  const int nw = gridDim.x;
  const int wave = (blockIdx.x % NXCD) * (nw / NXCD) + blockIdx.x / NXCD;
  const int item = wave * (64 >> lshift) + (lane >> lshift);
  if (wave * (64 >> lshift) >= nitem) return;  // uniform: the wave has no item
  const int itc = min(item, nitem - 1);
  const int shot = (itc % ngroup) * (1 << lshift) + (lane & ((1 << lshift) - 1));
  const bool live = item < nitem && shot < nshot;
  // 32-bit element offsets (the launcher checks that points x shots x 8 B stays below 4 GiB)
  const unsigned sh = (unsigned)min(shot, nshot - 1);
  const unsigned ps = (unsigned)nshot;
  const int4 d = desc[first_patch + itc / ngroup];
  const int ne = patch_ne(d.w);
#define U(j, i) s_u[(j) * NGLL + (i)][lane]
  float2 fc[NGLL];  // force carried along the shared column
#pragma unroll 1
  for (int k = 0; k < ne; k++) {
    const unsigned l0 = (unsigned)(d.x + k) * NGLL2;
#pragma unroll
    for (int j = 0; j < NGLL; j++) {
      if (k > 0) U(j, 0) = U(j, NGLL - 1);  // the shared column
#pragma unroll
      for (int i = (k > 0 ? 1 : 0); i < NGLL; i++) {
        const unsigned g = (unsigned)(ld_off(ids, 4u * (l0 + j * NGLL + i)) & ~Mesh::FIRST_TOUCH);
        U(j, i) = ld_off(displ, 8u * (g * ps + sh));
      }
    }
    const float4 ge = AFFINE ? ld_off(geom_elem, 16u * (l0 / NGLL2)) : make_float4(0.f, 0.f, 0.f, 0.f);
    float2 f[NGLL][NGLL];
#pragma unroll
    for (int j = 0; j < NGLL; j++)
#pragma unroll
      for (int i = 0; i < NGLL; i++) f[j][i] = make_float2(0.f, 0.f);
#pragma unroll
    for (int j = 0; j < NGLL; j++) {
      __builtin_amdgcn_sched_barrier(0);  // one row at a time: bounds the loads in flight, and the registers
      float2 row[NGLL], t1[NGLL];
#pragma unroll
      for (int m = 0; m < NGLL; m++) row[m] = U(j, m);
#pragma unroll
      for (int i = 0; i < NGLL; i++) {
        float dux_dxi = 0.f, duz_dxi = 0.f, dux_dga = 0.f, duz_dga = 0.f;
#pragma unroll
        for (int m = 0; m < NGLL; m++) {
          const float hi = basis.hp[i * NGLL + m], hj = basis.hp[j * NGLL + m];
          const float2 c = U(m, i);
          dux_dxi += row[m].x * hi;
          duz_dxi += row[m].y * hi;
          dux_dga += c.x * hj;
          duz_dga += c.y * hj;
        }
        const unsigned l = l0 + j * NGLL + i;
        const float4 gp = AFFINE ? ge : ld_off(geom, 16u * l);
        const float2 md = ld_off(moduli, 8u * l);
        const float xixl = gp.x, xizl = gp.y, gaxl = gp.z, gazl = gp.w;
        const float dux_dxl = dux_dxi * xixl + dux_dga * gaxl;
        const float dux_dzl = dux_dxi * xizl + dux_dga * gazl;
        const float duz_dxl = duz_dxi * xixl + duz_dga * gaxl;
        const float duz_dzl = duz_dxi * xizl + duz_dga * gazl;
        const float jac = 1.f / (xixl * gazl - gaxl * xizl);
        const float lambdal = md.x - md.y;
        const float lp2m = md.x + md.y;
        const float sxx = lp2m * dux_dxl + lambdal * duz_dzl;
        const float szz = lp2m * duz_dzl + lambdal * dux_dxl;
        const float sxz = md.y * (duz_dxl + dux_dzl);
        const float wj = basis.w[j] * jac, wi = basis.w[i] * jac;
        t1[i] = make_float2(wj * (sxx * xixl + sxz * xizl), wj * (sxz * xixl + szz * xizl));
        const float2 t2 = make_float2(wi * (sxx * gaxl + sxz * gazl), wi * (sxz * gaxl + szz * gazl));
#pragma unroll
        for (int jp = 0; jp < NGLL; jp++) {  // gamma-flux contraction, column i
          const float h = basis.hpw[j * NGLL + jp];
          f[jp][i].x += t2.x * h;
          f[jp][i].y += t2.y * h;
        }
      }
#pragma unroll
      for (int i = 0; i < NGLL; i++)  // xi-flux contraction, row j
#pragma unroll
        for (int m = 0; m < NGLL; m++) {
          const float h = basis.hpw[m * NGLL + i];
          f[j][i].x += t1[m].x * h;
          f[j][i].y += t1[m].y * h;
        }
    }
    if (k > 0)
#pragma unroll
      for (int j = 0; j < NGLL; j++) {
        f[j][0].x += fc[j].x;
        f[j][0].y += fc[j].y;
      }
    const bool last = k == ne - 1;
#pragma unroll
    for (int j = 0; j < NGLL; j++) {
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int i = 0; i < NGLL; i++) {
        if (i == NGLL - 1 && !last) continue;
        const int raw = ld_off(ids, 4u * (l0 + j * NGLL + i));
        const unsigned off = 8u * ((unsigned)(raw & ~Mesh::FIRST_TOUCH) * ps + sh);
        float2 a = (FIRST && (raw & Mesh::FIRST_TOUCH)) ? make_float2(0.f, 0.f) : ld_off((const float2*)accel, off);
        a.x -= f[j][i].x;
        a.y -= f[j][i].y;
        if (live) *reinterpret_cast<float2*>(reinterpret_cast<char*>(accel) + off) = a;
      }
    }
#pragma unroll
    for (int j = 0; j < NGLL; j++) fc[j] = f[j][NGLL - 1];
  }
#undef U
}

// every element affine (one geometry record per element): a register budget of three waves per SIMD
template <bool FIRST>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) void k_elastic_forces_walk_affine(
    int first_patch, int npatch, int nshot, int lshift, const int4* __restrict__ desc, const int* __restrict__ ids,
    const float4* __restrict__ geom, const float4* __restrict__ geom_elem, const float2* __restrict__ moduli,
    const float2* __restrict__ displ, float2* __restrict__ accel, const Basis basis) {
  __shared__ float2 s_u[NGLL2][64];  // each lane's element displacement: point p of lane l at s_u[p][l]
  walk_body<FIRST, true>(first_patch, npatch, nshot, lshift, desc, ids, geom, geom_elem, moduli, displ, accel, basis,
                         s_u);
}

// per-point geometry
template <bool FIRST>
__global__ __launch_bounds__(64) void k_elastic_forces_walk(
    int first_patch, int npatch, int nshot, int lshift, const int4* __restrict__ desc, const int* __restrict__ ids,
    const float4* __restrict__ geom, const float4* __restrict__ geom_elem, const float2* __restrict__ moduli,
    const float2* __restrict__ displ, float2* __restrict__ accel, const Basis basis) {
  __shared__ float2 s_u[NGLL2][64];
  walk_body<FIRST, false>(first_patch, npatch, nshot, lshift, desc, ids, geom, geom_elem, moduli, displ, accel, basis,
                          s_u);
}

// per element: geometry of point 0, and whether it stands for all 25 points: every value within 2^-22 of the
// element's largest geometry entry of the value at point 0.  Meshers leave cancellation residues (1e-18 against
// 2e-2) where a parallelogram's entries are constant, so exact equality would almost never hold; within this
// bound the substitution changes the element forces at the level of FP32 rounding.
__global__ __launch_bounds__(256) void k_affine_elements(int nspec, const float4* __restrict__ geom,
                                                         float4* __restrict__ geom_elem, int* __restrict__ affine) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nspec) return;
  const float4 g0 = geom[(size_t)e * NGLL2];
  float m = 0.f;
  for (int p = 0; p < NGLL2; p++) {
    const float4 g = geom[(size_t)e * NGLL2 + p];
    m = fmaxf(m, fmaxf(fmaxf(fabsf(g.x), fabsf(g.y)), fmaxf(fabsf(g.z), fabsf(g.w))));
  }
  const float tol = m * 0x1p-22f;
  int same = m > 0.f;  // false for NaN too
  for (int p = 0; p < NGLL2; p++) {
    const float4 g = geom[(size_t)e * NGLL2 + p];
    same &= (int)(fabsf(g.x - g0.x) <= tol) & (int)(fabsf(g.y - g0.y) <= tol) & (int)(fabsf(g.z - g0.z) <= tol) &
            (int)(fabsf(g.w - g0.w) <= tol);
  }
  // 2: affine and axis-aligned (xi_z and gamma_x zero to the same bound), 1: affine, 0: neither
  const int rect = same && fabsf(g0.y) <= tol && fabsf(g0.z) <= tol;
  geom_elem[e] = g0;
  affine[e] = same ? 1 + rect : 0;
}

// sets the affine flag (bit 7 of the counts word) of every patch whose elements are all affine
__global__ __launch_bounds__(256) void k_flag_affine_patches(int npatch, int4* __restrict__ desc,
                                                             const int* __restrict__ affine) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= npatch) return;
  int4 d = desc[q];
  const int ne = d.w & 0xf;
  int all = 1;
  for (int k = 0; k < ne; k++) all &= affine[d.x + k] != 0;
  d.w = all ? (d.w | 0x80) : (d.w & ~0x80);
  desc[q] = d;
}

// packs the patch kernel's per-point element data: geom = {xix, xiz, gammax, gammaz}, moduli = {kappa, mu}
__global__ __launch_bounds__(256) void k_pack_elements(size_t n, const float* __restrict__ xix, const float* __restrict__ xiz,
                                                       const float* __restrict__ gammax, const float* __restrict__ gammaz,
                                                       const float* __restrict__ kappav, const float* __restrict__ muv,
                                                       float4* __restrict__ geom, float2* __restrict__ moduli) {
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x) {
    geom[k] = make_float4(xix[k], xiz[k], gammax[k], gammaz[k]);
    moduli[k] = make_float2(kappav[k], muv[k]);
  }
}

// gather per-element arrays into the device element order: dst[pos] = src[ispec_of_pos[pos]]
template <typename T>
__global__ __launch_bounds__(256) void k_permute(const T* __restrict__ src, T* __restrict__ dst,
                                                 const int* __restrict__ ispec_of_pos, int nspec) {
  const size_t n = (size_t)nspec * NGLL2;
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x) {
    const size_t pos = k / NGLL2, p = k - pos * NGLL2;
    dst[k] = src[(size_t)ispec_of_pos[pos] * NGLL2 + p];
  }
}

// ---------------------------------------------------------------------------
// Stacey absorbing traction (src/gpu/kernels/compute_stacey_elastic_kernel.cu:37-150,
// src/specfem2D/compute_stacey_elastic.f90:34-1001)
__global__ __launch_bounds__(64) void k_stacey(int nabs, const int* __restrict__ abs_ispec,
                                               const int* __restrict__ abs_ij, const float* __restrict__ normal,
                                               const float* __restrict__ jacw, const int* __restrict__ edge_abs,
                                               const int* __restrict__ ib, const int* __restrict__ ibool,
                                               const int* __restrict__ is_elastic, const float* __restrict__ rho_vp,
                                               const float* __restrict__ rho_vs, const float2* __restrict__ veloc,
                                               float* __restrict__ accel, int write_store, float* store_bottom,
                                               float* store_right, float* store_top, float* store_left,
                                               long shot_stride, int ps) {
  veloc += blockIdx.y * shot_stride;
  accel += 2 * blockIdx.y * shot_stride;
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  const int f = idx / NGLL;
  const int k = idx - f * NGLL;
  if (f >= nabs) return;
  const int i = abs_ij[(f * NGLL + k) * 2];
  const int j = abs_ij[(f * NGLL + k) * 2 + 1];
  if (i < 0 || j < 0) return;  // duplicated corner point, handled by the neighbouring edge
  const int ispec = abs_ispec[f];
  if (!is_elastic[ispec]) return;
  const size_t l = (size_t)ispec * NGLL2 + j * NGLL + i;
  const long g = (long)ibool[l] * ps;
  const float2 v = veloc[g];
  const float nx = normal[(f * NGLL + k) * 2];
  const float nz = normal[(f * NGLL + k) * 2 + 1];
  const float vn = nx * v.x + nz * v.y;
  const float rvp = rho_vp[l], rvs = rho_vs[l];
  const float tx = rvp * vn * nx + rvs * (v.x - vn * nx);
  const float tz = rvp * vn * nz + rvs * (v.y - vn * nz);
  const float w = jacw[f * NGLL + k];
  const float fx = tx * w, fz = tz * w;
  atomic_add_f32(&accel[2 * (size_t)g], -fx);
  atomic_add_f32(&accel[2 * (size_t)g + 1], -fz);
  if (write_store) {
    const int side = edge_abs[f];
    float* st = side == 1 ? store_bottom : side == 2 ? store_right : side == 3 ? store_top : store_left;
    if (st) {
      const int o = (ib[f] * NGLL + k) * 2;
      st[o] = fx;
      st[o + 1] = fz;
    }
  }
}

// backward wavefield: subtract the stored forward traction (compute_stacey_elastic_kernel.cu:155-218)
__global__ __launch_bounds__(64) void k_stacey_backward(int nabs, const int* __restrict__ abs_ispec,
                                                        const int* __restrict__ abs_ij,
                                                        const int* __restrict__ edge_abs, const int* __restrict__ ib,
                                                        const int* __restrict__ ibool,
                                                        const int* __restrict__ is_elastic, float* __restrict__ b_accel,
                                                        const float* sb, const float* sr, const float* stp,
                                                        const float* sl) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  const int f = idx / NGLL;
  const int k = idx - f * NGLL;
  if (f >= nabs) return;
  const int i = abs_ij[(f * NGLL + k) * 2];
  const int j = abs_ij[(f * NGLL + k) * 2 + 1];
  if (i < 0 || j < 0) return;
  const int ispec = abs_ispec[f];
  if (!is_elastic[ispec]) return;
  const int g = ibool[(size_t)ispec * NGLL2 + j * NGLL + i];
  const int side = edge_abs[f];
  const float* st = side == 1 ? sb : side == 2 ? sr : side == 3 ? stp : sl;
  if (!st) return;
  const int o = (ib[f] * NGLL + k) * 2;
  atomic_add_f32(&b_accel[2 * (size_t)g], -st[o]);
  atomic_add_f32(&b_accel[2 * (size_t)g + 1], -st[o + 1]);
}

// ---------------------------------------------------------------------------
// collocated force sources (src/gpu/kernels/compute_add_sources_kernel.cu:37-70)
__global__ __launch_bounds__(64) void k_sources(int nsrc, int it, const int* __restrict__ ispec_src,
                                                const int* __restrict__ is_elastic, const int* __restrict__ ibool,
                                                const float* __restrict__ sourcearrays, const float* __restrict__ stf,
                                                float* __restrict__ accel, const int* __restrict__ src_shot,
                                                long shot_stride, int ps, const int* __restrict__ step) {
  const int s = blockIdx.x;
  const int p = threadIdx.x;
  if (s >= nsrc || p >= NGLL2) return;
  if (step) it = *step - 1;  // graph replays: the 1-based step from device memory
  if (src_shot) accel += 2 * (long)src_shot[s] * shot_stride;  // each source drives its own shot's field
  const int ispec = ispec_src[s];
  if (!is_elastic[ispec]) return;
  const long g = (long)ibool[(size_t)ispec * NGLL2 + p] * ps;
  const float f = stf[(size_t)it * nsrc + s];
  atomic_add_f32(&accel[2 * (size_t)g], sourcearrays[(s * NGLL2 + p) * 2] * f);
  atomic_add_f32(&accel[2 * (size_t)g + 1], sourcearrays[(s * NGLL2 + p) * 2 + 1] * f);
}

// adjoint sources at the receivers (src/gpu/kernels/add_sources_el_SIM_TYPE_2_OR_3_kernel.cu:37-74)
__global__ __launch_bounds__(64) void k_adjoint_sources(int nadj, int it, int nstep, const int* __restrict__ ispec_rec,
                                                        const int* __restrict__ is_elastic,
                                                        const int* __restrict__ ibool, const float* __restrict__ xir,
                                                        const float* __restrict__ gammar,
                                                        const float* __restrict__ adj, float* __restrict__ accel) {
  const int r = blockIdx.x;
  const int p = threadIdx.x;
  if (r >= nadj || p >= NGLL2) return;
  const int ispec = ispec_rec[r];
  if (!is_elastic[ispec]) return;
  const int i = p % NGLL, j = p / NGLL;
  const int g = ibool[(size_t)ispec * NGLL2 + p];
  const float h = xir[i * nadj + r] * gammar[j * nadj + r];
  const float ax = adj[((size_t)0 * nstep + it) * nadj + r];
  const float az = adj[((size_t)1 * nstep + it) * nadj + r];
  atomic_add_f32(&accel[2 * (size_t)g], ax * h);
  atomic_add_f32(&accel[2 * (size_t)g + 1], az * h);
}

// ---------------------------------------------------------------------------
// seismograms: Lagrange interpolation at the receivers, FP64 accumulation, rotation
// (src/gpu/kernels/compute_elastic_seismogram_kernel.cu:69-192; compute_interpolated_dva.f90)
__global__ __launch_bounds__(64) void k_seismograms(int nrec, int nlength, int cur, const int* __restrict__ ispec_rec,
                                                    const int* __restrict__ ibool, const float* __restrict__ xir,
                                                    const float* __restrict__ gammar, const float* __restrict__ cosrot,
                                                    const float* __restrict__ sinrot, const float2* __restrict__ field,
                                                    float* __restrict__ dst, long shot_stride, long dst_stride,
                                                    int ps, const int* __restrict__ step) {
  if (step) cur = *step - 1;  // graph replays: the 1-based step from device memory
  field += blockIdx.y * shot_stride;
  dst += blockIdx.y * dst_stride;
  const int r = blockIdx.x;
  const int p = threadIdx.x;
  if (r >= nrec) return;
  double vx = 0.0, vz = 0.0;
  if (p < NGLL2) {
    const int i = p % NGLL, j = p / NGLL;
    const long g = (long)ibool[(size_t)ispec_rec[r] * NGLL2 + p] * ps;
    const double h = (double)(xir[i * nrec + r] * gammar[j * nrec + r]);
    const float2 u = field[g];
    vx = (double)u.x * h;
    vz = (double)u.y * h;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    vx += __shfl_xor(vx, off, 64);
    vz += __shfl_xor(vz, off, 64);
  }
  if (p == 0) {
    const double c = cosrot[r], s = sinrot[r];
    dst[(size_t)r * nlength + cur] = (float)(c * vx + s * vz);
    dst[((size_t)nrec + r) * nlength + cur] = (float)(-s * vx + c * vz);
  }
}

// ---------------------------------------------------------------------------
// adjoint sensitivity kernels rho/mu/kappa (src/gpu/kernels/compute_kernels_cuda_kernel.cu:37-200,
// src/specfem2D/compute_kernels.f90)
__global__ __launch_bounds__(FORCE_THREADS) void k_kernels(int nspec, const int* __restrict__ ibool,
                                                           const int* __restrict__ is_elastic,
                                                           const float* __restrict__ xix,
                                                           const float* __restrict__ xiz,
                                                           const float* __restrict__ gammax,
                                                           const float* __restrict__ gammaz,
                                                           const float2* __restrict__ displ,
                                                           const float2* __restrict__ accel,
                                                           const float2* __restrict__ b_displ,
                                                           float* __restrict__ rho_kl, float* __restrict__ mu_kl,
                                                           float* __restrict__ kappa_kl, const Basis basis) {
  __shared__ float s_hp[NGLL2];
  __shared__ float s_ux[FORCE_THREADS], s_uz[FORCE_THREADS], s_bx[FORCE_THREADS], s_bz[FORCE_THREADS];
  const int tid = threadIdx.x;
  if (tid < NGLL2) s_hp[tid] = basis.hp[tid];
  const int le = tid / NGLL2;
  const int p = tid - le * NGLL2;
  const int i = p % NGLL, j = p / NGLL;
  const int ispec = blockIdx.x * FORCE_EPB + le;
  bool active = (tid < FORCE_EPB * NGLL2) && (ispec < nspec);
  if (active) active = is_elastic[ispec] != 0;
  int g = 0;
  size_t l = 0;
  if (active) {
    l = (size_t)ispec * NGLL2 + p;
    g = ibool[l];
    const float2 u = displ[g], b = b_displ[g];
    s_ux[tid] = u.x;
    s_uz[tid] = u.y;
    s_bx[tid] = b.x;
    s_bz[tid] = b.y;
  }
  __syncthreads();
  if (!active) return;
  const int base = le * NGLL2;
  float t1x = 0.f, t1z = 0.f, t3x = 0.f, t3z = 0.f, b1x = 0.f, b1z = 0.f, b3x = 0.f, b3z = 0.f;
#pragma unroll
  for (int k = 0; k < NGLL; k++) {
    const float hi = s_hp[i * NGLL + k], hj = s_hp[j * NGLL + k];
    const int o1 = base + j * NGLL + k, o3 = base + k * NGLL + i;
    t1x += s_ux[o1] * hi;
    t1z += s_uz[o1] * hi;
    b1x += s_bx[o1] * hi;
    b1z += s_bz[o1] * hi;
    t3x += s_ux[o3] * hj;
    t3z += s_uz[o3] * hj;
    b3x += s_bx[o3] * hj;
    b3z += s_bz[o3] * hj;
  }
  const float xl = xix[l], xzl = xiz[l], gxl = gammax[l], gzl = gammaz[l];
  const float duxdxl = xl * t1x + gxl * t3x, duxdzl = xzl * t1x + gzl * t3x;
  const float duzdxl = xl * t1z + gxl * t3z, duzdzl = xzl * t1z + gzl * t3z;
  const float bduxdxl = xl * b1x + gxl * b3x, bduxdzl = xzl * b1x + gzl * b3x;
  const float bduzdxl = xl * b1z + gxl * b3z, bduzdzl = xzl * b1z + gzl * b3z;
  const float dsxx = duxdxl, dszz = duzdzl, dsxz = 0.5f * (duzdxl + duxdzl);
  const float bdsxx = bduxdxl, bdszz = bduzdzl, bdsxz = 0.5f * (bduzdxl + bduxdzl);
  const float2 a = accel[g], b = b_displ[g];
  rho_kl[l] += a.x * b.x + a.y * b.y;
  const float prod = (dsxx + dszz) * (bdsxx + bdszz);
  mu_kl[l] += dsxx * bdsxx + dszz * bdszz + 2.0f * dsxz * bdsxz - prod / 3.0f;
  kappa_kl[l] += prod;
}

// ---------------------------------------------------------------------------
// max |u|^2 over the points (src/gpu/kernels/get_maximum_vector_kernel.cu:37)
__global__ __launch_bounds__(256) void k_max_norm(const float2* __restrict__ f, int n, unsigned int* __restrict__ out) {
  float m = 0.f;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    const float2 u = f[k];
    const float v = u.x * u.x + u.y * u.y;
    m = (v > m || v != v) ? v : m;  // a NaN propagates
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const float o = __shfl_xor(m, off, 64);
    m = (o > m || o != o) ? o : m;
  }
  if ((threadIdx.x & 63) == 0) atomicMax(out, __float_as_uint(m));  // non-negative floats order as uints
}

inline int blocks_for(long n, int threads, int cap = 1 << 20) {
  long b = (n + threads - 1) / threads;
  if (b < 1) b = 1;
  if (b > cap) b = cap;
  return (int)b;
}

}  // namespace

// step counter of graph replays (one thread)
__global__ void k_set_step(int* step, int value) { *step = value; }
__global__ void k_next_step(int* step) { *step += 1; }

// ---------------------------------------------------------------------------
// launchers
// Newmark, Stacey and seismogram launchers take nshot wavefields of shot_stride points each (shot-major),
// with the shot in blockIdx.y; nshot = 1 is the single-wavefield case of the reference entry points.
static dim3 shot_grid(long n, int threads, int cap, int nshot) {
  if (nshot < 1 || nshot > 65535) fail("shot count %d outside [1, 65535]", nshot);
  return dim3(blocks_for(n, threads, cap), nshot);
}

// shot-minor batches (ps = shots per point > 1): one flat pass over the points x shots entries
void launch_newmark_predictor(float* d, float* v, float* a, int nglob, float dt, float dt2o2, float dto2,
                              hipStream_t s, int nshot, long shot_stride, int zero_accel, int ps) {
  if (nglob <= 0) return;
  const long n = (long)nglob * ps;
  if (ps > 1) nshot = 1;
  hipLaunchKernelGGL(k_predictor, shot_grid(n, 256, 65536, nshot), dim3(256), 0, s, (float2*)d, (float2*)v,
                     (float2*)a, (int)n, dt, dt2o2, dto2, shot_stride, zero_accel);
  check_hip(hipGetLastError(), "k_predictor launch");
}

void launch_newmark_corrector(float* v, float* a, const float* rmx, const float* rmz, int nglob, float dto2,
                              int apply_mass, hipStream_t s, int nshot, long shot_stride, int ps) {
  if (nglob <= 0) return;
  const long n = (long)nglob * ps;
  if (ps > 1) nshot = 1;
  hipLaunchKernelGGL(k_corrector, shot_grid(n, 256, 65536, nshot), dim3(256), 0, s, (float2*)v, (float2*)a, rmx,
                     rmz, (int)n, dto2, apply_mass, shot_stride, ps);
  check_hip(hipGetLastError(), "k_corrector launch");
}

void launch_newmark_corrector_predictor(float* d, float* v, float* a, const float* rmx, const float* rmz, int nglob,
                                       float dto2_prev, float dt, float dt2o2, float dto2, hipStream_t s,
                                       int nshot, long shot_stride, int zero_accel, int ps) {
  if (nglob <= 0) return;
  const long n = (long)nglob * ps;
  if (ps > 1) nshot = 1;
  hipLaunchKernelGGL(k_corrector_predictor, shot_grid(n, 256, 65536, nshot), dim3(256), 0, s, (float2*)d,
                     (float2*)v, (float2*)a, rmx, rmz, (int)n, dto2_prev, dt, dt2o2, dto2, shot_stride, zero_accel,
                     ps);
  check_hip(hipGetLastError(), "k_corrector_predictor launch");
}

void launch_newmark_corrector_predictor_to(const float* d, const float* v, const float* a, float* d_out, float* v_out,
                                          float* a_out, const float* rmx, const float* rmz, int nglob,
                                          float dto2_prev, float dt, float dt2o2, float dto2, hipStream_t s,
                                          int nshot, long shot_stride, int zero_accel, int ps) {
  if (nglob <= 0) return;
  const long n = (long)nglob * ps;
  if (ps > 1) nshot = 1;
  hipLaunchKernelGGL(k_corrector_predictor_to, shot_grid(n, 256, 65536, nshot), dim3(256), 0, s,
                     (const float2*)d, (const float2*)v, (const float2*)a, (float2*)d_out, (float2*)v_out,
                     (float2*)a_out, rmx, rmz, (int)n, dto2_prev, dt, dt2o2, dto2, shot_stride, zero_accel, ps);
  check_hip(hipGetLastError(), "k_corrector_predictor_to launch");
}

void launch_zero_points(const Mesh* mp, float* accel, int nshot, long shot_stride, int ps) {
  if (mp->nspecial <= 0) return;
  const long n = (long)mp->nspecial * ps;
  if (ps > 1) nshot = 1;
  hipLaunchKernelGGL(k_zero_points, shot_grid(n, 256, 4096, nshot), dim3(256), 0, mp->stream, (float2*)accel,
                     mp->d_special, (int)n, shot_stride, ps);
  check_hip(hipGetLastError(), "k_zero_points launch");
}

void launch_elastic_forces(const Mesh* mp, int first, int nelem, int atomic, const float* displ, float* accel,
                           int att) {
  if (nelem <= 0) return;
  if (first < 0 || first + nelem > mp->nspec) fail("element range [%d, %d) outside the mesh", first, first + nelem);
  if (att && (!mp->d_att_idx || !mp->d_att_table || !mp->d_att_e))
    fail("attenuation arrays not prepared");
  const int nbatch = (nelem + FORCE_EPB - 1) / FORCE_EPB;
  const int per_xcd = (nbatch + NXCD - 1) / NXCD;
  const int nblocks = per_xcd * NXCD;
  using K = decltype(&k_elastic_forces<true, true, true>);
  const K kernels[2][2][2] = {
      {{k_elastic_forces<false, false, false>, k_elastic_forces<false, false, true>},
       {k_elastic_forces<false, true, false>, k_elastic_forces<false, true, true>}},
      {{k_elastic_forces<true, false, false>, k_elastic_forces<true, false, true>},
       {k_elastic_forces<true, true, false>, k_elastic_forces<true, true, true>}}};
  hipLaunchKernelGGL(kernels[atomic ? 1 : 0][att ? 1 : 0][mp->d_geom_rect ? 1 : 0], dim3(nblocks), dim3(FORCE_THREADS),
                     0, mp->stream, first, nelem, nbatch, per_xcd, mp->d_ibool, mp->d_xix, mp->d_xiz, mp->d_gammax,
                     mp->d_gammaz, mp->d_kappav, mp->d_muv, (const float2*)displ, accel, mp->basis, mp->d_att_idx,
                     mp->d_att_table, mp->d_att_e, mp->d_att_old, (long)mp->nspec * NGLL2,
                     (const float2*)mp->d_geom_rect, mp->d_aniso);
  check_hip(hipGetLastError(), "k_elastic_forces launch");
}

void launch_elastic_forces_patch(const Mesh* mp, int first_patch, int npatch, const float* displ, float* accel,
                                 int nshot, long shot_stride, int first_touch) {
  if (npatch <= 0) return;
  if (nshot < 1 || (nshot > 1 && shot_stride < mp->nglob) || (long)npatch * nshot > (1L << 30))
    fail("batched force launch: %d shots of stride %ld over %d patches", nshot, shot_stride, npatch);
  if (first_patch < 0 || first_patch + npatch > mp->npatch)
    fail("patch range [%d, %d) outside the %d patches", first_patch, first_patch + npatch, mp->npatch);
  // persistent grid: as many workgroups as stay resident on the device (a multiple of the XCD count);
  // SF2D_PATCH_WGS_PER_CU overrides the workgroups per CU (tuning), 0 gives one workgroup per patch
  static int resident = -1;
  if (resident < 0) {
    int dev = 0, cus = 0, per_cu = 0;
    check_hip(hipGetDevice(&dev), "hipGetDevice");
    check_hip(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "CU count");
    check_hip(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_elastic_forces_patch, FORCE_THREADS, 0),
              "occupancy");
    if (const char* env = getenv("SF2D_PATCH_WGS_PER_CU")) per_cu = atoi(env);
    resident = per_cu <= 0 ? 0 : std::max(NXCD, (cus * per_cu) / NXCD * NXCD);
    if (getenv("SF2D_VERBOSE")) fprintf(stderr, "k_elastic_forces_patch: %d CUs, %d workgroups resident\n", cus, resident);
  }
  const int all = (npatch * nshot + NXCD - 1) / NXCD * NXCD;
  if (mp->chain) {  // one wave per (patch, group of shots)
    if (first_touch && !mp->first_touch) fail("first-touch assembly is not set up for this handle");
    const int group = std::min(nshot, mp->shots_per_wave > 0 ? mp->shots_per_wave : CHAIN_SHOTS_PER_WAVE);
    const int items = npatch * ((nshot + group - 1) / group);
    const int grid = (items + NXCD - 1) / NXCD * NXCD;
    using K = decltype(&k_elastic_forces_chain<true, true, true>);
    const K kerns[2][2][2] = {
        {{k_elastic_forces_chain<false, false, false>, k_elastic_forces_chain<false, false, true>},
         {k_elastic_forces_chain<false, true, false>, k_elastic_forces_chain<false, true, true>}},
        {{k_elastic_forces_chain<true, false, false>, k_elastic_forces_chain<true, false, true>},
         {k_elastic_forces_chain<true, true, false>, k_elastic_forces_chain<true, true, true>}}};
    if (mp->d_col_ids) {  // lanes along columns, element data in lane order
      using KC = decltype(&k_elastic_forces_col<true, true, true>);
      const KC ckerns[2][2][2] = {
          {{k_elastic_forces_col<false, false, false>, k_elastic_forces_col<false, false, true>},
           {k_elastic_forces_col<false, true, false>, k_elastic_forces_col<false, true, true>}},
          {{k_elastic_forces_col<true, false, false>, k_elastic_forces_col<true, false, true>},
           {k_elastic_forces_col<true, true, false>, k_elastic_forces_col<true, true, true>}}};
      // viscoelastic attenuation of the forward wavefield: one shot, memory variables in lane order
      const bool att = mp->attenuation && accel == mp->d_accel;
      if (att && nshot != 1) fail("viscoelastic attenuation: one wavefield per call");
      const KC akerns[2][2] = {{k_elastic_forces_col<false, false, false, true>, k_elastic_forces_col<false, false, true, true>},
                               {k_elastic_forces_col<true, false, false, true>, k_elastic_forces_col<true, false, true, true>}};
      const ColAtt ca{mp->d_att_idx, mp->d_att_table, mp->d_att_e, mp->d_att_old, (long)mp->nspec * NGLL2};
      hipLaunchKernelGGL(att ? akerns[first_touch ? 1 : 0][mp->rect ? 1 : 0]
                             : ckerns[first_touch ? 1 : 0][group > 1 ? 1 : 0][mp->rect ? 1 : 0],
                         dim3(grid), dim3(CHAIN_LANES), 0, mp->stream, first_patch, npatch, nshot, group, shot_stride,
                         (const int4*)mp->d_patch_desc, mp->d_col_ids, mp->d_geom, mp->d_geom_elem, mp->d_col_mod,
                         (const float2*)displ, (float2*)accel, mp->basis, (const int4*)mp->d_col_rows, ca);
      check_hip(hipGetLastError(), "k_elastic_forces_col launch");
      return;
    }
    const K kern = kerns[first_touch ? 1 : 0][group > 1 ? 1 : 0][mp->rect ? 1 : 0];
    hipLaunchKernelGGL(kern, dim3(grid), dim3(CHAIN_LANES), 0, mp->stream, first_patch, npatch, nshot, group,
                       shot_stride, (const int4*)mp->d_patch_desc, mp->d_chain_ids, mp->d_geom, mp->d_geom_elem,
                       mp->d_moduli, (const float2*)displ, (float2*)accel, mp->basis);
    check_hip(hipGetLastError(), "k_elastic_forces_chain launch");
    return;
  }
  if (first_touch) fail("first-touch assembly needs the chain kernel");
  const int nwg = resident > 0 ? std::min(resident, all) : all;
  hipLaunchKernelGGL(k_elastic_forces_patch, dim3(nwg), dim3(FORCE_THREADS), 0, mp->stream, first_patch, npatch,
                     nshot, shot_stride, (const int4*)mp->d_patch_desc, mp->d_halo_nodes, mp->d_elem_slot, mp->d_geom,
                     mp->d_geom_elem, mp->d_moduli, (const float2*)displ, (float2*)accel, mp->basis);
  check_hip(hipGetLastError(), "k_elastic_forces_patch launch");
}

void launch_elastic_forces_walk(const Mesh* mp, int first_patch, int npatch, const float* displ, float* accel,
                                int nshot, int first_touch) {
  if (npatch <= 0) return;
  if (!mp->chain) fail("shot-minor batches need chain patches");
  if ((size_t)mp->nglob * nshot * sizeof(float2) >= ((size_t)1 << 32))
    fail("shot-minor batch fields of %d points x %d shots exceed the 4 GiB of 32-bit offsets", mp->nglob, nshot);
  if (first_touch && !mp->first_touch) fail("first-touch assembly is not set up for this handle");
  if (first_patch < 0 || first_patch + npatch > mp->npatch)
    fail("patch range [%d, %d) outside the %d patches", first_patch, first_patch + npatch, mp->npatch);
  int lshift = 0;
  while ((1 << lshift) < nshot && lshift < 6) lshift++;
  const long items = (long)npatch * ((nshot + (1 << lshift) - 1) >> lshift);
  const long waves = (items + (64 >> lshift) - 1) / (64 >> lshift);
  if (waves > (1L << 30)) fail("shot-minor force launch too large");
  const int grid = (int)((waves + NXCD - 1) / NXCD * NXCD);
  const bool affine = !mp->no_affine && mp->n_affine == mp->nspec;
  auto kern = affine ? (first_touch ? k_elastic_forces_walk_affine<true> : k_elastic_forces_walk_affine<false>)
                     : (first_touch ? k_elastic_forces_walk<true> : k_elastic_forces_walk<false>);
  hipLaunchKernelGGL(kern, dim3(grid), dim3(64), 0, mp->stream, first_patch, npatch, nshot, lshift,
                     (const int4*)mp->d_patch_desc, mp->d_chain_ids, mp->d_geom, mp->d_geom_elem, mp->d_moduli,
                     (const float2*)displ, (float2*)accel, mp->basis);
  check_hip(hipGetLastError(), "k_elastic_forces_walk launch");
}

void launch_set_step(int* step, int value, hipStream_t s) {
  hipLaunchKernelGGL(k_set_step, dim3(1), dim3(1), 0, s, step, value);
  check_hip(hipGetLastError(), "k_set_step launch");
}
void launch_next_step(int* step, hipStream_t s) {
  hipLaunchKernelGGL(k_next_step, dim3(1), dim3(1), 0, s, step);
  check_hip(hipGetLastError(), "k_next_step launch");
}

// the column-lane kernel's lane-order copies of a patch's point ids and moduli (one workgroup per patch)
__global__ __launch_bounds__(256) void k_col_pack(int npatch, const int4* __restrict__ desc, const int* __restrict__ ids,
                                                  const float2* __restrict__ mod, int* __restrict__ col_ids,
                                                  float2* __restrict__ col_mod) {
  const int q = blockIdx.x;
  if (q >= npatch) return;
  const int4 d = desc[q];
  const int ne = patch_ne(d.w);
  const size_t e0 = (size_t)d.x * NGLL2;
  for (int t = threadIdx.x; t < ne * NGLL2; t += blockDim.x) {
    const int k = t / NGLL2, p = t % NGLL2, j = p / NGLL, i = p % NGLL;
    const size_t dst = e0 + (size_t)j * NGLL * ne + NGLL * k + i;
    col_ids[dst] = ids[e0 + t];
    col_mod[dst] = mod[e0 + t];
  }
}

void launch_col_pack(const Mesh* mp) {
  if (!mp->chain || !mp->d_col_ids || !mp->d_chain_ids || !mp->d_moduli || mp->npatch <= 0) return;
  hipLaunchKernelGGL(k_col_pack, dim3(mp->npatch), dim3(256), 0, mp->stream, mp->npatch,
                     (const int4*)mp->d_patch_desc, mp->d_chain_ids, mp->d_moduli, mp->d_col_ids, mp->d_col_mod);
  check_hip(hipGetLastError(), "k_col_pack launch");
}

void launch_pack_elements(const Mesh* mp) {
  const size_t n = (size_t)mp->nspec * NGLL2;
  hipLaunchKernelGGL(k_pack_elements, dim3(blocks_for((long)n, 256, 65536)), dim3(256), 0, mp->stream, n, mp->d_xix,
                     mp->d_xiz, mp->d_gammax, mp->d_gammaz, mp->d_kappav, mp->d_muv, mp->d_geom, mp->d_moduli);
  check_hip(hipGetLastError(), "k_pack_elements launch");
  launch_col_pack(mp);
}

int launch_flag_affine(Mesh* mp) {
  int* d_aff = device_zeros<int>((size_t)mp->nspec);
  hipLaunchKernelGGL(k_affine_elements, dim3((mp->nspec + 255) / 256), dim3(256), 0, mp->stream, mp->nspec, mp->d_geom,
                     mp->d_geom_elem, d_aff);
  if (mp->npatch > 0 && !mp->no_affine)
    hipLaunchKernelGGL(k_flag_affine_patches, dim3((mp->npatch + 255) / 256), dim3(256), 0, mp->stream, mp->npatch,
                       mp->d_patch_desc, d_aff);
  check_hip(hipGetLastError(), "affine flags launch");
  std::vector<int> h((size_t)mp->nspec);
  check_hip(hipMemcpyAsync(h.data(), d_aff, h.size() * sizeof(int), hipMemcpyDeviceToHost, mp->stream), "affine D2H");
  check_hip(hipStreamSynchronize(mp->stream), "affine flags");
  device_free(d_aff);
  int n = 0, nrect = 0;
  for (int v : h) {
    n += v != 0;
    nrect += v == 2;
  }
  // every element affine and axis-aligned: the chain kernel's rectangle form (SF2D_NO_RECT keeps the general one)
  mp->rect = !mp->no_affine && nrect == mp->nspec && !getenv("SF2D_NO_RECT");
  return n;
}

void launch_permute_elements(const float* src, float* dst, const int* ispec_of_pos, int nspec, hipStream_t s) {
  hipLaunchKernelGGL(k_permute<float>, dim3(blocks_for((long)nspec * NGLL2, 256, 65536)), dim3(256), 0, s, src, dst,
                     ispec_of_pos, nspec);
  check_hip(hipGetLastError(), "k_permute launch");
}

void launch_permute_elements_int(const int* src, int* dst, const int* ispec_of_pos, int nspec, hipStream_t s) {
  hipLaunchKernelGGL(k_permute<int>, dim3(blocks_for((long)nspec * NGLL2, 256, 65536)), dim3(256), 0, s, src, dst,
                     ispec_of_pos, nspec);
  check_hip(hipGetLastError(), "k_permute launch");
}

void launch_stacey(const Mesh* mp, const float* veloc, float* accel, int write_store, int nshot, long shot_stride,
                   int ps) {
  if (mp->num_abs <= 0) return;
  if (nshot > 1 && write_store) fail("the Stacey boundary store holds one wavefield");
  hipLaunchKernelGGL(k_stacey, shot_grid((long)mp->num_abs * NGLL, 64, 1 << 20, nshot), dim3(64), 0, mp->stream,
                     mp->num_abs, mp->d_abs_ispec, mp->d_abs_ij, mp->d_abs_normal, mp->d_abs_jacw, mp->d_edge_abs,
                     mp->d_ib, mp->d_ibool, mp->d_ispec_is_elastic, mp->d_rho_vp, mp->d_rho_vs, (const float2*)veloc,
                     accel, write_store, mp->d_b_absorb[1], mp->d_b_absorb[2], mp->d_b_absorb[3], mp->d_b_absorb[4],
                     shot_stride, ps);
  check_hip(hipGetLastError(), "k_stacey launch");
}

void launch_stacey_backward(const Mesh* mp, float* b_accel) {
  if (mp->num_abs <= 0) return;
  hipLaunchKernelGGL(k_stacey_backward, dim3(blocks_for((long)mp->num_abs * NGLL, 64)), dim3(64), 0, mp->stream,
                     mp->num_abs, mp->d_abs_ispec, mp->d_abs_ij, mp->d_edge_abs, mp->d_ib, mp->d_ibool,
                     mp->d_ispec_is_elastic, b_accel, mp->d_b_absorb[1], mp->d_b_absorb[2], mp->d_b_absorb[3],
                     mp->d_b_absorb[4]);
  check_hip(hipGetLastError(), "k_stacey_backward launch");
}

void launch_sources(const Mesh* mp, float* accel, int it, const int* src_shot, long shot_stride, int ps,
                    const int* step) {
  if (mp->nsources_local <= 0) return;
  hipLaunchKernelGGL(k_sources, dim3(mp->nsources_local), dim3(64), 0, mp->stream, mp->nsources_local, it,
                     mp->d_ispec_selected_source, mp->d_ispec_is_elastic, mp->d_ibool, mp->d_sourcearrays, mp->d_stf,
                     accel, src_shot, shot_stride, ps, step);
  check_hip(hipGetLastError(), "k_sources launch");
}

void launch_adjoint_sources(const Mesh* mp, float* accel, int it) {
  if (mp->nadj_rec_local <= 0) return;
  hipLaunchKernelGGL(k_adjoint_sources, dim3(mp->nadj_rec_local), dim3(64), 0, mp->stream, mp->nadj_rec_local, it,
                     mp->nstep, mp->d_ispec_selected_rec, mp->d_ispec_is_elastic, mp->d_ibool, mp->d_xir,
                     mp->d_gammar, mp->d_source_adjoint, accel);
  check_hip(hipGetLastError(), "k_adjoint_sources launch");
}

void launch_seismograms(const Mesh* mp, const float* field, float* dst, int seismo_current, int nshot,
                        long shot_stride, long dst_stride, int ps, const int* step) {
  if (mp->nrec_local <= 0) return;
  hipLaunchKernelGGL(k_seismograms, dim3(mp->nrec_local, nshot), dim3(64), 0, mp->stream, mp->nrec_local,
                     mp->nlength_seismogram, seismo_current, mp->d_ispec_selected_rec, mp->d_ibool, mp->d_xir,
                     mp->d_gammar, mp->d_cosrot, mp->d_sinrot, (const float2*)field, dst, shot_stride, dst_stride,
                     ps, step);
  check_hip(hipGetLastError(), "k_seismograms launch");
}

void launch_kernels(const Mesh* mp) {
  if (mp->nspec <= 0) return;
  const int nblocks = (mp->nspec + FORCE_EPB - 1) / FORCE_EPB;
  hipLaunchKernelGGL(k_kernels, dim3(nblocks), dim3(FORCE_THREADS), 0, mp->stream, mp->nspec, mp->d_ibool,
                     mp->d_ispec_is_elastic, mp->d_xix, mp->d_xiz, mp->d_gammax, mp->d_gammaz,
                     (const float2*)mp->d_displ, (const float2*)mp->d_accel, (const float2*)mp->d_b_displ,
                     mp->d_rho_kl, mp->d_mu_kl, mp->d_kappa_kl, mp->basis);
  check_hip(hipGetLastError(), "k_kernels launch");
}

void launch_max_norm(const Mesh* mp, const float* field, unsigned int* result) {
  hipLaunchKernelGGL(k_max_norm, dim3(blocks_for(mp->nglob, 256, 4096)), dim3(256), 0, mp->stream,
                     (const float2*)field, mp->nglob, result);
  check_hip(hipGetLastError(), "k_max_norm launch");
}

}  // namespace sf2d
