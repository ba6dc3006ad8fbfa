// CPML absorbing layers (PML_BOUNDARY_CONDITIONS, Newmark time scheme) for elastic and acoustic elements.
//
// What is computed is the reference's CPU path: the stretched gradients through memory variables of the current
// and previous-step fields (src/specfem2D/pml_compute_memory_variables.f90:42-540), the non-symmetric stress of the
// stretched equations (compute_forces_viscoelastic.F90:596-611), the damping terms of the acceleration
// (pml_compute_accel_contribution.f90:42-333), the coefficients of pml_compute.f90:34-318 and the zero condition on
// the outer boundary of the layer (pml_compute.f90:520-605).  The damping profiles (pml_init.F90) and the layer's
// mass terms (invert_mass_matrix.F90:133-345) are the caller's.  The reference's own GPU code handles acoustic
// layers only (src/gpu/kernels/Kernel_2_acoustic_impl.cu:346, prepare_mesh_constants_cuda.cu:843) and stops on
// elastic ones (src/specfem2D/prepare_gpu.f90:74).
//
// Design.  The layer's elements stay in the main assembly plan with zero moduli (zero J/rho for fluid elements), so
// the main force kernels add nothing for them and keep their launch shapes and first-touch schedule; the kernels
// here then add the whole CPML force of every layer element, one colour of layer elements per launch (greedy
// colouring, plain read-modify-write, so the sums are the same from run to run), after the main launches of the
// element's phase.  Per layer element point they stream the point id, the geometry, the moduli (fluid: J/rho), the
// damping weight w_i w_j rho J (fluid: w_i w_j J/kappa), a 4-byte index into a table of the distinct coefficient
// sets (the profiles depend on the distance into the layer only, so a few thousand sets serve a layer of millions
// of points), the previous-step field of the point (written per element point by a pass before the Newmark
// predictor) and the memory variables (12 planes for a solid point, 6 for a fluid one), all stored
// [plane][layer element point] so that consecutive lanes touch consecutive words.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <map>
#include <vector>

#include "sf2d_mesh.h"
#include "specfem2d_gpu.h"
#include "specfem2d_gpu_ext.h"

namespace sf2d {

namespace {

constexpr int PML_EPB = 10;                // layer elements per 256-thread workgroup (250 active lanes)
constexpr unsigned int ON_EDGE = 1u << 31; // coefficient index flag: the point lies on the layer's outer boundary
constexpr int SET_F4 = 6;                  // float4 words per coefficient set
constexpr int EL_PLANES = 12, AC_PLANES = 6;
// viscoelastic layer (3 standard linear solids): per point an index into a table of Zener coefficient sets (NO_ZENER:
// the elastic layer formulation), 18 more memory-variable planes; a set is 11 float4:
//   per solid k (3 float4): {A31k/Mk, A13k/Mk, A31m/Mm, A13m/Mm}, {ck_0, ck_1, ck_2, cm_0}, {cm_1, cm_2, phi_k, phi_m}
//   then {A2k/Mk, A3k/Mk, A4k/Mk, A5k/Mk}, {A2m/Mm, A3m/Mm, A4m/Mm, A5m/Mm}
// (A..k/m: the Zener terms of lik_parameter_computation_viscoelastic for the bulk and shear relaxation, divided by
// Mu = N_SLS Mu_nu; ck, cm: the recursive-convolution coefficients of 1/tau_sigma; A2..A5: the sums over the solids
// of the stretched-gradient coefficients)
constexpr int ZSLS = 3, ZSET_F4 = 3 * ZSLS + 2, ZPLANES = 6 * ZSLS;
constexpr unsigned int NO_ZENER = 0xffffffffu;

// One coefficient set (6 float4): the stretched-gradient coefficients A0..A2 of the x derivatives (index 31) and of
// the z derivatives (index 13), A1..A4 of the acceleration's damping terms, and the recursive-convolution
// coefficients (c0, c1, c2) of alpha_x, alpha_z, beta_x, beta_z:
//   t[0] = {A31_0, A31_1, A31_2, A13_0}   t[1] = {A13_1, A13_2, Al_1, Al_2}   t[2] = {Al_3, Al_4, ax_0, ax_1}
//   t[3] = {ax_2, az_0, az_1, az_2}       t[4] = {bx_0, bx_1, bx_2, bz_0}     t[5] = {bz_1, bz_2, 0, 0}
struct Coefs {
  float a31[3], a13[3], al[4], cax[3], caz[3], cbx[3], cbz[3];
};

__device__ __forceinline__ Coefs load_coefs(const float4* __restrict__ table, unsigned int c) {
  const float4* t = table + (size_t)c * SET_F4;
  const float4 t0 = t[0], t1 = t[1], t2 = t[2], t3 = t[3], t4 = t[4], t5 = t[5];
  Coefs k;
  k.a31[0] = t0.x; k.a31[1] = t0.y; k.a31[2] = t0.z;
  k.a13[0] = t0.w; k.a13[1] = t1.x; k.a13[2] = t1.y;
  k.al[0] = t1.z; k.al[1] = t1.w; k.al[2] = t2.x; k.al[3] = t2.y;
  k.cax[0] = t2.z; k.cax[1] = t2.w; k.cax[2] = t3.x;
  k.caz[0] = t3.y; k.caz[1] = t3.z; k.caz[2] = t3.w;
  k.cbx[0] = t4.x; k.cbx[1] = t4.y; k.cbx[2] = t4.z;
  k.cbz[0] = t4.w; k.cbz[1] = t5.x; k.cbz[2] = t5.y;
  return k;
}

// recursive convolution r_n = c0 r_{n-1} + c1 f_n + c2 f_{n-1}
__device__ __forceinline__ float conv(const float c[3], float r, float f, float f_old) {
  return c[0] * r + c[1] * f + c[2] * f_old;
}

// Solid layer elements of one colour, first .. first+n-1 of the packed layer order: accel -= the element's CPML
// force (compute_forces_viscoelastic.F90:247-270, 596-611 with pml_compute_memory_variables.f90:230-540 and
// pml_compute_accel_contribution.f90:187-333; Newmark, not rotated).
// VISCO: points with a Zener set take the viscoelastic layer formulation (pml_compute_memory_variables.f90:546-894,
// compute_forces_viscoelastic.F90:545-556: the bulk and shear parts of the stretched gradients convolved with the
// standard linear solids, stress with the unrelaxed moduli, kappa = lambda + mu).
template <bool VISCO>
__global__ __launch_bounds__(256) void k_pml_elastic_forces(
    int first, int n, long nq, const int* __restrict__ ids, const float4* __restrict__ geom,
    const float2* __restrict__ moduli, const float* __restrict__ dampw, const unsigned int* __restrict__ cidx,
    const float4* __restrict__ table, const float2* __restrict__ old, float* __restrict__ rm,
    const float2* __restrict__ displ, const float2* __restrict__ veloc, float2* __restrict__ accel, const Basis basis,
    const unsigned int* __restrict__ zidx, const float4* __restrict__ ztable, float* __restrict__ zrm) {
  __shared__ float s_hp[NGLL2], s_hpw[NGLL2], s_w[NGLL];
  __shared__ float s_ux[256], s_uz[256], s_ox[256], s_oz[256];
  __shared__ float s_t1x[256], s_t1z[256], s_t2x[256], s_t2z[256];
  const int tid = threadIdx.x;
  if (tid < NGLL2) {
    s_hp[tid] = basis.hp[tid];
    s_hpw[tid] = basis.hpw[tid];
  }
  if (tid < NGLL) s_w[tid] = basis.w[tid];
  const int le = tid / NGLL2;
  const int p = tid - le * NGLL2;
  const int i = p % NGLL, j = p / NGLL;
  const int e = blockIdx.x * PML_EPB + le;
  const bool active = le < PML_EPB && e < n;
  const size_t q = (size_t)(first + (active ? e : 0)) * NGLL2 + p;
  int g = 0;
  float4 gm = make_float4(0.f, 0.f, 0.f, 0.f);
  float2 md = make_float2(0.f, 0.f), u = md, o = md, v = md, a_old = md;
  float w = 0.f;
  float r[EL_PLANES];
  Coefs c;
  if (active) {
    g = ids[q];
    u = displ[g];
    a_old = accel[g];  // no other element of this colour has the point: read with the fields, in one round trip
    o = old[q];
    v = veloc[g];
    gm = geom[q];
    md = moduli[q];
    w = dampw[q];
    c = load_coefs(table, cidx[q] & ~ON_EDGE);
    {  // the 12 memory variables of the point as three float4 planes
      const float4* rm4 = reinterpret_cast<const float4*>(rm);
#pragma unroll
      for (int k = 0; k < EL_PLANES / 4; k++) {
        const float4 v4 = rm4[(size_t)k * nq + q];
        r[4 * k] = v4.x;
        r[4 * k + 1] = v4.y;
        r[4 * k + 2] = v4.z;
        r[4 * k + 3] = v4.w;
      }
    }
    s_ux[tid] = u.x;
    s_uz[tid] = u.y;
    s_ox[tid] = o.x;
    s_oz[tid] = o.y;
  }
  __syncthreads();
  const int base = le * NGLL2;
  float2 pacc = make_float2(0.f, 0.f);
  if (active) {
    float ux_xi = 0.f, uz_xi = 0.f, ux_ga = 0.f, uz_ga = 0.f, ox_xi = 0.f, oz_xi = 0.f, ox_ga = 0.f, oz_ga = 0.f;
#pragma unroll
    for (int k = 0; k < NGLL; k++) {
      const float hi = s_hp[i * NGLL + k], hj = s_hp[j * NGLL + k];
      const int a = base + j * NGLL + k, b = base + k * NGLL + i;
      ux_xi += s_ux[a] * hi;
      uz_xi += s_uz[a] * hi;
      ux_ga += s_ux[b] * hj;
      uz_ga += s_uz[b] * hj;
      ox_xi += s_ox[a] * hi;
      oz_xi += s_oz[a] * hi;
      ox_ga += s_ox[b] * hj;
      oz_ga += s_oz[b] * hj;
    }
    // the unmodified gradient (the reference's PML_dux_dxl & co.) and that of the previous-step field
    const float p_ux_x = ux_xi * gm.x + ux_ga * gm.z, p_ux_z = ux_xi * gm.y + ux_ga * gm.w;
    const float p_uz_x = uz_xi * gm.x + uz_ga * gm.z, p_uz_z = uz_xi * gm.y + uz_ga * gm.w;
    const float o_ux_x = ox_xi * gm.x + ox_ga * gm.z, o_ux_z = ox_xi * gm.y + ox_ga * gm.w;
    const float o_uz_x = oz_xi * gm.x + oz_ga * gm.z, o_uz_z = oz_xi * gm.y + oz_ga * gm.w;
    // memory variables: x derivatives convolved with alpha_z and beta_x, z derivatives with alpha_x and beta_z
    r[0] = conv(c.caz, r[0], p_ux_x, o_ux_x);
    r[1] = conv(c.cbx, r[1], p_ux_x, o_ux_x);
    r[2] = conv(c.caz, r[2], p_uz_x, o_uz_x);
    r[3] = conv(c.cbx, r[3], p_uz_x, o_uz_x);
    r[4] = conv(c.cax, r[4], p_ux_z, o_ux_z);
    r[5] = conv(c.cbz, r[5], p_ux_z, o_ux_z);
    r[6] = conv(c.cax, r[6], p_uz_z, o_uz_z);
    r[7] = conv(c.cbz, r[7], p_uz_z, o_uz_z);
    float sxx, szz, szx, sxz;
    const unsigned int zi = VISCO ? zidx[q] : NO_ZENER;
    if (VISCO && zi != NO_ZENER) {
      const float4* z = ztable + (size_t)zi * ZSET_F4;
      float ks_xx = 0.f, ks_zz = 0.f, ms_xx = 0.f, ms_zz = 0.f, ms_xz = 0.f, ms_zx = 0.f;
      float sk_xx = 0.f, sk_zz = 0.f, sm_xx = 0.f, sm_zz = 0.f, sm_xz = 0.f, sm_zx = 0.f;
#pragma unroll
      for (int k = 0; k < ZSLS; k++) {
        const float4 za = z[3 * k], zb = z[3 * k + 1], zc = z[3 * k + 2];
        const float ck[3] = {zb.x, zb.y, zb.z}, cm[3] = {zb.w, zc.x, zc.y};
        // solid k's 6 memory variables: a float4 plane {kxx, kzz, mxx, mzz} of [3][nq] and a float2 plane {mxz, mzx}
        // of [3][nq] after the float4 planes
        float4* v4 = reinterpret_cast<float4*>(zrm) + (size_t)k * nq + q;
        float2* v2 = reinterpret_cast<float2*>(zrm + (size_t)12 * nq) + (size_t)k * nq + q;
        const float4 a4 = *v4;
        const float2 a2 = *v2;
        const float kxx = conv(ck, a4.x, p_ux_x, o_ux_x);
        const float kzz = conv(ck, a4.y, p_uz_z, o_uz_z);
        const float mxx = conv(cm, a4.z, p_ux_x, o_ux_x);
        const float mzz = conv(cm, a4.w, p_uz_z, o_uz_z);
        const float mxz = conv(cm, a2.x, p_ux_z, o_ux_z);
        const float mzx = conv(cm, a2.y, p_uz_x, o_uz_x);
        *v4 = make_float4(kxx, kzz, mxx, mzz);
        *v2 = make_float2(mxz, mzx);
        ks_xx += za.x * kxx;
        ks_zz += za.y * kzz;
        ms_xx += za.z * mxx;
        ms_zz += za.w * mzz;
        ms_xz += za.w * mxz;
        ms_zx += za.z * mzx;
        sk_xx += zc.z * kxx;
        sk_zz += zc.z * kzz;
        sm_xx += zc.w * mxx;
        sm_zz += zc.w * mzz;
        sm_xz += zc.w * mxz;
        sm_zx += zc.w * mzx;
      }
      const float4 sa = z[3 * ZSLS], sb = z[3 * ZSLS + 1];
      const float vk_xx = c.a31[0] * p_ux_x + ks_xx + sa.x * r[0] + sa.y * r[1];
      const float vk_zz = c.a13[0] * p_uz_z + ks_zz + sa.z * r[6] + sa.w * r[7];
      const float vm_xx = c.a31[0] * p_ux_x + ms_xx + sb.x * r[0] + sb.y * r[1];
      const float vm_zz = c.a13[0] * p_uz_z + ms_zz + sb.z * r[6] + sb.w * r[7];
      const float vm_xz = c.a13[0] * p_ux_z + ms_xz + sb.z * r[4] + sb.w * r[5];
      const float vm_zx = c.a31[0] * p_uz_x + ms_zx + sb.x * r[2] + sb.y * r[3];
      const float kk_xx = p_ux_x + sk_xx, kk_zz = p_uz_z + sk_zz;
      const float mm_xx = p_ux_x + sm_xx, mm_zz = p_uz_z + sm_zz, mm_xz = p_ux_z + sm_xz, mm_zx = p_uz_x + sm_zx;
      const float kap = md.x, mu = md.y;
      sxx = (kap * vk_xx + mu * vm_xx) + (kap * kk_zz - mu * mm_zz);
      sxz = mu * vm_zx + mu * mm_xz;
      szx = mu * mm_zx + mu * vm_xz;
      szz = (kap * kk_xx - mu * mm_xx) + (kap * vk_zz + mu * vm_zz);
    } else {
      const float ux_x = c.a31[0] * p_ux_x + c.a31[1] * r[0] + c.a31[2] * r[1];
      const float uz_x = c.a31[0] * p_uz_x + c.a31[1] * r[2] + c.a31[2] * r[3];
      const float ux_z = c.a13[0] * p_ux_z + c.a13[1] * r[4] + c.a13[2] * r[5];
      const float uz_z = c.a13[0] * p_uz_z + c.a13[1] * r[6] + c.a13[2] * r[7];
      // non-symmetric stress of the stretched equations: lambda = kappa - mu, lambda + 2 mu = kappa + mu
      const float lam = md.x - md.y, lp2m = md.x + md.y;
      sxx = lp2m * ux_x + lam * p_uz_z;
      szz = lp2m * uz_z + lam * p_ux_x;
      szx = md.y * (p_uz_x + ux_z);
      sxz = md.y * (p_ux_z + uz_x);
    }
    const float jac = 1.f / (gm.x * gm.w - gm.z * gm.y);
    s_t1x[tid] = jac * (sxx * gm.x + szx * gm.y);
    s_t1z[tid] = jac * (sxz * gm.x + szz * gm.y);
    s_t2x[tid] = jac * (sxx * gm.z + szx * gm.w);
    s_t2z[tid] = jac * (sxz * gm.z + szz * gm.w);
    // damping terms: displacement convolved with alpha_x and alpha_z
    r[8] = conv(c.cax, r[8], u.x, o.x);
    r[9] = conv(c.cax, r[9], u.y, o.y);
    r[10] = conv(c.caz, r[10], u.x, o.x);
    r[11] = conv(c.caz, r[11], u.y, o.y);
    pacc.x = w * (c.al[0] * v.x + c.al[1] * u.x + c.al[2] * r[8] + c.al[3] * r[10]);
    pacc.y = w * (c.al[0] * v.y + c.al[1] * u.y + c.al[2] * r[9] + c.al[3] * r[11]);
    {
      float4* rm4 = reinterpret_cast<float4*>(rm);
#pragma unroll
      for (int k = 0; k < EL_PLANES / 4; k++)
        rm4[(size_t)k * nq + q] = make_float4(r[4 * k], r[4 * k + 1], r[4 * k + 2], r[4 * k + 3]);
    }
  }
  __syncthreads();
  if (active) {
    float t1x = 0.f, t1z = 0.f, t2x = 0.f, t2z = 0.f;
#pragma unroll
    for (int k = 0; k < NGLL; k++) {
      const float wi = s_hpw[k * NGLL + i], wj = s_hpw[k * NGLL + j];
      t1x += s_t1x[base + j * NGLL + k] * wi;
      t1z += s_t1z[base + j * NGLL + k] * wi;
      t2x += s_t2x[base + k * NGLL + i] * wj;
      t2z += s_t2z[base + k * NGLL + i] * wj;
    }
    float2 a = a_old;
    a.x = (a.x - (s_w[j] * t1x + s_w[i] * t2x)) - pacc.x;
    a.y = (a.y - (s_w[j] * t1z + s_w[i] * t2z)) - pacc.y;
    accel[g] = a;
  }
}

// Fluid layer elements of one colour: pdd -= the element's CPML term (compute_forces_viscoacoustic.f90 with
// pml_compute_memory_variables.f90:42-224 and pml_compute_accel_contribution.f90:42-181; Newmark).
__global__ __launch_bounds__(256) void k_pml_acoustic_forces(
    int first, int n, long nq, const int* __restrict__ ids, const float4* __restrict__ geom,
    const float2* __restrict__ fac, const float* __restrict__ dampw, const unsigned int* __restrict__ cidx,
    const float4* __restrict__ table, const float* __restrict__ old, float* __restrict__ rm,
    const float* __restrict__ pot, const float* __restrict__ pd, float* __restrict__ pdd, const Basis basis) {
  __shared__ float s_hp[NGLL2], s_hpw[NGLL2], s_w[NGLL];
  __shared__ float s_u[256], s_o[256], s_t1[256], s_t2[256];
  const int tid = threadIdx.x;
  if (tid < NGLL2) {
    s_hp[tid] = basis.hp[tid];
    s_hpw[tid] = basis.hpw[tid];
  }
  if (tid < NGLL) s_w[tid] = basis.w[tid];
  const int le = tid / NGLL2;
  const int p = tid - le * NGLL2;
  const int i = p % NGLL, j = p / NGLL;
  const int e = blockIdx.x * PML_EPB + le;
  const bool active = le < PML_EPB && e < n;
  const size_t q = (size_t)(first + (active ? e : 0)) * NGLL2 + p;
  int g = 0;
  float4 gm = make_float4(0.f, 0.f, 0.f, 0.f);
  float u = 0.f, o = 0.f, v = 0.f, f = 0.f, w = 0.f, pdd_old = 0.f;
  float r[AC_PLANES];
  Coefs c;
  if (active) {
    g = ids[q];
    u = pot[g];
    v = pd[g];
    pdd_old = pdd[g];  // no other element of this colour has the point
    o = old[q];
    gm = geom[q];
    f = fac[q].x;
    w = dampw[q];
    c = load_coefs(table, cidx[q] & ~ON_EDGE);
    {  // the 6 memory variables of the point as three float2 planes
      const float2* rm2 = reinterpret_cast<const float2*>(rm);
#pragma unroll
      for (int k = 0; k < AC_PLANES / 2; k++) {
        const float2 v2 = rm2[(size_t)k * nq + q];
        r[2 * k] = v2.x;
        r[2 * k + 1] = v2.y;
      }
    }
    s_u[tid] = u;
    s_o[tid] = o;
  }
  __syncthreads();
  const int base = le * NGLL2;
  float pacc = 0.f;
  if (active) {
    float u_xi = 0.f, u_ga = 0.f, o_xi = 0.f, o_ga = 0.f;
#pragma unroll
    for (int k = 0; k < NGLL; k++) {
      const float hi = s_hp[i * NGLL + k], hj = s_hp[j * NGLL + k];
      u_xi += s_u[base + j * NGLL + k] * hi;
      u_ga += s_u[base + k * NGLL + i] * hj;
      o_xi += s_o[base + j * NGLL + k] * hi;
      o_ga += s_o[base + k * NGLL + i] * hj;
    }
    const float p_x = u_xi * gm.x + u_ga * gm.z, p_z = u_xi * gm.y + u_ga * gm.w;
    const float o_x = o_xi * gm.x + o_ga * gm.z, o_z = o_xi * gm.y + o_ga * gm.w;
    r[0] = conv(c.caz, r[0], p_x, o_x);
    r[1] = conv(c.cbx, r[1], p_x, o_x);
    r[2] = conv(c.cax, r[2], p_z, o_z);
    r[3] = conv(c.cbz, r[3], p_z, o_z);
    const float dx = c.a31[0] * p_x + c.a31[1] * r[0] + c.a31[2] * r[1];
    const float dz = c.a13[0] * p_z + c.a13[1] * r[2] + c.a13[2] * r[3];
    s_t1[tid] = f * (gm.x * dx + gm.y * dz);
    s_t2[tid] = f * (gm.z * dx + gm.w * dz);
    r[4] = conv(c.cax, r[4], u, o);
    r[5] = conv(c.caz, r[5], u, o);
    pacc = w * (c.al[0] * v + c.al[1] * u + c.al[2] * r[4] + c.al[3] * r[5]);
    {
      float2* rm2 = reinterpret_cast<float2*>(rm);
#pragma unroll
      for (int k = 0; k < AC_PLANES / 2; k++) rm2[(size_t)k * nq + q] = make_float2(r[2 * k], r[2 * k + 1]);
    }
  }
  __syncthreads();
  if (active) {
    float a1 = 0.f, a2 = 0.f;
#pragma unroll
    for (int k = 0; k < NGLL; k++) {
      a1 += s_t1[base + j * NGLL + k] * s_hpw[k * NGLL + i];
      a2 += s_t2[base + k * NGLL + i] * s_hpw[k * NGLL + j];
    }
    pdd[g] = (pdd_old - (s_w[j] * a1 + s_w[i] * a2)) - pacc;
  }
}

// previous-step field of each layer element point, before the Newmark predictor (update_displacement_Newmark.F90:
// 373-380, 428-433): u_old = u + dt^2/4 a, with a the corrected acceleration (the deferred corrector's mass
// scaling applied here when it is pending); zero on the layer's outer boundary, which the zero condition clears
// before the field is used
__global__ __launch_bounds__(256) void k_pml_old_elastic(long nq, const int* __restrict__ ids,
                                                         const unsigned int* __restrict__ cidx,
                                                         const float2* __restrict__ d, const float2* __restrict__ a,
                                                         const float* __restrict__ rmx, const float* __restrict__ rmz,
                                                         int pending, float dt2o4, float2* __restrict__ old) {
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < nq; q += (long)gridDim.x * blockDim.x) {
    float2 out = make_float2(0.f, 0.f);
    if (!(cidx[q] & ON_EDGE)) {
      const int g = ids[q];
      float2 acc = a[g];
      if (pending) {
        acc.x *= rmx[g];
        acc.y *= rmz[g];
      }
      const float2 u = d[g];
      out = make_float2(u.x + dt2o4 * acc.x, u.y + dt2o4 * acc.y);
    }
    old[q] = out;
  }
}

__global__ __launch_bounds__(256) void k_pml_old_acoustic(long nq, const int* __restrict__ ids,
                                                          const unsigned int* __restrict__ cidx,
                                                          const float* __restrict__ pot, const float* __restrict__ pdd,
                                                          float dt2o4, float* __restrict__ old) {
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < nq; q += (long)gridDim.x * blockDim.x) {
    const int g = ids[q];
    old[q] = (cidx[q] & ON_EDGE) ? 0.f : pot[g] + dt2o4 * pdd[g];
  }
}

// zero condition on the outer boundary of the layer (pml_boundary_elastic / pml_boundary_acoustic)
__global__ __launch_bounds__(256) void k_pml_zero_elastic(int n, const int* __restrict__ pts, float2* __restrict__ d,
                                                          float2* __restrict__ v, float2* __restrict__ a) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int g = pts[k];
  const float2 z = make_float2(0.f, 0.f);
  d[g] = z;
  v[g] = z;
  a[g] = z;
}

__global__ __launch_bounds__(256) void k_pml_zero_acoustic(int n, const int* __restrict__ pts, float* __restrict__ pot,
                                                           float* __restrict__ pd, float* __restrict__ pdd) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int g = pts[k];
  pot[g] = 0.f;
  pd[g] = 0.f;
  pdd[g] = 0.f;
}

// Packs the layer elements' per-point operands from the handle's element arrays (device element order): point id,
// geometry, J = 1/(xi_x gamma_z - gamma_x xi_z) as the device solver forms it, and
//   solid: moduli {kappa, mu} and damping weight (w_i w_j)(rho J);  fluid: {J/rho, 0} and (w_i w_j)(J/kappa)
__global__ __launch_bounds__(256) void k_pml_pack(long nq, int fluid, const int* __restrict__ pos,
                                                  const float* __restrict__ rho, const int* __restrict__ ibool,
                                                  const float* __restrict__ xix, const float* __restrict__ xiz,
                                                  const float* __restrict__ gammax, const float* __restrict__ gammaz,
                                                  const float* __restrict__ kappav, const float* __restrict__ muv,
                                                  const Basis basis, int* __restrict__ ids, float4* __restrict__ geom,
                                                  float2* __restrict__ moduli, float* __restrict__ dampw) {
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < nq; q += (long)gridDim.x * blockDim.x) {
    const long e = q / NGLL2;
    const int p = (int)(q - e * NGLL2);
    const size_t l = (size_t)pos[e] * NGLL2 + p;
    const float4 gm = make_float4(xix[l], xiz[l], gammax[l], gammaz[l]);
    const float jac = 1.f / (gm.x * gm.w - gm.z * gm.y);
    const float wij = basis.w[p % NGLL] * basis.w[p / NGLL];
    ids[q] = ibool[l];
    geom[q] = gm;
    if (fluid) {
      moduli[q] = make_float2(jac / rho[q], 0.f);
      dampw[q] = wij * (jac / kappav[l]);
    } else {
      moduli[q] = make_float2(kappav[l], muv[l]);
      dampw[q] = wij * (rho[q] * jac);
    }
  }
}

// the main force kernels skip the solid layer elements: zero moduli at their element points
__global__ __launch_bounds__(256) void k_pml_clear_moduli(long nq, const int* __restrict__ pos, float* __restrict__ kappav,
                                                          float* __restrict__ muv) {
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < nq; q += (long)gridDim.x * blockDim.x) {
    const long e = q / NGLL2;
    const size_t l = (size_t)pos[e] * NGLL2 + (q - e * NGLL2);
    kappav[l] = 0.f;
    muv[l] = 0.f;
  }
}

int grid_for(long n) { return (int)std::max(1L, std::min((n + 255) / 256, 16384L)); }

// ---------------------------------------------------------------------------
// coefficients (host, double precision as the reference computes them)

// compute_coef_convolution (pml_compute.f90:34-89): exp(-b dt), (1 - exp(-b dt/2))/b, that times exp(-b dt/2)
void coef_convolution(double b, double dt, double c[3]) {
  const double t = std::exp(-0.5 * b * dt);
  c[0] = t * t;
  if (std::fabs(b) > 1e-5) {
    c[1] = (1.0 - t) / b;
    c[2] = c[1] * t;
  } else {
    c[1] = c[2] = 0.5 * dt;
  }
}

// lik_parameter_computation (pml_compute.f90:93-169): the partial-fraction coefficients of the stretched gradient;
// index 31 passes the z quantities in the x slots and swaps the roles of the x-only and z-only regions
void lik_coefficients(double kx, double bx, double ax, double kz, double bz, double az, int region, int index_ik,
                      double A[3]) {
  const int xonly = index_ik == 13 ? 1 : 2, zonly = index_ik == 13 ? 2 : 1;
  if (region == 3) {
    const double a0 = kx / kz;
    const double gx = (ax * bz + ax * ax + 2.0 * bx * az - 2.0 * ax * (bx + az)) / (bz - ax);
    const double gz = (ax * bz + bz * bz + 2.0 * bx * az - 2.0 * bz * (bx + az)) / (ax - bz);
    A[0] = a0;
    A[1] = a0 / 2.0 * (gx - ax);
    A[2] = a0 / 2.0 * (gz - bz);
  } else if (region == xonly) {
    A[0] = kx;
    A[1] = -kx * (ax - bx);
    A[2] = 0.0;
  } else if (region == zonly) {
    A[0] = 1.0 / kz;
    A[1] = 0.0;
    A[2] = -A[0] * (bz - az);
  } else {
    A[0] = A[1] = A[2] = 0.0;
  }
}

// l_parameter_computation (pml_compute.f90:244-318): (i w)^2 s_x s_z = A0 (i w)^2 + A1 i w + A2 + A3/(i w + alpha_x)
// + A4/(i w + alpha_z)
void l_coefficients(double kx, double bx, double ax, double kz, double bz, double az, int region, double A[5]) {
  if (region == 3) {
    const double a0 = kx * kz;
    const double gx = (ax * az + ax * ax + 2.0 * bx * bz - 2.0 * ax * (bx + bz)) / (az - ax);
    const double gz = (ax * az + az * az + 2.0 * bx * bz - 2.0 * az * (bx + bz)) / (ax - az);
    A[0] = a0;
    A[1] = a0 / 2.0 * (gx - ax + gz - az);
    A[2] = a0 / 2.0 * (ax * ax - gx * ax + az * az - gz * az);
    A[3] = a0 / 2.0 * (ax * ax) * (gx - ax);
    A[4] = a0 / 2.0 * (az * az) * (gz - az);
  } else if (region == 1) {
    A[0] = kx;
    A[1] = kx * (bx - ax);
    A[2] = -kx * ax * (bx - ax);
    A[3] = kx * (ax * ax) * (bx - ax);
    A[4] = 0.0;
  } else if (region == 2) {
    A[0] = kz;
    A[1] = kz * (bz - az);
    A[2] = -kz * az * (bz - az);
    A[3] = 0.0;
    A[4] = kz * (az * az) * (bz - az);
  } else {
    for (int k = 0; k < 5; k++) A[k] = 0.0;
  }
}

// lik_parameter_computation_viscoelastic (pml_compute.f90:173-280): A_0, A_1, A_2 and the Zener term of one standard
// linear solid (1/tau_sigma itau, tau_epsilon te); the corner elements' partial fractions as in lik_coefficients
void lik_visco_coefficients(double kx, double bx, double ax, double kz, double bz, double az, int region,
                            int index_ik, double itau, double te, double A[4]) {
  const int xonly = index_ik == 13 ? 1 : 2, zonly = index_ik == 13 ? 2 : 1;
  auto gamma_eta = [&](double al, double be, double* ga, double* et) {
    *ga = (be * (itau - al * te * itau) + al * al * (te * itau - 1.0)) / (itau - al);
    *et = itau * (be - al) / (itau - al);
  };
  if (region == 3) {
    const double a0 = kx / kz;
    const double al[2] = {ax, bz};
    const double be[2] = {(ax * bz + ax * ax + 2.0 * bx * az - 2.0 * ax * (bx + az)) / (bz - ax),
                          (ax * bz + bz * bz + 2.0 * bx * az - 2.0 * bz * (bx + az)) / (ax - bz)};
    double ga[2], et[2];
    for (int i = 0; i < 2; i++) gamma_eta(al[i], be[i], &ga[i], &et[i]);
    A[0] = a0;
    A[1] = a0 / 2.0 * (ga[0] - al[0]);
    A[2] = a0 / 2.0 * (ga[1] - al[1]);
    A[3] = a0 * (te * itau - 1.0) * ((et[0] + et[1]) / 2.0 - itau);
  } else if (region == xonly || region == zonly) {
    const bool x = region == xonly;
    const double a0 = x ? kx : 1.0 / kz;
    double ga, et;
    gamma_eta(x ? ax : bz, x ? bx : az, &ga, &et);
    A[0] = a0;
    A[1] = x ? a0 * (ga - ax) : 0.0;
    A[2] = x ? 0.0 : a0 * (ga - bz);
    A[3] = a0 * (te * itau - 1.0) * (et - itau);
  } else {
    A[0] = A[1] = A[2] = A[3] = 0.0;
  }
}

using ZenerKey = std::array<float, 4 * ZSET_F4>;

// the Zener coefficient set of one layer point (layout above), double precision as the reference forms it
ZenerKey zener_coefficients(double dt, const double prof[6], int region, const double itk[ZSLS],
                            const double tek[ZSLS], const double phk[ZSLS], double mk, const double itm[ZSLS],
                            const double tem[ZSLS], const double phm[ZSLS], double mm) {
  const double kx = prof[0], kz = prof[1], dx = prof[2], dz = prof[3], ax = prof[4], az = prof[5];
  const double bx = ax + dx / kx, bz = az + dz / kz;
  ZenerKey z{};
  double a2k = 0, a3k = 0, a4k = 0, a5k = 0, a2m = 0, a3m = 0, a4m = 0, a5m = 0;
  for (int k = 0; k < ZSLS; k++) {
    double K31[4], M31[4], K13[4], M13[4], ck[3], cm[3];
    lik_visco_coefficients(kz, bz, az, kx, bx, ax, region, 31, itk[k], tek[k], K31);
    lik_visco_coefficients(kz, bz, az, kx, bx, ax, region, 31, itm[k], tem[k], M31);
    lik_visco_coefficients(kx, bx, ax, kz, bz, az, region, 13, itk[k], tek[k], K13);
    lik_visco_coefficients(kx, bx, ax, kz, bz, az, region, 13, itm[k], tem[k], M13);
    coef_convolution(itk[k], dt, ck);
    coef_convolution(itm[k], dt, cm);
    const double v[12] = {K31[3] / mk, K13[3] / mk, M31[3] / mm, M13[3] / mm, ck[0], ck[1], ck[2], cm[0], cm[1], cm[2],
                          phk[k], phm[k]};
    for (int m = 0; m < 12; m++) z[12 * k + m] = (float)v[m];
    a2k += K31[1]; a3k += K31[2]; a4k += K13[1]; a5k += K13[2];
    a2m += M31[1]; a3m += M31[2]; a4m += M13[1]; a5m += M13[2];
  }
  const double sums[8] = {a2k / mk, a3k / mk, a4k / mk, a5k / mk, a2m / mm, a3m / mm, a4m / mm, a5m / mm};
  for (int m = 0; m < 8; m++) z[12 * ZSLS + m] = (float)sums[m];
  return z;
}

using CoefKey = std::array<float, 22>;

CoefKey point_coefficients(double dt, double kx, double dx, double ax, double kz, double dz, double az, int region) {
  const double bx = ax + dx / kx, bz = az + dz / kz;  // beta = alpha + d / K
  double a31[3], a13[3], al[5], cax[3], caz[3], cbx[3], cbz[3];
  lik_coefficients(kz, bz, az, kx, bx, ax, region, 31, a31);
  lik_coefficients(kx, bx, ax, kz, bz, az, region, 13, a13);
  l_coefficients(kx, bx, ax, kz, bz, az, region, al);
  coef_convolution(ax, dt, cax);
  coef_convolution(az, dt, caz);
  coef_convolution(bx, dt, cbx);
  coef_convolution(bz, dt, cbz);
  const double v[22] = {a31[0], a31[1], a31[2], a13[0], a13[1], a13[2], al[1], al[2], al[3], al[4], cax[0],
                        cax[1], cax[2], caz[0], caz[1], caz[2], cbx[0], cbx[1], cbx[2], cbz[0], cbz[1], cbz[2]};
  CoefKey k;
  for (int m = 0; m < 22; m++) k[m] = (float)v[m];
  return k;
}

}  // namespace

// ---------------------------------------------------------------------------
// state

struct PmlLayer {  // the solid's or the fluid's layer elements, packed in (phase, colour) order
  int nelem = 0;
  int ncolour[2] = {0, 0};
  int colour_start[2][Mesh::MAX_COLOURS] = {};
  int colour_count[2][Mesh::MAX_COLOURS] = {};
  int* d_ids = nullptr;
  float4* d_geom = nullptr;
  float2* d_mod = nullptr;  // solid {kappa, mu}; fluid {J/rho, 0}
  float* d_dampw = nullptr;
  unsigned int* d_cidx = nullptr;
  float* d_old = nullptr;   // previous-step field per element point (float2 solid, float fluid)
  float* d_rm = nullptr;    // memory variables: solid [3][nelem * 25] float4, fluid [3][nelem * 25] float2
  unsigned int* d_zidx = nullptr;  // solid, viscoelastic layer: Zener set per element point (NO_ZENER: none)
  float* d_zrm = nullptr;          // its memory variables: [3][nelem * 25] float4, then [3][nelem * 25] float2
  int nbnd = 0;
  int* d_bnd = nullptr;     // outer-boundary point ids (device numbering)
  long nq() const { return (long)nelem * NGLL2; }
};

struct PmlState {
  std::vector<char> region;  // [nspec] caller order: 0 outside, 1 X only, 2 Z only, 3 XZ
  int ntable = 0;
  float4* d_table = nullptr;
  PmlLayer el, ac;
  // viscoelastic attenuation in the solid layer: its Zener coefficient sets arrive by prepare_cpml_attenuation_device_
  // (needed when the handle has attenuation; the time loop refuses to run without them); the solid layer's caller
  // elements in packed order and their profiles (K_x, K_z, d_x, d_z, alpha_x, alpha_z per point) are kept until then
  int visco_needed = 0, visco_ready = 0, nztable = 0;
  float4* d_ztable = nullptr;
  double dt = 0.0;
  std::vector<int> el_order;
  std::vector<double> el_prof;  // [nq][6]
};

namespace {

void release_layer(PmlLayer& L) {
  device_free(L.d_ids); device_free(L.d_geom); device_free(L.d_mod); device_free(L.d_dampw);
  device_free(L.d_cidx); device_free(L.d_old); device_free(L.d_rm); device_free(L.d_bnd);
  device_free(L.d_zidx); device_free(L.d_zrm);
}

// greedy colouring of a layer element list (device point ids), then the packed order per phase
void order_layer(PmlLayer& L, const std::vector<int> elems[2], const std::vector<int>& ibool_dev, const Mesh* mp,
                 std::vector<int>& order) {
  std::vector<unsigned int> used((size_t)mp->nglob, 0u);
  for (int ph = 0; ph < 2; ph++) {
    const std::vector<int>& el = elems[ph];
    std::vector<int> col(el.size(), 0);
    int ncol = 0;
    std::fill(used.begin(), used.end(), 0u);
    for (size_t k = 0; k < el.size(); k++) {
      const int row = mp->sorted ? mp->pos_of_ispec[el[k]] : el[k];
      const int* ib = ibool_dev.data() + (size_t)row * NGLL2;
      unsigned int forbidden = 0;
      for (int p = 0; p < NGLL2; p++) forbidden |= used[ib[p]];
      if (forbidden == 0xffffffffu) fail("CPML layer colouring needs more than %d colours", Mesh::MAX_COLOURS);
      int c = 0;
      while (forbidden & (1u << c)) c++;
      col[k] = c;
      ncol = std::max(ncol, c + 1);
      for (int p = 0; p < NGLL2; p++) used[ib[p]] |= 1u << c;
    }
    L.ncolour[ph] = ncol;
    for (int c = 0; c < ncol; c++) {
      L.colour_start[ph][c] = (int)order.size();
      for (size_t k = 0; k < el.size(); k++)
        if (col[k] == c) order.push_back(el[k]);
      L.colour_count[ph][c] = (int)order.size() - L.colour_start[ph][c];
    }
  }
  L.nelem = (int)order.size();
}

// the set-up common to both entry points.  region, row: per element (caller order), row into the profile arrays
// (-1 outside the layer); prof[6] = K_x, K_z, d_x, d_z, alpha_x, alpha_z [rows][25] double; rho [nspec][25]
// caller order; boundary point lists 0-based caller numbering
void pml_setup(Mesh* mp, const std::vector<int>& region, const std::vector<int>& row, const double* const prof[6],
               const float* rho, double dt, const std::vector<int>& bnd_el, const std::vector<int>& bnd_ac) {
  if (mp->pml_state) fail("prepare_pml_device: the CPML layer was already prepared");
  if (mp->nspec_elastic > 0 && !mp->sorted)
    fail("prepare_pml_device: call after prepare_fields_elastic_device (the layer follows the solid's element order)");
  if (mp->simulation_type != 1) fail("CPML layers: forward simulations (SIMULATION_TYPE=1) only");
  if (mp->batch) fail("CPML layers: shot batches are not supported");
  PmlState* S = new PmlState();
  mp->pml_state = S;
  S->dt = dt;
  const int nspec = mp->nspec;
  S->region.assign(nspec, 0);
  // coefficient sets: a table of the distinct ones, an index per layer element point (caller order)
  std::map<CoefKey, unsigned int> ids;
  std::vector<float4> table;
  std::vector<unsigned int> idx((size_t)nspec * NGLL2, 0u);
  std::vector<int> layer_elems[2][2];  // [solid 0 / fluid 1][phase]
  for (int e = 0; e < nspec; e++) {
    const int r = region[e];
    if (r == 0) continue;
    if (r < 0 || r > 3 || row[e] < 0) fail("prepare_pml_device: element %d has CPML region %d, row %d", e + 1, r, row[e]);
    S->region[e] = (char)r;
    const bool fluid = !mp->h_is_acoustic.empty() && mp->h_is_acoustic[e];
    int ph;
    if (fluid) {
      ph = acoustic_phase_of(mp, e);
    } else {
      if (!mp->sorted) fail("prepare_pml_device: solid CPML element %d outside the elastic phase lists", e + 1);
      const int pos = mp->pos_of_ispec[e];
      ph = pos < mp->phase_count[0] ? 0 : pos < mp->phase_count[0] + mp->phase_count[1] ? 1 : -1;
    }
    if (ph < 0) fail("prepare_pml_device: CPML element %d is in no phase list", e + 1);
    layer_elems[fluid ? 1 : 0][ph].push_back(e);
    CoefKey last{};
    unsigned int last_id = 0;
    for (int p = 0; p < NGLL2; p++) {
      const size_t k = (size_t)row[e] * NGLL2 + p;
      const double kx = prof[0][k], kz = prof[1][k], dx = prof[2][k], dz = prof[3][k], ax = prof[4][k], az = prof[5][k];
      if (!(kx > 0.0) || !(kz > 0.0) || !std::isfinite(dx) || !std::isfinite(dz) || !std::isfinite(ax) ||
          !std::isfinite(az))
        fail("prepare_pml_device: invalid profile at element %d point %d (K %g %g, d %g %g, alpha %g %g)", e + 1, p,
             kx, kz, dx, dz, ax, az);
      const CoefKey key = point_coefficients(dt, kx, dx, ax, kz, dz, az, r);
      for (float v : key)
        if (!std::isfinite(v))
          fail("prepare_pml_device: non-finite coefficient at element %d point %d (alpha_x = alpha_z or beta_x = "
               "alpha_z in a corner element?)", e + 1, p);
      unsigned int id;
      if (p > 0 && key == last) {
        id = last_id;
      } else {
        auto it = ids.find(key);
        if (it == ids.end()) {
          it = ids.emplace(key, (unsigned int)(table.size() / SET_F4)).first;
          table.push_back(make_float4(key[0], key[1], key[2], key[3]));
          table.push_back(make_float4(key[4], key[5], key[6], key[7]));
          table.push_back(make_float4(key[8], key[9], key[10], key[11]));
          table.push_back(make_float4(key[12], key[13], key[14], key[15]));
          table.push_back(make_float4(key[16], key[17], key[18], key[19]));
          table.push_back(make_float4(key[20], key[21], 0.f, 0.f));
        }
        id = it->second;
      }
      idx[(size_t)e * NGLL2 + p] = id;
      last = key;
      last_id = id;
    }
  }
  S->ntable = (int)(table.size() / SET_F4);
  if (S->ntable == 0) return;  // no layer element: nothing to do
  S->d_table = device_copy(table.data(), table.size());

  std::vector<int> ibool_dev((size_t)nspec * NGLL2);
  check_hip(hipMemcpy(ibool_dev.data(), mp->d_ibool, ibool_dev.size() * sizeof(int), hipMemcpyDeviceToHost),
            "ibool D2H");
  auto dev_id = [&](int g) { return mp->point_id.empty() ? g : mp->point_id[g]; };
  for (int side = 0; side < 2; side++) {
    PmlLayer& L = side ? S->ac : S->el;
    std::vector<int> order;
    order_layer(L, layer_elems[side], ibool_dev, mp, order);
    const std::vector<int>& bnd = side ? bnd_ac : bnd_el;
    std::vector<char> on_edge((size_t)mp->nglob, 0);
    std::vector<int> bids;
    for (int g : bnd) {
      if (g < 0 || g >= mp->nglob) fail("prepare_pml_device: boundary point %d outside 1..%d", g + 1, mp->nglob);
      on_edge[dev_id(g)] = 1;
      bids.push_back(dev_id(g));
    }
    L.nbnd = (int)bids.size();
    if (L.nbnd) L.d_bnd = device_copy(bids.data(), bids.size());
    if (side == 0) mp->h_pml_points = bids;
    if (L.nelem == 0) continue;
    const long nq = L.nq();
    std::vector<int> pos(L.nelem);
    std::vector<float> rq((size_t)nq);
    std::vector<unsigned int> cq((size_t)nq);
    for (int k = 0; k < L.nelem; k++) {
      const int e = order[k];
      pos[k] = mp->sorted ? mp->pos_of_ispec[e] : e;
      for (int p = 0; p < NGLL2; p++) {
        const size_t q = (size_t)k * NGLL2 + p;
        rq[q] = rho[(size_t)e * NGLL2 + p];
        if (!(rq[q] > 0.f)) fail("prepare_pml_device: density %g at element %d point %d", rq[q], e + 1, p);
        cq[q] = idx[(size_t)e * NGLL2 + p] | (on_edge[ibool_dev[(size_t)pos[k] * NGLL2 + p]] ? ON_EDGE : 0u);
      }
    }
    if (side == 0 && mp->attenuation) {  // kept for the Zener coefficient sets of the viscoelastic layer
      S->visco_needed = 1;
      S->el_order = order;
      S->el_prof.resize((size_t)nq * 6);
      for (int k = 0; k < L.nelem; k++)
        for (int p = 0; p < NGLL2; p++)
          for (int m = 0; m < 6; m++)
            S->el_prof[((size_t)k * NGLL2 + p) * 6 + m] = prof[m][(size_t)row[order[k]] * NGLL2 + p];
    }
    int* d_pos = device_copy(pos.data(), pos.size());
    float* d_rho = device_copy(rq.data(), rq.size());
    L.d_ids = device_alloc<int>((size_t)nq);
    L.d_geom = device_alloc<float4>((size_t)nq);
    L.d_mod = device_alloc<float2>((size_t)nq);
    L.d_dampw = device_alloc<float>((size_t)nq);
    L.d_cidx = device_copy(cq.data(), cq.size());
    L.d_old = device_zeros<float>((size_t)nq * (side ? 1 : 2));
    L.d_rm = device_zeros<float>((size_t)nq * (side ? AC_PLANES : EL_PLANES));
    hipLaunchKernelGGL(k_pml_pack, dim3(grid_for(nq)), dim3(256), 0, mp->stream, nq, side, d_pos, d_rho, mp->d_ibool,
                       mp->d_xix, mp->d_xiz, mp->d_gammax, mp->d_gammaz, mp->d_kappav, mp->d_muv, mp->basis, L.d_ids,
                       L.d_geom, L.d_mod, L.d_dampw);
    check_hip(hipGetLastError(), "k_pml_pack launch");
    if (side == 0) {
      // the main force kernels skip the solid layer elements
      hipLaunchKernelGGL(k_pml_clear_moduli, dim3(grid_for(nq)), dim3(256), 0, mp->stream, nq, d_pos, mp->d_kappav,
                         mp->d_muv);
      check_hip(hipGetLastError(), "k_pml_clear_moduli launch");
      if (mp->d_moduli) launch_pack_elements(mp);
    }
    check_hip(hipStreamSynchronize(mp->stream), "CPML set-up");
    device_free(d_pos);
    device_free(d_rho);
  }
  // the points of the solid's outer boundary are cleared by the predictor and updated by read-modify-writes
  upload_chain_ids(mp);
  check_hip(hipDeviceSynchronize(), "prepare_pml_device");
}

}  // namespace

// ---------------------------------------------------------------------------
// hooks of the time step

bool pml_element(const Mesh* mp, int ispec) {
  return mp->pml_state && !mp->pml_state->region.empty() && mp->pml_state->region[ispec] != 0;
}

void pml_release(Mesh* mp) {
  PmlState* S = mp->pml_state;
  if (!S) return;
  release_layer(S->el);
  release_layer(S->ac);
  device_free(S->d_table);
  device_free(S->d_ztable);
  delete S;
  mp->pml_state = nullptr;
}

void pml_reset(Mesh* mp) {
  PmlState* S = mp->pml_state;
  if (!S) return;
  for (PmlLayer* L : {&S->el, &S->ac}) {
    if (L->nelem == 0) continue;
    const bool fluid = L == &S->ac;
    check_hip(hipMemsetAsync(L->d_old, 0, (size_t)L->nq() * (fluid ? 1 : 2) * sizeof(float), mp->stream), "memset");
    check_hip(hipMemsetAsync(L->d_rm, 0, (size_t)L->nq() * (fluid ? AC_PLANES : EL_PLANES) * sizeof(float),
                             mp->stream), "memset");
    if (L->d_zrm)
      check_hip(hipMemsetAsync(L->d_zrm, 0, (size_t)L->nq() * ZPLANES * sizeof(float), mp->stream), "memset");
  }
}

void pml_elastic_before_predictor(Mesh* mp, float deltatsqover2) {
  PmlState* S = mp->pml_state;
  if (!S || S->el.nelem == 0) return;
  PmlLayer& L = S->el;
  hipLaunchKernelGGL(k_pml_old_elastic, dim3(grid_for(L.nq())), dim3(256), 0, mp->stream, L.nq(), L.d_ids, L.d_cidx,
                     (const float2*)mp->d_displ, (const float2*)mp->d_accel, mp->d_rmassx, mp->d_rmassz,
                     mp->pending_corrector[0], 0.5f * deltatsqover2, (float2*)L.d_old);
  check_hip(hipGetLastError(), "k_pml_old_elastic launch");
}

void pml_elastic_forces(Mesh* mp, int ph) {
  PmlState* S = mp->pml_state;
  if (!S || S->el.nelem == 0) return;
  PmlLayer& L = S->el;
  if (S->visco_needed && !S->visco_ready)
    fail("CPML with viscoelastic attenuation: call prepare_cpml_attenuation_device before the time loop");
  for (int c = 0; c < L.ncolour[ph]; c++) {
    const int n = L.colour_count[ph][c];
    if (n <= 0) continue;
    if (L.d_zidx)
      hipLaunchKernelGGL(k_pml_elastic_forces<true>, dim3((n + PML_EPB - 1) / PML_EPB), dim3(256), 0, mp->stream,
                         L.colour_start[ph][c], n, L.nq(), L.d_ids, L.d_geom, L.d_mod, L.d_dampw, L.d_cidx, S->d_table,
                         (const float2*)L.d_old, L.d_rm, (const float2*)mp->d_displ, (const float2*)mp->d_veloc,
                         (float2*)mp->d_accel, mp->basis, L.d_zidx, S->d_ztable, L.d_zrm);
    else
      hipLaunchKernelGGL(k_pml_elastic_forces<false>, dim3((n + PML_EPB - 1) / PML_EPB), dim3(256), 0, mp->stream,
                         L.colour_start[ph][c], n, L.nq(), L.d_ids, L.d_geom, L.d_mod, L.d_dampw, L.d_cidx, S->d_table,
                         (const float2*)L.d_old, L.d_rm, (const float2*)mp->d_displ, (const float2*)mp->d_veloc,
                         (float2*)mp->d_accel, mp->basis, nullptr, nullptr, nullptr);
    check_hip(hipGetLastError(), "k_pml_elastic_forces launch");
  }
}

void pml_acoustic_before_predictor(Mesh* mp, const float* pot, const float* pdd, float deltatsqover2) {
  PmlState* S = mp->pml_state;
  if (!S || S->ac.nelem == 0) return;
  PmlLayer& L = S->ac;
  hipLaunchKernelGGL(k_pml_old_acoustic, dim3(grid_for(L.nq())), dim3(256), 0, mp->stream, L.nq(), L.d_ids, L.d_cidx,
                     pot, pdd, 0.5f * deltatsqover2, L.d_old);
  check_hip(hipGetLastError(), "k_pml_old_acoustic launch");
}

void pml_acoustic_forces(Mesh* mp, int ph, const float* pot, const float* pd, float* pdd) {
  PmlState* S = mp->pml_state;
  if (!S || S->ac.nelem == 0) return;
  PmlLayer& L = S->ac;
  for (int c = 0; c < L.ncolour[ph]; c++) {
    const int n = L.colour_count[ph][c];
    if (n <= 0) continue;
    hipLaunchKernelGGL(k_pml_acoustic_forces, dim3((n + PML_EPB - 1) / PML_EPB), dim3(256), 0, mp->stream,
                       L.colour_start[ph][c], n, L.nq(), L.d_ids, L.d_geom, L.d_mod, L.d_dampw, L.d_cidx, S->d_table,
                       L.d_old, L.d_rm, pot, pd, pdd, mp->basis);
    check_hip(hipGetLastError(), "k_pml_acoustic_forces launch");
  }
}

void pml_acoustic_boundary(Mesh* mp, float* pot, float* pd, float* pdd) {
  PmlState* S = mp->pml_state;
  if (!S || S->ac.nbnd == 0) return;
  hipLaunchKernelGGL(k_pml_zero_acoustic, dim3((S->ac.nbnd + 255) / 256), dim3(256), 0, mp->stream, S->ac.nbnd,
                     S->ac.d_bnd, pot, pd, pdd);
  check_hip(hipGetLastError(), "k_pml_zero_acoustic launch");
}

}  // namespace sf2d

using namespace sf2d;

extern "C" {

// prepare_mesh_constants_cuda.cu:843 (called by prepare_gpu.f90:216): the reference's GPU CPML set-up, which
// carries acoustic layers only.  Its arguments describe the profiles as the reference's GPU kernel rebuilds them
// (Kernel_2_acoustic_impl.cu:497-505): K = 1 everywhere; in X-only and Z-only elements alpha = ALPHA_MAX_PML (1 -
// a), d = 2 d0_max a^2 with a the normalised distance into the layer (abs_normalized); in XZ elements alpha and
// beta = alpha + d along both axes from the *_store_GPU arrays.  spec_to_pml numbers the X-only elements first,
// then the Z-only, then the XZ ones (prepare_gpu.f90:393-420).
void prepare_pml_device_(long* Mesh_pointer, int* NSPEC_PML, int* NSPEC_PML_X, int* NSPEC_PML_Z, int* NSPEC_PML_XZ,
                         int* h_spec_to_pml, realw* h_abs_normalized, realw* ALPHA_MAX_PML, realw* d0_max,
                         realw* deltat, realw* h_alphax_store, realw* h_alphaz_store, realw* h_betax_store,
                         realw* h_betaz_store, int* PML_nglob_abs_acoustic_f, int* h_PML_abs_points_acoustic) {
  Mesh* mp = handle(Mesh_pointer);
  if (!mp->pml) return;  // as the reference: nothing unless PML_BOUNDARY_CONDITIONS was set at prepare time
  const int n = *NSPEC_PML, nx = *NSPEC_PML_X, nz = *NSPEC_PML_Z, nxz = *NSPEC_PML_XZ;
  if (n < 0 || nx < 0 || nz < 0 || nxz < 0 || nx + nz + nxz != n)
    fail("prepare_pml_device: NSPEC_PML %d != X %d + Z %d + XZ %d", n, nx, nz, nxz);
  const size_t np = (size_t)std::max(1, n) * NGLL2;
  std::vector<double> prof[6];
  for (auto& v : prof) v.assign(np, 0.0);
  std::vector<int> region(mp->nspec, 0), row(mp->nspec, -1);
  for (int e = 0; e < mp->nspec; e++) {
    const int k = h_spec_to_pml[e];
    if (k == 0) continue;
    if (k < 0 || k > n) fail("prepare_pml_device: spec_to_pml(%d) = %d outside 0..%d", e + 1, k, n);
    const int r = k <= nx ? 1 : k <= nx + nz ? 2 : 3;
    region[e] = r;
    row[e] = k - 1;
    for (int p = 0; p < NGLL2; p++) {
      const size_t q = (size_t)(k - 1) * NGLL2 + p;
      prof[0][q] = prof[1][q] = 1.0;
      if (r == 3) {
        const size_t qx = (size_t)(k - 1 - nx - nz) * NGLL2 + p;
        prof[4][q] = h_alphax_store[qx];
        prof[5][q] = h_alphaz_store[qx];
        prof[2][q] = (double)h_betax_store[qx] - (double)h_alphax_store[qx];
        prof[3][q] = (double)h_betaz_store[qx] - (double)h_alphaz_store[qx];
      } else {
        const double a = h_abs_normalized[q];
        const double alpha = (double)*ALPHA_MAX_PML * (1.0 - a), d = 2.0 * (double)*d0_max * a * a;
        prof[r == 1 ? 4 : 5][q] = alpha;
        prof[r == 1 ? 2 : 3][q] = d;
      }
    }
    if (mp->h_is_acoustic.empty() || !mp->h_is_acoustic[e])
      fail("prepare_pml_device: element %d is an elastic CPML element; the reference's GPU set-up carries acoustic "
           "layers only (use prepare_cpml_device)", e + 1);
  }
  std::vector<float> rho((size_t)mp->nspec * NGLL2, 1.f);
  for (int e = 0; e < mp->nspec; e++)
    if (region[e]) acoustic_rho_row(mp, e, rho.data() + (size_t)e * NGLL2);
  std::vector<int> bnd_ac;
  for (int k = 0; k < *PML_nglob_abs_acoustic_f; k++) bnd_ac.push_back(h_PML_abs_points_acoustic[k] - 1);
  const double* const pp[6] = {prof[0].data(), prof[1].data(), prof[2].data(), prof[3].data(), prof[4].data(),
                               prof[5].data()};
  pml_setup(mp, region, row, pp, rho.data(), (double)*deltat, std::vector<int>(), bnd_ac);
}

// The CPML layer in the CPU path's own terms (pml_init.F90:34-1366), for elastic and acoustic elements.
void prepare_cpml_device_(long* Mesh_pointer, int* NSPEC_PML, int* region_CPML, int* spec_to_PML, double* K_x_store,
                          double* K_z_store, double* d_x_store, double* d_z_store, double* alpha_x_store,
                          double* alpha_z_store, realw* rhostore, double* DT, int* nglob_abs_elastic,
                          int* abs_points_elastic, int* nglob_abs_acoustic, int* abs_points_acoustic) {
  Mesh* mp = handle(Mesh_pointer);
  if (!mp->pml) fail("prepare_cpml_device: PML_BOUNDARY_CONDITIONS was not set in prepare_constants_device");
  const int n = *NSPEC_PML;
  std::vector<int> region(mp->nspec, 0), row(mp->nspec, -1);
  for (int e = 0; e < mp->nspec; e++) {
    const int k = spec_to_PML[e];
    if (k == 0 && region_CPML[e] == 0) continue;
    if (k < 1 || k > n || region_CPML[e] < 1 || region_CPML[e] > 3)
      fail("prepare_cpml_device: element %d has spec_to_PML %d, region %d (NSPEC_PML %d)", e + 1, k, region_CPML[e], n);
    region[e] = region_CPML[e];
    row[e] = k - 1;
  }
  std::vector<int> bnd_el, bnd_ac;
  for (int k = 0; k < *nglob_abs_elastic; k++) bnd_el.push_back(abs_points_elastic[k] - 1);
  for (int k = 0; k < *nglob_abs_acoustic; k++) bnd_ac.push_back(abs_points_acoustic[k] - 1);
  const double* const pp[6] = {K_x_store, K_z_store, d_x_store, d_z_store, alpha_x_store, alpha_z_store};
  pml_setup(mp, region, row, pp, rhostore, *DT, bnd_el, bnd_ac);
}

// pml_boundary_elastic (pml_compute.f90:565-605): displ, veloc and accel are zero on the outer boundary of the
// layer's elastic elements
void pml_boundary_elastic_cuda_(long* Mesh_pointer, int* compute_wavefield_1, int* compute_wavefield_2) {
  Mesh* mp = handle(Mesh_pointer);
  PmlState* S = mp->pml_state;
  if (!S || S->el.nbnd == 0) return;
  if (*compute_wavefield_2) fail("pml_boundary_elastic_cuda: backward wavefields with CPML are not supported");
  if (!*compute_wavefield_1) return;
  flush_correctors(mp);
  hipLaunchKernelGGL(k_pml_zero_elastic, dim3((S->el.nbnd + 255) / 256), dim3(256), 0, mp->stream, S->el.nbnd,
                     S->el.d_bnd, (float2*)mp->d_displ, (float2*)mp->d_veloc, (float2*)mp->d_accel);
  check_hip(hipGetLastError(), "k_pml_zero_elastic launch");
}

// The viscoelastic layer: the Zener parameters of the points (attenuation_model.f90:145-161 as the CPU path stores
// them) for the attenuating solid layer elements.  Arrays in the reference's layout: inv_tau_sigma_nu*, phi_nu*,
// tau_epsilon_nu* (NGLLX, NGLLZ, nspec, N_SLS), Mu_nu*, qkappa/qmu_attenuation_store (NGLLX, NGLLZ, nspec).
void prepare_cpml_attenuation_device_(long* Mesh_pointer, int* N_SLS, realw* inv_tau_sigma_nu1, realw* phi_nu1,
                                      realw* tau_epsilon_nu1, realw* Mu_nu1, realw* inv_tau_sigma_nu2, realw* phi_nu2,
                                      realw* tau_epsilon_nu2, realw* Mu_nu2, realw* qkappa_attenuation_store,
                                      realw* qmu_attenuation_store) {
  Mesh* mp = handle(Mesh_pointer);
  PmlState* S = mp->pml_state;
  if (!S) fail("prepare_cpml_attenuation_device: call after prepare_cpml_device");
  if (!mp->attenuation) fail("prepare_cpml_attenuation_device: the handle has no viscoelastic attenuation");
  if (*N_SLS != ZSLS) fail("prepare_cpml_attenuation_device: N_SLS = %d (this library: %d)", *N_SLS, ZSLS);
  if (S->visco_ready) fail("prepare_cpml_attenuation_device: already prepared");
  PmlLayer& L = S->el;
  S->visco_ready = 1;
  if (L.nelem == 0) return;
  const size_t ne = (size_t)mp->nspec * NGLL2;
  const long nq = L.nq();
  std::map<ZenerKey, unsigned int> ids;
  std::vector<float4> table;
  std::vector<unsigned int> zidx((size_t)nq, NO_ZENER);
  for (int k = 0; k < L.nelem; k++) {
    const int e = S->el_order[k];
    // the element rule of compute_forces_viscoelastic.F90:178-181, 249-252: both quality factors below 9999 at every
    // point, else the elastic layer formulation
    float qk = 0.f, qm = 0.f;
    for (int p = 0; p < NGLL2; p++) {
      qk = std::max(qk, qkappa_attenuation_store[(size_t)e * NGLL2 + p]);
      qm = std::max(qm, qmu_attenuation_store[(size_t)e * NGLL2 + p]);
    }
    if (!(qk < 9998.999f && qm < 9998.999f)) continue;
    for (int p = 0; p < NGLL2; p++) {
      const size_t l = (size_t)e * NGLL2 + p;
      double itk[ZSLS], tek[ZSLS], phk[ZSLS], itm[ZSLS], tem[ZSLS], phm[ZSLS];
      for (int s2 = 0; s2 < ZSLS; s2++) {
        itk[s2] = inv_tau_sigma_nu1[l + ne * s2];
        tek[s2] = tau_epsilon_nu1[l + ne * s2];
        phk[s2] = phi_nu1[l + ne * s2];
        itm[s2] = inv_tau_sigma_nu2[l + ne * s2];
        tem[s2] = tau_epsilon_nu2[l + ne * s2];
        phm[s2] = phi_nu2[l + ne * s2];
        if (!(itk[s2] > 0.0) || !(itm[s2] > 0.0) || !std::isfinite(tek[s2]) || !std::isfinite(tem[s2]) ||
            !std::isfinite(phk[s2]) || !std::isfinite(phm[s2]))
          fail("prepare_cpml_attenuation_device: invalid Zener parameters at element %d point %d", e + 1, p + 1);
      }
      const double mk = (double)Mu_nu1[l] * (double)ZSLS, mm = (double)Mu_nu2[l] * (double)ZSLS;
      if (!(mk > 0.0) || !(mm > 0.0)) fail("prepare_cpml_attenuation_device: Mu_nu <= 0 at element %d", e + 1);
      const ZenerKey key = zener_coefficients(S->dt, &S->el_prof[((size_t)k * NGLL2 + p) * 6], S->region[e], itk,
                                              tek, phk, mk, itm, tem, phm, mm);
      for (float v : key)
        if (!std::isfinite(v)) fail("prepare_cpml_attenuation_device: non-finite coefficient at element %d", e + 1);
      auto it = ids.find(key);
      if (it == ids.end()) {
        it = ids.emplace(key, (unsigned int)(table.size() / ZSET_F4)).first;
        for (int m = 0; m < ZSET_F4; m++)
          table.push_back(make_float4(key[4 * m], key[4 * m + 1], key[4 * m + 2], key[4 * m + 3]));
      }
      zidx[(size_t)k * NGLL2 + p] = it->second;
    }
  }
  S->el_prof.clear();
  S->el_prof.shrink_to_fit();
  S->nztable = (int)(table.size() / ZSET_F4);
  if (S->nztable == 0) return;  // no attenuating layer element: the elastic layer kernel serves
  S->d_ztable = device_copy(table.data(), table.size());
  L.d_zidx = device_copy(zidx.data(), zidx.size());
  L.d_zrm = device_zeros<float>((size_t)nq * ZPLANES);
  check_hip(hipDeviceSynchronize(), "prepare_cpml_attenuation_device");
}

// layer statistics: elements (solid, fluid), coefficient sets, outer-boundary points (solid, fluid)
void sf2d_pml_stats(long* Mesh_pointer, int* stats) {
  Mesh* mp = handle(Mesh_pointer);
  const PmlState* S = mp->pml_state;
  for (int k = 0; k < 5; k++) stats[k] = 0;
  if (!S) return;
  stats[0] = S->el.nelem;
  stats[1] = S->ac.nelem;
  stats[2] = S->ntable;
  stats[3] = S->el.nbnd;
  stats[4] = S->ac.nbnd;
}

}  // extern "C"
