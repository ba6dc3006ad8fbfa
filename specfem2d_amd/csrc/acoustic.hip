// Acoustic (fluid) elements on MI355X: the potential formulation of SPECFEM2D's fluid, its free surface, its
// Stacey edges, pressure sources, pressure seismograms and the fluid-solid coupling terms, behind the reference's
// Fortran-callable entry points (src/gpu/compute_forces_acoustic_cuda.cu, compute_stacey_acoustic_cuda.cu,
// enforce_acoustic_free_surface_cuda.cu, compute_add_sources_viscoacoustic_cuda.cu, compute_coupling_cuda.cu,
// update_displacement_cuda.cu, transfer_fields_cuda.cu, prepare_mesh_constants_cuda.cu:319-436).
//
// Layout.  The fluid's state lives beside the solid's in the same point numbering (the device numbering of the
// handle, see points_to_device): potential, its first and second time derivatives and the inverse fluid mass,
// one float per point.  The fluid elements are coloured on their own (greedy, no two elements of a colour share
// a point) and their per-point operands are packed once in colour order -- {xix, xiz, gammax, gammaz}, J/rho and
// the point id -- so one colour launch streams them with unit stride and assembles with plain read-modify-writes.
// Boundary terms (free surface, Stacey, fluid-solid edges, sources, receivers) are flattened at set-up into
// point lists with their weights, so each is one small launch that reads only what it updates.  The Newmark
// updates visit only the points of fluid elements (an id range plus a short list, see ac_ready): elsewhere the
// potential has no source of change and keeps the value last transferred (zero in any physical state).
//
// Set-up is two-phase: the entry points of the preparation record the caller's arrays; the first call of the
// time loop resolves them into device ids (after the solid's preparation has fixed the numbering) and uploads
// them.  Unsupported cases (backward or adjoint fluid wavefields, attenuation, MPI interfaces, SAVE_FORWARD
// storage of the fluid's Stacey terms) stop the run with a message, as the reference does for its own limits.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "sf2d_mesh.h"
#include "specfem2d_gpu.h"

namespace sf2d {

namespace {
constexpr int AC_EPB = 10;  // fluid elements per 256-thread workgroup (250 active lanes)

__device__ __forceinline__ void atomic_add(float* p, float v) { unsafeAtomicAdd(p, v); }

// One colour of fluid elements: pdd -= K pot over elements first .. first+n-1 of the packed colour order
// (compute_forces_viscoacoustic.f90:107-330; Kernel_2_acoustic_impl in src/gpu/kernels).  RECT: every element is
// an axis-aligned rectangle (xi_z = gamma_x = 0, xi_x and gamma_z constant over the element), whose geometry is one
// float2 {xi_x, gamma_z} per element instead of a float4 per point.
// Threads go GLL row by GLL row across the workgroup's nb elements (thread t: row j = t / 5 nb, element
// (t % 5 nb) / 5, column t % 5), and the packed per-point arrays follow the same order (row_block_pos), so that a
// wave's gather of the potential reads GLL row j of ten neighbouring elements: a few cache lines instead of one or
// two per element.
template <bool RECT>
__global__ __launch_bounds__(256) void k_acoustic_forces(int first, int n, const int* __restrict__ ids,
                                                         const float4* __restrict__ geom,
                                                         const float2* __restrict__ geom_rect,
                                                         const float* __restrict__ fac, Basis b,
                                                         const float* __restrict__ pot, float* __restrict__ pdd) {
  __shared__ float s_u[AC_EPB][NGLL2];
  __shared__ float s_t1[AC_EPB][NGLL2];
  __shared__ float s_t2[AC_EPB][NGLL2];
  __shared__ float s_hp[NGLL2], s_hpw[NGLL2], s_w[NGLL];  // the basis, read per thread from LDS (not kernarg memory)
  const int t = threadIdx.x;
  if (t < NGLL2) {
    s_hp[t] = b.hp[t];
    s_hpw[t] = b.hpw[t];
  }
  if (t < NGLL) s_w[t] = b.w[t];
  const int e0 = first + blockIdx.x * AC_EPB;  // the workgroup's first element
  const int nb = min(AC_EPB, first + n - e0);  // its elements
  const int rw = NGLL * nb;                    // threads per GLL row
  const bool active = t < NGLL2 * nb;
  const int j = active ? t / rw : 0;
  const int le = active ? (t - j * rw) / NGLL : 0;
  const int i = active ? t % NGLL : 0;
  const int p = j * NGLL + i;
  const int e = e0 + le;
  int g = 0;
  float4 gm = make_float4(0.f, 0.f, 0.f, 0.f);
  float f = 0.f, pdd_old = 0.f;
  if (active) {
    const size_t l = (size_t)e0 * NGLL2 + t;  // row-block order
    g = ids[l];
    pdd_old = pdd[g];  // no other element of this colour has the point: read with the potential, in one round trip
    if (RECT) {
      const float2 r = geom_rect[e];
      gm = make_float4(r.x, 0.f, 0.f, r.y);
    } else {
      gm = geom[l];
    }
    f = fac[l];
    s_u[le][p] = pot[g];
  }
  __syncthreads();
  if (active) {
    const float* u = s_u[le];
    const float dxi = u[j * NGLL + 0] * s_hp[i * NGLL + 0] + u[j * NGLL + 1] * s_hp[i * NGLL + 1] +
                      u[j * NGLL + 2] * s_hp[i * NGLL + 2] + u[j * NGLL + 3] * s_hp[i * NGLL + 3] +
                      u[j * NGLL + 4] * s_hp[i * NGLL + 4];
    const float dga = u[0 * NGLL + i] * s_hp[j * NGLL + 0] + u[1 * NGLL + i] * s_hp[j * NGLL + 1] +
                      u[2 * NGLL + i] * s_hp[j * NGLL + 2] + u[3 * NGLL + i] * s_hp[j * NGLL + 3] +
                      u[4 * NGLL + i] * s_hp[j * NGLL + 4];
    if (RECT) {
      s_t1[le][p] = f * (gm.x * (dxi * gm.x));
      s_t2[le][p] = f * (gm.w * (dga * gm.w));
    } else {
      const float dxl = dxi * gm.x + dga * gm.z;  // d/dx: xix, gammax
      const float dzl = dxi * gm.y + dga * gm.w;  // d/dz: xiz, gammaz
      s_t1[le][p] = f * (gm.x * dxl + gm.y * dzl);
      s_t2[le][p] = f * (gm.z * dxl + gm.w * dzl);
    }
  }
  __syncthreads();
  if (active) {
    float a1 = 0.f, a2 = 0.f;
#pragma unroll
    for (int k = 0; k < NGLL; k++) {
      a1 = a1 + s_t1[le][j * NGLL + k] * s_hpw[k * NGLL + i];
      a2 = a2 + s_t2[le][k * NGLL + i] * s_hpw[k * NGLL + j];
    }
    pdd[g] = pdd_old - (s_w[j] * a1 + s_w[i] * a2);
  }
}

// Chains of fluid elements, lanes along GLL columns (the scalar twin of the solid's column-lane kernel): one wave per
// chain of up to 12 elements, lane 5 k + i holding column i of element k; gamma contractions in registers, xi
// contractions through two LDS transposes; the shared column's term moves to the next element's lane by a one-lane
// shuffle, and each point of the chain is updated by one lane with a read-modify-write of pdd (no point is shared
// between the chains of one colour).  Same element operator as k_acoustic_forces.
constexpr int AC_CHAIN = 12;
constexpr int AC_NXCD = 8;  // XCD-contiguous blocks of chains, as the solid's kernels
template <bool RECT>
__global__ __launch_bounds__(64) void k_acoustic_forces_col(int first_patch, int npatch,
                                                            const int4* __restrict__ desc,
                                                            const int* __restrict__ ids,
                                                            const float* __restrict__ fac,
                                                            const float4* __restrict__ geom,
                                                            const float2* __restrict__ rect, Basis b,
                                                            const float* __restrict__ pot,
                                                            float* __restrict__ pdd) {
  __shared__ float s_t[AC_CHAIN + 1][NGLL][NGLL + 1];
  __shared__ float s_v[AC_CHAIN + 1][NGLL][NGLL + 1];
  __shared__ float s_hp[NGLL2], s_hpw[NGLL2], s_w[NGLL];
  const int lane = threadIdx.x;
  if (lane < NGLL2) {
    s_hp[lane] = b.hp[lane];
    s_hpw[lane] = b.hpw[lane];
  }
  if (lane < NGLL) s_w[lane] = b.w[lane];
  const int nw = gridDim.x;
  const int item = (blockIdx.x % AC_NXCD) * (nw / AC_NXCD) + blockIdx.x / AC_NXCD;
  if (item >= npatch) return;  // uniform over the wave
  const int4 d = desc[first_patch + item];
  const int ne = d.w & 0xf;
  const int k = min(lane / NGLL, AC_CHAIN);
  const int i = lane % NGLL;
  const bool active = lane / NGLL < ne;
  const int e = d.x + min(k, ne - 1);
  const int row = NGLL * ne;
  const size_t c0 = (size_t)d.x * NGLL2 + (size_t)min(lane, row - 1);
  const bool writer = active && (i < NGLL - 1 || k == ne - 1);
  int g[NGLL];
  float u[NGLL], f[NGLL], old[NGLL];
  float4 gm[NGLL];
#pragma unroll
  for (int j = 0; j < NGLL; j++) {
    g[j] = ids[c0 + (size_t)j * row];
    f[j] = fac[c0 + (size_t)j * row];
    u[j] = pot[g[j]];
    old[j] = writer ? pdd[g[j]] : 0.f;  // no other chain of this colour has the point
    if (RECT) {
      const float2 r = rect[e];
      gm[j] = make_float4(r.x, 0.f, 0.f, r.y);
    } else {
      gm[j] = geom[c0 + (size_t)j * row];
    }
  }
  // xi-derivatives through the tile: this lane takes row r = i
#pragma unroll
  for (int j = 0; j < NGLL; j++) s_t[k][j][i] = u[j];
  __syncthreads();
  {
    const int r = i;
    float rw[NGLL];
#pragma unroll
    for (int m = 0; m < NGLL; m++) rw[m] = s_t[k][r][m];
#pragma unroll
    for (int c = 0; c < NGLL; c++) {
      float dx = 0.f;
#pragma unroll
      for (int m = 0; m < NGLL; m++) dx += rw[m] * s_hp[c * NGLL + m];
      s_v[k][r][c] = dx;
    }
  }
  __syncthreads();
  float t1[NGLL], t2[NGLL];
#pragma unroll
  for (int j = 0; j < NGLL; j++) {
    float dga = 0.f;
#pragma unroll
    for (int m = 0; m < NGLL; m++) dga += u[m] * s_hp[j * NGLL + m];
    const float dxi = s_v[k][j][i];
    if (RECT) {
      t1[j] = f[j] * (gm[j].x * (dxi * gm[j].x));
      t2[j] = f[j] * (gm[j].w * (dga * gm[j].w));
    } else {
      const float dxl = dxi * gm[j].x + dga * gm[j].z;
      const float dzl = dxi * gm[j].y + dga * gm[j].w;
      t1[j] = f[j] * (gm[j].x * dxl + gm[j].y * dzl);
      t2[j] = f[j] * (gm[j].z * dxl + gm[j].w * dzl);
    }
  }
  __syncthreads();  // every lane has read its derivatives
#pragma unroll
  for (int j = 0; j < NGLL; j++) s_t[k][j][i] = t1[j];
  __syncthreads();
  {
    const int r = i;
    float rw[NGLL];
#pragma unroll
    for (int m = 0; m < NGLL; m++) rw[m] = s_t[k][r][m];
#pragma unroll
    for (int c = 0; c < NGLL; c++) {
      float sx = 0.f;
#pragma unroll
      for (int m = 0; m < NGLL; m++) sx += rw[m] * s_hpw[m * NGLL + c];
      s_v[k][r][c] = sx;
    }
  }
  __syncthreads();
  float fo[NGLL];
  const float wi = s_w[i];
#pragma unroll
  for (int j = 0; j < NGLL; j++) {
    float a2 = 0.f;
#pragma unroll
    for (int m = 0; m < NGLL; m++) a2 += t2[m] * s_hpw[m * NGLL + j];
    fo[j] = s_w[j] * s_v[k][j][i] + wi * a2;
  }
#pragma unroll
  for (int j = 0; j < NGLL; j++) {
    const float fl = __shfl_up(fo[j], 1, 64);
    if (i == 0 && k > 0) fo[j] += fl;
  }
  if (writer) {
#pragma unroll
    for (int j = 0; j < NGLL; j++) pdd[g[j]] = old[j] - fo[j];
  }
}

// Newmark predictor on the potential (update_displacement_Newmark.F90:190-240) over the fluid's points: the id
// range [r0, r0 + n) and the listed ids below it (list[0 .. nlist))
__global__ __launch_bounds__(256) void k_ac_predictor(float* __restrict__ pot, float* __restrict__ pd,
                                                      float* __restrict__ pdd, int r0, int n,
                                                      const int* __restrict__ list, int nlist, float dt, float dt2o2,
                                                      float dto2) {
  const int stride = gridDim.x * blockDim.x;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n + nlist; q += stride) {
    const int k = q < n ? r0 + q : list[q - n];
    const float a = pdd[k], v = pd[k];
    pot[k] = pot[k] + dt * v + dt2o2 * a;
    pd[k] = v + dto2 * a;
    pdd[k] = 0.f;
  }
}

// mass and corrector (compute_forces_viscoacoustic_calling_routine.F90:183-205, update_veloc_acoustic_Newmark),
// same point set
__global__ __launch_bounds__(256) void k_ac_corrector(float* __restrict__ pd, float* __restrict__ pdd,
                                                      const float* __restrict__ rm, int r0, int n,
                                                      const int* __restrict__ list, int nlist, float dto2) {
  const int stride = gridDim.x * blockDim.x;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n + nlist; q += stride) {
    const int k = q < n ? r0 + q : list[q - n];
    const float a = pdd[k] * rm[k];
    pdd[k] = a;
    pd[k] = pd[k] + dto2 * a;
  }
}

// enforce_acoustic_free_surface.f90:34-80
__global__ __launch_bounds__(256) void k_ac_free_surface(const int* __restrict__ pts, int n, float* __restrict__ pot,
                                                         float* __restrict__ pd, float* __restrict__ pdd) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int g = pts[k];
  pot[g] = 0.f;
  pd[g] = 0.f;
  pdd[g] = 0.f;
}

// compute_stacey_acoustic.f90:34-160: pdd -= pd w / (rho vp) along the fluid's absorbing edges; with store, each
// term also goes to its edge slot (the SAVE_FORWARD copy b_absorb_acoustic_*, :46-48, 65-67, 86-88, 107-109)
__global__ __launch_bounds__(256) void k_ac_stacey(int n, const int* __restrict__ pts, const float* __restrict__ w,
                                                   const float* __restrict__ rvp, const float* __restrict__ pd,
                                                   float* __restrict__ pdd, const int* __restrict__ slot,
                                                   float* __restrict__ store) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int g = pts[k];
  const float t = pd[g] * w[k] / rvp[k];
  atomic_add(&pdd[g], -t);
  if (store) store[slot[k]] = t;
}

// reconstructed field: b_pdd -= the stored term of step NSTEP - it + 1 (compute_stacey_acoustic.f90:162-240)
__global__ __launch_bounds__(256) void k_ac_stacey_backward(int n, const int* __restrict__ pts,
                                                            const int* __restrict__ slot,
                                                            const float* __restrict__ store, float* __restrict__ bpdd) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  atomic_add(&bpdd[pts[k]], -store[slot[k]]);
}

// fluid side of a fluid-solid edge: pdd += w (u . n) (compute_coupling_acoustic_el.f90:34-150); sign -1 gives the
// adjoint run's term, u = minus the adjoint acceleration (:104-109)
__global__ __launch_bounds__(256) void k_coupling_ac_el(int n, const int* __restrict__ pts,
                                                        const float2* __restrict__ nrm, const float* __restrict__ w,
                                                        const float2* __restrict__ displ, float sign,
                                                        float* __restrict__ pdd) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int g = pts[k];
  const float2 u = displ[g], nv = nrm[k];
  const float dn = (sign * u.x) * nv.x + (sign * u.y) * nv.y;
  atomic_add(&pdd[g], w[k] * dn);
}

// solid side: accel += w n p with the fluid pressure p = -pdd (compute_coupling_viscoelastic_ac.f90:34-130); with
// sign -1 and the adjoint potential in place of pdd, the adjoint run's pressure p = potential (:98-102)
__global__ __launch_bounds__(256) void k_coupling_el_ac(int n, const int* __restrict__ pts,
                                                        const float2* __restrict__ nrm, const float* __restrict__ w,
                                                        const float* __restrict__ pdd, float sign,
                                                        float* __restrict__ accel) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int g = pts[k];
  const float pr = sign * -pdd[g];
  const float2 nv = nrm[k];
  const float wk = w[k];
  atomic_add(&accel[2 * (size_t)g], wk * nv.x * pr);
  atomic_add(&accel[2 * (size_t)g + 1], wk * nv.y * pr);
}

// pressure sources in fluid elements: pdd += h stf / kappa (compute_add_sources_acoustic.f90:34-75)
__global__ __launch_bounds__(64) void k_ac_sources(int nas, int it, int nsrc, const int* __restrict__ src,
                                                   const int* __restrict__ pts, const float* __restrict__ sa,
                                                   const float* __restrict__ kap, const float* __restrict__ stf,
                                                   float* __restrict__ pdd) {
  const int s = blockIdx.x, p = threadIdx.x;
  if (s >= nas || p >= NGLL2) return;
  const size_t l = (size_t)s * NGLL2 + p;
  const float f = stf[(size_t)it * nsrc + src[s]];
  atomic_add(&pdd[pts[l]], sa[l] * f / kap[l]);
}

// adjoint sources of receivers in fluid elements: pdd += h_i h_j adj(NSTEP - it + 1), without a 1/kappa factor
// (compute_add_sources_acoustic.f90:247-290); adj is source_adjoint(nrec, NSTEP, 2), component 1; rows of
// receivers in solid elements (pts = -1) are skipped
__global__ __launch_bounds__(64) void k_ac_adjoint_sources(int nrec, int it, int nstep, const int* __restrict__ pts,
                                                           const float* __restrict__ xir,
                                                           const float* __restrict__ gammar,
                                                           const float* __restrict__ adj, float* __restrict__ pdd) {
  const int r = blockIdx.x, p = threadIdx.x;
  if (r >= nrec || p >= NGLL2) return;
  const int g = pts[(size_t)r * NGLL2 + p];
  if (g < 0) return;
  const int i = p % NGLL, j = p / NGLL;
  const float h = xir[i * nrec + r] * gammar[j * nrec + r];
  atomic_add(&pdd[g], h * adj[(size_t)it * nrec + r]);
}

// fluid kernels (compute_kernels.f90:293-449, NTSTEP_BETWEEN_COMPUTE_KERNELS = 1): per point of a fluid element the
// gradients of the adjoint and of the reconstructed potential (xi/gamma derivatives in single precision, the
// jacobian products in double), rho_kl += 1/rho (grad pot . grad b_pot) dt, kappa_kl += 1/kappa pdd b_pot dt.
// One block of 256 threads takes 10 elements (rows of the fluid elements, 25 points each); products and sums are
// rounded one at a time as the CPU path forms them.
constexpr int KL_EPB = 10;
__global__ __launch_bounds__(256) void k_ac_kernels(int nrows, const int* __restrict__ ids,
                                                    const float4* __restrict__ geom, const float2* __restrict__ rk,
                                                    const Basis basis, const float* __restrict__ pot,
                                                    const float* __restrict__ pdd, const float* __restrict__ bpot,
                                                    float dt, float* __restrict__ rho_kl,
                                                    float* __restrict__ kappa_kl) {
  __shared__ float s_p[KL_EPB][NGLL2], s_b[KL_EPB][NGLL2], s_h[NGLL2];
  const int t = threadIdx.x, le = t / NGLL2, p = t - le * NGLL2;
  const long r = (long)blockIdx.x * KL_EPB + le;
  const bool act = le < KL_EPB && r < nrows;
  if (t < NGLL2) s_h[t] = basis.hp[t];
  int g = 0;
  if (act) {
    g = ids[r * NGLL2 + p];
    s_p[le][p] = pot[g];
    s_b[le][p] = bpot[g];
  }
  __syncthreads();
  if (!act) return;
  const int i = p % NGLL, j = p / NGLL;
  float t1 = 0.f, t2 = 0.f, u1 = 0.f, u2 = 0.f;
#pragma unroll
  for (int k = 0; k < NGLL; k++) {
    t1 = __fadd_rn(t1, __fmul_rn(s_p[le][j * NGLL + k], s_h[i * NGLL + k]));
    u1 = __fadd_rn(u1, __fmul_rn(s_b[le][j * NGLL + k], s_h[i * NGLL + k]));
    t2 = __fadd_rn(t2, __fmul_rn(s_p[le][k * NGLL + i], s_h[j * NGLL + k]));
    u2 = __fadd_rn(u2, __fmul_rn(s_b[le][k * NGLL + i], s_h[j * NGLL + k]));
  }
  const float4 G = geom[r * NGLL2 + p];  // {xi_x, xi_z, gamma_x, gamma_z}
  const float a1 = (float)__dadd_rn(__dmul_rn(t1, G.x), __dmul_rn(t2, G.z));
  const float a2 = (float)__dadd_rn(__dmul_rn(t1, G.y), __dmul_rn(t2, G.w));
  const float b1 = (float)__dadd_rn(__dmul_rn(u1, G.x), __dmul_rn(u2, G.z));
  const float b2 = (float)__dadd_rn(__dmul_rn(u1, G.y), __dmul_rn(u2, G.w));
  const float dot = __fadd_rn(__fmul_rn(a1, b1), __fmul_rn(a2, b2));
  const float2 m = rk[r * NGLL2 + p];  // {rho, kappa}
  const long l = r * NGLL2 + p;
  rho_kl[l] = __fadd_rn(rho_kl[l], __fmul_rn(__fmul_rn(1.f / m.x, dot), dt));
  kappa_kl[l] = __fadd_rn(kappa_kl[l], __fmul_rn(__fmul_rn(__fmul_rn(1.f / m.y, pdd[g]), s_b[le][p]), dt));
}

// pressure (sign = -1: -potential or -potential_dot_dot) or potential (sign = +1) seismograms of the receivers in
// fluid elements, interpolated with the receivers' Lagrange weights and accumulated in double
// (write_seismograms.F90:124-170, compute_pressure.f90:460-490, compute_interpolated_dva.f90); receivers in solid
// elements (pts = -1) record 0
// displacement, velocity or acceleration seismograms (seismotypes 1, 2, 3) of the receivers in fluid elements: the
// gradient over the density of the potential or of its first or second time derivative at each point of the element, interpolated with the receivers' Lagrange weights and rotated
// (compute_elastic_seismogram_kernel.cu:120-175, compute_vector_field.f90:143-210); one block per receiver, blocks
// of receivers in solid elements (pts = -1) leave the displacement recorded there untouched
__global__ __launch_bounds__(64) void k_ac_vector_seismograms(int nrec, int nlength, int cur,
                                                              const int* __restrict__ pts,
                                                              const float4* __restrict__ geom,
                                                              const float* __restrict__ rho, Basis b,
                                                              const float* __restrict__ xir,
                                                              const float* __restrict__ gammar,
                                                              const float* __restrict__ cosrot,
                                                              const float* __restrict__ sinrot,
                                                              const float* __restrict__ field,
                                                              float* __restrict__ dst) {
  __shared__ float s[NGLL2];
  const int r = blockIdx.x, p = threadIdx.x;
  if (r >= nrec || pts[(size_t)r * NGLL2] < 0) return;  // uniform over the block
  const size_t l = (size_t)r * NGLL2 + p;
  if (p < NGLL2) s[p] = field[pts[l]];
  __syncthreads();
  double vx = 0.0, vz = 0.0;
  if (p < NGLL2) {
    const int i = p % NGLL, j = p / NGLL;
    float t1 = 0.f, t2 = 0.f;
#pragma unroll
    for (int k = 0; k < NGLL; k++) {
      t1 += s[j * NGLL + k] * b.hp[i * NGLL + k];
      t2 += s[k * NGLL + i] * b.hp[j * NGLL + k];
    }
    const float4 g = geom[l];
    const float fx = (t1 * g.x + t2 * g.z) / rho[l];
    const float fz = (t1 * g.y + t2 * g.w) / rho[l];
    const double h = (double)(xir[i * nrec + r] * gammar[j * nrec + r]);
    vx = (double)fx * h;
    vz = (double)fz * h;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    vx += __shfl_xor(vx, off, 64);
    vz += __shfl_xor(vz, off, 64);
  }
  if (p == 0) {
    const double c = cosrot[r], sn = sinrot[r];
    dst[(size_t)r * nlength + cur] = (float)(c * vx + sn * vz);
    dst[((size_t)nrec + r) * nlength + cur] = (float)(-sn * vx + c * vz);
  }
}

__global__ __launch_bounds__(64) void k_ac_seismograms(int nrec, int nlength, int cur, const int* __restrict__ pts,
                                                       const float* __restrict__ xir, const float* __restrict__ gammar,
                                                       const float* __restrict__ field, float sign,
                                                       float* __restrict__ dst) {
  const int r = blockIdx.x, p = threadIdx.x;
  if (r >= nrec) return;
  double v = 0.0;
  if (p < NGLL2) {
    const int g = pts[(size_t)r * NGLL2 + p];
    if (g >= 0) {
      const int i = p % NGLL, j = p / NGLL;
      const double h = (double)(xir[i * nrec + r] * gammar[j * nrec + r]);
      v = (double)(sign * field[g]) * h;
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  if (p == 0) {
    dst[(size_t)r * nlength + cur] = (float)v;
    dst[((size_t)nrec + r) * nlength + cur] = 0.f;
  }
}

// max |field| as the bit pattern of a non-negative float (ordered like the value)
__global__ __launch_bounds__(256) void k_ac_max_abs(const float* __restrict__ f, int n, unsigned int* out) {
  float m = 0.f;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) m = fmaxf(m, fabsf(f[k]));
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
  if ((threadIdx.x & 63) == 0) atomicMax(out, __float_as_uint(m));
}

int grid_for(long n, int block, int cap) { return (int)std::max<long>(1, std::min<long>((n + block - 1) / block, cap)); }

}  // namespace

// ---------------------------------------------------------------------------
// state
struct AcousticState {
  int ready = 0;
  // recorded at preparation (caller numbering, 0-based elements and (i, j))
  std::vector<int> phase_elems[2];        // fluid elements of each phase list
  std::vector<float> rho_rows, kappa_rows;  // [fluid element][25], fluid elements in ispec order
  std::vector<int> row_of_ispec;          // fluid row of each element, -1 for solid elements
  std::vector<float> h_rm;                // [nglob] inverse fluid mass, caller numbering
  std::vector<int> free_e, free_ij;       // free-surface points: element, i + 5 j
  std::vector<int> coup_e, coup_ij;       // fluid-solid points: fluid element, i + 5 j
  std::vector<float> coup_n, coup_w;      // [2] normal out of the fluid, weight
  std::vector<int> st_e, st_ij;           // fluid Stacey points
  std::vector<float> st_w, st_rvp;
  std::vector<int> st_slot;               // slot of each Stacey point in the edge storage (below)
  // SAVE_FORWARD storage of the Stacey terms, the four sides' slices one after the other (bottom, right, top, left;
  // a side's slice is Fortran (NGLLX, nspec_side): slot ib * 5 + k)
  int abs_count[5] = {0, 0, 0, 0, 0}, abs_base[5] = {0, 0, 0, 0, 0}, abs_total = 0;
  int* d_st_slot = nullptr;
  float* d_absorb = nullptr;
  // device
  float *d_pot = nullptr, *d_pd = nullptr, *d_pdd = nullptr, *d_rm = nullptr;
  int ncolour[2] = {0, 0};
  int colour_start[2][Mesh::MAX_COLOURS] = {};
  int colour_count[2][Mesh::MAX_COLOURS] = {};
  int nelem = 0;
  // the fluid's points for the Newmark updates: the id range [r0, r0 + nrange) and the ids below it
  int r0 = 0, nrange = 0, nlist = 0;
  int* d_list = nullptr;
  int* d_ids = nullptr;      // [nelem][25] point ids, colour order
  float4* d_geom = nullptr;  // [nelem][25]
  float2* d_geom_rect = nullptr;  // [nelem] {xi_x, gamma_z} when every fluid element is an axis-aligned rectangle
  int rect = 0;
  float* d_fac = nullptr;    // [nelem][25] J / rho
  // chains of fluid elements (the column-lane fluid kernel): patches of the fluid phase lists, coloured; per patch
  // an int4 {first position, -, -, ne}; ids, J / rho and per-point geometry in lane order (the block of a chain of
  // ne elements at position e0 holds point (k, i, j) at e0 * 25 + j * 5 ne + 5 k + i); rectangles per element
  int chain = 0;
  int cn_colour[2] = {0, 0};
  int cc_start[2][Mesh::MAX_COLOURS] = {};
  int cc_count[2][Mesh::MAX_COLOURS] = {};
  int4* d_cdesc = nullptr;
  int* d_cids = nullptr;
  float* d_cfac = nullptr;
  float4* d_cgeom = nullptr;
  float2* d_crect = nullptr;
  int nfree = 0;
  int* d_free = nullptr;
  int ncoup = 0;
  int* d_coup = nullptr;
  float2* d_coup_n = nullptr;
  float* d_coup_w = nullptr;
  int nst = 0;
  int* d_st = nullptr;
  float *d_st_w = nullptr, *d_st_rvp = nullptr;
  int nas = 0;               // sources in fluid elements
  int *d_src = nullptr, *d_src_pts = nullptr;
  float *d_src_sa = nullptr, *d_src_kap = nullptr;
  int nrec_fluid = 0;
  int* d_rec_pts = nullptr;  // [nrec][25], -1 rows for receivers in solid elements
  float4* d_rec_geom = nullptr;  // [nrec][25] {xi_x, xi_z, gamma_x, gamma_z} of the receiver's element (fluid rows)
  float* d_rec_rho = nullptr;    // [nrec][25] density (fluid rows)
  // SIMULATION_TYPE = 3: the reconstructed potential and the fluid kernels, per fluid element row [row][25]
  // (rows in element order): point ids, geometry, {rho, kappa}, rho_ac_kl, kappa_ac_kl
  float *d_bpot = nullptr, *d_bpd = nullptr, *d_bpdd = nullptr;
  int* d_kl_ids = nullptr;
  float4* d_kl_geom = nullptr;
  float2* d_kl_rk = nullptr;
  float *d_rho_kl = nullptr, *d_kappa_kl = nullptr;
};

void acoustic_release(Mesh* mp) {
  AcousticState* ac = mp->ac;
  if (!ac) return;
  device_free(ac->d_pot); device_free(ac->d_pd); device_free(ac->d_pdd); device_free(ac->d_rm);
  device_free(ac->d_list); device_free(ac->d_ids); device_free(ac->d_geom); device_free(ac->d_geom_rect); device_free(ac->d_fac); device_free(ac->d_free);
  device_free(ac->d_cdesc); device_free(ac->d_cids); device_free(ac->d_cfac); device_free(ac->d_cgeom); device_free(ac->d_crect);
  device_free(ac->d_coup); device_free(ac->d_coup_n); device_free(ac->d_coup_w);
  device_free(ac->d_st); device_free(ac->d_st_w); device_free(ac->d_st_rvp);
  device_free(ac->d_src); device_free(ac->d_src_pts); device_free(ac->d_src_sa); device_free(ac->d_src_kap);
  device_free(ac->d_rec_pts); device_free(ac->d_rec_geom); device_free(ac->d_rec_rho);
  device_free(ac->d_st_slot); device_free(ac->d_absorb);
  device_free(ac->d_bpot); device_free(ac->d_bpd); device_free(ac->d_bpdd);
  device_free(ac->d_kl_ids); device_free(ac->d_kl_geom); device_free(ac->d_kl_rk);
  device_free(ac->d_rho_kl); device_free(ac->d_kappa_kl);
  delete ac;
  mp->ac = nullptr;
}

// the fluid's Stacey points, recorded by prepare_stacey_device_ (ispec 0-based caller order; ij 0-based with -1 for
// points the edge does not own; rho_vp caller order [nspec][25])
void acoustic_note_stacey(Mesh* mp, int nf, const int* ispec, const int* ij, const float* jacw, const float* rho_vp,
                          const int* edge, const int* ib, const int* counts) {
  AcousticState* ac = mp->ac;
  if (!ac) fail("prepare_stacey_device: fluid elements without prepare_fields_acoustic_device");
  ac->abs_total = 0;
  for (int side = 1; side <= 4; side++) {
    ac->abs_count[side] = counts[side];
    ac->abs_base[side] = ac->abs_total;
    ac->abs_total += NGLL * counts[side];
  }
  for (int f = 0; f < nf; f++) {
    const int e = ispec[f];
    if (ac->row_of_ispec[e] < 0) continue;
    for (int k = 0; k < NGLL; k++) {
      const int i = ij[((size_t)f * NGLL + k) * 2], j = ij[((size_t)f * NGLL + k) * 2 + 1];
      if (i < 0 || j < 0) continue;
      ac->st_e.push_back(e);
      ac->st_ij.push_back(i + NGLL * j);
      ac->st_w.push_back(jacw[(size_t)f * NGLL + k]);
      ac->st_rvp.push_back(rho_vp[(size_t)e * NGLL2 + j * NGLL + i]);
      ac->st_slot.push_back(ac->abs_base[edge[f]] + ib[f] * NGLL + k);
    }
  }
}

namespace {

// device point ids of (element, i + 5 j) pairs, from the handle's current ibool
struct IboolView {
  std::vector<int> ib;  // device element order, device ids
  const Mesh* mp;
  explicit IboolView(const Mesh* m) : mp(m) {
    ib.resize((size_t)m->nspec * NGLL2);
    check_hip(hipMemcpy(ib.data(), m->d_ibool, ib.size() * sizeof(int), hipMemcpyDeviceToHost), "ibool D2H");
  }
  int row(int e) const { return mp->sorted ? mp->pos_of_ispec[e] : e; }
  int id(int e, int p) const { return ib[(size_t)row(e) * NGLL2 + p]; }
};

std::vector<float> download_rows(const Mesh* mp, const float* d, const std::vector<int>& elems) {
  std::vector<float> all((size_t)mp->nspec * NGLL2), out(elems.size() * NGLL2);
  check_hip(hipMemcpy(all.data(), d, all.size() * sizeof(float), hipMemcpyDeviceToHost), "geometry D2H");
  for (size_t q = 0; q < elems.size(); q++) {
    const int r = mp->sorted ? mp->pos_of_ispec[elems[q]] : elems[q];
    std::copy(all.begin() + (size_t)r * NGLL2, all.begin() + (size_t)(r + 1) * NGLL2, out.begin() + q * NGLL2);
  }
  return out;
}

// resolves the recorded set-up into device arrays (first call of the time loop)
AcousticState* ac_ready(Mesh* mp, const char* who) {
  AcousticState* ac = mp->ac;
  if (!ac) fail("%s: prepare_fields_acoustic_device was not called", who);
  if (ac->ready) return ac;
  if (mp->nspec_elastic > 0 && !mp->sorted)
    fail("%s: prepare_fields_elastic_device must run before the first acoustic step of a coupled mesh", who);
  IboolView iv(mp);
  // colour the fluid elements of each phase: greedy, first free colour over the points of the element
  std::vector<int> order;
  std::vector<unsigned int> used((size_t)mp->nglob, 0u);
  for (int ph = 0; ph < 2; ph++) {
    const std::vector<int>& el = ac->phase_elems[ph];
    std::vector<int> col(el.size(), 0);
    int ncol = 0;
    std::fill(used.begin(), used.end(), 0u);
    for (size_t q = 0; q < el.size(); q++) {
      unsigned int forbidden = 0;
      for (int p = 0; p < NGLL2; p++) forbidden |= used[iv.id(el[q], p)];
      if (forbidden == 0xffffffffu) fail("%s: fluid elements need more than 32 colours", who);
      int c = 0;
      while (forbidden & (1u << c)) c++;
      col[q] = c;
      ncol = std::max(ncol, c + 1);
      for (int p = 0; p < NGLL2; p++) used[iv.id(el[q], p)] |= 1u << c;
    }
    ac->ncolour[ph] = ncol;
    for (int c = 0; c < ncol; c++) {
      ac->colour_start[ph][c] = (int)order.size();
      for (size_t q = 0; q < el.size(); q++)
        if (col[q] == c) order.push_back(el[q]);
      ac->colour_count[ph][c] = (int)order.size() - ac->colour_start[ph][c];
    }
  }
  ac->nelem = (int)order.size();
  {
    const size_t n = order.size() * NGLL2;
    std::vector<float> xix = download_rows(mp, mp->d_xix, order), xiz = download_rows(mp, mp->d_xiz, order);
    std::vector<float> gax = download_rows(mp, mp->d_gammax, order), gaz = download_rows(mp, mp->d_gammaz, order);
    std::vector<float4> geom(n);
    std::vector<float> fac(n);
    std::vector<int> ids(n);
    for (size_t q = 0; q < order.size(); q++) {
      const int e = order[q];
      const float* rho = ac->rho_rows.data() + (size_t)ac->row_of_ispec[e] * NGLL2;
      for (int p = 0; p < NGLL2; p++) {
        const size_t l = q * NGLL2 + p;
        geom[l] = make_float4(xix[l], xiz[l], gax[l], gaz[l]);
        // the jacobian from the inverse map, as the device solver forms it (no jacobian crosses the C-ABI)
        const float jac = 1.f / (xix[l] * gaz[l] - gax[l] * xiz[l]);
        // elements of a CPML layer add nothing here: their whole term comes from the layer's own kernel
        fac[l] = pml_element(mp, e) ? 0.f : jac / rho[p];
        ids[l] = iv.id(e, p);
      }
    }
    {
      // per-point arrays in the force kernel's row-block order: workgroup blocks of AC_EPB elements of one colour,
      // inside a block of nb elements point (j, i) of element le at j * 5 nb + 5 le + i
      std::vector<int> ids_b(n);
      std::vector<float4> geom_b(n);
      std::vector<float> fac_b(n);
      for (int ph = 0; ph < 2; ph++)
        for (int c = 0; c < ac->ncolour[ph]; c++) {
          const int cs = ac->colour_start[ph][c], cn = ac->colour_count[ph][c];
          for (int b0 = 0; b0 < cn; b0 += AC_EPB) {
            const int nb = std::min(AC_EPB, cn - b0);
            for (int le = 0; le < nb; le++)
              for (int p = 0; p < NGLL2; p++) {
                const size_t src = (size_t)(cs + b0 + le) * NGLL2 + p;
                const size_t dst = (size_t)(cs + b0) * NGLL2 + (size_t)(p / NGLL) * NGLL * nb + NGLL * le + p % NGLL;
                ids_b[dst] = ids[src];
                geom_b[dst] = geom[src];
                fac_b[dst] = fac[src];
              }
          }
        }
      ac->d_ids = device_copy(ids_b.data(), n);
      ac->d_geom = device_copy(geom_b.data(), n);
      ac->d_fac = device_copy(fac_b.data(), n);
    }
    // rectangles: xi_z, gamma_x zero and xi_x, gamma_z constant over each element, to 2^-22 of the largest entry
    // (meshers leave cancellation residues) -- SF2D_NO_RECT keeps the per-point geometry
    bool rect = !getenv("SF2D_NO_RECT") && !order.empty();
    std::vector<float2> gr(order.size());
    for (size_t q = 0; q < order.size() && rect; q++) {
      const float4* g4 = geom.data() + q * NGLL2;
      float m = 0.f;
      for (int p = 0; p < NGLL2; p++)
        m = std::max({m, std::fabs(g4[p].x), std::fabs(g4[p].y), std::fabs(g4[p].z), std::fabs(g4[p].w)});
      const float tol = m * 0x1p-22f;
      for (int p = 0; p < NGLL2 && rect; p++)
        rect = std::fabs(g4[p].y) <= tol && std::fabs(g4[p].z) <= tol && std::fabs(g4[p].x - g4[0].x) <= tol &&
               std::fabs(g4[p].w - g4[0].w) <= tol;
      gr[q] = make_float2(g4[0].x, g4[0].w);
    }
    ac->rect = rect;
    if (rect) ac->d_geom_rect = device_copy(gr.data(), gr.size());
    // chains of fluid elements for the column-lane kernel (SF2D_NO_FLUID_CHAIN keeps the per-point kernel)
    if (!getenv("SF2D_NO_FLUID_CHAIN") && !order.empty()) {
      const int np = (int)std::max(ac->phase_elems[0].size(), ac->phase_elems[1].size());
      std::vector<int> lists((size_t)2 * np, 0);
      for (int ph = 0; ph < 2; ph++)
        for (size_t q = 0; q < ac->phase_elems[ph].size(); q++) lists[(size_t)ph * np + q] = ac->phase_elems[ph][q] + 1;
      std::vector<int> ibd((size_t)mp->nspec * NGLL2);
      for (int e = 0; e < mp->nspec; e++)
        for (int p = 0; p < NGLL2; p++) ibd[(size_t)e * NGLL2 + p] = iv.id(e, p);
      std::vector<int> corder, cdesc;
      if (chain_patches(ibd, mp->nspec, mp->nglob, lists.data(), np, corder, cdesc, ac->cn_colour, ac->cc_start,
                        ac->cc_count)) {
        // the packed (colour-order) index of each fluid element, for its operands computed above
        std::vector<int> packed((size_t)mp->nspec, -1);
        for (size_t q = 0; q < order.size(); q++) packed[order[q]] = (int)q;
        const size_t nf = order.size();
        std::vector<int> cids(nf * NGLL2);
        std::vector<float> cfac(nf * NGLL2);
        std::vector<float4> cgeom(nf * NGLL2);
        std::vector<float2> crect(nf);
        const int npatch = (int)cdesc.size() / 4;
        bool ok = true;
        for (int q = 0; q < npatch && ok; q++) {
          const int e0 = cdesc[4 * q], nel = cdesc[4 * q + 3] & 0xf;
          for (int kk = 0; kk < nel && ok; kk++) {
            const int src = packed[corder[e0 + kk]];
            if (src < 0 || e0 + kk >= (int)nf) { ok = false; break; }
            crect[e0 + kk] = gr[src];
            for (int p = 0; p < NGLL2; p++) {
              const size_t dst = (size_t)e0 * NGLL2 + (size_t)(p / NGLL) * NGLL * nel + NGLL * kk + p % NGLL;
              cids[dst] = ids[(size_t)src * NGLL2 + p];
              cfac[dst] = fac[(size_t)src * NGLL2 + p];
              cgeom[dst] = geom[(size_t)src * NGLL2 + p];
            }
          }
        }
        if (ok) {
          for (int q = 0; q < npatch; q++) cdesc[4 * q + 1] = cdesc[4 * q + 2] = 0;
          ac->d_cdesc = reinterpret_cast<int4*>(device_copy(cdesc.data(), cdesc.size()));
          ac->d_cids = device_copy(cids.data(), cids.size());
          ac->d_cfac = device_copy(cfac.data(), cfac.size());
          if (rect) ac->d_crect = device_copy(crect.data(), crect.size());
          else ac->d_cgeom = device_copy(cgeom.data(), cgeom.size());
          ac->chain = 1;
        }
      }
    }
    // the Newmark updates visit the fluid's points only (the potential is zero elsewhere): the solid's numbering
    // puts the points of fluid elements alone after all others, so one id range holds them and a short list the
    // ones on fluid-solid edges; the split minimises range length + 3 x list length
    std::vector<int> pts(ids);
    std::sort(pts.begin(), pts.end());
    pts.erase(std::unique(pts.begin(), pts.end()), pts.end());
    size_t best = 0;
    long best_cost = -1;
    for (size_t k = 0; k < pts.size(); k++) {
      const long cost = (long)(pts.back() - pts[k] + 1) + 3L * (long)k;
      if (best_cost < 0 || cost < best_cost) {
        best_cost = cost;
        best = k;
      }
    }
    if (!pts.empty()) {
      ac->r0 = pts[best];
      ac->nrange = pts.back() - pts[best] + 1;
      ac->nlist = (int)best;
      if (best) ac->d_list = device_copy(pts.data(), best);
    }
  }
  ac->d_rm = device_zeros<float>((size_t)mp->nglob);
  points_to_device(mp, ac->d_rm, ac->h_rm.data(), 1, "rmass_acoustic H2D");
  {
    std::vector<int> pts;
    for (size_t k = 0; k < ac->free_e.size(); k++) pts.push_back(iv.id(ac->free_e[k], ac->free_ij[k]));
    std::sort(pts.begin(), pts.end());
    pts.erase(std::unique(pts.begin(), pts.end()), pts.end());
    ac->nfree = (int)pts.size();
    if (ac->nfree) ac->d_free = device_copy(pts.data(), pts.size());
  }
  {
    std::vector<int> pts;
    for (size_t k = 0; k < ac->coup_e.size(); k++) pts.push_back(iv.id(ac->coup_e[k], ac->coup_ij[k]));
    ac->ncoup = (int)pts.size();
    if (ac->ncoup) {
      ac->d_coup = device_copy(pts.data(), pts.size());
      ac->d_coup_n = reinterpret_cast<float2*>(device_copy(ac->coup_n.data(), ac->coup_n.size()));
      ac->d_coup_w = device_copy(ac->coup_w.data(), ac->coup_w.size());
    }
  }
  {
    std::vector<int> pts;
    for (size_t k = 0; k < ac->st_e.size(); k++) pts.push_back(iv.id(ac->st_e[k], ac->st_ij[k]));
    ac->nst = (int)pts.size();
    if (ac->nst) {
      ac->d_st = device_copy(pts.data(), pts.size());
      ac->d_st_w = device_copy(ac->st_w.data(), ac->st_w.size());
      ac->d_st_rvp = device_copy(ac->st_rvp.data(), ac->st_rvp.size());
      ac->d_st_slot = device_copy(ac->st_slot.data(), ac->st_slot.size());
    }
    if (ac->abs_total && (mp->simulation_type == 3 || mp->save_forward))
      ac->d_absorb = device_zeros<float>((size_t)ac->abs_total);
  }
  if (mp->simulation_type == 3) {  // fluid kernels: per fluid element row, element order
    std::vector<int> rows;
    for (int e = 0; e < mp->nspec; e++)
      if (ac->row_of_ispec[e] >= 0) rows.push_back(e);
    std::vector<float> xix = download_rows(mp, mp->d_xix, rows), xiz = download_rows(mp, mp->d_xiz, rows);
    std::vector<float> gax = download_rows(mp, mp->d_gammax, rows), gaz = download_rows(mp, mp->d_gammaz, rows);
    std::vector<int> ids(rows.size() * NGLL2);
    std::vector<float4> geom(ids.size());
    std::vector<float2> rk(ids.size());
    for (size_t q = 0; q < rows.size(); q++)
      for (int p = 0; p < NGLL2; p++) {
        const size_t l = q * NGLL2 + p;
        ids[l] = iv.id(rows[q], p);
        geom[l] = make_float4(xix[l], xiz[l], gax[l], gaz[l]);
        rk[l] = make_float2(ac->rho_rows[l], ac->kappa_rows[l]);
      }
    ac->d_kl_ids = device_copy(ids.data(), ids.size());
    ac->d_kl_geom = device_copy(geom.data(), geom.size());
    ac->d_kl_rk = device_copy(rk.data(), rk.size());
    ac->d_rho_kl = device_zeros<float>(ids.size());
    ac->d_kappa_kl = device_zeros<float>(ids.size());
  }
  if (mp->nsources_local > 0) {
    std::vector<float> sa((size_t)mp->nsources_local * NGLL2 * 2);
    check_hip(hipMemcpy(sa.data(), mp->d_sourcearrays, sa.size() * sizeof(float), hipMemcpyDeviceToHost), "D2H");
    std::vector<int> src, pts;
    std::vector<float> a, kap;
    for (int s = 0; s < mp->nsources_local; s++) {
      const int e = mp->h_src_ispec[s];
      const int row = ac->row_of_ispec[e];
      if (row < 0) continue;
      src.push_back(s);
      for (int p = 0; p < NGLL2; p++) {
        pts.push_back(iv.id(e, p));
        a.push_back(sa[((size_t)s * NGLL2 + p) * 2]);
        kap.push_back(ac->kappa_rows[(size_t)row * NGLL2 + p]);
      }
    }
    ac->nas = (int)src.size();
    if (ac->nas) {
      ac->d_src = device_copy(src.data(), src.size());
      ac->d_src_pts = device_copy(pts.data(), pts.size());
      ac->d_src_sa = device_copy(a.data(), a.size());
      ac->d_src_kap = device_copy(kap.data(), kap.size());
    }
  }
  if (mp->nrec_local > 0) {
    std::vector<int> pts((size_t)mp->nrec_local * NGLL2, -1), rec_el;
    for (int r = 0; r < mp->nrec_local; r++) {
      const int e = mp->h_rec_ispec[r];
      rec_el.push_back(e);
      if (ac->row_of_ispec[e] < 0) continue;
      ac->nrec_fluid++;
      for (int p = 0; p < NGLL2; p++) pts[(size_t)r * NGLL2 + p] = iv.id(e, p);
    }
    ac->d_rec_pts = device_copy(pts.data(), pts.size());
    if (ac->nrec_fluid > 0) {
      std::vector<float> xix = download_rows(mp, mp->d_xix, rec_el), xiz = download_rows(mp, mp->d_xiz, rec_el);
      std::vector<float> gax = download_rows(mp, mp->d_gammax, rec_el), gaz = download_rows(mp, mp->d_gammaz, rec_el);
      std::vector<float4> geom(pts.size(), make_float4(0.f, 0.f, 0.f, 0.f));
      std::vector<float> rho(pts.size(), 1.f);
      for (int r = 0; r < mp->nrec_local; r++) {
        const int row = ac->row_of_ispec[rec_el[r]];
        if (row < 0) continue;
        for (int p = 0; p < NGLL2; p++) {
          const size_t l = (size_t)r * NGLL2 + p;
          geom[l] = make_float4(xix[l], xiz[l], gax[l], gaz[l]);
          rho[l] = ac->rho_rows[(size_t)row * NGLL2 + p];
        }
      }
      ac->d_rec_geom = device_copy(geom.data(), geom.size());
      ac->d_rec_rho = device_copy(rho.data(), rho.size());
    }
  }
  ac->ready = 1;
  return ac;
}

// the reconstructed (backward) potential exists only in SIMULATION_TYPE = 3 runs
void require_backward(const Mesh* mp, const AcousticState* ac, const char* who) {
  if (!ac->d_bpot)
    fail("%s: backward fluid wavefield requested, but the run is not SIMULATION_TYPE = 3 or "
         "prepare_fields_acoustic_adj_dev was not called", who);
  if (mp->pml) fail("%s: backward fluid wavefields with CPML are not supported by this library", who);
}

}  // namespace

// ---------------------------------------------------------------------------
// launchers used by other translation units
void launch_acoustic_pressure_seismograms(Mesh* mp, float* dst, int cur, int use_trick, int potential) {
  AcousticState* ac = ac_ready(mp, "compute_seismograms_cuda");
  if (mp->nrec_local == 0) return;
  if (ac->nrec_fluid != mp->nrec_local)
    fail("compute_seismograms_cuda: pressure or potential of receivers in solid elements is not supported by this "
         "library");
  hipLaunchKernelGGL(k_ac_seismograms, dim3(mp->nrec_local), dim3(64), 0, mp->stream, mp->nrec_local,
                     mp->nlength_seismogram, cur, ac->d_rec_pts, mp->d_xir, mp->d_gammar,
                     (potential || use_trick) ? ac->d_pot : ac->d_pdd, potential ? 1.f : -1.f, dst);
  check_hip(hipGetLastError(), "k_ac_seismograms launch");
}

void launch_acoustic_vector_seismograms(Mesh* mp, float* dst, int cur, int seismotype) {
  AcousticState* ac = ac_ready(mp, "compute_seismograms_cuda");
  if (ac->nrec_fluid == 0) return;
  const float* chi = seismotype == 2 ? ac->d_pd : seismotype == 3 ? ac->d_pdd : ac->d_pot;
  hipLaunchKernelGGL(k_ac_vector_seismograms, dim3(mp->nrec_local), dim3(64), 0, mp->stream, mp->nrec_local,
                     mp->nlength_seismogram, cur, ac->d_rec_pts, ac->d_rec_geom, ac->d_rec_rho, mp->basis,
                     mp->d_xir, mp->d_gammar, mp->d_cosrot, mp->d_sinrot, chi, dst);
  check_hip(hipGetLastError(), "k_ac_vector_seismograms launch");
}

// number of receivers in fluid elements
int acoustic_receivers(Mesh* mp) { return mp->ac ? ac_ready(mp, "compute_seismograms_cuda")->nrec_fluid : 0; }

int acoustic_phase_of(const Mesh* mp, int ispec) {
  const AcousticState* ac = mp->ac;
  if (!ac) fail("fluid element %d without prepare_fields_acoustic_device", ispec + 1);
  if (ac->ready) fail("prepare_pml_device: the CPML layer must be prepared before the first acoustic step");
  for (int ph = 0; ph < 2; ph++)
    if (std::find(ac->phase_elems[ph].begin(), ac->phase_elems[ph].end(), ispec) != ac->phase_elems[ph].end())
      return ph;
  return -1;
}

void acoustic_rho_row(const Mesh* mp, int ispec, float* rho) {
  const AcousticState* ac = mp->ac;
  const int r = ac ? ac->row_of_ispec[ispec] : -1;
  if (r < 0) fail("element %d is not a fluid element", ispec + 1);
  std::memcpy(rho, &ac->rho_rows[(size_t)r * NGLL2], NGLL2 * sizeof(float));
}

}  // namespace sf2d

using namespace sf2d;

extern "C" {

// 1 when the fluid's element forces run on the chain kernel (resolved at the first acoustic step)
int sf2d_mesh_fluid_chain(long* Mesh_pointer) {
  const Mesh* mp = handle(Mesh_pointer);
  return mp->ac && mp->ac->ready && mp->ac->chain;
}

void prepare_fields_acoustic_device_(long* Mesh_pointer, realw* rmass_acoustic, realw* rhostore, realw* kappastore,
                                     int* num_phase_ispec_acoustic, int* phase_ispec_inner_acoustic,
                                     int* num_free_surface_faces, int* free_surface_ispec, int* free_surface_ijk,
                                     int* ELASTIC_SIMULATION, int* num_coupling_ac_el_faces, int* coupling_ac_el_ispec,
                                     int* coupling_ac_el_ijk, realw* coupling_ac_el_normal,
                                     realw* coupling_ac_el_jacobian2Dw, int* h_ninterface_acoustic,
                                     int* h_inum_interfaces_acoustic, int* ATTENUATION_VISCOACOUSTIC,
                                     realw* h_A_newmark, realw* h_B_newmark, int* NO_BACKWARD_RECONSTRUCTION,
                                     realw* h_no_backward_acoustic_buffer) {  // prepare_mesh_constants_cuda.cu:319
  (void)h_inum_interfaces_acoustic; (void)h_A_newmark; (void)h_B_newmark; (void)NO_BACKWARD_RECONSTRUCTION;
  (void)h_no_backward_acoustic_buffer;
  Mesh* mp = handle(Mesh_pointer);
  if (*ATTENUATION_VISCOACOUSTIC) fail("viscoacoustic attenuation is not supported by this library yet");
  if (*h_ninterface_acoustic > 0) fail("MPI acoustic interfaces passed: one GPU, one partition only");
  if (mp->ac) fail("prepare_fields_acoustic_device called twice");
  if (mp->h_is_acoustic.empty()) fail("prepare_fields_acoustic_device: no fluid element (ispec_is_acoustic)");
  AcousticState* ac = new AcousticState();
  mp->ac = ac;
  const int nglob = mp->nglob, nspec = mp->nspec;
  ac->d_pot = device_zeros<float>((size_t)nglob);
  ac->d_pd = device_zeros<float>((size_t)nglob);
  ac->d_pdd = device_zeros<float>((size_t)nglob);
  ac->h_rm.assign(rmass_acoustic, rmass_acoustic + nglob);
  ac->row_of_ispec.assign(nspec, -1);
  int nrows = 0;
  for (int e = 0; e < nspec; e++)
    if (mp->h_is_acoustic[e]) ac->row_of_ispec[e] = nrows++;
  ac->rho_rows.resize((size_t)nrows * NGLL2);
  ac->kappa_rows.resize((size_t)nrows * NGLL2);
  for (int e = 0; e < nspec; e++) {
    const int r = ac->row_of_ispec[e];
    if (r < 0) continue;
    std::memcpy(&ac->rho_rows[(size_t)r * NGLL2], rhostore + (size_t)e * NGLL2, NGLL2 * sizeof(float));
    std::memcpy(&ac->kappa_rows[(size_t)r * NGLL2], kappastore + (size_t)e * NGLL2, NGLL2 * sizeof(float));
  }
  // phase lists (Fortran (num_phase, 2)): outer then inner
  const int np = *num_phase_ispec_acoustic;
  for (int ph = 0; ph < 2; ph++)
    for (int q = 0; q < np; q++) {
      const int e1 = phase_ispec_inner_acoustic[(size_t)ph * np + q];
      if (e1 <= 0) continue;
      if (e1 > nspec) fail("phase_ispec_inner_acoustic: element %d outside 1..%d", e1, nspec);
      if (ac->row_of_ispec[e1 - 1] < 0) fail("phase_ispec_inner_acoustic: element %d is not acoustic", e1);
      ac->phase_elems[ph].push_back(e1 - 1);
    }
  if (!ac->phase_elems[0].empty()) fail("outer acoustic elements passed: one GPU, one partition only");
  auto ij_of = [&](const int* ijk, size_t k, const char* what) {
    const int i = ijk[2 * k], j = ijk[2 * k + 1];
    if (i < 1 || i > NGLL || j < 1 || j > NGLL) fail("%s: GLL index (%d, %d) out of range", what, i, j);
    return (i - 1) + NGLL * (j - 1);
  };
  for (int f = 0; f < *num_free_surface_faces; f++) {
    const int e1 = free_surface_ispec[f];
    if (e1 < 1 || e1 > nspec) fail("free_surface_ispec: element %d outside 1..%d", e1, nspec);
    for (int k = 0; k < NGLL; k++) {
      ac->free_e.push_back(e1 - 1);
      ac->free_ij.push_back(ij_of(free_surface_ijk, (size_t)f * NGLL + k, "free_surface_ij"));
    }
  }
  if (*ELASTIC_SIMULATION)
    for (int f = 0; f < *num_coupling_ac_el_faces; f++) {
      const int e1 = coupling_ac_el_ispec[f];
      if (e1 < 1 || e1 > nspec || ac->row_of_ispec[e1 - 1] < 0)
        fail("coupling_ac_el_ispec: element %d is not a fluid element", e1);
      for (int k = 0; k < NGLL; k++) {
        const size_t q = (size_t)f * NGLL + k;
        ac->coup_e.push_back(e1 - 1);
        ac->coup_ij.push_back(ij_of(coupling_ac_el_ijk, q, "coupling_ac_el_ij"));
        ac->coup_n.push_back(coupling_ac_el_normal[2 * q]);
        ac->coup_n.push_back(coupling_ac_el_normal[2 * q + 1]);
        ac->coup_w.push_back(coupling_ac_el_jacobian2Dw[q]);
      }
    }
  if (!ac->coup_e.empty()) {
    // first-touch assembly of the solid treats these points as special (cleared by the predictor, added to)
    if (mp->sorted) fail("prepare_fields_acoustic_device must run before prepare_fields_elastic_device");
    for (size_t k = 0; k < ac->coup_e.size(); k++)
      mp->h_coupling_points.push_back(mp->h_ibool[(size_t)ac->coup_e[k] * NGLL2 + ac->coup_ij[k]]);
  }
  check_hip(hipDeviceSynchronize(), "prepare_fields_acoustic_device");
}

void prepare_fields_acoustic_adj_dev_(long* Mesh_pointer, int* APPROXIMATE_HESS_KL, int* ATTENUATION_VISCOACOUSTIC,
                                      int* NO_BACKWARD_RECONSTRUCTION) {  // prepare_mesh_constants_cuda.cu:442
  Mesh* mp = handle(Mesh_pointer);
  AcousticState* ac = mp->ac;
  if (!ac) fail("prepare_fields_acoustic_adj_dev: prepare_fields_acoustic_device was not called");
  if (*APPROXIMATE_HESS_KL) fail("approximate Hessian kernels are not supported by this library");
  if (*ATTENUATION_VISCOACOUSTIC) fail("viscoacoustic attenuation is not supported by this library yet");
  if (*NO_BACKWARD_RECONSTRUCTION) fail("NO_BACKWARD_RECONSTRUCTION is not supported by this library");
  if (mp->simulation_type != 3) return;
  if (mp->pml) fail("adjoint runs of fluid elements with CPML are not supported by this library");
  if (ac->d_bpot) return;
  ac->d_bpot = device_zeros<float>((size_t)mp->nglob);
  ac->d_bpd = device_zeros<float>((size_t)mp->nglob);
  ac->d_bpdd = device_zeros<float>((size_t)mp->nglob);
}

void update_displacement_ac_cuda_(long* Mesh_pointer, realw* deltat_F, realw* deltatsqover2_F, realw* deltatover2_F,
                                  realw* b_deltat_F, realw* b_deltatsqover2_F, realw* b_deltatover2_F,
                                  int* compute_b_wavefield, int* UNDO_ATTENUATION_AND_OR_PML) {  // update_displacement_cuda.cu:118
  // the Fortran passes compute_b_wavefield = .true. unless NO_BACKWARD_RECONSTRUCTION, also in forward runs: the
  // reconstructed potential moves only in SIMULATION_TYPE = 3 runs (update_displacement_cuda.cu:172-188)
  Mesh* mp = handle(Mesh_pointer);
  AcousticState* ac = ac_ready(mp, "update_displacement_ac_cuda");
  if (*UNDO_ATTENUATION_AND_OR_PML) fail("UNDO_ATTENUATION_AND_OR_PML is not supported by this library");
  // CPML: the previous-step potential of the layer's points, before the predictor
  pml_acoustic_before_predictor(mp, ac->d_pot, ac->d_pdd, *deltatsqover2_F);
  const long m = (long)ac->nrange + ac->nlist;
  if (m == 0) return;
  hipLaunchKernelGGL(k_ac_predictor, dim3(grid_for(m, 256, 65536)), dim3(256), 0, mp->stream, ac->d_pot, ac->d_pd,
                     ac->d_pdd, ac->r0, ac->nrange, ac->d_list, ac->nlist, *deltat_F, *deltatsqover2_F,
                     *deltatover2_F);
  check_hip(hipGetLastError(), "k_ac_predictor launch");
  if (mp->simulation_type == 3 && *compute_b_wavefield) {
    require_backward(mp, ac, "update_displacement_ac_cuda");
    hipLaunchKernelGGL(k_ac_predictor, dim3(grid_for(m, 256, 65536)), dim3(256), 0, mp->stream, ac->d_bpot,
                       ac->d_bpd, ac->d_bpdd, ac->r0, ac->nrange, ac->d_list, ac->nlist, *b_deltat_F,
                       *b_deltatsqover2_F, *b_deltatover2_F);
    check_hip(hipGetLastError(), "k_ac_predictor launch");
  }
}

void compute_forces_acoustic_cuda_(long* Mesh_pointer, int* iphase, int* nspec_outer_acoustic,
                                   int* nspec_inner_acoustic, int* ATTENUATION_VISCOACOUSTIC, int* compute_wavefield_1,
                                   int* compute_wavefield_2) {  // compute_forces_acoustic_cuda.cu:248
  (void)nspec_outer_acoustic; (void)nspec_inner_acoustic;
  Mesh* mp = handle(Mesh_pointer);
  if (*ATTENUATION_VISCOACOUSTIC) fail("viscoacoustic attenuation is not supported by this library yet");
  AcousticState* ac = ac_ready(mp, "compute_forces_acoustic_cuda");
  const int ph = *iphase - 1;
  if (ph < 0 || ph > 1) fail("compute_forces_acoustic_cuda: iphase %d", *iphase);
  if (*compute_wavefield_2) require_backward(mp, ac, "compute_forces_acoustic_cuda");
  for (int w = 0; w < 2; w++) {
    if (!(w == 0 ? *compute_wavefield_1 : *compute_wavefield_2)) continue;
    const float* pot = w == 0 ? ac->d_pot : ac->d_bpot;
    float* pdd = w == 0 ? ac->d_pdd : ac->d_bpdd;
    if (ac->chain) {  // chains of fluid elements, one wave per chain
      for (int c = 0; c < ac->cn_colour[ph]; c++) {
        const int n = ac->cc_count[ph][c];
        if (n == 0) continue;
        hipLaunchKernelGGL(ac->rect ? k_acoustic_forces_col<true> : k_acoustic_forces_col<false>,
                           dim3((n + AC_NXCD - 1) / AC_NXCD * AC_NXCD), dim3(64), 0, mp->stream, ac->cc_start[ph][c],
                           n, ac->d_cdesc, ac->d_cids, ac->d_cfac, ac->d_cgeom, ac->d_crect, mp->basis, pot, pdd);
        check_hip(hipGetLastError(), "k_acoustic_forces_col launch");
      }
    } else {
      for (int c = 0; c < ac->ncolour[ph]; c++) {
        const int n = ac->colour_count[ph][c];
        if (n == 0) continue;
        hipLaunchKernelGGL(ac->rect ? k_acoustic_forces<true> : k_acoustic_forces<false>,
                           dim3((n + AC_EPB - 1) / AC_EPB), dim3(256), 0, mp->stream, ac->colour_start[ph][c], n,
                           ac->d_ids, ac->d_geom, ac->d_geom_rect, ac->d_fac, mp->basis, pot, pdd);
        check_hip(hipGetLastError(), "k_acoustic_forces launch");
      }
    }
    // the CPML layer's elements of this phase (forward field only: backward fields refuse CPML)
    if (w == 0) pml_acoustic_forces(mp, ph, ac->d_pot, ac->d_pd, ac->d_pdd);
  }
}

// pml_compute_cuda.cu:47: potential, its time derivatives and the previous-step potential are zero on the outer
// boundary of the layer's fluid elements
void pml_boundary_acoustic_cuda_(long* Mesh_pointer, int* compute_wavefield_1, int* compute_wavefield_2) {
  Mesh* mp = handle(Mesh_pointer);
  if (*compute_wavefield_2) fail("pml_boundary_acoustic_cuda: backward fluid wavefields with CPML are not supported");
  if (!*compute_wavefield_1) return;
  AcousticState* ac = ac_ready(mp, "pml_boundary_acoustic_cuda");
  pml_acoustic_boundary(mp, ac->d_pot, ac->d_pd, ac->d_pdd);
}

void compute_stacey_acoustic_cuda_(long* Mesh_pointer, int* iphasef, realw* h_b_absorb_potential_left,
                                   realw* h_b_absorb_potential_right, realw* h_b_absorb_potential_top,
                                   realw* h_b_absorb_potential_bottom, int* compute_wavefield_1,
                                   int* compute_wavefield_2, int* UNDO_ATTENUATION_AND_OR_PML) {  // compute_stacey_acoustic_cuda.cu:40
  Mesh* mp = handle(Mesh_pointer);
  if (*iphasef != 1) return;
  if (*UNDO_ATTENUATION_AND_OR_PML) fail("UNDO_ATTENUATION_AND_OR_PML is not supported by this library");
  AcousticState* ac = ac_ready(mp, "compute_stacey_acoustic_cuda");
  // the caller's slices of one step, Fortran (NGLLX, nspec_side) each
  realw* host[5] = {nullptr, h_b_absorb_potential_bottom, h_b_absorb_potential_right, h_b_absorb_potential_top,
                    h_b_absorb_potential_left};
  const int write_abs = mp->simulation_type == 1 && mp->save_forward && ac->d_absorb;
  if (*compute_wavefield_1 && ac->nst > 0) {
    hipLaunchKernelGGL(k_ac_stacey, dim3((ac->nst + 255) / 256), dim3(256), 0, mp->stream, ac->nst, ac->d_st,
                       ac->d_st_w, ac->d_st_rvp, ac->d_pd, ac->d_pdd, ac->d_st_slot,
                       write_abs ? ac->d_absorb : nullptr);
    check_hip(hipGetLastError(), "k_ac_stacey launch");
  }
  if (write_abs) {
    for (int side = 1; side <= 4; side++)
      if (ac->abs_count[side] > 0)
        check_hip(hipMemcpyAsync(host[side], ac->d_absorb + ac->abs_base[side],
                                 sizeof(float) * NGLL * ac->abs_count[side], hipMemcpyDeviceToHost, mp->stream),
                  "b_absorb_potential D2H");
    check_hip(hipStreamSynchronize(mp->stream), "compute_stacey_acoustic_cuda");
  }
  if (*compute_wavefield_2) {
    require_backward(mp, ac, "compute_stacey_acoustic_cuda");
    if (ac->nst == 0) return;
    for (int side = 1; side <= 4; side++)
      if (ac->abs_count[side] > 0)
        check_hip(hipMemcpyAsync(ac->d_absorb + ac->abs_base[side], host[side],
                                 sizeof(float) * NGLL * ac->abs_count[side], hipMemcpyHostToDevice, mp->stream),
                  "b_absorb_potential H2D");
    hipLaunchKernelGGL(k_ac_stacey_backward, dim3((ac->nst + 255) / 256), dim3(256), 0, mp->stream, ac->nst,
                       ac->d_st, ac->d_st_slot, ac->d_absorb, ac->d_bpdd);
    check_hip(hipGetLastError(), "k_ac_stacey_backward launch");
    check_hip(hipStreamSynchronize(mp->stream), "compute_stacey_acoustic_cuda");  // the host slice is reused
  }
}

void acoustic_enforce_free_surf_cuda_(long* Mesh_pointer, int* compute_wavefield_1,
                                      int* compute_wavefield_2) {  // enforce_acoustic_free_surface_cuda.cu:46
  Mesh* mp = handle(Mesh_pointer);
  AcousticState* ac = ac_ready(mp, "acoustic_enforce_free_surf_cuda");
  if (*compute_wavefield_2) require_backward(mp, ac, "acoustic_enforce_free_surf_cuda");
  if (ac->nfree == 0) return;
  if (*compute_wavefield_1) {
    hipLaunchKernelGGL(k_ac_free_surface, dim3((ac->nfree + 255) / 256), dim3(256), 0, mp->stream, ac->d_free,
                       ac->nfree, ac->d_pot, ac->d_pd, ac->d_pdd);
    check_hip(hipGetLastError(), "k_ac_free_surface launch");
  }
  if (*compute_wavefield_2) {
    hipLaunchKernelGGL(k_ac_free_surface, dim3((ac->nfree + 255) / 256), dim3(256), 0, mp->stream, ac->d_free,
                       ac->nfree, ac->d_bpot, ac->d_bpd, ac->d_bpdd);
    check_hip(hipGetLastError(), "k_ac_free_surface launch");
  }
}

void compute_add_sources_ac_cuda_(long* Mesh_pointer, int* iphasef, int* itf) {  // compute_add_sources_viscoacoustic_cuda.cu:46
  Mesh* mp = handle(Mesh_pointer);
  if (*iphasef != 1) return;
  AcousticState* ac = ac_ready(mp, "compute_add_sources_ac_cuda");
  if (ac->nas == 0) return;
  const int it = *itf - 1;
  if (it < 0 || it >= mp->nstep) fail("compute_add_sources_ac_cuda: it %d outside 1..%d", *itf, mp->nstep);
  hipLaunchKernelGGL(k_ac_sources, dim3(ac->nas), dim3(64), 0, mp->stream, ac->nas, it, mp->nsources_local,
                     ac->d_src, ac->d_src_pts, ac->d_src_sa, ac->d_src_kap, mp->d_stf, ac->d_pdd);
  check_hip(hipGetLastError(), "k_ac_sources launch");
}

// FORWARD_OR_ADJOINT 1: the forward field (SIMULATION_TYPE 1) or the adjoint field (SIMULATION_TYPE 3, coupled
// through minus the adjoint acceleration); 3: the reconstructed field, coupled through the reconstructed displacement
// the reconstructed field's pressure sources at step it (compute_add_sources_acoustic(b_pdd, NSTEP - it + 1))
void compute_add_sources_ac_s3_cuda_(long* Mesh_pointer, int* iphasef, int* itf) {  // compute_add_sources_viscoacoustic_cuda.cu:91
  Mesh* mp = handle(Mesh_pointer);
  if (*iphasef != 1) return;
  AcousticState* ac = ac_ready(mp, "compute_add_sources_ac_s3_cuda");
  require_backward(mp, ac, "compute_add_sources_ac_s3_cuda");
  if (ac->nas == 0) return;
  const int it = *itf - 1;
  if (it < 0 || it >= mp->nstep) fail("compute_add_sources_ac_s3_cuda: it %d outside 1..%d", *itf, mp->nstep);
  hipLaunchKernelGGL(k_ac_sources, dim3(ac->nas), dim3(64), 0, mp->stream, ac->nas, it, mp->nsources_local,
                     ac->d_src, ac->d_src_pts, ac->d_src_sa, ac->d_src_kap, mp->d_stf, ac->d_bpdd);
  check_hip(hipGetLastError(), "k_ac_sources launch");
}

// adjoint sources of the receivers in fluid elements at step it (= NSTEP - it_forward + 1), into the adjoint field
void add_sources_ac_sim_2_or_3_cuda_(long* Mesh_pointer, int* iphasef, int* itf, int* nadj_rec_local,
                                     int* NSTEP) {  // compute_add_sources_viscoacoustic_cuda.cu:208
  Mesh* mp = handle(Mesh_pointer);
  if (*iphasef != 1) return;
  AcousticState* ac = ac_ready(mp, "add_sources_ac_sim_2_or_3_cuda");
  if (*nadj_rec_local != mp->nadj_rec_local) fail("add_sources_ac_sim_2_or_3_cuda: nadj_rec_local not equal");
  if (mp->nadj_rec_local == 0 || ac->nrec_fluid == 0) return;
  if (*NSTEP != mp->nstep) fail("add_sources_ac_sim_2_or_3_cuda: NSTEP mismatch");
  if (*itf < 1 || *itf > mp->nstep) fail("add_sources_ac_sim_2_or_3_cuda: it=%d outside 1..%d", *itf, mp->nstep);
  hipLaunchKernelGGL(k_ac_adjoint_sources, dim3(mp->nadj_rec_local), dim3(64), 0, mp->stream, mp->nadj_rec_local,
                     *itf - 1, mp->nstep, ac->d_rec_pts, mp->d_xir, mp->d_gammar, mp->d_source_adjoint, ac->d_pdd);
  check_hip(hipGetLastError(), "k_ac_adjoint_sources launch");
}

// rho_ac_kl, kappa_ac_kl += this step's increments (with deltat; the caller multiplies by
// NTSTEP_BETWEEN_COMPUTE_KERNELS, iterate_time.F90:489-497)
void compute_kernels_acoustic_cuda_(long* Mesh_pointer, realw* deltat) {  // compute_kernels_cuda.cu:92
  Mesh* mp = handle(Mesh_pointer);
  AcousticState* ac = ac_ready(mp, "compute_kernels_acoustic_cuda");
  require_backward(mp, ac, "compute_kernels_acoustic_cuda");
  const int nrows = (int)ac->rho_rows.size() / NGLL2;
  if (nrows == 0) return;
  hipLaunchKernelGGL(k_ac_kernels, dim3((nrows + KL_EPB - 1) / KL_EPB), dim3(256), 0, mp->stream, nrows,
                     ac->d_kl_ids, ac->d_kl_geom, ac->d_kl_rk, mp->basis, ac->d_pot, ac->d_pdd, ac->d_bpot, *deltat,
                     ac->d_rho_kl, ac->d_kappa_kl);
  check_hip(hipGetLastError(), "k_ac_kernels launch");
}

void transfer_kernels_ac_to_host_(long* Mesh_pointer, realw* h_rho_ac_kl, realw* h_kappa_ac_kl,
                                  int* NSPEC_AB) {  // transfer_fields_cuda.cu:391
  Mesh* mp = handle(Mesh_pointer);
  if (*NSPEC_AB != mp->nspec) fail("transfer_kernels_ac_to_host: NSPEC_AB %d != nspec %d", *NSPEC_AB, mp->nspec);
  AcousticState* ac = ac_ready(mp, "transfer_kernels_ac_to_host");
  require_backward(mp, ac, "transfer_kernels_ac_to_host");
  std::vector<float> r(ac->rho_rows.size()), k(ac->rho_rows.size());
  check_hip(hipStreamSynchronize(mp->stream), "transfer_kernels_ac_to_host");
  check_hip(hipMemcpy(r.data(), ac->d_rho_kl, r.size() * sizeof(float), hipMemcpyDeviceToHost), "rho_ac_kl D2H");
  check_hip(hipMemcpy(k.data(), ac->d_kappa_kl, k.size() * sizeof(float), hipMemcpyDeviceToHost), "kappa_ac_kl D2H");
  for (int e = 0; e < mp->nspec; e++) {
    const int row = ac->row_of_ispec[e];
    for (int p = 0; p < NGLL2; p++) {
      h_rho_ac_kl[(size_t)e * NGLL2 + p] = row < 0 ? 0.f : r[(size_t)row * NGLL2 + p];
      h_kappa_ac_kl[(size_t)e * NGLL2 + p] = row < 0 ? 0.f : k[(size_t)row * NGLL2 + p];
    }
  }
}

void compute_coupling_ac_el_cuda_(long* Mesh_pointer, int* iphasef, int* num_coupling_ac_el_facesf,
                                  int* FORWARD_OR_ADJOINT) {  // compute_coupling_cuda.cu:47
  Mesh* mp = handle(Mesh_pointer);
  if (*iphasef != 1) return;
  const int which = *FORWARD_OR_ADJOINT;
  if (which != 1 && which != 3) fail("compute_coupling_ac_el_cuda: FORWARD_OR_ADJOINT = %d", which);
  AcousticState* ac = ac_ready(mp, "compute_coupling_ac_el_cuda");
  if (*num_coupling_ac_el_facesf * NGLL != ac->ncoup)
    fail("compute_coupling_ac_el_cuda: %d faces, %d prepared points", *num_coupling_ac_el_facesf, ac->ncoup);
  if (ac->ncoup == 0) return;
  if (!mp->d_displ) fail("compute_coupling_ac_el_cuda: the solid fields are not prepared");
  const float* u = mp->d_displ;
  float sign = 1.f;
  float* pdd = ac->d_pdd;
  if (which == 3) {
    require_backward(mp, ac, "compute_coupling_ac_el_cuda");
    u = mp->d_b_displ;
    pdd = ac->d_bpdd;
  } else if (mp->simulation_type == 3) {
    flush_correctors(mp);  // the adjoint acceleration after its mass and corrector
    u = mp->d_accel;
    sign = -1.f;
  }
  hipLaunchKernelGGL(k_coupling_ac_el, dim3((ac->ncoup + 255) / 256), dim3(256), 0, mp->stream, ac->ncoup,
                     ac->d_coup, ac->d_coup_n, ac->d_coup_w, reinterpret_cast<const float2*>(u), sign, pdd);
  check_hip(hipGetLastError(), "k_coupling_ac_el launch");
}

// FORWARD_OR_ADJOINT 1: forward field (pressure -pdd) or, in SIMULATION_TYPE 3, the adjoint field (pressure = the
// adjoint potential); 3: the reconstructed field
void compute_coupling_el_ac_cuda_(long* Mesh_pointer, int* iphasef, int* num_coupling_ac_el_facesf,
                                  int* FORWARD_OR_ADJOINT) {  // compute_coupling_cuda.cu:125
  Mesh* mp = handle(Mesh_pointer);
  if (*iphasef != 1) return;
  const int which = *FORWARD_OR_ADJOINT;
  if (which != 1 && which != 3) fail("compute_coupling_el_ac_cuda: FORWARD_OR_ADJOINT = %d", which);
  AcousticState* ac = ac_ready(mp, "compute_coupling_el_ac_cuda");
  if (*num_coupling_ac_el_facesf * NGLL != ac->ncoup)
    fail("compute_coupling_el_ac_cuda: %d faces, %d prepared points", *num_coupling_ac_el_facesf, ac->ncoup);
  if (ac->ncoup == 0) return;
  const float* p = ac->d_pdd;
  float sign = 1.f;
  float* acc = mp->d_accel;
  if (which == 3) {
    require_backward(mp, ac, "compute_coupling_el_ac_cuda");
    if (!mp->d_b_accel) fail("compute_coupling_el_ac_cuda: the backward solid field is not prepared");
    p = ac->d_bpdd;
    acc = mp->d_b_accel;
  } else if (mp->simulation_type == 3) {
    p = ac->d_pot;
    sign = -1.f;
  }
  hipLaunchKernelGGL(k_coupling_el_ac, dim3((ac->ncoup + 255) / 256), dim3(256), 0, mp->stream, ac->ncoup,
                     ac->d_coup, ac->d_coup_n, ac->d_coup_w, p, sign, acc);
  check_hip(hipGetLastError(), "k_coupling_el_ac launch");
}

void kernel_3_acoustic_cuda_(long* Mesh_pointer, realw* deltatover2, realw* b_deltatover2, int* compute_wavefield_1,
                             int* compute_wavefield_2) {  // update_displacement_cuda.cu:243
  Mesh* mp = handle(Mesh_pointer);
  AcousticState* ac = ac_ready(mp, "kernel_3_acoustic_cuda");
  if (*compute_wavefield_2) require_backward(mp, ac, "kernel_3_acoustic_cuda");
  const long m = (long)ac->nrange + ac->nlist;
  if (m == 0) return;
  if (*compute_wavefield_1) {
    hipLaunchKernelGGL(k_ac_corrector, dim3(grid_for(m, 256, 65536)), dim3(256), 0, mp->stream, ac->d_pd, ac->d_pdd,
                       ac->d_rm, ac->r0, ac->nrange, ac->d_list, ac->nlist, *deltatover2);
    check_hip(hipGetLastError(), "k_ac_corrector launch");
  }
  if (*compute_wavefield_2) {
    hipLaunchKernelGGL(k_ac_corrector, dim3(grid_for(m, 256, 65536)), dim3(256), 0, mp->stream, ac->d_bpd,
                       ac->d_bpdd, ac->d_rm, ac->r0, ac->nrange, ac->d_list, ac->nlist, *b_deltatover2);
    check_hip(hipGetLastError(), "k_ac_corrector launch");
  }
}

void transfer_fields_ac_to_device_(int* size, realw* potential_acoustic, realw* potential_dot_acoustic,
                                   realw* potential_dot_dot_acoustic, long* Mesh_pointer) {  // transfer_fields_cuda.cu:229
  Mesh* mp = handle(Mesh_pointer);
  if (*size != mp->nglob) fail("transfer_fields_ac_to_device: size %d != nglob %d", *size, mp->nglob);
  AcousticState* ac = ac_ready(mp, "transfer_fields_ac_to_device");
  points_to_device(mp, ac->d_pot, potential_acoustic, 1, "potential H2D");
  points_to_device(mp, ac->d_pd, potential_dot_acoustic, 1, "potential_dot H2D");
  points_to_device(mp, ac->d_pdd, potential_dot_dot_acoustic, 1, "potential_dot_dot H2D");
}

void transfer_fields_ac_from_device_(int* size, realw* potential_acoustic, realw* potential_dot_acoustic,
                                     realw* potential_dot_dot_acoustic, long* Mesh_pointer) {  // transfer_fields_cuda.cu:281
  Mesh* mp = handle(Mesh_pointer);
  if (*size != mp->nglob) fail("transfer_fields_ac_from_device: size %d != nglob %d", *size, mp->nglob);
  AcousticState* ac = ac_ready(mp, "transfer_fields_ac_from_device");
  points_from_device(mp, potential_acoustic, ac->d_pot, 1, "potential D2H");
  points_from_device(mp, potential_dot_acoustic, ac->d_pd, 1, "potential_dot D2H");
  points_from_device(mp, potential_dot_dot_acoustic, ac->d_pdd, 1, "potential_dot_dot D2H");
}

void transfer_dot_dot_from_device_(int* size, realw* potential_dot_dot_acoustic, long* Mesh_pointer) {  // transfer_fields_cuda.cu:357
  Mesh* mp = handle(Mesh_pointer);
  if (*size != mp->nglob) fail("transfer_dot_dot_from_device: size %d != nglob %d", *size, mp->nglob);
  AcousticState* ac = ac_ready(mp, "transfer_dot_dot_from_device");
  points_from_device(mp, potential_dot_dot_acoustic, ac->d_pdd, 1, "potential_dot_dot D2H");
}

void transfer_b_fields_ac_to_device_(int* size, realw* b_potential_acoustic, realw* b_potential_dot_acoustic,
                                     realw* b_potential_dot_dot_acoustic, long* Mesh_pointer) {  // transfer_fields_cuda.cu:248
  Mesh* mp = handle(Mesh_pointer);
  if (*size != mp->nglob) fail("transfer_b_fields_ac_to_device: size %d != nglob %d", *size, mp->nglob);
  AcousticState* ac = ac_ready(mp, "transfer_b_fields_ac_to_device");
  require_backward(mp, ac, "transfer_b_fields_ac_to_device");
  points_to_device(mp, ac->d_bpot, b_potential_acoustic, 1, "b_potential H2D");
  points_to_device(mp, ac->d_bpd, b_potential_dot_acoustic, 1, "b_potential_dot H2D");
  points_to_device(mp, ac->d_bpdd, b_potential_dot_dot_acoustic, 1, "b_potential_dot_dot H2D");
}

void transfer_b_fields_ac_from_device_(int* size, realw* b_potential_acoustic, realw* b_potential_dot_acoustic,
                                       realw* b_potential_dot_dot_acoustic, long* Mesh_pointer) {  // transfer_fields_cuda.cu:300
  Mesh* mp = handle(Mesh_pointer);
  if (*size != mp->nglob) fail("transfer_b_fields_ac_from_device: size %d != nglob %d", *size, mp->nglob);
  AcousticState* ac = ac_ready(mp, "transfer_b_fields_ac_from_device");
  require_backward(mp, ac, "transfer_b_fields_ac_from_device");
  points_from_device(mp, b_potential_acoustic, ac->d_bpot, 1, "b_potential D2H");
  points_from_device(mp, b_potential_dot_acoustic, ac->d_bpd, 1, "b_potential_dot D2H");
  points_from_device(mp, b_potential_dot_dot_acoustic, ac->d_bpdd, 1, "b_potential_dot_dot D2H");
}

void transfer_b_potential_ac_from_device_(int* size, realw* b_potential_acoustic, long* Mesh_pointer) {  // transfer_fields_cuda.cu:324
  Mesh* mp = handle(Mesh_pointer);
  if (*size != mp->nglob) fail("transfer_b_potential_ac_from_device: size %d != nglob %d", *size, mp->nglob);
  AcousticState* ac = ac_ready(mp, "transfer_b_potential_ac_from_device");
  require_backward(mp, ac, "transfer_b_potential_ac_from_device");
  points_from_device(mp, b_potential_acoustic, ac->d_bpot, 1, "b_potential D2H");
}

void transfer_b_potential_ac_to_device_(int* size, realw* b_potential_acoustic, long* Mesh_pointer) {  // transfer_fields_cuda.cu:341
  Mesh* mp = handle(Mesh_pointer);
  if (*size != mp->nglob) fail("transfer_b_potential_ac_to_device: size %d != nglob %d", *size, mp->nglob);
  AcousticState* ac = ac_ready(mp, "transfer_b_potential_ac_to_device");
  require_backward(mp, ac, "transfer_b_potential_ac_to_device");
  points_to_device(mp, ac->d_bpot, b_potential_acoustic, 1, "b_potential H2D");
}

void transfer_b_dot_dot_from_device_(int* size, realw* b_potential_dot_dot_acoustic, long* Mesh_pointer) {  // transfer_fields_cuda.cu:374
  Mesh* mp = handle(Mesh_pointer);
  if (*size != mp->nglob) fail("transfer_b_dot_dot_from_device: size %d != nglob %d", *size, mp->nglob);
  AcousticState* ac = ac_ready(mp, "transfer_b_dot_dot_from_device");
  require_backward(mp, ac, "transfer_b_dot_dot_from_device");
  points_from_device(mp, b_potential_dot_dot_acoustic, ac->d_bpdd, 1, "b_potential_dot_dot D2H");
}

void get_norm_acoustic_from_device_(realw* norm, long* Mesh_pointer, const int* FORWARD_OR_ADJOINT) {  // check_fields_cuda.cu:240
  Mesh* mp = handle(Mesh_pointer);
  AcousticState* ac = ac_ready(mp, "get_norm_acoustic_from_device");
  if (*FORWARD_OR_ADJOINT == 3) require_backward(mp, ac, "get_norm_acoustic_from_device");
  const float* f = *FORWARD_OR_ADJOINT == 3 ? ac->d_bpot : ac->d_pot;
  check_hip(hipMemsetAsync(mp->d_scratch_u32, 0, sizeof(unsigned int), mp->stream), "memset");
  hipLaunchKernelGGL(k_ac_max_abs, dim3(grid_for(mp->nglob, 256, 2048)), dim3(256), 0, mp->stream, f,
                     mp->nglob, mp->d_scratch_u32);
  check_hip(hipGetLastError(), "k_ac_max_abs launch");
  unsigned int bits = 0;
  check_hip(hipMemcpyAsync(&bits, mp->d_scratch_u32, sizeof(bits), hipMemcpyDeviceToHost, mp->stream), "D2H");
  check_hip(hipStreamSynchronize(mp->stream), "get_norm_acoustic_from_device");
  float v;
  std::memcpy(&v, &bits, sizeof(v));
  *norm = v;
}

}  // extern "C"
