// Device-resident state of one SPECFEM2D simulation on one MI355X (gfx950).
//
// One handle per simulation (the reference's `Mesh`, src/gpu/mesh_constants_cuda.h:389-665),
// created by prepare_constants_device_ and released by prepare_cleanup_device_.  The layout
// differs from the reference on purpose:
//   * per-element arrays are unpadded [nspec][25] (i fastest): 25-point elements are read by
//     flat, fully coalesced 64-lane loads instead of one 32-wide padded block per element;
//   * ibool and element lists are stored 0-based;
//   * the GLL basis travels as a kernel argument (no __constant__ symbols), so several
//     handles can live in one process;
//   * Stacey storage, seismogram buffers and reduction scratch are allocated once.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace sf2d {

constexpr int NGLL = 5;
constexpr int NGLL2 = 25;

struct Basis {
  float hp[NGLL2];   // hp[i*5+k]  = h'_k(xi_i)
  float hpw[NGLL2];  // hpw[k*5+i] = w_k h'_i(xi_k)
  float w[NGLL];     // GLL weights
};

struct Mesh {
  int myrank = 0;
  int device = 0;
  int nspec = 0, nglob = 0;
  int simulation_type = 1, save_forward = 0, p_sv = 1;
  int stacey = 0, pml = 0;
  int nstep = 0;
  hipStream_t stream = nullptr;
  Basis basis{};

  // per-element geometry / material [nspec][25]
  int* d_ibool = nullptr;
  float *d_xix = nullptr, *d_xiz = nullptr, *d_gammax = nullptr, *d_gammaz = nullptr;
  float *d_kappav = nullptr, *d_muv = nullptr;
  int* d_ispec_is_elastic = nullptr;
  int nspec_elastic = 0, nspec_acoustic = 0;
  int all_elastic = 1;

  // elastic wavefields [nglob][2] and mass matrix
  float *d_displ = nullptr, *d_veloc = nullptr, *d_accel = nullptr;
  float *d_b_displ = nullptr, *d_b_veloc = nullptr, *d_b_accel = nullptr;
  float *d_rmassx = nullptr, *d_rmassz = nullptr;

  // element phase lists (0-based), [2][num_phase]; identity lists are not stored
  int num_phase_ispec_elastic = 0;
  int* d_phase_ispec[2] = {nullptr, nullptr};
  int phase_is_identity[2] = {0, 0};

  // sources
  int nsources_local = 0;
  float* d_sourcearrays = nullptr;  // [nsrc][25][2]
  float* d_stf = nullptr;           // [nstep][nsrc]
  int* d_ispec_selected_source = nullptr;

  // receivers / seismograms
  int nrec_local = 0;
  int nsigtype = 0;
  int nlength_seismogram = 0;
  int* h_seismotype = nullptr;  // retained caller pointer, as the reference does
  float* d_xir = nullptr;       // [5][nrec] (receiver fastest)
  float* d_gammar = nullptr;    // [5][nrec]
  float *d_cosrot = nullptr, *d_sinrot = nullptr;
  int* d_ispec_selected_rec = nullptr;
  float** d_seismo = nullptr;   // per i_sig: [2][nrec][nlength]
  float** h_seismo = nullptr;   // pinned staging buffers

  // adjoint sources [2][NSTEP][nadj]
  int nadj_rec_local = 0;
  float* d_source_adjoint = nullptr;

  // Stacey
  int num_abs = 0;
  int nspec_left = 0, nspec_right = 0, nspec_top = 0, nspec_bottom = 0;
  int* d_abs_ispec = nullptr;   // 0-based
  int* d_abs_ij = nullptr;      // [F][5][2] 0-based, -1 marks a skipped point
  float* d_abs_normal = nullptr;// [F][5][2]
  float* d_abs_jacw = nullptr;  // [F][5]
  int* d_edge_abs = nullptr;
  int* d_ib = nullptr;          // index within its side (0-based)
  float *d_rho_vp = nullptr, *d_rho_vs = nullptr;  // [nspec][25]
  float* d_b_absorb[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};  // by edge code 1..4

  // kernels
  float *d_rho_kl = nullptr, *d_mu_kl = nullptr, *d_kappa_kl = nullptr;

  // scratch
  unsigned int* d_scratch_u32 = nullptr;  // reductions
};

// Errors follow the reference's convention: message, error file, exit(EXIT_FAILURE).
[[noreturn]] void fail(const char* fmt, ...);
void check_hip(hipError_t err, const char* what);

inline Mesh* handle(long* p) {
  Mesh* mp = reinterpret_cast<Mesh*>(*p);
  if (!mp) fail("invalid Mesh_pointer (not prepared)");
  return mp;
}

template <typename T>
T* device_copy(const T* host, size_t n, hipStream_t s = nullptr) {
  T* d = nullptr;
  size_t bytes = (n ? n : 1) * sizeof(T);
  check_hip(hipMalloc(&d, bytes), "hipMalloc");
  if (host && n) check_hip(hipMemcpy(d, host, n * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy H2D");
  (void)s;
  return d;
}

template <typename T>
T* device_zeros(size_t n) {
  T* d = nullptr;
  size_t bytes = (n ? n : 1) * sizeof(T);
  check_hip(hipMalloc(&d, bytes), "hipMalloc");
  check_hip(hipMemset(d, 0, bytes), "hipMemset");
  return d;
}

template <typename T>
void device_free(T*& p) {
  if (p) {
    (void)hipFree(p);
    p = nullptr;
  }
}

// ---------------------------------------------------------------------------
// kernel launchers (defined in elastic.hip), all asynchronous on mp->stream
void launch_newmark_predictor(float* d, float* v, float* a, int nglob, float dt, float dt2o2, float dto2,
                              hipStream_t s);
void launch_newmark_corrector(float* v, float* a, const float* rmx, const float* rmz, int nglob, float dto2,
                              int apply_mass, hipStream_t s);
void launch_elastic_forces(const Mesh* mp, const int* elist, int nelem, const float* displ, float* accel);
void launch_stacey(const Mesh* mp, const float* veloc, float* accel, int write_store);
void launch_stacey_backward(const Mesh* mp, float* b_accel);
void launch_sources(const Mesh* mp, float* accel, int it);
void launch_adjoint_sources(const Mesh* mp, float* accel, int it);
void launch_seismograms(const Mesh* mp, const float* field, float* dst, int seismo_current);
void launch_kernels(const Mesh* mp);
void launch_max_norm(const Mesh* mp, const float* field, unsigned int* result);

}  // namespace sf2d
