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
#include <vector>

namespace sf2d {

constexpr int NGLL = 5;
constexpr int NGLL2 = 25;

struct Basis {
  float hp[NGLL2];   // hp[i*5+k]  = h'_k(xi_i)
  float hpw[NGLL2];  // hpw[k*5+i] = w_k h'_i(xi_k)
  float w[NGLL];     // GLL weights
};

struct AcousticState;  // fluid elements (acoustic.hip)
struct PmlState;       // CPML layer (pml.hip)

struct Mesh {
  int myrank = 0;
  int device = 0;
  int nspec = 0, nglob = 0;
  // the solid's Newmark updates run over internal ids 0 .. nglob_el-1: in patch mode the plan numbers the points of
  // elastic elements first, and nothing changes the solid's fields at the others (SF2D_ALL_POINTS: all points)
  int nglob_el = 0;
  int same_mass = 0;  // rmassx == rmassz at every point: the Newmark kernels read rmassx for both components
  int simulation_type = 1, save_forward = 0, p_sv = 1;
  int stacey = 0, pml = 0;
  int nstep = 0;
  hipStream_t stream = nullptr;
  Basis basis{};

  // per-element geometry / material [nspec][25]
  int* d_ibool = nullptr;
  float *d_xix = nullptr, *d_xiz = nullptr, *d_gammax = nullptr, *d_gammaz = nullptr;
  float *d_kappav = nullptr, *d_muv = nullptr;
  // viscoelastic attenuation (3 standard linear solids), device element order, ne = nspec * 25: per point an
  // index into a table of distinct coefficient sets (3 x {A_kappa, B_kappa, A_mu, B_mu}), memory variables
  // e1, e11, e13 x 3 solids (9 [ne] planes), previous strains dux_dxl, duz_dzl, dux_dzl + duz_dxl (3 planes)
  int attenuation = 0, att_ntable = 0;
  unsigned int* d_att_idx = nullptr;
  float4* d_att_table = nullptr;
  // attenuation state: [3][nspec * 25] float4 (e1, e11, e13 of the 3 solids, then the 3 previous strains), in the
  // force kernel's point order; d_att_old is unused (null)
  float *d_att_e = nullptr, *d_att_old = nullptr;
  int* d_ispec_is_elastic = nullptr;
  // anisotropic elements: per point of every element (device element order) {c11, c13, c15, c33}, {c35, c55, 0, 0};
  // isotropic elements carry their equivalent constants (c11 = c33 = kappa + mu, c13 = kappa - mu, c55 = mu)
  int anisotropy = 0;
  float4* d_aniso = nullptr;
  int nspec_elastic = 0, nspec_acoustic = 0;
  int all_elastic = 1;
  std::vector<char> h_is_acoustic;  // [nspec] caller order (empty: no fluid element)
  AcousticState* ac = nullptr;      // fluid potential and its set-up, from prepare_fields_acoustic_device_

  // elastic wavefields [nglob][2] and mass matrix
  float *d_displ = nullptr, *d_veloc = nullptr, *d_accel = nullptr;
  float *d_b_displ = nullptr, *d_b_veloc = nullptr, *d_b_accel = nullptr;
  float *d_rmassx = nullptr, *d_rmassz = nullptr;

  // Element order on the device.  prepare_fields_elastic_device_ sorts the elements by (phase, colour):
  // a greedy colouring of each phase list in which no two elements of one colour share a GLL point, so
  // the force kernel assembles with plain read-modify-writes, one launch per colour.  All per-element
  // device arrays are stored in this order; pos_of_ispec / ispec_of_pos translate.
  static constexpr int MAX_COLOURS = 32;
  int num_phase_ispec_elastic = 0;
  int phase_count[2] = {0, 0};
  int ncolour[2] = {0, 0};
  int colour_start[2][MAX_COLOURS] = {};
  int colour_count[2][MAX_COLOURS] = {};
  // assembly: 0 coloured patches (default), 1 coloured elements, 2 FP32 atomics (SF2D_ASSEMBLY=patch|colour|atomic)
  int assembly = 0;
  // coloured patches: connected groups of <= PATCH_ELEMS consecutive elements of a phase list; one workgroup
  // gathers a patch's unique points into LDS and writes each back once.  In patch mode the colour ranges above
  // count patches, the element order is (phase, colour, patch, element in patch), and the points are
  // renumbered patch-major: each patch owns a contiguous range of point ids and lists the few points it
  // shares with earlier patches (its halo).  Every point array on the device uses that numbering.
  static constexpr int PATCH_ELEMS = 10;
  static constexpr int CHAIN_ELEMS = 12;  // patches of meshes whose patches are all chains (the chain kernel)
  static constexpr int PATCH_SHARE = 4;  // element points of one patch that may share a point
  int npatch = 0;
  int4* d_patch_desc = nullptr;      // [npatch] {first element, first owned id, halo offset, ne | nown<<8 | nhalo<<20}
  int* d_halo_nodes = nullptr;       // halo point ids of each patch (sorted), plus one spare entry
  uint16_t* d_elem_slot = nullptr;   // [nspec][25] slot of each element point in its patch's point list | rank << 8
  float4* d_geom = nullptr;          // [nspec][25] {xix, xiz, gammax, gammaz}, device element order (patch mode)
  float2* d_moduli = nullptr;        // [nspec][25] {kappa, mu}, device element order (patch mode)
  float4* d_geom_elem = nullptr;     // [nspec] geometry of point 0; the whole element's when it is affine
  // coloured and atomic assembly (the per-point kernel): {xi_x, gamma_z} per element, device element order, when
  // every element is an axis-aligned rectangle (SF2D_NO_RECT keeps the per-point geometry)
  float* d_geom_rect = nullptr;
  int n_affine = 0;                  // elements whose 25 points carry bitwise the same geometry
  int no_affine = 0;                 // SF2D_NO_AFFINE: always read the per-point geometry
  int rect = 0;                      // every element affine and axis-aligned (chain kernel's rectangle form)
  // tuning: elements per chain patch (SF2D_CHAIN_ELEMS at set-up, 1..CHAIN_ELEMS), shots per wave of the
  // batched chain kernel (SF2D_SHOTS_PER_WAVE or sf2d_mesh_set_shots_per_wave; 0 = built-in default)
  int chain_elems = 0;
  int chain_request = 0;  // elements per chain patch asked for by sf2d_mesh_set_chain_elems (0: default)
  int shots_per_wave = 0;
  int chain = 0;                     // every patch is a chain of elements: the row-lane force kernel (SF2D_NO_CHAIN
                                     // keeps the general patch kernel)
  int* d_chain_ids = nullptr;        // chain kernel: [nspec][25] device order, internal point id | FIRST_TOUCH
  std::vector<int> h_chain_ids;      // the same ids without flags, and the first-touch flags of the plan
  // the column-lane chain kernel's copies of the point ids and moduli in lane order: the block of a patch of ne
  // elements starting at position e0 holds element point (k, i, j) at e0 * 25 + j * 5 ne + 5 k + i, so that the
  // wave's load of GLL row j reads consecutive words (SF2D_ROW_CHAIN keeps the row-lane kernel)
  int* d_col_ids = nullptr;
  float2* d_col_mod = nullptr;
  // the column-lane kernel's point ids as rows: when GLL row j of every chain numbers its points x = 1 .. 4 ne
  // consecutively (id = base_j + x; the numbering of the plan makes it so), a patch needs per row only base_j, the
  // id of its point x = 0 and a 64-bit mask of the first-touch flags: int4 {base_j, id(0, j), mask low, mask high}
  // per (patch, row), 80 B per chain instead of 1200 B of ids (SF2D_COL_IDS keeps the per-point ids)
  int4* d_col_rows = nullptr;
  std::vector<int> h_patch_desc;
  std::vector<uint8_t> h_first;
  // First-touch assembly (forward wavefield of SIMULATION_TYPE=1 and shot batches; SF2D_NO_FIRST_TOUCH turns it
  // off): the predictor leaves accel as it is except at the special points (Stacey edges, source elements),
  // which it clears; the first colour that updates a point in a step writes its force there instead of adding.
  // Between the predictor and the last force launch of the step accel is therefore incomplete (accel_partial).
  static constexpr int FIRST_TOUCH = 1 << 30;
  int first_touch = 0;
  std::vector<int> h_stacey_points;  // internal ids of the points of the absorbing edges
  std::vector<int> h_uncovered_points;  // internal ids of the points of no solid element (fluid-only points)
  std::vector<int> h_coupling_points;   // caller ids of the fluid-solid edge points (the fluid adds to accel there)
  std::vector<int> h_pml_points;        // internal ids of the outer-boundary points of the solid's CPML layer
  int* d_special = nullptr;
  int nspecial = 0;
  int accel_partial = 0;
  std::vector<int> point_id;         // device id of each caller point (empty: identity)
  std::vector<int> h_ibool;          // 0-based, ispec order, kept until the elements are sorted
  std::vector<int> pos_of_ispec, ispec_of_pos;
  std::vector<int> h_src_ispec, h_rec_ispec;  // 0-based ispec, as passed
  int sorted = 0;

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

  // Deferred Newmark corrector: kernel_3_a_cuda_ records it and the next update_displacement_cuda_
  // applies corrector and predictor in one pass over the points; any entry point that reads veloc or
  // accel (or writes the fields) applies it first, so callers observe the reference's sequence.
  int pending_corrector[2] = {0, 0};  // forward/adjoint wavefield, backward wavefield
  float pending_dto2[2] = {0.f, 0.f};

  // CPML layer (prepare_pml_device_ / prepare_cpml_device_)
  PmlState* pml_state = nullptr;

  // scratch
  unsigned int* d_scratch_u32 = nullptr;  // reductions

  // exact discrete adjoint (gradient.hip), allocated by sf2d_grad_prepare / sf2d_batch_grad_prepare
  struct GradState* grad = nullptr;
  int grad_batch = 0;  // the gradient buffers follow the shot batch
  // shot batch (batch.hip), allocated by sf2d_batch_prepare
  struct BatchState* batch = nullptr;
};

// nshot wavefields of one mesh and model, point g of shot s at float2 g * ps + s * stride (shot-major: ps = 1;
// shot-minor: ps = nshot, stride = 1); every source of the handle belongs to one shot (src_shot).  The Newmark
// corrector is deferred as for the handle's own fields.
struct BatchState {
  int nshot = 0;
  // layout: shot-major (ps = 1; shot s at s * stride points, stride >= nglob, a multiple of 64) or shot-minor
  // (ps = nshot; point g of shot s at g * nshot + s, stride = 1), chosen by sf2d_batch_prepare
  long stride = 0;
  int ps = 1;
  float *d = nullptr, *v = nullptr, *a = nullptr;
  int* src_shot = nullptr;           // [nsources_local]
  float* seismo = nullptr;           // [nshot][2][nrec][nstep] displacement seismograms
  int pending = 0;
  float pending_dto2 = 0.f;
  // one captured step (corrector-predictor, forces, Stacey, sources, seismograms, step counter), replayed for
  // the steps after the first of a run; d_step is its 1-based step index in device memory
  int* d_step = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  float graph_dt[3] = {0.f, 0.f, 0.f};
  int graph_record = -1;
};

// shot batch fields and the re-arming of its deferred corrector (batch.hip)
void batch_fields(Mesh* mp, float** d, float** v, float** a, int* nshot, long* stride, int* ps);
void batch_set_pending(Mesh* mp, float dto2);
// time step it of every shot from state (d, v, a) -- corrector pending with dto2_prev, or at rest when d is
// null -- into state (d_out, v_out, a_out), each [nshot][stride][2]
void batch_step_into(Mesh* mp, int it, const float* d, const float* v, const float* a, float dto2_prev, float* d_out,
                     float* v_out, float* a_out, float deltat, float deltatsqover2, float deltatover2);

// applies a deferred corrector now (no-op when none is pending)
void flush_correctors(Mesh* mp);
// acc -= K u over the elements of phase ph (0 outer, 1 inner), with the handle's assembly scheme, for nshot
// wavefields shot_stride points apart
// first_touch: the step's accel was left by a predictor with zero_accel = 0 (first-touch assembly)
// ps > 1: shot-minor fields of ps shots per point (nshot and shot_stride unused)
void apply_element_forces(const Mesh* mp, int ph, const float* u, float* acc, int nshot = 1, long shot_stride = 0,
                          int first_touch = 0, int ps = 1);
// fluid elements (acoustic.hip): release, the fluid's Stacey points (recorded by prepare_stacey_device_: ispec
// 0-based caller order, ij 0-based with -1 for points an edge does not own, rho_vp [nspec][25] caller order),
// pressure seismograms (-potential with USE_TRICK_FOR_BETTER_PRESSURE, else -potential_dot_dot) or potential
// seismograms (potential != 0) and the number of receivers in fluid elements
void acoustic_release(Mesh* mp);
// edge: side code of each face (1 bottom, 2 right, 3 top, 4 left), ib: 0-based slot of the face in its side's
// storage, counts[side]: faces per side (the SAVE_FORWARD slices)
void acoustic_note_stacey(Mesh* mp, int nf, const int* ispec, const int* ij, const float* jacw, const float* rho_vp,
                          const int* edge, const int* ib, const int* counts);
void launch_acoustic_pressure_seismograms(Mesh* mp, float* dst, int cur, int use_trick, int potential);
void launch_acoustic_vector_seismograms(Mesh* mp, float* dst, int cur, int seismotype);
int acoustic_receivers(Mesh* mp);
// phase (0 outer, 1 inner, -1 none) of fluid element ispec, and its density row ([25], caller's rhostore)
int acoustic_phase_of(const Mesh* mp, int ispec);
void acoustic_rho_row(const Mesh* mp, int ispec, float* rho);
// CPML layer (pml.hip): whether element ispec (caller order, 0-based) is in it; release; zero its memory variables;
// the previous-step fields (before the Newmark predictors); the layer's element forces of phase ph (after the main
// force launches of that phase); the fluid's zero condition on the outer boundary
bool pml_element(const Mesh* mp, int ispec);
void pml_release(Mesh* mp);
void pml_reset(Mesh* mp);
void pml_elastic_before_predictor(Mesh* mp, float deltatsqover2);
void pml_elastic_forces(Mesh* mp, int ph);
void pml_acoustic_before_predictor(Mesh* mp, const float* pot, const float* pdd, float deltatsqover2);
void pml_acoustic_forces(Mesh* mp, int ph, const float* pot, const float* pd, float* pdd);
void pml_acoustic_boundary(Mesh* mp, float* pot, float* pd, float* pdd);
// chain-kernel point ids with first-touch flags and the special points, uploaded again when the special points change
void upload_chain_ids(Mesh* mp);
bool chain_row_records(const std::vector<int>& desc, const std::vector<int>& ids, const std::vector<int>& flagged,
                       std::vector<int4>& rows);
bool chain_patches(const std::vector<int>& ibool, int nspec, int nglob, const int* phase_lists, int np,
                   std::vector<int>& order, std::vector<int>& desc, int ncolour[2],
                   int colour_start[2][Mesh::MAX_COLOURS], int colour_count[2][Mesh::MAX_COLOURS]);
// fills the column-lane copies of the chain ids and moduli from d_chain_ids and d_moduli (when both are allocated)
void launch_col_pack(const Mesh* mp);
// point arrays in the caller's numbering <-> the device numbering (n2 values per point)
void points_to_device(const Mesh* mp, float* dev, const float* host, int n2, const char* what);
void points_from_device(const Mesh* mp, float* host, const float* dev, int n2, const char* what);
// per-element arrays: caller's ispec order <-> device element order
std::vector<float> to_device_order(const Mesh* mp, const float* a);
void to_ispec_order(const Mesh* mp, const float* dev_order, float* out);

// device address of the no-allocation entry point's scratch array on the current device (noalloc.hip), looked up
// once per device
unsigned int* noalloc_scratch();

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
T* device_alloc(size_t n) {  // uninitialised
  T* d = nullptr;
  check_hip(hipMalloc(&d, (n ? n : 1) * sizeof(T)), "hipMalloc");
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
// zero_accel = 0 leaves accel as it is (first-touch assembly)
void launch_newmark_predictor(float* d, float* v, float* a, int nglob, float dt, float dt2o2, float dto2,
                              hipStream_t s, int nshot = 1, long shot_stride = 0, int zero_accel = 1, int ps = 1);
void launch_newmark_corrector(float* v, float* a, const float* rmx, const float* rmz, int nglob, float dto2,
                              int apply_mass, hipStream_t s, int nshot = 1, long shot_stride = 0, int ps = 1);
void launch_newmark_corrector_predictor(float* d, float* v, float* a, const float* rmx, const float* rmz, int nglob,
                                       float dto2_prev, float dt, float dt2o2, float dto2, hipStream_t s,
                                       int nshot = 1, long shot_stride = 0, int zero_accel = 1, int ps = 1);
// out of place: (d, v, a) with the corrector pending -> (d_out, v_out, 0); the two states must not overlap
void launch_newmark_corrector_predictor_to(const float* d, const float* v, const float* a, float* d_out, float* v_out,
                                          float* a_out, const float* rmx, const float* rmz, int nglob,
                                          float dto2_prev, float dt, float dt2o2, float dto2, hipStream_t s,
                                          int nshot = 1, long shot_stride = 0, int zero_accel = 1, int ps = 1);
// accel = 0 at the special points of first-touch assembly
void launch_zero_points(const Mesh* mp, float* accel, int nshot = 1, long shot_stride = 0, int ps = 1);
void launch_elastic_forces(const Mesh* mp, int first, int nelem, int atomic, const float* displ, float* accel,
                           int att = 0);
// nshot wavefields at a stride of shot_stride points (float2) in displ and accel
void launch_elastic_forces_patch(const Mesh* mp, int first_patch, int npatch, const float* displ, float* accel,
                                 int nshot = 1, long shot_stride = 0, int first_touch = 0);
// shot-minor batches: fields [point][nshot] (chain patches only)
void launch_elastic_forces_walk(const Mesh* mp, int first_patch, int npatch, const float* displ, float* accel,
                                int nshot, int first_touch);
void launch_permute_elements(const float* src, float* dst, const int* ispec_of_pos, int nspec, hipStream_t s);
void launch_pack_elements(const Mesh* mp);  // geom / moduli from the per-element arrays (patch mode)
// per-element geometry and the affine flags of the patches; returns the number of affine elements
int launch_flag_affine(Mesh* mp);
void launch_permute_elements_int(const int* src, int* dst, const int* ispec_of_pos, int nspec, hipStream_t s);
// ps: point stride (float2) of the fields -- the shots per point of a shot-minor batch, 1 otherwise
void launch_stacey(const Mesh* mp, const float* veloc, float* accel, int write_store, int nshot = 1,
                   long shot_stride = 0, int ps = 1);
void launch_stacey_backward(const Mesh* mp, float* b_accel);
// src_shot (device, one entry per source) sends each source to its own wavefield; null: all into accel
// step (device, 1-based) replaces it when not null (graph replays)
void launch_sources(const Mesh* mp, float* accel, int it, const int* src_shot = nullptr, long shot_stride = 0,
                    int ps = 1, const int* step = nullptr);
void launch_adjoint_sources(const Mesh* mp, float* accel, int it);
void launch_seismograms(const Mesh* mp, const float* field, float* dst, int seismo_current, int nshot = 1,
                        long shot_stride = 0, long dst_stride = 0, int ps = 1, const int* step = nullptr);
void launch_set_step(int* step, int value, hipStream_t s);
void launch_next_step(int* step, hipStream_t s);
void launch_kernels(const Mesh* mp);
void launch_max_norm(const Mesh* mp, const float* field, unsigned int* result);

}  // namespace sf2d
