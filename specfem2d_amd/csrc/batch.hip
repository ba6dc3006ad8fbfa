// Shot batches: nshot wavefields of one mesh and one model advanced together, each driven by its own
// sources and recorded at the same receivers.  No counterpart in the reference, which runs one source set per
// simulation (src/specfem2D/iterate_time.F90); a batch step applies the reference's GPU sequence
// (update_displacement_cuda, compute_forces_viscoelastic_cuda, compute_stacey_viscoelastic_cuda,
// compute_add_sources_el_cuda, kernel_3_a_cuda, compute_seismograms_cuda) to every shot in one launch per
// kernel.  The patch force kernel walks (patch, shot) items with the shots of a patch adjacent, so a patch's
// geometry and moduli are fetched from HBM once per step for the whole batch and re-read from L2.
#include <algorithm>

#include "sf2d_mesh.h"
#include "specfem2d_gpu_ext.h"

namespace sf2d {
namespace {

BatchState* need_batch(Mesh* mp) {
  if (!mp->batch) fail("shot batch not prepared (sf2d_batch_prepare)");
  return mp->batch;
}

void flush_batch(Mesh* mp) {
  BatchState* B = mp->batch;
  if (!B || !B->pending) return;
  launch_newmark_corrector(B->v, B->a, mp->d_rmassx, mp->d_rmassz, mp->nglob, B->pending_dto2, 1, mp->stream,
                           B->nshot, B->stride);
  B->pending = 0;
}

}  // namespace

// device addresses of the batch fields (gradient.hip checkpoints them)
void batch_fields(Mesh* mp, float** d, float** v, float** a, int* nshot, long* stride) {
  BatchState* B = need_batch(mp);
  *d = B->d;
  *v = B->v;
  *a = B->a;
  *nshot = B->nshot;
  *stride = B->stride;
}

void batch_step_into(Mesh* mp, int it, const float* d, const float* v, const float* a, float dto2_prev, float* d_out,
                     float* v_out, float* a_out, float deltat, float deltatsqover2, float deltatover2) {
  BatchState* B = need_batch(mp);
  if (it < 1 || it > mp->nstep) fail("batch step %d outside 1..%d", it, mp->nstep);
  const size_t n = (size_t)2 * B->stride * B->nshot;
  const int keep = mp->first_touch;  // first-touch assembly: the force kernels define accel
  if (!d) {  // the predictor of a state at rest is the state at rest
    for (float* p : {d_out, v_out, a_out}) check_hip(hipMemsetAsync(p, 0, n * sizeof(float), mp->stream), "zero");
  } else {
    launch_newmark_corrector_predictor_to(d, v, a, d_out, v_out, a_out, mp->d_rmassx, mp->d_rmassz, mp->nglob,
                                          dto2_prev, deltat, deltatsqover2, deltatover2, mp->stream, B->nshot,
                                          B->stride, !keep);
    if (keep) launch_zero_points(mp, a_out, B->nshot, B->stride);
  }
  for (int ph = 0; ph < 2; ph++) {
    apply_element_forces(mp, ph, d_out, a_out, B->nshot, B->stride, keep);
    if (ph == 0) {
      if (mp->stacey) launch_stacey(mp, v_out, a_out, 0, B->nshot, B->stride);
      launch_sources(mp, a_out, it - 1, B->src_shot, B->stride);
    }
  }
}

void batch_set_pending(Mesh* mp, float dto2) {
  BatchState* B = need_batch(mp);
  B->pending = 1;
  B->pending_dto2 = dto2;
}

}  // namespace sf2d

using namespace sf2d;

extern "C" {

int sf2d_batch_prepare(long* Mesh_pointer, int nshot, const int* src_shot) {
  Mesh* mp = handle(Mesh_pointer);
  if (!mp->sorted) fail("sf2d_batch_prepare: prepare_fields_elastic_device was not called");
  if (nshot < 1 || nshot > 65535) fail("sf2d_batch_prepare: %d shots outside [1, 65535]", nshot);
  std::vector<int> owner((size_t)std::max(1, mp->nsources_local), 0);
  for (int k = 0; k < mp->nsources_local; k++) {
    if (!src_shot || src_shot[k] < 0 || src_shot[k] >= nshot)
      fail("sf2d_batch_prepare: source %d belongs to shot %d outside 0..%d", k + 1, src_shot ? src_shot[k] : -1,
           nshot - 1);
    owner[k] = src_shot[k];
  }
  const long stride = ((long)mp->nglob + 63) / 64 * 64;
  const size_t n = (size_t)2 * stride * nshot;
  if (n * 3 * sizeof(float) > ((size_t)1 << 38)) fail("sf2d_batch_prepare: %d shots exceed 256 GiB of fields", nshot);
  if (mp->batch) sf2d_batch_release(Mesh_pointer);
  BatchState* B = new BatchState();
  B->nshot = nshot;
  B->stride = stride;
  B->d = device_zeros<float>(n);
  B->v = device_zeros<float>(n);
  B->a = device_zeros<float>(n);
  B->src_shot = device_copy(owner.data(), owner.size());
  B->seismo = device_zeros<float>((size_t)nshot * 2 * std::max(1, mp->nrec_local) * mp->nstep);
  mp->batch = B;
  return 0;
}

void sf2d_batch_release(long* Mesh_pointer) {
  Mesh* mp = handle(Mesh_pointer);
  BatchState* B = mp->batch;
  if (!B) return;
  if (mp->grad && mp->grad_batch) sf2d_grad_release(Mesh_pointer);
  check_hip(hipStreamSynchronize(mp->stream), "sf2d_batch_release");
  device_free(B->d);
  device_free(B->v);
  device_free(B->a);
  device_free(B->src_shot);
  device_free(B->seismo);
  delete B;
  mp->batch = nullptr;
}

// zeroes the fields and the seismograms of every shot (a run from rest)
void sf2d_batch_reset(long* Mesh_pointer) {
  Mesh* mp = handle(Mesh_pointer);
  BatchState* B = need_batch(mp);
  const size_t n = (size_t)2 * B->stride * B->nshot;
  for (float* p : {B->d, B->v, B->a}) check_hip(hipMemsetAsync(p, 0, n * sizeof(float), mp->stream), "zero");
  check_hip(hipMemsetAsync(B->seismo, 0, sizeof(float) * B->nshot * 2 * std::max(1, mp->nrec_local) * mp->nstep,
                           mp->stream),
            "zero");
  B->pending = 0;
}

// time step it (1-based) of every shot
void sf2d_batch_step(long* Mesh_pointer, int it, float deltat, float deltatsqover2, float deltatover2) {
  Mesh* mp = handle(Mesh_pointer);
  BatchState* B = need_batch(mp);
  if (it < 1 || it > mp->nstep) fail("sf2d_batch_step: step %d outside 1..%d", it, mp->nstep);
  const int keep = mp->first_touch;  // first-touch assembly: the force kernels define accel
  if (B->pending)
    launch_newmark_corrector_predictor(B->d, B->v, B->a, mp->d_rmassx, mp->d_rmassz, mp->nglob, B->pending_dto2,
                                       deltat, deltatsqover2, deltatover2, mp->stream, B->nshot, B->stride, !keep);
  else
    launch_newmark_predictor(B->d, B->v, B->a, mp->nglob, deltat, deltatsqover2, deltatover2, mp->stream, B->nshot,
                             B->stride, !keep);
  if (keep) launch_zero_points(mp, B->a, B->nshot, B->stride);
  B->pending = 0;
  for (int ph = 0; ph < 2; ph++) {
    apply_element_forces(mp, ph, B->d, B->a, B->nshot, B->stride, keep);
    if (ph == 0) {
      if (mp->stacey) launch_stacey(mp, B->v, B->a, 0, B->nshot, B->stride);
      launch_sources(mp, B->a, it - 1, B->src_shot, B->stride);
    }
  }
  B->pending = 1;
  B->pending_dto2 = deltatover2;
}

// displacement seismograms of step it (1-based) for every shot
void sf2d_batch_record(long* Mesh_pointer, int it) {
  Mesh* mp = handle(Mesh_pointer);
  BatchState* B = need_batch(mp);
  if (it < 1 || it > mp->nstep) fail("sf2d_batch_record: step %d outside 1..%d", it, mp->nstep);
  launch_seismograms(mp, B->d, B->seismo, it - 1, B->nshot, B->stride, (long)2 * mp->nrec_local * mp->nstep);
}

// seismograms [nshot][2][nrec][nstep] (x then z component, receiver-major)
void sf2d_batch_seismograms(long* Mesh_pointer, float* out) {
  Mesh* mp = handle(Mesh_pointer);
  BatchState* B = need_batch(mp);
  check_hip(hipStreamSynchronize(mp->stream), "sf2d_batch_seismograms");
  check_hip(hipMemcpy(out, B->seismo, sizeof(float) * B->nshot * 2 * mp->nrec_local * mp->nstep,
                      hipMemcpyDeviceToHost),
            "seismograms D2H");
}

// applies a pending corrector (the fields then hold the state at the end of the last step)
void sf2d_batch_flush(long* Mesh_pointer) {
  Mesh* mp = handle(Mesh_pointer);
  need_batch(mp);
  flush_batch(mp);
}

// displacement, velocity, acceleration [nglob][2] of one shot, in the caller's numbering
void sf2d_batch_get_fields(long* Mesh_pointer, int shot, float* displ, float* veloc, float* accel) {
  Mesh* mp = handle(Mesh_pointer);
  BatchState* B = need_batch(mp);
  if (shot < 0 || shot >= B->nshot) fail("sf2d_batch_get_fields: shot %d outside 0..%d", shot, B->nshot - 1);
  flush_batch(mp);
  check_hip(hipStreamSynchronize(mp->stream), "sf2d_batch_get_fields");
  const size_t off = (size_t)2 * B->stride * shot;
  if (displ) points_from_device(mp, displ, B->d + off, 2, "batch displ D2H");
  if (veloc) points_from_device(mp, veloc, B->v + off, 2, "batch veloc D2H");
  if (accel) points_from_device(mp, accel, B->a + off, 2, "batch accel D2H");
}

int sf2d_batch_shots(long* Mesh_pointer) {
  Mesh* mp = handle(Mesh_pointer);
  return mp->batch ? mp->batch->nshot : 0;
}

}  // extern "C"
