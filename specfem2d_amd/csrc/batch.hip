// Shot batches: nshot wavefields of one mesh and one model advanced together, each driven by its own
// sources and recorded at the same receivers.  No counterpart in the reference, which runs one source set per
// simulation (src/specfem2D/iterate_time.F90); a batch step applies the reference's GPU sequence
// (update_displacement_cuda, compute_forces_viscoelastic_cuda, compute_stacey_viscoelastic_cuda,
// compute_add_sources_el_cuda, kernel_3_a_cuda, compute_seismograms_cuda) to every shot in one launch per
// kernel.  Two layouts: shot-major (shot s at s * stride, the default) runs the patch kernels over (patch, shot
// group) items, a patch's element data fetched from HBM once per step and re-read from L2; shot-minor (point g of
// shot s at g * nshot + s, SF2D_SHOT_MINOR) runs the lane-per-shot kernels, whose element data is shared by the
// lanes of an item and whose point accesses are nshot consecutive values.
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
                           B->nshot, B->stride, B->ps);
  B->pending = 0;
}

size_t field_floats(const Mesh* mp, const BatchState* B) {
  return B->ps > 1 ? (size_t)2 * mp->nglob * B->nshot : (size_t)2 * B->stride * B->nshot;
}

// forces, Stacey and sources of step it (or *step on the device) on state (d, v, a) whose accel holds the
// predictor's output
void batch_forces(Mesh* mp, const BatchState* B, int it, const float* d, const float* v, float* a, int keep,
                  const int* step = nullptr) {
  for (int ph = 0; ph < 2; ph++) {
    apply_element_forces(mp, ph, d, a, B->nshot, B->stride, keep, B->ps);
    if (ph == 0) {
      if (mp->stacey) launch_stacey(mp, v, a, 0, B->nshot, B->stride, B->ps);
      launch_sources(mp, a, it - 1, B->src_shot, B->stride, B->ps, step);
    }
  }
}

// one step of every shot on the batch fields: predictor (with the pending corrector), forces, Stacey, sources
void batch_step_fields(Mesh* mp, BatchState* B, int it, float deltat, float deltatsqover2, float deltatover2,
                       const int* step) {
  const int keep = mp->first_touch;  // first-touch assembly: the force kernels define accel
  if (B->pending)
    launch_newmark_corrector_predictor(B->d, B->v, B->a, mp->d_rmassx, mp->d_rmassz, mp->nglob, B->pending_dto2,
                                       deltat, deltatsqover2, deltatover2, mp->stream, B->nshot, B->stride, !keep,
                                       B->ps);
  else
    launch_newmark_predictor(B->d, B->v, B->a, mp->nglob, deltat, deltatsqover2, deltatover2, mp->stream, B->nshot,
                             B->stride, !keep, B->ps);
  if (keep) launch_zero_points(mp, B->a, B->nshot, B->stride, B->ps);
  batch_forces(mp, B, it, B->d, B->v, B->a, keep, step);
  B->pending = 1;
  B->pending_dto2 = deltatover2;
}

void drop_graph(BatchState* B) {
  if (B->graph_exec) (void)hipGraphExecDestroy(B->graph_exec);
  if (B->graph) (void)hipGraphDestroy(B->graph);
  B->graph_exec = nullptr;
  B->graph = nullptr;
}

// captures one steady-state step (corrector pending with dto2 = deltatover2) reading its index from d_step
void capture_step(Mesh* mp, BatchState* B, float deltat, float deltatsqover2, float deltatover2, int record) {
  drop_graph(B);
  if (!B->d_step) B->d_step = device_zeros<int>(1);
  const int pending = B->pending;
  const float pdto2 = B->pending_dto2;
  check_hip(hipStreamBeginCapture(mp->stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
  B->pending = 1;
  B->pending_dto2 = deltatover2;
  batch_step_fields(mp, B, 1, deltat, deltatsqover2, deltatover2, B->d_step);
  if (record && mp->nrec_local > 0)
    launch_seismograms(mp, B->d, B->seismo, 0, B->nshot, B->stride, (long)2 * mp->nrec_local * mp->nstep, B->ps,
                       B->d_step);
  launch_next_step(B->d_step, mp->stream);
  check_hip(hipStreamEndCapture(mp->stream, &B->graph), "hipStreamEndCapture");
  check_hip(hipGraphInstantiate(&B->graph_exec, B->graph, nullptr, nullptr, 0), "hipGraphInstantiate");
  B->pending = pending;
  B->pending_dto2 = pdto2;
  B->graph_dt[0] = deltat;
  B->graph_dt[1] = deltatsqover2;
  B->graph_dt[2] = deltatover2;
  B->graph_record = record;
}

}  // namespace

// device addresses of the batch fields (gradient.hip checkpoints them)
void batch_fields(Mesh* mp, float** d, float** v, float** a, int* nshot, long* stride, int* ps) {
  BatchState* B = need_batch(mp);
  *d = B->d;
  *v = B->v;
  *a = B->a;
  *nshot = B->nshot;
  *stride = B->stride;
  *ps = B->ps;
}

void batch_step_into(Mesh* mp, int it, const float* d, const float* v, const float* a, float dto2_prev, float* d_out,
                     float* v_out, float* a_out, float deltat, float deltatsqover2, float deltatover2) {
  BatchState* B = need_batch(mp);
  if (it < 1 || it > mp->nstep) fail("batch step %d outside 1..%d", it, mp->nstep);
  const size_t n = field_floats(mp, B);
  const int keep = mp->first_touch;  // first-touch assembly: the force kernels define accel
  if (!d) {  // the predictor of a state at rest is the state at rest
    for (float* p : {d_out, v_out, a_out}) check_hip(hipMemsetAsync(p, 0, n * sizeof(float), mp->stream), "zero");
  } else {
    launch_newmark_corrector_predictor_to(d, v, a, d_out, v_out, a_out, mp->d_rmassx, mp->d_rmassz, mp->nglob,
                                          dto2_prev, deltat, deltatsqover2, deltatover2, mp->stream, B->nshot,
                                          B->stride, !keep, B->ps);
    if (keep) launch_zero_points(mp, a_out, B->nshot, B->stride, B->ps);
  }
  batch_forces(mp, B, it, d_out, v_out, a_out, keep);
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
  if (mp->nspec_acoustic > 0) fail("sf2d_batch_prepare: shot batches of meshes with fluid elements are not supported");
  if (nshot < 1 || nshot > 65535) fail("sf2d_batch_prepare: %d shots outside [1, 65535]", nshot);
  std::vector<int> owner((size_t)std::max(1, mp->nsources_local), 0);
  for (int k = 0; k < mp->nsources_local; k++) {
    if (!src_shot || src_shot[k] < 0 || src_shot[k] >= nshot)
      fail("sf2d_batch_prepare: source %d belongs to shot %d outside 0..%d", k + 1, src_shot ? src_shot[k] : -1,
           nshot - 1);
    owner[k] = src_shot[k];
  }
  // shot-minor on request (SF2D_SHOT_MINOR) when the lane-per-shot kernels apply: chain patches, 32-bit offsets,
  // a power of two of at most 64 shots (the reverse sweep sums the shots of a point across lanes).  Shot-major is
  // the default: with one wave walking a whole patch, the lane-per-shot force kernel has too few waves per SIMD
  // to hide its load latency and runs slower per shot than the patch kernels.
  const bool minor = getenv("SF2D_SHOT_MINOR") && !getenv("SF2D_SHOT_MAJOR") && nshot > 1 && nshot <= 64 &&
                     (nshot & (nshot - 1)) == 0 && mp->chain && mp->assembly == 0 &&
                     (size_t)mp->nglob * nshot * 8 < ((size_t)1 << 32);
  const long stride = minor ? 1 : ((long)mp->nglob + 63) / 64 * 64;
  const size_t n = minor ? (size_t)2 * mp->nglob * nshot : (size_t)2 * stride * nshot;
  if (n * 3 * sizeof(float) > ((size_t)1 << 38)) fail("sf2d_batch_prepare: %d shots exceed 256 GiB of fields", nshot);
  if (mp->batch) sf2d_batch_release(Mesh_pointer);
  BatchState* B = new BatchState();
  B->nshot = nshot;
  B->stride = stride;
  B->ps = minor ? nshot : 1;
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
  drop_graph(B);
  device_free(B->d_step);
  delete B;
  mp->batch = nullptr;
}

// zeroes the fields and the seismograms of every shot (a run from rest)
void sf2d_batch_reset(long* Mesh_pointer) {
  Mesh* mp = handle(Mesh_pointer);
  BatchState* B = need_batch(mp);
  const size_t n = field_floats(mp, B);
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
  batch_step_fields(mp, B, it, deltat, deltatsqover2, deltatover2, nullptr);
}

// steps it_begin..it_end of every shot, recording the seismograms when record is set and copying the fields into
// gradient checkpoint it / ckpt_every - 1 after every ckpt_every-th step (ckpt_every > 0; not after NSTEP).
// With SF2D_BATCH_GRAPH set, each step after the first of a run is one replay of a captured graph.
void sf2d_batch_run(long* Mesh_pointer, int it_begin, int it_end, float deltat, float deltatsqover2,
                    float deltatover2, int record, int ckpt_every) {
  Mesh* mp = handle(Mesh_pointer);
  BatchState* B = need_batch(mp);
  if (it_begin < 1 || it_end > mp->nstep || it_end < it_begin)
    fail("sf2d_batch_run: steps %d..%d outside 1..%d", it_begin, it_end, mp->nstep);
  // graph replays on request: at the sizes measured (80x60 elements and up, 8 shots and more) a batch step is
  // device-bound and a replay per step costs more than the launches it replaces
  const bool use_graph = getenv("SF2D_BATCH_GRAPH") != nullptr;
  bool counting = false;  // d_step holds the index of the next replay
  for (int it = it_begin; it <= it_end; it++) {
    if (!use_graph || !B->pending || B->pending_dto2 != deltatover2) {
      batch_step_fields(mp, B, it, deltat, deltatsqover2, deltatover2, nullptr);
      if (record && mp->nrec_local > 0)
        launch_seismograms(mp, B->d, B->seismo, it - 1, B->nshot, B->stride, (long)2 * mp->nrec_local * mp->nstep,
                           B->ps);
      counting = false;
    } else {
      if (!B->graph_exec || B->graph_record != record || B->graph_dt[0] != deltat ||
          B->graph_dt[1] != deltatsqover2 || B->graph_dt[2] != deltatover2)
        capture_step(mp, B, deltat, deltatsqover2, deltatover2, record);
      if (!counting) launch_set_step(B->d_step, it, mp->stream);
      counting = true;
      check_hip(hipGraphLaunch(B->graph_exec, mp->stream), "hipGraphLaunch");
      B->pending = 1;
      B->pending_dto2 = deltatover2;
    }
    if (ckpt_every > 0 && it % ckpt_every == 0 && it < mp->nstep) sf2d_grad_save_state(Mesh_pointer, 0, it / ckpt_every - 1);
  }
}

// displacement seismograms of step it (1-based) for every shot
void sf2d_batch_record(long* Mesh_pointer, int it) {
  Mesh* mp = handle(Mesh_pointer);
  BatchState* B = need_batch(mp);
  if (it < 1 || it > mp->nstep) fail("sf2d_batch_record: step %d outside 1..%d", it, mp->nstep);
  launch_seismograms(mp, B->d, B->seismo, it - 1, B->nshot, B->stride, (long)2 * mp->nrec_local * mp->nstep, B->ps);
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
  if (B->ps == 1) {
    const size_t off = (size_t)2 * B->stride * shot;
    if (displ) points_from_device(mp, displ, B->d + off, 2, "batch displ D2H");
    if (veloc) points_from_device(mp, veloc, B->v + off, 2, "batch veloc D2H");
    if (accel) points_from_device(mp, accel, B->a + off, 2, "batch accel D2H");
    return;
  }
  // shot-minor: the shot's values are every nshot-th float2
  std::vector<float> all(field_floats(mp, B));
  const float* dev[3] = {B->d, B->v, B->a};
  float* out[3] = {displ, veloc, accel};
  for (int f = 0; f < 3; f++) {
    if (!out[f]) continue;
    check_hip(hipMemcpy(all.data(), dev[f], all.size() * sizeof(float), hipMemcpyDeviceToHost), "batch field D2H");
    for (int g = 0; g < mp->nglob; g++) {
      const size_t id = mp->point_id.empty() ? (size_t)g : (size_t)mp->point_id[g];
      const size_t q = (id * B->nshot + shot) * 2;
      out[f][2 * (size_t)g] = all[q];
      out[f][2 * (size_t)g + 1] = all[q + 1];
    }
  }
}

int sf2d_batch_shots(long* Mesh_pointer) {
  Mesh* mp = handle(Mesh_pointer);
  return mp->batch ? mp->batch->nshot : 0;
}

// 1 when the batch fields are shot-minor (the lane-per-shot force kernel), 0 shot-major, -1 without a batch
int sf2d_batch_shot_minor(long* Mesh_pointer) {
  Mesh* mp = handle(Mesh_pointer);
  return mp->batch ? (mp->batch->ps > 1 ? 1 : 0) : -1;
}

}  // extern "C"
