/*
 * specfem2d_gpu_ext.h -- entry points of the GPU library that have no
 * counterpart in the reference's C-ABI.  They serve the host driver, the
 * tests and the benchmark: access to the handle's HIP stream (for event
 * timing and graph capture) and raw device pointers of the wavefields.
 */
#ifndef SPECFEM2D_GPU_EXT_H
#define SPECFEM2D_GPU_EXT_H

#ifdef __cplusplus
extern "C" {
#endif

/* the hipStream_t every kernel of this handle is launched on */
void* sf2d_mesh_stream(long* Mesh_pointer);
/* applies any deferred Newmark corrector (queued by kernel_3_a_cuda_ and normally fused into the next
   update_displacement_cuda_) on the handle's stream, without waiting */
void sf2d_mesh_flush(long* Mesh_pointer);
/* sf2d_mesh_flush, then waits for the handle's stream */
void sf2d_mesh_synchronize(long* Mesh_pointer);
/* zeroes the memory variables and previous strains of viscoelastic attenuation (a run restarted from rest);
   no effect without attenuation */
void sf2d_attenuation_reset(long* Mesh_pointer);
/* elements whose geometry (xix, xiz, gammax, gammaz) has the same values at all 25 points; in patch assembly,
   patches made only of such elements read one geometry record per element (SF2D_NO_AFFINE disables it) */
int sf2d_mesh_affine_elements(long* Mesh_pointer);
/* number of colours of the element list of iphase (1 outer, 2 inner) used for assembly */
int sf2d_mesh_colours(long* Mesh_pointer, int iphase);
/* host-only check of the element assembly plan built for ibool (1-based, [nspec][25]) and the phase
   lists (Fortran phase_ispec_inner_elastic(np,2), zero-padded); mode 0 coloured patches, 1 coloured
   elements, 2 atomics.  stats[6] = {colours phase 1, colours phase 2, patches, max points per patch,
   halo entries of all patches, points owned by patches}.  Returns 0 when the plan is valid; a positive
   code names the violated invariant (1-2 element order, 3 patch size, 4 slots, 5 colour conflict,
   6 point numbering not a permutation, 7 owned ranges overlap, 8 halo list, 9 unused patch point). */
int sf2d_plan_validate(int nspec, int nglob, const int* ibool_1based, int np, const int* phase_lists, int mode,
                       int* stats);
/* 1 when every patch of the coloured-patch plan for this mesh is a chain (element k's right edge is element
   k+1's left edge, no other point shared inside the patch): such meshes use the row-lane force kernel.
   Host only, same arguments as sf2d_plan_validate; first_writes (if not NULL) receives the number of element
   points whose update is the first of its point in a step (first-touch assembly: nglob when every point has
   exactly one, 0 when the plan cannot use it). */
int sf2d_plan_chains(int nspec, int nglob, const int* ibool_1based, int np, const int* phase_lists,
                     int* first_writes);
/* 1 when the handle's element forces run on a chain kernel for chain patches (0 with SF2D_NO_CHAIN set or for
   other meshes): the column-lane kernel by default, the row-lane kernel with SF2D_ROW_CHAIN */
int sf2d_mesh_chain(long* Mesh_pointer);
/* 1 when the column-lane kernel reads its point ids as one {base, id at x = 0, first-touch mask} record per chain
   row (every chain row numbered consecutively; SF2D_COL_IDS keeps the per-point ids) */
int sf2d_mesh_col_rows(long* Mesh_pointer);
/* host only: 1 when the row records hold for the chain plan of a mesh (ibool 1-based [nspec][25], phase lists as
   for sf2d_plan_validate) and reproduce every writer's point id and first-touch flag, 0 when they do not, -1 when
   the plan is not all chains; nrecords: the records (5 per chain) */
int sf2d_plan_row_records(int nspec, int nglob, const int* ibool_1based, int np, const int* phase_lists,
                          int* nrecords);
/* 1 when the fluid's element forces run on the fluid chain kernel (every patch of the fluid elements a chain;
   SF2D_NO_FLUID_CHAIN keeps the per-point kernel); known after the first acoustic step, 0 before */
int sf2d_mesh_fluid_chain(long* Mesh_pointer);
/* 1 when every element is affine and axis-aligned (xi_z = gamma_x = 0 to 2^-22 of the element's largest geometry
   entry): the chain kernel then uses its rectangle form (SF2D_NO_RECT keeps the general form) */
int sf2d_mesh_rect(long* Mesh_pointer);
/* tuning knobs (specfem2d_amd/autotune.py picks them by timing): elements per chain patch, chosen at set-up
   from SF2D_CHAIN_ELEMS (1..12, default 12), and shots per wave of the batched chain kernel (0 = default 8;
   SF2D_SHOTS_PER_WAVE at set-up) */
int sf2d_mesh_chain_elems(long* Mesh_pointer);
void sf2d_mesh_set_shots_per_wave(long* Mesh_pointer, int shots);
/* Elements per chain patch (1..12; 0 = the default 12) of the handle's element plan: call between
   prepare_constants_device_ and prepare_fields_elastic_device_; takes precedence over SF2D_CHAIN_ELEMS.  No
   counterpart in the reference (tuning of this library's assembly plan). */
void sf2d_mesh_set_chain_elems(long* Mesh_pointer, int n);
/* CPML layer in the terms of the reference's CPU path (pml_init.F90:34-1366, pml_compute*.f90), for elastic and
   acoustic elements.  The reference's GPU set-up prepare_pml_device_ carries acoustic layers only (K = 1, one d0),
   and its Fortran stops on elastic ones (prepare_gpu.f90:74).  Call after prepare_fields_elastic_device_ and
   prepare_fields_acoustic_device_, before the first time step, on a handle prepared with PML_BOUNDARY_CONDITIONS
   (forward runs; no shot batches or gradients; with viscoelastic attenuation see prepare_cpml_attenuation_device_).
     region_CPML[nspec]: 0 outside the layer, 1 CPML_X_ONLY, 2 CPML_Z_ONLY, 3 CPML_XZ;
     spec_to_PML[nspec]: 1-based row of the element in the profile arrays (0 outside);
     K_x_store .. alpha_z_store: double [NSPEC_PML][5][5] (i fastest), as pml_init.F90 leaves them (after the
       separation of the corner parameters); d = 0, K = 1, alpha = 0 along an axis the region does not damp;
     rhostore: [nspec][5][5] density (the damping terms of solid elements need it);
     DT: the time step in double precision, as the reference's coefficient routines take it;
     abs_points_elastic / _acoustic: 1-based points of the layer's outer boundary that belong to elastic / acoustic
       layer elements (determine_boundary_abs_points_PML, pml_init.F90:306-421).
   The mass matrices passed to prepare_fields_*_device_ carry the layer's terms (invert_mass_matrix.F90:133-345). */
void prepare_cpml_device_(long* Mesh_pointer, int* NSPEC_PML, int* region_CPML, int* spec_to_PML, double* K_x_store,
                          double* K_z_store, double* d_x_store, double* d_z_store, double* alpha_x_store,
                          double* alpha_z_store, float* rhostore, double* DT, int* nglob_abs_elastic,
                          int* abs_points_elastic, int* nglob_abs_acoustic, int* abs_points_acoustic);
/* the viscoelastic layer (pml_compute_memory_variables.f90:546-894, compute_forces_viscoelastic.F90:545-556): on a
   handle with ATTENUATION_VISCOELASTIC and a CPML layer, call after prepare_cpml_device_ and before the first time
   step with the Zener parameters the CPU path keeps (attenuation_model.f90:145-161), in its layout:
   inv_tau_sigma_nu1/2, phi_nu1/2, tau_epsilon_nu1/2 (NGLLX, NGLLZ, nspec, N_SLS), Mu_nu1/2 and the quality factors
   qkappa/qmu_attenuation_store (NGLLX, NGLLZ, nspec); N_SLS = 3.  Solid layer elements whose Qkappa and Qmu are below
   9999 at every point take the viscoelastic formulation, the others the elastic one (the reference's element rule).
   The time loop refuses to run a layer with attenuation without this call. */
void prepare_cpml_attenuation_device_(long* Mesh_pointer, int* N_SLS, float* inv_tau_sigma_nu1, float* phi_nu1,
                                      float* tau_epsilon_nu1, float* Mu_nu1, float* inv_tau_sigma_nu2, float* phi_nu2,
                                      float* tau_epsilon_nu2, float* Mu_nu2, float* qkappa_attenuation_store,
                                      float* qmu_attenuation_store);
/* the elastic twin of pml_boundary_acoustic_cuda_ (pml_boundary_elastic, pml_compute.f90:565-605): displ, veloc
   and accel are zero on the layer's outer boundary; the CPU path calls it after the Stacey step of iphase 1
   (compute_forces_viscoelastic_calling_routine.F90:74-77) */
void pml_boundary_elastic_cuda_(long* Mesh_pointer, int* compute_wavefield_1, int* compute_wavefield_2);
/* zeroes the CPML memory variables and previous-step fields (a run restarted from rest) */
void sf2d_pml_reset(long* Mesh_pointer);
/* stats[5] = {solid layer elements, fluid layer elements, distinct coefficient sets, solid outer-boundary points,
   fluid outer-boundary points} (zeros without a layer) */
void sf2d_pml_stats(long* Mesh_pointer, int* stats);
/* get_norm_elastic_from_device_ (same arguments and result: max over the points of |displ|, or of |b_displ| when
   *FORWARD_OR_ADJOINT == 3) that allocates nothing: no counterpart in the reference.  It is for callers that cannot
   allocate device memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
   Its reduction scratch is a file-scope device array of 2049 words (8 KiB per GPU, held until the process ends)
   instead of the handle's buffer; any field size fits it (grid-stride first pass, at most 2048 partial maxima).
   Calls that overlap on one GPU (several host threads or streams) race on that array. */
void get_norm_elastic_from_device_noalloc_(float* norm, long* Mesh_pointer, const int* FORWARD_OR_ADJOINT);
/* device pointer of a field: 0 displ 1 veloc 2 accel 3 b_displ 4 b_veloc 5 b_accel
   6 rho_kl 7 mu_kl 8 kappa_kl (NULL when not allocated).  On the device, point arrays are in the
   library's internal point numbering and kernels in its element order; the transfer entry points
   return the caller's numbering. */
void* sf2d_mesh_field(long* Mesh_pointer, int which);

/* ---------------------------------------------------------------------------
 * Exact discrete gradient of the L2 waveform misfit (reverse mode of the elastic time loop with
 * checkpointed recomputation).  No counterpart in the reference, whose SIMULATION_TYPE=3 path reconstructs
 * the forward field backwards instead; the host driver (specfem2d_driver.h) sequences these calls. */
/* allocates checkpoints every `chunk` steps, one segment of `chunk` step states and the accumulators;
   returns the number of checkpoints */
int sf2d_grad_prepare(long* Mesh_pointer, int chunk, int nrec);
void sf2d_grad_release(long* Mesh_pointer);
/* copies the fields after a forward step (d_n, predicted v, unscaled force r_n; corrector pending) into
   checkpoint (kind 0) or segment (kind 1) slot */
void sf2d_grad_save_state(long* Mesh_pointer, int kind, int slot);
/* restores a checkpoint into the fields and re-arms the corrector of that step */
void sf2d_grad_restore_checkpoint(long* Mesh_pointer, int slot, float dto2);
/* residuals dt (s - o) [nstep][nrec][2]; zeroes the adjoint state and the accumulators */
void sf2d_grad_begin_reverse(long* Mesh_pointer, const float* res);
/* reverse step n (1-based) from segment slot; first = 1 for n = NSTEP */
void sf2d_grad_reverse_step(long* Mesh_pointer, int n, int slot, int first, float dt);
/* adds the FP32 accumulators of the segment into the FP64 totals */
void sf2d_grad_flush(long* Mesh_pointer);
/* d chi / d kappa, d mu [nspec][25] (ispec order), d chi / d M [nglob][2] (caller numbering), Stacey traction
   parts d chi / d (rho vp), d (rho vs) per absorbing face point [nabs][5] */
void sf2d_grad_result(long* Mesh_pointer, double* g_kappa, double* g_mu, double* g_mass, double* g_rho_vp,
                      double* g_rho_vs);
/* gradient buffers over the shot batch (sf2d_batch_prepare): the calls above then checkpoint and reverse all
   shots together, residuals are [nstep][nshot][nrec][2] and the results are sums over the shots */
int sf2d_batch_grad_prepare(long* Mesh_pointer, int chunk);
/* shot batch: recomputes steps n0+1..n1 directly into segment slots 0..n1-n0-1, starting from checkpoint slot
   ckpt (the state after step n0) or from rest (ckpt = -1 with n0 = 0); replaces restore + forward + saves */
void sf2d_grad_recompute_segment(long* Mesh_pointer, int ckpt, int n0, int n1, float dt);

/* ---------------------------------------------------------------------------
 * Shot batches: nshot wavefields of one mesh and model, each driven by its own sources and recorded at the
 * handle's receivers, advanced with one launch per kernel for the whole batch.  No counterpart in the
 * reference, which runs one source set per simulation (iterate_time.F90); a step applies the sequence of
 * update_displacement_cuda, compute_forces_viscoelastic_cuda, compute_stacey_viscoelastic_cuda,
 * compute_add_sources_el_cuda and kernel_3_a_cuda to every shot.  The handle is prepared as usual with the
 * sources of all shots (prepare_constants_device). */
/* src_shot[k] = shot (0-based) of source k; allocates the fields at rest */
int sf2d_batch_prepare(long* Mesh_pointer, int nshot, const int* src_shot);
void sf2d_batch_release(long* Mesh_pointer);
/* all fields and seismograms to zero (a run from rest) */
void sf2d_batch_reset(long* Mesh_pointer);
/* time step it (1-based) of every shot, with the reference's Newmark constants */
void sf2d_batch_step(long* Mesh_pointer, int it, float deltat, float deltatsqover2, float deltatover2);
/* steps it_begin..it_end (1-based) of every shot, with the displacement seismograms when record != 0 and a copy
   of the fields into gradient checkpoint it / ckpt_every - 1 after every ckpt_every-th step (ckpt_every > 0;
   none after NSTEP).  With SF2D_BATCH_GRAPH set, each step after the first of a run replays one captured HIP
   graph of the step (its index read from device memory) instead of launching its kernels. */
void sf2d_batch_run(long* Mesh_pointer, int it_begin, int it_end, float deltat, float deltatsqover2,
                    float deltatover2, int record, int ckpt_every);
/* records the displacement seismograms of step it (1-based) */
void sf2d_batch_record(long* Mesh_pointer, int it);
/* seismograms [nshot][2][nrec][nstep] */
void sf2d_batch_seismograms(long* Mesh_pointer, float* out);
/* applies the pending Newmark corrector */
void sf2d_batch_flush(long* Mesh_pointer);
/* fields of one shot, [nglob][2] each in the caller's numbering (NULL skips one) */
void sf2d_batch_get_fields(long* Mesh_pointer, int shot, float* displ, float* veloc, float* accel);
/* shots of the batch (0 without one) */
int sf2d_batch_shots(long* Mesh_pointer);
/* 1 when the batch fields are shot-minor (point-major, nshot values per point: the lane-per-shot kernels, on
   request with SF2D_SHOT_MINOR for chain meshes and 2, 4, ..., 64 shots), 0 when shot-major (the default), -1
   without a batch */
int sf2d_batch_shot_minor(long* Mesh_pointer);

/* replaces kappa, mu, rho vp, rho vs ([nspec][25], ispec order) and rmassx, rmassz ([nglob], caller numbering);
   NULL leaves an array unchanged */
void sf2d_mesh_set_materials(long* Mesh_pointer, const float* kappav, const float* muv, const float* rho_vp,
                             const float* rho_vs, const float* rmassx, const float* rmassz);

#ifdef __cplusplus
}
#endif
#endif
