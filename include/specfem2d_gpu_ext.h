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
/* number of colours of the element list of iphase (1 outer, 2 inner) used for assembly */
int sf2d_mesh_colours(long* Mesh_pointer, int iphase);
/* host-only check of the element assembly plan built for ibool (1-based, [nspec][25]) and the phase
   lists (Fortran phase_ispec_inner_elastic(np,2), zero-padded); mode 0 coloured patches, 1 coloured
   elements, 2 atomics.  stats = {colours phase 1, colours phase 2, patches, max points per patch}.
   Returns 0 when the plan is valid. */
int sf2d_plan_validate(int nspec, int nglob, const int* ibool_1based, int np, const int* phase_lists, int mode,
                       int* stats);
/* device pointer of a field: 0 displ 1 veloc 2 accel 3 b_displ 4 b_veloc 5 b_accel
   6 rho_kl 7 mu_kl 8 kappa_kl (NULL when not allocated) */
void* sf2d_mesh_field(long* Mesh_pointer, int which);

#ifdef __cplusplus
}
#endif
#endif
