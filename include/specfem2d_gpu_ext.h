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
/* device pointer of a field: 0 displ 1 veloc 2 accel 3 b_displ 4 b_veloc 5 b_accel
   6 rho_kl 7 mu_kl 8 kappa_kl (NULL when not allocated) */
void* sf2d_mesh_field(long* Mesh_pointer, int which);

#ifdef __cplusplus
}
#endif
#endif
