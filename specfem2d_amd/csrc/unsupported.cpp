// Entry points of the reference GPU layer whose physics this library does not implement.
//
// They exist so that a SPECFEM2D binary built with GPU_MODE links against this library unchanged; each
// terminates the run with a clear message (the reference's error convention) instead of silently doing
// nothing.  Viscoacoustic attenuation, transfers of attenuation memory variables, MPI halo exchange, kernel
// smoothing, Hessians and moving sources fall in this group.
#include "specfem2d_gpu.h"

namespace sf2d {
[[noreturn]] void fail(const char* fmt, ...);
}

#define SF2D_UNSUPPORTED(name, what) sf2d::fail("%s: %s is not supported by this library", #name, what)

extern "C" {

void transfer_boun_pot_from_device_(long*, realw*, const int*) { SF2D_UNSUPPORTED(transfer_boun_pot_from_device, "MPI halo exchange"); }
void transfer_asmbl_pot_to_device_(long*, realw*, const int*) { SF2D_UNSUPPORTED(transfer_asmbl_pot_to_device, "MPI halo exchange"); }
void transfer_boun_accel_from_device_(long*, realw*, const int*) { SF2D_UNSUPPORTED(transfer_boun_accel_from_device, "MPI halo exchange"); }
void transfer_boundary_from_device_a_(long*) { SF2D_UNSUPPORTED(transfer_boundary_from_device_a, "MPI halo exchange"); }
void prepare_boundary_on_device_(long*) { SF2D_UNSUPPORTED(prepare_boundary_on_device, "MPI halo exchange"); }
void transfer_boundary_to_device_a_(long*, realw*, const int*) { SF2D_UNSUPPORTED(transfer_boundary_to_device_a, "MPI halo exchange"); }
void transfer_asmbl_accel_to_device_(long*, realw*, const int*, const int*, const int*, const int*) { SF2D_UNSUPPORTED(transfer_asmbl_accel_to_device, "MPI halo exchange"); }
void sync_copy_from_device_(long*, int*, realw*) { SF2D_UNSUPPORTED(sync_copy_from_device, "MPI halo exchange"); }
void compute_add_moving_sources_ac_cuda_(long*, int*, int*, int*, int*, int*) { SF2D_UNSUPPORTED(compute_add_moving_sources_ac_cuda, "moving sources"); }
void compute_kernels_hess_cuda_(long*, int*, int*) { SF2D_UNSUPPORTED(compute_kernels_hess_cuda, "approximate Hessian kernels"); }
void prepare_moving_sources_cuda_(long*, int*, int*, realw*, int*, int*, realw*) { SF2D_UNSUPPORTED(prepare_moving_sources_cuda, "moving sources"); }
void prepare_arrays_gpu_(long*, realw*, realw*, realw*, realw*, realw*, realw*, int*, int*, realw*) { SF2D_UNSUPPORTED(prepare_arrays_gpu, "kernel smoothing"); }
void compute_smooth_(long*, realw*, realw*, realw*, realw*, const int*) { SF2D_UNSUPPORTED(compute_smooth, "kernel smoothing"); }
void get_smooth_(long*, realw*) { SF2D_UNSUPPORTED(get_smooth, "kernel smoothing"); }
void transfer_kernels_hess_el_tohost_(long*, realw*, int*) { SF2D_UNSUPPORTED(transfer_kernels_hess_el_tohost, "approximate Hessian kernels"); }
void transfer_kernels_hess_ac_tohost_(long*, realw*, int*) { SF2D_UNSUPPORTED(transfer_kernels_hess_ac_tohost, "approximate Hessian kernels"); }
void transfer_viscoacoustic_b_var_to_device_(int*, realw*, realw*, long*) { SF2D_UNSUPPORTED(transfer_viscoacoustic_b_var_to_device, "viscoacoustic attenuation"); }
void transfer_viscoacoustic_var_from_device_(int*, realw*, realw*, long*) { SF2D_UNSUPPORTED(transfer_viscoacoustic_var_from_device, "viscoacoustic attenuation"); }
void transfer_viscoelastic_b_var_to_device_(int*, realw*, realw*, realw*, realw*, realw*, realw*, long*) { SF2D_UNSUPPORTED(transfer_viscoelastic_b_var_to_device, "viscoelastic attenuation"); }
void transfer_viscoelastic_var_from_device_(int*, realw*, realw*, realw*, realw*, realw*, realw*, long*) { SF2D_UNSUPPORTED(transfer_viscoelastic_var_from_device, "viscoelastic attenuation"); }
void transfer_async_pot_ac_from_device_(realw*, long*) { SF2D_UNSUPPORTED(transfer_async_pot_ac_from_device, "the acoustic potential"); }
void transfer_async_pot_ac_to_device_(realw*, long*) { SF2D_UNSUPPORTED(transfer_async_pot_ac_to_device, "the acoustic potential"); }

}  // extern "C"
