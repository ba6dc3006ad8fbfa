/*
 * specfem2d_gpu.h -- the C-ABI of the MI355X (gfx950) time-stepping library.
 *
 * These are the entry points that the SPECFEM2D Fortran solver calls when
 * GPU_MODE = .true.  Names carry the Fortran mangling of the reference build
 * (FC_FUNC_(name,NAME) = name##_), every argument is passed by reference,
 * logical flags arrive as int (non-zero = true), array indices stored inside
 * arrays are 1-based and arrays are column-major (i fastest), exactly as the
 * reference's src/gpu/specfem2D_gpu_cuda_method_stubs.c declares them.  The
 * handle is an opaque host pointer written into the caller's integer(kind=8)
 * `Mesh_pointer`.
 *
 * Errors: like the reference (src/gpu/helper_functions.cu:95-121), a failing
 * call prints a message, appends it to OUTPUT_FILES/error_message_*.txt when
 * that directory exists, and terminates the process with EXIT_FAILURE.
 * Implemented: forward and adjoint (SIMULATION_TYPE 1 and 3) runs of solid,
 * fluid (acoustic) and coupled fluid-solid meshes with Stacey edges, free
 * surfaces, sources, seismograms, the stored Stacey terms of SAVE_FORWARD,
 * the reconstructed fields and the solid and fluid kernels; anisotropic
 * elements (forward runs); viscoelastic attenuation (forward runs); CPML layers (forward, elastic and acoustic
 * elements: prepare_pml_device_, pml_boundary_acoustic_cuda_, and the
 * extensions of specfem2d_gpu_ext.h).  Entry points of physics this library
 * does not implement (viscoacoustic attenuation, transfers of attenuation
 * memory variables, MPI halo exchange, kernel smoothing, Hessians, moving
 * sources) are exported so that the reference binary links, and fail loudly
 * when called.
 *
 * Each declaration cites the reference definition it replaces.
 */
#ifndef SPECFEM2D_GPU_H
#define SPECFEM2D_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

typedef float realw;

/* ---------------------------------------------------------------------------
 * device set-up                                                               */
/* src/gpu/initialize_cuda.cu:50 */
void initialize_cuda_device_(int* myrank_f, int* ncuda_devices);
/* src/gpu/initialize_cuda.cu:264 (no-op: one process per GPU, no MPI here) */
void initialize_cuda_aware_mpi_(void);
/* src/gpu/prepare_mesh_constants_cuda.cu:83 */
void prepare_constants_device_(long* Mesh_pointer, int* h_NGLLX, int* NSPEC_AB, int* NGLOB_AB,
                               realw* h_xix, realw* h_xiz, realw* h_gammax, realw* h_gammaz,
                               realw* h_kappav, realw* h_muv, int* h_ibool,
                               int* num_interfaces_ext_mesh, int* max_nibool_interfaces_ext_mesh,
                               int* h_nibool_interfaces_ext_mesh, int* h_ibool_interfaces_ext_mesh,
                               realw* h_hprime_xx, realw* h_hprimewgll_xx, realw* h_wxgll,
                               int* STACEY_BOUNDARY_CONDITIONS, int* PML_BOUNDARY_CONDITIONS,
                               int* h_ispec_is_inner, int* nsources_local_f, realw* h_sourcearrays,
                               realw* h_source_time_function, int* NSTEP, int* h_ispec_selected_source,
                               int* h_ispec_selected_rec_loc, int* nrec_local, realw* h_cosrot, realw* h_sinrot,
                               int* SIMULATION_TYPE, int* P_SV, int* nspec_acoustic, int* nspec_elastic,
                               int* ispec_is_acoustic, int* ispec_is_elastic, int* h_myrank, int* SAVE_FORWARD,
                               realw* h_xir_store, realw* h_gammar_store, int* h_NSIGTYPE, int* h_seismotypeVec,
                               int* nlength_seismogram);
/* src/gpu/prepare_mesh_constants_cuda.cu:527 */
void prepare_fields_elastic_device_(long* Mesh_pointer, realw* rmassx, realw* rmassz,
                                    int* num_phase_ispec_elastic, int* phase_ispec_inner_elastic,
                                    int* ispec_is_anisotropic, int* ANISOTROPY, realw* c11store, realw* c12store,
                                    realw* c13store, realw* c15store, realw* c23store, realw* c25store,
                                    realw* c33store, realw* c35store, realw* c55store, int* h_ninterface_elastic,
                                    int* h_inum_interfaces_elastic, int* ATTENUATION_VISCOELASTIC,
                                    realw* h_A_newmark_mu, realw* h_B_newmark_mu, realw* h_A_newmark_kappa,
                                    realw* h_B_newmark_kappa);
/* src/gpu/prepare_mesh_constants_cuda.cu:711 */
void prepare_fields_elastic_adj_dev_(long* Mesh_pointer, int* size_f, int* APPROXIMATE_HESS_KL,
                                     int* ATTENUATION_VISCOELASTIC, int* NO_BACKWARD_RECONSTRUCTION);
/* src/gpu/prepare_mesh_constants_cuda.cu:926 */
void prepare_stacey_device_(long* Mesh_pointer, int* ACOUSTIC_SIMULATION, int* ELASTIC_SIMULATION, realw* rho_vp,
                            realw* rho_vs, int* h_nspec_bottom, int* h_nspec_left, int* h_nspec_right,
                            int* h_nspec_top, int* h_abs_boundary_ispec, int* h_abs_boundary_ij,
                            realw* h_abs_boundary_normal, realw* h_abs_boundary_jacobian1Dw,
                            int* h_num_abs_boundary_faces, int* h_edge_abs, int* h_ib_bottom, int* h_ib_left,
                            int* h_ib_right, int* h_ib_top);
/* src/gpu/prepare_mesh_constants_cuda.cu:815 */
void prepare_sim2_or_3_const_device_(long* Mesh_pointer, int* nadj_rec_local, realw* h_source_adjoint, int* NSTEP);
/* src/gpu/prepare_mesh_constants_cuda.cu:1139 */
void prepare_cleanup_device_(long* Mesh_pointer, int* ACOUSTIC_SIMULATION, int* ELASTIC_SIMULATION, int* ANISOTROPY,
                             int* APPROXIMATE_HESS_KL, int* ATTENUATION_VISCOACOUSTIC,
                             int* ATTENUATION_VISCOELASTIC, int* NO_BACKWARD_RECONSTRUCTION,
                             realw* h_no_backward_acoustic_buffer);

/* ---------------------------------------------------------------------------
 * field transfers (size = NDIM*NGLOB floats, interleaved x,z)                 */
/* src/gpu/transfer_fields_cuda.cu:53 */
void transfer_fields_el_to_device_(int* size, realw* displ, realw* veloc, realw* accel, long* Mesh_pointer);
/* src/gpu/transfer_fields_cuda.cu:69 */
void transfer_fields_el_from_device_(int* size, realw* displ, realw* veloc, realw* accel, long* Mesh_pointer);
/* src/gpu/transfer_fields_cuda.cu:85 */
void transfer_b_fields_to_device_(int* size, realw* b_displ, realw* b_veloc, realw* b_accel, long* Mesh_pointer);
/* src/gpu/transfer_fields_cuda.cu:102 */
void transfer_b_fields_from_device_(int* size, realw* b_displ, realw* b_veloc, realw* b_accel, long* Mesh_pointer);
/* src/gpu/transfer_fields_cuda.cu:119 */
void transfer_accel_to_device_(int* size, realw* accel, long* Mesh_pointer);
/* src/gpu/transfer_fields_cuda.cu:133 */
void transfer_accel_from_device_(int* size, realw* accel, long* Mesh_pointer);
/* src/gpu/transfer_fields_cuda.cu:147 */
void transfer_b_accel_from_device_(int* size, realw* b_accel, long* Mesh_pointer);
/* src/gpu/transfer_fields_cuda.cu:162 */
void transfer_b_displ_from_device_(int* size, realw* displ, long* Mesh_pointer);
/* src/gpu/transfer_fields_cuda.cu:176 */
void transfer_displ_from_device_(int* size, realw* displ, long* Mesh_pointer);
/* src/gpu/transfer_fields_cuda.cu:194 */
void transfer_kernels_el_to_host_(long* Mesh_pointer, realw* h_rho_kl, realw* h_mu_kl, realw* h_kappa_kl,
                                  int* NSPEC_AB);
/* src/gpu/transfer_fields_cuda.cu:598 */
void transfer_compute_kernel_answers_from_device_(long* Mesh_pointer, realw* rho_kl, int* size_rho, realw* mu_kl,
                                                  int* size_mu, realw* kappa_kl, int* size_kappa);
/* src/gpu/transfer_fields_cuda.cu:617 */
void transfer_compute_kernel_fields_from_device_(
    long* Mesh_pointer, realw* accel, int* size_accel, realw* b_displ, int* size_b_displ, realw* epsilondev_xx,
    realw* epsilondev_yy, realw* epsilondev_xy, realw* epsilondev_xz, realw* epsilondev_yz, int* size_epsilondev,
    realw* b_epsilondev_xx, realw* b_epsilondev_yy, realw* b_epsilondev_xy, realw* b_epsilondev_xz,
    realw* b_epsilondev_yz, int* size_b_epsilondev, realw* rho_kl, int* size_rho, realw* mu_kl, int* size_mu,
    realw* kappa_kl, int* size_kappa, realw* epsilon_trace_over_3, realw* b_epsilon_trace_over_3,
    int* size_epsilon_trace_over_3);

/* ---------------------------------------------------------------------------
 * elastic time step                                                           */
/* src/gpu/update_displacement_cuda.cu:52 -- Newmark predictor (both wavefields when SIMULATION_TYPE=3) */
void update_displacement_cuda_(long* Mesh_pointer, realw* deltat_F, realw* deltatsqover2_F, realw* deltatover2_F,
                               realw* b_deltat_F, realw* b_deltatsqover2_F, realw* b_deltatover2_F);
/* src/gpu/compute_forces_viscoelastic_cuda.cu:327 -- element forces, assembled into accel */
void compute_forces_viscoelastic_cuda_(long* Mesh_pointer, int* iphase, realw* deltat, int* nspec_outer_elastic,
                                       int* nspec_inner_elastic, int* ANISOTROPY, int* ATTENUATION_VISCOELASTIC,
                                       int* compute_wavefield_1, int* compute_wavefield_2);
/* src/gpu/compute_stacey_viscoelastic_cuda.cu:40 */
void compute_stacey_viscoelastic_cuda_(long* Mesh_pointer, int* iphasef, realw* h_b_absorb_elastic_left,
                                       realw* h_b_absorb_elastic_right, realw* h_b_absorb_elastic_top,
                                       realw* h_b_absorb_elastic_bottom, int* compute_wavefield_1,
                                       int* compute_wavefield_2, int* UNDO_ATTENUATION_AND_OR_PML);
/* src/gpu/compute_add_sources_viscoelastic_cuda.cu:46 */
void compute_add_sources_el_cuda_(long* Mesh_pointer, int* iphasef, int* itf);
/* src/gpu/compute_add_sources_viscoelastic_cuda.cu:86 */
void compute_add_sources_el_s3_cuda_(long* Mesh_pointer, int* iphasef, int* itf);
/* src/gpu/compute_add_sources_viscoelastic_cuda.cu:131 */
void add_sources_el_sim_type_2_or_3_(long* Mesh_pointer, int* iphasef, int* itf, int* nadj_rec_local, int* NSTEP);
/* src/gpu/update_displacement_cuda.cu:221 -- mass matrix and Newmark corrector */
void kernel_3_a_cuda_(long* Mesh_pointer, realw* deltatover2_F, realw* b_deltatover2_F, int* compute_wavefield_1,
                      int* compute_wavefield_2);
/* src/gpu/update_displacement_cuda.cu:273 -- corrector without the mass multiply (unused by the reference) */
void kernel_3_b_cuda_(long* Mesh_pointer, realw* deltatover2_F, realw* b_deltatover2_F);
/* src/gpu/write_seismograms_cuda.cu:41 */
void compute_seismograms_cuda_(long* Mesh_pointer_f, int* i_sigf, double* sisux, double* sisuz, int* seismo_currentf,
                               int* nlength_seismogramf, int* ELASTIC_SIMULATION, int* ACOUSTIC_SIMULATION,
                               int* USE_TRICK_FOR_BETTER_PRESSURE, int* ATTENUATION_VISCOELASTIC, int* itf,
                               int* it_endf);
/* src/gpu/compute_kernels_cuda.cu:47 -- rho/mu/kappa kernel accumulation */
void compute_kernels_elastic_cuda_(long* Mesh_pointer, realw* deltat);

/* ---------------------------------------------------------------------------
 * housekeeping and checks                                                     */
/* src/gpu/check_fields_cuda.cu:358 */
void get_norm_elastic_from_device_(realw* norm, long* Mesh_pointer, const int* FORWARD_OR_ADJOINT);
/* src/gpu/check_fields_cuda.cu:165 */
void get_free_device_memory_(realw* free, realw* used, realw* total);
/* src/gpu/check_fields_cuda.cu:149 */
void output_free_device_memory_(int* myrank_f);
/* src/gpu/check_fields_cuda.cu:484 */
void get_max_accel_(int* itf, int* sizef, long* Mesh_pointer);
/* src/gpu/check_fields_cuda.cu:705 */
void check_max_norm_displ_gpu_(int* size, realw* displ, long* Mesh_pointer, int* announceID);
/* src/gpu/check_fields_cuda.cu:725 */
void check_max_norm_vector_(int* size, realw* vector1, int* announceID);
/* src/gpu/check_fields_cuda.cu:752 */
void check_max_norm_displ_(int* size, realw* displ, int* announceID);
/* src/gpu/check_fields_cuda.cu:770 */
void check_max_norm_b_displ_gpu_(int* size, realw* b_displ, long* Mesh_pointer, int* announceID);
/* src/gpu/check_fields_cuda.cu:799 */
void check_max_norm_b_accel_gpu_(int* size, realw* b_accel, long* Mesh_pointer, int* announceID);
/* src/gpu/check_fields_cuda.cu:821 */
void check_max_norm_b_veloc_gpu_(int* size, realw* b_veloc, long* Mesh_pointer, int* announceID);
/* src/gpu/check_fields_cuda.cu:843 */
void check_max_norm_b_displ_(int* size, realw* b_displ, int* announceID);
/* src/gpu/check_fields_cuda.cu:861 */
void check_max_norm_b_accel_(int* size, realw* b_accel, int* announceID);
/* src/gpu/check_fields_cuda.cu:879 */
void check_error_vectors_(int* sizef, realw* vector1, realw* vector2);
/* src/gpu/helper_functions.cu:281 */
void pause_for_debug_(void);

/* ---------------------------------------------------------------------------
 * exported, fail loudly: physics and parallel paths outside this library      */
/* src/gpu/assemble_MPI_scalar_cuda.cu:48 */
void transfer_boun_pot_from_device_(long* Mesh_pointer, realw* h_send_potential_dot_dot_buffer,
                                    const int* FORWARD_OR_ADJOINT);
/* src/gpu/assemble_MPI_scalar_cuda.cu:118 */
void transfer_asmbl_pot_to_device_(long* Mesh_pointer, realw* h_buffer_recv_scalar_gpu, const int* FORWARD_OR_ADJOINT);
/* src/gpu/assemble_MPI_vector_cuda.cu:49 */
void transfer_boun_accel_from_device_(long* Mesh_pointer, realw* h_send_accel_buffer, const int* FORWARD_OR_ADJOINT);
/* src/gpu/assemble_MPI_vector_cuda.cu:115 */
void transfer_boundary_from_device_a_(long* Mesh_pointer);
/* src/gpu/assemble_MPI_vector_cuda.cu:157 */
void prepare_boundary_on_device_(long* Mesh_pointer);
/* src/gpu/assemble_MPI_vector_cuda.cu:197 */
void transfer_boundary_to_device_a_(long* Mesh_pointer, realw* buffer_recv_vector_gpu,
                                    const int* max_nibool_interfaces_ext_mesh);
/* src/gpu/assemble_MPI_vector_cuda.cu:230 */
void transfer_asmbl_accel_to_device_(long* Mesh_pointer, realw* buffer_recv_vector_gpu,
                                     const int* max_nibool_interfaces_ext_mesh, const int* nibool_interfaces_ext_mesh,
                                     const int* ibool_interfaces_ext_mesh, const int* FORWARD_OR_ADJOINT);
/* src/gpu/assemble_MPI_vector_cuda.cu:314 */
void sync_copy_from_device_(long* Mesh_pointer, int* iphase, realw* send_buffer);
/* src/gpu/check_fields_cuda.cu:240 */
void get_norm_acoustic_from_device_(realw* norm, long* Mesh_pointer, const int* FORWARD_OR_ADJOINT);
/* src/gpu/compute_add_sources_viscoacoustic_cuda.cu:46 */
void compute_add_sources_ac_cuda_(long* Mesh_pointer, int* iphasef, int* itf);
/* src/gpu/compute_add_sources_viscoacoustic_cuda.cu:91 */
void compute_add_sources_ac_s3_cuda_(long* Mesh_pointer, int* iphasef, int* itf);
/* src/gpu/compute_add_sources_viscoacoustic_cuda.cu:137 */
void compute_add_moving_sources_ac_cuda_(long* Mesh_pointer, int* iphase_f, int* nsources_local_moving, int* itf,
                                         int* NSTEP_f, int* nsources_f);
/* src/gpu/compute_add_sources_viscoacoustic_cuda.cu:208 */
void add_sources_ac_sim_2_or_3_cuda_(long* Mesh_pointer, int* iphasef, int* itf, int* nadj_rec_local, int* NSTEP);
/* src/gpu/compute_coupling_cuda.cu:47 */
void compute_coupling_ac_el_cuda_(long* Mesh_pointer, int* iphasef, int* num_coupling_ac_el_facesf,
                                  int* FORWARD_OR_ADJOINT);
/* src/gpu/compute_coupling_cuda.cu:125 */
void compute_coupling_el_ac_cuda_(long* Mesh_pointer, int* iphasef, int* num_coupling_ac_el_facesf,
                                  int* FORWARD_OR_ADJOINT);
/* src/gpu/compute_forces_acoustic_cuda.cu:248 */
void compute_forces_acoustic_cuda_(long* Mesh_pointer, int* iphase, int* nspec_outer_acoustic,
                                   int* nspec_inner_acoustic, int* ATTENUATION_VISCOACOUSTIC, int* compute_wavefield_1,
                                   int* compute_wavefield_2);
/* src/gpu/compute_kernels_cuda.cu:92 */
void compute_kernels_acoustic_cuda_(long* Mesh_pointer, realw* deltat);
/* src/gpu/compute_kernels_cuda.cu:133 */
void compute_kernels_hess_cuda_(long* Mesh_pointer, int* ELASTIC_SIMULATION, int* ACOUSTIC_SIMULATION);
/* src/gpu/compute_stacey_acoustic_cuda.cu:40 */
void compute_stacey_acoustic_cuda_(long* Mesh_pointer, int* iphasef, realw* h_b_absorb_potential_left,
                                   realw* h_b_absorb_potential_right, realw* h_b_absorb_potential_top,
                                   realw* h_b_absorb_potential_bottom, int* compute_wavefield_1,
                                   int* compute_wavefield_2, int* UNDO_ATTENUATION_AND_OR_PML);
/* src/gpu/enforce_acoustic_free_surface_cuda.cu:46 */
void acoustic_enforce_free_surf_cuda_(long* Mesh_pointer, int* compute_wavefield_1, int* compute_wavefield_2);
/* src/gpu/pml_compute_cuda.cu:47 */
void pml_boundary_acoustic_cuda_(long* Mesh_pointer, int* compute_wavefield_1, int* compute_wavefield_2);
/* src/gpu/prepare_mesh_constants_cuda.cu:319 */
void prepare_fields_acoustic_device_(long* Mesh_pointer, realw* rmass_acoustic, realw* rhostore, realw* kappastore,
                                     int* num_phase_ispec_acoustic, int* phase_ispec_inner_acoustic,
                                     int* num_free_surface_faces, int* free_surface_ispec, int* free_surface_ijk,
                                     int* ELASTIC_SIMULATION, int* num_coupling_ac_el_faces, int* coupling_ac_el_ispec,
                                     int* coupling_ac_el_ijk, realw* coupling_ac_el_normal,
                                     realw* coupling_ac_el_jacobian2Dw, int* h_ninterface_acoustic,
                                     int* h_inum_interfaces_acoustic, int* ATTENUATION_VISCOACOUSTIC,
                                     realw* h_A_newmark, realw* h_B_newmark, int* NO_BACKWARD_RECONSTRUCTION,
                                     realw* h_no_backward_acoustic_buffer);
/* src/gpu/prepare_mesh_constants_cuda.cu:442 */
void prepare_fields_acoustic_adj_dev_(long* Mesh_pointer, int* APPROXIMATE_HESS_KL, int* ATTENUATION_VISCOACOUSTIC,
                                      int* NO_BACKWARD_RECONSTRUCTION);
/* src/gpu/prepare_mesh_constants_cuda.cu:843 */
void prepare_pml_device_(long* Mesh_pointer, int* NSPEC_PML, int* NSPEC_PML_X, int* NSPEC_PML_Z, int* NSPEC_PML_XZ,
                         int* h_spec_to_pml, realw* h_abs_normalized, realw* ALPHA_MAX_PML, realw* d0_max,
                         realw* deltat, realw* h_alphax_store, realw* h_alphaz_store, realw* h_betax_store,
                         realw* h_betaz_store, int* PML_nglob_abs_acoustic_f, int* h_PML_abs_points_acoustic);
/* src/gpu/prepare_mesh_constants_cuda.cu:1034 */
void prepare_moving_sources_cuda_(long* Mesh_pointer, int* h_nsources_local_f_moving, int* NSOURCES,
                                  realw* h_sourcearrays_moving, int* h_ispec_selected_source_moving, int* NSTEP,
                                  realw* h_source_time_function_moving);
/* src/gpu/smooth_cuda.cu:41 */
void prepare_arrays_gpu_(long* Container, realw* xstore_me, realw* zstore_me, realw* sigma_h2_inv,
                         realw* sigma_v2_inv, realw* h_criterion, realw* v_criterion, int* nspec_me, int* nker,
                         realw* wgll_sq);
/* src/gpu/smooth_cuda.cu:77 */
void compute_smooth_(long* smooth_pointer, realw* jacobian, realw* xstore_other, realw* zstore_other,
                     realw* data_other, const int* nspec_other);
/* src/gpu/smooth_cuda.cu:128 */
void get_smooth_(long* smooth_pointer, realw* data_smooth);
/* src/gpu/transfer_fields_cuda.cu:223 */
void transfer_fields_ac_to_device_(int* size, realw* potential_acoustic, realw* potential_dot_acoustic,
                                   realw* potential_dot_dot_acoustic, long* Mesh_pointer);
/* src/gpu/transfer_fields_cuda.cu:248 */
void transfer_b_fields_ac_to_device_(int* size, realw* b_potential_acoustic, realw* b_potential_dot_acoustic,
                                     realw* b_potential_dot_dot_acoustic, long* Mesh_pointer);
/* src/gpu/transfer_fields_cuda.cu:274 */
void transfer_fields_ac_from_device_(int* size, realw* potential_acoustic, realw* potential_dot_acoustic,
                                     realw* potential_dot_dot_acoustic, long* Mesh_pointer);
/* src/gpu/transfer_fields_cuda.cu:300 */
void transfer_b_fields_ac_from_device_(int* size, realw* b_potential_acoustic, realw* b_potential_dot_acoustic,
                                       realw* b_potential_dot_dot_acoustic, long* Mesh_pointer);
/* src/gpu/transfer_fields_cuda.cu:324 */
void transfer_b_potential_ac_from_device_(int* size, realw* b_potential_acoustic, long* Mesh_pointer);
/* src/gpu/transfer_fields_cuda.cu:341 */
void transfer_b_potential_ac_to_device_(int* size, realw* b_potential_acoustic, long* Mesh_pointer);
/* src/gpu/transfer_fields_cuda.cu:358 */
void transfer_dot_dot_from_device_(int* size, realw* potential_dot_dot_acoustic, long* Mesh_pointer);
/* src/gpu/transfer_fields_cuda.cu:374 */
void transfer_b_dot_dot_from_device_(int* size, realw* b_potential_dot_dot_acoustic, long* Mesh_pointer);
/* src/gpu/transfer_fields_cuda.cu:391 */
void transfer_kernels_ac_to_host_(long* Mesh_pointer, realw* h_rho_ac_kl, realw* h_kappa_ac_kl, int* NSPEC_AB);
/* src/gpu/transfer_fields_cuda.cu:415 */
void transfer_kernels_hess_el_tohost_(long* Mesh_pointer, realw* h_hess_kl, int* NSPEC_AB);
/* src/gpu/transfer_fields_cuda.cu:430 */
void transfer_kernels_hess_ac_tohost_(long* Mesh_pointer, realw* h_hess_ac_kl, int* NSPEC_AB);
/* src/gpu/transfer_fields_cuda.cu:447 */
void transfer_viscoacoustic_b_var_to_device_(int* size, realw* b_e1_acous_sf, realw* b_sum_forces_old,
                                             long* Mesh_pointer);
/* src/gpu/transfer_fields_cuda.cu:463 */
void transfer_viscoacoustic_var_from_device_(int* size, realw* e1_acous_sf, realw* sum_forces_old,
                                             long* Mesh_pointer);
/* src/gpu/transfer_fields_cuda.cu:480 */
void transfer_viscoelastic_b_var_to_device_(int* size, realw* b_e1, realw* b_e11, realw* b_e13,
                                            realw* b_dux_dxl_old, realw* b_duz_dzl_old,
                                            realw* b_dux_dzl_plus_duz_dxl_old, long* Mesh_pointer);
/* src/gpu/transfer_fields_cuda.cu:510 */
void transfer_viscoelastic_var_from_device_(int* size, realw* e1, realw* e11, realw* e13, realw* dux_dxl_old,
                                            realw* duz_dzl_old, realw* dux_dzl_plus_duz_dxl_old,
                                            long* Mesh_pointer);
/* src/gpu/transfer_fields_cuda.cu:539 */
void transfer_async_pot_ac_from_device_(realw* pot_buffer, long* Mesh_pointer);
/* src/gpu/transfer_fields_cuda.cu:567 */
void transfer_async_pot_ac_to_device_(realw* pot_buffer, long* Mesh_pointer);
/* src/gpu/update_displacement_cuda.cu:133 */
void update_displacement_ac_cuda_(long* Mesh_pointer, realw* deltat_F, realw* deltatsqover2_F,
                                  realw* deltatover2_F, realw* b_deltat_F, realw* b_deltatsqover2_F,
                                  realw* b_deltatover2_F, int* compute_b_wavefield,
                                  int* UNDO_ATTENUATION_AND_OR_PML);
/* src/gpu/update_displacement_cuda.cu:322 */
void kernel_3_acoustic_cuda_(long* Mesh_pointer, realw* deltatover2, realw* b_deltatover2, int* compute_wavefield_1,
                             int* compute_wavefield_2);

#ifdef __cplusplus
}
#endif
#endif /* SPECFEM2D_GPU_H */
