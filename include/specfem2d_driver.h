/*
 * specfem2d_driver.h -- host side above the C-ABI of specfem2d_gpu.h.
 *
 * The driver plays the role of the reference's xspecfem2D on one GPU for
 * structured internal meshes: it builds what the Fortran setup builds
 * (src/specfem2D/setup_mesh.F90, createnum_fast.f90 + get_global.f90,
 * define_derivation_matrices.f90, invert_mass_matrix.F90,
 * prepare_timerun.F90, locate_source.F90, locate_receivers.F90,
 * setup_sources_receivers.F90, prepare_source_time_function.f90), hands it to
 * the GPU library through prepare_constants_device_ & co. exactly as
 * src/specfem2D/prepare_gpu.f90:80-347 does, and then replays the GPU branch
 * of the time loop (src/specfem2D/iterate_time.F90:146-414,
 * compute_gpu_elastic.f90:36-311).  It never computes physics itself.
 */
#ifndef SPECFEM2D_DRIVER_H
#define SPECFEM2D_DRIVER_H

#ifdef __cplusplus
extern "C" {
#endif

#define SF2D_MAXLAYERS 32

/* Par_file subset for the internal mesher and the elastic solver */
typedef struct {
  int nx, nz;                          /* elements along x and z */
  int nlayers;                         /* flat layers, bottom first */
  int nz_layer[SF2D_MAXLAYERS];
  double xmin, xmax;
  double zinterface[SF2D_MAXLAYERS + 1];
  int absorbbottom, absorbright, absorbtop, absorbleft;
  int stacey;                          /* STACEY_ABSORBING_CONDITIONS */
  int nstep;                           /* NSTEP */
  double dt;                           /* DT */
  double user_t0;                      /* USER_T0 */
  double anglerec;                     /* degrees */
  int material_mode;                   /* 0: material table (rho,vp,vs per point), 1: MODEL=gll */
  int seismotype;                      /* 1 displacement, 2 velocity, 3 acceleration (in fluid elements the gradient of
                                          the potential or of its time derivatives over the density), 4 pressure, 6
                                          potential (receivers in fluid elements); shot batches, adjoint runs and
                                          gradients record displacement only */
  int use_trick;                       /* USE_TRICK_FOR_BETTER_PRESSURE */
  int pml;                             /* PML_BOUNDARY_CONDITIONS: CPML layer on the absorbing sides (stacey = 0) */
  int nelem_pml;                       /* NELEM_PML_THICKNESS */
  double k_min_pml, k_max_pml;         /* K_MIN_PML, K_MAX_PML */
  double damping_ac, damping_el;       /* damping_change_factor_acoustic, damping_change_factor_elastic */
} sf2d_params;

/* DATA/SOURCE subset: collocated force in a solid element, pressure source in a fluid element (vs = 0), or a
   moment tensor in a solid element */
typedef struct {
  double x, z, f0, factor, anglesource, tshift;
  int time_function_type;              /* 1 Ricker, 2 first derivative of a Gaussian, 3 Gaussian */
  int source_type;                     /* 1 (or 0) force / pressure, 2 moment tensor */
  double Mxx, Mzz, Mxz;                /* moment tensor (source_type 2) */
} sf2d_source;

/* Builds the host-side setup of a simulation (no device call).  rho/vp/vs are
   per GLL point [nspec][25] (i fastest).  simulation_type 1 (forward) or 3 (adjoint);
   save_forward stores the Stacey traction per step for a later adjoint run.
   Returns NULL on invalid input (message on stderr). */
void* sf2d_sim_create(const sf2d_params* p, const double* rho, const double* vp, const double* vs, int nsrc,
                      const sf2d_source* src, int nrec, const double* rec_xz, int simulation_type,
                      int save_forward);
/* the same for a mesh read from files (read_external_mesh, src/meshfem2D/read_external_mesh_files.F90): nspec 4-node
   elements over npgeo control nodes (knods [nspec][4] 0-based, counter-clockwise; coorg [npgeo][2]), the sides of the
   free-surface file (only those of fluid elements count) and of the absorbing file (Stacey edges when p->stacey) as
   0-based element and side code (1 bottom, 2 right, 3 top, 4 left); p's internal-mesher fields (nx, nz, layers,
   absorb flags) are unused and CPML layers are refused.  The GLL points are numbered by coordinates, then by first
   touch in element order (createnum_fast.f90, get_global.f90). */
void* sf2d_sim_create_external(const sf2d_params* p, int nspec, int npgeo, const int* knods, const double* coorg,
                               int nfree, const int* free_elem, const int* free_code, int nabs, const int* abs_elem,
                               const int* abs_code, const double* rho, const double* vp, const double* vs, int nsrc,
                               const sf2d_source* src, int nrec, const double* rec_xz, int simulation_type,
                               int save_forward);
/* An uneven batch (no counterpart in the reference): the disjoint union of n host-only simulations made by
   sf2d_sim_create / sf2d_sim_create_external (+ sf2d_sim_set_attenuation) -- different meshes, sizes and physics
   (fluid, elastic, viscoelastic, CPML or Stacey), one DT, NSTEP and seismotype -- as one simulation whose kernels
   cover every instance in each launch.  Elements, points, sources and receivers follow in instance order; receiver
   r of instance k is union receiver sum_{j<k} nrec_j + r.  The instances stay valid and are released by the caller;
   returns NULL when they do not fit together. */
void* sf2d_sim_create_union(int n, void* const* sims);
/* A shot batch (no counterpart in the reference): nshot forward simulations of one mesh, model and receiver
   set, shot k driven by the next shot_nsrc[k] sources of src (timed as that shot's own simulation: its t0 and
   time shifts are those of a run with only its sources).  On the device all shots advance together
   (sf2d_batch_* of specfem2d_gpu_ext.h); the gradient calls below then take obs[nshot][nrec][2][NSTEP],
   return the misfit summed over the shots and the gradient of that sum. */
void* sf2d_sim_create_shots(const sf2d_params* p, const double* rho, const double* vp, const double* vs, int nshot,
                            const int* shot_nsrc, const sf2d_source* src, int nrec, const double* rec_xz);
/* all NSTEP steps of every shot from rest; seis[nshot][nrec][2][NSTEP] displacement; timing[0] = device ms */
int sf2d_sim_batch_forward(void* sim, double* seis, double* timing);
int sf2d_sim_nshot(void* sim);                      /* 0 for a single simulation */
void sf2d_sim_shot_t0(void* sim, double* t0);       /* [nshot] */
void sf2d_sim_shot_misfits(void* sim, double* chi); /* [nshot], from the last gradient forward */
void sf2d_sim_batch_get_fields(void* sim, int shot, float* displ, float* veloc, float* accel);
/* ATTENUATION_VISCOELASTIC (before sf2d_sim_prepare_device): quality factors Qkappa, Qmu per GLL point
   ([nspec][25]; both above 9998.999 = no attenuation at that point), n_sls standard linear solids (3, as the
   reference's GPU code requires), reference frequency the first source's f0.  The relaxation times come from the
   least-squares Zener fit of attenuation_model.f90 (USE_SOLVOPT false) and the coefficients from
   prepare_attenuation.f90; forward runs of solid meshes only (as prepare_gpu.f90:69-72).  Returns 0 on success. */
int sf2d_sim_set_attenuation(void* sim, const double* qkappa, const double* qmu, int n_sls);
/* the attenuation coefficients handed to prepare_fields_elastic_device_, [nspec][25][3] each (for parity checks) */
void sf2d_sim_get_attenuation(void* sim, float* A_kappa, float* B_kappa, float* A_mu, float* B_mu);
/* Hands the setup to the GPU library in the order of prepare_gpu.f90:80-347. Returns 0 on success. */
int sf2d_sim_prepare_device(void* sim);
void sf2d_sim_destroy(void* sim);

int sf2d_sim_nspec(void* sim);
int sf2d_sim_nglob(void* sim);
int sf2d_sim_nabs(void* sim);
double sf2d_sim_t0(void* sim);
/* CPML layer of the setup (pml_init.F90): counts[3] = {layer elements, outer-boundary points of elastic layer
   elements, of acoustic ones}; region[nspec] (0, 1 X only, 2 Z only, 3 XZ), prof[6][NSPEC_PML][25] = K_x, K_z, d_x, d_z,
   alpha_x, alpha_z (layer elements in element order), the outer-boundary points (1-based, increasing) */
void sf2d_sim_pml_counts(void* sim, int* counts);
void sf2d_sim_get_pml(void* sim, int* region, double* prof, int* pts_el, int* pts_ac);
/* the opaque Fortran-style handle (value of Mesh_pointer) */
long sf2d_sim_mesh_pointer(void* sim);

/* setup products, for parity checks */
void sf2d_sim_get_ibool(void* sim, int* out);                   /* [nspec][25], 1-based */
void sf2d_sim_get_rmass_inverse(void* sim, float* out);         /* [nglob][2] */
void sf2d_sim_get_elem_array(void* sim, int which, float* out); /* 0 xix 1 xiz 2 gammax 3 gammaz 4 jacobian
                                                                   5 kappa 6 mu 7 rho_vp 8 rho_vs 9 rho */
void sf2d_sim_get_sources(void* sim, int* ispec, double* xi, double* gamma, float* sourcearrays, float* stf);
void sf2d_sim_get_receivers(void* sim, int* ispec, double* xi, double* gamma, float* xir, float* gammar);
void sf2d_sim_get_stacey(void* sim, int* ispec, int* ij, float* normal, float* jacw, int* edge, int* ib);

/* Forward time loop (SIMULATION_TYPE=1): runs the next nstep_run steps (continuing after the steps
   already run, up to NSTEP), displacement seismograms of those steps into seis[nrec][2][nstep_run].
   With save_forward the Stacey term per step is copied to absorb[nstep_run][nabs][5][2] (may be NULL;
   the traction of a solid face, the potential term of a fluid face in component 0).  If timing is not NULL, HIP events on the library stream give {loop ms, predictor ms,
   forces ms, boundary+sources ms, corrector ms, seismogram ms}: the loop and the element-force
   launches always, the other phases only when timing[0] < 0 on entry.  Returns 0 on success. */
int sf2d_sim_forward(void* sim, int nstep_run, double* seis, float* absorb, double* timing);
/* Adjoint run (SIMULATION_TYPE=3, simulation created with simulation_type 3): adj[nrec][2][NSTEP]
   adjoint source traces (x then z), absorb[NSTEP][nabs][5][2] and the last frame from a save_forward
   run of the same model.  Replays the GPU branch of iterate_time.F90 for SIMULATION_TYPE=3 with both
   wavefields; records the adjoint displacement at the receivers into seis[nrec][2][NSTEP] and the
   kernels K_rho, K_mu, K_kappa (scaled as save_adjoint_kernels.f90 writes them) into
   kernels[3][nspec][25] in ispec order.  Returns 0 on success. */
int sf2d_sim_adjoint(void* sim, const float* adj, const float* absorb, const float* last_displ,
                     const float* last_veloc, const float* last_accel, double* seis, float* kernels);
/* Adjoint run over fluid, solid and coupled meshes (no CPML, no attenuation): as sf2d_sim_adjoint, plus the
   last fluid frame (potential and its first and second time derivatives, [nglob] each; NULL without fluid
   elements; the solid frame is NULL without solid elements).  A receiver in a fluid element takes component x of
   adj as its source (source_adjoint(:,:,1)) and records the adjoint field's seismotype.  kernels[3][nspec][25]:
   solid K_rho, K_mu, K_kappa as sf2d_sim_adjoint (zero in fluid elements); kernels_ac[4][nspec][25] (may be NULL):
   the fluid's rho_ac_kl, kappa_ac_kl, rho' = rho + kappa and c = 2 kappa (compute_kernels.f90:426-439, zero in
   solid elements).  Coupled meshes run the adjoint solid, the adjoint fluid, then the reconstructed fluid and
   solid (iterate_time.F90:227-358).  Returns 0 on success. */
int sf2d_sim_adjoint_full(void* sim, const float* adj, const float* absorb, const float* last_displ,
                          const float* last_veloc, const float* last_accel, const float* last_pot,
                          const float* last_pot_dot, const float* last_pot_dd, double* seis, float* kernels,
                          float* kernels_ac);
/* Anisotropic elements (material table icodemat 2, read_materials.f90:160-200), before the device is prepared:
   is_aniso[nspec] flags, c[6][nspec][25] = c11, c13, c15, c33, c35, c55 per GLL point (the element's rho, vp =
   sqrt(c33/rho), vs = sqrt(c55/rho) of the creation call give its mass and Stacey impedances).  Forward runs without
   attenuation, CPML or shot batches.  Returns 0 on success. */
int sf2d_sim_set_anisotropy(void* sim, const int* is_aniso, const double* c);
/* Elements per chain patch (1..12, 0 = default) of the device plan, before the device is prepared (per handle;
   the autotuner's knob).  Returns 0 on success. */
int sf2d_sim_set_chain_elems(void* sim, int n);
/* replaces the model (MODEL=gll parameters rho, vp, vs per GLL point, [nspec][25]) of a simulation: moduli,
   Stacey impedances and mass matrix are recomputed and uploaded.  Returns 0 on success. */
int sf2d_sim_set_model(void* sim, const double* rho, const double* vp, const double* vs);
/* Exact discrete gradient of chi = 1/2 dt sum_n sum_r |s_n - o_n|^2 (obs[nrec][2][NSTEP], displacement)
   over a full run from rest.  The forward call runs all NSTEP steps keeping a checkpoint every `chunk` steps in
   device memory and returns chi (and the seismograms into seis if not NULL); the backward call recomputes
   each segment from its checkpoint and runs the reverse sweep, writing d chi / d (rho, vp, vs) into
   grad[3][nspec][25] and, if parts is not NULL, d chi / d (kappa, mu, rho through the mass, rho vp, rho vs)
   into parts[5][nspec][25].  Return 0 on success. */
int sf2d_sim_gradient_forward(void* sim, const double* obs, int chunk, double* chi, double* seis);
int sf2d_sim_gradient_backward(void* sim, double* grad, double* parts);
/* zero the wavefields and restart the step counter */
void sf2d_sim_reset(void* sim);
int sf2d_sim_steps_done(void* sim);

/* device fields (interleaved x,z; 2*nglob floats); NULL pointers are skipped */
/* fluid elements (vs = 0): their count, and the potential, its first and second time derivatives [nglob] */
int sf2d_sim_nspec_acoustic(void* sim);
int sf2d_sim_ncoupling(void* sim);                       /* fluid-solid edges */
void sf2d_sim_get_rmass_acoustic(void* sim, float* out); /* [nglob] inverse fluid mass (1 off the fluid) */
/* fluid-solid edges: fluid element (1-based) and side (1 bottom, 2 right, 3 top, 4 left), solid element and side */
void sf2d_sim_get_coupling(void* sim, int* ac_ispec, int* ac_code, int* el_ispec, int* el_code);
void sf2d_sim_get_acoustic_fields(void* sim, float* pot, float* pot_dot, float* pot_dot_dot);
void sf2d_sim_set_acoustic_fields(void* sim, const float* pot, const float* pot_dot, const float* pot_dot_dot);
void sf2d_sim_get_fields(void* sim, float* displ, float* veloc, float* accel);
void sf2d_sim_set_fields(void* sim, const float* displ, const float* veloc, const float* accel);
void sf2d_sim_get_b_fields(void* sim, float* displ, float* veloc, float* accel);
void sf2d_sim_set_b_fields(void* sim, const float* displ, const float* veloc, const float* accel);

#ifdef __cplusplus
}
#endif
#endif
