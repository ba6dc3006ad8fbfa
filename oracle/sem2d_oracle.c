/*
 * CPU oracle for the SPECFEM2D elastic time-stepping path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is a plain-C restatement of the
 * reference Fortran solver (SPECFEM2D 8.1.0, CPU path) for an elastic,
 * isotropic, P-SV, Newmark simulation on a structured internal mesh with
 * Stacey absorbing edges, extended to acoustic (fluid) elements: the
 * potential formulation, the acoustic free surface, acoustic Stacey edges,
 * pressure sources, pressure seismograms and fluid-solid coupling, and to CPML absorbing layers (elastic and
 * acoustic, Newmark, convolutional memory variables).  It is the checker for the HIP implementation and
 * the CPU baseline timed beside it; the product never links or calls it.
 *
 * It is compiled twice: -DREAL=float (CUSTOM_REAL=real, the reference
 * default) and -DREAL=double (the reference --enable-double-precision
 * build).  Operation order follows the Fortran so that the float build
 * reproduces the reference's rounding; compile with -ffp-contract=off.
 *
 * Every routine cites the reference file it restates.  Parity status of
 * each part (what pins it) is listed in the test suite's header
 * (tests/test_oracle_golden.py, tests/test_oracle_adjoint.py).
 */
#ifdef _OPENMP
#include <omp.h>
#endif
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <xmmintrin.h>
#include <pmmintrin.h>

#ifndef REAL
#define REAL float
#endif

#define NGLL 5
#define NGLL2 25
#define MAXLAYERS 32

/* ------------------------------------------------------------------ */
/* public parameter blocks (mirrored by oracle/oracle.py via ctypes)    */
/* ------------------------------------------------------------------ */
typedef struct {
  int nx, nz;                 /* elements along x and z (nxread, nzread) */
  int nlayers;                /* flat layers, bottom to top */
  int nz_layer[MAXLAYERS];    /* elements per layer */
  double xmin, xmax;
  double zinterface[MAXLAYERS + 1]; /* interface heights, bottom first */
  int absorb_bottom, absorb_right, absorb_top, absorb_left;
  int stacey;                 /* STACEY_ABSORBING_CONDITIONS */
  int nstep;
  double dt;
  double user_t0;             /* USER_T0 (0 = automatic) */
  double anglerec_deg;        /* anglerec */
  int material_mode;          /* 0: Par_file material table, 1: MODEL=gll */
  int seismotype;             /* 1 displacement, 2 velocity, 3 acceleration, 4 pressure, 6 potential */
  int use_trick;              /* USE_TRICK_FOR_BETTER_PRESSURE */
  int pml;                    /* PML_BOUNDARY_CONDITIONS (CPML on the absorbing sides) */
  int nelem_pml;              /* NELEM_PML_THICKNESS */
  double k_min_pml, k_max_pml;/* K_MIN_PML, K_MAX_PML */
  double damping_ac, damping_el; /* damping_change_factor_acoustic, damping_change_factor_elastic */
} orc_params;

typedef struct {
  double x, z;
  double f0;
  double factor;
  double angle_deg;           /* anglesource, degrees */
  double tshift;
  int time_function_type;     /* 1 Ricker, 2 first derivative of Gaussian, 3 Gaussian */
  int source_type;            /* 1 (or 0) collocated force / pressure, 2 moment tensor */
  double Mxx, Mzz, Mxz;       /* moment tensor (source_type 2) */
} orc_source;

typedef struct {
  orc_params p;
  int nspec, nglob, npgeo;
  int *knods;                 /* [nspec][4] 0-based control nodes */
  /* external meshes (read_external_mesh): the sides of the files' free surface and absorbing boundary as
     (0-based element, side code 1 bottom 2 right 3 top 4 left); internal meshes derive them from nx, nz */
  int external;
  int ext_nfree, ext_nabs;
  int *ext_free_elem, *ext_free_code, *ext_abs_elem, *ext_abs_code;
  double *coorg;              /* [npgeo][2] */
  int *ibool;                 /* [nspec][25] 1-based (i fastest, then j) */
  double *coord;              /* [nglob][2] */
  double xigll[NGLL];
  REAL wxgll[NGLL];
  REAL hprime[NGLL][NGLL];    /* hprime[i][k] = h'_k(xi_i)  (Fortran hprime_xx(i,k)) */
  REAL hprimewgll[NGLL][NGLL];/* hprimewgll[k][i] = w_k h'_i(xi_k) (Fortran hprimewgll_xx(k,i)) */
  REAL *xix, *xiz, *gammax, *gammaz, *jacobian; /* [nspec][25] */
  REAL *rhostore, *mustore, *kappastore, *rho_vp, *rho_vs;
  REAL *rmass_inv;            /* [nglob][2] */
  /* absorbing edges */
  int nabs;
  int *abs_ispec;             /* 1-based */
  int *abs_code;              /* 1 bottom, 2 right, 3 top, 4 left */
  int *abs_corner;            /* [nabs][4] */
  int *ib;                    /* index within its own side, 1-based */
  int nspec_side[5];
  /* sources */
  int nsrc;
  orc_source *src;
  int *src_ispec;
  double *src_xi, *src_gamma;
  REAL *sourcearrays;         /* [nsrc][25][2]  (c fastest) */
  REAL *stf;                  /* [nsrc][nstep] */
  double t0;
  /* receivers */
  int nrec;
  double *rec_xz;
  int *rec_ispec;
  double *rec_xi, *rec_gamma;
  REAL *xir, *gammar;         /* [nrec][5] */
  double cosrot, sinrot;
  /* domains: an element is acoustic when vs = 0 at all its GLL points */
  int *is_acoustic;           /* [nspec] */
  int *is_aniso;              /* [nspec] anisotropic elements (NULL: none) */
  REAL *aniso_c;              /* [nspec*25][6] c11 c13 c15 c33 c35 c55 */
  int nspec_acoustic, any_acoustic, any_elastic;
  REAL *rmass_ac_inv;         /* [nglob] acoustic inverse mass */
  int nfree;                  /* acoustic free-surface edges: [nfree] element, side code */
  int *free_ispec, *free_code;
  int ncoup;                  /* fluid-solid edges: acoustic element/side, elastic element/side */
  int *coup_ac_ispec, *coup_ac_code, *coup_el_ispec, *coup_el_code;
  int *abs_jbegin, *abs_jend; /* acoustic Stacey point range along each absorbing edge, 0-based inclusive */
  /* wavefields */
  REAL *displ, *veloc, *accel;/* [nglob][2] */
  REAL *pot, *pot_dot, *pot_dd;/* [nglob] acoustic potential and its time derivatives */
  int ftz;                    /* run the time loop with flush-to-zero */
  /* viscoelastic attenuation (ATTENUATION_VISCOELASTIC, Newmark): nsls standard linear solids per point, the
     convolution coefficients of prepare_attenuation.f90:368-397 for the bulk (1) and shear (2) modes, memory
     variables and the previous step's strains (compute_attenuation_viscoelastic.f90) */
  int att, att_live, nsls;
  char *att_on;               /* [nspec*25] Qkappa or Qmu below 9999 */
  REAL *att_A1, *att_B1, *att_A2, *att_B2; /* [nspec*25][nsls] */
  REAL *e1, *e11, *e13;       /* [nspec*25][nsls] */
  REAL *old_xx, *old_zz, *old_xz; /* [nspec*25] */
  /* the Zener model's per-point parameters as attenuation_model.f90:145-161 stores them (single precision):
     1/tau_sigma, tau_epsilon, phi [nspec*25][nsls] and Mu [nspec*25] for the bulk (1) and shear (2) modes; the
     viscoelastic CPML uses them (pml_compute_memory_variables.f90:546-894) in elements where both quality factors
     are finite at every point (visco_pml), with memory variables of the bulk and shear convolutions
     [nspec_pml*25][nsls]: ka_xx, ka_zz (dux_dx, duz_dz) and mu_xx, mu_zz, mu_xz, mu_zx (dux_dx, duz_dz, dux_dz,
     duz_dx) */
  REAL *att_itau1, *att_itau2, *att_te1, *att_te2, *att_phi1, *att_phi2, *att_mu1, *att_mu2;
  char *visco_pml;            /* [nspec] */
  REAL *vp_ka_xx, *vp_ka_zz, *vp_mu_xx, *vp_mu_zz, *vp_mu_xz, *vp_mu_zx;
  int it_done;                /* forward steps already run */
  /* CPML (pml_init.F90, prepare_pml.f90): PML elements, their region (1 X only, 2 Z only, 3 XZ), per-point
     damping profiles in double precision (K, d, alpha along x and z), the outer-boundary points that are held
     at zero, the memory variables of the recursive convolutions and the previous-step fields */
  int pml_on, nspec_pml;
  int *is_pml, *region, *spec_to_pml; /* [nspec]; spec_to_pml 1-based, 0 outside */
  double *pml_kx, *pml_kz, *pml_dx, *pml_dz, *pml_ax, *pml_az; /* [nspec_pml][25] */
  int npml_el, npml_ac;       /* outer-boundary points of elastic / acoustic PML elements */
  int *pml_pts_el, *pml_pts_ac; /* 1-based iglob, increasing */
  REAL *rm_dux_dx, *rm_duz_dx, *rm_dux_dz, *rm_duz_dz; /* [nspec_pml*25][2] */
  REAL *rm_displ;             /* [nspec_pml*25][2 series][2 components] */
  REAL *rm_ac_dx, *rm_ac_dz, *rm_pot; /* [nspec_pml*25][2] */
  REAL *displ_old, *pot_old;  /* [nglob][2], [nglob] */
  double pml_alpha_max;       /* ALPHA_MAX_PML = pi f0max */
  /* OpenMP build only (the CPU baseline): elements ordered by colour of a greedy colouring in which no two
     elements of one colour share a point, so each colour's elements assemble in parallel without atomics */
  int ncolour;
  int *colour_start;          /* [ncolour + 1] */
  int *colour_elems;          /* [nspec] */
} orc_sim;

/* ------------------------------------------------------------------ */
/* GLL basis  (src/shared/gll_library.f90, lagrange_poly.f90,          */
/*             src/specfem2D/define_derivation_matrices.f90:34-82)     */
/* ------------------------------------------------------------------ */
static double pnleg(double z, int n) {
  /* Legendre polynomial by the three-term recurrence (lagrange_poly.f90 pnleg) */
  double p1 = 1.0, p2 = z, p3 = p2;
  if (n == 0) return 1.0;
  for (int k = 1; k < n; k++) {
    double fk = (double)k;
    p3 = ((2.0 * fk + 1.0) * z * p2 - fk * p1) / (fk + 1.0);
    p1 = p2;
    p2 = p3;
  }
  return p3;
}

static double pndleg(double z, int n) {
  /* derivative of the Legendre polynomial (lagrange_poly.f90 pndleg) */
  double p1 = 1.0, p2 = z, p1d = 0.0, p2d = 1.0, p3d = 1.0;
  if (n == 0) return 0.0;
  for (int k = 1; k < n; k++) {
    double fk = (double)k;
    double p3 = ((2.0 * fk + 1.0) * z * p2 - fk * p1) / (fk + 1.0);
    p3d = ((2.0 * fk + 1.0) * p2 + (2.0 * fk + 1.0) * z * p2d - fk * p1d) / (fk + 1.0);
    p1 = p2; p2 = p3; p1d = p2d; p2d = p3d;
  }
  return p3d;
}

static double lagrange_deriv_gll(int i, int j, const double *z, int nz) {
  /* h'_i(z_j)  (lagrange_poly.f90:102-136) */
  int deg = nz - 1;
  if (i == 0 && j == 0) return -(double)deg * ((double)deg + 1.0) * 0.25;
  if (i == deg && j == deg) return (double)deg * ((double)deg + 1.0) * 0.25;
  if (i == j) return 0.0;
  return pnleg(z[j], deg) / (pnleg(z[i], deg) * (z[j] - z[i])) +
         (1.0 - z[j] * z[j]) * pndleg(z[j], deg) /
             ((double)deg * ((double)deg + 1.0) * pnleg(z[i], deg) * (z[j] - z[i]) * (z[j] - z[i]));
}

static void setup_gll(orc_sim *s) {
  /* nodes and weights of 5-point Gauss-Lobatto-Legendre quadrature: roots of
     (1-x^2) P'_4(x); weights 2/(N(N+1) P_N(x)^2) with N=4 */
  const double r = sqrt(3.0 / 7.0);
  double x[NGLL] = {-1.0, -r, 0.0, r, 1.0};
  for (int i = 0; i < NGLL; i++) {
    double pn = pnleg(x[i], NGLL - 1);
    double w = 2.0 / ((double)(NGLL - 1) * (double)NGLL * pn * pn);
    s->xigll[i] = x[i];
    s->wxgll[i] = (REAL)w;
  }
  s->xigll[2] = 0.0; /* define_derivation_matrices.f90:57 */
  for (int i1 = 0; i1 < NGLL; i1++)
    for (int i2 = 0; i2 < NGLL; i2++) {
      s->hprime[i2][i1] = (REAL)lagrange_deriv_gll(i1, i2, s->xigll, NGLL);
      s->hprimewgll[i2][i1] = s->wxgll[i2] * s->hprime[i2][i1];
    }
}

/* Lagrange interpolants and their derivatives at xi (lagrange_poly.f90:34-93) */
static void lagrange_any_deriv(double xi, const double *xigll, double *h, double *hp) {
  for (int d = 0; d < NGLL; d++) {
    double prod1 = 1.0, prod2 = 1.0, x0 = xigll[d];
    for (int i = 0; i < NGLL; i++)
      if (i != d) {
        prod1 = prod1 * (xi - xigll[i]);
        prod2 = prod2 * (x0 - xigll[i]);
      }
    double prod2_inv = 1.0 / prod2;
    h[d] = prod1 * prod2_inv;
    double sum = 0.0;
    for (int i = 0; i < NGLL; i++) {
      if (i == d) continue;
      double prod3 = 1.0;
      for (int j = 0; j < NGLL; j++)
        if (j != d && j != i) prod3 = prod3 * (xi - xigll[j]);
      sum = sum + prod3;
    }
    hp[d] = sum * prod2_inv;
  }
}

static void lagrange_any(double xi, const double *xigll, double *h) {
  /* Lagrange interpolants at xi (lagrange_poly.f90:34-93) */
  for (int d = 0; d < NGLL; d++) {
    double prod1 = 1.0, prod2 = 1.0, x0 = xigll[d];
    for (int i = 0; i < NGLL; i++)
      if (i != d) {
        prod1 *= (xi - xigll[i]);
        prod2 *= (x0 - xigll[i]);
      }
    h[d] = prod1 * (1.0 / prod2);
  }
}

/* ------------------------------------------------------------------ */
/* mesher: structured grid with flat interfaces                         */
/* (src/meshfem2D/meshfem2D.F90:133-141, read_mesh_files.F90:169-338)  */
/* ------------------------------------------------------------------ */
static void mesh_internal(orc_sim *s) {
  const orc_params *p = &s->p;
  int nx = p->nx, nz = p->nz;
  s->nspec = nx * nz;
  s->npgeo = (nx + 1) * (nz + 1);
  s->coorg = (double *)calloc((size_t)2 * s->npgeo, sizeof(double));
  s->knods = (int *)malloc(sizeof(int) * 4 * (size_t)s->nspec);
  int ioff = 0;
  for (int il = 0; il < p->nlayers; il++) {
    double bot = p->zinterface[il], top = p->zinterface[il + 1];
    for (int ix = 0; ix <= nx; ix++) {
      double absx = p->xmin + (p->xmax - p->xmin) * (double)ix / (double)nx;
      for (int iz = 0; iz <= p->nz_layer[il]; iz++) {
        double g = (double)iz / (double)p->nz_layer[il];
        int node = (iz + ioff) * (nx + 1) + ix;
        s->coorg[2 * node] = absx;
        s->coorg[2 * node + 1] = (1.0 - g) * bot + g * top;
      }
    }
    ioff += p->nz_layer[il];
  }
  int e = 0;
  for (int j = 1; j <= nz; j++)
    for (int i = 1; i <= nx; i++) {
      s->knods[4 * e + 0] = (j - 1) * (nx + 1) + (i - 1);
      s->knods[4 * e + 1] = (j - 1) * (nx + 1) + (i - 1) + 1;
      s->knods[4 * e + 2] = j * (nx + 1) + (i - 1) + 1;
      s->knods[4 * e + 3] = j * (nx + 1) + (i - 1);
      e++;
    }
}

static void shape4(double xi, double ga, double *sh, double *dsh /* [2][4] */) {
  /* bilinear shape functions (src/shared/define_shape_functions.f90, NGNOD=4) */
  double sp = xi + 1.0, sm = xi - 1.0, tp = ga + 1.0, tm = ga - 1.0;
  sh[0] = 0.25 * sm * tm;
  sh[1] = -0.25 * sp * tm;
  sh[2] = 0.25 * sp * tp;
  sh[3] = -0.25 * sm * tp;
  dsh[0] = 0.25 * tm;
  dsh[1] = -0.25 * tm;
  dsh[2] = 0.25 * tp;
  dsh[3] = -0.25 * tp;
  dsh[4] = 0.25 * sm;
  dsh[5] = -0.25 * sp;
  dsh[6] = 0.25 * sp;
  dsh[7] = -0.25 * sm;
}

static void recompute_jacobian(const orc_sim *s, int ispec, double xi, double ga, double *x, double *z,
                               double *xix, double *xiz, double *gax, double *gaz, double *jac) {
  /* src/specfem2D/recompute_jacobian.f90:37-161 */
  double sh[4], dsh[8];
  shape4(xi, ga, sh, dsh);
  double xx = 0, zz = 0, xxi = 0, zxi = 0, xga = 0, zga = 0;
  for (int a = 0; a < 4; a++) {
    int n = s->knods[4 * ispec + a];
    double xe = s->coorg[2 * n], ze = s->coorg[2 * n + 1];
    xx = xx + sh[a] * xe;
    zz = zz + sh[a] * ze;
    xxi = xxi + dsh[a] * xe;
    zxi = zxi + dsh[a] * ze;
    xga = xga + dsh[4 + a] * xe;
    zga = zga + dsh[4 + a] * ze;
  }
  double j = xxi * zga - xga * zxi;
  *x = xx;
  *z = zz;
  *jac = j;
  *xix = zga / j;
  *gax = -zxi / j;
  *xiz = -xga / j;
  *gaz = xxi / j;
}

/* ------------------------------------------------------------------ */
/* global numbering: coordinate classes, then first-touch renumbering  */
/* (src/specfem2D/createnum_fast.f90, get_global.f90:34-72)            */
/* ------------------------------------------------------------------ */
static const double *g_sortkey;
static int cmp_by_x(const void *a, const void *b) {
  double xa = g_sortkey[2 * *(const int *)a], xb = g_sortkey[2 * *(const int *)b];
  return (xa > xb) - (xa < xb);
}
static int cmp_by_z(const void *a, const void *b) {
  double za = g_sortkey[2 * *(const int *)a + 1], zb = g_sortkey[2 * *(const int *)b + 1];
  return (za > zb) - (za < zb);
}

static void setup_numbering(orc_sim *s) {
  int ntot = s->nspec * NGLL2;
  double *xp = (double *)malloc(sizeof(double) * 2 * (size_t)ntot);
  for (int e = 0; e < s->nspec; e++)
    for (int j = 0; j < NGLL; j++)
      for (int i = 0; i < NGLL; i++) {
        double x, z, a, b, c, d, jac;
        recompute_jacobian(s, e, s->xigll[i], s->xigll[j], &x, &z, &a, &b, &c, &d, &jac);
        int l = e * NGLL2 + j * NGLL + i;
        xp[2 * l] = x;
        xp[2 * l + 1] = z;
      }
  /* tolerance relative to the mesh extent (createnum_fast: SMALLVALTOL * typical distance) */
  double xmin = 1e300, xmax = -1e300, zmin = 1e300, zmax = -1e300;
  for (int l = 0; l < ntot; l++) {
    if (xp[2 * l] < xmin) xmin = xp[2 * l];
    if (xp[2 * l] > xmax) xmax = xp[2 * l];
    if (xp[2 * l + 1] < zmin) zmin = xp[2 * l + 1];
    if (xp[2 * l + 1] > zmax) zmax = xp[2 * l + 1];
  }
  double ext = fmax(xmax - xmin, zmax - zmin);
  double tol = 1e-10 * ext;
  int *idx = (int *)malloc(sizeof(int) * (size_t)ntot);
  int *cls = (int *)malloc(sizeof(int) * (size_t)ntot);
  for (int l = 0; l < ntot; l++) idx[l] = l;
  g_sortkey = xp;
  qsort(idx, (size_t)ntot, sizeof(int), cmp_by_x);
  int ncls = 0;
  int a = 0;
  while (a < ntot) {
    int b = a + 1;
    while (b < ntot && xp[2 * idx[b]] - xp[2 * idx[b - 1]] <= tol) b++;
    qsort(idx + a, (size_t)(b - a), sizeof(int), cmp_by_z);
    int c = a;
    while (c < b) {
      int d = c + 1;
      while (d < b && xp[2 * idx[d] + 1] - xp[2 * idx[d - 1] + 1] <= tol) d++;
      for (int k = c; k < d; k++) cls[idx[k]] = ncls;
      ncls++;
      c = d;
    }
    a = b;
  }
  /* first-touch renumbering, element-major, i fastest */
  int *mask = (int *)malloc(sizeof(int) * (size_t)ncls);
  for (int k = 0; k < ncls; k++) mask[k] = -1;
  s->ibool = (int *)malloc(sizeof(int) * (size_t)ntot);
  int inum = 0;
  for (int l = 0; l < ntot; l++) {
    int c = cls[l];
    if (mask[c] == -1) {
      inum++;
      mask[c] = inum;
    }
    s->ibool[l] = mask[c];
  }
  s->nglob = inum;
  s->coord = (double *)calloc((size_t)2 * s->nglob, sizeof(double));
  for (int l = 0; l < ntot; l++) {
    int g = s->ibool[l] - 1;
    s->coord[2 * g] = xp[2 * l];
    s->coord[2 * g + 1] = xp[2 * l + 1];
  }
  free(mask);
  free(cls);
  free(idx);
  free(xp);
}

/* ------------------------------------------------------------------ */
/* geometry and materials (setup_mesh.F90:195-236, 680-792;            */
/* read_materials.f90:138-141)                                         */
/* ------------------------------------------------------------------ */
static void setup_geometry_materials(orc_sim *s, const double *rho, const double *vp, const double *vs) {
  size_t n = (size_t)s->nspec * NGLL2;
  s->xix = (REAL *)malloc(sizeof(REAL) * n);
  s->xiz = (REAL *)malloc(sizeof(REAL) * n);
  s->gammax = (REAL *)malloc(sizeof(REAL) * n);
  s->gammaz = (REAL *)malloc(sizeof(REAL) * n);
  s->jacobian = (REAL *)malloc(sizeof(REAL) * n);
  s->rhostore = (REAL *)malloc(sizeof(REAL) * n);
  s->mustore = (REAL *)malloc(sizeof(REAL) * n);
  s->kappastore = (REAL *)malloc(sizeof(REAL) * n);
  s->rho_vp = (REAL *)malloc(sizeof(REAL) * n);
  s->rho_vs = (REAL *)malloc(sizeof(REAL) * n);
  for (int e = 0; e < s->nspec; e++)
    for (int j = 0; j < NGLL; j++)
      for (int i = 0; i < NGLL; i++) {
        size_t l = (size_t)e * NGLL2 + j * NGLL + i;
        double x, z, a, b, c, d, jac;
        recompute_jacobian(s, e, s->xigll[i], s->xigll[j], &x, &z, &a, &b, &c, &d, &jac);
        s->xix[l] = (REAL)a;
        s->xiz[l] = (REAL)b;
        s->gammax[l] = (REAL)c;
        s->gammaz[l] = (REAL)d;
        s->jacobian[l] = (REAL)jac;
        double rhol = rho[l], vpl = vp[l], vsl = vs[l], mul, kappal;
        if (s->p.material_mode == 0) {
          double lp2m = rhol * vpl * vpl;
          mul = rhol * vsl * vsl;
          double lambdal = lp2m - 2.0 * mul;
          kappal = lambdal + mul;
          vpl = sqrt((kappal + mul) / rhol);
        } else {
          mul = rhol * vsl * vsl;
          kappal = rhol * vpl * vpl - mul;
        }
        s->rhostore[l] = (REAL)rhol;
        s->mustore[l] = (REAL)mul;
        s->kappastore[l] = (REAL)kappal;
        double vsr = sqrt(mul / rhol);
        s->rho_vp[l] = (REAL)(rhol * vpl);
        s->rho_vs[l] = (REAL)(rhol * vsr);
      }
}

/* ------------------------------------------------------------------ */
/* absorbing edges (src/meshfem2D/determine_abs_surface.f90,           */
/* src/specfem2D/read_mesh_databases.F90:957-1030)                     */
/* ------------------------------------------------------------------ */
static void setup_absorbing(orc_sim *s) {
  const orc_params *p = &s->p;
  int n = 0;
  if (s->external) {
    n = s->ext_nabs;
  } else {
    if (p->absorb_bottom) n += p->nx;
    if (p->absorb_top) n += p->nx;
    if (p->absorb_left) n += p->nz;
    if (p->absorb_right) n += p->nz;
  }
  s->nabs = n;
  s->abs_ispec = (int *)calloc((size_t)(n + 1), sizeof(int));
  s->abs_code = (int *)calloc((size_t)(n + 1), sizeof(int));
  s->abs_corner = (int *)calloc((size_t)4 * (n + 1), sizeof(int));
  s->ib = (int *)calloc((size_t)(n + 1), sizeof(int));
  n = 0;
  /* an external mesh lists its absorbing sides in the file's order */
  for (int a = 0; s->external && a < s->ext_nabs; a++) {
    s->abs_ispec[a] = s->ext_abs_elem[a] + 1;
    s->abs_code[a] = s->ext_abs_code[a];
  }
  for (int iz = 1; iz <= p->nz && !s->external; iz++)
    for (int ix = 1; ix <= p->nx; ix++) {
      int el = (iz - 1) * p->nx + ix;
      if (p->absorb_bottom && iz == 1) { s->abs_ispec[n] = el; s->abs_code[n] = 1; n++; }
      if (p->absorb_right && ix == p->nx) { s->abs_ispec[n] = el; s->abs_code[n] = 2; n++; }
      if (p->absorb_top && iz == p->nz) { s->abs_ispec[n] = el; s->abs_code[n] = 3; n++; }
      if (p->absorb_left && ix == 1) { s->abs_ispec[n] = el; s->abs_code[n] = 4; n++; }
    }
  for (int a = 0; a < s->nabs; a++) {
    if (s->abs_code[a] != 1 && s->abs_code[a] != 3) continue;
    for (int b = 0; b < s->nabs; b++) {
      if (a == b || s->abs_ispec[a] != s->abs_ispec[b]) continue;
      int off = (s->abs_code[a] == 1) ? 0 : 2;
      if (s->abs_code[b] == 4) s->abs_corner[4 * a + off + 0] = 1;
      if (s->abs_code[b] == 2) s->abs_corner[4 * a + off + 1] = 1;
    }
  }
  memset(s->nspec_side, 0, sizeof(s->nspec_side));
  for (int a = 0; a < s->nabs; a++) s->ib[a] = ++s->nspec_side[s->abs_code[a]];
}

/* ------------------------------------------------------------------ */
/* domains, acoustic free surface, fluid-solid edges                    */
/* (src/meshfem2D/determine_acoustic_surface.f90:34-160,              */
/*  repartition_coupling.f90:38-106, src/specfem2D/get_coupling_edges.f90) */
/* ------------------------------------------------------------------ */
static const long *g_side_key;
static int cmp_side_key(const void *a, const void *b) {
  const long ka = g_side_key[*(const int *)a], kb = g_side_key[*(const int *)b];
  if (ka != kb) return (ka > kb) - (ka < kb);
  return (*(const int *)a > *(const int *)b) - (*(const int *)a < *(const int *)b);
}

static void setup_domains(orc_sim *s) {
  const orc_params *p = &s->p;
  const int nx = p->nx, nz = p->nz;
  s->is_acoustic = (int *)calloc((size_t)s->nspec + 1, sizeof(int));
  s->nspec_acoustic = 0;
  for (int e = 0; e < s->nspec; e++) {
    int nzero = 0;
    for (int l = 0; l < NGLL2; l++) nzero += s->mustore[(size_t)e * NGLL2 + l] == (REAL)0;
    s->is_acoustic[e] = nzero == NGLL2;
    s->nspec_acoustic += s->is_acoustic[e];
  }
  s->any_acoustic = s->nspec_acoustic > 0;
  s->any_elastic = s->nspec_acoustic < s->nspec;
  /* free surface: the non-absorbing outer sides of acoustic elements, in the mesher's order (top, bottom, left,
     right) */
  const size_t nfree_max = s->external ? (size_t)s->ext_nfree : (size_t)2 * (nx + nz);
  s->free_ispec = (int *)calloc(nfree_max + 1, sizeof(int));
  s->free_code = (int *)calloc(nfree_max + 1, sizeof(int));
  s->nfree = 0;
  /* an external mesh's free-surface file: the sides of fluid elements count (read_external_mesh_files.F90:398-416) */
  for (int k = 0; s->external && k < s->ext_nfree; k++) {
    const int e = s->ext_free_elem[k];
    if (s->is_acoustic[e]) { s->free_ispec[s->nfree] = e; s->free_code[s->nfree++] = s->ext_free_code[k]; }
  }
  if (!s->external && !p->absorb_top)
    for (int i = 0; i < nx; i++) {
      int e = (nz - 1) * nx + i;
      if (s->is_acoustic[e]) { s->free_ispec[s->nfree] = e; s->free_code[s->nfree++] = 3; }
    }
  if (!s->external && !p->absorb_bottom)
    for (int i = 0; i < nx; i++) {
      int e = i;
      if (s->is_acoustic[e]) { s->free_ispec[s->nfree] = e; s->free_code[s->nfree++] = 1; }
    }
  if (!s->external && !p->absorb_left)
    for (int j = 0; j < nz; j++) {
      int e = j * nx;
      if (s->is_acoustic[e]) { s->free_ispec[s->nfree] = e; s->free_code[s->nfree++] = 4; }
    }
  if (!s->external && !p->absorb_right)
    for (int j = 0; j < nz; j++) {
      int e = j * nx + nx - 1;
      if (s->is_acoustic[e]) { s->free_ispec[s->nfree] = e; s->free_code[s->nfree++] = 2; }
    }
  /* fluid-solid edges: every acoustic element with an elastic edge neighbour, acoustic elements in order;
     side codes 1 bottom, 2 right, 3 top, 4 left */
  s->coup_ac_ispec = (int *)calloc((size_t)4 * s->nspec + 1, sizeof(int));
  s->coup_ac_code = (int *)calloc((size_t)4 * s->nspec + 1, sizeof(int));
  s->coup_el_ispec = (int *)calloc((size_t)4 * s->nspec + 1, sizeof(int));
  s->coup_el_code = (int *)calloc((size_t)4 * s->nspec + 1, sizeof(int));
  s->ncoup = 0;
  if (s->external && s->any_acoustic && s->any_elastic) {
    /* the elastic element sharing each side of an acoustic element: sides keyed by their sorted node pair */
    const int ns = 4 * s->nspec;
    long *key = (long *)malloc(sizeof(long) * (size_t)ns);
    int *ord = (int *)malloc(sizeof(int) * (size_t)ns);
    for (int e = 0; e < s->nspec; e++)
      for (int c = 1; c <= 4; c++) {
        const int a = s->knods[4 * e + c - 1], b = s->knods[4 * e + c % 4];
        key[4 * e + c - 1] = (long)(a < b ? a : b) * (long)s->npgeo + (a < b ? b : a);
        ord[4 * e + c - 1] = 4 * e + c - 1;
      }
    g_side_key = key;
    qsort(ord, (size_t)ns, sizeof(int), cmp_side_key);
    for (int e = 0; e < s->nspec; e++) {
      if (!s->is_acoustic[e]) continue;
      for (int c = 1; c <= 4; c++) {
        const long k = key[4 * e + c - 1];
        int lo = 0, hi = ns;  /* first sorted side with this key */
        while (lo < hi) {
          const int mid = (lo + hi) / 2;
          if (key[ord[mid]] < k) lo = mid + 1; else hi = mid;
        }
        for (int q = lo; q < ns && key[ord[q]] == k; q++) {
          const int f = ord[q] / 4;
          if (f == e || s->is_acoustic[f]) continue;
          s->coup_ac_ispec[s->ncoup] = e;
          s->coup_ac_code[s->ncoup] = c;
          s->coup_el_ispec[s->ncoup] = f;
          s->coup_el_code[s->ncoup] = ord[q] % 4 + 1;
          s->ncoup++;
        }
      }
    }
    free(ord);
    free(key);
  }
  if (!s->external && s->any_acoustic && s->any_elastic)
    for (int e = 0; e < s->nspec; e++) {
      if (!s->is_acoustic[e]) continue;
      const int ix = e % nx, iz = e / nx;
      const int nb[4] = {iz > 0 ? e - nx : -1, ix > 0 ? e - 1 : -1, ix < nx - 1 ? e + 1 : -1,
                         iz < nz - 1 ? e + nx : -1};
      const int ac_code[4] = {1, 4, 2, 3}, el_code[4] = {3, 2, 4, 1};
      for (int k = 0; k < 4; k++) {
        if (nb[k] < 0 || s->is_acoustic[nb[k]]) continue;
        s->coup_ac_ispec[s->ncoup] = e;
        s->coup_ac_code[s->ncoup] = ac_code[k];
        s->coup_el_ispec[s->ncoup] = nb[k];
        s->coup_el_code[s->ncoup] = el_code[k];
        s->ncoup++;
      }
    }
  /* acoustic Stacey point ranges: an absorbing edge of an acoustic element leaves out the point it shares with
     a fluid-solid edge of the same element (get_coupling_edges.f90:528-563) and, on bottom/top edges, the
     corner points of codeabs_corner (compute_stacey_acoustic.f90) */
  s->abs_jbegin = (int *)calloc((size_t)s->nabs + 1, sizeof(int));
  s->abs_jend = (int *)calloc((size_t)s->nabs + 1, sizeof(int));
  for (int a = 0; a < s->nabs; a++) {
    const int e = s->abs_ispec[a] - 1, code = s->abs_code[a];
    int b = 0, en = NGLL - 1;
    for (int c = 0; c < s->ncoup; c++) {
      if (s->coup_ac_ispec[c] != e) continue;
      const int ce = s->coup_ac_code[c];
      if (code == 4 || code == 2) {
        if (ce == 1) b = 1;
        if (ce == 3) en = NGLL - 2;
      } else {
        if (ce == 4) b = 1;
        if (ce == 2) en = NGLL - 2;
      }
    }
    if (code == 1 && s->abs_corner[4 * a + 0]) b = 1;
    if (code == 1 && s->abs_corner[4 * a + 1]) en = NGLL - 2;
    if (code == 3 && s->abs_corner[4 * a + 2]) b = 1;
    if (code == 3 && s->abs_corner[4 * a + 3]) en = NGLL - 2;
    s->abs_jbegin[a] = b;
    s->abs_jend[a] = en;
  }
}

/* (i, j) of point k of side code of an element, in increasing i or j */
static void side_point(int code, int k, int *i, int *j) {
  if (code == 4) { *i = 0; *j = k; }
  else if (code == 2) { *i = NGLL - 1; *j = k; }
  else if (code == 1) { *i = k; *j = 0; }
  else { *i = k; *j = NGLL - 1; }
}

/* edge geometry at GLL point (i,j) of an absorbing element: outward normal
   and 1D jacobian (compute_stacey_elastic.f90 / prepare_timerun.F90:1316-1460) */
/* ------------------------------------------------------------------ */
/* CPML set-up for the internal mesher (PML_BOUNDARY_CONDITIONS with   */
/* PML_PARAMETER_ADJUSTMENT and ROTATE_PML_ACTIVATE false)             */
/* ------------------------------------------------------------------ */
static void setup_pml(orc_sim *s) {
  const int nspec = s->nspec, nglob = s->nglob;
  s->is_pml = (int *)calloc((size_t)nspec + 1, sizeof(int));
  s->region = (int *)calloc((size_t)nspec + 1, sizeof(int));
  s->spec_to_pml = (int *)calloc((size_t)nspec + 1, sizeof(int));
  s->nspec_pml = 0;
  s->pml_on = 0;
  if (!s->p.pml || s->nabs == 0) return; /* prepare_pml.f90:49: nothing without absorbing edges */
  /* element detection, pml_init.F90:67-160: per side, the elements of its absorbing edges (an element listed
     again for another side takes that entry's flag), then NELEM_PML_THICKNESS-1 times every element sharing a
     corner with the side's elements so far; corner lists are kept as a point mask (same membership test) */
  char *which[5];
  for (int b = 1; b <= 4; b++) which[b] = (char *)calloc((size_t)nspec + 1, 1);
  char *corner = (char *)calloc((size_t)nglob + 1, 1);
  const int cl[4] = {0, NGLL - 1, (NGLL - 1) * NGLL, NGLL2 - 1}; /* (1,1) (5,1) (1,5) (5,5) */
  for (int b = 1; b <= 4; b++) {
    memset(corner, 0, (size_t)nglob);
    for (int a = 0; a < s->nabs; a++) {
      const int e = s->abs_ispec[a] - 1;
      which[b][e] = (s->abs_code[a] == b);
      if (s->abs_code[a] == b)
        for (int c = 0; c < 4; c++) corner[s->ibool[(size_t)e * NGLL2 + cl[c]] - 1] = 1;
    }
    for (int icoef = 2; icoef <= s->p.nelem_pml; icoef++) {
      for (int e = 0; e < nspec; e++) {
        if (which[b][e]) continue;
        for (int c = 0; c < 4; c++)
          if (corner[s->ibool[(size_t)e * NGLL2 + cl[c]] - 1]) which[b][e] = 1;
      }
      memset(corner, 0, (size_t)nglob);
      for (int e = 0; e < nspec; e++) {
        if (!which[b][e]) continue;
        s->is_pml[e] = 1;
        for (int c = 0; c < 4; c++) corner[s->ibool[(size_t)e * NGLL2 + cl[c]] - 1] = 1;
      }
    }
  }
  /* regions, pml_init.F90:201-248 */
  for (int e = 0; e < nspec; e++) {
    if (!s->is_pml[e]) continue;
    const int L = which[4][e], R = which[2][e], T = which[3][e], B = which[1][e];
    int r = 0;
    if (L && !R && !T && !B) r = 1;
    else if (!L && R && !T && !B) r = 1;
    else if (!L && !R && T && !B) r = 2;
    else if (!L && !R && !T && B) r = 2;
    else if (L && !R && T && !B) r = 3;
    else if (!L && R && T && !B) r = 3;
    else if (L && !R && !T && B) r = 3;
    else if (!L && R && !T && B) r = 3;
    s->region[e] = r;
  }
  for (int e = 0; e < nspec; e++)
    if (s->is_pml[e]) s->spec_to_pml[e] = ++s->nspec_pml;
  for (int b = 1; b <= 4; b++) free(which[b]);
  free(corner);
  if (s->nspec_pml == 0) return;
  s->pml_on = 1;

  /* damping profiles, define_PML_coefficients (pml_init.F90:592-1366) */
  const double NPOWER = 2.0, Rcoef = 0.001, PI = 3.141592653589793;
  double xmin = 1e300, zmin = 1e300, xmax = -1e300, zmax = -1e300;
  for (int g = 0; g < nglob; g++) {
    const double x = s->coord[2 * g], z = s->coord[2 * g + 1];
    if (x < xmin) xmin = x;
    if (z < zmin) zmin = z;
    if (x > xmax) xmax = x;
    if (z > zmax) zmax = z;
  }
  const double xorigin = xmin + (xmax - xmin) / 2.0, zorigin = zmin + (zmax - zmin) / 2.0;
  double zminb = 1e30, zmaxb = -1e30, xminr = 1e30, xmaxr = -1e30, zmint = 1e30, zmaxt = -1e30, xminl = 1e30,
         xmaxl = -1e30;
  for (int e = 0; e < nspec; e++) {
    if (!s->is_pml[e]) continue;
    const int r = s->region[e];
    for (int j = 0; j < NGLL; j++)
      for (int i = 0; i < NGLL; i++) {
        const int g = s->ibool[(size_t)e * NGLL2 + j * NGLL + i] - 1;
        const double x = s->coord[2 * g], z = s->coord[2 * g + 1];
        if (z < zorigin && (r == 2 || r == 3)) { zmaxb = fmax(z, zmaxb); zminb = fmin(z, zminb); }
        if (x > xorigin && (r == 1 || r == 3)) { xmaxr = fmax(x, xmaxr); xminr = fmin(x, xminr); }
        if (z > zorigin && (r == 2 || r == 3)) { zmaxt = fmax(z, zmaxt); zmint = fmin(z, zmint); }
        if (x < xorigin && (r == 1 || r == 3)) { xmaxl = fmax(x, xmaxl); xminl = fmin(x, xminl); }
      }
  }
  double thl = xmaxl - xminl, thr = xmaxr - xminr, thb = zmaxb - zminb, tht = zmaxt - zmint;
  if (fabs(thl) > 1e30) thl = 0.0;
  if (fabs(thr) > 1e30) thr = 0.0;
  if (fabs(thb) > 1e30) thb = 0.0;
  if (fabs(tht) > 1e30) tht = 0.0;
  const double xoleft = thl + xmin, xoright = xmax - thr, zobottom = thb + zmin, zotop = zmax - tht;
  double vpmax_ac = 0.0, vpmax_el = 0.0;
  for (int e = 0; e < nspec; e++) {
    if (!s->is_pml[e]) continue;
    for (int l = 0; l < NGLL2; l++) {
      const size_t q = (size_t)e * NGLL2 + l;
      const double vp = (double)(s->rho_vp[q] / s->rhostore[q]);
      if (s->is_acoustic[e]) vpmax_ac = fmax(vpmax_ac, vp);
      else vpmax_el = fmax(vpmax_el, vp);
    }
  }
  double f0max = -1e300;
  for (int k = 0; k < s->nsrc; k++) f0max = fmax(f0max, s->src[k].f0);
  const double alpha_max = PI * f0max;
  s->pml_alpha_max = alpha_max;
  double d0[4][2]; /* [left, right, bottom, top][acoustic, elastic] */
  const double th[4] = {thl, thr, thb, tht};
  for (int k = 0; k < 4; k++) {
    if (th[k] > 0.0) {
      d0[k][0] = -(NPOWER + 1) * vpmax_ac * log(Rcoef) / (2.0 * th[k]);
      d0[k][1] = -(NPOWER + 1) * vpmax_el * log(Rcoef) / (2.0 * th[k]);
    } else {
      d0[k][0] = d0[k][1] = 0.0;
    }
  }
  const size_t np = (size_t)s->nspec_pml * NGLL2;
  s->pml_kx = (double *)calloc(np, sizeof(double));
  s->pml_kz = (double *)calloc(np, sizeof(double));
  s->pml_dx = (double *)calloc(np, sizeof(double));
  s->pml_dz = (double *)calloc(np, sizeof(double));
  s->pml_ax = (double *)calloc(np, sizeof(double));
  s->pml_az = (double *)calloc(np, sizeof(double));
  const double kmin = s->p.k_min_pml, kmax = s->p.k_max_pml;
  for (int e = 0; e < nspec; e++) {
    if (!s->is_pml[e]) continue;
    const int r = s->region[e], ac = s->is_acoustic[e];
    const double damp = ac ? s->p.damping_ac : s->p.damping_el;
    for (int j = 0; j < NGLL; j++)
      for (int i = 0; i < NGLL; i++) {
        double dx = 0.0, dz = 0.0, kx = 1.0, kz = 1.0, ax = 0.0, az = 0.0;
        const int g = s->ibool[(size_t)e * NGLL2 + j * NGLL + i] - 1;
        const double x = s->coord[2 * g], z = s->coord[2 * g + 1];
        /* bottom, top, right, left edges in this order (pml_init.F90:908-1012) */
        for (int side = 0; side < 4; side++) {
          double a;
          int zside = side < 2, apply;
          if (side == 0) { apply = z < zorigin; a = zobottom - z; }
          else if (side == 1) { apply = z > zorigin; a = z - zotop; }
          else if (side == 2) { apply = x > xorigin; a = x - xoright; }
          else { apply = x < xorigin; a = xoleft - x; }
          if (!apply) continue;
          if (zside ? !(r == 2 || r == 3) : !(r == 1 || r == 3)) continue;
          const int kd = side == 0 ? 2 : side == 1 ? 3 : side == 2 ? 1 : 0;
          double dd = 0.0, kk = 1.0, aa = 0.0;
          if (a >= 0.0) {
            const double an = a / th[kd];
            dd = d0[kd][ac ? 0 : 1] / damp * (an * an);
            kk = kmin + (kmax - 1.0) * (an * an);
            aa = alpha_max * (1.0 - an);
          }
          if (zside) {
            dz = dd; kz = kk; az = aa;
            if (r == 2) { dx = 0.0; kx = 1.0; ax = 0.0; }
          } else {
            dx = dd; kx = kk; ax = aa;
            if (r == 1) { dz = 0.0; kz = 1.0; az = 0.0; }
          }
        }
        const size_t q = (size_t)(s->spec_to_pml[e] - 1) * NGLL2 + j * NGLL + i;
        s->pml_dx[q] = dx; s->pml_dz[q] = dz; s->pml_kx[q] = kx; s->pml_kz[q] = kz;
        s->pml_ax[q] = ax; s->pml_az[q] = az;
      }
  }
  /* separation of the corner parameters (pml_init.F90:1269-1356) */
  double dmin = 1.0e30;
  for (int e = 0; e < nspec; e++) {
    if (!s->is_pml[e]) continue;
    for (int j = 0; j < NGLL - 1; j++)
      for (int i = 0; i < NGLL - 1; i++) {
        const int *ib = s->ibool + (size_t)e * NGLL2;
        const int g1 = ib[j * NGLL + i] - 1, g2 = ib[j * NGLL + i + 1] - 1, g3 = ib[(j + 1) * NGLL + i] - 1;
        const double x1 = s->coord[2 * g1], z1 = s->coord[2 * g1 + 1];
        double x2 = s->coord[2 * g2], z2 = s->coord[2 * g2 + 1];
        double dist = (x1 - x2) * (x1 - x2) + (z1 - z2) * (z1 - z2);
        if (dist < dmin) dmin = dist;
        x2 = s->coord[2 * g3]; z2 = s->coord[2 * g3 + 1];
        dist = (x1 - x2) * (x1 - x2) + (z1 - z2) * (z1 - z2);
        if (dist < dmin) dmin = dist;
      }
  }
  dmin = sqrt(dmin);
  const double thzmax = fmax(thb, tht), thxmax = fmax(thl, thr);
  const double mind = alpha_max / 8.0 * dmin / fmax(thxmax, thzmax);
  const double two = mind * 2.0;
  for (int e = 0; e < nspec; e++) {
    if (!s->is_pml[e] || s->region[e] != 3) continue;
    for (int l = 0; l < NGLL2; l++) {
      const size_t q = (size_t)(s->spec_to_pml[e] - 1) * NGLL2 + l;
      const double kx = s->pml_kx[q], kz = s->pml_kz[q];
      double dx = s->pml_dx[q], dz = s->pml_dz[q], ax = s->pml_ax[q], az = s->pml_az[q];
      if (fabs(ax - az) < mind) {
        if (ax > az) ax = az + two;
        else az = ax + two;
      }
      double bx = ax + dx / kx;
      if (fabs(bx - az) < mind) bx = az + two;
      double bz = az + dz / kz;
      if (fabs(bz - ax) < mind) bz = ax + two;
      dx = (bx - ax) * kx;
      dz = (bz - az) * kz;
      s->pml_dx[q] = dx; s->pml_ax[q] = ax; s->pml_dz[q] = dz; s->pml_az[q] = az;
    }
  }
  /* Dirichlet points on the outer boundary of the layer, determine_boundary_abs_points_PML
     (pml_init.F90:306-421): every point of an absorbing edge of a PML element, in increasing iglob */
  char *mel = (char *)calloc((size_t)nglob + 1, 1), *mac = (char *)calloc((size_t)nglob + 1, 1);
  for (int a = 0; a < s->nabs; a++) {
    const int e = s->abs_ispec[a] - 1, code = s->abs_code[a];
    if (!s->is_pml[e]) continue;
    for (int k = 0; k < NGLL; k++) {
      int i, j;
      side_point(code, k, &i, &j);
      const int g = s->ibool[(size_t)e * NGLL2 + j * NGLL + i] - 1;
      if (s->is_acoustic[e]) mac[g] = 1;
      else mel[g] = 1;
    }
  }
  s->npml_el = s->npml_ac = 0;
  for (int g = 0; g < nglob; g++) { s->npml_el += mel[g]; s->npml_ac += mac[g]; }
  s->pml_pts_el = (int *)calloc((size_t)s->npml_el + 1, sizeof(int));
  s->pml_pts_ac = (int *)calloc((size_t)s->npml_ac + 1, sizeof(int));
  int ne = 0, na = 0;
  for (int g = 0; g < nglob; g++) {
    if (mel[g]) s->pml_pts_el[ne++] = g + 1;
    if (mac[g]) s->pml_pts_ac[na++] = g + 1;
  }
  free(mel);
  free(mac);
  s->rm_dux_dx = (REAL *)calloc(np * 2, sizeof(REAL));
  s->rm_duz_dx = (REAL *)calloc(np * 2, sizeof(REAL));
  s->rm_dux_dz = (REAL *)calloc(np * 2, sizeof(REAL));
  s->rm_duz_dz = (REAL *)calloc(np * 2, sizeof(REAL));
  s->rm_displ = (REAL *)calloc(np * 4, sizeof(REAL));
  s->rm_ac_dx = (REAL *)calloc(np * 2, sizeof(REAL));
  s->rm_ac_dz = (REAL *)calloc(np * 2, sizeof(REAL));
  s->rm_pot = (REAL *)calloc(np * 2, sizeof(REAL));
}

/* A deliberate mistake planted in the CPML terms, for the tests that show each term is pinned by the reference's
   traces (0 = none): 1 previous-step fields without the factor 1/2 on the acceleration term, 2 no Newmark damping
   term in the mass matrix, 3 no zero condition on the outer boundary, 4 symmetric stress in PML elements, 5 no
   damping terms in the acceleration, 6 the elastic layer formulation in attenuating layer elements, 7 no Zener
   convolution terms in the viscoelastic layer (A_zener and the phi sums dropped). */
static int g_pml_mistake = 0;
void orc_set_pml_mistake(int which) { g_pml_mistake = which; }

/* recursive convolution coefficients exp(-b dt), (1 - exp(-b dt/2))/b, (1 - exp(-b dt/2)) exp(-b dt/2)/b
   (pml_compute.f90:34-89, compute_coef_convolution) */
static void pml_coef(double bb, double DT, double *c0, double *c1, double *c2) {
  const double temp = exp(-0.5 * bb * DT);
  *c0 = temp * temp;
  if (fabs(bb) > 1e-5) {
    *c1 = (1.0 - temp) / bb;
    *c2 = *c1 * temp;
  } else {
    *c1 = 0.5 * DT;
    *c2 = *c1;
  }
}

/* coefficients of the stretched gradient (pml_compute.f90:93-172, lik_parameter_computation); index_ik 13 or 31
   swaps the roles of the x-only and z-only regions */
static void pml_lik(double DT, double kx, double bx, double ax, double kz, double bz, double az, int region,
                    int index_ik, double A[3], double c1[3], double c2[3]) {
  const int xonly = index_ik == 13 ? 1 : 2, zonly = index_ik == 13 ? 2 : 1;
  if (region == 3) {
    const double A0 = kx / kz;
    const double gx = (ax * bz + ax * ax + 2.0 * bx * az - 2.0 * ax * (bx + az)) / (bz - ax);
    const double gz = (ax * bz + bz * bz + 2.0 * bx * az - 2.0 * bz * (bx + az)) / (ax - bz);
    A[0] = A0;
    A[1] = A0 / 2.0 * (gx - ax);
    A[2] = A0 / 2.0 * (gz - bz);
  } else if (region == xonly) {
    A[0] = kx;
    A[1] = -(kx * (ax - bx));
    A[2] = 0.0;
  } else if (region == zonly) {
    const double A0 = 1.0 / kz;
    A[0] = A0;
    A[1] = 0.0;
    A[2] = -(A0 * (bz - az));
  } else {
    A[0] = A[1] = A[2] = 0.0; /* the reference leaves them undefined outside the three regions */
  }
  pml_coef(ax, DT, &c1[0], &c1[1], &c1[2]);
  pml_coef(bz, DT, &c2[0], &c2[1], &c2[2]);
}

/* coefficients of the damping terms of the acceleration (pml_compute.f90:244-318, l_parameter_computation) */
static void pml_l(double DT, double kx, double bx, double ax, double kz, double bz, double az, int region,
                  double A[5], double c1[3], double c2[3]) {
  if (region == 3) {
    const double A0 = kx * kz;
    const double gx = (ax * az + ax * ax + 2.0 * bx * bz - 2.0 * ax * (bx + bz)) / (az - ax);
    const double gz = (ax * az + az * az + 2.0 * bx * bz - 2.0 * az * (bx + bz)) / (ax - az);
    A[0] = A0;
    A[1] = A0 / 2.0 * (gx - ax + gz - az);
    A[2] = A0 / 2.0 * (ax * ax - gx * ax + az * az - gz * az);
    A[3] = A0 / 2.0 * (ax * ax) * (gx - ax);
    A[4] = A0 / 2.0 * (az * az) * (gz - az);
  } else if (region == 1) {
    A[0] = kx;
    A[1] = kx * (bx - ax);
    A[2] = -(kx * ax * (bx - ax));
    A[3] = kx * (ax * ax) * (bx - ax);
    A[4] = 0.0;
  } else if (region == 2) {
    A[0] = kz;
    A[1] = kz * (bz - az);
    A[2] = -(kz * az * (bz - az));
    A[3] = 0.0;
    A[4] = kz * (az * az) * (bz - az);
  } else {
    A[0] = A[1] = A[2] = A[3] = A[4] = 0.0;
  }
  pml_coef(ax, DT, &c1[0], &c1[1], &c1[2]);
  pml_coef(az, DT, &c2[0], &c2[1], &c2[2]);
}

/* the per-point inputs of the coefficient routines: K, beta = alpha + d/K and alpha along x and z */
static void pml_point(const orc_sim *s, size_t q, double *kx, double *bx, double *ax, double *kz, double *bz,
                      double *az) {
  *kx = s->pml_kx[q]; *kz = s->pml_kz[q];
  *ax = s->pml_ax[q]; *az = s->pml_az[q];
  *bx = *ax + s->pml_dx[q] / *kx;
  *bz = *az + s->pml_dz[q] / *kz;
}

/* zero fields on the outer boundary of the layer (pml_compute.f90:520-605, pml_boundary_acoustic and
   pml_boundary_elastic) */
static void pml_boundary_elastic(orc_sim *s) {
  for (int k = 0; k < s->npml_el; k++) {
    const int g = s->pml_pts_el[k] - 1;
    for (int c = 0; c < 2; c++) {
      s->displ_old[2 * g + c] = (REAL)0;
      s->displ[2 * g + c] = (REAL)0;
      s->veloc[2 * g + c] = (REAL)0;
      s->accel[2 * g + c] = (REAL)0;
    }
  }
}

static void pml_boundary_acoustic(orc_sim *s) {
  for (int k = 0; k < s->npml_ac; k++) {
    const int g = s->pml_pts_ac[k] - 1;
    s->pot_old[g] = (REAL)0;
    s->pot[g] = (REAL)0;
    s->pot_dot[g] = (REAL)0;
    s->pot_dd[g] = (REAL)0;
  }
}

static void edge_normal(const orc_sim *s, int code, size_t l, REAL *nx, REAL *nz, REAL *jac1d) {
  REAL xxi, zxi, xgamma, zgamma;
  if (code == 4 || code == 2) {
    xgamma = -s->xiz[l] * s->jacobian[l];
    zgamma = +s->xix[l] * s->jacobian[l];
    *jac1d = (REAL)sqrt(xgamma * xgamma + zgamma * zgamma);
    if (code == 4) { *nx = -zgamma / *jac1d; *nz = +xgamma / *jac1d; }
    else { *nx = +zgamma / *jac1d; *nz = -xgamma / *jac1d; }
  } else {
    xxi = +s->gammaz[l] * s->jacobian[l];
    zxi = -s->gammax[l] * s->jacobian[l];
    *jac1d = (REAL)sqrt(xxi * xxi + zxi * zxi);
    if (code == 1) { *nx = +zxi / *jac1d; *nz = -xxi / *jac1d; }
    else { *nx = -zxi / *jac1d; *nz = +xxi / *jac1d; }
  }
}

/* ------------------------------------------------------------------ */
/* mass matrix (invert_mass_matrix.F90:34-464, 470-792, 798-848)       */
/* ------------------------------------------------------------------ */
static void setup_mass(orc_sim *s) {
  s->rmass_inv = (REAL *)calloc((size_t)2 * s->nglob, sizeof(REAL));
  s->rmass_ac_inv = (REAL *)calloc((size_t)s->nglob, sizeof(REAL));
  REAL *rm = s->rmass_inv, *rma = s->rmass_ac_inv;
  for (int e = 0; e < s->nspec; e++)
    for (int j = 0; j < NGLL; j++)
      for (int i = 0; i < NGLL; i++) {
        size_t l = (size_t)e * NGLL2 + j * NGLL + i;
        int g = s->ibool[l] - 1;
        REAL wxz = s->wxgll[i] * s->wxgll[j];
        const int r = (s->pml_on && s->is_pml[e]) ? s->region[e] : -1;
        if (s->is_acoustic[e]) {
          double kappal = (double)s->kappastore[l];
          if (r < 0) {
            /* w_i w_j J (single precision) divided by the double kappa (invert_mass_matrix.F90:424-427) */
            rma[g] = (REAL)((double)rma[g] + (double)(wxz * s->jacobian[l]) / kappal);
          } else if (r > 0) {
            /* CPML: w_i w_j / kappa J times K, then the Newmark term dt/2 d (invert_mass_matrix.F90:262-345) */
            const size_t q = (size_t)(s->spec_to_pml[e] - 1) * NGLL2 + j * NGLL + i;
            const double base = (double)wxz / kappal * (double)s->jacobian[l];
            const double kx = s->pml_kx[q], kz = s->pml_kz[q], dx = s->pml_dx[q], dz = s->pml_dz[q];
            const double dto2 = (double)(REAL)(0.5 * (double)(REAL)s->p.dt);
            const double kterm = r == 1 ? kx : r == 3 ? kx * kz : kz;
            const double dterm = r == 1 ? dx * dto2 : r == 3 ? (dx * kz + dz * kx) * dto2 : dz * dto2;
            rma[g] = (REAL)((double)rma[g] + base * kterm);
            rma[g] = (REAL)((double)rma[g] + base * dterm);
          }
          continue;
        }
        double rhol = (double)s->rhostore[l];
        if (r < 0) {
          rm[2 * g] = (REAL)((double)rm[2 * g] + (double)wxz * rhol * (double)s->jacobian[l]);
        } else if (r > 0) {
          /* CPML: w_i w_j rho J times K, then the Newmark term dt/2 d (invert_mass_matrix.F90:133-233) */
          const size_t q = (size_t)(s->spec_to_pml[e] - 1) * NGLL2 + j * NGLL + i;
          const double base = (double)wxz * rhol * (double)s->jacobian[l];
          const double kx = s->pml_kx[q], kz = s->pml_kz[q], dx = s->pml_dx[q], dz = s->pml_dz[q];
          const double dto2 = (double)(REAL)(0.5 * (double)(REAL)s->p.dt);
          const double kterm = r == 1 ? kx : r == 3 ? kx * kz : kz;
          const double dterm = r == 1 ? dx * dto2 : r == 3 ? (dx * kz + dz * kx) * dto2 : dz * dto2;
          rm[2 * g] = (REAL)((double)rm[2 * g] + base * kterm);
          if (g_pml_mistake != 2) rm[2 * g] = (REAL)((double)rm[2 * g] + base * dterm);
        }
        rm[2 * g + 1] = rm[2 * g];
      }
  if (s->p.stacey) {
    REAL deltatover2 = (REAL)(0.5 * (double)(REAL)s->p.dt);
    for (int a = 0; a < s->nabs; a++) {
      int e = s->abs_ispec[a] - 1, code = s->abs_code[a];
      if (s->is_acoustic[e]) {
        /* acoustic Newmark Stacey term dt/2 w / rho_vp (invert_mass_matrix.F90:709-790) */
        for (int k = s->abs_jbegin[a]; k <= s->abs_jend[a]; k++) {
          int i, j;
          side_point(code, k, &i, &j);
          size_t l = (size_t)e * NGLL2 + j * NGLL + i;
          int g = s->ibool[l] - 1;
          REAL nx, nz, j1d;
          edge_normal(s, code, l, &nx, &nz, &j1d);
          REAL weight = j1d * ((code == 4 || code == 2) ? s->wxgll[j] : s->wxgll[i]);
          rma[g] = rma[g] + deltatover2 * weight / s->rho_vp[l];
        }
        continue;
      }
      for (int k = 0; k < NGLL; k++) {
        int i, j;
        if (code == 4) { i = 0; j = k; }
        else if (code == 2) { i = NGLL - 1; j = k; }
        else if (code == 1) { i = k; j = 0; }
        else { i = k; j = NGLL - 1; }
        if (code == 1 && ((s->abs_corner[4 * a + 0] && i == 0) || (s->abs_corner[4 * a + 1] && i == NGLL - 1))) continue;
        if (code == 3 && ((s->abs_corner[4 * a + 2] && i == 0) || (s->abs_corner[4 * a + 3] && i == NGLL - 1))) continue;
        size_t l = (size_t)e * NGLL2 + j * NGLL + i;
        int g = s->ibool[l] - 1;
        REAL nx, nz, j1d;
        edge_normal(s, code, l, &nx, &nz, &j1d);
        REAL weight = j1d * ((code == 4 || code == 2) ? s->wxgll[j] : s->wxgll[i]);
        REAL vx = deltatover2, vz = deltatover2;
        REAL vn = nx * vx + nz * vz;
        REAL tx = s->rho_vp[l] * vn * nx + s->rho_vs[l] * (vx - vn * nx);
        REAL tz = s->rho_vp[l] * vn * nz + s->rho_vs[l] * (vz - vn * nz);
        rm[2 * g] = rm[2 * g] + tx * weight;
        rm[2 * g + 1] = rm[2 * g + 1] + tz * weight;
      }
    }
  }
  for (int g = 0; g < 2 * s->nglob; g++) {
    if (rm[g] <= (REAL)0) rm[g] = (REAL)1;
    rm[g] = (REAL)1 / rm[g];
  }
  for (int g = 0; g < s->nglob; g++) {
    if (rma[g] <= (REAL)0) rma[g] = (REAL)1;
    rma[g] = (REAL)1 / rma[g];
  }
}

/* ------------------------------------------------------------------ */
/* point location (locate_receivers.F90:148-340, locate_source.F90)    */
/* ------------------------------------------------------------------ */
static void locate_point(const orc_sim *s, double px, double pz, int newton_iter, int *ispec_out, double *xi_out,
                         double *gamma_out) {
  double distmin = 1e30;
  int isel = 0, ixg = 0, izg = 0;
  for (int e = 0; e < s->nspec; e++)
    for (int j = 0; j < NGLL; j++)
      for (int i = 0; i < NGLL; i++) {
        int g = s->ibool[e * NGLL2 + j * NGLL + i] - 1;
        double dx = px - s->coord[2 * g], dz = pz - s->coord[2 * g + 1];
        double d2 = dx * dx + dz * dz;
        if (d2 < distmin) { distmin = d2; isel = e; ixg = i; izg = j; }
      }
  /* elements sharing a corner with the selected one, in element order */
  int corners[4] = {s->ibool[isel * NGLL2 + 0], s->ibool[isel * NGLL2 + NGLL - 1],
                    s->ibool[isel * NGLL2 + NGLL2 - 1], s->ibool[isel * NGLL2 + (NGLL - 1) * NGLL]};
  int *cand = (int *)malloc(sizeof(int) * (size_t)(s->nspec + 1));
  int nc = 0;
  cand[nc++] = isel;
  for (int e = 0; e < s->nspec; e++) {
    if (e == isel) continue;
    int hit = 0;
    for (int j = 0; j < NGLL && !hit; j += NGLL - 1)
      for (int i = 0; i < NGLL && !hit; i += NGLL - 1) {
        int g = s->ibool[e * NGLL2 + j * NGLL + i];
        for (int c = 0; c < 4; c++)
          if (g == corners[c]) hit = 1;
      }
    if (hit) cand[nc++] = e;
  }
  double best = 1e30;
  for (int c = 0; c < nc; c++) {
    if (c > 0) { ixg = NGLL / 2 - 1; izg = NGLL / 2 - 1; } /* Fortran int(NGLLX/2.0)=2, 1-based */
    int e = cand[c];
    double xi = s->xigll[ixg], ga = s->xigll[izg];
    double x, z, a, b, cc, d, jac;
    for (int it = 0; it < newton_iter; it++) {
      recompute_jacobian(s, e, xi, ga, &x, &z, &a, &b, &cc, &d, &jac);
      double dx = -(x - px), dz = -(z - pz);
      double dxi = a * dx + b * dz, dga = cc * dx + d * dz;
      xi = xi + dxi;
      ga = ga + dga;
      if (xi > 1.01) xi = 1.01;
      if (xi < -1.01) xi = -1.01;
      if (ga > 1.01) ga = 1.01;
      if (ga < -1.01) ga = -1.01;
    }
    recompute_jacobian(s, e, xi, ga, &x, &z, &a, &b, &cc, &d, &jac);
    double dist = sqrt((px - x) * (px - x) + (pz - z) * (pz - z));
    if (dist < best) {
      best = dist;
      *ispec_out = e + 1;
      *xi_out = xi;
      *gamma_out = ga;
    }
  }
  free(cand);
}

static void setup_sources(orc_sim *s) {
  const double PI = 3.141592653589793;
  int n = s->nsrc;
  s->src_ispec = (int *)calloc((size_t)n + 1, sizeof(int));
  s->src_xi = (double *)calloc((size_t)n + 1, sizeof(double));
  s->src_gamma = (double *)calloc((size_t)n + 1, sizeof(double));
  s->sourcearrays = (REAL *)calloc((size_t)(n + 1) * NGLL2 * 2, sizeof(REAL));
  s->stf = (REAL *)calloc((size_t)(n + 1) * s->p.nstep, sizeof(REAL));
  /* t0 and time shifts (set_source_parameters.f90:116-250) */
  double *t0s = (double *)calloc((size_t)n + 1, sizeof(double));
  double *tsh = (double *)calloc((size_t)n + 1, sizeof(double));
  for (int k = 0; k < n; k++) {
    double hdur = 1.0 / s->src[k].f0;
    tsh[k] = s->src[k].tshift;
    t0s[k] = 1.20 * hdur + tsh[k];
  }
  double t0, min_tshift;
  if (n == 1) {
    t0 = t0s[0];
    min_tshift = tsh[0];
    tsh[0] = 0.0;
  } else {
    t0 = t0s[0];
    min_tshift = tsh[0];
    for (int k = 1; k < n; k++) {
      if (t0s[k] < t0) t0 = t0s[k];
      if (tsh[k] < min_tshift) min_tshift = tsh[k];
    }
    for (int k = 0; k < n; k++) tsh[k] = t0s[k] - t0;
  }
  if (s->p.user_t0 > 0.0 && t0 <= s->p.user_t0 + min_tshift) {
    t0 = s->p.user_t0;
    for (int k = 0; k < n; k++) tsh[k] = t0s[k] - 1.20 / s->src[k].f0;
  }
  s->t0 = t0;
  for (int k = 0; k < n; k++) {
    locate_point(s, s->src[k].x, s->src[k].z, 5, &s->src_ispec[k], &s->src_xi[k], &s->src_gamma[k]);
    double hx[NGLL], hg[NGLL];
    lagrange_any(s->src_xi[k], s->xigll, hx);
    lagrange_any(s->src_gamma[k], s->xigll, hg);
    double ang = s->src[k].angle_deg * PI / 180.0;
    const int acoustic = s->is_acoustic[s->src_ispec[k] - 1];
    if (s->src[k].source_type == 2) {
      /* moment tensor (compute_arrays_source.f90:34-160): the gradient of the Lagrange weights at the source,
         mapped by the element's inverse jacobian interpolated there */
      double hpx[NGLL], hpg[NGLL];
      lagrange_any_deriv(s->src_xi[k], s->xigll, hx, hpx);
      lagrange_any_deriv(s->src_gamma[k], s->xigll, hg, hpg);
      const int e = s->src_ispec[k] - 1;
      double dxis_dx = 0.0, dxis_dz = 0.0, dgam_dx = 0.0, dgam_dz = 0.0;
      for (int m = 0; m < NGLL; m++)
        for (int kk = 0; kk < NGLL; kk++) {
          size_t l = (size_t)e * NGLL2 + m * NGLL + kk;
          double hl = hx[kk] * hg[m];
          dxis_dx = dxis_dx + hl * (double)s->xix[l];
          dxis_dz = dxis_dz + hl * (double)s->xiz[l];
          dgam_dx = dgam_dx + hl * (double)s->gammax[l];
          dgam_dz = dgam_dz + hl * (double)s->gammaz[l];
        }
      const double Mxx = s->src[k].Mxx, Mzz = s->src[k].Mzz, Mxz = s->src[k].Mxz;
      for (int m = 0; m < NGLL; m++)
        for (int kk = 0; kk < NGLL; kk++) {
          double dsx = (hpx[kk] * dxis_dx) * hg[m] + hx[kk] * (hpg[m] * dgam_dx);
          double dsz = (hpx[kk] * dxis_dz) * hg[m] + hx[kk] * (hpg[m] * dgam_dz);
          s->sourcearrays[((size_t)k * NGLL2 + m * NGLL + kk) * 2 + 0] = (REAL)(Mxx * dsx + Mxz * dsz);
          s->sourcearrays[((size_t)k * NGLL2 + m * NGLL + kk) * 2 + 1] = (REAL)(Mxz * dsx + Mzz * dsz);
        }
    } else
    for (int j = 0; j < NGLL; j++)
      for (int i = 0; i < NGLL; i++) {
        double hl = hx[i] * hg[j];
        /* setup_sources_receivers.F90:994-1010: a pressure source in a fluid element, a force in a solid one */
        s->sourcearrays[((size_t)k * NGLL2 + j * NGLL + i) * 2 + 0] = (REAL)(acoustic ? hl : +sin(ang) * hl);
        s->sourcearrays[((size_t)k * NGLL2 + j * NGLL + i) * 2 + 1] = (REAL)(acoustic ? hl : -cos(ang) * hl);
      }
    /* source time function (prepare_source_time_function.f90:200-262, comp_source_time_function.f90:52-200);
       with USE_TRICK_FOR_BETTER_PRESSURE every type is replaced by its second time derivative */
    double f0 = s->src[k].f0, a = (PI * PI) * (f0 * f0); /* PI**2 * f0**2 */
    const int type = s->src[k].time_function_type, trick = s->p.use_trick;
    for (int it = 1; it <= s->p.nstep; it++) {
      double timeval = (double)(it - 1) * s->p.dt;
      double t = timeval - t0 - tsh[k];
      double v, t2 = t * t; /* t**2, in the Fortran's evaluation order below */
      if (type == 1 && !trick) /* Ricker = second derivative of the Gaussian */
        v = (1.0 - 2.0 * a * t2) * exp(-a * t2);
      else if (type == 1) /* fourth derivative of the Gaussian */
        v = -(2.0 * a * (3.0 - 12.0 * a * t * t + 4.0 * (a * a) * t * t * t * t) * exp(-a * t * t));
      else if (type == 2 && !trick) /* first derivative of the Gaussian */
        v = t * exp(-a * t2);
      else if (type == 2) /* third derivative */
        v = 2.0 * a * t * exp(-a * t2) * (-3.0 + 2.0 * a * t2);
      else if (!trick) /* Gaussian */
        v = -exp(-a * t2) / (2.0 * a);
      else /* second derivative of the Gaussian */
        v = (1.0 - 2.0 * a * t2) * exp(-a * t2);
      s->stf[(size_t)k * s->p.nstep + (it - 1)] = (REAL)(-s->src[k].factor * v);
    }
  }
  free(t0s);
  free(tsh);
}

static void setup_receivers(orc_sim *s) {
  const double PI = 3.141592653589793;
  int n = s->nrec;
  s->rec_ispec = (int *)calloc((size_t)n + 1, sizeof(int));
  s->rec_xi = (double *)calloc((size_t)n + 1, sizeof(double));
  s->rec_gamma = (double *)calloc((size_t)n + 1, sizeof(double));
  s->xir = (REAL *)calloc((size_t)(n + 1) * NGLL, sizeof(REAL));
  s->gammar = (REAL *)calloc((size_t)(n + 1) * NGLL, sizeof(REAL));
  for (int r = 0; r < n; r++) {
    locate_point(s, s->rec_xz[2 * r], s->rec_xz[2 * r + 1], 5, &s->rec_ispec[r], &s->rec_xi[r], &s->rec_gamma[r]);
    double hx[NGLL], hg[NGLL];
    lagrange_any(s->rec_xi[r], s->xigll, hx);
    lagrange_any(s->rec_gamma[r], s->xigll, hg);
    for (int i = 0; i < NGLL; i++) {
      s->xir[r * NGLL + i] = (REAL)hx[i];
      s->gammar[r * NGLL + i] = (REAL)hg[i];
    }
  }
  double ang = s->p.anglerec_deg * PI / 180.0;
  s->cosrot = cos(ang);
  s->sinrot = sin(ang);
}

/* ------------------------------------------------------------------ */
/* time step pieces                                                     */
/* ------------------------------------------------------------------ */
static void pml_lik_visco(double kx, double bx, double ax, double kz, double bz, double az, int region, int index_ik,
                          double itau, double te, double A[4]);

static void element_force(orc_sim *s, int e, const REAL *displ, REAL *accel) {
  /* compute_forces_viscoelastic.F90:34-885, isotropic P-SV, no AXISYM; CPML; viscoelastic attenuation in forward
     runs (att_live) -- one element */
  const REAL (*H)[NGLL] = (const REAL(*)[NGLL])s->hprime;
  const REAL (*HW)[NGLL] = (const REAL(*)[NGLL])s->hprimewgll;
  REAL u[2][NGLL][NGLL]; /* u[c][j][i] */
  REAL t1x[NGLL][NGLL], t1z[NGLL][NGLL], t2x[NGLL][NGLL], t2z[NGLL][NGLL];
  REAL pml_acc[NGLL2][2];
  const double DT = s->p.dt;
  {
    if (s->is_acoustic[e]) return;
    const int *ib = s->ibool + (size_t)e * NGLL2;
    const int pml = s->pml_on && s->is_pml[e];
    const int vpml = pml && s->att_live && s->visco_pml && s->visco_pml[e] && g_pml_mistake != 6;
    for (int j = 0; j < NGLL; j++)
      for (int i = 0; i < NGLL; i++) {
        int g = ib[j * NGLL + i] - 1;
        u[0][j][i] = displ[2 * g];
        u[1][j][i] = displ[2 * g + 1];
      }
    for (int j = 0; j < NGLL; j++)
      for (int i = 0; i < NGLL; i++) {
        size_t l = (size_t)e * NGLL2 + j * NGLL + i;
        /* mxm_4comp_singleA: left-to-right sums of five products */
        REAL dux_dxi = u[0][j][0] * H[i][0] + u[0][j][1] * H[i][1] + u[0][j][2] * H[i][2] + u[0][j][3] * H[i][3] +
                       u[0][j][4] * H[i][4];
        REAL duz_dxi = u[1][j][0] * H[i][0] + u[1][j][1] * H[i][1] + u[1][j][2] * H[i][2] + u[1][j][3] * H[i][3] +
                       u[1][j][4] * H[i][4];
        REAL dux_dga = u[0][0][i] * H[j][0] + u[0][1][i] * H[j][1] + u[0][2][i] * H[j][2] + u[0][3][i] * H[j][3] +
                       u[0][4][i] * H[j][4];
        REAL duz_dga = u[1][0][i] * H[j][0] + u[1][1][i] * H[j][1] + u[1][2][i] * H[j][2] + u[1][3][i] * H[j][3] +
                       u[1][4][i] * H[j][4];
        REAL xixl = s->xix[l], xizl = s->xiz[l], gaxl = s->gammax[l], gazl = s->gammaz[l], jacl = s->jacobian[l];
        REAL dux_dxl = dux_dxi * xixl + dux_dga * gaxl;
        REAL dux_dzl = dux_dxi * xizl + dux_dga * gazl;
        REAL duz_dxl = duz_dxi * xixl + duz_dga * gaxl;
        REAL duz_dzl = duz_dxi * xizl + duz_dga * gazl;
        /* CPML: the unmodified gradient (PML_du*) and the stretched one through the memory variables of the
           current and previous-step gradients (pml_compute_memory_variables.f90:230-540, Newmark, not rotated) */
        REAL p_dux_dxl = dux_dxl, p_dux_dzl = dux_dzl, p_duz_dxl = duz_dxl, p_duz_dzl = duz_dzl;
        /* viscoelastic CPML: the bulk (k) and shear (m) stretched gradients with their Zener convolutions, and the
           attenuated unstretched ones (pml_compute_memory_variables.f90:546-894) */
        REAL vk_xx = 0, vk_zz = 0, vm_xx = 0, vm_zz = 0, vm_xz = 0, vm_zx = 0;
        REAL kk_xx = 0, kk_zz = 0, mm_xx = 0, mm_zz = 0, mm_xz = 0, mm_zx = 0;
        if (pml) {
          const REAL *dold = s->displ_old;
          REAL oxi = 0, ozi = 0, oxg = 0, ozg = 0;
          for (int k = 0; k < NGLL; k++) {
            oxi = oxi + dold[2 * (ib[j * NGLL + k] - 1)] * H[i][k];
            ozi = ozi + dold[2 * (ib[j * NGLL + k] - 1) + 1] * H[i][k];
            oxg = oxg + dold[2 * (ib[k * NGLL + i] - 1)] * H[j][k];
            ozg = ozg + dold[2 * (ib[k * NGLL + i] - 1) + 1] * H[j][k];
          }
          const REAL o_dux_dxl = oxi * xixl + oxg * gaxl, o_dux_dzl = oxi * xizl + oxg * gazl;
          const REAL o_duz_dxl = ozi * xixl + ozg * gaxl, o_duz_dzl = ozi * xizl + ozg * gazl;
          const size_t q = (size_t)(s->spec_to_pml[e] - 1) * NGLL2 + j * NGLL + i;
          double kx, bx, ax, kz, bz, az, Azx[3], Axz[3], czx1[3], czx2[3], cxz1[3], cxz2[3];
          pml_point(s, q, &kx, &bx, &ax, &kz, &bz, &az);
          pml_lik(DT, kz, bz, az, kx, bx, ax, s->region[e], 31, Azx, czx1, czx2);
          pml_lik(DT, kx, bx, ax, kz, bz, az, s->region[e], 13, Axz, cxz1, cxz2);
          REAL *rxx = s->rm_dux_dx + 2 * q, *rzx = s->rm_duz_dx + 2 * q;
          REAL *rxz = s->rm_dux_dz + 2 * q, *rzz = s->rm_duz_dz + 2 * q;
          rxx[0] = (REAL)(czx1[0] * rxx[0] + czx1[1] * p_dux_dxl + czx1[2] * o_dux_dxl);
          rzx[0] = (REAL)(czx1[0] * rzx[0] + czx1[1] * p_duz_dxl + czx1[2] * o_duz_dxl);
          rxx[1] = (REAL)(czx2[0] * rxx[1] + czx2[1] * p_dux_dxl + czx2[2] * o_dux_dxl);
          rzx[1] = (REAL)(czx2[0] * rzx[1] + czx2[1] * p_duz_dxl + czx2[2] * o_duz_dxl);
          rxz[0] = (REAL)(cxz1[0] * rxz[0] + cxz1[1] * p_dux_dzl + cxz1[2] * o_dux_dzl);
          rzz[0] = (REAL)(cxz1[0] * rzz[0] + cxz1[1] * p_duz_dzl + cxz1[2] * o_duz_dzl);
          rxz[1] = (REAL)(cxz2[0] * rxz[1] + cxz2[1] * p_dux_dzl + cxz2[2] * o_dux_dzl);
          rzz[1] = (REAL)(cxz2[0] * rzz[1] + cxz2[1] * p_duz_dzl + cxz2[2] * o_duz_dzl);
          if (vpml) {
            const int N = s->nsls;
            const double mk = (double)s->att_mu1[l] * (double)N, mm = (double)s->att_mu2[l] * (double)N;
            REAL ks_xx = 0, ks_zz = 0, ms_xx = 0, ms_zz = 0, ms_xz = 0, ms_zx = 0;  /* kappa_sum_*, mu_sum_* */
            REAL sk_xx = 0, sk_zz = 0, sm_xx = 0, sm_zz = 0, sm_xz = 0, sm_zx = 0;  /* sum_kappa_*, sum_mu_* */
            double a2k = 0, a2m = 0, a3k = 0, a3m = 0, a4k = 0, a4m = 0, a5k = 0, a5m = 0;
            for (int k = 0; k < N; k++) {
              const size_t la = l * N + k, qa = q * N + k;
              const double itk = s->att_itau1[la], itm = s->att_itau2[la];
              const double tek = s->att_te1[la], tem = s->att_te2[la];
              const double phk = s->att_phi1[la], phm = s->att_phi2[la];
              double K31[4], M31[4], K13[4], M13[4], clk[3], clm[3];
              pml_lik_visco(kz, bz, az, kx, bx, ax, s->region[e], 31, itk, tek, K31);
              pml_lik_visco(kz, bz, az, kx, bx, ax, s->region[e], 31, itm, tem, M31);
              pml_lik_visco(kx, bx, ax, kz, bz, az, s->region[e], 13, itk, tek, K13);
              pml_lik_visco(kx, bx, ax, kz, bz, az, s->region[e], 13, itm, tem, M13);
              pml_coef(itk, DT, &clk[0], &clk[1], &clk[2]);
              pml_coef(itm, DT, &clm[0], &clm[1], &clm[2]);
              REAL *kxx = s->vp_ka_xx + qa, *kzz = s->vp_ka_zz + qa;
              REAL *mxx = s->vp_mu_xx + qa, *mzz = s->vp_mu_zz + qa, *mxz = s->vp_mu_xz + qa, *mzx = s->vp_mu_zx + qa;
              *kxx = (REAL)(clk[0] * *kxx + clk[1] * p_dux_dxl + clk[2] * o_dux_dxl);
              *kzz = (REAL)(clk[0] * *kzz + clk[1] * p_duz_dzl + clk[2] * o_duz_dzl);
              *mxx = (REAL)(clm[0] * *mxx + clm[1] * p_dux_dxl + clm[2] * o_dux_dxl);
              *mzz = (REAL)(clm[0] * *mzz + clm[1] * p_duz_dzl + clm[2] * o_duz_dzl);
              *mxz = (REAL)(clm[0] * *mxz + clm[1] * p_dux_dzl + clm[2] * o_dux_dzl);
              *mzx = (REAL)(clm[0] * *mzx + clm[1] * p_duz_dxl + clm[2] * o_duz_dxl);
              a2k = a2k + K31[1]; a2m = a2m + M31[1];
              a3k = a3k + K31[2]; a3m = a3m + M31[2];
              a4k = a4k + K13[1]; a4m = a4m + M13[1];
              a5k = a5k + K13[2]; a5m = a5m + M13[2];
              if (g_pml_mistake == 7) K31[3] = K13[3] = M31[3] = M13[3] = 0.0;
              const double phk_ = g_pml_mistake == 7 ? 0.0 : phk, phm_ = g_pml_mistake == 7 ? 0.0 : phm;
              ks_xx = (REAL)((double)ks_xx + K31[3] * (double)*kxx / mk);
              ks_zz = (REAL)((double)ks_zz + K13[3] * (double)*kzz / mk);
              ms_xx = (REAL)((double)ms_xx + M31[3] * (double)*mxx / mm);
              ms_zz = (REAL)((double)ms_zz + M13[3] * (double)*mzz / mm);
              ms_xz = (REAL)((double)ms_xz + M13[3] * (double)*mxz / mm);
              ms_zx = (REAL)((double)ms_zx + M31[3] * (double)*mzx / mm);
              sk_xx = (REAL)((double)sk_xx + phk_ * (double)*kxx);
              sk_zz = (REAL)((double)sk_zz + phk_ * (double)*kzz);
              sm_xx = (REAL)((double)sm_xx + phm_ * (double)*mxx);
              sm_zz = (REAL)((double)sm_zz + phm_ * (double)*mzz);
              sm_xz = (REAL)((double)sm_xz + phm_ * (double)*mxz);
              sm_zx = (REAL)((double)sm_zx + phm_ * (double)*mzx);
            }
            const double A01 = Azx[0], A10 = Axz[0];
            vk_xx = (REAL)(A01 * p_dux_dxl + ks_xx + a2k * rxx[0] / mk + a3k * rxx[1] / mk);
            vk_zz = (REAL)(A10 * p_duz_dzl + ks_zz + a4k * rzz[0] / mk + a5k * rzz[1] / mk);
            vm_xx = (REAL)(A01 * p_dux_dxl + ms_xx + a2m * rxx[0] / mm + a3m * rxx[1] / mm);
            vm_zz = (REAL)(A10 * p_duz_dzl + ms_zz + a4m * rzz[0] / mm + a5m * rzz[1] / mm);
            vm_xz = (REAL)(A10 * p_dux_dzl + ms_xz + a4m * rxz[0] / mm + a5m * rxz[1] / mm);
            vm_zx = (REAL)(A01 * p_duz_dxl + ms_zx + a2m * rzx[0] / mm + a3m * rzx[1] / mm);
            kk_xx = p_dux_dxl + sk_xx;
            kk_zz = p_duz_dzl + sk_zz;
            mm_xx = p_dux_dxl + sm_xx;
            mm_zz = p_duz_dzl + sm_zz;
            mm_xz = p_dux_dzl + sm_xz;
            mm_zx = p_duz_dxl + sm_zx;
          } else {
            dux_dxl = (REAL)(Azx[0] * p_dux_dxl + Azx[1] * rxx[0] + Azx[2] * rxx[1]);
            duz_dxl = (REAL)(Azx[0] * p_duz_dxl + Azx[1] * rzx[0] + Azx[2] * rzx[1]);
            dux_dzl = (REAL)(Axz[0] * p_dux_dzl + Axz[1] * rxz[0] + Axz[2] * rxz[1]);
            duz_dzl = (REAL)(Axz[0] * p_duz_dzl + Axz[1] * rzz[0] + Axz[2] * rzz[1]);
          }
        }
        REAL mul = s->mustore[l], rhol = s->rhostore[l];
        REAL cpl = s->rho_vp[l] / rhol;
        REAL lambdal = rhol * cpl * cpl - (REAL)2 * mul;
        REAL lp2m = lambdal + (REAL)2 * mul;
        REAL sxx = lp2m * dux_dxl + lambdal * duz_dzl;
        REAL szz = lp2m * duz_dzl + lambdal * dux_dxl;
        REAL sxz = mul * (duz_dxl + dux_dzl);
        if (s->is_aniso && s->is_aniso[e]) {
          /* compute_forces_viscoelastic.F90:372-447 (P-SV): the stress of an anisotropic element */
          const REAL *c = s->aniso_c + l * 6;
          const REAL sh = duz_dxl + dux_dzl;
          sxx = c[0] * dux_dxl + c[1] * duz_dzl + c[2] * sh;
          szz = c[1] * dux_dxl + c[3] * duz_dzl + c[4] * sh;
          sxz = c[2] * dux_dxl + c[4] * duz_dzl + c[5] * sh;
        }
        if (s->att_live && s->att_on[l] && !pml) {  /* nothing in CPML elements (compute_attenuation_viscoelastic.f90:80) */
          /* compute_attenuation_viscoelastic.f90:93-121 (Newmark, convolution memory variables), then the
             unrelaxed-moduli stress correction of compute_forces_viscoelastic.F90:481-496 */
          const int N = s->nsls;
          REAL th = dux_dxl + duz_dzl, th_old = s->old_xx[l] + s->old_zz[l];
          REAL e1s = 0, e11s = 0, e13s = 0;
          for (int k = 0; k < N; k++) {
            size_t q = l * N + k;
            REAL an = s->att_A1[q];
            s->e1[q] = an * an * s->e1[q] + s->att_B1[q] * (th + an * th_old);
            an = s->att_A2[q];
            s->e11[q] = an * an * s->e11[q] +
                        s->att_B2[q] * (dux_dxl - th / (REAL)2 + an * (s->old_xx[l] - th_old / (REAL)2));
            s->e13[q] = an * an * s->e13[q] + s->att_B2[q] * (dux_dzl + duz_dxl + an * s->old_xz[l]);
          }
          for (int k = 0; k < N; k++) {
            e1s = e1s + s->e1[l * N + k];
            e11s = e11s + s->e11[l * N + k];
            e13s = e13s + s->e13[l * N + k];
          }
          s->old_xx[l] = dux_dxl;
          s->old_zz[l] = duz_dzl;
          s->old_xz[l] = dux_dzl + duz_dxl;
          REAL lpm = lambdal + mul;
          sxx = sxx + lpm * e1s + (REAL)2 * mul * e11s;
          sxz = sxz + mul * e13s;
          szz = szz + lpm * e1s - (REAL)2 * mul * e11s;
        }
        REAL szx = sxz;
        if (vpml) {
          /* compute_forces_viscoelastic.F90:545-556 with the unrelaxed moduli, lambda + mu = kappa */
          const REAL lpm = lambdal + mul;
          sxx = (lpm * vk_xx + mul * vm_xx) + (lpm * kk_zz - mul * mm_zz);
          sxz = (mul * vm_zx + mul * mm_xz);
          szx = (mul * mm_zx + mul * vm_xz);
          szz = (lpm * kk_xx - mul * mm_xx) + (lpm * vk_zz + mul * vm_zz);
        } else if (pml) {
          /* non-symmetric stress of the stretched equations (compute_forces_viscoelastic.F90:596-611) */
          sxx = lp2m * dux_dxl + lambdal * p_duz_dzl;
          szz = lp2m * duz_dzl + lambdal * p_dux_dxl;
          szx = mul * (p_duz_dxl + dux_dzl);
          sxz = mul * (p_dux_dzl + duz_dxl);
          if (g_pml_mistake == 4) szx = sxz = mul * (duz_dxl + dux_dzl);
        }
        t1x[j][i] = jacl * (sxx * xixl + szx * xizl);
        t1z[j][i] = jacl * (sxz * xixl + szz * xizl);
        t2x[j][i] = jacl * (sxx * gaxl + szx * gazl);
        t2z[j][i] = jacl * (sxz * gaxl + szz * gazl);
      }
    if (pml) {
      /* damping terms of the acceleration with the memory variables of the displacement
         (pml_compute_accel_contribution.f90:187-333, Newmark) */
      for (int j = 0; j < NGLL; j++)
        for (int i = 0; i < NGLL; i++) {
          const size_t l = (size_t)e * NGLL2 + j * NGLL + i;
          const int g = ib[j * NGLL + i] - 1;
          const REAL fac = s->rhostore[l] * s->jacobian[l];
          const size_t q = (size_t)(s->spec_to_pml[e] - 1) * NGLL2 + j * NGLL + i;
          double kx, bx, ax, kz, bz, az, A[5], c1[3], c2[3];
          pml_point(s, q, &kx, &bx, &ax, &kz, &bz, &az);
          pml_l(DT, kx, bx, ax, kz, bz, az, s->region[e], A, c1, c2);
          REAL *rd = s->rm_displ + 4 * q; /* [series][component] */
          for (int c = 0; c < 2; c++) {
            rd[c] = (REAL)(c1[0] * rd[c] + c1[1] * u[c][j][i] + c1[2] * s->displ_old[2 * g + c]);
            rd[2 + c] = (REAL)(c2[0] * rd[2 + c] + c2[1] * u[c][j][i] + c2[2] * s->displ_old[2 * g + c]);
          }
          const REAL w = s->wxgll[i] * s->wxgll[j] * fac;
          for (int c = 0; c < 2; c++)
            pml_acc[j * NGLL + i][c] = (REAL)(w * (A[1] * s->veloc[2 * g + c] + A[2] * u[c][j][i] + A[3] * rd[c] +
                                                   A[4] * rd[2 + c]));
        }
    }
    for (int j = 0; j < NGLL; j++)
      for (int i = 0; i < NGLL; i++) {
        REAL a1 = 0, a2 = 0, a3 = 0, a4 = 0;
        for (int k = 0; k < NGLL; k++) {
          a1 = a1 + t1x[j][k] * HW[k][i];
          a2 = a2 + t2x[k][i] * HW[k][j];
          a3 = a3 + t1z[j][k] * HW[k][i];
          a4 = a4 + t2z[k][i] * HW[k][j];
        }
        int g = ib[j * NGLL + i] - 1;
        accel[2 * g] = accel[2 * g] - (s->wxgll[j] * a1 + s->wxgll[i] * a2);
        accel[2 * g + 1] = accel[2 * g + 1] - (s->wxgll[j] * a3 + s->wxgll[i] * a4);
      }
    if (pml && g_pml_mistake != 5)
      for (int l = 0; l < NGLL2; l++) {
        const int g = ib[l] - 1;
        accel[2 * g] = accel[2 * g] - pml_acc[l][0];
        accel[2 * g + 1] = accel[2 * g + 1] - pml_acc[l][1];
      }
  }
}

#ifdef _OPENMP
/* greedy element colouring (OpenMP build): each element takes the lowest colour that no element sharing one of its
   points has; a structured quad mesh gets the 2 x 2 checkerboard */
static void setup_colours(orc_sim *s) {
  if (s->colour_elems) return;
  unsigned int *used = (unsigned int *)calloc((size_t)s->nglob, sizeof(unsigned int));
  int *col = (int *)malloc((size_t)s->nspec * sizeof(int));
  int ncol = 0;
  for (int e = 0; e < s->nspec; e++) {
    unsigned int m = 0;
    for (int l = 0; l < NGLL2; l++) m |= used[s->ibool[(size_t)e * NGLL2 + l] - 1];
    int c = 0;
    while (c < 31 && (m >> c & 1u)) c++;
    col[e] = c;
    if (c + 1 > ncol) ncol = c + 1;
    for (int l = 0; l < NGLL2; l++) used[s->ibool[(size_t)e * NGLL2 + l] - 1] |= 1u << c;
  }
  s->ncolour = ncol;
  s->colour_start = (int *)calloc((size_t)ncol + 1, sizeof(int));
  s->colour_elems = (int *)malloc((size_t)s->nspec * sizeof(int));
  for (int e = 0; e < s->nspec; e++) s->colour_start[col[e] + 1]++;
  for (int c = 0; c < ncol; c++) s->colour_start[c + 1] += s->colour_start[c];
  int *fill = (int *)malloc((size_t)ncol * sizeof(int));
  for (int c = 0; c < ncol; c++) fill[c] = s->colour_start[c];
  for (int e = 0; e < s->nspec; e++) s->colour_elems[fill[col[e]]++] = e;
  free(fill);
  free(col);
  free(used);
}
#endif

/* the serial build visits the elements in order, as the reference's CPU loop does (parity); the OpenMP build
   (CPU baseline) runs the elements of each colour in parallel */
static void element_forces(orc_sim *s, const REAL *displ, REAL *accel) {
#ifdef _OPENMP
  setup_colours(s);
  for (int c = 0; c < s->ncolour; c++) {
#pragma omp parallel for schedule(static)
    for (int k = s->colour_start[c]; k < s->colour_start[c + 1]; k++) element_force(s, s->colour_elems[k], displ, accel);
  }
#else
  for (int e = 0; e < s->nspec; e++) element_force(s, e, displ, accel);
#endif
}

static void stacey_forces(orc_sim *s, const REAL *veloc, REAL *accel, REAL *store /* may be NULL */, int it,
                          int nstep_store) {
  /* compute_stacey_elastic.f90:34-1001 (P-SV, no Bielak), SAVE_FORWARD storage */
  for (int a = 0; a < s->nabs; a++) {
    int e = s->abs_ispec[a] - 1, code = s->abs_code[a];
    if (s->is_acoustic[e]) continue;
    for (int k = 0; k < NGLL; k++) {
      int i, j;
      if (code == 4) { i = 0; j = k; }
      else if (code == 2) { i = NGLL - 1; j = k; }
      else if (code == 1) { i = k; j = 0; }
      else { i = k; j = NGLL - 1; }
      if (code == 1 && ((s->abs_corner[4 * a + 0] && i == 0) || (s->abs_corner[4 * a + 1] && i == NGLL - 1))) continue;
      if (code == 3 && ((s->abs_corner[4 * a + 2] && i == 0) || (s->abs_corner[4 * a + 3] && i == NGLL - 1))) continue;
      size_t l = (size_t)e * NGLL2 + j * NGLL + i;
      int g = s->ibool[l] - 1;
      REAL nx, nz, j1d;
      edge_normal(s, code, l, &nx, &nz, &j1d);
      REAL vx = veloc[2 * g], vz = veloc[2 * g + 1];
      REAL vn = nx * vx + nz * vz;
      REAL tx = s->rho_vp[l] * vn * nx + s->rho_vs[l] * (vx - vn * nx);
      REAL tz = s->rho_vp[l] * vn * nz + s->rho_vs[l] * (vz - vn * nz);
      REAL weight = j1d * ((code == 4 || code == 2) ? s->wxgll[j] : s->wxgll[i]);
      /* the reference adds a double-precision traction offset, so the product and
         the subtraction are evaluated in double before rounding */
      double fx = (double)tx * (double)weight, fz = (double)tz * (double)weight;
      accel[2 * g] = (REAL)((double)accel[2 * g] - fx);
      accel[2 * g + 1] = (REAL)((double)accel[2 * g + 1] - fz);
      if (store) {
        (void)nstep_store;
        store[(((size_t)it * s->nabs + a) * NGLL + k) * 2 + 0] = (REAL)fx;
        store[(((size_t)it * s->nabs + a) * NGLL + k) * 2 + 1] = (REAL)fz;
      }
    }
  }
}

static void add_sources(orc_sim *s, REAL *accel, int it /* 1-based */) {
  /* compute_add_sources_viscoelastic.f90:34-100 */
  for (int k = 0; k < s->nsrc; k++) {
    int e = s->src_ispec[k] - 1;
    if (s->is_acoustic[e]) continue;
    REAL stf = s->stf[(size_t)k * s->p.nstep + (it - 1)];
    for (int j = 0; j < NGLL; j++)
      for (int i = 0; i < NGLL; i++) {
        int g = s->ibool[e * NGLL2 + j * NGLL + i] - 1;
        const REAL *sa = s->sourcearrays + ((size_t)k * NGLL2 + j * NGLL + i) * 2;
        accel[2 * g] = accel[2 * g] + sa[0] * stf;
        accel[2 * g + 1] = accel[2 * g + 1] + sa[1] * stf;
      }
  }
}

/* acoustic element forces: compute_forces_viscoacoustic.f90:34-330 (no PML, attenuation or AXISYM) */
static void acoustic_force(orc_sim *s, int e, const REAL *pot, REAL *pot_dd) {
  const REAL (*H)[NGLL] = (const REAL(*)[NGLL])s->hprime;
  const REAL (*HW)[NGLL] = (const REAL(*)[NGLL])s->hprimewgll;
  REAL u[NGLL][NGLL], t1[NGLL][NGLL], t2[NGLL][NGLL]; /* [j][i] */
  REAL pml_acc[NGLL2];
  const double DT = s->p.dt;
  {
    if (!s->is_acoustic[e]) return;
    const int *ib = s->ibool + (size_t)e * NGLL2;
    const int pml = s->pml_on && s->is_pml[e];
    for (int j = 0; j < NGLL; j++)
      for (int i = 0; i < NGLL; i++) u[j][i] = pot[ib[j * NGLL + i] - 1];
    for (int j = 0; j < NGLL; j++)
      for (int i = 0; i < NGLL; i++) {
        size_t l = (size_t)e * NGLL2 + j * NGLL + i;
        /* mxm_2comp_singleA (compute_forces_viscoacoustic.f90:418-477) */
        REAL dxi = u[j][0] * H[i][0] + u[j][1] * H[i][1] + u[j][2] * H[i][2] + u[j][3] * H[i][3] + u[j][4] * H[i][4];
        REAL dga = u[0][i] * H[j][0] + u[1][i] * H[j][1] + u[2][i] * H[j][2] + u[3][i] * H[j][3] + u[4][i] * H[j][4];
        REAL xixl = s->xix[l], xizl = s->xiz[l], gaxl = s->gammax[l], gazl = s->gammaz[l];
        REAL fac = s->jacobian[l] / s->rhostore[l];
        REAL dxl = dxi * xixl + dga * gaxl;
        REAL dzl = dxi * xizl + dga * gazl;
        if (pml) {
          /* stretched gradient of the potential (pml_compute_memory_variables.f90:42-224, Newmark) */
          const REAL *pold = s->pot_old;
          REAL oxi = 0, oga = 0;
          for (int k = 0; k < NGLL; k++) {
            oxi = oxi + pold[ib[j * NGLL + k] - 1] * H[i][k];
            oga = oga + pold[ib[k * NGLL + i] - 1] * H[j][k];
          }
          const REAL odxl = oxi * xixl + oga * gaxl, odzl = oxi * xizl + oga * gazl;
          const REAL pdxl = dxl, pdzl = dzl;
          const size_t q = (size_t)(s->spec_to_pml[e] - 1) * NGLL2 + j * NGLL + i;
          double kx, bx, ax, kz, bz, az, Azx[3], Axz[3], czx1[3], czx2[3], cxz1[3], cxz2[3];
          pml_point(s, q, &kx, &bx, &ax, &kz, &bz, &az);
          pml_lik(DT, kz, bz, az, kx, bx, ax, s->region[e], 31, Azx, czx1, czx2);
          pml_lik(DT, kx, bx, ax, kz, bz, az, s->region[e], 13, Axz, cxz1, cxz2);
          REAL *rx = s->rm_ac_dx + 2 * q, *rz = s->rm_ac_dz + 2 * q;
          rx[0] = (REAL)(czx1[0] * rx[0] + czx1[1] * pdxl + czx1[2] * odxl);
          rx[1] = (REAL)(czx2[0] * rx[1] + czx2[1] * pdxl + czx2[2] * odxl);
          rz[0] = (REAL)(cxz1[0] * rz[0] + cxz1[1] * pdzl + cxz1[2] * odzl);
          rz[1] = (REAL)(cxz2[0] * rz[1] + cxz2[1] * pdzl + cxz2[2] * odzl);
          dxl = (REAL)(Azx[0] * pdxl + Azx[1] * rx[0] + Azx[2] * rx[1]);
          dzl = (REAL)(Axz[0] * pdzl + Axz[1] * rz[0] + Axz[2] * rz[1]);
        }
        t1[j][i] = fac * (xixl * dxl + xizl * dzl);
        t2[j][i] = fac * (gaxl * dxl + gazl * dzl);
      }
    if (pml)
      /* damping terms with the memory variables of the potential (pml_compute_accel_contribution.f90:42-181) */
      for (int j = 0; j < NGLL; j++)
        for (int i = 0; i < NGLL; i++) {
          const size_t l = (size_t)e * NGLL2 + j * NGLL + i;
          const int g = ib[j * NGLL + i] - 1;
          const REAL fac = s->jacobian[l] / s->kappastore[l];
          const size_t q = (size_t)(s->spec_to_pml[e] - 1) * NGLL2 + j * NGLL + i;
          double kx, bx, ax, kz, bz, az, A[5], c1[3], c2[3];
          pml_point(s, q, &kx, &bx, &ax, &kz, &bz, &az);
          pml_l(DT, kx, bx, ax, kz, bz, az, s->region[e], A, c1, c2);
          REAL *rp = s->rm_pot + 2 * q;
          rp[0] = (REAL)(c1[0] * rp[0] + c1[1] * pot[g] + c1[2] * s->pot_old[g]);
          rp[1] = (REAL)(c2[0] * rp[1] + c2[1] * pot[g] + c2[2] * s->pot_old[g]);
          const REAL w = s->wxgll[i] * s->wxgll[j] * fac;
          pml_acc[j * NGLL + i] = (REAL)(w * (A[1] * s->pot_dot[g] + A[2] * pot[g] + A[3] * rp[0] + A[4] * rp[1]));
        }
    for (int j = 0; j < NGLL; j++)
      for (int i = 0; i < NGLL; i++) {
        REAL a1 = 0, a2 = 0;
        for (int k = 0; k < NGLL; k++) {
          a1 = a1 + t1[j][k] * HW[k][i];
          a2 = a2 + t2[k][i] * HW[k][j];
        }
        int g = ib[j * NGLL + i] - 1;
        pot_dd[g] = pot_dd[g] - (s->wxgll[j] * a1 + s->wxgll[i] * a2);
      }
    if (pml)
      for (int l = 0; l < NGLL2; l++) pot_dd[ib[l] - 1] = pot_dd[ib[l] - 1] - pml_acc[l];
  }
}

static void acoustic_forces(orc_sim *s, const REAL *pot, REAL *pot_dd) {
#ifdef _OPENMP
  setup_colours(s);
  for (int c = 0; c < s->ncolour; c++) {
#pragma omp parallel for schedule(static)
    for (int k = s->colour_start[c]; k < s->colour_start[c + 1]; k++) acoustic_force(s, s->colour_elems[k], pot, pot_dd);
  }
#else
  for (int e = 0; e < s->nspec; e++) acoustic_force(s, e, pot, pot_dd);
#endif
}

/* compute_stacey_acoustic.f90:34-160; with store, the SAVE_FORWARD copy of each subtracted term
   (b_absorb_acoustic_*, :47-108) in component 0 of store[it][nabs][5][2] */
static void acoustic_stacey(orc_sim *s, const REAL *pot_dot, REAL *pot_dd, REAL *store, int it) {
  for (int a = 0; a < s->nabs; a++) {
    int e = s->abs_ispec[a] - 1, code = s->abs_code[a];
    if (!s->is_acoustic[e]) continue;
    for (int k = s->abs_jbegin[a]; k <= s->abs_jend[a]; k++) {
      int i, j;
      side_point(code, k, &i, &j);
      size_t l = (size_t)e * NGLL2 + j * NGLL + i;
      int g = s->ibool[l] - 1;
      REAL nx, nz, j1d;
      edge_normal(s, code, l, &nx, &nz, &j1d);
      REAL weight = j1d * ((code == 4 || code == 2) ? s->wxgll[j] : s->wxgll[i]);
      REAL term = pot_dot[g] * weight / s->rho_vp[l];
      pot_dd[g] = pot_dd[g] - term;
      if (store) store[(((size_t)it * s->nabs + a) * NGLL + k) * 2] = term;
    }
  }
}

/* compute_add_sources_acoustic.f90:34-75 (pressure sources, source_type 1) */
static void add_sources_acoustic(orc_sim *s, REAL *pot_dd, int it /* 1-based */) {
  for (int k = 0; k < s->nsrc; k++) {
    int e = s->src_ispec[k] - 1;
    if (!s->is_acoustic[e]) continue;
    REAL stf = s->stf[(size_t)k * s->p.nstep + (it - 1)];
    for (int j = 0; j < NGLL; j++)
      for (int i = 0; i < NGLL; i++) {
        size_t l = (size_t)e * NGLL2 + j * NGLL + i;
        int g = s->ibool[l] - 1;
        pot_dd[g] = pot_dd[g] + s->sourcearrays[((size_t)k * NGLL2 + j * NGLL + i) * 2] * stf / s->kappastore[l];
      }
  }
}

/* point k (ipoin1D - 1) of a fluid-solid edge: the reference walks the acoustic side with ivalue_inverse and the
   elastic side with ivalue, which meet at the same global point (get_coupling_edges.f90:93-111) */
static void coupling_point(int code, int k, int inverse, int *i, int *j) {
  switch (code) {
    case 1: *i = inverse ? NGLL - 1 - k : k; *j = 0; break;              /* bottom */
    case 2: *i = NGLL - 1; *j = inverse ? NGLL - 1 - k : k; break;      /* right */
    case 3: *i = inverse ? k : NGLL - 1 - k; *j = NGLL - 1; break;      /* top */
    default: *i = 0; *j = inverse ? k : NGLL - 1 - k; break;            /* left */
  }
}

/* normal (out of the fluid) and weight at point (i,j) of side code of an acoustic element
   (compute_coupling_acoustic_el.f90:80-140) */
static void coupling_normal(const orc_sim *s, int e, int code, int i, int j, REAL *nx, REAL *nz, REAL *weight) {
  size_t l = (size_t)e * NGLL2 + j * NGLL + i;
  REAL j1d;
  edge_normal(s, code, l, nx, nz, &j1d);
  *weight = j1d * ((code == 4 || code == 2) ? s->wxgll[j] : s->wxgll[i]);
}

/* fluid side: pot_dd += w (u . n) with the solid displacement (compute_coupling_acoustic_el.f90:34-150) */
static void coupling_acoustic_el(orc_sim *s, const REAL *displ, REAL *pot_dd) {
  for (int c = 0; c < s->ncoup; c++) {
    const int ea = s->coup_ac_ispec[c], ca = s->coup_ac_code[c];
    const int ee = s->coup_el_ispec[c], ce = s->coup_el_code[c];
    for (int k = 0; k < NGLL; k++) {
      int i, j;
      coupling_point(ce, k, 1, &i, &j);
      int g = s->ibool[(size_t)ee * NGLL2 + j * NGLL + i] - 1;
      REAL dx = displ[2 * g], dz = displ[2 * g + 1];
      coupling_point(ca, k, 0, &i, &j);
      g = s->ibool[(size_t)ea * NGLL2 + j * NGLL + i] - 1;
      REAL nx, nz, w;
      coupling_normal(s, ea, ca, i, j, &nx, &nz, &w);
      REAL dn = dx * nx + dz * nz;
      pot_dd[g] = pot_dd[g] + w * dn;
    }
  }
}

/* solid side: accel += w n p with the fluid pressure p = -pot_dd (compute_coupling_viscoelastic_ac.f90:34-130) */
static void coupling_elastic_ac(orc_sim *s, const REAL *pot_dd, REAL *accel) {
  for (int c = 0; c < s->ncoup; c++) {
    const int ea = s->coup_ac_ispec[c], ca = s->coup_ac_code[c];
    const int ee = s->coup_el_ispec[c], ce = s->coup_el_code[c];
    for (int k = 0; k < NGLL; k++) {
      int i, j, ii, jj;
      coupling_point(ca, k, 1, &i, &j);
      int g = s->ibool[(size_t)ea * NGLL2 + j * NGLL + i] - 1;
      REAL pressure = -pot_dd[g];
      coupling_point(ce, k, 0, &ii, &jj);
      g = s->ibool[(size_t)ee * NGLL2 + jj * NGLL + ii] - 1;
      REAL nx, nz, w;
      coupling_normal(s, ea, ca, i, j, &nx, &nz, &w);
      accel[2 * g] = accel[2 * g] + w * nx * pressure;
      accel[2 * g + 1] = accel[2 * g + 1] + w * nz * pressure;
    }
  }
}

/* enforce_acoustic_free_surface.f90:34-80 */
static void free_surface_fields(orc_sim *s, REAL *pot, REAL *pot_dot, REAL *pot_dd) {
  for (int f = 0; f < s->nfree; f++)
    for (int k = 0; k < NGLL; k++) {
      int i, j;
      side_point(s->free_code[f], k, &i, &j);
      int g = s->ibool[(size_t)s->free_ispec[f] * NGLL2 + j * NGLL + i] - 1;
      pot[g] = (REAL)0;
      pot_dot[g] = (REAL)0;
      pot_dd[g] = (REAL)0;
    }
}
static void enforce_free_surface(orc_sim *s) { free_surface_fields(s, s->pot, s->pot_dot, s->pot_dd); }

static void record_seismograms(orc_sim *s, const REAL *d, const REAL *v, const REAL *a, double *seis, int it,
                               int nstep) {
  /* write_seismograms.F90:60-200, compute_interpolated_dva.f90: seismotype 1, 2, 3 interpolate the displacement,
     velocity or acceleration (in fluid elements the gradient over the density of the potential, its first or its
     second time derivative), 4 the pressure, 6 the potential */
  const int st = s->p.seismotype;
  const REAL *displ = st == 2 ? v : st == 3 ? a : d;
  const REAL *chi = st == 2 ? s->pot_dot : st == 3 ? s->pot_dd : s->pot;
  if (s->p.seismotype == 4 || s->p.seismotype == 6) {
    /* pressure in fluid elements: -pot with USE_TRICK_FOR_BETTER_PRESSURE, else -pot_dd
       (compute_pressure.f90:460-490); seismotype 6, the potential itself (write_seismograms.F90:147-170);
       solid receivers record zero here */
    const int potential = s->p.seismotype == 6;
    const REAL *f = potential ? s->pot : s->p.use_trick ? s->pot : s->pot_dd;
    for (int r = 0; r < s->nrec; r++) {
      int e = s->rec_ispec[r] - 1;
      double vx = 0.0;
      if (s->is_acoustic[e])
        for (int j = 0; j < NGLL; j++)
          for (int i = 0; i < NGLL; i++) {
            int g = s->ibool[e * NGLL2 + j * NGLL + i] - 1;
            double hl = (double)(s->xir[r * NGLL + i] * s->gammar[r * NGLL + j]);
            vx = vx + (double)(potential ? f[g] : -f[g]) * hl;
          }
      seis[((size_t)r * 2 + 0) * nstep + it] = vx;
      seis[((size_t)r * 2 + 1) * nstep + it] = 0.0;
    }
    return;
  }
  for (int r = 0; r < s->nrec; r++) {
    int e = s->rec_ispec[r] - 1;
    double vx = 0.0, vz = 0.0;
    if (s->is_acoustic[e]) {
      /* vector in a fluid element: grad(chi) / rho, per GLL point in double, stored in single precision, then
         interpolated (compute_vector_field.f90:143-210) */
      const REAL (*H)[NGLL] = (const REAL(*)[NGLL])s->hprime;
      for (int j = 0; j < NGLL; j++)
        for (int i = 0; i < NGLL; i++) {
          size_t l = (size_t)e * NGLL2 + j * NGLL + i;
          double t1 = 0.0, t2 = 0.0;
          for (int k = 0; k < NGLL; k++) t1 = t1 + (double)chi[s->ibool[e * NGLL2 + j * NGLL + k] - 1] * (double)H[i][k];
          for (int k = 0; k < NGLL; k++) t2 = t2 + (double)chi[s->ibool[e * NGLL2 + k * NGLL + i] - 1] * (double)H[j][k];
          double dchi_dx = t1 * (double)s->xix[l] + t2 * (double)s->gammax[l];
          double dchi_dz = t1 * (double)s->xiz[l] + t2 * (double)s->gammaz[l];
          double rhol = (double)s->rhostore[l];
          REAL fx = (REAL)(dchi_dx / rhol), fz = (REAL)(dchi_dz / rhol);
          double hl = (double)(s->xir[r * NGLL + i] * s->gammar[r * NGLL + j]);
          vx = vx + (double)fx * hl;
          vz = vz + (double)fz * hl;
        }
    } else
    for (int j = 0; j < NGLL; j++)
      for (int i = 0; i < NGLL; i++) {
        int g = s->ibool[e * NGLL2 + j * NGLL + i] - 1;
        double hl = (double)(s->xir[r * NGLL + i] * s->gammar[r * NGLL + j]);
        vx = vx + (double)displ[2 * g] * hl;
        vz = vz + (double)displ[2 * g + 1] * hl;
      }
    seis[((size_t)r * 2 + 0) * nstep + it] = s->cosrot * vx + s->sinrot * vz;
    seis[((size_t)r * 2 + 1) * nstep + it] = -s->sinrot * vx + s->cosrot * vz;
  }
}

/* ------------------------------------------------------------------ */
/* public API                                                           */
/* ------------------------------------------------------------------ */
static orc_sim *orc_build(orc_sim *s, const orc_params *p, const double *rho, const double *vp, const double *vs,
                          int nsrc, const orc_source *src, int nrec, const double *rec_xz);

void *orc_new(const orc_params *p, const double *rho, const double *vp, const double *vs, int nsrc,
              const orc_source *src, int nrec, const double *rec_xz) {
  orc_sim *s = (orc_sim *)calloc(1, sizeof(orc_sim));
  s->p = *p;
  setup_gll(s);
  mesh_internal(s);
  return orc_build(s, p, rho, vp, vs, nsrc, src, nrec, rec_xz);
}

/* an external mesh (read_external_mesh): nspec 4-node elements (knods 0-based, counter-clockwise), npgeo control
   nodes, the free-surface and absorbing sides as (0-based element, side code); nx, nz and the absorb flags of the
   parameters are unused, and a CPML layer is not set up */
void *orc_new_external(const orc_params *p, int nspec, int npgeo, const int *knods, const double *coorg, int nfree,
                       const int *free_elem, const int *free_code, int nabs, const int *abs_elem, const int *abs_code,
                       const double *rho, const double *vp, const double *vs, int nsrc, const orc_source *src,
                       int nrec, const double *rec_xz) {
  if (p->pml) return NULL;
  orc_sim *s = (orc_sim *)calloc(1, sizeof(orc_sim));
  s->p = *p;
  setup_gll(s);
  s->external = 1;
  s->nspec = nspec;
  s->npgeo = npgeo;
  s->knods = (int *)malloc(sizeof(int) * 4 * (size_t)nspec);
  memcpy(s->knods, knods, sizeof(int) * 4 * (size_t)nspec);
  s->coorg = (double *)malloc(sizeof(double) * 2 * (size_t)npgeo);
  memcpy(s->coorg, coorg, sizeof(double) * 2 * (size_t)npgeo);
  s->ext_nfree = nfree;
  s->ext_nabs = nabs;
  s->ext_free_elem = (int *)calloc((size_t)nfree + 1, sizeof(int));
  s->ext_free_code = (int *)calloc((size_t)nfree + 1, sizeof(int));
  s->ext_abs_elem = (int *)calloc((size_t)nabs + 1, sizeof(int));
  s->ext_abs_code = (int *)calloc((size_t)nabs + 1, sizeof(int));
  if (nfree) memcpy(s->ext_free_elem, free_elem, sizeof(int) * (size_t)nfree);
  if (nfree) memcpy(s->ext_free_code, free_code, sizeof(int) * (size_t)nfree);
  if (nabs) memcpy(s->ext_abs_elem, abs_elem, sizeof(int) * (size_t)nabs);
  if (nabs) memcpy(s->ext_abs_code, abs_code, sizeof(int) * (size_t)nabs);
  return orc_build(s, p, rho, vp, vs, nsrc, src, nrec, rec_xz);
}

static orc_sim *orc_build(orc_sim *s, const orc_params *p, const double *rho, const double *vp, const double *vs,
                          int nsrc, const orc_source *src, int nrec, const double *rec_xz) {
  setup_numbering(s);
  setup_geometry_materials(s, rho, vp, vs);
  if (p->stacey || p->pml) setup_absorbing(s);
  setup_domains(s);
  s->nsrc = nsrc;
  s->src = (orc_source *)calloc((size_t)nsrc + 1, sizeof(orc_source));
  memcpy(s->src, src, sizeof(orc_source) * (size_t)nsrc);
  setup_pml(s);
  setup_mass(s);
  setup_sources(s);
  s->nrec = nrec;
  s->rec_xz = (double *)calloc((size_t)2 * nrec + 2, sizeof(double));
  memcpy(s->rec_xz, rec_xz, sizeof(double) * 2 * (size_t)nrec);
  setup_receivers(s);
  s->displ = (REAL *)calloc((size_t)2 * s->nglob, sizeof(REAL));
  s->veloc = (REAL *)calloc((size_t)2 * s->nglob, sizeof(REAL));
  s->accel = (REAL *)calloc((size_t)2 * s->nglob, sizeof(REAL));
  s->pot = (REAL *)calloc((size_t)s->nglob, sizeof(REAL));
  s->pot_dot = (REAL *)calloc((size_t)s->nglob, sizeof(REAL));
  s->pot_dd = (REAL *)calloc((size_t)s->nglob, sizeof(REAL));
  s->displ_old = (REAL *)calloc((size_t)2 * s->nglob, sizeof(REAL));
  s->pot_old = (REAL *)calloc((size_t)s->nglob, sizeof(REAL));
  return s;
}

void orc_free(void *h) {
  orc_sim *s = (orc_sim *)h;
  if (!s) return;
  free(s->knods); free(s->coorg); free(s->ibool); free(s->coord);
  free(s->is_aniso); free(s->aniso_c);
  free(s->ext_free_elem); free(s->ext_free_code); free(s->ext_abs_elem); free(s->ext_abs_code);
  free(s->xix); free(s->xiz); free(s->gammax); free(s->gammaz); free(s->jacobian);
  free(s->rhostore); free(s->mustore); free(s->kappastore); free(s->rho_vp); free(s->rho_vs);
  free(s->rmass_inv);
  free(s->abs_ispec); free(s->abs_code); free(s->abs_corner); free(s->ib);
  free(s->src); free(s->src_ispec); free(s->src_xi); free(s->src_gamma); free(s->sourcearrays); free(s->stf);
  free(s->rec_xz); free(s->rec_ispec); free(s->rec_xi); free(s->rec_gamma); free(s->xir); free(s->gammar);
  free(s->displ); free(s->veloc); free(s->accel);
  free(s->is_acoustic); free(s->rmass_ac_inv); free(s->free_ispec); free(s->free_code);
  free(s->coup_ac_ispec); free(s->coup_ac_code); free(s->coup_el_ispec); free(s->coup_el_code);
  free(s->abs_jbegin); free(s->abs_jend);
  free(s->pot); free(s->pot_dot); free(s->pot_dd);
  free(s->att_on); free(s->att_A1); free(s->att_B1); free(s->att_A2); free(s->att_B2);
  free(s->e1); free(s->e11); free(s->e13); free(s->old_xx); free(s->old_zz); free(s->old_xz);
  free(s->att_itau1); free(s->att_itau2); free(s->att_te1); free(s->att_te2); free(s->att_phi1); free(s->att_phi2);
  free(s->att_mu1); free(s->att_mu2); free(s->visco_pml);
  free(s->vp_ka_xx); free(s->vp_ka_zz); free(s->vp_mu_xx); free(s->vp_mu_zz); free(s->vp_mu_xz); free(s->vp_mu_zx);
  free(s->is_pml); free(s->region); free(s->spec_to_pml);
  free(s->pml_kx); free(s->pml_kz); free(s->pml_dx); free(s->pml_dz); free(s->pml_ax); free(s->pml_az);
  free(s->pml_pts_el); free(s->pml_pts_ac);
  free(s->rm_dux_dx); free(s->rm_duz_dx); free(s->rm_dux_dz); free(s->rm_duz_dz); free(s->rm_displ);
  free(s->rm_ac_dx); free(s->rm_ac_dz); free(s->rm_pot); free(s->displ_old); free(s->pot_old);
  free(s->colour_start); free(s->colour_elems);
  free(s);
}

/* Relaxation times of N standard linear solids whose 1/Q fits 1/Qref over [fmin, fmax] (Hz): the linear least
   squares step of attenuation_model.f90:443-486 / 762-799 (USE_SOLVOPT false): relaxation frequencies
   p_l = 2 pi fmin (fmax/fmin)^((l-1)/(N-1)) (exponent in single precision, as remplit_point computes it), 2N-1
   sample frequencies spaced the same way, model 1/Q(w) = sum_l y_l w (p_l - w/Qref) / (p_l^2 + w^2), normal
   equations solved by Gaussian elimination with partial pivoting; tau_sigma = 1/p_l,
   tau_epsilon = tau_sigma (1 + N y_l). */
static void zener_fit(int N, double Qref, double fmin, double fmax, double *tau_eps, double *tau_sig) {
  const double two_pi = 6.283185307179586476925286766559005768;
  double p[16], y[16], A[16][17];
  if (N == 1) p[0] = sqrt(fmin * fmax);
  else
    for (int l = 0; l < N; l++) p[l] = two_pi * pow(fmax / fmin, (double)((float)l / (float)(N - 1))) * fmin;
  const int m = 2 * N - 1;
  for (int a = 0; a < N; a++)
    for (int b = 0; b <= N; b++) A[a][b] = 0.0;
  for (int k = 0; k < m; k++) {
    double w = pow(fmax / fmin, (double)((float)k / (float)(m - 1)));
    w = two_pi * fmin * w;
    double f[16];
    for (int l = 0; l < N; l++) f[l] = w * (p[l] - w / Qref) / (p[l] * p[l] + w * w);
    for (int a = 0; a < N; a++) {
      for (int b = 0; b < N; b++) A[a][b] += f[a] * f[b];
      A[a][N] += (1.0 / Qref) * f[a];
    }
  }
  for (int c = 0; c < N; c++) {
    int piv = c;
    for (int r = c + 1; r < N; r++)
      if (fabs(A[r][c]) > fabs(A[piv][c])) piv = r;
    if (piv != c)
      for (int b = 0; b <= N; b++) {
        double t = A[c][b];
        A[c][b] = A[piv][b];
        A[piv][b] = t;
      }
    for (int r = c + 1; r < N; r++) {
      double f = A[r][c] / A[c][c];
      for (int b = c; b <= N; b++) A[r][b] -= f * A[c][b];
    }
  }
  for (int r = N - 1; r >= 0; r--) {
    double t = A[r][N];
    for (int b = r + 1; b < N; b++) t -= A[r][b] * y[b];
    y[r] = t / A[r][r];
  }
  for (int l = 0; l < N; l++) {
    tau_sig[l] = 1.0 / p[l];
    tau_eps[l] = tau_sig[l] * (1.0 + N * y[l]);
  }
}

/* convolution coefficients [nspec*25][nsls] of the bulk (1) and shear (2) modes */
void orc_get_attenuation(void *h, double *A1, double *B1, double *A2, double *B2) {
  orc_sim *s = (orc_sim *)h;
  const size_t n = (size_t)s->nspec * NGLL2 * s->nsls;
  for (size_t k = 0; k < n && s->att; k++) {
    A1[k] = (double)s->att_A1[k];
    B1[k] = (double)s->att_B1[k];
    A2[k] = (double)s->att_A2[k];
    B2[k] = (double)s->att_B2[k];
  }
}

/* ATTENUATION_VISCOELASTIC with Qkappa, Qmu per GLL point ([nspec][25]) and nsls solids, reference frequency f0
   (the first source's f0 unless it is a Dirac or Heaviside, prepare_attenuation.f90:319-322); fit band
   [f0/10, 10 f0] (attenuation_model.f90:95-104).  Points with both Q above 9998.999 do not attenuate.  Returns 0,
   or 1 for invalid input (nsls outside 1..16, a fluid element, negative relaxation times). */
int orc_set_attenuation(void *h, const double *qkappa, const double *qmu, int nsls, double f0) {
  orc_sim *s = (orc_sim *)h;
  if (nsls < 1 || nsls > 16 || !(f0 > 0.0)) return 1;
  const size_t ne = (size_t)s->nspec * NGLL2;
  s->nsls = nsls;
  s->att_on = (char *)calloc(ne, 1);
  REAL **arr[] = {&s->att_A1, &s->att_B1, &s->att_A2, &s->att_B2, &s->e1, &s->e11, &s->e13};
  for (int k = 0; k < 7; k++) *arr[k] = (REAL *)calloc(ne * nsls, sizeof(REAL));
  s->old_xx = (REAL *)calloc(ne, sizeof(REAL));
  s->old_zz = (REAL *)calloc(ne, sizeof(REAL));
  s->old_xz = (REAL *)calloc(ne, sizeof(REAL));
  REAL **par[] = {&s->att_itau1, &s->att_itau2, &s->att_te1, &s->att_te2, &s->att_phi1, &s->att_phi2};
  for (int k = 0; k < 6; k++) *par[k] = (REAL *)calloc(ne * nsls, sizeof(REAL));
  s->att_mu1 = (REAL *)calloc(ne, sizeof(REAL));
  s->att_mu2 = (REAL *)calloc(ne, sizeof(REAL));
  s->visco_pml = (char *)calloc((size_t)s->nspec + 1, 1);
  const double fmin = f0 / 10.0, fmax = f0 * 10.0;
  double te[2][16], ts[2][16];
  for (size_t l = 0; l < ne; l++) {
    if (qkappa[l] > 9998.999 && qmu[l] > 9998.999) continue;
    if (s->is_acoustic[l / NGLL2]) return 1;
    s->att_on[l] = 1;
    zener_fit(nsls, qkappa[l], fmin, fmax, te[0], ts[0]);
    zener_fit(nsls, qmu[l], fmin, fmax, te[1], ts[1]);
    for (int mode = 0; mode < 2; mode++) {
      double ratio = 0.0;
      for (int k = 0; k < nsls; k++) {
        if (te[mode][k] < 0.0 || ts[mode][k] < 0.0) return 1;
        ratio += te[mode][k] / ts[mode][k];
      }
      REAL *A = mode ? s->att_A2 : s->att_A1, *B = mode ? s->att_B2 : s->att_B1;
      for (int k = 0; k < nsls; k++) {
        /* attenuation_model.f90:152-159, then prepare_attenuation.f90:385-396 in single precision */
        REAL inv_tau = (REAL)(1.0 / ts[mode][k]);
        REAL phi = (REAL)((1.0 - te[mode][k] / ts[mode][k]) / ts[mode][k] / ratio);
        REAL temp = (REAL)exp(-0.5 * (double)inv_tau * s->p.dt);
        REAL coef = (REAL)((1.0 - (double)temp) / (double)inv_tau);
        A[l * nsls + k] = temp;
        B[l * nsls + k] = phi * coef;
        (mode ? s->att_itau2 : s->att_itau1)[l * nsls + k] = inv_tau;
        (mode ? s->att_te2 : s->att_te1)[l * nsls + k] = (REAL)te[mode][k];
        (mode ? s->att_phi2 : s->att_phi1)[l * nsls + k] = phi;
      }
      (mode ? s->att_mu2 : s->att_mu1)[l] = (REAL)(ratio / (double)nsls);
    }
  }
  /* CPML elements whose quality factors are all finite take the viscoelastic layer formulation; the reference
     decides per element from the largest Qkappa and Qmu (compute_forces_viscoelastic.F90:178-181, 249-252) */
  if (s->pml_on) {
    const size_t np = (size_t)s->nspec_pml * NGLL2 * nsls;
    REAL **mv[] = {&s->vp_ka_xx, &s->vp_ka_zz, &s->vp_mu_xx, &s->vp_mu_zz, &s->vp_mu_xz, &s->vp_mu_zx};
    for (int k = 0; k < 6; k++) *mv[k] = (REAL *)calloc(np + 1, sizeof(REAL));
    for (int e = 0; e < s->nspec; e++) {
      if (!s->is_pml[e] || s->is_acoustic[e]) continue;
      double qk = 0.0, qm = 0.0;
      for (int p = 0; p < NGLL2; p++) {
        qk = (qkappa[(size_t)e * NGLL2 + p] > qk ? qkappa[(size_t)e * NGLL2 + p] : qk);
        qm = (qmu[(size_t)e * NGLL2 + p] > qm ? qmu[(size_t)e * NGLL2 + p] : qm);
      }
      s->visco_pml[e] = qk < 9998.999 && qm < 9998.999;
    }
  }
  s->att = 1;
  return 0;
}

/* coefficients of the stretched gradient with a standard linear solid (pml_compute.f90:173-280,
   lik_parameter_computation_viscoelastic): A_0, A_1, A_2 and the Zener term for relaxation rate itau and
   tau_epsilon te; index_ik 13 or 31 as in pml_lik */
static void pml_lik_visco(double kx, double bx, double ax, double kz, double bz, double az, int region, int index_ik,
                          double itau, double te, double A[4]) {
  const int xonly = index_ik == 13 ? 1 : 2, zonly = index_ik == 13 ? 2 : 1;
  if (region == 3) {
    const double A0 = kx / kz;
    /* the partial fractions of (s + beta_x)(s + alpha_z) / ((s + alpha_x)(s + beta_z)): poles alpha_x, beta_z,
       residue terms from decompose_rational_fraction_PML (pml_compute.f90:578-610) */
    const double al[2] = {ax, bz};
    const double be[2] = {(ax * bz + ax * ax + 2.0 * bx * az - 2.0 * ax * (bx + az)) / (bz - ax),
                          (ax * bz + bz * bz + 2.0 * bx * az - 2.0 * bz * (bx + az)) / (ax - bz)};
    double ga[2], et[2];
    for (int i = 0; i < 2; i++) {
      ga[i] = (be[i] * (itau - al[i] * te * itau) + al[i] * al[i] * (te * itau - 1.0)) / (itau - al[i]);
      et[i] = itau * (be[i] - al[i]) / (itau - al[i]);
    }
    A[0] = A0;
    A[1] = A0 / 2.0 * (ga[0] - al[0]);
    A[2] = A0 / 2.0 * (ga[1] - al[1]);
    A[3] = A0 * (te * itau - 1.0) * ((et[0] + et[1]) / 2.0 - itau);
  } else if (region == xonly || region == zonly) {
    const int x = region == xonly;
    const double A0 = x ? kx : 1.0 / kz, al = x ? ax : bz, be = x ? bx : az;
    const double ga = (be * (itau - al * te * itau) + al * al * (te * itau - 1.0)) / (itau - al);
    const double et = itau * (be - al) / (itau - al);
    A[0] = A0;
    A[1] = x ? A0 * (ga - al) : 0.0;
    A[2] = x ? 0.0 : A0 * (ga - al);
    A[3] = A0 * (te * itau - 1.0) * (et - itau);
  } else {
    A[0] = A[1] = A[2] = A[3] = 0.0;
  }
}

/* anisotropic elements (setup_mesh.F90:824-870: c_ij of the material table per GLL point, single precision):
   is_aniso[nspec], c[6][nspec*25] = c11, c13, c15, c33, c35, c55.  Returns -1 with attenuation or CPML (the
   anisotropic branch of the reference has neither here). */
int orc_set_anisotropy(void *h, const int *is_aniso, const double *c) {
  orc_sim *s = (orc_sim *)h;
  if (s->att || s->pml_on) return -1;
  const size_t ne = (size_t)s->nspec * NGLL2;
  free(s->is_aniso);
  free(s->aniso_c);
  s->is_aniso = (int *)calloc((size_t)s->nspec, sizeof(int));
  s->aniso_c = (REAL *)calloc(ne * 6, sizeof(REAL));
  for (int e = 0; e < s->nspec; e++) s->is_aniso[e] = is_aniso[e] != 0;
  for (size_t l = 0; l < ne; l++)
    for (int k = 0; k < 6; k++) s->aniso_c[l * 6 + k] = (REAL)c[k * ne + l];
  return 0;
}

int orc_real_size(void) { return (int)sizeof(REAL); }
int orc_nspec(void *h) { return ((orc_sim *)h)->nspec; }
int orc_nglob(void *h) { return ((orc_sim *)h)->nglob; }
int orc_nabs(void *h) { return ((orc_sim *)h)->nabs; }
double orc_t0(void *h) { return ((orc_sim *)h)->t0; }

void orc_get_ibool(void *h, int *out) {
  orc_sim *s = (orc_sim *)h;
  memcpy(out, s->ibool, sizeof(int) * (size_t)s->nspec * NGLL2);
}
void orc_get_coord(void *h, double *out) {
  orc_sim *s = (orc_sim *)h;
  memcpy(out, s->coord, sizeof(double) * 2 * (size_t)s->nglob);
}
/* which: 0 xix 1 xiz 2 gammax 3 gammaz 4 jacobian 5 rho 6 mu 7 kappa 8 rho_vp 9 rho_vs */
void orc_get_elem_array(void *h, int which, double *out) {
  orc_sim *s = (orc_sim *)h;
  REAL *arr[10] = {s->xix, s->xiz, s->gammax, s->gammaz, s->jacobian,
                   s->rhostore, s->mustore, s->kappastore, s->rho_vp, s->rho_vs};
  size_t n = (size_t)s->nspec * NGLL2;
  for (size_t l = 0; l < n; l++) out[l] = (double)arr[which][l];
}
void orc_get_rmass_inverse(void *h, double *out) {
  orc_sim *s = (orc_sim *)h;
  for (int g = 0; g < 2 * s->nglob; g++) out[g] = (double)s->rmass_inv[g];
}
void orc_get_basis(void *h, double *xigll, double *wxgll, double *hprime, double *hprimewgll) {
  orc_sim *s = (orc_sim *)h;
  for (int i = 0; i < NGLL; i++) {
    xigll[i] = s->xigll[i];
    wxgll[i] = (double)s->wxgll[i];
    for (int k = 0; k < NGLL; k++) {
      hprime[i * NGLL + k] = (double)s->hprime[i][k];
      hprimewgll[i * NGLL + k] = (double)s->hprimewgll[i][k];
    }
  }
}
void orc_get_absorbing(void *h, int *ispec, int *code, int *corner, int *ib) {
  orc_sim *s = (orc_sim *)h;
  for (int a = 0; a < s->nabs; a++) {
    ispec[a] = s->abs_ispec[a];
    code[a] = s->abs_code[a];
    ib[a] = s->ib[a];
    for (int c = 0; c < 4; c++) corner[4 * a + c] = s->abs_corner[4 * a + c];
  }
}
void orc_get_sources(void *h, int *ispec, double *xi, double *gamma, double *sourcearrays, double *stf) {
  orc_sim *s = (orc_sim *)h;
  for (int k = 0; k < s->nsrc; k++) {
    ispec[k] = s->src_ispec[k];
    xi[k] = s->src_xi[k];
    gamma[k] = s->src_gamma[k];
  }
  for (size_t l = 0; l < (size_t)s->nsrc * NGLL2 * 2; l++) sourcearrays[l] = (double)s->sourcearrays[l];
  for (size_t l = 0; l < (size_t)s->nsrc * s->p.nstep; l++) stf[l] = (double)s->stf[l];
}
void orc_get_receivers(void *h, int *ispec, double *xi, double *gamma, double *xir, double *gammar) {
  orc_sim *s = (orc_sim *)h;
  for (int r = 0; r < s->nrec; r++) {
    ispec[r] = s->rec_ispec[r];
    xi[r] = s->rec_xi[r];
    gamma[r] = s->rec_gamma[r];
    for (int i = 0; i < NGLL; i++) {
      xir[r * NGLL + i] = (double)s->xir[r * NGLL + i];
      gammar[r * NGLL + i] = (double)s->gammar[r * NGLL + i];
    }
  }
}

/* Forward time loop (iterate_time.F90:146-414 for SIMULATION_TYPE=1, Newmark).
   Runs the next nstep_run steps (continuing after the steps already run, up to NSTEP) and records displacement
   seismograms into seis[nrec][2][nstep_run].  If stacey_store is not NULL it
   receives the Stacey traction*weight per step [nstep_run][nabs][5][2]. */
void orc_forward(void *h, int nstep_run, double *seis, double *stacey_store_d) {
  orc_sim *s = (orc_sim *)h;
  if (nstep_run < 1 || s->it_done + nstep_run > s->p.nstep) return;
  unsigned int csr = _mm_getcsr();
  if (s->ftz) _mm_setcsr(csr | 0x8040); /* flush-to-zero + denormals-are-zero */
  REAL deltat = (REAL)s->p.dt;
  REAL deltatover2 = (REAL)(0.5 * (double)deltat);
  REAL deltatsqover2 = (REAL)(0.5 * (double)deltat * (double)deltat);
  size_t n2 = (size_t)2 * s->nglob;
  REAL *d = s->displ, *v = s->veloc, *a = s->accel;
  REAL *store = NULL;
  s->att_live = s->att;
  if (stacey_store_d && s->nabs > 0) store = (REAL *)calloc((size_t)nstep_run * s->nabs * NGLL * 2, sizeof(REAL));
  REAL *pa = s->pot, *pv = s->pot_dot, *pdd = s->pot_dd;
  const size_t n1 = (size_t)s->nglob;
  for (int step = 0; step < nstep_run; step++) {
    const int it = s->it_done + step + 1;
    /* update_displacement_newmark_acoustic / _elastic (update_displacement_Newmark.F90:190-454) */
    /* CPML keeps the previous step's fields, advanced by half the acceleration term
       (update_displacement_Newmark.F90:373-380, 428-433) */
    if (s->pml_on && s->any_acoustic)
#pragma omp parallel for
      for (size_t k = 0; k < n1; k++)
        s->pot_old[k] = (REAL)((double)pa[k] + (double)deltatsqover2 / 2.0 * (double)pdd[k]);
    if (s->pml_on && s->any_elastic)
#pragma omp parallel for
      for (size_t k = 0; k < n2; k++)
        s->displ_old[k] = (REAL)((double)d[k] + (double)deltatsqover2 / (g_pml_mistake == 1 ? 1.0 : 2.0) *
                                 (double)a[k]);
    if (s->any_acoustic)
#pragma omp parallel for
      for (size_t k = 0; k < n1; k++) {
        pa[k] = pa[k] + deltat * pv[k] + deltatsqover2 * pdd[k];
        pv[k] = pv[k] + deltatover2 * pdd[k];
        pdd[k] = (REAL)0;
      }
    if (s->any_elastic)
#pragma omp parallel for
      for (size_t k = 0; k < n2; k++) {
        d[k] = d[k] + deltat * v[k] + deltatsqover2 * a[k];
        v[k] = v[k] + deltatover2 * a[k];
        a[k] = (REAL)0;
      }
    /* compute_forces_viscoacoustic_main (compute_forces_viscoacoustic_calling_routine.F90:34-230): free
       surface; iphase 1 (Stacey, coupling with the predicted solid displacement, sources); iphase 2
       (elements); mass; corrector; free surface */
    if (s->any_acoustic) {
      enforce_free_surface(s);
      if (s->pml_on) pml_boundary_acoustic(s);
      if (s->p.stacey && s->nabs > 0) acoustic_stacey(s, pv, pdd, store, step);
      if (s->ncoup > 0) coupling_acoustic_el(s, d, pdd);
      add_sources_acoustic(s, pdd, it);
      acoustic_forces(s, pa, pdd);
      if (s->pml_on) pml_boundary_acoustic(s);
#pragma omp parallel for
      for (size_t k = 0; k < n1; k++) {
        pdd[k] = pdd[k] * s->rmass_ac_inv[k];
        pv[k] = pv[k] + deltatover2 * pdd[k];
      }
      enforce_free_surface(s);
    }
    /* compute_forces_viscoelastic_main: iphase 1 (Stacey, coupling with the new fluid pressure, sources),
       iphase 2 (elements) */
    if (s->any_elastic) {
      if (s->p.stacey && s->nabs > 0) stacey_forces(s, v, a, store, step, nstep_run);
      if (s->pml_on && g_pml_mistake != 3) pml_boundary_elastic(s);
      if (s->ncoup > 0) coupling_elastic_ac(s, pdd, a);
      add_sources(s, a, it);
      element_forces(s, d, a);
      /* mass, then update_veloc_elastic_Newmark */
#pragma omp parallel for
      for (size_t k = 0; k < n2; k++) {
        a[k] = a[k] * s->rmass_inv[k];
        v[k] = v[k] + deltatover2 * a[k];
      }
    }
    if (seis && s->nrec > 0) record_seismograms(s, d, v, a, seis, step, nstep_run);
  }
  s->it_done += nstep_run;
  if (store) {
    size_t n = (size_t)nstep_run * s->nabs * NGLL * 2;
    for (size_t k = 0; k < n; k++) stacey_store_d[k] = (double)store[k];
    free(store);
  }
  s->att_live = 0;
  _mm_setcsr(csr);
}

void orc_set_ftz(void *h, int on) { ((orc_sim *)h)->ftz = on; }
/* threads of the OpenMP build (CPU baseline); returns the count in effect, 1 for the serial builds */
int orc_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
#else
  (void)n;
  return 1;
#endif
}

/* The CPML coefficient routines for one point: out[0:9] lik 31 (A, coef of bb_1, coef of bb_2), out[9:18] lik 13,
   out[18:28] l (A0..A4, coef of alpha_x, coef of alpha_z); inputs K, d, alpha along x and z. */
void orc_pml_coefficients(double DT, double kx, double dx, double ax, double kz, double dz, double az, int region,
                          double *out) {
  const double bx = ax + dx / kx, bz = az + dz / kz;
  pml_lik(DT, kz, bz, az, kx, bx, ax, region, 31, out, out + 3, out + 6);
  pml_lik(DT, kx, bx, ax, kz, bz, az, region, 13, out + 9, out + 12, out + 15);
  pml_l(DT, kx, bx, ax, kz, bz, az, region, out + 18, out + 23, out + 26);
}

/* CPML set-up products: PML flag and region per element; K, d, alpha along x and z per point of each PML element
   ([nspec_pml][25], PML elements in element order); the outer-boundary points (1-based iglob) */
int orc_nspec_pml(void *h) { return ((orc_sim *)h)->nspec_pml; }
int orc_npml_points(void *h, int elastic) {
  orc_sim *s = (orc_sim *)h;
  return elastic ? s->npml_el : s->npml_ac;
}
void orc_get_pml(void *h, int *is_pml, int *region, double *prof /* [6][nspec_pml*25]: kx kz dx dz ax az */,
                 int *pts_el, int *pts_ac) {
  orc_sim *s = (orc_sim *)h;
  for (int e = 0; e < s->nspec; e++) {
    is_pml[e] = s->is_pml[e];
    region[e] = s->region[e];
  }
  const size_t np = (size_t)s->nspec_pml * NGLL2;
  const double *src[6] = {s->pml_kx, s->pml_kz, s->pml_dx, s->pml_dz, s->pml_ax, s->pml_az};
  for (int k = 0; k < 6; k++)
    for (size_t q = 0; q < np; q++) prof[k * np + q] = src[k][q];
  for (int k = 0; k < s->npml_el; k++) pts_el[k] = s->pml_pts_el[k];
  for (int k = 0; k < s->npml_ac; k++) pts_ac[k] = s->pml_pts_ac[k];
}

/* acoustic domain (test accessors) */
int orc_nspec_acoustic(void *h) { return ((orc_sim *)h)->nspec_acoustic; }
int orc_ncoupling(void *h) { return ((orc_sim *)h)->ncoup; }
int orc_nfree(void *h) { return ((orc_sim *)h)->nfree; }
void orc_get_is_acoustic(void *h, int *out) {
  orc_sim *s = (orc_sim *)h;
  memcpy(out, s->is_acoustic, sizeof(int) * (size_t)s->nspec);
}
void orc_get_rmass_acoustic(void *h, double *out) {
  orc_sim *s = (orc_sim *)h;
  for (int g = 0; g < s->nglob; g++) out[g] = (double)s->rmass_ac_inv[g];
}
/* fluid-solid edges: 0-based elements and side codes (1 bottom, 2 right, 3 top, 4 left) */
void orc_get_coupling(void *h, int *ac_ispec, int *ac_code, int *el_ispec, int *el_code) {
  orc_sim *s = (orc_sim *)h;
  for (int c = 0; c < s->ncoup; c++) {
    ac_ispec[c] = s->coup_ac_ispec[c];
    ac_code[c] = s->coup_ac_code[c];
    el_ispec[c] = s->coup_el_ispec[c];
    el_code[c] = s->coup_el_code[c];
  }
}
void orc_get_free_surface(void *h, int *ispec, int *code) {
  orc_sim *s = (orc_sim *)h;
  for (int f = 0; f < s->nfree; f++) {
    ispec[f] = s->free_ispec[f];
    code[f] = s->free_code[f];
  }
}
void orc_get_acoustic_stacey_range(void *h, int *jbegin, int *jend) {
  orc_sim *s = (orc_sim *)h;
  for (int a = 0; a < s->nabs; a++) {
    jbegin[a] = s->abs_jbegin[a];
    jend[a] = s->abs_jend[a];
  }
}
void orc_get_acoustic_fields(void *h, double *pot, double *pot_dot, double *pot_dd) {
  orc_sim *s = (orc_sim *)h;
  for (int g = 0; g < s->nglob; g++) {
    if (pot) pot[g] = (double)s->pot[g];
    if (pot_dot) pot_dot[g] = (double)s->pot_dot[g];
    if (pot_dd) pot_dd[g] = (double)s->pot_dd[g];
  }
}
void orc_set_acoustic_fields(void *h, const double *pot, const double *pot_dot, const double *pot_dd) {
  orc_sim *s = (orc_sim *)h;
  for (int g = 0; g < s->nglob; g++) {
    s->pot[g] = pot ? (REAL)pot[g] : (REAL)0;
    s->pot_dot[g] = pot_dot ? (REAL)pot_dot[g] : (REAL)0;
    s->pot_dd[g] = pot_dd ? (REAL)pot_dd[g] : (REAL)0;
  }
}
/* one application of the assembled acoustic operator: out = -K pot (element forces only) */
void orc_apply_acoustic_stiffness(void *h, const double *pot, double *out) {
  orc_sim *s = (orc_sim *)h;
  REAL *u = (REAL *)calloc((size_t)s->nglob, sizeof(REAL)), *a = (REAL *)calloc((size_t)s->nglob, sizeof(REAL));
  for (int g = 0; g < s->nglob; g++) u[g] = (REAL)pot[g];
  acoustic_forces(s, u, a);
  for (int g = 0; g < s->nglob; g++) out[g] = (double)a[g];
  free(u);
  free(a);
}

void orc_get_fields(void *h, double *displ, double *veloc, double *accel) {
  orc_sim *s = (orc_sim *)h;
  for (int g = 0; g < 2 * s->nglob; g++) {
    if (displ) displ[g] = (double)s->displ[g];
    if (veloc) veloc[g] = (double)s->veloc[g];
    if (accel) accel[g] = (double)s->accel[g];
  }
}

void orc_set_fields(void *h, const double *displ, const double *veloc, const double *accel) {
  orc_sim *s = (orc_sim *)h;
  for (int g = 0; g < 2 * s->nglob; g++) {
    s->displ[g] = displ ? (REAL)displ[g] : (REAL)0;
    s->veloc[g] = veloc ? (REAL)veloc[g] : (REAL)0;
    s->accel[g] = accel ? (REAL)accel[g] : (REAL)0;
  }
}

/* One application of the assembled stiffness operator: out = -K u (element forces
   only, no mass matrix), used to test the operator against brute force. */
void orc_apply_stiffness(void *h, const double *u, double *out) {
  orc_sim *s = (orc_sim *)h;
  size_t n2 = (size_t)2 * s->nglob;
  REAL *d = (REAL *)calloc(n2, sizeof(REAL)), *a = (REAL *)calloc(n2, sizeof(REAL));
  for (size_t k = 0; k < n2; k++) d[k] = (REAL)u[k];
  element_forces(s, d, a);
  for (size_t k = 0; k < n2; k++) out[k] = (double)a[k];
  free(d);
  free(a);
}


/* ------------------------------------------------------------------ */
/* adjoint simulation, SIMULATION_TYPE = 3 with backward reconstruction */
/* ------------------------------------------------------------------ */
static void kernels_accumulate(orc_sim *s, const REAL *displ, const REAL *accel, const REAL *b_displ, REAL *rho_kl,
                               REAL *mu_kl, REAL *kappa_kl) {
  /* compute_kernels.f90 compute_kernels_el (P-SV, isotropic); the jacobian terms are double there */
  const REAL (*H)[NGLL] = (const REAL(*)[NGLL])s->hprime;
  for (int e = 0; e < s->nspec; e++) {
    if (s->is_acoustic[e]) continue;  /* elastic elements only (compute_kernels.f90:119) */
    const int *ib = s->ibool + (size_t)e * NGLL2;
    for (int j = 0; j < NGLL; j++)
      for (int i = 0; i < NGLL; i++) {
        REAL dux_dxi = 0, duz_dxi = 0, dux_dga = 0, duz_dga = 0;
        REAL bdux_dxi = 0, bduz_dxi = 0, bdux_dga = 0, bduz_dga = 0;
        for (int k = 0; k < NGLL; k++) {
          int g1 = ib[j * NGLL + k] - 1, g2 = ib[k * NGLL + i] - 1;
          dux_dxi = dux_dxi + displ[2 * g1] * H[i][k];
          duz_dxi = duz_dxi + displ[2 * g1 + 1] * H[i][k];
          dux_dga = dux_dga + displ[2 * g2] * H[j][k];
          duz_dga = duz_dga + displ[2 * g2 + 1] * H[j][k];
          bdux_dxi = bdux_dxi + b_displ[2 * g1] * H[i][k];
          bduz_dxi = bduz_dxi + b_displ[2 * g1 + 1] * H[i][k];
          bdux_dga = bdux_dga + b_displ[2 * g2] * H[j][k];
          bduz_dga = bduz_dga + b_displ[2 * g2 + 1] * H[j][k];
        }
        size_t l = (size_t)e * NGLL2 + j * NGLL + i;
        double xixl = s->xix[l], xizl = s->xiz[l], gaxl = s->gammax[l], gazl = s->gammaz[l];
        REAL dux_dxl = (REAL)(dux_dxi * xixl + dux_dga * gaxl);
        REAL dux_dzl = (REAL)(dux_dxi * xizl + dux_dga * gazl);
        REAL duz_dxl = (REAL)(duz_dxi * xixl + duz_dga * gaxl);
        REAL duz_dzl = (REAL)(duz_dxi * xizl + duz_dga * gazl);
        REAL bdux_dxl = (REAL)(bdux_dxi * xixl + bdux_dga * gaxl);
        REAL bdux_dzl = (REAL)(bdux_dxi * xizl + bdux_dga * gazl);
        REAL bduz_dxl = (REAL)(bduz_dxi * xixl + bduz_dga * gaxl);
        REAL bduz_dzl = (REAL)(bduz_dxi * xizl + bduz_dga * gazl);
        REAL dsxx = dux_dxl, dsxz = (REAL)0.5 * (duz_dxl + dux_dzl), dszz = duz_dzl;
        REAL bdsxx = bdux_dxl, bdsxz = (REAL)0.5 * (bduz_dxl + bdux_dzl), bdszz = bduz_dzl;
        REAL kap = (dsxx + dszz) * (bdsxx + bdszz);
        REAL mu = dsxx * bdsxx + dszz * bdszz + (REAL)2 * dsxz * bdsxz - (REAL)1 / (REAL)3 * kap;
        int g = ib[j * NGLL + i] - 1;
        REAL rho = accel[2 * g] * b_displ[2 * g] + accel[2 * g + 1] * b_displ[2 * g + 1];
        rho_kl[l] = rho_kl[l] + rho;
        mu_kl[l] = mu_kl[l] + mu;
        kappa_kl[l] = kappa_kl[l] + kap;
      }
  }
}

/* Adjoint run (iterate_time.F90:146-414 with SIMULATION_TYPE=3, Newmark, backward reconstruction):
   adj[nrec][2][nstep] are the adjoint source traces (SEM/ .adj files, component x then z);
   stacey_store[nstep][nabs][5][2] and the last forward frame (lastframe_elastic) come from the
   SAVE_FORWARD run.  Records the adjoint displacement at the receivers into seis[nrec][2][nstep] and
   returns the scaled kernels (save_adjoint_kernels.f90:100-140) in kernels[3][nspec*25] (rho, mu, kappa). */
void orc_adjoint(void *h, const double *adj, const double *stacey_store, const double *last_d, const double *last_v,
                 const double *last_a, double *seis, double *kernels, int ndump, const int *dump_steps,
                 double *b_displ_dump) {
  orc_sim *s = (orc_sim *)h;
  unsigned int csr = _mm_getcsr();
  if (s->ftz) _mm_setcsr(csr | 0x8040);
  const int nstep = s->p.nstep;
  REAL deltat = (REAL)s->p.dt;
  REAL deltatover2 = (REAL)(0.5 * (double)deltat);
  REAL deltatsqover2 = (REAL)(0.5 * (double)deltat * (double)deltat);
  REAL b_deltat = -deltat;
  REAL b_deltatover2 = (REAL)(0.5 * (double)b_deltat);
  REAL b_deltatsqover2 = (REAL)(0.5 * (double)b_deltat * (double)b_deltat);
  size_t n2 = (size_t)2 * s->nglob, ne = (size_t)s->nspec * NGLL2;
  REAL *d = s->displ, *v = s->veloc, *a = s->accel;
  memset(d, 0, sizeof(REAL) * n2); memset(v, 0, sizeof(REAL) * n2); memset(a, 0, sizeof(REAL) * n2);
  REAL *bd = (REAL *)calloc(n2, sizeof(REAL)), *bv = (REAL *)calloc(n2, sizeof(REAL)), *ba = (REAL *)calloc(n2, sizeof(REAL));
  REAL *rho_kl = (REAL *)calloc(ne, sizeof(REAL)), *mu_kl = (REAL *)calloc(ne, sizeof(REAL)), *kappa_kl = (REAL *)calloc(ne, sizeof(REAL));
  REAL *sadj = (REAL *)calloc((size_t)(s->nrec + 1) * 2 * nstep, sizeof(REAL));
  for (size_t k = 0; k < (size_t)s->nrec * 2 * nstep; k++) sadj[k] = (REAL)adj[k];
  for (int it = 1; it <= nstep; it++) {
    int it_tmp = nstep - it + 1;
    for (size_t k = 0; k < n2; k++) {
      d[k] = d[k] + deltat * v[k] + deltatsqover2 * a[k];
      v[k] = v[k] + deltatover2 * a[k];
      a[k] = (REAL)0;
    }
    for (size_t k = 0; k < n2; k++) {
      bd[k] = bd[k] + b_deltat * bv[k] + b_deltatsqover2 * ba[k];
      bv[k] = bv[k] + b_deltatover2 * ba[k];
      ba[k] = (REAL)0;
    }
    /* adjoint wavefield: Stacey, adjoint sources, elements, mass, corrector */
    if (s->p.stacey && s->nabs > 0) stacey_forces(s, v, a, NULL, 0, 0);
    for (int r = 0; r < s->nrec; r++) {
      int e = s->rec_ispec[r] - 1;
      for (int j = 0; j < NGLL; j++)
        for (int i = 0; i < NGLL; i++) {
          int g = s->ibool[e * NGLL2 + j * NGLL + i] - 1;
          REAL hl = s->xir[r * NGLL + i] * s->gammar[r * NGLL + j];
          REAL stfx = hl * sadj[((size_t)r * 2 + 0) * nstep + (it_tmp - 1)];
          REAL stfz = hl * sadj[((size_t)r * 2 + 1) * nstep + (it_tmp - 1)];
          a[2 * g] = a[2 * g] + stfx;
          a[2 * g + 1] = a[2 * g + 1] + stfz;
        }
    }
    element_forces(s, d, a);
    for (size_t k = 0; k < n2; k++) a[k] = a[k] * s->rmass_inv[k];
    for (size_t k = 0; k < n2; k++) v[k] = v[k] + deltatover2 * a[k];
    /* backward (reconstructed) wavefield: stored Stacey, source at NSTEP-it+1, elements, mass, corrector */
    if (s->p.stacey && s->nabs > 0) {
      for (int f = 0; f < s->nabs; f++) {
        int e = s->abs_ispec[f] - 1, code = s->abs_code[f];
        for (int k = 0; k < NGLL; k++) {
          int i, j;
          if (code == 4) { i = 0; j = k; }
          else if (code == 2) { i = NGLL - 1; j = k; }
          else if (code == 1) { i = k; j = 0; }
          else { i = k; j = NGLL - 1; }
          int g = s->ibool[e * NGLL2 + j * NGLL + i] - 1;
          const double *st = stacey_store + (((size_t)(it_tmp - 1) * s->nabs + f) * NGLL + k) * 2;
          ba[2 * g] = ba[2 * g] - (REAL)st[0];
          ba[2 * g + 1] = ba[2 * g + 1] - (REAL)st[1];
        }
      }
    }
    add_sources(s, ba, it_tmp);
    element_forces(s, bd, ba);
    for (size_t k = 0; k < n2; k++) ba[k] = ba[k] * s->rmass_inv[k];
    for (size_t k = 0; k < n2; k++) bv[k] = bv[k] + b_deltatover2 * ba[k];
    if (it == 1) {
      for (size_t k = 0; k < n2; k++) {
        bd[k] = (REAL)last_d[k];
        bv[k] = (REAL)last_v[k];
        ba[k] = (REAL)last_a[k];
      }
    }
    if (seis && s->nrec > 0) record_seismograms(s, d, v, a, seis, it - 1, nstep);
    kernels_accumulate(s, d, a, bd, rho_kl, mu_kl, kappa_kl);
    for (int q = 0; q < ndump; q++)
      if (dump_steps[q] == it)
        for (size_t k = 0; k < n2; k++) b_displ_dump[(size_t)q * n2 + k] = (double)bd[k];
  }
  for (size_t l = 0; l < ne; l++) {
    REAL rhol = s->rhostore[l], mul = s->mustore[l], kappal = s->kappastore[l];
    kernels[l] = (double)(-rhol * (deltat * (REAL)1) * rho_kl[l]);
    kernels[ne + l] = (double)(REAL)(-2.0 * (double)mul * (double)(deltat * (REAL)1) * (double)mu_kl[l]);
    kernels[2 * ne + l] = (double)(-kappal * (deltat * (REAL)1) * kappa_kl[l]);
  }
  free(bd); free(bv); free(ba); free(rho_kl); free(mu_kl); free(kappa_kl); free(sadj);
  _mm_setcsr(csr);
}

/* compute_kernels.f90:293-449 compute_kernels_ac (CPU path, NTSTEP_BETWEEN_COMPUTE_KERNELS = 1): gradients of
   the adjoint potential and of the reconstructed potential (the xi/gamma derivatives in single precision, the
   jacobian products in double, rounded to single precision), then
   rho_ac_kl += 1/rho (grad pot . grad b_pot) deltat and kappa_ac_kl += 1/kappa pot_dd b_pot deltat */
static void kernels_ac_accumulate(orc_sim *s, const REAL *pot, const REAL *pot_dd, const REAL *b_pot, REAL deltat,
                                  REAL *rho_ac_kl, REAL *kappa_ac_kl) {
  const REAL (*H)[NGLL] = (const REAL(*)[NGLL])s->hprime;
  for (int e = 0; e < s->nspec; e++) {
    if (!s->is_acoustic[e]) continue;
    const int *ib = s->ibool + (size_t)e * NGLL2;
    for (int j = 0; j < NGLL; j++)
      for (int i = 0; i < NGLL; i++) {
        const size_t l = (size_t)e * NGLL2 + j * NGLL + i;
        const int g = ib[j * NGLL + i] - 1;
        REAL t1 = 0, t2 = 0, bt1 = 0, bt2 = 0;
        for (int k = 0; k < NGLL; k++) {
          const int g1 = ib[j * NGLL + k] - 1, g2 = ib[k * NGLL + i] - 1;
          t1 = t1 + pot[g1] * H[i][k];
          bt1 = bt1 + b_pot[g1] * H[i][k];
          t2 = t2 + pot[g2] * H[j][k];
          bt2 = bt2 + b_pot[g2] * H[j][k];
        }
        const double xixl = s->xix[l], xizl = s->xiz[l], gaxl = s->gammax[l], gazl = s->gammaz[l];
        const REAL a1 = (REAL)((double)t1 * xixl + (double)t2 * gaxl);
        const REAL a2 = (REAL)((double)t1 * xizl + (double)t2 * gazl);
        const REAL b1 = (REAL)((double)bt1 * xixl + (double)bt2 * gaxl);
        const REAL b2 = (REAL)((double)bt1 * xizl + (double)bt2 * gazl);
        const REAL dot = a1 * b1 + a2 * b2;
        rho_ac_kl[l] = rho_ac_kl[l] + (REAL)1 / s->rhostore[l] * dot * deltat;
        kappa_ac_kl[l] = kappa_ac_kl[l] + (REAL)1 / s->kappastore[l] * pot_dd[g] * b_pot[g] * deltat;
      }
  }
}

/* SIMULATION_TYPE=3 run with backward reconstruction over fluid, solid and coupled meshes (iterate_time.F90:155-390,
   Newmark, no CPML, no attenuation).  Per step: the predictors of the adjoint and of the reconstructed fields;
   the adjoint solid (compute_forces_viscoelastic_calling_routine.F90:34-205: Stacey, coupling with the adjoint
   potential as the pressure, compute_coupling_viscoelastic_ac.f90:98-102, adjoint sources, elements, mass,
   corrector), then the adjoint fluid (compute_forces_viscoacoustic_calling_routine.F90:34-228: Stacey, coupling
   with minus the new adjoint solid acceleration, compute_coupling_acoustic_el.f90:104-109, adjoint sources
   without 1/kappa, compute_add_sources_acoustic.f90:247-290, elements, mass, corrector); then the reconstructed
   fluid and solid (:234-374 and compute_forces_viscoelastic_calling_routine.F90:214-352: stored Stacey terms of step
   NSTEP-it+1, forward coupling, sources at NSTEP-it+1); after step 1 the last forward frames replace the
   reconstructed fields (read_forward_arrays.f90:52-98); seismograms; kernels.
   adj[nrec][2][nstep]: a receiver in a fluid element takes component 0 (source_adjoint(:,:,1)).
   stacey_store[nstep][nabs][5][2] from the SAVE_FORWARD run (fluid faces in component 0).
   last_el: displ, veloc, accel [3][2*nglob] (may be NULL without solids); last_ac: potential and its two time
   derivatives [3][nglob] (NULL without fluids).  kernels_el[3][ne]: rho, mu, kappa as save_adjoint_kernels.f90
   scales them; kernels_ac[4][ne]: rho_ac_kl, kappa_ac_kl, rhop_ac_kl, alpha_ac_kl (compute_kernels.f90:426-439).
   Returns 0, or -1 for a set-up it does not cover. */
int orc_adjoint_full(void *h, const double *adj, const double *stacey_store, const double *last_el,
                     const double *last_ac, double *seis, double *kernels_el, double *kernels_ac) {
  orc_sim *s = (orc_sim *)h;
  if (s->pml_on || s->att) return -1;
  const int any_ac = s->nspec_acoustic > 0, any_el = s->nspec_acoustic < s->nspec;
  if ((any_el && !last_el) || (any_ac && !last_ac)) return -1;
  unsigned int csr = _mm_getcsr();
  if (s->ftz) _mm_setcsr(csr | 0x8040);
  const int nstep = s->p.nstep;
  REAL deltat = (REAL)s->p.dt;
  REAL deltatover2 = (REAL)(0.5 * (double)deltat);
  REAL deltatsqover2 = (REAL)(0.5 * (double)deltat * (double)deltat);
  REAL b_deltat = -deltat;
  REAL b_deltatover2 = (REAL)(0.5 * (double)b_deltat);
  REAL b_deltatsqover2 = (REAL)(0.5 * (double)b_deltat * (double)b_deltat);
  const size_t n1 = (size_t)s->nglob, n2 = 2 * n1, ne = (size_t)s->nspec * NGLL2;
  REAL *d = s->displ, *v = s->veloc, *a = s->accel;
  REAL *pa = s->pot, *pv = s->pot_dot, *pdd = s->pot_dd;
  memset(d, 0, sizeof(REAL) * n2); memset(v, 0, sizeof(REAL) * n2); memset(a, 0, sizeof(REAL) * n2);
  memset(pa, 0, sizeof(REAL) * n1); memset(pv, 0, sizeof(REAL) * n1); memset(pdd, 0, sizeof(REAL) * n1);
  REAL *bd = (REAL *)calloc(n2, sizeof(REAL)), *bv = (REAL *)calloc(n2, sizeof(REAL)), *ba = (REAL *)calloc(n2, sizeof(REAL));
  REAL *bp = (REAL *)calloc(n1, sizeof(REAL)), *bpv = (REAL *)calloc(n1, sizeof(REAL)), *bpdd = (REAL *)calloc(n1, sizeof(REAL));
  REAL *tmp1 = (REAL *)calloc(n1, sizeof(REAL)), *tmp2 = (REAL *)calloc(n2, sizeof(REAL));
  REAL *rho_kl = (REAL *)calloc(ne, sizeof(REAL)), *mu_kl = (REAL *)calloc(ne, sizeof(REAL)), *kappa_kl = (REAL *)calloc(ne, sizeof(REAL));
  REAL *rho_ac = (REAL *)calloc(ne, sizeof(REAL)), *kappa_ac = (REAL *)calloc(ne, sizeof(REAL));
  REAL *sadj = (REAL *)calloc((size_t)(s->nrec + 1) * 2 * nstep, sizeof(REAL));
  for (size_t k = 0; k < (size_t)s->nrec * 2 * nstep; k++) sadj[k] = (REAL)adj[k];
  for (int it = 1; it <= nstep; it++) {
    const int it_tmp = nstep - it + 1;
    /* update_displ_Newmark, then update_displ_Newmark_backward (update_displacement_Newmark.F90:58-120) */
    for (size_t k = 0; k < n1 && any_ac; k++) {
      pa[k] = pa[k] + deltat * pv[k] + deltatsqover2 * pdd[k];
      pv[k] = pv[k] + deltatover2 * pdd[k];
      pdd[k] = (REAL)0;
    }
    for (size_t k = 0; k < n2 && any_el; k++) {
      d[k] = d[k] + deltat * v[k] + deltatsqover2 * a[k];
      v[k] = v[k] + deltatover2 * a[k];
      a[k] = (REAL)0;
    }
    for (size_t k = 0; k < n1 && any_ac; k++) {
      bp[k] = bp[k] + b_deltat * bpv[k] + b_deltatsqover2 * bpdd[k];
      bpv[k] = bpv[k] + b_deltatover2 * bpdd[k];
      bpdd[k] = (REAL)0;
    }
    for (size_t k = 0; k < n2 && any_el; k++) {
      bd[k] = bd[k] + b_deltat * bv[k] + b_deltatsqover2 * ba[k];
      bv[k] = bv[k] + b_deltatover2 * ba[k];
      ba[k] = (REAL)0;
    }
    /* adjoint solid first (iterate_time.F90:227-246) */
    if (any_el) {
      if (s->p.stacey && s->nabs > 0) stacey_forces(s, v, a, NULL, 0, 0);
      if (s->ncoup > 0) {
        for (size_t k = 0; k < n1; k++) tmp1[k] = -pa[k];  /* pressure = potential (coupling negates its input) */
        coupling_elastic_ac(s, tmp1, a);
      }
      for (int r = 0; r < s->nrec; r++) {
        const int e = s->rec_ispec[r] - 1;
        if (s->is_acoustic[e]) continue;
        for (int j = 0; j < NGLL; j++)
          for (int i = 0; i < NGLL; i++) {
            const int g = s->ibool[e * NGLL2 + j * NGLL + i] - 1;
            const REAL hl = s->xir[r * NGLL + i] * s->gammar[r * NGLL + j];
            a[2 * g] = a[2 * g] + hl * sadj[((size_t)r * 2 + 0) * nstep + (it_tmp - 1)];
            a[2 * g + 1] = a[2 * g + 1] + hl * sadj[((size_t)r * 2 + 1) * nstep + (it_tmp - 1)];
          }
      }
      element_forces(s, d, a);
      for (size_t k = 0; k < n2; k++) a[k] = a[k] * s->rmass_inv[k];
      for (size_t k = 0; k < n2; k++) v[k] = v[k] + deltatover2 * a[k];
    }
    /* then the adjoint fluid */
    if (any_ac) {
      free_surface_fields(s, pa, pv, pdd);
      if (s->p.stacey && s->nabs > 0) acoustic_stacey(s, pv, pdd, NULL, 0);
      if (s->ncoup > 0) {
        for (size_t k = 0; k < n2; k++) tmp2[k] = -a[k];  /* displacement = minus the adjoint acceleration */
        coupling_acoustic_el(s, tmp2, pdd);
      }
      for (int r = 0; r < s->nrec; r++) {
        const int e = s->rec_ispec[r] - 1;
        if (!s->is_acoustic[e]) continue;
        for (int j = 0; j < NGLL; j++)
          for (int i = 0; i < NGLL; i++) {
            const int g = s->ibool[e * NGLL2 + j * NGLL + i] - 1;
            const REAL stf = s->xir[r * NGLL + i] * s->gammar[r * NGLL + j] * sadj[((size_t)r * 2) * nstep + (it_tmp - 1)];
            pdd[g] = pdd[g] + stf;
          }
      }
      acoustic_forces(s, pa, pdd);
      for (size_t k = 0; k < n1; k++) {
        pdd[k] = pdd[k] * s->rmass_ac_inv[k];
        pv[k] = pv[k] + deltatover2 * pdd[k];
      }
      free_surface_fields(s, pa, pv, pdd);
    }
    /* reconstructed fluid, then solid, in the forward order */
    if (any_ac) {
      free_surface_fields(s, bp, bpv, bpdd);
      if (s->p.stacey && s->nabs > 0)
        for (int f = 0; f < s->nabs; f++) {
          const int e = s->abs_ispec[f] - 1, code = s->abs_code[f];
          if (!s->is_acoustic[e]) continue;
          for (int k = s->abs_jbegin[f]; k <= s->abs_jend[f]; k++) {
            int i, j;
            side_point(code, k, &i, &j);
            const int g = s->ibool[(size_t)e * NGLL2 + j * NGLL + i] - 1;
            bpdd[g] = bpdd[g] - (REAL)stacey_store[(((size_t)(it_tmp - 1) * s->nabs + f) * NGLL + k) * 2];
          }
        }
      if (s->ncoup > 0) coupling_acoustic_el(s, bd, bpdd);
      add_sources_acoustic(s, bpdd, it_tmp);
      acoustic_forces(s, bp, bpdd);
      for (size_t k = 0; k < n1; k++) {
        bpdd[k] = bpdd[k] * s->rmass_ac_inv[k];
        bpv[k] = bpv[k] + b_deltatover2 * bpdd[k];
      }
      free_surface_fields(s, bp, bpv, bpdd);
    }
    if (any_el) {
      if (s->p.stacey && s->nabs > 0)
        for (int f = 0; f < s->nabs; f++) {
          const int e = s->abs_ispec[f] - 1, code = s->abs_code[f];
          if (s->is_acoustic[e]) continue;
          for (int k = 0; k < NGLL; k++) {
            int i, j;
            if (code == 4) { i = 0; j = k; }
            else if (code == 2) { i = NGLL - 1; j = k; }
            else if (code == 1) { i = k; j = 0; }
            else { i = k; j = NGLL - 1; }
            const int g = s->ibool[e * NGLL2 + j * NGLL + i] - 1;
            const double *st = stacey_store + (((size_t)(it_tmp - 1) * s->nabs + f) * NGLL + k) * 2;
            ba[2 * g] = ba[2 * g] - (REAL)st[0];
            ba[2 * g + 1] = ba[2 * g + 1] - (REAL)st[1];
          }
        }
      if (s->ncoup > 0) coupling_elastic_ac(s, bpdd, ba);
      add_sources(s, ba, it_tmp);
      element_forces(s, bd, ba);
      for (size_t k = 0; k < n2; k++) ba[k] = ba[k] * s->rmass_inv[k];
      for (size_t k = 0; k < n2; k++) bv[k] = bv[k] + b_deltatover2 * ba[k];
    }
    if (it == 1) {
      for (size_t k = 0; k < n2 && any_el; k++) {
        bd[k] = (REAL)last_el[k];
        bv[k] = (REAL)last_el[n2 + k];
        ba[k] = (REAL)last_el[2 * n2 + k];
      }
      for (size_t k = 0; k < n1 && any_ac; k++) {
        bp[k] = (REAL)last_ac[k];
        bpv[k] = (REAL)last_ac[n1 + k];
        bpdd[k] = (REAL)last_ac[2 * n1 + k];
      }
      if (any_ac) free_surface_fields(s, bp, bpv, bpdd);
    }
    if (seis && s->nrec > 0) record_seismograms(s, d, v, a, seis, it - 1, nstep);
    if (any_el) kernels_accumulate(s, d, a, bd, rho_kl, mu_kl, kappa_kl);
    if (any_ac) kernels_ac_accumulate(s, pa, pdd, bp, deltat * (REAL)1, rho_ac, kappa_ac);
  }
  for (size_t l = 0; l < ne && kernels_el; l++) {
    REAL rhol = s->rhostore[l], mul = s->mustore[l], kappal = s->kappastore[l];
    kernels_el[l] = (double)(-rhol * (deltat * (REAL)1) * rho_kl[l]);
    kernels_el[ne + l] = (double)(REAL)(-2.0 * (double)mul * (double)(deltat * (REAL)1) * (double)mu_kl[l]);
    kernels_el[2 * ne + l] = (double)(-kappal * (deltat * (REAL)1) * kappa_kl[l]);
  }
  for (size_t l = 0; l < ne && kernels_ac; l++) {
    kernels_ac[l] = rho_ac[l];
    kernels_ac[ne + l] = kappa_ac[l];
    kernels_ac[2 * ne + l] = (double)(rho_ac[l] + kappa_ac[l]);
    kernels_ac[3 * ne + l] = (double)((REAL)2 * kappa_ac[l]);
  }
  free(bd); free(bv); free(ba); free(bp); free(bpv); free(bpdd); free(tmp1); free(tmp2);
  free(rho_kl); free(mu_kl); free(kappa_kl); free(rho_ac); free(kappa_ac); free(sadj);
  _mm_setcsr(csr);
  return 0;
}

/* One accumulation step of the fluid kernels for given fields (adjoint potential and its second time derivative,
   reconstructed potential) and time step dt; out[2][nspec*25]: rho_ac_kl, kappa_ac_kl increments. */
void orc_kernel_ac_increment(void *h, const double *pot, const double *pot_dd, const double *b_pot, double dt,
                             double *out) {
  orc_sim *s = (orc_sim *)h;
  const size_t n1 = (size_t)s->nglob, ne = (size_t)s->nspec * NGLL2;
  REAL *p = (REAL *)malloc(sizeof(REAL) * n1), *a = (REAL *)malloc(sizeof(REAL) * n1), *b = (REAL *)malloc(sizeof(REAL) * n1);
  for (size_t k = 0; k < n1; k++) { p[k] = (REAL)pot[k]; a[k] = (REAL)pot_dd[k]; b[k] = (REAL)b_pot[k]; }
  REAL *r = (REAL *)calloc(ne, sizeof(REAL)), *kk = (REAL *)calloc(ne, sizeof(REAL));
  kernels_ac_accumulate(s, p, a, b, (REAL)dt, r, kk);
  for (size_t l = 0; l < ne; l++) { out[l] = r[l]; out[ne + l] = kk[l]; }
  free(p); free(a); free(b); free(r); free(kk);
}

/* One accumulation step of the kernels (unscaled) for given fields; out[3][nspec*25]. */
void orc_kernel_increment(void *h, const double *displ, const double *accel, const double *b_displ, double *out) {
  orc_sim *s = (orc_sim *)h;
  size_t n2 = (size_t)2 * s->nglob, ne = (size_t)s->nspec * NGLL2;
  REAL *d = (REAL *)malloc(sizeof(REAL) * n2), *a = (REAL *)malloc(sizeof(REAL) * n2), *b = (REAL *)malloc(sizeof(REAL) * n2);
  for (size_t k = 0; k < n2; k++) { d[k] = (REAL)displ[k]; a[k] = (REAL)accel[k]; b[k] = (REAL)b_displ[k]; }
  REAL *r = (REAL *)calloc(ne, sizeof(REAL)), *m = (REAL *)calloc(ne, sizeof(REAL)), *kk = (REAL *)calloc(ne, sizeof(REAL));
  kernels_accumulate(s, d, a, b, r, m, kk);
  for (size_t l = 0; l < ne; l++) { out[l] = r[l]; out[ne + l] = m[l]; out[2 * ne + l] = kk[l]; }
  free(d); free(a); free(b); free(r); free(m); free(kk);
}

/* ------------------------------------------------------------------ */
/* Exact discrete gradient of the L2 waveform misfit (reverse mode of  */
/* the forward time loop above).  No counterpart in the reference: the */
/* reference's SIMULATION_TYPE=3 kernels approximate the same gradient */
/* by a continuous adjoint with backward reconstruction.               */
/* ------------------------------------------------------------------ */
/* field gradient (d/dx, d/dz of ux and uz) at the element points of element e, as the force kernel forms it */
static void point_gradients(const orc_sim *s, const REAL *u, int e, REAL (*gr)[4] /* [25][4]: ux_x ux_z uz_x uz_z */) {
  const REAL (*H)[NGLL] = (const REAL(*)[NGLL])s->hprime;
  const int *ib = s->ibool + (size_t)e * NGLL2;
  for (int j = 0; j < NGLL; j++)
    for (int i = 0; i < NGLL; i++) {
      REAL dux_dxi = 0, duz_dxi = 0, dux_dga = 0, duz_dga = 0;
      for (int k = 0; k < NGLL; k++) {
        int g1 = ib[j * NGLL + k] - 1, g2 = ib[k * NGLL + i] - 1;
        dux_dxi = dux_dxi + u[2 * g1] * H[i][k];
        duz_dxi = duz_dxi + u[2 * g1 + 1] * H[i][k];
        dux_dga = dux_dga + u[2 * g2] * H[j][k];
        duz_dga = duz_dga + u[2 * g2 + 1] * H[j][k];
      }
      size_t l = (size_t)e * NGLL2 + j * NGLL + i;
      REAL xixl = s->xix[l], xizl = s->xiz[l], gaxl = s->gammax[l], gazl = s->gammaz[l];
      gr[j * NGLL + i][0] = dux_dxi * xixl + dux_dga * gaxl;
      gr[j * NGLL + i][1] = dux_dxi * xizl + dux_dga * gazl;
      gr[j * NGLL + i][2] = duz_dxi * xixl + duz_dga * gaxl;
      gr[j * NGLL + i][3] = duz_dxi * xizl + duz_dga * gazl;
    }
}

/* Misfit chi = 1/2 dt sum_n sum_r |s_n - o_n|^2 of the forward run of orc_forward (from rest, all NSTEP steps;
   seismotype displacement), with obs[nrec][2][nstep].  Returns chi and the exact gradient of the discrete
   forward map with respect to the per-GLL-point parameters, parts[5][nspec*25] = d chi / d (kappa, mu, rho
   through the mass matrix, rho*vp and rho*vs through the Stacey traction and its mass term), and with respect
   to the MODEL=gll parameters grad[3][nspec*25] = d chi / d (rho, vp, vs) (mu = rho vs^2,
   kappa = rho vp^2 - mu).  The forward states are kept in memory (small meshes only). */
double orc_gradient(void *h, const double *obs, double *parts, double *grad) {
  orc_sim *s = (orc_sim *)h;
  const int N = s->p.nstep;
  const size_t n2 = (size_t)2 * s->nglob, ne = (size_t)s->nspec * NGLL2;
  unsigned int csr = _mm_getcsr();
  if (s->ftz) _mm_setcsr(csr | 0x8040);
  REAL deltat = (REAL)s->p.dt;
  REAL deltatover2 = (REAL)(0.5 * (double)deltat);
  REAL deltatsqover2 = (REAL)(0.5 * (double)deltat * (double)deltat);
  /* forward: d_n, predicted veloc v'_n and the unscaled force r_n for n = 1..N */
  REAL *Dn = (REAL *)calloc(n2 * N, sizeof(REAL)), *Vp = (REAL *)calloc(n2 * N, sizeof(REAL));
  REAL *Rn = (REAL *)calloc(n2 * N, sizeof(REAL));
  double *seis = (double *)calloc((size_t)(s->nrec > 0 ? s->nrec : 1) * 2 * N, sizeof(double));
  REAL *d = (REAL *)calloc(n2, sizeof(REAL)), *v = (REAL *)calloc(n2, sizeof(REAL)), *a = (REAL *)calloc(n2, sizeof(REAL));
  for (int n = 1; n <= N; n++) {
    for (size_t k = 0; k < n2; k++) {
      d[k] = d[k] + deltat * v[k] + deltatsqover2 * a[k];
      v[k] = v[k] + deltatover2 * a[k];
      a[k] = (REAL)0;
    }
    if (s->p.stacey && s->nabs > 0) stacey_forces(s, v, a, NULL, 0, 0);
    add_sources(s, a, n);
    element_forces(s, d, a);
    memcpy(Dn + n2 * (n - 1), d, n2 * sizeof(REAL));
    memcpy(Vp + n2 * (n - 1), v, n2 * sizeof(REAL));
    memcpy(Rn + n2 * (n - 1), a, n2 * sizeof(REAL));
    for (size_t k = 0; k < n2; k++) a[k] = a[k] * s->rmass_inv[k];
    for (size_t k = 0; k < n2; k++) v[k] = v[k] + deltatover2 * a[k];
    if (s->nrec > 0) record_seismograms(s, d, d, d, seis, n - 1, N); /* the misfit is on the displacement */
  }
  double chi = 0.0;
  for (int r = 0; r < s->nrec; r++)
    for (int c = 0; c < 2; c++)
      for (int n = 0; n < N; n++) {
        double q = seis[((size_t)r * 2 + c) * N + n] - obs[((size_t)r * 2 + c) * N + n];
        chi += 0.5 * (double)deltat * q * q;
      }
  /* reverse: (D, V, A) are the cotangents of (d_n, v_n, a_n) */
  REAL *D = (REAL *)calloc(n2, sizeof(REAL)), *V = (REAL *)calloc(n2, sizeof(REAL)), *A = (REAL *)calloc(n2, sizeof(REAL));
  REAL *Rb = (REAL *)calloc(n2, sizeof(REAL)), *Vb = (REAL *)calloc(n2, sizeof(REAL));
  double *gk = parts, *gm = parts + ne, *gr = parts + 2 * ne, *gvp = parts + 3 * ne, *gvs = parts + 4 * ne;
  memset(parts, 0, sizeof(double) * 5 * ne);
  double *pm = (double *)calloc(n2, sizeof(double)); /* d chi / d M_c(g) summed over steps */
  REAL (*gd)[4] = (REAL(*)[4])malloc(sizeof(REAL) * 4 * NGLL2), (*gb)[4] = (REAL(*)[4])malloc(sizeof(REAL) * 4 * NGLL2);
  for (int n = N; n >= 1; n--) {
    const REAL *dn = Dn + n2 * (n - 1), *vpn = Vp + n2 * (n - 1), *rn = Rn + n2 * (n - 1);
    /* seismogram s_n = R d_n: D += R^T dt (s_n - o_n) */
    for (int r = 0; r < s->nrec; r++) {
      double px = (double)deltat * (seis[((size_t)r * 2 + 0) * N + n - 1] - obs[((size_t)r * 2 + 0) * N + n - 1]);
      double pz = (double)deltat * (seis[((size_t)r * 2 + 1) * N + n - 1] - obs[((size_t)r * 2 + 1) * N + n - 1]);
      double cx = s->cosrot * px - s->sinrot * pz, cz = s->sinrot * px + s->cosrot * pz;
      int e = s->rec_ispec[r] - 1;
      for (int j = 0; j < NGLL; j++)
        for (int i = 0; i < NGLL; i++) {
          int g = s->ibool[e * NGLL2 + j * NGLL + i] - 1;
          double hl = (double)(s->xir[r * NGLL + i] * s->gammar[r * NGLL + j]);
          D[2 * g] = (REAL)((double)D[2 * g] + hl * cx);
          D[2 * g + 1] = (REAL)((double)D[2 * g + 1] + hl * cz);
        }
    }
    /* v_n = v' + dt/2 a_n */
    for (size_t k = 0; k < n2; k++) {
      Vb[k] = V[k];
      A[k] = A[k] + deltatover2 * V[k];
    }
    /* a_n = W r_n (W = rmass_inverse): r-bar = W A; d chi / d M = -A r W^2 */
    for (size_t k = 0; k < n2; k++) {
      Rb[k] = s->rmass_inv[k] * A[k];
      pm[k] -= (double)A[k] * (double)rn[k] * (double)s->rmass_inv[k] * (double)s->rmass_inv[k];
    }
    /* r_n = -K d_n - B v'_n + f_n */
    element_forces(s, Rb, D);
    if (s->p.stacey && s->nabs > 0) stacey_forces(s, Rb, Vb, NULL, 0, 0);
    for (int e = 0; e < s->nspec; e++) {
      point_gradients(s, dn, e, gd);
      point_gradients(s, Rb, e, gb);
      for (int p = 0; p < NGLL2; p++) {
        size_t l = (size_t)e * NGLL2 + p;
        int i = p % NGLL, j = p / NGLL;
        double wj = (double)s->wxgll[i] * (double)s->wxgll[j] * (double)s->jacobian[l];
        double ux = gd[p][0], uxz = gd[p][1], uzx = gd[p][2], uz = gd[p][3];
        double wx = gb[p][0], wxz = gb[p][1], wzx = gb[p][2], wz = gb[p][3];
        gk[l] -= wj * (ux + uz) * (wx + wz);
        gm[l] -= wj * (ux * wx + uz * wz - ux * wz - uz * wx + (uxz + uzx) * (wxz + wzx));
      }
    }
    for (int ab = 0; ab < s->nabs; ab++) {
      int e = s->abs_ispec[ab] - 1, code = s->abs_code[ab];
      for (int k = 0; k < NGLL; k++) {
        int i, j;
        if (code == 4) { i = 0; j = k; }
        else if (code == 2) { i = NGLL - 1; j = k; }
        else if (code == 1) { i = k; j = 0; }
        else { i = k; j = NGLL - 1; }
        if (code == 1 && ((s->abs_corner[4 * ab + 0] && i == 0) || (s->abs_corner[4 * ab + 1] && i == NGLL - 1))) continue;
        if (code == 3 && ((s->abs_corner[4 * ab + 2] && i == 0) || (s->abs_corner[4 * ab + 3] && i == NGLL - 1))) continue;
        size_t l = (size_t)e * NGLL2 + j * NGLL + i;
        int g = s->ibool[l] - 1;
        REAL nx, nz, j1d;
        edge_normal(s, code, l, &nx, &nz, &j1d);
        double weight = (double)j1d * (double)((code == 4 || code == 2) ? s->wxgll[j] : s->wxgll[i]);
        double vx = vpn[2 * g], vz = vpn[2 * g + 1], bx = Rb[2 * g], bz = Rb[2 * g + 1];
        double vn = nx * vx + nz * vz, bn = nx * bx + nz * bz;
        gvp[l] -= weight * vn * bn;
        gvs[l] -= weight * (vx * bx + vz * bz - vn * bn);
      }
    }
    /* d_n = d + dt v + dt^2/2 a,  v' = v + dt/2 a */
    for (size_t k = 0; k < n2; k++) {
      REAL vb = Vb[k], db = D[k];
      V[k] = vb + deltat * db;
      A[k] = deltatover2 * vb + deltatsqover2 * db;
    }
  }
  /* mass matrix M_c(g) = sum w_i w_j J rho + Stacey term dt/2 weight T_c(rho vp, rho vs) (setup_mass above) */
  for (int e = 0; e < s->nspec; e++)
    for (int p = 0; p < NGLL2; p++) {
      size_t l = (size_t)e * NGLL2 + p;
      int i = p % NGLL, j = p / NGLL, g = s->ibool[l] - 1;
      double wj = (double)s->wxgll[i] * (double)s->wxgll[j] * (double)s->jacobian[l];
      gr[l] += wj * (pm[2 * g] + pm[2 * g + 1]);
    }
  if (s->p.stacey)
    for (int ab = 0; ab < s->nabs; ab++) {
      int e = s->abs_ispec[ab] - 1, code = s->abs_code[ab];
      for (int k = 0; k < NGLL; k++) {
        int i, j;
        if (code == 4) { i = 0; j = k; }
        else if (code == 2) { i = NGLL - 1; j = k; }
        else if (code == 1) { i = k; j = 0; }
        else { i = k; j = NGLL - 1; }
        if (code == 1 && ((s->abs_corner[4 * ab + 0] && i == 0) || (s->abs_corner[4 * ab + 1] && i == NGLL - 1))) continue;
        if (code == 3 && ((s->abs_corner[4 * ab + 2] && i == 0) || (s->abs_corner[4 * ab + 3] && i == NGLL - 1))) continue;
        size_t l = (size_t)e * NGLL2 + j * NGLL + i;
        int g = s->ibool[l] - 1;
        REAL nx, nz, j1d;
        edge_normal(s, code, l, &nx, &nz, &j1d);
        double weight = (double)j1d * (double)((code == 4 || code == 2) ? s->wxgll[j] : s->wxgll[i]);
        double h2 = 0.5 * (double)deltat, sn = (double)nx + (double)nz;
        /* T_x = dt/2 [rho_vp (nx+nz) nx + rho_vs (1 - (nx+nz) nx)], T_z likewise with nz */
        gvp[l] += weight * h2 * (pm[2 * g] * sn * nx + pm[2 * g + 1] * sn * nz);
        gvs[l] += weight * h2 * (pm[2 * g] * (1.0 - sn * nx) + pm[2 * g + 1] * (1.0 - sn * nz));
      }
    }
  /* chain to (rho, vp, vs): mu = rho vs^2, kappa = rho vp^2 - mu, rho_vp = rho vp, rho_vs = rho vs */
  for (size_t l = 0; l < ne; l++) {
    double rho = s->rhostore[l], vp = (double)s->rho_vp[l] / rho, vs = (double)s->rho_vs[l] / rho;
    grad[l] = gk[l] * (vp * vp - vs * vs) + gm[l] * vs * vs + gr[l] + gvp[l] * vp + gvs[l] * vs;
    grad[ne + l] = gk[l] * 2.0 * rho * vp + gvp[l] * rho;
    grad[2 * ne + l] = -gk[l] * 2.0 * rho * vs + gm[l] * 2.0 * rho * vs + gvs[l] * rho;
  }
  free(Dn); free(Vp); free(Rn); free(seis); free(d); free(v); free(a);
  free(D); free(V); free(A); free(Rb); free(Vb); free(pm); free(gd); free(gb);
  _mm_setcsr(csr);
  return chi;
}
