// Host driver: builds the solver's setup for a structured internal mesh and replays the GPU branch of the
// SPECFEM2D time loop through the Fortran-style C-ABI (specfem2d_gpu.h).  See specfem2d_driver.h.
//
// Structured meshes let several setup steps use direct formulas where the reference searches:
//   * global numbering: GLL point (i,j) of element (ex,ez) sits on node (4ex+i, 4ez+j) of the global GLL lattice,
//     then the first-touch renumbering of src/specfem2D/get_global.f90:34-72 gives the reference's ibool;
//   * point location: the closest-GLL-point search of locate_receivers.F90:148-200 only visits the lattice
//     neighbourhood of the target, with the reference's tie-break (first element, then j, then i).
// Everything handed to the GPU has the reference's types, layouts and 1-based indices.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <map>
#include <type_traits>
#include <vector>

#include "specfem2d_driver.h"
#include "specfem2d_gpu.h"
#include "specfem2d_gpu_ext.h"

namespace {

constexpr int NGLL = 5;
constexpr int NGLL2 = 25;
constexpr double PI = 3.141592653589793;

void hc(hipError_t e) {
  if (e != hipSuccess) {
    fprintf(stderr, "specfem2d driver: HIP error %s\n", hipGetErrorString(e));
    exit(EXIT_FAILURE);
  }
}

// ---------------------------------------------------------------------------
// GLL basis (src/specfem2D/define_derivation_matrices.f90:34-82, src/shared/gll_library.f90)
double legendre(int n, double x) {
  double p0 = 1.0, p1 = x;
  if (n == 0) return p0;
  for (int k = 1; k < n; k++) {
    double p2 = ((2.0 * k + 1.0) * x * p1 - k * p0) / (k + 1.0);
    p0 = p1;
    p1 = p2;
  }
  return p1;
}

struct Basis {
  double xi[NGLL];
  float w[NGLL];
  float hprime_f[NGLL * NGLL];     // Fortran hprime_xx(i,k) at [(k)*5+(i)] = h'_k(xi_i)
  float hprimewgll_f[NGLL * NGLL]; // Fortran hprimewgll_xx(k,i) at [(i)*5+(k)] = w_k h'_i(xi_k)
};

Basis make_basis() {
  Basis b;
  const int N = NGLL - 1;
  const double r = std::sqrt(3.0 / 7.0);
  const double x[NGLL] = {-1.0, -r, 0.0, r, 1.0};
  for (int i = 0; i < NGLL; i++) {
    b.xi[i] = x[i];
    double pn = legendre(N, x[i]);
    b.w[i] = (float)(2.0 / (N * (N + 1.0) * pn * pn));
  }
  // GLL differentiation matrix D(i,k) = h'_k(xi_i): P_N(x_i)/(P_N(x_k)(x_i-x_k)) off the diagonal,
  // -/+ N(N+1)/4 at the two end points, 0 elsewhere on the diagonal
  for (int i = 0; i < NGLL; i++)
    for (int k = 0; k < NGLL; k++) {
      double d;
      if (i == k)
        d = (i == 0) ? -N * (N + 1.0) / 4.0 : (i == N ? N * (N + 1.0) / 4.0 : 0.0);
      else
        d = legendre(N, x[i]) / (legendre(N, x[k]) * (x[i] - x[k]));
      b.hprime_f[k * NGLL + i] = (float)d;
    }
  for (int i = 0; i < NGLL; i++)
    for (int k = 0; k < NGLL; k++) b.hprimewgll_f[i * NGLL + k] = b.w[k] * b.hprime_f[i * NGLL + k];
  return b;
}

// Lagrange interpolants at xi, and their derivatives when hp is given (src/shared/lagrange_poly.f90:34-93)
void lagrange(double xi, const double* xg, double* h, double* hp = nullptr) {
  for (int d = 0; d < NGLL; d++) {
    double num = 1.0, den = 1.0;
    for (int i = 0; i < NGLL; i++)
      if (i != d) {
        num *= (xi - xg[i]);
        den *= (xg[d] - xg[i]);
      }
    const double inv = 1.0 / den;
    h[d] = num * inv;
    if (!hp) continue;
    double sum = 0.0;
    for (int i = 0; i < NGLL; i++) {
      if (i == d) continue;
      double prod = 1.0;
      for (int j = 0; j < NGLL; j++)
        if (j != d && j != i) prod *= (xi - xg[j]);
      sum += prod;
    }
    hp[d] = sum * inv;
  }
}

// ---------------------------------------------------------------------------
struct Sim {
  sf2d_params p{};
  int nspec = 0, nglob = 0;
  // viscoelastic attenuation: convolution coefficients per point and solid, [nspec][25][3] (solid fastest)
  int attenuation = 0;
  std::vector<float> att_A_kappa, att_B_kappa, att_A_mu, att_B_mu;
  // the Zener parameters in the CPU path's layout (the viscoelastic CPML layer): 1/tau_sigma, tau_epsilon, phi
  // [3][nspec][25] (solid slowest) and Mu [nspec][25] per mode (1 bulk, 2 shear); the quality factors per point
  std::vector<float> zen_itau1, zen_te1, zen_phi1, zen_mu1, zen_itau2, zen_te2, zen_phi2, zen_mu2, zen_qk, zen_qm;
  int simulation_type = 1, save_forward = 0;
  Basis basis{};
  // control nodes of the structured grid
  std::vector<double> gx, gz;  // [(nz+1)][(nx+1)]
  // an external mesh (read_external_mesh): 4-node elements (0-based, counter-clockwise) over control nodes, and the
  // free-surface and absorbing sides of its files as (0-based element, side code 1 bottom 2 right 3 top 4 left)
  bool external = false;
  int npgeo = 0;
  std::vector<int> knods;
  std::vector<double> coorg;
  std::vector<int> ext_free_elem, ext_free_code, ext_abs_elem, ext_abs_code;
  std::vector<int> ibool;      // 1-based [nspec][25]
  std::vector<double> coord;   // [nglob][2]
  std::vector<float> xix, xiz, gammax, gammaz, jacobian, kappa, mu, rho_vp, rho_vs, rho;
  std::vector<float> rmass_inverse;  // [nglob][2]
  std::vector<float> rmassx, rmassz;
  // fluid elements (vs = 0 at all 25 points) and their boundaries
  std::vector<int> is_acoustic;      // [nspec]
  std::vector<int> aniso_elem;       // [nspec] anisotropic elements (empty: none)
  int chain_request = 0;             // elements per chain patch for the device plan (0: the library's default)
  std::vector<float> aniso_c[6];     // c11, c13, c15, c33, c35, c55 [nspec][25]
  int nspec_ac = 0;
  std::vector<float> rmass_ac;       // [nglob] inverse fluid mass
  std::vector<int> free_ispec, free_ij;            // acoustic free-surface edges: 1-based element, [nf][5][2] (i, j)
  std::vector<int> coup_ac_ispec, coup_ac_code;    // fluid-solid edges: fluid element (1-based) and its side
  std::vector<int> coup_el_ispec, coup_el_code;
  std::vector<int> coup_ij;                        // [n][5][2] 1-based (i, j) in the fluid element
  std::vector<float> coup_normal, coup_jacw;       // [n][5][2], [n][5]
  // Stacey (prepare_timerun.F90:1270-1460)
  int nabs = 0;
  int nspec_side[5] = {0, 0, 0, 0, 0};
  std::vector<int> abs_ispec, abs_ij, edge_abs, ib_bottom, ib_left, ib_right, ib_top, ib0;
  std::vector<float> abs_normal, abs_jacw;
  // CPML layer (pml_init.F90): region per element (0, 1 X only, 2 Z only, 3 XZ), row of each layer element in the
  // profile arrays (1-based, 0 outside), profiles K, d, alpha along x and z [nspec_pml][25], the outer-boundary
  // points of elastic and acoustic layer elements (1-based, increasing)
  int nspec_pml = 0;
  std::vector<int> region_cpml, spec_to_pml;
  std::vector<double> pml_kx, pml_kz, pml_dx, pml_dz, pml_ax, pml_az;
  std::vector<int> pml_pts_el, pml_pts_ac;
  // sources
  int nsrc = 0;
  std::vector<sf2d_source> src;
  std::vector<int> src_ispec;
  std::vector<double> src_xi, src_gamma;
  std::vector<float> sourcearrays, stf;  // [nsrc][5][5][2], [nstep][nsrc]
  double t0 = 0.0;
  // receivers
  int nrec = 0;
  std::vector<int> rec_ispec;
  std::vector<double> rec_xi, rec_gamma;
  std::vector<float> xir, gammar;  // [5][nrec] receiver fastest
  std::vector<float> cosrot, sinrot;
  int seismotype[1] = {1};
  // phase lists
  std::vector<int> phase_ispec;  // [2][nspec]
  long mesh = 0;
  int it_done = 0;                 // time steps already run
  // model in the MODEL=gll parameterisation, double (gradient chain rule)
  std::vector<double> model_rho, model_vp, model_vs;
  // gradient: state saves inside the forward loop (0 none, 1 checkpoints every grad_chunk steps, 2 every step
  // of the segment that starts after step grad_seg_base)
  int grad_save = 0, grad_chunk = 0, grad_seg_base = 0, grad_ready = 0;
  std::vector<float> grad_res;     // [nstep][nrec][2] dt (s - o)
  std::vector<double> sisux, sisuz;  // host seismograms (nlength = NSTEP), as the Fortran caller keeps them
  // shot batch: nshot > 0 runs every shot's sources in its own wavefield (src_shot[k] = shot of source k)
  int nshot = 0;
  std::vector<int> src_shot, shot_nsrc;
  std::vector<double> shot_t0;
  std::vector<double> shot_chi;
  bool bad_source = false;
};

// bilinear map of element e (src/shared/define_shape_functions.f90 NGNOD=4, recompute_jacobian.f90)
void map_point(const Sim& s, int e, double xi, double ga, double* x, double* z, double* dxdxi, double* dzdxi,
               double* dxdga, double* dzdga) {
  int nodes[4];
  if (s.external) {
    for (int k = 0; k < 4; k++) nodes[k] = s.knods[(size_t)e * 4 + k];
  } else {
    const int nx = s.p.nx;
    const int ex = e % nx, ez = e / nx;
    nodes[0] = ez * (nx + 1) + ex;
    nodes[1] = nodes[0] + 1;
    nodes[2] = nodes[0] + (nx + 1) + 1;
    nodes[3] = nodes[0] + (nx + 1);
  }
  const std::vector<double>& cx = s.external ? s.coorg : s.gx;
  const double sp = xi + 1.0, sm = xi - 1.0, tp = ga + 1.0, tm = ga - 1.0;
  const double sh[4] = {0.25 * sm * tm, -0.25 * sp * tm, 0.25 * sp * tp, -0.25 * sm * tp};
  const double dxi[4] = {0.25 * tm, -0.25 * tm, 0.25 * tp, -0.25 * tp};
  const double dga[4] = {0.25 * sm, -0.25 * sp, 0.25 * sp, -0.25 * sm};
  double X = 0, Z = 0, a = 0, b = 0, c = 0, d = 0;
  for (int k = 0; k < 4; k++) {
    const double xe = s.external ? cx[2 * (size_t)nodes[k]] : s.gx[nodes[k]];
    const double ze = s.external ? cx[2 * (size_t)nodes[k] + 1] : s.gz[nodes[k]];
    X += sh[k] * xe;
    Z += sh[k] * ze;
    a += dxi[k] * xe;
    b += dxi[k] * ze;
    c += dga[k] * xe;
    d += dga[k] * ze;
  }
  *x = X;
  *z = Z;
  *dxdxi = a;
  *dzdxi = b;
  *dxdga = c;
  *dzdga = d;
}

// numbering of an external mesh: the GLL points fall into classes of equal coordinates (to 1e-10 of the mesh extent:
// sorted by x, runs of close x sorted by z), then numbered by first touch in element order, i fastest
// (createnum_fast.f90, get_global.f90:34-72)
void number_by_coordinates(Sim& s) {
  const size_t ntot = (size_t)s.nspec * NGLL2;
  std::vector<double> xp(2 * ntot);
  double xmin = 1e300, xmax = -1e300, zmin = 1e300, zmax = -1e300;
  for (int e = 0; e < s.nspec; e++)
    for (int j = 0; j < NGLL; j++)
      for (int i = 0; i < NGLL; i++) {
        double x, z, a, b, c, d;
        map_point(s, e, s.basis.xi[i], s.basis.xi[j], &x, &z, &a, &b, &c, &d);
        const size_t l = (size_t)e * NGLL2 + j * NGLL + i;
        xp[2 * l] = x;
        xp[2 * l + 1] = z;
        xmin = std::min(xmin, x); xmax = std::max(xmax, x);
        zmin = std::min(zmin, z); zmax = std::max(zmax, z);
      }
  const double tol = 1e-10 * std::max(xmax - xmin, zmax - zmin);
  std::vector<int> idx(ntot), cls(ntot);
  for (size_t l = 0; l < ntot; l++) idx[l] = (int)l;
  std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return xp[2 * (size_t)a] < xp[2 * (size_t)b]; });
  int ncls = 0;
  for (size_t a = 0; a < ntot;) {
    size_t b = a + 1;
    while (b < ntot && xp[2 * (size_t)idx[b]] - xp[2 * (size_t)idx[b - 1]] <= tol) b++;
    std::stable_sort(idx.begin() + a, idx.begin() + b,
                     [&](int u, int v) { return xp[2 * (size_t)u + 1] < xp[2 * (size_t)v + 1]; });
    for (size_t c = a; c < b;) {
      size_t d = c + 1;
      while (d < b && xp[2 * (size_t)idx[d] + 1] - xp[2 * (size_t)idx[d - 1] + 1] <= tol) d++;
      for (size_t k = c; k < d; k++) cls[idx[k]] = ncls;
      ncls++;
      c = d;
    }
    a = b;
  }
  std::vector<int> mask(ncls, 0);
  s.ibool.resize(ntot);
  int next = 0;
  for (size_t l = 0; l < ntot; l++) {
    if (mask[cls[l]] == 0) mask[cls[l]] = ++next;
    s.ibool[l] = mask[cls[l]];
  }
  s.nglob = next;
}

// geometry at the GLL points (setup_mesh.F90:195-236)
bool build_geometry(Sim& s) {
  const size_t ne = (size_t)s.nspec * NGLL2;
  s.coord.assign((size_t)2 * s.nglob, 0.0);
  s.xix.resize(ne); s.xiz.resize(ne); s.gammax.resize(ne); s.gammaz.resize(ne); s.jacobian.resize(ne);
  for (int e = 0; e < s.nspec; e++)
    for (int j = 0; j < NGLL; j++)
      for (int i = 0; i < NGLL; i++) {
        double x, z, xxi, zxi, xga, zga;
        map_point(s, e, s.basis.xi[i], s.basis.xi[j], &x, &z, &xxi, &zxi, &xga, &zga);
        const double jac = xxi * zga - xga * zxi;
        if (!(jac > 0.0)) return false;
        const size_t l = (size_t)e * NGLL2 + j * NGLL + i;
        s.xix[l] = (float)(zga / jac);
        s.gammax[l] = (float)(-zxi / jac);
        s.xiz[l] = (float)(-xga / jac);
        s.gammaz[l] = (float)(xxi / jac);
        s.jacobian[l] = (float)jac;
        const int g = s.ibool[l] - 1;
        s.coord[2 * g] = x;
        s.coord[2 * g + 1] = z;
      }
  return true;
}

bool build_mesh(Sim& s) {
  const sf2d_params& p = s.p;
  if (s.external) {
    if (s.nspec < 1 || s.npgeo < 4) return false;
    for (int n : s.knods)
      if (n < 0 || n >= s.npgeo) return false;
    number_by_coordinates(s);
    return build_geometry(s);
  }
  if (p.nx < 1 || p.nz < 1 || p.nlayers < 1 || p.nlayers > SF2D_MAXLAYERS) return false;
  int nzsum = 0;
  for (int l = 0; l < p.nlayers; l++) nzsum += p.nz_layer[l];
  if (nzsum != p.nz) return false;
  s.nspec = p.nx * p.nz;
  s.gx.assign((size_t)(p.nx + 1) * (p.nz + 1), 0.0);
  s.gz.assign(s.gx.size(), 0.0);
  // node grid (src/meshfem2D/read_mesh_files.F90:169-338, flat interfaces)
  int row0 = 0;
  for (int l = 0; l < p.nlayers; l++) {
    const double bot = p.zinterface[l], top = p.zinterface[l + 1];
    for (int iz = 0; iz <= p.nz_layer[l]; iz++) {
      const double g = (double)iz / (double)p.nz_layer[l];
      for (int ix = 0; ix <= p.nx; ix++) {
        const size_t n = (size_t)(row0 + iz) * (p.nx + 1) + ix;
        s.gx[n] = p.xmin + (p.xmax - p.xmin) * (double)ix / (double)p.nx;
        s.gz[n] = (1.0 - g) * bot + g * top;
      }
    }
    row0 += p.nz_layer[l];
  }
  // numbering: lattice index, then first-touch renumbering in element order
  const long lx = 4L * p.nx + 1, lz = 4L * p.nz + 1;
  std::vector<int> mask((size_t)(lx * lz), 0);
  s.ibool.resize((size_t)s.nspec * NGLL2);
  int next = 0;
  for (int e = 0; e < s.nspec; e++) {
    const int ex = e % p.nx, ez = e / p.nx;
    for (int j = 0; j < NGLL; j++)
      for (int i = 0; i < NGLL; i++) {
        const size_t L = (size_t)(4L * ez + j) * lx + (4L * ex + i);
        if (mask[L] == 0) mask[L] = ++next;
        s.ibool[(size_t)e * NGLL2 + j * NGLL + i] = mask[L];
      }
  }
  s.nglob = next;
  return build_geometry(s);
}

// materials (read_materials.f90:138-141, setup_mesh.F90:680-792)
void build_materials(Sim& s, const double* rho, const double* vp, const double* vs) {
  const size_t ne = (size_t)s.nspec * NGLL2;
  s.kappa.resize(ne); s.mu.resize(ne); s.rho_vp.resize(ne); s.rho_vs.resize(ne); s.rho.resize(ne);
  for (size_t l = 0; l < ne; l++) {
    const double r = rho[l];
    double v = vp[l];
    const double mu = r * vs[l] * vs[l];
    double kap;
    if (s.p.material_mode == 0) {
      const double lambda = r * v * v - 2.0 * mu;
      kap = lambda + mu;
      v = std::sqrt((kap + mu) / r);
    } else {
      kap = r * v * v - mu;
    }
    s.rho[l] = (float)r;
    s.mu[l] = (float)mu;
    s.kappa[l] = (float)kap;
    s.rho_vp[l] = (float)(r * v);
    s.rho_vs[l] = (float)(r * std::sqrt(mu / r));
  }
}

// fluid elements, the acoustic free surface (every non-absorbing outer side of a fluid element,
// src/meshfem2D/determine_acoustic_surface.f90:34-160) and the fluid-solid edges (fluid elements in order, then their
// solid neighbours below, left, right, above: src/meshfem2D/repartition_coupling.f90:38-106), with the points,
// normals and weights of src/specfem2D/prepare_gpu.f90:545-660 (ivalue order: get_coupling_edges.f90:93-111)
void build_domains(Sim& s) {
  const int nx = s.p.nx, nz = s.p.nz;
  s.is_acoustic.assign(s.nspec, 0);
  s.nspec_ac = 0;
  for (int e = 0; e < s.nspec; e++) {
    int nzero = 0;
    for (int p = 0; p < NGLL2; p++) nzero += s.mu[(size_t)e * NGLL2 + p] == 0.f;
    s.is_acoustic[e] = nzero == NGLL2;
    s.nspec_ac += s.is_acoustic[e];
  }
  s.free_ispec.clear(); s.free_ij.clear();
  s.coup_ac_ispec.clear(); s.coup_ac_code.clear(); s.coup_el_ispec.clear(); s.coup_el_code.clear();
  s.coup_ij.clear(); s.coup_normal.clear(); s.coup_jacw.clear();
  if (s.nspec_ac == 0) return;
  // (i, j) of point k along side code in ivalue order (1 bottom, 2 right, 3 top, 4 left)
  auto ivalue = [](int code, int k, int* i, int* j) {
    switch (code) {
      case 1: *i = k; *j = 0; break;
      case 2: *i = NGLL - 1; *j = k; break;
      case 3: *i = NGLL - 1 - k; *j = NGLL - 1; break;
      default: *i = 0; *j = NGLL - 1 - k; break;
    }
  };
  auto add_free = [&](int e, int code) {
    if (!s.is_acoustic[e]) return;
    s.free_ispec.push_back(e + 1);
    for (int k = 0; k < NGLL; k++) {
      int i, j;
      ivalue(code, k, &i, &j);
      s.free_ij.push_back(i + 1);
      s.free_ij.push_back(j + 1);
    }
  };
  auto add_coupling = [&](int e, int code, int f, int fcode) {
    s.coup_ac_ispec.push_back(e + 1);
    s.coup_ac_code.push_back(code);
    s.coup_el_ispec.push_back(f + 1);
    s.coup_el_code.push_back(fcode);
    for (int k = 0; k < NGLL; k++) {
      int i, j;
      ivalue(code, k, &i, &j);
      const size_t l = (size_t)e * NGLL2 + j * NGLL + i;
      float nxv, nzv, j1, w;
      if (code == 1 || code == 3) {
        const float xx = +s.gammaz[l] * s.jacobian[l];
        const float zx = -s.gammax[l] * s.jacobian[l];
        j1 = std::sqrt(xx * xx + zx * zx);
        nxv = (code == 3 ? -zx : zx) / j1;
        nzv = (code == 3 ? xx : -xx) / j1;
        w = j1 * s.basis.w[i];
      } else {
        const float xg = -s.xiz[l] * s.jacobian[l];
        const float zg = +s.xix[l] * s.jacobian[l];
        j1 = std::sqrt(xg * xg + zg * zg);
        nxv = (code == 4 ? -zg : zg) / j1;
        nzv = (code == 4 ? xg : -xg) / j1;
        w = j1 * s.basis.w[j];
      }
      s.coup_ij.push_back(i + 1);
      s.coup_ij.push_back(j + 1);
      s.coup_normal.push_back(nxv);
      s.coup_normal.push_back(nzv);
      s.coup_jacw.push_back(w);
    }
  };
  if (s.external) {
    // the free-surface file's sides of fluid elements (read_external_mesh_files.F90:398-416); the fluid-solid
    // edges: each side of a fluid element (in element order, sides 1..4) shared with a solid element, found by
    // the side's sorted pair of control nodes
    for (size_t k = 0; k < s.ext_free_elem.size(); k++) add_free(s.ext_free_elem[k], s.ext_free_code[k]);
    if (s.nspec_ac == s.nspec) return;
    std::vector<std::pair<long, int>> sides((size_t)4 * s.nspec);
    for (int e = 0; e < s.nspec; e++)
      for (int c = 1; c <= 4; c++) {
        const long a = s.knods[(size_t)e * 4 + c - 1], b = s.knods[(size_t)e * 4 + c % 4];
        sides[(size_t)e * 4 + c - 1] = {std::min(a, b) * (long)s.npgeo + std::max(a, b), e * 4 + c - 1};
      }
    std::sort(sides.begin(), sides.end());
    for (int e = 0; e < s.nspec; e++) {
      if (!s.is_acoustic[e]) continue;
      for (int c = 1; c <= 4; c++) {
        const long a = s.knods[(size_t)e * 4 + c - 1], b = s.knods[(size_t)e * 4 + c % 4];
        const long key = std::min(a, b) * (long)s.npgeo + std::max(a, b);
        for (auto it = std::lower_bound(sides.begin(), sides.end(), std::make_pair(key, -1));
             it != sides.end() && it->first == key; ++it) {
          const int f = it->second / 4;
          if (f == e || s.is_acoustic[f]) continue;
          add_coupling(e, c, f, it->second % 4 + 1);
        }
      }
    }
    return;
  }
  if (!s.p.absorbtop) for (int i = 0; i < nx; i++) add_free((nz - 1) * nx + i, 3);
  if (!s.p.absorbbottom) for (int i = 0; i < nx; i++) add_free(i, 1);
  if (!s.p.absorbleft) for (int j = 0; j < nz; j++) add_free(j * nx, 4);
  if (!s.p.absorbright) for (int j = 0; j < nz; j++) add_free(j * nx + nx - 1, 2);
  for (int e = 0; e < s.nspec; e++) {
    if (!s.is_acoustic[e]) continue;
    const int ix = e % nx, iz = e / nx;
    const int nb[4] = {iz > 0 ? e - nx : -1, ix > 0 ? e - 1 : -1, ix < nx - 1 ? e + 1 : -1, iz < nz - 1 ? e + nx : -1};
    const int ac_code[4] = {1, 4, 2, 3}, el_code[4] = {3, 2, 4, 1};
    for (int q = 0; q < 4; q++) {
      if (nb[q] < 0 || s.is_acoustic[nb[q]]) continue;
      add_coupling(e, ac_code[q], nb[q], el_code[q]);
    }
  }
}

// absorbing edges in the mesher's order (determine_abs_surface.f90) and their GPU arrays
void build_stacey(Sim& s) {
  const sf2d_params& p = s.p;
  s.nabs = 0;
  if (!p.stacey && !p.pml) return;  // the CPML layer grows from the absorbing edges too
  std::vector<int> el, code;
  for (size_t k = 0; s.external && k < s.ext_abs_elem.size(); k++) {  // the absorbing file's sides, in its order
    el.push_back(s.ext_abs_elem[k] + 1);
    code.push_back(s.ext_abs_code[k]);
  }
  for (int iz = 1; iz <= p.nz && !s.external; iz++)
    for (int ix = 1; ix <= p.nx; ix++) {
      const int e = (iz - 1) * p.nx + ix;
      if (p.absorbbottom && iz == 1) { el.push_back(e); code.push_back(1); }
      if (p.absorbright && ix == p.nx) { el.push_back(e); code.push_back(2); }
      if (p.absorbtop && iz == p.nz) { el.push_back(e); code.push_back(3); }
      if (p.absorbleft && ix == 1) { el.push_back(e); code.push_back(4); }
    }
  const int nf = (int)el.size();
  s.nabs = nf;
  s.abs_ispec = el;
  s.edge_abs = code;
  s.abs_ij.assign((size_t)nf * NGLL * 2, 0);
  s.abs_normal.assign((size_t)nf * NGLL * 2, 0.f);
  s.abs_jacw.assign((size_t)nf * NGLL, 0.f);
  s.ib_bottom.assign(nf, 0); s.ib_left.assign(nf, 0); s.ib_right.assign(nf, 0); s.ib_top.assign(nf, 0);
  s.ib0.assign(nf, 0);
  for (int a = 0; a < 5; a++) s.nspec_side[a] = 0;
  for (int f = 0; f < nf; f++) {
    const int c = code[f];
    const int k = ++s.nspec_side[c];
    (c == 1 ? s.ib_bottom : c == 2 ? s.ib_right : c == 3 ? s.ib_top : s.ib_left)[f] = k;
    s.ib0[f] = k;
    // corner points shared with a left/right edge of the same element (read_mesh_databases.F90:957-1030)
    bool cl = false, cr = false;
    if (c == 1 || c == 3)
      for (int g = 0; g < nf; g++)
        if (g != f && el[g] == el[f]) {
          if (code[g] == 4) cl = true;
          if (code[g] == 2) cr = true;
        }
    const int e = el[f] - 1;
    for (int k2 = 0; k2 < NGLL; k2++) {
      int i, j;
      if (c == 4) { i = 0; j = k2; }
      else if (c == 2) { i = NGLL - 1; j = k2; }
      else if (c == 1) { i = k2; j = 0; }
      else { i = k2; j = NGLL - 1; }
      const size_t l = (size_t)e * NGLL2 + j * NGLL + i;
      float nx, nz, j1;
      if (c == 4 || c == 2) {
        const float xg = -s.xiz[l] * s.jacobian[l];
        const float zg = +s.xix[l] * s.jacobian[l];
        j1 = std::sqrt(xg * xg + zg * zg);
        nx = (c == 4 ? -zg : zg) / j1;
        nz = (c == 4 ? xg : -xg) / j1;
        s.abs_jacw[(size_t)f * NGLL + k2] = j1 * s.basis.w[j];
      } else {
        const float xx = +s.gammaz[l] * s.jacobian[l];
        const float zx = -s.gammax[l] * s.jacobian[l];
        j1 = std::sqrt(xx * xx + zx * zx);
        nx = (c == 1 ? zx : -zx) / j1;
        nz = (c == 1 ? -xx : xx) / j1;
        s.abs_jacw[(size_t)f * NGLL + k2] = j1 * s.basis.w[i];
      }
      s.abs_normal[((size_t)f * NGLL + k2) * 2] = nx;
      s.abs_normal[((size_t)f * NGLL + k2) * 2 + 1] = nz;
      int ii = i + 1, jj = j + 1;
      if ((c == 1 || c == 3) && ((cl && i == 0) || (cr && i == NGLL - 1))) ii = NGLL + 1;
      // a fluid element's absorbing edge leaves out the point it shares with one of its fluid-solid edges
      // (get_coupling_edges.f90:528-563, prepare_timerun.F90:1360-1451)
      for (size_t q = 0; q < s.coup_ac_ispec.size(); q++) {
        if (s.coup_ac_ispec[q] != el[f]) continue;
        const int ce = s.coup_ac_code[q];
        if ((c == 4 || c == 2) && ((ce == 1 && j == 0) || (ce == 3 && j == NGLL - 1))) jj = NGLL + 1;
        if ((c == 1 || c == 3) && ((ce == 4 && i == 0) || (ce == 2 && i == NGLL - 1))) ii = NGLL + 1;
      }
      s.abs_ij[((size_t)f * NGLL + k2) * 2] = ii;
      s.abs_ij[((size_t)f * NGLL + k2) * 2 + 1] = jj;
    }
  }
}

// CPML layer of the internal mesher (pml_init.F90:34-1366 with PML_PARAMETER_ADJUSTMENT and ROTATE_PML_ACTIVATE false)
void build_pml(Sim& s) {
  const int nspec = s.nspec, nglob = s.nglob;
  s.nspec_pml = 0;
  s.region_cpml.assign(nspec, 0);
  s.spec_to_pml.assign(nspec, 0);
  if (!s.p.pml || s.nabs == 0) return;  // nothing without absorbing edges (prepare_pml.f90:49)
  // element detection (pml_init.F90:67-160), side by side (1 bottom, 2 right, 3 top, 4 left): the side's flag of
  // every absorbing-edge element is that of its edge entry (a corner element listed again for another side keeps
  // the flag of its later entry); then, NELEM_PML_THICKNESS - 1 times, every element with a corner point among the
  // corners of the side's elements so far joins the side
  std::vector<char> member((size_t)nspec * 5, 0), in_pml(nspec, 0), corner((size_t)nglob, 0);
  const int cl[4] = {0, NGLL - 1, (NGLL - 1) * NGLL, NGLL2 - 1};
  auto ib = [&](int e, int c) { return s.ibool[(size_t)e * NGLL2 + cl[c]] - 1; };
  for (int side = 1; side <= 4; side++) {
    std::fill(corner.begin(), corner.end(), 0);
    for (int f = 0; f < s.nabs; f++) {
      const int e = s.abs_ispec[f] - 1;
      const bool on = s.edge_abs[f] == side;
      member[(size_t)e * 5 + side] = on;
      if (on)
        for (int c = 0; c < 4; c++) corner[ib(e, c)] = 1;
    }
    for (int layer = 2; layer <= s.p.nelem_pml; layer++) {
      for (int e = 0; e < nspec; e++) {
        if (member[(size_t)e * 5 + side]) continue;
        for (int c = 0; c < 4; c++)
          if (corner[ib(e, c)]) member[(size_t)e * 5 + side] = 1;
      }
      std::fill(corner.begin(), corner.end(), 0);
      for (int e = 0; e < nspec; e++) {
        if (!member[(size_t)e * 5 + side]) continue;
        in_pml[e] = 1;
        for (int c = 0; c < 4; c++) corner[ib(e, c)] = 1;
      }
    }
  }
  // regions (pml_init.F90:201-248): one side alone, or two adjacent sides; anything else stays 0
  for (int e = 0; e < nspec; e++) {
    if (!in_pml[e]) continue;
    const bool B = member[(size_t)e * 5 + 1], R = member[(size_t)e * 5 + 2], T = member[(size_t)e * 5 + 3],
               L = member[(size_t)e * 5 + 4];
    const int nx_sides = (int)L + (int)R, nz_sides = (int)T + (int)B;
    int r = 0;
    if (nx_sides == 1 && nz_sides == 0) r = 1;
    else if (nx_sides == 0 && nz_sides == 1) r = 2;
    else if (nx_sides == 1 && nz_sides == 1) r = 3;
    s.region_cpml[e] = r;
  }
  for (int e = 0; e < nspec; e++)
    if (in_pml[e]) s.spec_to_pml[e] = ++s.nspec_pml;
  if (s.nspec_pml == 0) return;

  // damping profiles (define_PML_coefficients, pml_init.F90:592-1012): thickness per side from the extent of the
  // layer's points on that side of the mesh centre, d0 = -(N + 1) vp_max ln(R) / (2 L) per side and medium,
  // d = d0 / damping_change_factor a^N, K = K_min + (K_max - 1) a^N, alpha = pi f0_max (1 - a), a the normalised
  // distance into the layer (N = 2, R = 0.001)
  const double NPOWER = 2.0, RCOEF = 0.001, PI = 3.141592653589793;
  double xmin = 1e300, xmax = -1e300, zmin = 1e300, zmax = -1e300;
  for (int g = 0; g < nglob; g++) {
    xmin = std::min(xmin, s.coord[2 * g]);
    xmax = std::max(xmax, s.coord[2 * g]);
    zmin = std::min(zmin, s.coord[2 * g + 1]);
    zmax = std::max(zmax, s.coord[2 * g + 1]);
  }
  const double xo = xmin + (xmax - xmin) / 2.0, zo = zmin + (zmax - zmin) / 2.0;
  // [left, right, bottom, top] extents of the layer's points
  double lo[4] = {1e30, 1e30, 1e30, 1e30}, hi[4] = {-1e30, -1e30, -1e30, -1e30};
  for (int e = 0; e < nspec; e++) {
    if (!in_pml[e]) continue;
    const int r = s.region_cpml[e];
    const bool zr = r == 2 || r == 3, xr = r == 1 || r == 3;
    for (int p = 0; p < NGLL2; p++) {
      const int g = s.ibool[(size_t)e * NGLL2 + p] - 1;
      const double x = s.coord[2 * g], z = s.coord[2 * g + 1];
      auto ext = [&](int k, double v) { lo[k] = std::min(lo[k], v); hi[k] = std::max(hi[k], v); };
      if (z < zo && zr) ext(2, z);
      if (x > xo && xr) ext(1, x);
      if (z > zo && zr) ext(3, z);
      if (x < xo && xr) ext(0, x);
    }
  }
  double th[4];
  for (int k = 0; k < 4; k++) {
    th[k] = hi[k] - lo[k];
    if (std::fabs(th[k]) > 1e30) th[k] = 0.0;
  }
  const double x_left = th[0] + xmin, x_right = xmax - th[1], z_bottom = th[2] + zmin, z_top = zmax - th[3];
  double vpmax_ac = 0.0, vpmax_el = 0.0;
  for (int e = 0; e < nspec; e++) {
    if (!in_pml[e]) continue;
    const bool fluid = s.nspec_ac > 0 && s.is_acoustic[e];
    for (int p = 0; p < NGLL2; p++) {
      const size_t l = (size_t)e * NGLL2 + p;
      const double vp = (double)(s.rho_vp[l] / s.rho[l]);
      (fluid ? vpmax_ac : vpmax_el) = std::max(fluid ? vpmax_ac : vpmax_el, vp);
    }
  }
  double f0max = -1e300;
  for (const sf2d_source& src : s.src) f0max = std::max(f0max, src.f0);
  const double alpha_max = PI * f0max;
  double d0[4][2];  // [side][fluid]
  for (int k = 0; k < 4; k++)
    for (int m = 0; m < 2; m++)
      d0[k][m] = th[k] > 0.0 ? -(NPOWER + 1) * (m ? vpmax_ac : vpmax_el) * std::log(RCOEF) / (2.0 * th[k]) : 0.0;
  const size_t np = (size_t)s.nspec_pml * NGLL2;
  s.pml_kx.assign(np, 1.0); s.pml_kz.assign(np, 1.0);
  s.pml_dx.assign(np, 0.0); s.pml_dz.assign(np, 0.0);
  s.pml_ax.assign(np, 0.0); s.pml_az.assign(np, 0.0);
  for (int e = 0; e < nspec; e++) {
    if (!in_pml[e]) continue;
    const int r = s.region_cpml[e];
    const bool fluid = s.nspec_ac > 0 && s.is_acoustic[e];
    const double damp = fluid ? s.p.damping_ac : s.p.damping_el;
    for (int p = 0; p < NGLL2; p++) {
      const int g = s.ibool[(size_t)e * NGLL2 + p] - 1;
      const double x = s.coord[2 * g], z = s.coord[2 * g + 1];
      double dx = 0.0, dz = 0.0, kx = 1.0, kz = 1.0, ax = 0.0, az = 0.0;
      // the sides in the reference's order: bottom, top, right, left
      const int order[4] = {2, 3, 1, 0};
      for (int side : order) {
        const bool zside = side >= 2;
        const bool applies = side == 2 ? z < zo : side == 3 ? z > zo : side == 1 ? x > xo : x < xo;
        if (!applies || (zside ? !(r == 2 || r == 3) : !(r == 1 || r == 3))) continue;
        const double a = side == 2 ? z_bottom - z : side == 3 ? z - z_top : side == 1 ? x - x_right : x_left - x;
        double dd = 0.0, kk = 1.0, aa = 0.0;
        if (a >= 0.0) {
          const double an = a / th[side];
          dd = d0[side][fluid ? 1 : 0] / damp * (an * an);  // a^NPOWER with NPOWER = 2
          kk = s.p.k_min_pml + (s.p.k_max_pml - 1.0) * (an * an);
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
      const size_t q = (size_t)(s.spec_to_pml[e] - 1) * NGLL2 + p;
      s.pml_kx[q] = kx; s.pml_kz[q] = kz; s.pml_dx[q] = dx; s.pml_dz[q] = dz; s.pml_ax[q] = ax; s.pml_az[q] = az;
    }
  }
  // corner elements keep alpha_x, alpha_z, beta_x, beta_z apart (pml_init.F90:1269-1356): the minimum distance
  // between GLL points along an element's i and j directions sets the separation alpha_max / 8 dmin / max L
  double dmin2 = 1e30;
  for (int e = 0; e < nspec; e++) {
    if (!in_pml[e]) continue;
    for (int j = 0; j < NGLL - 1; j++)
      for (int i = 0; i < NGLL - 1; i++) {
        const int* b = s.ibool.data() + (size_t)e * NGLL2;
        const int g1 = b[j * NGLL + i] - 1;
        for (int g2 : {b[j * NGLL + i + 1] - 1, b[(j + 1) * NGLL + i] - 1}) {
          const double ddx = s.coord[2 * g1] - s.coord[2 * g2], ddz = s.coord[2 * g1 + 1] - s.coord[2 * g2 + 1];
          dmin2 = std::min(dmin2, ddx * ddx + ddz * ddz);
        }
      }
  }
  const double sep = alpha_max / 8.0 * std::sqrt(dmin2) / std::max(std::max(th[2], th[3]), std::max(th[0], th[1]));
  const double sep2 = sep * 2.0;
  for (int e = 0; e < nspec; e++) {
    if (!in_pml[e] || s.region_cpml[e] != 3) continue;
    for (int p = 0; p < NGLL2; p++) {
      const size_t q = (size_t)(s.spec_to_pml[e] - 1) * NGLL2 + p;
      const double kx = s.pml_kx[q], kz = s.pml_kz[q];
      double ax = s.pml_ax[q], az = s.pml_az[q];
      if (std::fabs(ax - az) < sep) {
        if (ax > az) ax = az + sep2;
        else az = ax + sep2;
      }
      double bx = ax + s.pml_dx[q] / kx;
      if (std::fabs(bx - az) < sep) bx = az + sep2;
      double bz = az + s.pml_dz[q] / kz;
      if (std::fabs(bz - ax) < sep) bz = ax + sep2;
      s.pml_dx[q] = (bx - ax) * kx;
      s.pml_dz[q] = (bz - az) * kz;
      s.pml_ax[q] = ax;
      s.pml_az[q] = az;
    }
  }
  // points of the absorbing edges of layer elements, by medium (determine_boundary_abs_points_PML,
  // pml_init.F90:306-421)
  std::vector<char> mel((size_t)nglob, 0), mac((size_t)nglob, 0);
  for (int f = 0; f < s.nabs; f++) {
    const int e = s.abs_ispec[f] - 1, c = s.edge_abs[f];
    if (!in_pml[e]) continue;
    const bool fluid = s.nspec_ac > 0 && s.is_acoustic[e];
    for (int k = 0; k < NGLL; k++) {
      const int i = c == 4 ? 0 : c == 2 ? NGLL - 1 : k, j = c == 1 ? 0 : c == 3 ? NGLL - 1 : k;
      (fluid ? mac : mel)[s.ibool[(size_t)e * NGLL2 + j * NGLL + i] - 1] = 1;
    }
  }
  s.pml_pts_el.clear();
  s.pml_pts_ac.clear();
  for (int g = 0; g < nglob; g++) {
    if (mel[g]) s.pml_pts_el.push_back(g + 1);
    if (mac[g]) s.pml_pts_ac.push_back(g + 1);
  }
}

// mass matrix with the Newmark Stacey term (invert_mass_matrix.F90:34-464, 470-792, 798-848)
void build_mass(Sim& s) {
  std::vector<float>& rm = s.rmass_inverse;
  std::vector<float>& rma = s.rmass_ac;
  rm.assign((size_t)2 * s.nglob, 0.f);
  rma.assign(s.nspec_ac > 0 ? (size_t)s.nglob : 0, 0.f);
  for (int e = 0; e < s.nspec; e++)
    for (int j = 0; j < NGLL; j++)
      for (int i = 0; i < NGLL; i++) {
        const size_t l = (size_t)e * NGLL2 + j * NGLL + i;
        const int g = s.ibool[l] - 1;
        const float wxz = s.basis.w[i] * s.basis.w[j];
        // CPML layer elements: w_i w_j J (rho or 1/kappa) times K_x, K_x K_z or K_z, then the Newmark damping term
        // (d_x, d_x K_z + d_z K_x or d_z) dt/2 (invert_mass_matrix.F90:133-233, 262-345); a layer element of no
        // region adds nothing
        const int r = s.nspec_pml > 0 && s.spec_to_pml[e] > 0 ? s.region_cpml[e] : -1;
        double kterm = 0.0, dterm = 0.0;
        if (r > 0) {
          const size_t q = (size_t)(s.spec_to_pml[e] - 1) * NGLL2 + j * NGLL + i;
          const double kx = s.pml_kx[q], kz = s.pml_kz[q], dx = s.pml_dx[q], dz = s.pml_dz[q];
          const double dto2 = (double)(float)(0.5 * (double)(float)s.p.dt);
          kterm = r == 1 ? kx : r == 3 ? kx * kz : kz;
          dterm = (r == 1 ? dx : r == 3 ? dx * kz + dz * kx : dz) * dto2;
        }
        if (s.nspec_ac > 0 && s.is_acoustic[e]) {
          if (r < 0) {
            // w_i w_j J in single precision over the double kappa (invert_mass_matrix.F90:424-427)
            rma[g] = (float)((double)rma[g] + (double)(wxz * s.jacobian[l]) / (double)s.kappa[l]);
          } else if (r > 0) {
            const double base = (double)wxz / (double)s.kappa[l] * (double)s.jacobian[l];
            rma[g] = (float)((double)rma[g] + base * kterm);
            rma[g] = (float)((double)rma[g] + base * dterm);
          }
          continue;
        }
        if (r < 0) {
          rm[2 * g] = (float)((double)rm[2 * g] + (double)wxz * (double)s.rho[l] * (double)s.jacobian[l]);
        } else if (r > 0) {
          const double base = (double)wxz * (double)s.rho[l] * (double)s.jacobian[l];
          rm[2 * g] = (float)((double)rm[2 * g] + base * kterm);
          rm[2 * g] = (float)((double)rm[2 * g] + base * dterm);
        }
        rm[2 * g + 1] = rm[2 * g];
      }
  if (s.p.stacey) {
    const float dto2 = (float)(0.5 * (double)(float)s.p.dt);
    for (int f = 0; f < s.nabs; f++) {
      const int e = s.abs_ispec[f] - 1;
      const bool fluid = s.nspec_ac > 0 && s.is_acoustic[e];
      for (int k = 0; k < NGLL; k++) {
        const int ii = s.abs_ij[((size_t)f * NGLL + k) * 2], jj = s.abs_ij[((size_t)f * NGLL + k) * 2 + 1];
        if (ii > NGLL || jj > NGLL) continue;
        const size_t l = (size_t)e * NGLL2 + (jj - 1) * NGLL + (ii - 1);
        const int g = s.ibool[l] - 1;
        if (fluid) {  // dt/2 w / (rho vp) (invert_mass_matrix.F90:709-790)
          rma[g] = rma[g] + dto2 * s.abs_jacw[(size_t)f * NGLL + k] / s.rho_vp[l];
          continue;
        }
        const float nx = s.abs_normal[((size_t)f * NGLL + k) * 2], nz = s.abs_normal[((size_t)f * NGLL + k) * 2 + 1];
        const float w = s.abs_jacw[(size_t)f * NGLL + k];
        const float vn = nx * dto2 + nz * dto2;
        const float tx = s.rho_vp[l] * vn * nx + s.rho_vs[l] * (dto2 - vn * nx);
        const float tz = s.rho_vp[l] * vn * nz + s.rho_vs[l] * (dto2 - vn * nz);
        rm[2 * g] = rm[2 * g] + tx * w;
        rm[2 * g + 1] = rm[2 * g + 1] + tz * w;
      }
    }
  }
  for (auto& v : rm) {
    if (v <= 0.f) v = 1.f;
    v = 1.f / v;
  }
  for (auto& v : rma) {
    if (v <= 0.f) v = 1.f;
    v = 1.f / v;
  }
  s.rmassx.resize(s.nglob);
  s.rmassz.resize(s.nglob);
  for (int g = 0; g < s.nglob; g++) {
    s.rmassx[g] = rm[2 * g];
    s.rmassz[g] = rm[2 * g + 1];
  }
}

// point location: closest GLL point (strict <, scan order ispec,j,i), then Newton from the selected element and
// every element sharing one of its corners (locate_receivers.F90:148-340, locate_source.F90)
void locate_newton(const Sim& s, double px, double pz, int isel, int ixg, int izg, std::vector<int>& cand,
                   int* ispec_out, double* xi_out, double* ga_out);

// an external mesh: the nearest GLL point over all elements (first in element order), then the elements sharing a
// corner with its element, in element order (locate_receivers.F90:197-329)
void locate_external(const Sim& s, double px, double pz, int* ispec_out, double* xi_out, double* ga_out) {
  double dmin = std::numeric_limits<double>::infinity();
  int isel = 0, ixg = 0, izg = 0;
  for (int e = 0; e < s.nspec; e++)
    for (int j = 0; j < NGLL; j++)
      for (int i = 0; i < NGLL; i++) {
        const int g = s.ibool[(size_t)e * NGLL2 + j * NGLL + i] - 1;
        const double dx = px - s.coord[2 * (size_t)g], dz = pz - s.coord[2 * (size_t)g + 1];
        const double d2 = dx * dx + dz * dz;
        if (d2 < dmin) { dmin = d2; isel = e; ixg = i; izg = j; }
      }
  const int* ib = &s.ibool[(size_t)isel * NGLL2];
  const int corners[4] = {ib[0], ib[NGLL - 1], ib[NGLL2 - 1], ib[(NGLL - 1) * NGLL]};
  std::vector<int> cand{isel};
  for (int e = 0; e < s.nspec; e++) {
    if (e == isel) continue;
    const int* ie = &s.ibool[(size_t)e * NGLL2];
    const int ec[4] = {ie[0], ie[NGLL - 1], ie[NGLL2 - 1], ie[(NGLL - 1) * NGLL]};
    bool hit = false;
    for (int a = 0; a < 4 && !hit; a++)
      for (int b = 0; b < 4; b++) hit = hit || ec[a] == corners[b];
    if (hit) cand.push_back(e);
  }
  locate_newton(s, px, pz, isel, ixg, izg, cand, ispec_out, xi_out, ga_out);
}

void locate(const Sim& s, double px, double pz, int* ispec_out, double* xi_out, double* ga_out) {
  if (s.external) {
    locate_external(s, px, pz, ispec_out, xi_out, ga_out);
    return;
  }
  const sf2d_params& p = s.p;
  // lattice lines of the GLL points: x along columns, z along rows
  const long lx = 4L * p.nx + 1, lz = 4L * p.nz + 1;
  auto lattice_x = [&](long c) {
    const long ex = std::min<long>(c / 4, p.nx - 1);
    const int i = (int)(c - 4 * ex);
    const int e = (int)ex;
    return s.coord[2 * (size_t)(s.ibool[(size_t)e * NGLL2 + i] - 1)];
  };
  auto lattice_z = [&](long r) {
    const long ez = std::min<long>(r / 4, p.nz - 1);
    const int j = (int)(r - 4 * ez);
    const int e = (int)(ez * p.nx);
    return s.coord[2 * (size_t)(s.ibool[(size_t)e * NGLL2 + j * NGLL] - 1) + 1];
  };
  // nearest lattice column/row by bisection, then a +-2 neighbourhood
  auto nearest = [](long n, double t, auto f) {
    long lo = 0, hi = n - 1;
    while (hi - lo > 1) {
      long mid = (lo + hi) / 2;
      if (f(mid) <= t) lo = mid; else hi = mid;
    }
    return lo;
  };
  const long c0 = nearest(lx, px, lattice_x), r0 = nearest(lz, pz, lattice_z);
  double dmin = std::numeric_limits<double>::infinity();
  long best_key = std::numeric_limits<long>::max();
  int isel = 0, ixg = 0, izg = 0;
  for (long r = std::max(0L, r0 - 2); r <= std::min(lz - 1, r0 + 3); r++)
    for (long c = std::max(0L, c0 - 2); c <= std::min(lx - 1, c0 + 3); c++) {
      // every element holding lattice point (c, r), first in scan order wins ties
      for (long ez = std::max(0L, (r - 1) / 4); ez <= std::min<long>(p.nz - 1, r / 4); ez++)
        for (long ex = std::max(0L, (c - 1) / 4); ex <= std::min<long>(p.nx - 1, c / 4); ex++) {
          const int i = (int)(c - 4 * ex), j = (int)(r - 4 * ez);
          if (i < 0 || i >= NGLL || j < 0 || j >= NGLL) continue;
          const int e = (int)(ez * p.nx + ex);
          const int g = s.ibool[(size_t)e * NGLL2 + j * NGLL + i] - 1;
          const double dx = px - s.coord[2 * (size_t)g], dz = pz - s.coord[2 * (size_t)g + 1];
          const double d2 = dx * dx + dz * dz;
          const long key = ((long)e * NGLL + j) * NGLL + i;
          if (d2 < dmin || (d2 == dmin && key < best_key)) {
            dmin = d2;
            best_key = key;
            isel = e;
            ixg = i;
            izg = j;
          }
        }
    }
  // candidates: the selected element first, then corner neighbours in element order
  std::vector<int> cand{isel};
  {
    const int ex0 = isel % p.nx, ez0 = isel / p.nx;
    for (int ez = std::max(0, ez0 - 1); ez <= std::min(p.nz - 1, ez0 + 1); ez++)
      for (int ex = std::max(0, ex0 - 1); ex <= std::min(p.nx - 1, ex0 + 1); ex++) {
        const int e = ez * p.nx + ex;
        if (e != isel) cand.push_back(e);
      }
    std::sort(cand.begin() + 1, cand.end());
  }
  locate_newton(s, px, pz, isel, ixg, izg, cand, ispec_out, xi_out, ga_out);
}

// Newton iterations from the nearest GLL point in the first candidate and from the element centre in the others;
// the closest result wins
void locate_newton(const Sim& s, double px, double pz, int isel, int ixg, int izg, std::vector<int>& cand,
                   int* ispec_out, double* xi_out, double* ga_out) {
  (void)isel;
  double best = std::numeric_limits<double>::infinity();
  for (size_t c = 0; c < cand.size(); c++) {
    if (c > 0) { ixg = 1; izg = 1; }  // int(NGLLX/2.0) = 2 in the reference's 1-based indexing
    const int e = cand[c];
    double xi = s.basis.xi[ixg], ga = s.basis.xi[izg];
    double x, z, a, b, cc, d;
    for (int it = 0; it < 5; it++) {
      map_point(s, e, xi, ga, &x, &z, &a, &b, &cc, &d);
      const double jac = a * d - cc * b;
      const double xixl = d / jac, gaxl = -b / jac, xizl = -cc / jac, gazl = a / jac;
      const double dx = -(x - px), dz = -(z - pz);
      xi = xi + (xixl * dx + xizl * dz);
      ga = ga + (gaxl * dx + gazl * dz);
      xi = std::min(1.01, std::max(-1.01, xi));
      ga = std::min(1.01, std::max(-1.01, ga));
    }
    map_point(s, e, xi, ga, &x, &z, &a, &b, &cc, &d);
    const double dist = std::sqrt((px - x) * (px - x) + (pz - z) * (pz - z));
    if (dist < best) {
      best = dist;
      *ispec_out = e + 1;
      *xi_out = xi;
      *ga_out = ga;
    }
  }
}

// sources (setup_sources_receivers.F90:979-1040, set_source_parameters.f90:116-250,
// prepare_source_time_function.f90, comp_source_time_function.f90)
// sources k0 .. k0+n-1 as one simulation's source set: their common t0 and time shifts (setup_sources_receivers.F90,
// compute_source_time_function), source arrays and time functions into the handle's [nsrc] layouts
double build_source_set(Sim& s, int k0, int n) {
  std::vector<double> t0s(n), tsh(n);
  for (int k = 0; k < n; k++) {
    tsh[k] = s.src[k0 + k].tshift;
    const double hdur = 1.0 / s.src[k0 + k].f0;
    t0s[k] = 1.20 * hdur + tsh[k];
  }
  double t0 = 0.0, min_tshift = 0.0;
  if (n == 1) {
    t0 = t0s[0];
    min_tshift = tsh[0];
    tsh[0] = 0.0;
  } else if (n > 1) {
    t0 = *std::min_element(t0s.begin(), t0s.end());
    min_tshift = *std::min_element(tsh.begin(), tsh.end());
    for (int k = 0; k < n; k++) tsh[k] = t0s[k] - t0;
  }
  if (s.p.user_t0 > 0.0 && t0 <= s.p.user_t0 + min_tshift) {
    t0 = s.p.user_t0;
    for (int k = 0; k < n; k++) tsh[k] = t0s[k] - 1.20 / s.src[k0 + k].f0;
  }
  const int ns = s.nsrc;
  for (int q = 0; q < n; q++) {
    const int k = k0 + q;
    locate(s, s.src[k].x, s.src[k].z, &s.src_ispec[k], &s.src_xi[k], &s.src_gamma[k]);
    double hx[NGLL], hg[NGLL];
    lagrange(s.src_xi[k], s.basis.xi, hx);
    lagrange(s.src_gamma[k], s.basis.xi, hg);
    const double ang = s.src[k].anglesource * PI / 180.0;
    const bool fluid = s.nspec_ac > 0 && s.is_acoustic[s.src_ispec[k] - 1];
    if (s.src[k].source_type == 2 && fluid) {
      fprintf(stderr, "sf2d_sim_create: source %d: a moment-tensor source cannot sit in a fluid element\n", k + 1);
      s.bad_source = true;
    } else if (s.src[k].source_type == 2) {
      // moment tensor (compute_arrays_source.f90:34-160): gradient of the Lagrange weights at the source through
      // the element's inverse jacobian interpolated there
      double hpx[NGLL], hpg[NGLL];
      lagrange(s.src_xi[k], s.basis.xi, hx, hpx);
      lagrange(s.src_gamma[k], s.basis.xi, hg, hpg);
      const size_t e0 = (size_t)(s.src_ispec[k] - 1) * NGLL2;
      double dxi_dx = 0.0, dxi_dz = 0.0, dga_dx = 0.0, dga_dz = 0.0;
      for (int m = 0; m < NGLL; m++)
        for (int i = 0; i < NGLL; i++) {
          const size_t l = e0 + m * NGLL + i;
          const double hl = hx[i] * hg[m];
          dxi_dx += hl * (double)s.xix[l];
          dxi_dz += hl * (double)s.xiz[l];
          dga_dx += hl * (double)s.gammax[l];
          dga_dz += hl * (double)s.gammaz[l];
        }
      const double mxx = s.src[k].Mxx, mzz = s.src[k].Mzz, mxz = s.src[k].Mxz;
      for (int m = 0; m < NGLL; m++)
        for (int i = 0; i < NGLL; i++) {
          const double dsx = (hpx[i] * dxi_dx) * hg[m] + hx[i] * (hpg[m] * dga_dx);
          const double dsz = (hpx[i] * dxi_dz) * hg[m] + hx[i] * (hpg[m] * dga_dz);
          s.sourcearrays[(((size_t)k * NGLL + m) * NGLL + i) * 2] = (float)(mxx * dsx + mxz * dsz);
          s.sourcearrays[(((size_t)k * NGLL + m) * NGLL + i) * 2 + 1] = (float)(mxz * dsx + mzz * dsz);
        }
    } else
    for (int j = 0; j < NGLL; j++)
      for (int i = 0; i < NGLL; i++) {
        // a pressure source in a fluid element, a force in a solid one (setup_sources_receivers.F90:994-1010)
        const double h = hx[i] * hg[j];
        s.sourcearrays[(((size_t)k * NGLL + j) * NGLL + i) * 2] = (float)(fluid ? h : std::sin(ang) * h);
        s.sourcearrays[(((size_t)k * NGLL + j) * NGLL + i) * 2 + 1] = (float)(fluid ? h : -std::cos(ang) * h);
      }
    // time functions (comp_source_time_function.f90:52-200); USE_TRICK_FOR_BETTER_PRESSURE replaces each by its
    // second time derivative (prepare_source_time_function.f90:200-262)
    const double a = (PI * PI) * (s.src[k].f0 * s.src[k].f0);
    const int type = s.src[k].time_function_type;
    const bool trick = s.p.use_trick != 0;
    for (int it = 0; it < s.p.nstep; it++) {
      const double t = (double)it * s.p.dt - t0 - tsh[q];
      const double t2 = t * t;
      double v;
      if (type == 1 && !trick)
        v = (1.0 - 2.0 * a * t2) * std::exp(-a * t2);
      else if (type == 1)
        v = -(2.0 * a * (3.0 - 12.0 * a * t * t + 4.0 * (a * a) * t * t * t * t) * std::exp(-a * t * t));
      else if (type == 2 && !trick)
        v = t * std::exp(-a * t2);
      else if (type == 2)
        v = 2.0 * a * t * std::exp(-a * t2) * (-3.0 + 2.0 * a * t2);
      else if (!trick)
        v = -std::exp(-a * t2) / (2.0 * a);
      else
        v = (1.0 - 2.0 * a * t2) * std::exp(-a * t2);
      s.stf[(size_t)it * ns + k] = (float)(-s.src[k].factor * v);
    }
  }
  return t0;
}

// one source set for the simulation, or one per shot of a batch (each shot timed as its own simulation)
void build_sources(Sim& s) {
  const int n = s.nsrc;
  s.src_ispec.assign(n, 0);
  s.src_xi.assign(n, 0.0);
  s.src_gamma.assign(n, 0.0);
  s.sourcearrays.assign((size_t)n * NGLL2 * 2, 0.f);
  s.stf.assign((size_t)n * s.p.nstep, 0.f);
  if (s.nshot == 0) {
    s.t0 = build_source_set(s, 0, n);
    return;
  }
  s.shot_t0.assign(s.nshot, 0.0);
  s.src_shot.assign(n, 0);
  for (int sh = 0, k0 = 0; sh < s.nshot; k0 += s.shot_nsrc[sh], sh++) {
    s.shot_t0[sh] = build_source_set(s, k0, s.shot_nsrc[sh]);
    for (int q = 0; q < s.shot_nsrc[sh]; q++) s.src_shot[k0 + q] = sh;
  }
  s.t0 = s.shot_t0[0];
}

void build_receivers(Sim& s, const double* rec_xz) {
  const int n = s.nrec;
  s.rec_ispec.assign(n, 0);
  s.rec_xi.assign(n, 0.0);
  s.rec_gamma.assign(n, 0.0);
  s.xir.assign((size_t)NGLL * std::max(n, 1), 0.f);
  s.gammar.assign((size_t)NGLL * std::max(n, 1), 0.f);
  s.cosrot.assign(std::max(n, 1), 1.f);
  s.sinrot.assign(std::max(n, 1), 0.f);
  const double ang = s.p.anglerec * PI / 180.0;
  for (int r = 0; r < n; r++) {
    locate(s, rec_xz[2 * r], rec_xz[2 * r + 1], &s.rec_ispec[r], &s.rec_xi[r], &s.rec_gamma[r]);
    double hx[NGLL], hg[NGLL];
    lagrange(s.rec_xi[r], s.basis.xi, hx);
    lagrange(s.rec_gamma[r], s.basis.xi, hg);
    for (int i = 0; i < NGLL; i++) {
      s.xir[(size_t)i * n + r] = (float)hx[i];
      s.gammar[(size_t)i * n + r] = (float)hg[i];
    }
    s.cosrot[r] = (float)std::cos(ang);
    s.sinrot[r] = (float)std::sin(ang);
  }
}

// device preparation, in the order of prepare_gpu.f90:80-347
void prepare_device(Sim& s) {
  int myrank = 0, ndev = 0;
  initialize_cuda_device_(&myrank, &ndev);
  int ngllx = NGLL, nspec = s.nspec, nglob = s.nglob, zero = 0;
  int stacey = s.p.stacey ? 1 : 0, pml = s.nspec_pml > 0 ? 1 : 0;
  std::vector<int> inner(s.nspec, 1), is_elastic(s.nspec, 1), is_acoustic(s.nspec, 0);
  for (int e = 0; e < s.nspec && s.nspec_ac > 0; e++) {
    is_acoustic[e] = s.is_acoustic[e];
    is_elastic[e] = !s.is_acoustic[e];
  }
  int nsrc = s.nsrc, nstep = s.p.nstep, nrec = s.nrec;
  int simtype = s.simulation_type, p_sv = 1, nspec_ac = s.nspec_ac, nspec_el = s.nspec - s.nspec_ac;
  int save_fwd = s.save_forward;
  s.seismotype[0] = s.p.seismotype;
  int nsig = 1, nlength = s.p.nstep;
  int dummy_i = 0;
  std::vector<int> src_ispec = s.src_ispec, rec_ispec = s.rec_ispec;
  if (src_ispec.empty()) src_ispec.push_back(1);
  if (rec_ispec.empty()) rec_ispec.push_back(1);
  std::vector<float> sa = s.sourcearrays, stf = s.stf;
  if (sa.empty()) sa.assign(2 * NGLL2, 0.f);
  if (stf.empty()) stf.assign(1, 0.f);
  prepare_constants_device_(&s.mesh, &ngllx, &nspec, &nglob, s.xix.data(), s.xiz.data(), s.gammax.data(),
                            s.gammaz.data(), s.kappa.data(), s.mu.data(), s.ibool.data(), &zero, &zero, &dummy_i,
                            &dummy_i, s.basis.hprime_f, s.basis.hprimewgll_f, s.basis.w, &stacey, &pml, inner.data(),
                            &nsrc, sa.data(), stf.data(), &nstep, src_ispec.data(), rec_ispec.data(), &nrec,
                            s.cosrot.data(), s.sinrot.data(), &simtype, &p_sv, &nspec_ac, &nspec_el,
                            is_acoustic.data(), is_elastic.data(), &myrank, &save_fwd, s.xir.data(), s.gammar.data(),
                            &nsig, s.seismotype, &nlength);
  if (s.chain_request > 0) sf2d_mesh_set_chain_elems(&s.mesh, s.chain_request);
  // single partition: every element is "inner" (iphase 2); the outer lists are empty
  float dummy_f = 0.f;
  if (s.nspec_ac > 0) {
    int num_phase = s.nspec_ac, nfree = (int)s.free_ispec.size(), ncoup = (int)s.coup_ac_ispec.size();
    int any_el = nspec_el > 0 ? 1 : 0;
    std::vector<int> phase((size_t)2 * num_phase, 0);
    for (int e = 0, q = 0; e < s.nspec; e++)
      if (s.is_acoustic[e]) phase[(size_t)num_phase + q++] = e + 1;
    std::vector<int> fi = s.free_ispec, fij = s.free_ij, ci = s.coup_ac_ispec, cij = s.coup_ij;
    std::vector<float> cn = s.coup_normal, cw = s.coup_jacw;
    if (fi.empty()) { fi.push_back(1); fij.assign(2 * NGLL, 1); }
    if (ci.empty()) { ci.push_back(1); cij.assign(2 * NGLL, 1); cn.assign(2 * NGLL, 0.f); cw.assign(NGLL, 0.f); }
    prepare_fields_acoustic_device_(&s.mesh, s.rmass_ac.data(), s.rho.data(), s.kappa.data(), &num_phase,
                                    phase.data(), &nfree, fi.data(), fij.data(), &any_el, &ncoup, ci.data(),
                                    cij.data(), cn.data(), cw.data(), &zero, &dummy_i, &zero, &dummy_f, &dummy_f,
                                    &zero, &dummy_f);
    if (s.simulation_type == 3) prepare_fields_acoustic_adj_dev_(&s.mesh, &zero, &zero, &zero);
  }
  if (nspec_el > 0) {
    int num_phase = nspec_el;
    s.phase_ispec.assign((size_t)2 * num_phase, 0);
    for (int e = 0, q = 0; e < s.nspec; e++)
      if (is_elastic[e]) s.phase_ispec[(size_t)num_phase + q++] = e + 1;
    std::vector<int> aniso = s.aniso_elem.empty() ? std::vector<int>(s.nspec, 0) : s.aniso_elem;
    int att = s.attenuation, any_aniso = s.aniso_elem.empty() ? 0 : 1;
    std::vector<float> c0((size_t)s.nspec * NGLL2, 0.f);  // c12, c23, c25: AXISYM and SH only
    float* c[6];
    for (int k = 0; k < 6; k++) c[k] = any_aniso ? s.aniso_c[k].data() : &dummy_f;
    // Fortran argument order c11, c12, c13, c15, c23, c25, c33, c35, c55 (prepare_gpu.f90:189-199)
    prepare_fields_elastic_device_(&s.mesh, s.rmassx.data(), s.rmassz.data(), &num_phase, s.phase_ispec.data(),
                                   aniso.data(), &any_aniso, c[0], any_aniso ? c0.data() : &dummy_f, c[1], c[2],
                                   any_aniso ? c0.data() : &dummy_f, any_aniso ? c0.data() : &dummy_f, c[3], c[4],
                                   c[5], &zero, &dummy_i, &att,
                                   att ? s.att_A_mu.data() : &dummy_f, att ? s.att_B_mu.data() : &dummy_f,
                                   att ? s.att_A_kappa.data() : &dummy_f, att ? s.att_B_kappa.data() : &dummy_f);
  }
  if (s.simulation_type == 3 && nspec_el > 0) {
    int size = 2 * s.nglob;
    prepare_fields_elastic_adj_dev_(&s.mesh, &size, &zero, &zero, &zero);
  }
  if (s.p.stacey) {
    int nf = s.nabs, ac_sim = s.nspec_ac > 0 ? 1 : 0, el_sim = nspec_el > 0 ? 1 : 0;
    prepare_stacey_device_(&s.mesh, &ac_sim, &el_sim, s.rho_vp.data(), s.rho_vs.data(), &s.nspec_side[1], &s.nspec_side[4],
                           &s.nspec_side[2], &s.nspec_side[3], s.abs_ispec.data(), s.abs_ij.data(),
                           s.abs_normal.data(), s.abs_jacw.data(), &nf, s.edge_abs.data(), s.ib_bottom.data(),
                           s.ib_left.data(), s.ib_right.data(), s.ib_top.data());
  }
  if (s.nspec_pml > 0) {
    // the layer in the CPU path's terms (prepare_gpu.f90:215-226 hands the GPU the acoustic subset only)
    int npml = s.nspec_pml, nel = (int)s.pml_pts_el.size(), nac_pts = (int)s.pml_pts_ac.size();
    double dt = s.p.dt;
    std::vector<int> pe = s.pml_pts_el, pa = s.pml_pts_ac;
    if (pe.empty()) pe.push_back(1);
    if (pa.empty()) pa.push_back(1);
    prepare_cpml_device_(&s.mesh, &npml, s.region_cpml.data(), s.spec_to_pml.data(), s.pml_kx.data(), s.pml_kz.data(),
                         s.pml_dx.data(), s.pml_dz.data(), s.pml_ax.data(), s.pml_az.data(), s.rho.data(), &dt, &nel,
                         pe.data(), &nac_pts, pa.data());
    if (s.attenuation) {
      int nsls = 3;
      prepare_cpml_attenuation_device_(&s.mesh, &nsls, s.zen_itau1.data(), s.zen_phi1.data(), s.zen_te1.data(),
                                       s.zen_mu1.data(), s.zen_itau2.data(), s.zen_phi2.data(), s.zen_te2.data(),
                                       s.zen_mu2.data(), s.zen_qk.data(), s.zen_qm.data());
    }
  }
  std::vector<float> zeros((size_t)2 * s.nglob, 0.f);
  int size = 2 * s.nglob;
  if (s.nspec_ac > 0) {
    int nac = s.nglob;
    transfer_fields_ac_to_device_(&nac, zeros.data(), zeros.data(), zeros.data(), &s.mesh);
    if (s.simulation_type == 3) transfer_b_fields_ac_to_device_(&nac, zeros.data(), zeros.data(), zeros.data(), &s.mesh);
  }
  if (nspec_el > 0) {
    transfer_fields_el_to_device_(&size, zeros.data(), zeros.data(), zeros.data(), &s.mesh);
    if (s.simulation_type == 3) transfer_b_fields_to_device_(&size, zeros.data(), zeros.data(), zeros.data(), &s.mesh);
  }
  if (s.nshot > 0) sf2d_batch_prepare(&s.mesh, s.nshot, s.src_shot.data());
}

// Relaxation times of n standard linear solids fitting 1/Q over [fmin, fmax] Hz by linear least squares, as
// attenuation_model.f90:443-486, 636-799 does with USE_SOLVOPT false: relaxation angular frequencies and 2n-1
// sample frequencies log-spaced over the band (exponents formed in single precision, as remplit_point and
// classical_linear_least_squares form them), basis w (p - w/Q) / (p^2 + w^2) per solid, the normal equations
// solved here by a Cholesky factorisation.  Returns false when the fit is not positive definite.
bool zener_relaxation(int n, double q, double fmin, double fmax, double* tau_eps, double* tau_sig) {
  const double two_pi = 2.0 * 3.141592653589793;
  double p[3], m[3][3] = {}, r[3] = {};
  for (int l = 0; l < n; l++) p[l] = two_pi * std::pow(fmax / fmin, (double)((float)l / (float)(n - 1))) * fmin;
  const int nk = 2 * n - 1;
  for (int k = 0; k < nk; k++) {
    const double w = two_pi * fmin * std::pow(fmax / fmin, (double)((float)k / (float)(nk - 1)));
    double b[3];
    for (int l = 0; l < n; l++) b[l] = w * (p[l] - w / q) / (p[l] * p[l] + w * w);
    for (int a = 0; a < n; a++) {
      r[a] += b[a] / q;
      for (int c = 0; c <= a; c++) m[a][c] += b[a] * b[c];
    }
  }
  double L[3][3] = {};
  for (int a = 0; a < n; a++)
    for (int c = 0; c <= a; c++) {
      double t = m[a][c];
      for (int k = 0; k < c; k++) t -= L[a][k] * L[c][k];
      if (a == c) {
        if (!(t > 0.0)) return false;
        L[a][a] = std::sqrt(t);
      } else {
        L[a][c] = t / L[c][c];
      }
    }
  double z[3], y[3];
  for (int a = 0; a < n; a++) {
    double t = r[a];
    for (int k = 0; k < a; k++) t -= L[a][k] * z[k];
    z[a] = t / L[a][a];
  }
  for (int a = n - 1; a >= 0; a--) {
    double t = z[a];
    for (int k = a + 1; k < n; k++) t -= L[k][a] * y[k];
    y[a] = t / L[a][a];
  }
  for (int l = 0; l < n; l++) {
    tau_sig[l] = 1.0 / p[l];
    tau_eps[l] = tau_sig[l] * (1.0 + n * y[l]);
    if (tau_sig[l] < 0.0 || tau_eps[l] < 0.0) return false;
  }
  return true;
}

// the single-precision recursive-convolution coefficients of one mode (prepare_attenuation.f90:368-397 after
// attenuation_model.f90:152-159): A = exp(-dt / (2 tau_sigma)), B = phi (1 - A) tau_sigma; and the model's
// single-precision parameters 1/tau_sigma, tau_epsilon, phi and Mu = sum(tau_epsilon / tau_sigma) / 3
// (attenuation_model.f90:145-161), in Z[10] = {itau[3], te[3], phi[3], Mu}
bool newmark_coefficients(double q, double f0, double dt, float* A, float* B, float* Z) {
  double te[3], ts[3];
  if (!zener_relaxation(3, q, f0 / 10.0, f0 * 10.0, te, ts)) return false;
  double ratio = 0.0;
  for (int k = 0; k < 3; k++) ratio += te[k] / ts[k];
  for (int k = 0; k < 3; k++) {
    const float inv_tau = (float)(1.0 / ts[k]);
    const float phi = (float)((1.0 - te[k] / ts[k]) / ts[k] / ratio);
    const float a = (float)std::exp(-0.5 * (double)inv_tau * dt);
    const float c = (float)((1.0 - (double)a) / (double)inv_tau);
    A[k] = a;
    B[k] = phi * c;
    Z[k] = inv_tau;
    Z[3 + k] = (float)te[k];
    Z[6 + k] = phi;
  }
  Z[9] = (float)(ratio / 3.0);
  return true;
}

struct ExternalMeshIn {
  int nspec, npgeo;
  const int* knods;
  const double* coorg;
  int nfree;
  const int *free_elem, *free_code;
  int nabs;
  const int *abs_elem, *abs_code;
};

void* create(const sf2d_params* p, const double* rho, const double* vp, const double* vs, int nsrc,
             const sf2d_source* src, int nrec, const double* rec_xz, int simulation_type, int save_forward, int nshot,
             const int* shot_nsrc, const ExternalMeshIn* ext = nullptr) {
  const int st = p->seismotype;
  if (st != 1 && st != 2 && st != 3 && st != 4 && st != 6) {
    fprintf(stderr, "sf2d_sim_create: seismotype %d is not supported (1, 2, 3, 4 or 6)\n", st);
    return nullptr;
  }
  if (st != 1 && nshot > 0) {
    fprintf(stderr, "sf2d_sim_create: shot batches record displacement (seismotype 1)\n");
    return nullptr;
  }
  if (p->pml && (p->stacey || nshot > 0 || simulation_type != 1 || save_forward || p->nelem_pml < 1 ||
                 !(p->k_min_pml > 0.0) || !(p->damping_ac > 0.0) || !(p->damping_el > 0.0))) {
    fprintf(stderr, "sf2d_sim_create: a CPML layer replaces Stacey edges, in forward runs without shot batches or "
                    "SAVE_FORWARD, with NELEM_PML_THICKNESS >= 1, K_MIN_PML > 0 and positive damping factors\n");
    return nullptr;
  }
  if (ext && p->pml) {
    fprintf(stderr, "sf2d_sim_create_external: CPML layers of external meshes are not supported\n");
    return nullptr;
  }
  Sim* s = new Sim();
  s->p = *p;
  if (ext) {
    s->external = true;
    s->nspec = ext->nspec;
    s->npgeo = ext->npgeo;
    s->knods.assign(ext->knods, ext->knods + (size_t)4 * ext->nspec);
    s->coorg.assign(ext->coorg, ext->coorg + (size_t)2 * ext->npgeo);
    s->ext_free_elem.assign(ext->free_elem, ext->free_elem + ext->nfree);
    s->ext_free_code.assign(ext->free_code, ext->free_code + ext->nfree);
    s->ext_abs_elem.assign(ext->abs_elem, ext->abs_elem + ext->nabs);
    s->ext_abs_code.assign(ext->abs_code, ext->abs_code + ext->nabs);
    for (int k = 0; k < ext->nfree; k++)
      if (ext->free_elem[k] < 0 || ext->free_elem[k] >= ext->nspec || ext->free_code[k] < 1 || ext->free_code[k] > 4)
        s->npgeo = -1;
    for (int k = 0; k < ext->nabs; k++)
      if (ext->abs_elem[k] < 0 || ext->abs_elem[k] >= ext->nspec || ext->abs_code[k] < 1 || ext->abs_code[k] > 4)
        s->npgeo = -1;
  }
  s->simulation_type = simulation_type;
  s->save_forward = save_forward;
  s->basis = make_basis();
  if (!build_mesh(*s)) {
    fprintf(stderr, "sf2d_sim_create: invalid mesh parameters\n");
    delete s;
    return nullptr;
  }
  build_materials(*s, rho, vp, vs);
  build_domains(*s);
  const size_t ne = (size_t)s->nspec * NGLL2;
  s->model_rho.assign(rho, rho + ne);
  s->model_vp.assign(vp, vp + ne);
  s->model_vs.assign(vs, vs + ne);
  s->nsrc = nsrc;
  s->src.assign(src, src + nsrc);  // the CPML profiles use the sources' highest f0
  build_stacey(*s);
  build_pml(*s);
  build_mass(*s);
  s->nshot = nshot;
  if (nshot > 0) s->shot_nsrc.assign(shot_nsrc, shot_nsrc + nshot);
  build_sources(*s);
  if (s->bad_source) {
    delete s;
    return nullptr;
  }
  s->nrec = nrec;
  build_receivers(*s, rec_xz);
  return s;
}

// steps it_begin..it_end of every shot of a batch, with the gradient's state saves; seismograms when record
void batch_steps(Sim* s, int it_begin, int it_end, bool record) {
  const float deltat = (float)s->p.dt;
  const float deltatover2 = (float)(0.5 * (double)deltat);
  const float deltatsqover2 = (float)(0.5 * (double)deltat * (double)deltat);
  sf2d_batch_run(&s->mesh, it_begin, it_end, deltat, deltatsqover2, deltatover2, record && s->nrec > 0,
                 s->grad_save == 1 ? s->grad_chunk : 0);
  s->it_done = it_end;
}

// device seismograms [nshot][2][nrec][nstep] -> [nshot][nrec][2][nstep] doubles
void batch_seismograms(Sim* s, double* out) {
  const int N = s->p.nstep, R = s->nrec;
  std::vector<float> dev((size_t)s->nshot * 2 * std::max(1, R) * N);
  if (R > 0) sf2d_batch_seismograms(&s->mesh, dev.data());
  for (int sh = 0; sh < s->nshot; sh++)
    for (int r = 0; r < R; r++)
      for (int c = 0; c < 2; c++)
        for (int n = 0; n < N; n++)
          out[(((size_t)sh * R + r) * 2 + c) * N + n] = dev[(((size_t)sh * 2 + c) * R + r) * N + n];
}

}  // namespace

extern "C" {

void* sf2d_sim_create(const sf2d_params* p, const double* rho, const double* vp, const double* vs, int nsrc,
                      const sf2d_source* src, int nrec, const double* rec_xz, int simulation_type,
                      int save_forward) {
  return create(p, rho, vp, vs, nsrc, src, nrec, rec_xz, simulation_type, save_forward, 0, nullptr);
}

void* sf2d_sim_create_external(const sf2d_params* p, int nspec, int npgeo, const int* knods, const double* coorg,
                               int nfree, const int* free_elem, const int* free_code, int nabs, const int* abs_elem,
                               const int* abs_code, const double* rho, const double* vp, const double* vs, int nsrc,
                               const sf2d_source* src, int nrec, const double* rec_xz, int simulation_type,
                               int save_forward) {
  if (!knods || !coorg || nspec < 1 || nfree < 0 || nabs < 0) return nullptr;
  const ExternalMeshIn m{nspec, npgeo, knods, coorg, nfree, free_elem, free_code, nabs, abs_elem, abs_code};
  return create(p, rho, vp, vs, nsrc, src, nrec, rec_xz, simulation_type, save_forward, 0, nullptr, &m);
}

// An uneven batch: the disjoint union of n host-only simulations (different meshes, sizes and physics, one time step
// and one recording kind) as one simulation -- their elements, points, boundaries, layers, sources and receivers
// concatenated in instance order with the element, point, layer-row and side offsets applied -- so that every kernel
// launch of a time step covers all instances.  No counterpart in the reference (one mesh per run; its multi-run mode
// is MPI sub-communicators, src/shared/parallel.F90:95-116).  The union runs the longest instance's NSTEP; the
// instances are set up with that NSTEP (their source time functions then cover it) and record their own first steps.
void* sf2d_sim_create_union(int n, void* const* sims) {
  if (n < 1 || !sims) return nullptr;
  const Sim* f = (const Sim*)sims[0];
  for (int k = 0; k < n; k++) {
    const Sim* a = (const Sim*)sims[k];
    if (!a || a->mesh || a->simulation_type != 1 || a->nshot > 0 || a->save_forward || a->p.dt != f->p.dt ||
        a->p.nstep != f->p.nstep || a->p.seismotype != f->p.seismotype || a->p.use_trick != f->p.use_trick) {
      fprintf(stderr, "sf2d_sim_create_union: instances must be host-only forward simulations with one time step, "
                      "NSTEP and recording kind\n");
      return nullptr;
    }
  }
  Sim* u = new Sim();
  u->p = f->p;
  u->basis = f->basis;
  u->simulation_type = 1;
  long ne_tot = 0, ng_tot = 0;
  for (int k = 0; k < n; k++) {
    const Sim* a = (const Sim*)sims[k];
    ne_tot += a->nspec;
    ng_tot += a->nglob;
    u->p.stacey = u->p.stacey || a->p.stacey;
    u->p.pml = u->p.pml || a->p.pml;
    u->attenuation = u->attenuation || a->attenuation;
  }
  if (ne_tot * NGLL2 > (long)INT32_MAX || ng_tot * 2 > (long)INT32_MAX) {
    fprintf(stderr, "sf2d_sim_create_union: %ld elements, %ld points: too many for 32-bit indices\n", ne_tot, ng_tot);
    delete u;
    return nullptr;
  }
  u->nspec = (int)ne_tot;
  u->nglob = (int)ng_tot;
  const size_t ne = (size_t)u->nspec * NGLL2;
  // per-side counters of the absorbing faces
  int side_count[5] = {0, 0, 0, 0, 0};
  int eoff = 0, goff = 0, poff = 0, nrec = 0;
  for (int k = 0; k < n; k++) nrec += ((const Sim*)sims[k])->nrec;
  u->nrec = nrec;
  u->xir.assign((size_t)NGLL * std::max(nrec, 1), 0.f);
  u->gammar.assign(u->xir.size(), 0.f);
  // Zener parameters: only the viscoelastic CPML layer uses them
  const bool zener = u->attenuation && u->p.pml;
  if (zener) {
    u->zen_itau1.assign(3 * ne, -1.f); u->zen_te1.assign(3 * ne, -1.f); u->zen_phi1.assign(3 * ne, 0.f);
    u->zen_itau2.assign(3 * ne, -1.f); u->zen_te2.assign(3 * ne, -1.f); u->zen_phi2.assign(3 * ne, 0.f);
    u->zen_mu1.assign(ne, 1.f); u->zen_mu2.assign(ne, 1.f);
    u->zen_qk.assign(ne, 9999.f); u->zen_qm.assign(ne, 9999.f);
  }
  // reserve the concatenated arrays once: they hold tens of GB for large batches
  auto reserve_sum = [&](auto member) {
    size_t len = 0;
    for (int k = 0; k < n; k++) len += (((const Sim*)sims[k])->*member).size();
    (u->*member).reserve(len);
  };
  reserve_sum(&Sim::ibool); reserve_sum(&Sim::coord); reserve_sum(&Sim::xix); reserve_sum(&Sim::xiz);
  reserve_sum(&Sim::gammax); reserve_sum(&Sim::gammaz); reserve_sum(&Sim::jacobian); reserve_sum(&Sim::kappa);
  reserve_sum(&Sim::mu); reserve_sum(&Sim::rho_vp); reserve_sum(&Sim::rho_vs); reserve_sum(&Sim::rho);
  reserve_sum(&Sim::model_rho); reserve_sum(&Sim::model_vp); reserve_sum(&Sim::model_vs);
  if (u->attenuation)
    for (auto m : {&Sim::att_A_kappa, &Sim::att_B_kappa, &Sim::att_A_mu, &Sim::att_B_mu}) (u->*m).reserve(3 * ne);
  int roff = 0;
  for (int k = 0; k < n; k++) {
    const Sim& a = *(const Sim*)sims[k];
    const size_t ae = (size_t)a.nspec * NGLL2;
    auto cat = [](auto& dst, const auto& src) { dst.insert(dst.end(), src.begin(), src.end()); };
    auto cat_or_zero = [](auto& dst, const auto& src, size_t len) {
      if (src.size() == len) dst.insert(dst.end(), src.begin(), src.end());
      else dst.insert(dst.end(), len, typename std::decay_t<decltype(dst)>::value_type{});
    };
    for (int v : a.ibool) u->ibool.push_back(v + goff);
    cat(u->coord, a.coord);
    cat(u->xix, a.xix); cat(u->xiz, a.xiz); cat(u->gammax, a.gammax); cat(u->gammaz, a.gammaz);
    cat(u->jacobian, a.jacobian); cat(u->kappa, a.kappa); cat(u->mu, a.mu); cat(u->rho_vp, a.rho_vp);
    cat(u->rho_vs, a.rho_vs); cat(u->rho, a.rho);
    cat_or_zero(u->rmass_inverse, a.rmass_inverse, (size_t)2 * a.nglob);
    cat_or_zero(u->rmassx, a.rmassx, (size_t)a.nglob);
    cat_or_zero(u->rmassz, a.rmassz, (size_t)a.nglob);
    cat_or_zero(u->rmass_ac, a.rmass_ac, (size_t)a.nglob);
    cat_or_zero(u->is_acoustic, a.is_acoustic, (size_t)a.nspec);
    u->nspec_ac += a.nspec_ac;
    for (int v : a.free_ispec) u->free_ispec.push_back(v + eoff);
    cat(u->free_ij, a.free_ij);
    for (int v : a.coup_ac_ispec) u->coup_ac_ispec.push_back(v + eoff);
    for (int v : a.coup_el_ispec) u->coup_el_ispec.push_back(v + eoff);
    cat(u->coup_ac_code, a.coup_ac_code); cat(u->coup_el_code, a.coup_el_code); cat(u->coup_ij, a.coup_ij);
    cat(u->coup_normal, a.coup_normal); cat(u->coup_jacw, a.coup_jacw);
    // Stacey faces: element offset, indices within each side continued across the instances (a CPML instance's
    // absorbing sides only seeded its layer, which comes merged below)
    const int nabs_a = a.p.stacey ? a.nabs : 0;
    for (int q = 0; q < nabs_a; q++) {
      const int c = a.edge_abs[q];
      const int kk = ++side_count[c];
      u->abs_ispec.push_back(a.abs_ispec[q] + eoff);
      u->edge_abs.push_back(c);
      u->ib0.push_back(kk);
      u->ib_bottom.push_back(c == 1 ? kk : 0);
      u->ib_right.push_back(c == 2 ? kk : 0);
      u->ib_top.push_back(c == 3 ? kk : 0);
      u->ib_left.push_back(c == 4 ? kk : 0);
    }
    u->nabs += nabs_a;
    if (nabs_a) {
      cat(u->abs_ij, a.abs_ij); cat(u->abs_normal, a.abs_normal); cat(u->abs_jacw, a.abs_jacw);
    }
    // CPML layer: rows of the profile arrays continued, outer-boundary points offset
    cat_or_zero(u->region_cpml, a.region_cpml, (size_t)a.nspec);
    for (int e = 0; e < a.nspec; e++)
      u->spec_to_pml.push_back(a.spec_to_pml.empty() || a.spec_to_pml[e] == 0 ? 0 : a.spec_to_pml[e] + poff);
    cat(u->pml_kx, a.pml_kx); cat(u->pml_kz, a.pml_kz); cat(u->pml_dx, a.pml_dx); cat(u->pml_dz, a.pml_dz);
    cat(u->pml_ax, a.pml_ax); cat(u->pml_az, a.pml_az);
    for (int g : a.pml_pts_el) u->pml_pts_el.push_back(g + goff);
    for (int g : a.pml_pts_ac) u->pml_pts_ac.push_back(g + goff);
    u->nspec_pml += a.nspec_pml;
    // attenuation: coefficients per point ([nspec][25][3]), Zener parameters ([3][nspec][25]), quality factors
    if (u->attenuation) {
      cat_or_zero(u->att_A_kappa, a.att_A_kappa, 3 * ae); cat_or_zero(u->att_B_kappa, a.att_B_kappa, 3 * ae);
      cat_or_zero(u->att_A_mu, a.att_A_mu, 3 * ae); cat_or_zero(u->att_B_mu, a.att_B_mu, 3 * ae);
    }
    if (zener && a.attenuation && !a.zen_itau1.empty()) {
      const size_t o = (size_t)eoff * NGLL2;
      for (int sl = 0; sl < 3; sl++)
        for (size_t l = 0; l < ae; l++) {
          u->zen_itau1[sl * ne + o + l] = a.zen_itau1[sl * ae + l];
          u->zen_te1[sl * ne + o + l] = a.zen_te1[sl * ae + l];
          u->zen_phi1[sl * ne + o + l] = a.zen_phi1[sl * ae + l];
          u->zen_itau2[sl * ne + o + l] = a.zen_itau2[sl * ae + l];
          u->zen_te2[sl * ne + o + l] = a.zen_te2[sl * ae + l];
          u->zen_phi2[sl * ne + o + l] = a.zen_phi2[sl * ae + l];
        }
      std::copy(a.zen_mu1.begin(), a.zen_mu1.end(), u->zen_mu1.begin() + o);
      std::copy(a.zen_mu2.begin(), a.zen_mu2.end(), u->zen_mu2.begin() + o);
      std::copy(a.zen_qk.begin(), a.zen_qk.end(), u->zen_qk.begin() + o);
      std::copy(a.zen_qm.begin(), a.zen_qm.end(), u->zen_qm.begin() + o);
    }
    cat(u->model_rho, a.model_rho); cat(u->model_vp, a.model_vp); cat(u->model_vs, a.model_vs);
    // sources: element offset; their arrays and time functions (each instance timed as a run of its own)
    cat(u->src, a.src);
    for (int v : a.src_ispec) u->src_ispec.push_back(v + eoff);
    cat(u->src_xi, a.src_xi); cat(u->src_gamma, a.src_gamma); cat(u->sourcearrays, a.sourcearrays);
    u->nsrc += a.nsrc;
    // receivers: element offset, interpolation weights [5][nrec] re-laid for the union
    for (int r = 0; r < a.nrec; r++) {
      u->rec_ispec.push_back(a.rec_ispec[r] + eoff);
      u->rec_xi.push_back(a.rec_xi[r]);
      u->rec_gamma.push_back(a.rec_gamma[r]);
      u->cosrot.push_back(a.cosrot[r]);
      u->sinrot.push_back(a.sinrot[r]);
      for (int i = 0; i < NGLL; i++) {
        u->xir[(size_t)i * nrec + roff + r] = a.xir[(size_t)i * a.nrec + r];
        u->gammar[(size_t)i * nrec + roff + r] = a.gammar[(size_t)i * a.nrec + r];
      }
    }
    roff += a.nrec;
    eoff += a.nspec;
    goff += a.nglob;
    poff += a.nspec_pml;
  }
  for (int c = 1; c <= 4; c++) u->nspec_side[c] = side_count[c];
  if (u->is_acoustic.empty()) u->is_acoustic.assign(u->nspec, 0);
  if (u->cosrot.empty()) { u->cosrot.assign(1, 1.f); u->sinrot.assign(1, 0.f); }
  // time functions [nstep][nsrc], source fastest
  const int nstep = u->p.nstep;
  u->stf.assign((size_t)nstep * std::max(u->nsrc, 1), 0.f);
  for (int k = 0, soff = 0; k < n; k++) {
    const Sim& a = *(const Sim*)sims[k];
    for (int t = 0; t < nstep; t++)
      for (int q = 0; q < a.nsrc; q++) u->stf[(size_t)t * u->nsrc + soff + q] = a.stf[(size_t)t * a.nsrc + q];
    soff += a.nsrc;
  }
  u->t0 = f->t0;
  return u;
}

void* sf2d_sim_create_shots(const sf2d_params* p, const double* rho, const double* vp, const double* vs, int nshot,
                            const int* shot_nsrc, const sf2d_source* src, int nrec, const double* rec_xz) {
  if (nshot < 1 || !shot_nsrc) return nullptr;
  int nsrc = 0;
  for (int k = 0; k < nshot; k++) {
    if (shot_nsrc[k] < 1) return nullptr;
    nsrc += shot_nsrc[k];
  }
  return create(p, rho, vp, vs, nsrc, src, nrec, rec_xz, 1, 0, nshot, shot_nsrc);
}

int sf2d_sim_prepare_device(void* sim) {
  Sim* s = (Sim*)sim;
  if (!s) return 1;
  if (s->mesh) return 0;
  prepare_device(*s);
  return 0;
}

void sf2d_sim_destroy(void* sim) {
  Sim* s = (Sim*)sim;
  if (!s) return;
  if (s->mesh) {
    int one = 1, zero = 0;
    float dummy = 0.f;
    prepare_cleanup_device_(&s->mesh, &zero, &one, &zero, &zero, &zero, &zero, &zero, &dummy);
  }
  delete s;
}

void sf2d_sim_pml_counts(void* sim, int* counts) {
  const Sim* s = (const Sim*)sim;
  counts[0] = s->nspec_pml;
  counts[1] = (int)s->pml_pts_el.size();
  counts[2] = (int)s->pml_pts_ac.size();
}

void sf2d_sim_get_pml(void* sim, int* region, double* prof, int* pts_el, int* pts_ac) {
  const Sim* s = (const Sim*)sim;
  std::copy(s->region_cpml.begin(), s->region_cpml.end(), region);
  const size_t np = (size_t)s->nspec_pml * NGLL2;
  const std::vector<double>* src[6] = {&s->pml_kx, &s->pml_kz, &s->pml_dx, &s->pml_dz, &s->pml_ax, &s->pml_az};
  for (int k = 0; k < 6 && np > 0; k++) std::copy(src[k]->begin(), src[k]->end(), prof + k * np);
  std::copy(s->pml_pts_el.begin(), s->pml_pts_el.end(), pts_el);
  std::copy(s->pml_pts_ac.begin(), s->pml_pts_ac.end(), pts_ac);
}

int sf2d_sim_nspec(void* sim) { return ((Sim*)sim)->nspec; }
int sf2d_sim_nglob(void* sim) { return ((Sim*)sim)->nglob; }
int sf2d_sim_nabs(void* sim) { return ((Sim*)sim)->nabs; }
double sf2d_sim_t0(void* sim) { return ((Sim*)sim)->t0; }
long sf2d_sim_mesh_pointer(void* sim) { return ((Sim*)sim)->mesh; }

void sf2d_sim_get_ibool(void* sim, int* out) {
  Sim* s = (Sim*)sim;
  std::copy(s->ibool.begin(), s->ibool.end(), out);
}
void sf2d_sim_get_rmass_inverse(void* sim, float* out) {
  Sim* s = (Sim*)sim;
  std::copy(s->rmass_inverse.begin(), s->rmass_inverse.end(), out);
}
void sf2d_sim_get_elem_array(void* sim, int which, float* out) {
  Sim* s = (Sim*)sim;
  const std::vector<float>* a[10] = {&s->xix, &s->xiz, &s->gammax, &s->gammaz, &s->jacobian,
                                     &s->kappa, &s->mu, &s->rho_vp, &s->rho_vs, &s->rho};
  if (which < 0 || which > 9) return;
  std::copy(a[which]->begin(), a[which]->end(), out);
}
void sf2d_sim_get_sources(void* sim, int* ispec, double* xi, double* gamma, float* sourcearrays, float* stf) {
  Sim* s = (Sim*)sim;
  for (int k = 0; k < s->nsrc; k++) {
    ispec[k] = s->src_ispec[k];
    xi[k] = s->src_xi[k];
    gamma[k] = s->src_gamma[k];
  }
  std::copy(s->sourcearrays.begin(), s->sourcearrays.end(), sourcearrays);
  std::copy(s->stf.begin(), s->stf.end(), stf);
}
void sf2d_sim_get_receivers(void* sim, int* ispec, double* xi, double* gamma, float* xir, float* gammar) {
  Sim* s = (Sim*)sim;
  for (int r = 0; r < s->nrec; r++) {
    ispec[r] = s->rec_ispec[r];
    xi[r] = s->rec_xi[r];
    gamma[r] = s->rec_gamma[r];
    for (int i = 0; i < NGLL; i++) {
      xir[r * NGLL + i] = s->xir[(size_t)i * s->nrec + r];
      gammar[r * NGLL + i] = s->gammar[(size_t)i * s->nrec + r];
    }
  }
}
void sf2d_sim_get_stacey(void* sim, int* ispec, int* ij, float* normal, float* jacw, int* edge, int* ib) {
  Sim* s = (Sim*)sim;
  std::copy(s->abs_ispec.begin(), s->abs_ispec.end(), ispec);
  std::copy(s->abs_ij.begin(), s->abs_ij.end(), ij);
  std::copy(s->abs_normal.begin(), s->abs_normal.end(), normal);
  std::copy(s->abs_jacw.begin(), s->abs_jacw.end(), jacw);
  std::copy(s->edge_abs.begin(), s->edge_abs.end(), edge);
  std::copy(s->ib0.begin(), s->ib0.end(), ib);
}

int sf2d_sim_forward(void* sim, int nstep_run, double* seis, float* absorb, double* timing) {
  Sim* s = (Sim*)sim;
  if (!s || !s->mesh || nstep_run < 1 || s->it_done + nstep_run > s->p.nstep) return 1;
  // Newmark constants (prepare_timerun.F90:262-264)
  float deltat = (float)s->p.dt;
  float deltatover2 = (float)(0.5 * (double)deltat);
  float deltatsqover2 = (float)(0.5 * (double)deltat * (double)deltat);
  float b_deltat = 0.f, b_deltatover2 = 0.f, b_deltatsqover2 = 0.f;
  int nspec_outer = 0, nspec_inner = s->nspec, zero = 0, one = 1, att = s->attenuation;
  const int it_begin = s->it_done + 1;
  int i_sig = 1, nlength = s->p.nstep, it_end = s->it_done + nstep_run;
  const bool any_ac = s->nspec_ac > 0, any_el = s->nspec_ac < s->nspec;
  const bool pml = s->nspec_pml > 0;
  int elastic_sim = any_el ? 1 : 0, acoustic_sim = any_ac ? 1 : 0, trick = s->p.use_trick ? 1 : 0;
  int nspec_inner_ac = s->nspec_ac, nfluid_solid = (int)s->coup_ac_ispec.size();
  int aniso = s->aniso_elem.empty() ? 0 : 1;
  nspec_inner = s->nspec - s->nspec_ac;
  if (any_ac && s->grad_save) return 1;  // no fluid gradients
  if (s->sisux.empty()) {
    s->sisux.assign((size_t)std::max(1, s->nrec) * nlength, 0.0);
    s->sisuz.assign(s->sisux.size(), 0.0);
  }
  const int counts[5] = {0, s->nspec_side[1], s->nspec_side[2], s->nspec_side[3], s->nspec_side[4]};
  std::vector<float> slice[5], slice_ac[5];
  for (int c = 1; c <= 4; c++) slice[c].assign((size_t)2 * NGLL * std::max(1, counts[c]), 0.f);
  for (int c = 1; c <= 4; c++) slice_ac[c].assign((size_t)NGLL * std::max(1, counts[c]), 0.f);
  // timing: events are recorded on the library's stream and read once after the loop
  const bool timed = timing != nullptr;
  const bool phases = timed && timing[0] < 0.0;
  hipStream_t stream = (hipStream_t)sf2d_mesh_stream(&s->mesh);
  const int nev = phases ? 10 : 2;  // per step: forces (2) [+ predictor, boundary, corrector, seismograms]
  std::vector<hipEvent_t> ev(timed ? (size_t)nev * nstep_run : 0);
  for (auto& e : ev) hc(hipEventCreate(&e));
  hipEvent_t loop_start, loop_end;
  if (timed) {
    hc(hipEventCreate(&loop_start));
    hc(hipEventCreate(&loop_end));
    hc(hipEventRecord(loop_start, stream));
  }
  auto mark = [&](int step, int k) {
    if (timed && k < nev) hc(hipEventRecord(ev[(size_t)step * nev + k], stream));
  };
  for (int it = it_begin; it <= it_end; it++) {
    const int step = it - it_begin;
    mark(step, 2);
    // update_displ_Newmark (update_displacement_Newmark.F90:130-300): fluid, then solid
    // compute_b_wavefield = .true. as update_displ_acoustic_forward passes it (update_displacement_Newmark.F90:150-155)
    if (any_ac)
      update_displacement_ac_cuda_(&s->mesh, &deltat, &deltatsqover2, &deltatover2, &b_deltat, &b_deltatsqover2,
                                   &b_deltatover2, &one, &zero);
    if (any_el)
      update_displacement_cuda_(&s->mesh, &deltat, &deltatsqover2, &deltatover2, &b_deltat, &b_deltatsqover2,
                                &b_deltatover2);
    mark(step, 3);
    // the fluid first, coupled to the predicted solid displacement (compute_gpu_acoustic.f90:34-220)
    if (any_ac && !any_el) mark(step, 0);
    if (any_ac) {
      acoustic_enforce_free_surf_cuda_(&s->mesh, &one, &zero);
      for (int iphase = 1; iphase <= 2; iphase++) {
        compute_forces_acoustic_cuda_(&s->mesh, &iphase, &zero, &nspec_inner_ac, &zero, &one, &zero);
        if (pml) pml_boundary_acoustic_cuda_(&s->mesh, &one, &zero);
        if (iphase == 1) {
          if (s->p.stacey)
            compute_stacey_acoustic_cuda_(&s->mesh, &iphase, slice_ac[4].data(), slice_ac[2].data(),
                                          slice_ac[3].data(), slice_ac[1].data(), &one, &zero, &zero);
          if (any_el && nfluid_solid > 0) compute_coupling_ac_el_cuda_(&s->mesh, &iphase, &nfluid_solid, &one);
          compute_add_sources_ac_cuda_(&s->mesh, &iphase, &it);
        }
      }
      kernel_3_acoustic_cuda_(&s->mesh, &deltatover2, &b_deltatover2, &one, &zero);
      acoustic_enforce_free_surf_cuda_(&s->mesh, &one, &zero);
    }
    if (any_ac && !any_el) {
      mark(step, 1);
      mark(step, 4);
      mark(step, 5);
    }
    // then the solid, coupled to the new fluid pressure (compute_gpu_elastic.f90:34-215)
    for (int iphase = 1; iphase <= 2 && any_el; iphase++) {
      if (iphase == 2) mark(step, 0);
      compute_forces_viscoelastic_cuda_(&s->mesh, &iphase, &deltat, &nspec_outer, &nspec_inner, &aniso, &att, &one,
                                        &zero);
      if (iphase == 2) mark(step, 1);
      if (iphase == 1) {
        mark(step, 4);
        if (s->p.stacey)
          compute_stacey_viscoelastic_cuda_(&s->mesh, &iphase, slice[4].data(), slice[2].data(), slice[3].data(),
                                            slice[1].data(), &one, &zero, &zero);
        // the CPML layer's zero condition (compute_forces_viscoelastic_calling_routine.F90:74-77)
        if (pml) pml_boundary_elastic_cuda_(&s->mesh, &one, &zero);
        if (any_ac && nfluid_solid > 0) compute_coupling_el_ac_cuda_(&s->mesh, &iphase, &nfluid_solid, &one);
        compute_add_sources_el_cuda_(&s->mesh, &iphase, &it);
        mark(step, 5);
      }
    }
    if (absorb && s->save_forward && s->nabs > 0) {
      // the Fortran caller keeps the per-step slices (compute_gpu_elastic.f90:267-273, compute_gpu_acoustic.f90:
      // 162-167); a fluid face's term goes to component 0
      for (int f = 0; f < s->nabs; f++) {
        const int c = s->edge_abs[f];
        const int ib = s->ib0[f] - 1;
        const bool fluid = any_ac && s->is_acoustic[s->abs_ispec[f] - 1];
        for (int k = 0; k < NGLL; k++)
          for (int d = 0; d < 2; d++)
            absorb[(((size_t)step * s->nabs + f) * NGLL + k) * 2 + d] =
                fluid ? (d == 0 ? slice_ac[c][(size_t)ib * NGLL + k] : 0.f) : slice[c][((size_t)ib * NGLL + k) * 2 + d];
      }
    }
    mark(step, 6);
    if (any_el) kernel_3_a_cuda_(&s->mesh, &deltatover2, &b_deltatover2, &one, &zero);
    mark(step, 7);
    // the fields now hold (d_n, v', r_n) with the corrector pending: the state the reverse sweep needs
    if (s->grad_save == 1 && it % s->grad_chunk == 0 && it < s->p.nstep)
      sf2d_grad_save_state(&s->mesh, 0, it / s->grad_chunk - 1);
    else if (s->grad_save == 2)
      sf2d_grad_save_state(&s->mesh, 1, it - s->grad_seg_base - 1);
    mark(step, 8);
    if (s->nrec > 0) {
      int cur = it;
      compute_seismograms_cuda_(&s->mesh, &i_sig, s->sisux.data(), s->sisuz.data(), &cur, &nlength, &elastic_sim,
                                &acoustic_sim, &trick, &zero, &it, &it_end);
    }
    mark(step, 9);
  }
  s->it_done = it_end;
  // the last corrector is queued by kernel_3_a_cuda_; run it inside the measured loop
  if (any_el) sf2d_mesh_flush(&s->mesh);
  if (timed) {
    hc(hipEventRecord(loop_end, stream));
    hc(hipEventSynchronize(loop_end));
    float ms = 0.f;
    hc(hipEventElapsedTime(&ms, loop_start, loop_end));
    double acc[6] = {ms, 0, 0, 0, 0, 0};
    const int pairs[5][2] = {{2, 3}, {0, 1}, {4, 5}, {6, 7}, {8, 9}};  // predictor, forces, boundary, corrector, seismo
    for (int step = 0; step < nstep_run; step++)
      for (int q = 0; q < 5; q++) {
        if (pairs[q][1] >= nev) continue;
        float t = 0.f;
        hc(hipEventElapsedTime(&t, ev[(size_t)step * nev + pairs[q][0]], ev[(size_t)step * nev + pairs[q][1]]));
        acc[q + 1] += t;
      }
    for (int k = 0; k < 6; k++) timing[k] = acc[k];
    for (auto& e : ev) hc(hipEventDestroy(e));
    hc(hipEventDestroy(loop_start));
    hc(hipEventDestroy(loop_end));
  } else {
    hc(hipStreamSynchronize(stream));
  }
  if (seis)
    for (int r = 0; r < s->nrec; r++)
      for (int k = 0; k < nstep_run; k++) {
        seis[((size_t)r * 2 + 0) * nstep_run + k] = s->sisux[(size_t)r * nlength + it_begin - 1 + k];
        seis[((size_t)r * 2 + 1) * nstep_run + k] = s->sisuz[(size_t)r * nlength + it_begin - 1 + k];
      }
  return 0;
}

int sf2d_sim_batch_forward(void* sim, double* seis, double* timing) {
  Sim* s = (Sim*)sim;
  if (!s || !s->mesh || s->nshot < 1) return 1;
  hipStream_t stream = (hipStream_t)sf2d_mesh_stream(&s->mesh);
  sf2d_batch_reset(&s->mesh);
  hipEvent_t e0, e1;
  if (timing) {
    hc(hipEventCreate(&e0));
    hc(hipEventCreate(&e1));
    hc(hipEventRecord(e0, stream));
  }
  batch_steps(s, 1, s->p.nstep, true);
  sf2d_batch_flush(&s->mesh);
  if (timing) {
    hc(hipEventRecord(e1, stream));
    hc(hipEventSynchronize(e1));
    float ms = 0.f;
    hc(hipEventElapsedTime(&ms, e0, e1));
    timing[0] = ms;
    hc(hipEventDestroy(e0));
    hc(hipEventDestroy(e1));
  }
  if (seis) batch_seismograms(s, seis);
  return 0;
}

int sf2d_sim_nshot(void* sim) { return ((Sim*)sim)->nshot; }

void sf2d_sim_shot_t0(void* sim, double* t0) {
  Sim* s = (Sim*)sim;
  for (int k = 0; k < s->nshot; k++) t0[k] = s->shot_t0[k];
}

void sf2d_sim_shot_misfits(void* sim, double* chi) {
  Sim* s = (Sim*)sim;
  for (int k = 0; k < s->nshot && k < (int)s->shot_chi.size(); k++) chi[k] = s->shot_chi[k];
}

void sf2d_sim_batch_get_fields(void* sim, int shot, float* displ, float* veloc, float* accel) {
  Sim* s = (Sim*)sim;
  sf2d_batch_get_fields(&s->mesh, shot, displ, veloc, accel);
}

int sf2d_sim_adjoint(void* sim, const float* adj, const float* absorb, const float* last_displ,
                     const float* last_veloc, const float* last_accel, double* seis, float* kernels) {
  Sim* s = (Sim*)sim;
  if (!s || s->nspec_ac > 0) return 1;
  return sf2d_sim_adjoint_full(sim, adj, absorb, last_displ, last_veloc, last_accel, nullptr, nullptr, nullptr, seis,
                               kernels, nullptr);
}

int sf2d_sim_adjoint_full(void* sim, const float* adj, const float* absorb, const float* last_displ,
                          const float* last_veloc, const float* last_accel, const float* last_pot,
                          const float* last_pot_dot, const float* last_pot_dd, double* seis, float* kernels,
                          float* kernels_ac) {
  Sim* s = (Sim*)sim;
  if (!s || !s->mesh || s->simulation_type != 3 || s->nspec_pml > 0) return 1;
  const bool any_ac = s->nspec_ac > 0, any_el = s->nspec_ac < s->nspec;
  if ((any_el && (!last_displ || !last_veloc || !last_accel)) || (any_ac && (!last_pot || !last_pot_dot || !last_pot_dd)))
    return 1;
  const bool coupled = any_ac && any_el && !s->coup_ac_ispec.empty();
  const int nstep = s->p.nstep;
  // adjoint sources: Fortran source_adjoint(nrecloc, NSTEP, 2) (setup_sources_receivers.F90:578); a receiver in a
  // fluid element reads component 1 (compute_arrays_source.f90:415-424)
  {
    std::vector<float> sa((size_t)2 * nstep * std::max(1, s->nrec), 0.f);
    for (int r = 0; r < s->nrec; r++)
      for (int c = 0; c < 2; c++)
        for (int k = 0; k < nstep; k++) sa[((size_t)c * nstep + k) * s->nrec + r] = adj[((size_t)r * 2 + c) * nstep + k];
    int nadj = s->nrec, ns = nstep;
    prepare_sim2_or_3_const_device_(&s->mesh, &nadj, sa.data(), &ns);
  }
  // Newmark constants; the reconstructed field runs backwards (prepare_timerun.F90:262-277)
  float deltat = (float)s->p.dt;
  float deltatover2 = (float)(0.5 * (double)deltat);
  float deltatsqover2 = (float)(0.5 * (double)deltat * (double)deltat);
  float b_deltat = -deltat;
  float b_deltatover2 = (float)(0.5 * (double)b_deltat);
  float b_deltatsqover2 = (float)(0.5 * (double)b_deltat * (double)b_deltat);
  int nspec_outer = 0, nspec_inner = s->nspec - s->nspec_ac, nspec_inner_ac = s->nspec_ac, zero = 0, one = 1;
  int att = s->attenuation, nfluid_solid = (int)s->coup_ac_ispec.size(), three = 3;
  int i_sig = 1, nlength = nstep, it_end = nstep, nadj = s->nrec, ns = nstep;
  int elastic_sim = any_el ? 1 : 0, acoustic_sim = any_ac ? 1 : 0, trick = s->p.use_trick ? 1 : 0;
  s->sisux.assign((size_t)std::max(1, s->nrec) * nlength, 0.0);
  s->sisuz.assign(s->sisux.size(), 0.0);
  const int counts[5] = {0, s->nspec_side[1], s->nspec_side[2], s->nspec_side[3], s->nspec_side[4]};
  std::vector<float> slice[5], slice_ac[5];
  for (int c = 1; c <= 4; c++) slice[c].assign((size_t)2 * NGLL * std::max(1, counts[c]), 0.f);
  for (int c = 1; c <= 4; c++) slice_ac[c].assign((size_t)NGLL * std::max(1, counts[c]), 0.f);
  int size = 2 * s->nglob, nac = s->nglob;
  // the caller's slices of the stored Stacey terms of step it_tmp (compute_gpu_elastic.f90:252-258,
  // compute_gpu_acoustic.f90:148-153)
  auto load_slices = [&](int it_tmp) {
    for (int f = 0; f < s->nabs; f++) {
      const int c = s->edge_abs[f];
      const int ib = s->ib0[f] - 1;
      const bool fluid = any_ac && s->is_acoustic[s->abs_ispec[f] - 1];
      for (int k = 0; k < NGLL; k++) {
        const float* st = absorb + (((size_t)(it_tmp - 1) * s->nabs + f) * NGLL + k) * 2;
        if (fluid) {
          slice_ac[c][(size_t)ib * NGLL + k] = st[0];
        } else {
          slice[c][((size_t)ib * NGLL + k) * 2] = st[0];
          slice[c][((size_t)ib * NGLL + k) * 2 + 1] = st[1];
        }
      }
    }
  };
  // compute_forces_viscoacoustic_GPU (compute_gpu_acoustic.f90:34-220) for the wavefields w1 (adjoint), w2
  // (reconstructed)
  auto fluid_step = [&](int w1, int w2, int it_tmp) {
    acoustic_enforce_free_surf_cuda_(&s->mesh, &w1, &w2);
    for (int iphase = 1; iphase <= 2; iphase++) {
      compute_forces_acoustic_cuda_(&s->mesh, &iphase, &zero, &nspec_inner_ac, &zero, &w1, &w2);
      if (iphase == 1) {
        if (s->p.stacey)
          compute_stacey_acoustic_cuda_(&s->mesh, &iphase, slice_ac[4].data(), slice_ac[2].data(), slice_ac[3].data(),
                                        slice_ac[1].data(), &w1, &w2, &zero);
        if (coupled && w1) compute_coupling_ac_el_cuda_(&s->mesh, &iphase, &nfluid_solid, &one);
        if (coupled && w2) compute_coupling_ac_el_cuda_(&s->mesh, &iphase, &nfluid_solid, &three);
        int itt = it_tmp;
        if (s->nrec > 0 && w1) add_sources_ac_sim_2_or_3_cuda_(&s->mesh, &iphase, &itt, &nadj, &ns);
        if (w2) compute_add_sources_ac_s3_cuda_(&s->mesh, &iphase, &itt);
      }
    }
    kernel_3_acoustic_cuda_(&s->mesh, &deltatover2, &b_deltatover2, &w1, &w2);
    acoustic_enforce_free_surf_cuda_(&s->mesh, &w1, &w2);
  };
  // compute_forces_viscoelastic_GPU (compute_gpu_elastic.f90:34-215)
  auto solid_step = [&](int w1, int w2, int it_tmp) {
    for (int iphase = 1; iphase <= 2; iphase++) {
      compute_forces_viscoelastic_cuda_(&s->mesh, &iphase, &deltat, &nspec_outer, &nspec_inner, &zero, &att, &w1,
                                        &w2);
      if (iphase == 1) {
        if (s->p.stacey && s->nabs > 0)
          compute_stacey_viscoelastic_cuda_(&s->mesh, &iphase, slice[4].data(), slice[2].data(), slice[3].data(),
                                            slice[1].data(), &w1, &w2, &zero);
        if (coupled && w1) compute_coupling_el_ac_cuda_(&s->mesh, &iphase, &nfluid_solid, &one);
        if (coupled && w2) compute_coupling_el_ac_cuda_(&s->mesh, &iphase, &nfluid_solid, &three);
        int itt = it_tmp;
        if (s->nrec > 0 && w1) add_sources_el_sim_type_2_or_3_(&s->mesh, &iphase, &itt, &nadj, &ns);
        if (w2) compute_add_sources_el_s3_cuda_(&s->mesh, &iphase, &itt);
      }
    }
    kernel_3_a_cuda_(&s->mesh, &deltatover2, &b_deltatover2, &w1, &w2);
  };
  for (int it = 1; it <= nstep; it++) {
    const int it_tmp = nstep - it + 1;
    // update_displ_Newmark: both fields of each domain (update_displacement_Newmark.F90:126-160, 500-548)
    if (any_ac)
      update_displacement_ac_cuda_(&s->mesh, &deltat, &deltatsqover2, &deltatover2, &b_deltat, &b_deltatsqover2,
                                   &b_deltatover2, &one, &zero);
    if (any_el)
      update_displacement_cuda_(&s->mesh, &deltat, &deltatsqover2, &deltatover2, &b_deltat, &b_deltatsqover2,
                                &b_deltatover2);
    if (s->p.stacey && s->nabs > 0) load_slices(it_tmp);
    // iterate_time.F90:227-246, 328-358: the adjoint solid, then the adjoint fluid; on coupled meshes the
    // reconstructed fields follow in the forward order, fluid then solid
    if (coupled) {
      solid_step(1, 0, it_tmp);
      fluid_step(1, 0, it_tmp);
      fluid_step(0, 1, it_tmp);
      solid_step(0, 1, it_tmp);
    } else {
      if (any_el) solid_step(1, 1, it_tmp);
      if (any_ac) fluid_step(1, 1, it_tmp);
    }
    if (it == 1) {
      // read_forward_arrays.f90:52-98: the last forward frames replace the reconstructed fields
      if (any_ac) {
        std::vector<float> p(last_pot, last_pot + nac), v(last_pot_dot, last_pot_dot + nac),
            a(last_pot_dd, last_pot_dd + nac);
        transfer_b_fields_ac_to_device_(&nac, p.data(), v.data(), a.data(), &s->mesh);
      }
      if (any_el) {
        std::vector<float> d(last_displ, last_displ + size), v(last_veloc, last_veloc + size),
            a(last_accel, last_accel + size);
        transfer_b_fields_to_device_(&size, d.data(), v.data(), a.data(), &s->mesh);
      }
    }
    if (s->nrec > 0) {
      int cur = it;
      compute_seismograms_cuda_(&s->mesh, &i_sig, s->sisux.data(), s->sisuz.data(), &cur, &nlength, &elastic_sim,
                                &acoustic_sim, &trick, &zero, &it, &it_end);
    }
    // compute_kernels.f90:34-80: solid, then fluid
    if (any_el) compute_kernels_elastic_cuda_(&s->mesh, &deltat);
    if (any_ac) compute_kernels_acoustic_cuda_(&s->mesh, &deltat);
  }
  s->it_done = nstep;
  sf2d_mesh_synchronize(&s->mesh);
  if (seis)
    for (int r = 0; r < s->nrec; r++)
      for (int k = 0; k < nstep; k++) {
        seis[((size_t)r * 2 + 0) * nstep + k] = s->sisux[(size_t)r * nlength + k];
        seis[((size_t)r * 2 + 1) * nstep + k] = s->sisuz[(size_t)r * nlength + k];
      }
  const size_t ne = (size_t)s->nspec * NGLL2;
  int nsp = s->nspec;
  if (kernels) {
    // transfer_kernels_el_to_host + save_adjoint_kernels.f90:100-140 scaling (zero in fluid elements)
    std::vector<float> rk(ne, 0.f), mk(ne, 0.f), kk(ne, 0.f);
    if (any_el) transfer_kernels_el_to_host_(&s->mesh, rk.data(), mk.data(), kk.data(), &nsp);
    for (size_t l = 0; l < ne; l++) {
      kernels[l] = -s->rho[l] * deltat * rk[l];
      kernels[ne + l] = (float)(-2.0 * (double)s->mu[l] * (double)deltat * (double)mk[l]);
      kernels[2 * ne + l] = -s->kappa[l] * deltat * kk[l];
    }
  }
  if (kernels_ac) {
    // transfer_kernels_ac_to_host and it_transfer_from_GPU (iterate_time.F90:485-497): rho, kappa, rho' and c
    std::vector<float> rk(ne, 0.f), kk(ne, 0.f);
    if (any_ac) transfer_kernels_ac_to_host_(&s->mesh, rk.data(), kk.data(), &nsp);
    for (size_t l = 0; l < ne; l++) {
      kernels_ac[l] = rk[l];
      kernels_ac[ne + l] = kk[l];
      kernels_ac[2 * ne + l] = rk[l] + kk[l];
      kernels_ac[3 * ne + l] = (float)(2.0 * (double)kk[l]);
    }
  }
  return 0;
}

int sf2d_sim_set_chain_elems(void* sim, int n) {
  Sim* s = (Sim*)sim;
  if (!s || s->mesh || n < 0 || n > 12) return 1;
  s->chain_request = n;
  return 0;
}

int sf2d_sim_set_anisotropy(void* sim, const int* is_aniso, const double* c) {
  Sim* s = (Sim*)sim;
  if (!s || s->mesh || s->simulation_type != 1 || s->nshot > 0 || s->attenuation || s->nspec_pml > 0) return 1;
  const size_t ne = (size_t)s->nspec * NGLL2;
  for (int e = 0; e < s->nspec; e++)
    if (is_aniso[e] && s->nspec_ac > 0 && s->is_acoustic[e]) return 2;
  s->aniso_elem.assign(is_aniso, is_aniso + s->nspec);
  for (int k = 0; k < 6; k++) {
    s->aniso_c[k].resize(ne);
    for (size_t l = 0; l < ne; l++) s->aniso_c[k][l] = (float)c[k * ne + l];
  }
  return 0;
}

int sf2d_sim_set_model(void* sim, const double* rho, const double* vp, const double* vs) {
  Sim* s = (Sim*)sim;
  if (!s || s->nspec_pml > 0) return 1;  // the CPML layer's profiles and moduli are fixed at set-up
  const size_t ne = (size_t)s->nspec * NGLL2;
  for (size_t l = 0; l < ne; l++)
    if (!(rho[l] > 0.0) || !(vp[l] > 0.0) || !(vs[l] >= 0.0)) return 2;
  build_materials(*s, rho, vp, vs);
  s->model_rho.assign(rho, rho + ne);
  s->model_vp.assign(vp, vp + ne);
  s->model_vs.assign(vs, vs + ne);
  build_mass(*s);
  if (s->mesh)
    sf2d_mesh_set_materials(&s->mesh, s->kappa.data(), s->mu.data(), s->rho_vp.data(), s->rho_vs.data(),
                            s->rmassx.data(), s->rmassz.data());
  return 0;
}

int sf2d_sim_gradient_forward(void* sim, const double* obs, int chunk, double* chi, double* seis) {
  Sim* s = (Sim*)sim;
  if (!s || !s->mesh || s->simulation_type != 1 || chunk < 1 || s->nspec_ac > 0 || !s->aniso_elem.empty() ||
      s->p.seismotype != 1 ||
      s->attenuation || s->nspec_pml > 0)
    return 1;
  const int N = s->p.nstep, R = s->nrec, S = std::max(1, s->nshot);
  s->grad_chunk = std::min(chunk, N);
  std::vector<double> sm((size_t)S * std::max(1, R) * 2 * N, 0.0);
  s->grad_save = 1;
  if (s->nshot > 0) {
    sf2d_batch_grad_prepare(&s->mesh, chunk);
    sf2d_batch_reset(&s->mesh);
    batch_steps(s, 1, N, true);
    s->grad_save = 0;
    batch_seismograms(s, sm.data());
  } else {
    sf2d_grad_prepare(&s->mesh, chunk, R);
    sf2d_sim_reset(sim);
    const int rc = sf2d_sim_forward(sim, N, sm.data(), nullptr, nullptr);
    s->grad_save = 0;
    if (rc != 0) return rc;
  }
  // chi = 1/2 dt sum |s - o|^2 in double (per shot, then summed); residuals dt (s - o), [N][S][R][2]
  const double dt = (double)(float)s->p.dt;
  s->grad_res.assign((size_t)S * std::max(1, R) * 2 * N, 0.f);
  s->shot_chi.assign(S, 0.0);
  double total = 0.0;
  for (int sh = 0; sh < S; sh++) {
    double acc = 0.0;
    for (int r = 0; r < R; r++)
      for (int c = 0; c < 2; c++)
        for (int n = 0; n < N; n++) {
          const size_t at = (((size_t)sh * R + r) * 2 + c) * N + n;
          const double q = sm[at] - obs[at];
          acc += 0.5 * dt * q * q;
          s->grad_res[(((size_t)n * S + sh) * R + r) * 2 + c] = (float)(dt * q);
        }
    s->shot_chi[sh] = acc;
    total += acc;
  }
  *chi = total;
  if (seis) std::copy(sm.begin(), sm.begin() + (size_t)S * R * 2 * N, seis);
  s->grad_ready = 1;
  return 0;
}

int sf2d_sim_gradient_backward(void* sim, double* grad, double* parts) {
  Sim* s = (Sim*)sim;
  if (!s || !s->mesh || !s->grad_ready) return 1;
  const int N = s->p.nstep, C = s->grad_chunk;
  const float dt = (float)s->p.dt, dto2 = (float)(0.5 * (double)dt);
  sf2d_grad_begin_reverse(&s->mesh, s->grad_res.data());
  const int nseg = (N + C - 1) / C;
  for (int k = nseg - 1; k >= 0; k--) {
    const int n0 = k * C, n1 = std::min(N, n0 + C);
    if (s->nshot > 0) {
      // a batch recomputes the segment straight into its state slots
      sf2d_grad_recompute_segment(&s->mesh, k - 1, n0, n1, dt);
    } else {
      if (k == 0)
        sf2d_sim_reset(sim);
      else
        sf2d_grad_restore_checkpoint(&s->mesh, k - 1, dto2);
      s->it_done = n0;
      s->grad_save = 2;
      s->grad_seg_base = n0;
      const int rc = sf2d_sim_forward(sim, n1 - n0, nullptr, nullptr, nullptr);
      s->grad_save = 0;
      if (rc != 0) return rc;
    }
    for (int n = n1; n > n0; n--) sf2d_grad_reverse_step(&s->mesh, n, n - n0 - 1, n == N ? 1 : 0, dt);
    sf2d_grad_flush(&s->mesh);
  }
  const size_t ne = (size_t)s->nspec * NGLL2, nf = (size_t)s->nabs * NGLL;
  std::vector<double> gk(ne), gm(ne), pm((size_t)2 * s->nglob), fvp(std::max<size_t>(1, nf)), fvs(std::max<size_t>(1, nf));
  sf2d_grad_result(&s->mesh, gk.data(), gm.data(), pm.data(), fvp.data(), fvs.data());
  // mass term: M_c(g) = sum w_i w_j J rho (+ the Stacey term below)
  std::vector<double> gr(ne, 0.0), gvp(ne, 0.0), gvs(ne, 0.0);
  for (int e = 0; e < s->nspec; e++)
    for (int p = 0; p < NGLL2; p++) {
      const size_t l = (size_t)e * NGLL2 + p;
      const int g = s->ibool[l] - 1;
      const double w = (double)s->basis.w[p % NGLL] * (double)s->basis.w[p / NGLL] * (double)s->jacobian[l];
      gr[l] = w * (pm[2 * (size_t)g] + pm[2 * (size_t)g + 1]);
    }
  // Stacey traction and its Newmark mass term T_c = dt/2 w [rho_vp (nx+nz) n_c + rho_vs (1 - (nx+nz) n_c)]
  for (int f = 0; f < s->nabs; f++) {
    const int e = s->abs_ispec[f] - 1;
    for (int k = 0; k < NGLL; k++) {
      const size_t q = (size_t)f * NGLL + k;
      const int ii = s->abs_ij[2 * q], jj = s->abs_ij[2 * q + 1];
      if (ii > NGLL || jj > NGLL) continue;
      const size_t l = (size_t)e * NGLL2 + (jj - 1) * NGLL + (ii - 1);
      const int g = s->ibool[l] - 1;
      const double nx = s->abs_normal[2 * q], nz = s->abs_normal[2 * q + 1], w = s->abs_jacw[q];
      const double sn = nx + nz, px = pm[2 * (size_t)g], pz = pm[2 * (size_t)g + 1];
      gvp[l] += fvp[q] + w * (double)dto2 * (px * sn * nx + pz * sn * nz);
      gvs[l] += fvs[q] + w * (double)dto2 * (px * (1.0 - sn * nx) + pz * (1.0 - sn * nz));
    }
  }
  // chain to (rho, vp, vs): mu = rho vs^2, kappa = rho vp^2 - mu, rho_vp = rho vp, rho_vs = rho vs
  for (size_t l = 0; l < ne; l++) {
    const double rho = s->model_rho[l], vp = s->model_vp[l], vs = s->model_vs[l];
    grad[l] = gk[l] * (vp * vp - vs * vs) + gm[l] * vs * vs + gr[l] + gvp[l] * vp + gvs[l] * vs;
    grad[ne + l] = gk[l] * 2.0 * rho * vp + gvp[l] * rho;
    grad[2 * ne + l] = -gk[l] * 2.0 * rho * vs + gm[l] * 2.0 * rho * vs + gvs[l] * rho;
    if (parts) {
      parts[l] = gk[l];
      parts[ne + l] = gm[l];
      parts[2 * ne + l] = gr[l];
      parts[3 * ne + l] = gvp[l];
      parts[4 * ne + l] = gvs[l];
    }
  }
  return 0;
}

int sf2d_sim_set_attenuation(void* sim, const double* qkappa, const double* qmu, int n_sls) {
  Sim* s = (Sim*)sim;
  if (!s || s->mesh || !qkappa || !qmu) return 1;
  if (n_sls != 3 || s->simulation_type != 1 || s->nshot > 0 || s->src.empty()) {
    fprintf(stderr, "sf2d_sim_set_attenuation: 3 solids, forward runs with a source only\n");
    return 1;
  }
  const double f0 = s->src[0].f0;  // ATTENUATION_f0_REFERENCE (prepare_attenuation.f90:319-322)
  const size_t ne = (size_t)s->nspec * NGLL2;
  for (auto* v : {&s->att_A_kappa, &s->att_B_kappa, &s->att_A_mu, &s->att_B_mu}) v->assign(3 * ne, 0.f);
  // the CPU path's defaults at points without attenuation (prepare_attenuation.f90:285-289)
  for (auto* v : {&s->zen_itau1, &s->zen_te1, &s->zen_itau2, &s->zen_te2}) v->assign(3 * ne, -1.f);
  for (auto* v : {&s->zen_phi1, &s->zen_phi2}) v->assign(3 * ne, 0.f);
  for (auto* v : {&s->zen_mu1, &s->zen_mu2}) v->assign(ne, 1.f);
  s->zen_qk.assign(qkappa, qkappa + ne);
  s->zen_qm.assign(qmu, qmu + ne);
  std::map<double, std::array<float, 16>> cache;  // Q -> A[3], B[3], Z[10]
  auto last = cache.end();
  auto coefficients = [&](double q, float* A, float* B, size_t l, int mode) {
    auto it = (last != cache.end() && last->first == q) ? last : cache.find(q);
    if (it == cache.end()) {
      std::array<float, 16> c{};
      if (!newmark_coefficients(q, f0, s->p.dt, c.data(), c.data() + 3, c.data() + 6)) return false;
      it = cache.emplace(q, c).first;
    }
    last = it;
    const float* c = it->second.data();
    std::copy(c, c + 3, A);
    std::copy(c + 3, c + 6, B);
    std::vector<float>& itau = mode ? s->zen_itau2 : s->zen_itau1;
    std::vector<float>& te = mode ? s->zen_te2 : s->zen_te1;
    std::vector<float>& phi = mode ? s->zen_phi2 : s->zen_phi1;
    for (int k = 0; k < 3; k++) {
      itau[l + ne * k] = c[6 + k];
      te[l + ne * k] = c[9 + k];
      phi[l + ne * k] = c[12 + k];
    }
    (mode ? s->zen_mu2 : s->zen_mu1)[l] = c[15];
    return true;
  };
  for (size_t l = 0; l < ne; l++) {
    if (qkappa[l] > 9998.999 && qmu[l] > 9998.999) continue;  // no attenuation at this point: A = B = 0
    if (s->nspec_ac > 0 && s->is_acoustic[l / NGLL2]) {  // viscoacoustic attenuation is another model
      fprintf(stderr, "sf2d_sim_set_attenuation: finite quality factors at point %zu of a fluid element\n", l);
      return 1;
    }
    if (!(qkappa[l] > 0.0 && qmu[l] > 0.0) ||
        !coefficients(qkappa[l], &s->att_A_kappa[3 * l], &s->att_B_kappa[3 * l], l, 0) ||
        !coefficients(qmu[l], &s->att_A_mu[3 * l], &s->att_B_mu[3 * l], l, 1)) {
      fprintf(stderr, "sf2d_sim_set_attenuation: invalid quality factors at point %zu\n", l);
      return 1;
    }
  }
  s->attenuation = 1;
  return 0;
}

void sf2d_sim_get_attenuation(void* sim, float* A_kappa, float* B_kappa, float* A_mu, float* B_mu) {
  Sim* s = (Sim*)sim;
  if (!s || !s->attenuation) return;
  std::copy(s->att_A_kappa.begin(), s->att_A_kappa.end(), A_kappa);
  std::copy(s->att_B_kappa.begin(), s->att_B_kappa.end(), B_kappa);
  std::copy(s->att_A_mu.begin(), s->att_A_mu.end(), A_mu);
  std::copy(s->att_B_mu.begin(), s->att_B_mu.end(), B_mu);
}

void sf2d_sim_reset(void* sim) {
  Sim* s = (Sim*)sim;
  if (!s || !s->mesh) return;
  sf2d_attenuation_reset(&s->mesh);
  sf2d_pml_reset(&s->mesh);
  std::vector<float> zeros((size_t)2 * s->nglob, 0.f);
  int size = 2 * s->nglob, nac = s->nglob;
  if (s->nspec_ac > 0) transfer_fields_ac_to_device_(&nac, zeros.data(), zeros.data(), zeros.data(), &s->mesh);
  if (s->nspec_ac < s->nspec) transfer_fields_el_to_device_(&size, zeros.data(), zeros.data(), zeros.data(), &s->mesh);
  s->it_done = 0;
  std::fill(s->sisux.begin(), s->sisux.end(), 0.0);
  std::fill(s->sisuz.begin(), s->sisuz.end(), 0.0);
}

int sf2d_sim_steps_done(void* sim) { return ((Sim*)sim)->it_done; }

int sf2d_sim_nspec_acoustic(void* sim) { return ((Sim*)sim)->nspec_ac; }
int sf2d_sim_ncoupling(void* sim) { return (int)((Sim*)sim)->coup_ac_ispec.size(); }

void sf2d_sim_get_rmass_acoustic(void* sim, float* out) {
  Sim* s = (Sim*)sim;
  for (int g = 0; g < s->nglob; g++) out[g] = s->nspec_ac > 0 ? s->rmass_ac[g] : 1.f;
}

void sf2d_sim_get_coupling(void* sim, int* ac_ispec, int* ac_code, int* el_ispec, int* el_code) {
  Sim* s = (Sim*)sim;
  for (size_t q = 0; q < s->coup_ac_ispec.size(); q++) {
    ac_ispec[q] = s->coup_ac_ispec[q];
    ac_code[q] = s->coup_ac_code[q];
    el_ispec[q] = s->coup_el_ispec[q];
    el_code[q] = s->coup_el_code[q];
  }
}

void sf2d_sim_get_acoustic_fields(void* sim, float* pot, float* pot_dot, float* pot_dot_dot) {
  Sim* s = (Sim*)sim;
  int size = s->nglob;
  std::vector<float> p((size_t)size), v((size_t)size), a((size_t)size);
  transfer_fields_ac_from_device_(&size, p.data(), v.data(), a.data(), &s->mesh);
  if (pot) std::copy(p.begin(), p.end(), pot);
  if (pot_dot) std::copy(v.begin(), v.end(), pot_dot);
  if (pot_dot_dot) std::copy(a.begin(), a.end(), pot_dot_dot);
}

void sf2d_sim_set_acoustic_fields(void* sim, const float* pot, const float* pot_dot, const float* pot_dot_dot) {
  Sim* s = (Sim*)sim;
  int size = s->nglob;
  std::vector<float> p((size_t)size, 0.f), v((size_t)size, 0.f), a((size_t)size, 0.f);
  if (pot) std::copy(pot, pot + size, p.begin());
  if (pot_dot) std::copy(pot_dot, pot_dot + size, v.begin());
  if (pot_dot_dot) std::copy(pot_dot_dot, pot_dot_dot + size, a.begin());
  transfer_fields_ac_to_device_(&size, p.data(), v.data(), a.data(), &s->mesh);
}

void sf2d_sim_get_fields(void* sim, float* displ, float* veloc, float* accel) {
  Sim* s = (Sim*)sim;
  int size = 2 * s->nglob;
  std::vector<float> tmp((size_t)size);
  if (displ) transfer_displ_from_device_(&size, displ, &s->mesh);
  if (accel) transfer_accel_from_device_(&size, accel, &s->mesh);
  if (veloc) {
    transfer_fields_el_from_device_(&size, tmp.data(), veloc, tmp.data(), &s->mesh);
  }
}
void sf2d_sim_set_fields(void* sim, const float* displ, const float* veloc, const float* accel) {
  Sim* s = (Sim*)sim;
  int size = 2 * s->nglob;
  std::vector<float> d((size_t)size, 0.f), v((size_t)size, 0.f), a((size_t)size, 0.f);
  if (displ) std::copy(displ, displ + size, d.begin());
  if (veloc) std::copy(veloc, veloc + size, v.begin());
  if (accel) std::copy(accel, accel + size, a.begin());
  transfer_fields_el_to_device_(&size, d.data(), v.data(), a.data(), &s->mesh);
}
void sf2d_sim_get_b_fields(void* sim, float* displ, float* veloc, float* accel) {
  Sim* s = (Sim*)sim;
  int size = 2 * s->nglob;
  std::vector<float> d((size_t)size), v((size_t)size), a((size_t)size);
  transfer_b_fields_from_device_(&size, d.data(), v.data(), a.data(), &s->mesh);
  if (displ) std::copy(d.begin(), d.end(), displ);
  if (veloc) std::copy(v.begin(), v.end(), veloc);
  if (accel) std::copy(a.begin(), a.end(), accel);
}
void sf2d_sim_set_b_fields(void* sim, const float* displ, const float* veloc, const float* accel) {
  Sim* s = (Sim*)sim;
  int size = 2 * s->nglob;
  std::vector<float> d((size_t)size, 0.f), v((size_t)size, 0.f), a((size_t)size, 0.f);
  if (displ) std::copy(displ, displ + size, d.begin());
  if (veloc) std::copy(veloc, veloc + size, v.begin());
  if (accel) std::copy(accel, accel + size, a.begin());
  transfer_b_fields_to_device_(&size, d.data(), v.data(), a.data(), &s->mesh);
}

}  // extern "C"
