// Fortran-callable entry points of the elastic GPU path (the reference's src/gpu/*.cu host layer).
//
// Argument lists and meaning follow the reference's declarations (src/gpu/specfem2D_gpu_cuda_method_stubs.c);
// each function cites the reference definition it replaces.  Host arrays are copied at prepare time; the
// device layout is this library's own (unpadded [nspec][25], 0-based indices).
#include <sys/stat.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <array>
#include <map>
#include <vector>

#include "sf2d_mesh.h"
#include "specfem2d_gpu.h"
#include "specfem2d_gpu_ext.h"

namespace sf2d {

[[noreturn]] void fail(const char* fmt, ...) {
  char msg[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(msg, sizeof(msg), fmt, ap);
  va_end(ap);
  fprintf(stderr, "\nspecfem2d_gpu error: %s\n", msg);
  fflush(stderr);
  struct stat st;
  if (stat("OUTPUT_FILES", &st) == 0 && S_ISDIR(st.st_mode)) {
    FILE* f = fopen("OUTPUT_FILES/error_message_000000.txt", "a");
    if (f) {
      fprintf(f, "specfem2d_gpu error: %s\n", msg);
      fclose(f);
    }
  }
  exit(EXIT_FAILURE);
}

void check_hip(hipError_t err, const char* what) {
  if (err != hipSuccess) fail("%s: %s", what, hipGetErrorString(err));
}

// the z inverse mass for the Newmark kernels: null when it equals the x one at every point (one array read)
const float* rmass_z(const Mesh* mp) { return mp->same_mass ? nullptr : mp->d_rmassz; }

void flush_correctors(Mesh* mp) {
  if (mp->pending_corrector[0])
    launch_newmark_corrector(mp->d_veloc, mp->d_accel, mp->d_rmassx, rmass_z(mp), mp->nglob_el, mp->pending_dto2[0], 1,
                             mp->stream);
  if (mp->pending_corrector[1])
    launch_newmark_corrector(mp->d_b_veloc, mp->d_b_accel, mp->d_rmassx, rmass_z(mp), mp->nglob_el,
                             mp->pending_dto2[1], 1, mp->stream);
  mp->pending_corrector[0] = mp->pending_corrector[1] = 0;
}

namespace {

std::vector<int> to_zero_based(const int* a, size_t n, int lo, int hi, const char* what) {
  std::vector<int> out(n);
  for (size_t k = 0; k < n; k++) {
    int v = a[k];
    if (v < lo || v > hi) fail("%s: index %d at position %zu outside [%d, %d]", what, v, k, lo, hi);
    out[k] = v - 1;
  }
  return out;
}



// Greedy colouring of an element list: each element takes the lowest colour not yet used by any element
// sharing one of its GLL points (per-point colour bit masks).  Structured quad meshes get 4 colours.
int colour_elements(const std::vector<int>& ibool, const std::vector<int>& elems, int nglob, std::vector<int>& colour) {
  std::vector<uint32_t> used((size_t)nglob, 0u);
  colour.assign(elems.size(), 0);
  int ncol = 0;
  for (size_t k = 0; k < elems.size(); k++) {
    const int* ib = ibool.data() + (size_t)elems[k] * NGLL2;
    uint32_t forbidden = 0;
    for (int p = 0; p < NGLL2; p++) forbidden |= used[ib[p]];
    if (forbidden == 0xffffffffu) fail("mesh colouring needs more than %d colours", Mesh::MAX_COLOURS);
    int c = 0;
    while (forbidden & (1u << c)) c++;
    colour[k] = c;
    for (int p = 0; p < NGLL2; p++) used[ib[p]] |= (1u << c);
    if (c + 1 > ncol) ncol = c + 1;
  }
  return ncol;
}

// Groups a phase's element list into patches: runs of consecutive list entries, at most PATCH_ELEMS long,
// in which every element shares a point with the patch built so far.  For the internal mesher's row-major
// element order these are row segments of PATCH_ELEMS elements.
std::vector<int> make_patches(const std::vector<int>& ibool, const std::vector<int>& elems, int nglob,
                              int max_elems) {
  std::vector<int> start;  // index in elems of each patch's first element, plus elems.size()
  std::vector<int> mark((size_t)nglob, -1);
  int cur = -1, size = 0;
  for (size_t k = 0; k < elems.size(); k++) {
    const int* ib = ibool.data() + (size_t)elems[k] * NGLL2;
    bool touches = false;
    if (cur >= 0)
      for (int p = 0; p < NGLL2 && !touches; p++) touches = (mark[ib[p]] == cur);
    if (cur < 0 || size == max_elems || !touches) {
      cur = (int)start.size();
      start.push_back((int)k);
      size = 0;
    }
    for (int p = 0; p < NGLL2; p++) mark[ib[p]] = cur;
    size++;
  }
  start.push_back((int)elems.size());
  return start;
}

// Device element order and assembly schedule of both phase lists (host only, no device call):
//   patch  (mode 0): patches coloured greedily (patches of one colour share no point); order (phase, colour, patch).
//                    The points are renumbered patch-major (see below).
//   colour (mode 1): elements coloured greedily; order (phase, colour)
//   atomic (mode 2): the phase list order, one launch per phase with FP32 atomics
//
// Point renumbering of the patch mode.  Walking the patches in phase-list order, each patch owns the points
// that no earlier patch touched and numbers them consecutively, row by row (GLL row j outermost, then its
// elements, then i).  A patch therefore reads its owned points as one contiguous range, and only the points
// owned by earlier patches (its halo, sorted) through an index list.  On a strip of elements the halo is the
// strip's bottom row, which is the top row of the strip below and hence itself contiguous.
struct Plan {
  std::vector<int> order;  // ispec of each device position
  int phase_count[2] = {0, 0};
  int ncolour[2] = {0, 0};
  int colour_start[2][Mesh::MAX_COLOURS] = {};
  int colour_count[2][Mesh::MAX_COLOURS] = {};
  // patch mode
  std::vector<int> newid;             // internal id of every point (a permutation of 0..nglob-1)
  std::vector<int> desc;              // [npatch][4] = {first element position, first owned id, halo offset, counts}
  std::vector<int> halo;              // internal ids of the halo points of all patches, one sorted list per patch
  // per device position and GLL point: index of its point in the patch's point list (low 8 bits) and its
  // rank among the element points of the patch that share that point (bits 8-9)
  std::vector<uint16_t> slot;
  // every patch is a chain: element k's right edge (i = 4) is element k+1's left edge (i = 0), and no other
  // point is shared inside the patch
  bool chains = false;
  std::vector<int> uncovered;         // chains: points (caller numbering) of no element of the phase lists
  // chains: per device position and GLL point, whether the chain kernel's update of that point is the first of
  // the step (the earliest colour, in launch order, of the patches that contain the point)
  std::vector<uint8_t> first;
};

// whether elements elems[k0..k1) form a chain (see Plan::chains)
bool is_chain(const std::vector<int>& ibool, const std::vector<int>& elems, int k0, int k1) {
  std::vector<int> pts;
  for (int k = k0; k < k1; k++) {
    const int* a = ibool.data() + (size_t)elems[k] * NGLL2;
    if (k + 1 < k1) {
      const int* b = ibool.data() + (size_t)elems[k + 1] * NGLL2;
      for (int j = 0; j < NGLL; j++)
        if (a[j * NGLL + NGLL - 1] != b[j * NGLL]) return false;
    }
    pts.insert(pts.end(), a, a + NGLL2);
  }
  std::sort(pts.begin(), pts.end());
  const size_t distinct = std::unique(pts.begin(), pts.end()) - pts.begin();
  return distinct == (size_t)(NGLL * (4 * (k1 - k0) + 1));
}

int desc_counts(int ne, int nown, int nhalo) { return ne | (nown << 8) | (nhalo << 20); }

Plan build_plan(const std::vector<int>& ibool, int nspec, int nglob, const int* phase_lists, int np, int mode,
                int max_elems = Mesh::PATCH_ELEMS) {
  Plan P;
  std::vector<int>& order = P.order;
  order.reserve(nspec);
  std::vector<char> placed(nspec, 0);
  std::vector<std::vector<int>> phase_elems(2);
  for (int ph = 0; ph < 2; ph++) {
    std::vector<int>& elems = phase_elems[ph];
    for (int k = 0; k < np; k++) {
      const int v = phase_lists[(size_t)ph * np + k];
      if (v < 1) break;  // lists are padded with zeros after their used prefix
      if (v > nspec) fail("phase_ispec_inner_elastic: element %d outside 1..%d", v, nspec);
      if (placed[v - 1]) fail("phase_ispec_inner_elastic: element %d listed twice", v);
      placed[v - 1] = 1;
      elems.push_back(v - 1);
    }
    P.phase_count[ph] = (int)elems.size();
  }
  if (mode == 0) {
    P.slot.assign((size_t)nspec * NGLL2, 0);
    P.newid.assign((size_t)nglob, -1);
    P.chains = true;
  }
  int next_id = 0;
  for (int ph = 0; ph < 2; ph++) {
    const std::vector<int>& elems = phase_elems[ph];
    if (mode == 2) {  // atomics: one range in list order
      P.ncolour[ph] = elems.empty() ? 0 : 1;
      P.colour_start[ph][0] = (int)order.size();
      P.colour_count[ph][0] = (int)elems.size();
      order.insert(order.end(), elems.begin(), elems.end());
      continue;
    }
    if (mode == 1) {  // coloured elements
      std::vector<int> colour;
      const int ncol = colour_elements(ibool, elems, nglob, colour);
      P.ncolour[ph] = ncol;
      for (int c = 0; c < ncol; c++) {
        P.colour_start[ph][c] = (int)order.size();
        for (size_t k = 0; k < elems.size(); k++)
          if (colour[k] == c) order.push_back(elems[k]);
        P.colour_count[ph][c] = (int)order.size() - P.colour_start[ph][c];
      }
      continue;
    }
    // coloured patches; ownership and numbering in phase-list order
    const std::vector<int> pstart = make_patches(ibool, elems, nglob, max_elems);
    const int np_ph = (int)pstart.size() - 1;
    for (int q = 0; q < np_ph && P.chains; q++) P.chains = is_chain(ibool, elems, pstart[q], pstart[q + 1]);
    std::vector<std::vector<int>> pnodes(np_ph), phalo(np_ph);
    std::vector<int> own0(np_ph), nown(np_ph);
    for (int q = 0; q < np_ph; q++) {
      std::vector<int>& nodes = pnodes[q];
      own0[q] = next_id;
      for (int j = 0; j < NGLL; j++)
        for (int k = pstart[q]; k < pstart[q + 1]; k++)
          for (int i = 0; i < NGLL; i++) {
            const int g = ibool[(size_t)elems[k] * NGLL2 + j * NGLL + i];
            nodes.push_back(g);
            if (P.newid[g] < 0) P.newid[g] = next_id++;
          }
      nown[q] = next_id - own0[q];
      std::sort(nodes.begin(), nodes.end());
      nodes.erase(std::unique(nodes.begin(), nodes.end()), nodes.end());
      if (nodes.size() > 256) fail("patch with %zu points exceeds the 256-point workgroup", nodes.size());
      for (int g : nodes)
        if (P.newid[g] < own0[q]) phalo[q].push_back(P.newid[g]);
      std::sort(phalo[q].begin(), phalo[q].end());
    }
    std::vector<uint32_t> used((size_t)nglob, 0u);
    std::vector<int> pcol(np_ph, 0);
    int ncol = 0;
    for (int q = 0; q < np_ph; q++) {
      uint32_t forbidden = 0;
      for (int g : pnodes[q]) forbidden |= used[g];
      if (forbidden == 0xffffffffu) fail("patch colouring needs more than %d colours", Mesh::MAX_COLOURS);
      int c = 0;
      while (forbidden & (1u << c)) c++;
      pcol[q] = c;
      for (int g : pnodes[q]) used[g] |= (1u << c);
      ncol = std::max(ncol, c + 1);
    }
    P.ncolour[ph] = ncol;
    for (int c = 0; c < ncol; c++) {
      P.colour_start[ph][c] = (int)P.desc.size() / 4;
      for (int q = 0; q < np_ph; q++) {
        if (pcol[q] != c) continue;
        const int ne = pstart[q + 1] - pstart[q];
        const std::vector<int>& hl = phalo[q];
        P.desc.insert(P.desc.end(), {(int)order.size(), own0[q], (int)P.halo.size(),
                                     desc_counts(ne, nown[q], (int)hl.size())});
        std::vector<int> mult(nown[q] + hl.size(), 0);
        for (int k = pstart[q]; k < pstart[q + 1]; k++) {
          const int e = elems[k];
          const int pos = (int)order.size();
          order.push_back(e);
          const int* ib = ibool.data() + (size_t)e * NGLL2;
          for (int p = 0; p < NGLL2; p++) {
            const int id = P.newid[ib[p]];
            const int sl = id >= own0[q] ? id - own0[q]
                                         : nown[q] + (int)(std::lower_bound(hl.begin(), hl.end(), id) - hl.begin());
            const int rank = mult[sl]++;
            if (rank >= Mesh::PATCH_SHARE)
              fail("a point is shared by more than %d element points of one patch", Mesh::PATCH_SHARE);
            P.slot[(size_t)pos * NGLL2 + p] = (uint16_t)(sl | (rank << 8));
          }
        }
        P.halo.insert(P.halo.end(), hl.begin(), hl.end());
      }
      P.colour_count[ph][c] = (int)P.desc.size() / 4 - P.colour_start[ph][c];
    }
  }
  if (mode == 0 && P.chains) {
    // launch order of the colours: phase 1's, then phase 2's; a lane updates points i = 0..3 of its row, and
    // i = 4 in the patch's last element
    P.first.assign((size_t)nspec * NGLL2, 0);
    std::vector<int> earliest((size_t)nglob, 1 << 30);
    const int npatch = (int)P.desc.size() / 4;
    std::vector<int> launch(npatch, 0);
    for (int ph = 0, base = 0; ph < 2; base += P.ncolour[ph], ph++)
      for (int c = 0; c < P.ncolour[ph]; c++)
        for (int q = P.colour_start[ph][c]; q < P.colour_start[ph][c] + P.colour_count[ph][c]; q++) launch[q] = base + c;
    for (int pass = 0; pass < 2; pass++)
      for (int q = 0; q < npatch; q++) {
        const int pos0 = P.desc[4 * q], ne = P.desc[4 * q + 3] & 0xf;
        for (int k = 0; k < ne; k++)
          for (int j = 0; j < NGLL; j++)
            for (int i = 0; i < (k == ne - 1 ? NGLL : NGLL - 1); i++) {
              const size_t l = (size_t)(pos0 + k) * NGLL2 + j * NGLL + i;
              const int g = ibool[(size_t)order[pos0 + k] * NGLL2 + j * NGLL + i];
              if (pass == 0)
                earliest[g] = std::min(earliest[g], launch[q]);
              else
                P.first[l] = earliest[g] == launch[q];
            }
      }
    // points that no force launch writes (those of fluid elements only) join the special points, which the
    // predictor clears
    for (int g = 0; g < nglob; g++)
      if (earliest[g] == (1 << 30)) P.uncovered.push_back(g);
  }
  for (int e = 0; e < nspec; e++)
    if (!placed[e]) order.push_back(e);  // elements outside the elastic phase lists keep their relative order
  if (mode == 0) {
    for (int g = 0; g < nglob; g++)
      if (P.newid[g] < 0) P.newid[g] = next_id++;  // points of no elastic element
    P.halo.push_back(0);  // one spare entry: lanes without a halo point read a valid address
  }
  return P;
}

}  // namespace

// Patches of up to CHAIN_ELEMS elements over the given phase lists (1-based, zero-padded rows of np), coloured so
// that no two patches of a colour share a point; true when every patch is a chain.  For the fluid's chain kernel:
// only the element order, the patch descriptors and the colour ranges are used (not the plan's point numbering).
bool chain_patches(const std::vector<int>& ibool, int nspec, int nglob, const int* phase_lists, int np,
                   std::vector<int>& order, std::vector<int>& desc, int ncolour[2],
                   int colour_start[2][Mesh::MAX_COLOURS], int colour_count[2][Mesh::MAX_COLOURS]) {
  Plan P = build_plan(ibool, nspec, nglob, phase_lists, np, 0, Mesh::CHAIN_ELEMS);
  if (!P.chains) return false;
  order = std::move(P.order);
  desc = std::move(P.desc);
  for (int ph = 0; ph < 2; ph++) {
    ncolour[ph] = P.ncolour[ph];
    for (int c = 0; c < Mesh::MAX_COLOURS; c++) {
      colour_start[ph][c] = P.colour_start[ph][c];
      colour_count[ph][c] = P.colour_count[ph][c];
    }
  }
  return true;
}

// Row records of chain point ids: per patch (desc: {first position, -, -, ne}) and GLL row j, {base, id of column
// 0, first-touch mask low, high} such that column x > 0 of the row is point base + x.  ids: per element point (device
// order) the point id; flagged: the same with the FIRST_TOUCH flags.  False when some row is not numbered
// consecutively (or a shared column is not shared): the per-point ids are kept then.
bool chain_row_records(const std::vector<int>& desc, const std::vector<int>& ids, const std::vector<int>& flagged,
                       std::vector<int4>& rows) {
  const int npatch = (int)desc.size() / 4;
  rows.assign((size_t)npatch * NGLL, make_int4(0, 0, 0, 0));
  for (int q = 0; q < npatch; q++) {
    const int e0 = desc[4 * q], ne = desc[4 * q + 3] & 0xf;
    if (ne < 1 || 4 * ne + 1 > 64 || (size_t)(e0 + ne) * NGLL2 > ids.size()) return false;
    for (int j = 0; j < NGLL; j++) {
      // element point of the lane that writes column x of row j: (x / 4, x % 4), and (ne - 1, 4) for x = 4 ne
      auto at = [&](int x) {
        const int k = x == 4 * ne ? ne - 1 : x / 4;
        return (size_t)(e0 + k) * NGLL2 + j * NGLL + (x - 4 * k);
      };
      const int base = ids[at(1)] - 1;
      unsigned long long mask = 0ull;
      for (int x = 0; x <= 4 * ne; x++) {
        if (x > 0 && ids[at(x)] != base + x) return false;
        if (flagged[at(x)] & Mesh::FIRST_TOUCH) mask |= 1ull << x;
      }
      // the other element point of each shared column holds the same point (the chain property)
      for (int k = 1; k < ne; k++)
        if (ids[(size_t)(e0 + k - 1) * NGLL2 + j * NGLL + NGLL - 1] != ids[(size_t)(e0 + k) * NGLL2 + j * NGLL])
          return false;
      rows[(size_t)q * NGLL + j] =
          make_int4(base, ids[at(0)], (int)(unsigned)(mask & 0xffffffffull), (int)(unsigned)(mask >> 32));
    }
  }
  return true;
}

// Chain-kernel point ids with the first-touch flags, and the special points of first-touch assembly (points that
// the Stacey edges, the sources, the fluid and the CPML layer's zero condition also update: the predictor clears
// them and every force update adds to them).  Called when the plan is built and again when the Stacey edges or the
// CPML layer arrive.
void upload_chain_ids(Mesh* mp) {
  if (!mp->chain) return;
  const size_t ne = (size_t)mp->nspec * NGLL2;
  std::vector<char> special((size_t)mp->nglob, 0);
  for (int e : mp->h_src_ispec) {
    const int pos = mp->pos_of_ispec[e];
    for (int p = 0; p < NGLL2; p++) special[mp->h_chain_ids[(size_t)pos * NGLL2 + p]] = 1;
  }
  for (int g : mp->h_stacey_points) special[g] = 1;
  for (int g : mp->h_uncovered_points)
    if (g < mp->nglob_el) special[g] = 1;  // beyond nglob_el nothing touches the solid's fields
  for (int g : mp->h_coupling_points) special[mp->point_id.empty() ? g : mp->point_id[g]] = 1;
  for (int g : mp->h_pml_points) special[g] = 1;
  std::vector<int> ids(mp->h_chain_ids);
  std::vector<int> list;
  if (mp->first_touch) {
    for (size_t l = 0; l < ne; l++)
      if (mp->h_first[l] && !special[ids[l]]) ids[l] |= Mesh::FIRST_TOUCH;
    for (int g = 0; g < mp->nglob; g++)
      if (special[g]) list.push_back(g);
  }
  if (!mp->d_chain_ids) check_hip(hipMalloc(&mp->d_chain_ids, ne * sizeof(int)), "hipMalloc");
  check_hip(hipMemcpy(mp->d_chain_ids, ids.data(), ne * sizeof(int), hipMemcpyHostToDevice), "chain ids H2D");
  if (!getenv("SF2D_ROW_CHAIN") && !mp->d_col_ids) {
    mp->d_col_ids = device_zeros<int>(ne);
    mp->d_col_mod = device_zeros<float2>(ne);
  }
  launch_col_pack(mp);
  // the ids of each chain row as {base, id at x = 0, first-touch mask}, verified point by point
  device_free(mp->d_col_rows);
  mp->d_col_rows = nullptr;
  if (mp->d_col_ids && !getenv("SF2D_COL_IDS")) {
    std::vector<int4> rows;
    if (chain_row_records(mp->h_patch_desc, mp->h_chain_ids, ids, rows) && !rows.empty())
      mp->d_col_rows = device_copy(rows.data(), rows.size());
  }
  device_free(mp->d_special);
  mp->nspecial = (int)list.size();
  if (!list.empty()) mp->d_special = device_copy(list.data(), list.size());
}

namespace {

// Applies the plan: schedule into the handle, per-element device arrays moved into the plan's order, ibool
// rewritten in the plan's point numbering (patch mode).
void sort_elements_by_colour(Mesh* mp, const int* phase_lists, int np) {
  const int nspec = mp->nspec;
  if (mp->sorted) fail("prepare_fields_elastic_device called twice");
  // patch mode: chains of up to CHAIN_ELEMS elements when every patch is one (the chain kernel), otherwise
  // patches of up to PATCH_ELEMS for the general patch kernel
  const bool try_chain = mp->assembly == 0 && !getenv("SF2D_NO_CHAIN");
  mp->chain_elems = mp->chain_request > 0 ? mp->chain_request : Mesh::CHAIN_ELEMS;
  if (const char* env = mp->chain_request > 0 ? nullptr : getenv("SF2D_CHAIN_ELEMS")) {
    const int n = atoi(env);
    if (n < 1 || n > Mesh::CHAIN_ELEMS) fail("SF2D_CHAIN_ELEMS=%s outside 1..%d", env, Mesh::CHAIN_ELEMS);
    mp->chain_elems = n;
  }
  if (const char* env = getenv("SF2D_SHOTS_PER_WAVE")) mp->shots_per_wave = std::max(1, atoi(env));
  Plan P = build_plan(mp->h_ibool, nspec, mp->nglob, phase_lists, np, mp->assembly,
                      try_chain ? mp->chain_elems : Mesh::PATCH_ELEMS);
  if (try_chain && !P.chains) {
    if (mp->attenuation) {  // the memory variables ride in the chain kernel only: atomic assembly otherwise
      mp->assembly = 2;
      P = build_plan(mp->h_ibool, nspec, mp->nglob, phase_lists, np, 2);
    } else {
      P = build_plan(mp->h_ibool, nspec, mp->nglob, phase_lists, np, mp->assembly);
    }
  }
  std::vector<int>& order = P.order;
  for (int ph = 0; ph < 2; ph++) {
    mp->phase_count[ph] = P.phase_count[ph];
    mp->ncolour[ph] = P.ncolour[ph];
    for (int c = 0; c < Mesh::MAX_COLOURS; c++) {
      mp->colour_start[ph][c] = P.colour_start[ph][c];
      mp->colour_count[ph][c] = P.colour_count[ph][c];
    }
  }
  mp->ispec_of_pos = order;
  mp->pos_of_ispec.assign(nspec, 0);
  for (int pos = 0; pos < nspec; pos++) mp->pos_of_ispec[order[pos]] = pos;
  if (mp->assembly == 0) {
    mp->npatch = (int)P.desc.size() / 4;
    mp->h_patch_desc = P.desc;

    mp->d_patch_desc = reinterpret_cast<int4*>(device_copy(P.desc.data(), P.desc.size()));
    mp->d_halo_nodes = device_copy(P.halo.data(), P.halo.size());
    mp->d_elem_slot = device_copy(P.slot.data(), P.slot.size());
    mp->point_id = std::move(P.newid);
    mp->chain = P.chains && try_chain;
  }

  // the solid's Newmark updates visit ids 0 .. nglob_el-1: up to the largest point id of an elastic element (in
  // patch mode the plan numbers the points of elastic elements first, so the others form the tail)
  if (!getenv("SF2D_ALL_POINTS")) {
    int top = -1;
    for (int pos = 0; pos < P.phase_count[0] + P.phase_count[1]; pos++) {
      const int* src = mp->h_ibool.data() + (size_t)order[pos] * NGLL2;
      for (int p = 0; p < NGLL2; p++) top = std::max(top, mp->point_id.empty() ? src[p] : mp->point_id[src[p]]);
    }
    mp->nglob_el = top + 1;
  }
  // move the per-element device arrays
  int* d_map = device_copy(order.data(), order.size());
  const size_t ne = (size_t)nspec * NGLL2;
  float** farrs[6] = {&mp->d_xix, &mp->d_xiz, &mp->d_gammax, &mp->d_gammaz, &mp->d_kappav, &mp->d_muv};
  for (float** a : farrs) {
    float* dst = nullptr;
    check_hip(hipMalloc(&dst, ne * sizeof(float)), "hipMalloc");
    launch_permute_elements(*a, dst, d_map, nspec, mp->stream);
    check_hip(hipStreamSynchronize(mp->stream), "permute");
    device_free(*a);
    *a = dst;
  }
  device_free(d_map);
  if (mp->assembly == 0) {
    mp->d_geom = device_zeros<float4>(ne);
    mp->d_moduli = device_zeros<float2>(ne);
    mp->d_geom_elem = device_zeros<float4>((size_t)nspec);
    launch_pack_elements(mp);
    mp->no_affine = getenv("SF2D_NO_AFFINE") != nullptr;
    mp->n_affine = launch_flag_affine(mp);
  } else if (!getenv("SF2D_NO_RECT")) {
    // the per-point kernel's rectangle form: xi_z, gamma_x zero and xi_x, gamma_z constant over each element, to
    // 2^-22 of the element's largest entry (meshers leave cancellation residues)
    std::vector<float> gx(ne), gz(ne), ga(ne), gc(ne);
    check_hip(hipMemcpy(gx.data(), mp->d_xix, ne * sizeof(float), hipMemcpyDeviceToHost), "xix D2H");
    check_hip(hipMemcpy(gz.data(), mp->d_xiz, ne * sizeof(float), hipMemcpyDeviceToHost), "xiz D2H");
    check_hip(hipMemcpy(ga.data(), mp->d_gammax, ne * sizeof(float), hipMemcpyDeviceToHost), "gammax D2H");
    check_hip(hipMemcpy(gc.data(), mp->d_gammaz, ne * sizeof(float), hipMemcpyDeviceToHost), "gammaz D2H");
    std::vector<float> rect((size_t)2 * nspec);
    bool all = nspec > 0;
    for (int e = 0; e < nspec && all; e++) {
      const size_t l0 = (size_t)e * NGLL2;
      float m = 0.f;
      for (int p = 0; p < NGLL2; p++)
        m = std::max({m, std::fabs(gx[l0 + p]), std::fabs(gz[l0 + p]), std::fabs(ga[l0 + p]), std::fabs(gc[l0 + p])});
      const float tol = m * 0x1p-22f;
      for (int p = 0; p < NGLL2 && all; p++)
        all = std::fabs(gz[l0 + p]) <= tol && std::fabs(ga[l0 + p]) <= tol && std::fabs(gx[l0 + p] - gx[l0]) <= tol &&
              std::fabs(gc[l0 + p] - gc[l0]) <= tol;
      rect[2 * e] = gx[l0];
      rect[2 * e + 1] = gc[l0];
    }
    if (all) mp->d_geom_rect = device_copy(rect.data(), rect.size());
  }
  {
    // ibool in device element order and, in patch mode, internal point numbering
    std::vector<int> ib(ne);
    for (int pos = 0; pos < nspec; pos++) {
      const int* src = mp->h_ibool.data() + (size_t)order[pos] * NGLL2;
      int* dst = ib.data() + (size_t)pos * NGLL2;
      for (int p = 0; p < NGLL2; p++) dst[p] = mp->point_id.empty() ? src[p] : mp->point_id[src[p]];
    }
    check_hip(hipMemcpy(mp->d_ibool, ib.data(), ne * sizeof(int), hipMemcpyHostToDevice), "ibool H2D");
    if (mp->chain) {
      mp->h_chain_ids = std::move(ib);
      mp->h_first = std::move(P.first);
      for (int g : P.uncovered) mp->h_uncovered_points.push_back(mp->point_id.empty() ? g : mp->point_id[g]);
      mp->first_touch = mp->simulation_type == 1 && !mp->h_first.empty() && !getenv("SF2D_NO_FIRST_TOUCH");
      upload_chain_ids(mp);
    }
  }
  {
    std::vector<int> el(nspec);
    std::vector<int> h(nspec);
    check_hip(hipMemcpy(h.data(), mp->d_ispec_is_elastic, nspec * sizeof(int), hipMemcpyDeviceToHost), "D2H");
    for (int pos = 0; pos < nspec; pos++) el[pos] = h[order[pos]];
    check_hip(hipMemcpy(mp->d_ispec_is_elastic, el.data(), nspec * sizeof(int), hipMemcpyHostToDevice), "H2D");
  }
  // element indices of sources and receivers follow the new order
  if (mp->nsources_local > 0) {
    std::vector<int> v(mp->h_src_ispec.size());
    for (size_t k = 0; k < v.size(); k++) v[k] = mp->pos_of_ispec[mp->h_src_ispec[k]];
    check_hip(hipMemcpy(mp->d_ispec_selected_source, v.data(), v.size() * sizeof(int), hipMemcpyHostToDevice), "H2D");
  }
  if (mp->nrec_local > 0) {
    std::vector<int> v(mp->h_rec_ispec.size());
    for (size_t k = 0; k < v.size(); k++) v[k] = mp->pos_of_ispec[mp->h_rec_ispec[k]];
    check_hip(hipMemcpy(mp->d_ispec_selected_rec, v.data(), v.size() * sizeof(int), hipMemcpyHostToDevice), "H2D");
  }
  mp->sorted = 1;
  std::vector<int>().swap(mp->h_ibool);
}

}  // namespace

// Point arrays cross the C-ABI in the caller's numbering; on the device they are stored in the plan's
// numbering (identity unless the patch plan renumbered them).  n2 = values per point (1 or 2).
void points_to_device(const Mesh* mp, float* dev, const float* host, int n2, const char* what) {
  const size_t n = (size_t)mp->nglob * n2;
  if (mp->point_id.empty()) {
    check_hip(hipMemcpyAsync(dev, host, n * sizeof(float), hipMemcpyHostToDevice, mp->stream), what);
    check_hip(hipStreamSynchronize(mp->stream), what);
    return;
  }
  std::vector<float> tmp(n);
  for (int g = 0; g < mp->nglob; g++)
    for (int c = 0; c < n2; c++) tmp[(size_t)mp->point_id[g] * n2 + c] = host[(size_t)g * n2 + c];
  check_hip(hipMemcpyAsync(dev, tmp.data(), n * sizeof(float), hipMemcpyHostToDevice, mp->stream), what);
  check_hip(hipStreamSynchronize(mp->stream), what);
}
void points_from_device(const Mesh* mp, float* host, const float* dev, int n2, const char* what) {
  const size_t n = (size_t)mp->nglob * n2;
  if (mp->point_id.empty()) {
    check_hip(hipMemcpyAsync(host, dev, n * sizeof(float), hipMemcpyDeviceToHost, mp->stream), what);
    check_hip(hipStreamSynchronize(mp->stream), what);
    return;
  }
  std::vector<float> tmp(n);
  check_hip(hipMemcpyAsync(tmp.data(), dev, n * sizeof(float), hipMemcpyDeviceToHost, mp->stream), what);
  check_hip(hipStreamSynchronize(mp->stream), what);
  for (int g = 0; g < mp->nglob; g++)
    for (int c = 0; c < n2; c++) host[(size_t)g * n2 + c] = tmp[(size_t)mp->point_id[g] * n2 + c];
}

// per-element host array in ispec order -> device order (identity before sorting)
std::vector<float> to_device_order(const Mesh* mp, const float* a) {
  const size_t ne = (size_t)mp->nspec * NGLL2;
  std::vector<float> out(a, a + ne);
  if (mp->sorted)
    for (int pos = 0; pos < mp->nspec; pos++)
      memcpy(out.data() + (size_t)pos * NGLL2, a + (size_t)mp->ispec_of_pos[pos] * NGLL2, NGLL2 * sizeof(float));
  return out;
}

// device-order per-element array -> caller's ispec order
void to_ispec_order(const Mesh* mp, const float* dev_order, float* out) {
  for (int e = 0; e < mp->nspec; e++) {
    const int pos = mp->sorted ? mp->pos_of_ispec[e] : e;
    memcpy(out + (size_t)e * NGLL2, dev_order + (size_t)pos * NGLL2, NGLL2 * sizeof(float));
  }
}

// acc -= K u over the elements of phase ph (0 outer, 1 inner), with the handle's assembly scheme
void apply_element_forces(const Mesh* mp, int ph, const float* u, float* acc, int nshot, long shot_stride,
                          int first_touch, int ps) {
  if (mp->phase_count[ph] <= 0) return;
  if (ps > 1) {  // shot-minor batch: lanes are shots
    if (mp->assembly != 0 || !mp->chain) fail("shot-minor batches need chain patches");
    for (int c = 0; c < mp->ncolour[ph]; c++)
      launch_elastic_forces_walk(mp, mp->colour_start[ph][c], mp->colour_count[ph][c], u, acc, ps, first_touch);
    return;
  }
  if (mp->assembly == 0) {  // patches: every shot of a colour in one launch
    for (int c = 0; c < mp->ncolour[ph]; c++)
      launch_elastic_forces_patch(mp, mp->colour_start[ph][c], mp->colour_count[ph][c], u, acc, nshot, shot_stride,
                                  first_touch);
    return;
  }
  if (first_touch) fail("first-touch assembly needs the chain kernel");
  // the memory variables of viscoelastic attenuation follow the forward wavefield only
  const int att = mp->attenuation && acc == mp->d_accel;
  if (att && nshot != 1) fail("viscoelastic attenuation: one wavefield per call");
  for (int s = 0; s < nshot; s++) {
    const float* us = u + 2 * s * shot_stride;
    float* as = acc + 2 * s * shot_stride;
    if (mp->assembly == 2) {
      launch_elastic_forces(mp, mp->colour_start[ph][0], mp->phase_count[ph], 1, us, as, att);
      continue;
    }
    for (int c = 0; c < mp->ncolour[ph]; c++)
      launch_elastic_forces(mp, mp->colour_start[ph][c], mp->colour_count[ph][c], 0, us, as, att);
  }
}

}  // namespace sf2d

using namespace sf2d;

extern "C" {

// ---------------------------------------------------------------------------
void initialize_cuda_device_(int* myrank_f, int* ncuda_devices) {  // initialize_cuda.cu:50
  int count = 0;
  check_hip(hipGetDeviceCount(&count), "hipGetDeviceCount");
  if (count <= 0) fail("no HIP device found");
  *ncuda_devices = count;
  int device = *myrank_f % count;
  check_hip(hipSetDevice(device), "hipSetDevice");
  int got = -1;
  check_hip(hipGetDevice(&got), "hipGetDevice");
  if (got != device) fail("device id conflict: asked %d got %d", device, got);
  noalloc_scratch();  // the device address of get_norm_elastic_from_device_noalloc_'s scratch on this device
}

void initialize_cuda_aware_mpi_(void) {}  // initialize_cuda.cu:264 (nothing to set up on one GPU)

// ---------------------------------------------------------------------------
void prepare_constants_device_(long* Mesh_pointer, int* h_NGLLX, int* NSPEC_AB, int* NGLOB_AB, realw* h_xix,
                               realw* h_xiz, realw* h_gammax, realw* h_gammaz, realw* h_kappav, realw* h_muv,
                               int* h_ibool, int* num_interfaces_ext_mesh, int* max_nibool_interfaces_ext_mesh,
                               int* h_nibool_interfaces_ext_mesh, int* h_ibool_interfaces_ext_mesh,
                               realw* h_hprime_xx, realw* h_hprimewgll_xx, realw* h_wxgll,
                               int* STACEY_BOUNDARY_CONDITIONS, int* PML_BOUNDARY_CONDITIONS, int* h_ispec_is_inner,
                               int* nsources_local_f, realw* h_sourcearrays, realw* h_source_time_function,
                               int* NSTEP, int* h_ispec_selected_source, int* h_ispec_selected_rec_loc,
                               int* nrec_local, realw* h_cosrot, realw* h_sinrot, int* SIMULATION_TYPE, int* P_SV,
                               int* nspec_acoustic, int* nspec_elastic, int* ispec_is_acoustic,
                               int* ispec_is_elastic, int* h_myrank, int* SAVE_FORWARD, realw* h_xir_store,
                               realw* h_gammar_store, int* h_NSIGTYPE, int* h_seismotypeVec,
                               int* nlength_seismogram) {  // prepare_mesh_constants_cuda.cu:83
  (void)max_nibool_interfaces_ext_mesh;
  (void)h_nibool_interfaces_ext_mesh;
  (void)h_ibool_interfaces_ext_mesh;
  (void)h_ispec_is_inner;
  (void)ispec_is_acoustic;
  if (*h_NGLLX != NGLL) fail("NGLLX = %d, this library is built for NGLLX = %d", *h_NGLLX, NGLL);
  if (*num_interfaces_ext_mesh > 0)
    fail("MPI interfaces (%d) passed: this library drives one GPU with one mesh partition", *num_interfaces_ext_mesh);
  if (!*P_SV) fail("SH (membrane) waves are not supported by this library; use P_SV = .true.");

  Mesh* mp = new Mesh();
  *Mesh_pointer = (long)mp;
  check_hip(hipGetDevice(&mp->device), "hipGetDevice");
  check_hip(hipStreamCreate(&mp->stream), "hipStreamCreate");
  mp->myrank = *h_myrank;
  mp->nspec = *NSPEC_AB;
  mp->nglob = *NGLOB_AB;
  mp->nglob_el = mp->nglob;
  mp->simulation_type = *SIMULATION_TYPE;
  mp->save_forward = *SAVE_FORWARD;
  mp->p_sv = *P_SV;
  mp->stacey = *STACEY_BOUNDARY_CONDITIONS;
  mp->pml = *PML_BOUNDARY_CONDITIONS;
  mp->nstep = *NSTEP;
  if (mp->nspec <= 0 || mp->nglob <= 0) fail("empty mesh (nspec=%d, nglob=%d)", mp->nspec, mp->nglob);
  // the force kernels address element arrays and fields with 32-bit byte offsets
  if ((long long)mp->nspec * NGLL2 * 4 >= (1LL << 32) || (long long)mp->nglob * 8 >= (1LL << 32))
    fail("mesh too large for one handle (nspec=%d, nglob=%d): element arrays and fields must stay below 4 GiB",
         mp->nspec, mp->nglob);

  // basis: Fortran hprime_xx(i,k) = h'_k(xi_i) at h[(k-1)*5+(i-1)], hprimewgll_xx(k,i) at h[(i-1)*5+(k-1)]
  for (int i = 0; i < NGLL; i++) {
    for (int k = 0; k < NGLL; k++) {
      mp->basis.hp[i * NGLL + k] = h_hprime_xx[k * NGLL + i];
      mp->basis.hpw[k * NGLL + i] = h_hprimewgll_xx[i * NGLL + k];
    }
    mp->basis.w[i] = h_wxgll[i];
  }

  const size_t ne = (size_t)mp->nspec * NGLL2;
  mp->d_xix = device_copy(h_xix, ne);
  mp->d_xiz = device_copy(h_xiz, ne);
  mp->d_gammax = device_copy(h_gammax, ne);
  mp->d_gammaz = device_copy(h_gammaz, ne);
  mp->d_kappav = device_copy(h_kappav, ne);
  mp->d_muv = device_copy(h_muv, ne);
  mp->h_ibool = to_zero_based(h_ibool, ne, 1, mp->nglob, "ibool");
  mp->d_ibool = device_copy(mp->h_ibool.data(), ne);
  {
    const char* mode = getenv("SF2D_ASSEMBLY");
    mp->assembly = 0;
    if (mode && strcmp(mode, "colour") == 0) mp->assembly = 1;
    if (mode && strcmp(mode, "atomic") == 0) mp->assembly = 2;
    if (mode && mp->assembly == 0 && strcmp(mode, "patch") != 0) fail("SF2D_ASSEMBLY=%s: use patch, colour or atomic", mode);
  }
  std::vector<int> elastic(mp->nspec);
  for (int e = 0; e < mp->nspec; e++) elastic[e] = ispec_is_elastic[e] ? 1 : 0;
  mp->d_ispec_is_elastic = device_copy(elastic.data(), (size_t)mp->nspec);
  mp->nspec_elastic = *nspec_elastic;
  mp->nspec_acoustic = *nspec_acoustic;
  if (mp->nspec_acoustic > 0) {
    mp->h_is_acoustic.assign(mp->nspec, 0);
    int n = 0;
    for (int e = 0; e < mp->nspec; e++) {
      mp->h_is_acoustic[e] = ispec_is_acoustic[e] ? 1 : 0;
      n += mp->h_is_acoustic[e];
      if (ispec_is_acoustic[e] && ispec_is_elastic[e]) fail("element %d is both acoustic and elastic", e + 1);
    }
    if (n != mp->nspec_acoustic) fail("nspec_acoustic %d != %d flagged elements", mp->nspec_acoustic, n);
  }

  // sources
  mp->nsources_local = *nsources_local_f;
  if (mp->nsources_local > 0) {
    mp->d_sourcearrays = device_copy(h_sourcearrays, (size_t)mp->nsources_local * NGLL2 * 2);
    mp->d_stf = device_copy(h_source_time_function, (size_t)mp->nsources_local * mp->nstep);
    mp->h_src_ispec = to_zero_based(h_ispec_selected_source, mp->nsources_local, 1, mp->nspec, "ispec_selected_source");
    mp->d_ispec_selected_source = device_copy(mp->h_src_ispec.data(), mp->h_src_ispec.size());
  }

  // receivers and seismogram buffers
  mp->nrec_local = *nrec_local;
  mp->nsigtype = *h_NSIGTYPE;
  mp->h_seismotype = h_seismotypeVec;
  mp->nlength_seismogram = *nlength_seismogram;
  if (mp->nrec_local > 0) {
    mp->h_rec_ispec = to_zero_based(h_ispec_selected_rec_loc, mp->nrec_local, 1, mp->nspec, "ispec_selected_rec_loc");
    mp->d_ispec_selected_rec = device_copy(mp->h_rec_ispec.data(), mp->h_rec_ispec.size());
    mp->d_cosrot = device_copy(h_cosrot, (size_t)mp->nrec_local);
    mp->d_sinrot = device_copy(h_sinrot, (size_t)mp->nrec_local);
    mp->d_xir = device_copy(h_xir_store, (size_t)mp->nrec_local * NGLL);
    mp->d_gammar = device_copy(h_gammar_store, (size_t)mp->nrec_local * NGLL);
    mp->d_seismo = new float*[mp->nsigtype]();
    mp->h_seismo = new float*[mp->nsigtype]();
    const size_t n = (size_t)2 * mp->nrec_local * mp->nlength_seismogram;
    for (int s = 0; s < mp->nsigtype; s++) {
      if (h_seismotypeVec[s] == 0) continue;
      mp->d_seismo[s] = device_zeros<float>(n);
      check_hip(hipHostMalloc((void**)&mp->h_seismo[s], n * sizeof(float), hipHostMallocDefault), "hipHostMalloc");
    }
  }
  mp->d_scratch_u32 = device_zeros<unsigned int>(64);
  check_hip(hipDeviceSynchronize(), "prepare_constants_device");
}

void prepare_fields_elastic_device_(long* Mesh_pointer, realw* rmassx, realw* rmassz, int* num_phase_ispec_elastic,
                                    int* phase_ispec_inner_elastic, int* ispec_is_anisotropic, int* ANISOTROPY,
                                    realw* c11store, realw* c12store, realw* c13store, realw* c15store,
                                    realw* c23store, realw* c25store, realw* c33store, realw* c35store,
                                    realw* c55store, int* h_ninterface_elastic, int* h_inum_interfaces_elastic,
                                    int* ATTENUATION_VISCOELASTIC, realw* h_A_newmark_mu, realw* h_B_newmark_mu,
                                    realw* h_A_newmark_kappa,
                                    realw* h_B_newmark_kappa) {  // prepare_mesh_constants_cuda.cu:527
  (void)c12store; (void)c23store; (void)c25store; (void)h_inum_interfaces_elastic;  // AXISYM and SH terms
  Mesh* mp = handle(Mesh_pointer);
  if (*ANISOTROPY) {
    // anisotropic elements (Kernel_2_viscoelastic_impl.cu:616-853): the per-point kernel with coloured assembly
    if (*ATTENUATION_VISCOELASTIC || mp->simulation_type != 1 || mp->pml)
      fail("anisotropic elements: forward runs without attenuation or CPML only");
    if (!ispec_is_anisotropic || !c11store || !c13store || !c15store || !c33store || !c35store || !c55store)
      fail("anisotropic elements: stiffness arrays missing");
    mp->anisotropy = 1;
    mp->assembly = 1;
  }
  if (*h_ninterface_elastic > 0) fail("MPI elastic interfaces passed: one GPU, one partition only");
  if (*ATTENUATION_VISCOELASTIC) {
    // forward runs, 3 solids, in the solid elements (the reference's GPU code refuses attenuation with fluid elements,
    // prepare_gpu.f90:69-72; its CPU path runs them, and so does this library: the fluid is not attenuated)
    if (mp->simulation_type != 1) fail("viscoelastic attenuation: forward simulations (SIMULATION_TYPE=1) only");
    if (!h_A_newmark_mu || !h_B_newmark_mu || !h_A_newmark_kappa || !h_B_newmark_kappa)
      fail("viscoelastic attenuation: coefficient arrays missing");
    mp->attenuation = 1;
    // chain meshes: the column-lane chain kernel carries the memory variables in lane order; other meshes (and
    // SF2D_ATT_ATOMIC, SF2D_ROW_CHAIN, SF2D_NO_CHAIN) the per-point kernel in one launch over all elements with FP32
    // atomic assembly, which beats four coloured launches; SF2D_ASSEMBLY=colour keeps the deterministic coloured
    // order
    if (mp->assembly == 0 && (getenv("SF2D_ATT_ATOMIC") || getenv("SF2D_ROW_CHAIN") || getenv("SF2D_NO_CHAIN")))
      mp->assembly = 2;
  }
  const size_t n2 = (size_t)2 * mp->nglob;
  mp->d_displ = device_zeros<float>(n2);
  mp->d_veloc = device_zeros<float>(n2);
  mp->d_accel = device_zeros<float>(n2);
  const int np = *num_phase_ispec_elastic;
  mp->num_phase_ispec_elastic = np;
  sort_elements_by_colour(mp, phase_ispec_inner_elastic, np);
  if (mp->anisotropy) {
    // stiffness per point in device element order: the caller's c_ij (Fortran (NGLLX, NGLLZ, nspec)) for the
    // anisotropic elements, the isotropic equivalent of kappa and mu for the others
    const size_t ne = (size_t)mp->nspec * NGLL2;
    std::vector<float> kap(ne), mu(ne);
    check_hip(hipMemcpy(kap.data(), mp->d_kappav, ne * sizeof(float), hipMemcpyDeviceToHost), "kappa D2H");
    check_hip(hipMemcpy(mu.data(), mp->d_muv, ne * sizeof(float), hipMemcpyDeviceToHost), "mu D2H");
    std::vector<float4> c(2 * ne);
    for (int pos = 0; pos < mp->nspec; pos++) {
      const int e = mp->sorted ? mp->ispec_of_pos[pos] : pos;
      for (int q = 0; q < NGLL2; q++) {
        const size_t l = (size_t)pos * NGLL2 + q, h = (size_t)e * NGLL2 + q;
        if (ispec_is_anisotropic[e]) {
          c[2 * l] = make_float4(c11store[h], c13store[h], c15store[h], c33store[h]);
          c[2 * l + 1] = make_float4(c35store[h], c55store[h], 0.f, 0.f);
        } else {
          c[2 * l] = make_float4(kap[l] + mu[l], kap[l] - mu[l], 0.f, kap[l] + mu[l]);
          c[2 * l + 1] = make_float4(0.f, mu[l], 0.f, 0.f);
        }
      }
    }
    mp->d_aniso = device_copy(c.data(), c.size());
  }
  if (mp->attenuation) {
    // host arrays A(N_SLS, NGLLX, NGLLZ, nspec) (solid fastest, ispec order): the distinct coefficient sets of
    // the points go to a table (a few per model: one per distinct pair of quality factors), each point keeps a
    // 4-byte index in device element order -- 44 B less per point and step than per-point coefficients
    const size_t ne = (size_t)mp->nspec * NGLL2;
    const realw* src[4] = {h_A_newmark_kappa, h_B_newmark_kappa, h_A_newmark_mu, h_B_newmark_mu};
    std::map<std::array<float, 12>, unsigned int> ids;
    std::vector<float4> table;
    std::vector<unsigned int> idx(ne);
    std::array<float, 12> last_c{};
    unsigned int last_id = 0;
    for (int e = 0; e < mp->nspec; e++) {
      const size_t row = (size_t)mp->pos_of_ispec[e] * NGLL2;
      for (int p = 0; p < NGLL2; p++) {
        std::array<float, 12> c;
        for (int k = 0; k < 3; k++)
          for (int w = 0; w < 4; w++) c[4 * k + w] = src[w][((size_t)e * NGLL2 + p) * 3 + k];
        if (!table.empty() && c == last_c) {  // neighbouring points mostly share their set (no NaN: NaN != NaN)
          idx[row + p] = last_id;
          continue;
        }
        // A = exp(-dt / (2 tau_sigma)) lies in [0, 1] and B is finite (prepare_attenuation.f90:377-389); the
        // reference leaves the arrays unset at points without attenuation (Qkappa, Qmu >= 9999, :347), so a
        // caller must zero them there -- anything else would make the memory variables grow without bound
        for (int k = 0; k < 3; k++)
          for (int w = 0; w < 4; w++) {
            const float v = c[4 * k + w];
            if (!std::isfinite(v) || ((w == 0 || w == 2) && (v < 0.f || v > 1.f)))
              fail("viscoelastic attenuation: invalid %s_newmark_%s coefficient %g (solid %d, point %d of element %d);"
                   " A must lie in [0, 1] and B be finite (zero both where Qkappa and Qmu are 9999)",
                   (w & 1) ? "B" : "A", w < 2 ? "kappa" : "mu", (double)v, k + 1, p + 1, e + 1);
          }
        auto it = ids.find(c);
        if (it == ids.end()) {
          it = ids.emplace(c, (unsigned int)(table.size() / 3)).first;
          for (int k = 0; k < 3; k++) table.push_back(make_float4(c[4 * k], c[4 * k + 1], c[4 * k + 2], c[4 * k + 3]));
        }
        idx[row + p] = it->second;
        last_c = c;
        last_id = it->second;
      }
    }
    mp->att_ntable = (int)(table.size() / 3);
    mp->d_att_table = device_copy(table.data(), table.size());
    if (mp->chain) {
      // the column-lane kernel's lane order: the block of a chain of ne elements at position e0 holds point
      // (k, i, j) at e0 * 25 + j * 5 ne + 5 k + i
      std::vector<unsigned int> lane(idx);
      for (int q = 0; q < mp->npatch; q++) {
        const int e0 = mp->h_patch_desc[4 * q], nel = mp->h_patch_desc[4 * q + 3] & 0xf;
        for (int k = 0; k < nel; k++)
          for (int p = 0; p < NGLL2; p++)
            lane[(size_t)e0 * NGLL2 + (size_t)(p / NGLL) * NGLL * nel + NGLL * k + p % NGLL] =
                idx[(size_t)(e0 + k) * NGLL2 + p];
      }
      idx.swap(lane);
    }
    mp->d_att_idx = device_copy(idx.data(), ne);
    mp->d_att_e = device_zeros<float>(12 * ne);  // [3][ne] float4: 9 memory variables, 3 previous strains
  }
  mp->d_rmassx = device_zeros<float>((size_t)mp->nglob);
  mp->d_rmassz = device_zeros<float>((size_t)mp->nglob);
  points_to_device(mp, mp->d_rmassx, rmassx, 1, "rmassx H2D");
  points_to_device(mp, mp->d_rmassz, rmassz, 1, "rmassz H2D");
  // the two components' inverse masses differ only where the Stacey terms of the mass differ (absorbing edges);
  // without them the Newmark kernels read one array
  mp->same_mass = !getenv("SF2D_TWO_MASSES") &&
                  std::equal(rmassx, rmassx + mp->nglob, rmassz, [](float a, float b) { return a == b; });
  check_hip(hipDeviceSynchronize(), "prepare_fields_elastic_device");
}

void prepare_fields_elastic_adj_dev_(long* Mesh_pointer, int* size_f, int* APPROXIMATE_HESS_KL,
                                     int* ATTENUATION_VISCOELASTIC,
                                     int* NO_BACKWARD_RECONSTRUCTION) {  // prepare_mesh_constants_cuda.cu:711
  Mesh* mp = handle(Mesh_pointer);
  if (*APPROXIMATE_HESS_KL) fail("approximate Hessian kernels are not supported on the GPU (as in the reference)");
  if (*ATTENUATION_VISCOELASTIC)
    fail("viscoelastic attenuation with adjoint simulations is not supported (as in the reference's GPU code)");
  if (*NO_BACKWARD_RECONSTRUCTION) fail("NO_BACKWARD_RECONSTRUCTION is not supported by this library");
  if (*size_f != 2 * mp->nglob) fail("prepare_fields_elastic_adj_dev: size %d != 2*nglob", *size_f);
  const size_t n2 = (size_t)2 * mp->nglob;
  mp->d_b_displ = device_zeros<float>(n2);
  mp->d_b_veloc = device_zeros<float>(n2);
  mp->d_b_accel = device_zeros<float>(n2);
  const size_t ne = (size_t)mp->nspec * NGLL2;
  mp->d_rho_kl = device_zeros<float>(ne);
  mp->d_mu_kl = device_zeros<float>(ne);
  mp->d_kappa_kl = device_zeros<float>(ne);
}

void prepare_stacey_device_(long* Mesh_pointer, int* ACOUSTIC_SIMULATION, int* ELASTIC_SIMULATION, realw* rho_vp,
                            realw* rho_vs, int* h_nspec_bottom, int* h_nspec_left, int* h_nspec_right,
                            int* h_nspec_top, int* h_abs_boundary_ispec, int* h_abs_boundary_ij,
                            realw* h_abs_boundary_normal, realw* h_abs_boundary_jacobian1Dw,
                            int* h_num_abs_boundary_faces, int* h_edge_abs, int* h_ib_bottom, int* h_ib_left,
                            int* h_ib_right, int* h_ib_top) {  // prepare_mesh_constants_cuda.cu:926
  Mesh* mp = handle(Mesh_pointer);
  if (!mp->stacey) fail("Stacey flag inconsistent with prepare_stacey_device call");
  const int nf = *h_num_abs_boundary_faces;
  mp->num_abs = nf;
  mp->nspec_bottom = *h_nspec_bottom;
  mp->nspec_left = *h_nspec_left;
  mp->nspec_right = *h_nspec_right;
  mp->nspec_top = *h_nspec_top;
  if (nf <= 0) return;
  std::vector<int> isp = to_zero_based(h_abs_boundary_ispec, nf, 1, mp->nspec, "abs_boundary_ispec");
  if (mp->sorted)
    for (auto& e : isp) e = mp->pos_of_ispec[e];
  std::vector<int> ij((size_t)nf * NGLL * 2);
  for (size_t k = 0; k < ij.size(); k++) {
    int v = h_abs_boundary_ij[k];
    if (v < 1 || v > NGLL + 1) fail("abs_boundary_ij: value %d out of range", v);
    ij[k] = (v == NGLL + 1) ? -1 : v - 1;  // NGLL+1 marks a point that the edge does not own
  }
  // after marking, a skipped point has -1 in one coordinate: make both negative
  for (int k = 0; k < nf * NGLL; k++)
    if (ij[2 * k] < 0 || ij[2 * k + 1] < 0) ij[2 * k] = ij[2 * k + 1] = -1;
  std::vector<int> ib(nf);
  const int counts[5] = {0, mp->nspec_bottom, mp->nspec_right, mp->nspec_top, mp->nspec_left};
  for (int f = 0; f < nf; f++) {
    int side = h_edge_abs[f];
    int v = side == 1 ? h_ib_bottom[f] : side == 2 ? h_ib_right[f] : side == 3 ? h_ib_top[f] : side == 4 ? h_ib_left[f] : -1;
    if (side < 1 || side > 4) fail("edge_abs(%d) = %d is not 1..4", f + 1, side);
    if (v < 1 || v > counts[side]) fail("ib index %d of face %d outside 1..%d", v, f + 1, counts[side]);
    ib[f] = v - 1;
  }
  if (*ACOUSTIC_SIMULATION && mp->nspec_acoustic > 0) {
    // the fluid's points keep the caller's element numbering until the first acoustic step resolves them
    std::vector<int> isp0 = to_zero_based(h_abs_boundary_ispec, nf, 1, mp->nspec, "abs_boundary_ispec");
    acoustic_note_stacey(mp, nf, isp0.data(), ij.data(), h_abs_boundary_jacobian1Dw, rho_vp, h_edge_abs, ib.data(),
                         counts);
  }
  if (mp->chain) {  // points of the absorbing edges, for first-touch assembly
    for (int f = 0; f < nf; f++)
      for (int k = 0; k < NGLL; k++) {
        const int ii = ij[((size_t)f * NGLL + k) * 2], jj = ij[((size_t)f * NGLL + k) * 2 + 1];
        if (ii >= 0 && jj >= 0) mp->h_stacey_points.push_back(mp->h_chain_ids[(size_t)isp[f] * NGLL2 + jj * NGLL + ii]);
      }
    upload_chain_ids(mp);
  }
  mp->d_abs_ispec = device_copy(isp.data(), isp.size());
  mp->d_abs_ij = device_copy(ij.data(), ij.size());
  mp->d_abs_normal = device_copy(h_abs_boundary_normal, (size_t)nf * NGLL * 2);
  mp->d_abs_jacw = device_copy(h_abs_boundary_jacobian1Dw, (size_t)nf * NGLL);
  mp->d_edge_abs = device_copy(h_edge_abs, (size_t)nf);
  mp->d_ib = device_copy(ib.data(), ib.size());
  if (*ELASTIC_SIMULATION) {
    std::vector<float> rvp = to_device_order(mp, rho_vp), rvs = to_device_order(mp, rho_vs);
    mp->d_rho_vp = device_copy(rvp.data(), rvp.size());
    mp->d_rho_vs = device_copy(rvs.data(), rvs.size());
    if (mp->simulation_type == 3 || (mp->simulation_type == 1 && mp->save_forward)) {
      for (int side = 1; side <= 4; side++)
        if (counts[side] > 0) mp->d_b_absorb[side] = device_zeros<float>((size_t)counts[side] * NGLL * 2);
    }
  }
}

void prepare_sim2_or_3_const_device_(long* Mesh_pointer, int* nadj_rec_local, realw* h_source_adjoint,
                                     int* NSTEP) {  // prepare_mesh_constants_cuda.cu:815
  Mesh* mp = handle(Mesh_pointer);
  device_free(mp->d_source_adjoint);  // a new adjoint run may replace the sources
  mp->nadj_rec_local = *nadj_rec_local;
  if (mp->nadj_rec_local > 0) {
    if (mp->nadj_rec_local != mp->nrec_local) fail("nadj_rec_local %d != nrec_local %d", mp->nadj_rec_local, mp->nrec_local);
    mp->d_source_adjoint = device_copy(h_source_adjoint, (size_t)2 * (*NSTEP) * mp->nadj_rec_local);
  }
}

void prepare_cleanup_device_(long* Mesh_pointer, int* ACOUSTIC_SIMULATION, int* ELASTIC_SIMULATION, int* ANISOTROPY,
                             int* APPROXIMATE_HESS_KL, int* ATTENUATION_VISCOACOUSTIC,
                             int* ATTENUATION_VISCOELASTIC, int* NO_BACKWARD_RECONSTRUCTION,
                             realw* h_no_backward_acoustic_buffer) {  // prepare_mesh_constants_cuda.cu:1139
  (void)ACOUSTIC_SIMULATION; (void)ELASTIC_SIMULATION; (void)ANISOTROPY; (void)APPROXIMATE_HESS_KL;
  (void)ATTENUATION_VISCOACOUSTIC; (void)ATTENUATION_VISCOELASTIC; (void)NO_BACKWARD_RECONSTRUCTION;
  (void)h_no_backward_acoustic_buffer;
  Mesh* mp = handle(Mesh_pointer);
  (void)hipStreamSynchronize(mp->stream);
  acoustic_release(mp);
  pml_release(mp);
  device_free(mp->d_ibool); device_free(mp->d_xix); device_free(mp->d_xiz); device_free(mp->d_gammax);
  device_free(mp->d_gammaz); device_free(mp->d_kappav); device_free(mp->d_muv); device_free(mp->d_att_idx); device_free(mp->d_att_table); device_free(mp->d_att_e); device_free(mp->d_att_old); device_free(mp->d_ispec_is_elastic);
  device_free(mp->d_displ); device_free(mp->d_veloc); device_free(mp->d_accel);
  device_free(mp->d_b_displ); device_free(mp->d_b_veloc); device_free(mp->d_b_accel);
  device_free(mp->d_rmassx); device_free(mp->d_rmassz);
  device_free(mp->d_sourcearrays); device_free(mp->d_stf); device_free(mp->d_ispec_selected_source);
  device_free(mp->d_xir); device_free(mp->d_gammar); device_free(mp->d_cosrot); device_free(mp->d_sinrot);
  device_free(mp->d_ispec_selected_rec);
  if (mp->d_seismo) {
    for (int s = 0; s < mp->nsigtype; s++) {
      device_free(mp->d_seismo[s]);
      if (mp->h_seismo[s]) (void)hipHostFree(mp->h_seismo[s]);
    }
    delete[] mp->d_seismo;
    delete[] mp->h_seismo;
  }
  device_free(mp->d_source_adjoint);
  device_free(mp->d_abs_ispec); device_free(mp->d_abs_ij); device_free(mp->d_abs_normal); device_free(mp->d_abs_jacw);
  device_free(mp->d_edge_abs); device_free(mp->d_ib); device_free(mp->d_rho_vp); device_free(mp->d_rho_vs);
  for (int s = 0; s < 5; s++) device_free(mp->d_b_absorb[s]);
  device_free(mp->d_rho_kl); device_free(mp->d_mu_kl); device_free(mp->d_kappa_kl);
  device_free(mp->d_scratch_u32);
  device_free(mp->d_patch_desc); device_free(mp->d_halo_nodes);
  device_free(mp->d_elem_slot);
  device_free(mp->d_geom);
  device_free(mp->d_moduli);
  device_free(mp->d_geom_elem);
  device_free(mp->d_geom_rect);
  device_free(mp->d_aniso);
  device_free(mp->d_chain_ids);
  device_free(mp->d_col_ids);
  device_free(mp->d_col_mod);
  device_free(mp->d_col_rows);
  device_free(mp->d_special);
  if (mp->grad) sf2d_grad_release(Mesh_pointer);
  if (mp->batch) sf2d_batch_release(Mesh_pointer);
  (void)hipStreamDestroy(mp->stream);
  delete mp;
  *Mesh_pointer = 0;
}

// ---------------------------------------------------------------------------
// transfers (transfer_fields_cuda.cu)
// First-touch assembly leaves accel incomplete from update_displacement_cuda until the last force launch of the
// step; the reference's sequence never reads it there.
static void check_accel_complete(const Mesh* mp, const char* what) {
  if (mp->accel_partial)
    fail("%s: accel is read between update_displacement_cuda and compute_forces_viscoelastic_cuda (iphase 2); "
         "set SF2D_NO_FIRST_TOUCH for this call sequence", what);
}

static void copy_field(float* dst_dev, const realw* src_host, int size, Mesh* mp, const char* what) {
  if (size != 2 * mp->nglob) fail("%s: size %d != 2*nglob = %d", what, size, 2 * mp->nglob);
  if (!dst_dev) fail("%s: field not allocated on the device", what);
  points_to_device(mp, dst_dev, src_host, 2, what);
}
static void fetch_field(realw* dst_host, const float* src_dev, int size, Mesh* mp, const char* what) {
  if (size != 2 * mp->nglob) fail("%s: size %d != 2*nglob = %d", what, size, 2 * mp->nglob);
  if (!src_dev) fail("%s: field not allocated on the device", what);
  points_from_device(mp, dst_host, src_dev, 2, what);
}

void transfer_fields_el_to_device_(int* size, realw* displ, realw* veloc, realw* accel, long* Mesh_pointer) {
  Mesh* mp = handle(Mesh_pointer);
  flush_correctors(mp);
  mp->accel_partial = 0;
  copy_field(mp->d_displ, displ, *size, mp, "transfer_fields_el_to_device");
  copy_field(mp->d_veloc, veloc, *size, mp, "transfer_fields_el_to_device");
  copy_field(mp->d_accel, accel, *size, mp, "transfer_fields_el_to_device");
}
void transfer_fields_el_from_device_(int* size, realw* displ, realw* veloc, realw* accel, long* Mesh_pointer) {
  Mesh* mp = handle(Mesh_pointer);
  flush_correctors(mp);
  check_accel_complete(mp, "transfer_fields_el_from_device");
  fetch_field(displ, mp->d_displ, *size, mp, "transfer_fields_el_from_device");
  fetch_field(veloc, mp->d_veloc, *size, mp, "transfer_fields_el_from_device");
  fetch_field(accel, mp->d_accel, *size, mp, "transfer_fields_el_from_device");
}
void transfer_b_fields_to_device_(int* size, realw* b_displ, realw* b_veloc, realw* b_accel, long* Mesh_pointer) {
  Mesh* mp = handle(Mesh_pointer);
  flush_correctors(mp);
  copy_field(mp->d_b_displ, b_displ, *size, mp, "transfer_b_fields_to_device");
  copy_field(mp->d_b_veloc, b_veloc, *size, mp, "transfer_b_fields_to_device");
  copy_field(mp->d_b_accel, b_accel, *size, mp, "transfer_b_fields_to_device");
}
void transfer_b_fields_from_device_(int* size, realw* b_displ, realw* b_veloc, realw* b_accel, long* Mesh_pointer) {
  Mesh* mp = handle(Mesh_pointer);
  flush_correctors(mp);
  fetch_field(b_displ, mp->d_b_displ, *size, mp, "transfer_b_fields_from_device");
  fetch_field(b_veloc, mp->d_b_veloc, *size, mp, "transfer_b_fields_from_device");
  fetch_field(b_accel, mp->d_b_accel, *size, mp, "transfer_b_fields_from_device");
}
void transfer_accel_to_device_(int* size, realw* accel, long* Mesh_pointer) {
  Mesh* mp = handle(Mesh_pointer);
  flush_correctors(mp);
  mp->accel_partial = 0;
  copy_field(mp->d_accel, accel, *size, mp, "transfer_accel_to_device");
}
void transfer_accel_from_device_(int* size, realw* accel, long* Mesh_pointer) {
  Mesh* mp = handle(Mesh_pointer);
  flush_correctors(mp);
  check_accel_complete(mp, "transfer_accel_from_device");
  fetch_field(accel, mp->d_accel, *size, mp, "transfer_accel_from_device");
}
void transfer_b_accel_from_device_(int* size, realw* b_accel, long* Mesh_pointer) {
  Mesh* mp = handle(Mesh_pointer);
  flush_correctors(mp);
  fetch_field(b_accel, mp->d_b_accel, *size, mp, "transfer_b_accel_from_device");
}
void transfer_b_displ_from_device_(int* size, realw* displ, long* Mesh_pointer) {
  Mesh* mp = handle(Mesh_pointer);
  fetch_field(displ, mp->d_b_displ, *size, mp, "transfer_b_displ_from_device");
}
void transfer_displ_from_device_(int* size, realw* displ, long* Mesh_pointer) {
  Mesh* mp = handle(Mesh_pointer);
  fetch_field(displ, mp->d_displ, *size, mp, "transfer_displ_from_device");
}
void transfer_kernels_el_to_host_(long* Mesh_pointer, realw* h_rho_kl, realw* h_mu_kl, realw* h_kappa_kl,
                                  int* NSPEC_AB) {  // transfer_fields_cuda.cu:194
  Mesh* mp = handle(Mesh_pointer);
  if (!mp->d_rho_kl) fail("transfer_kernels_el_to_host: kernels not allocated (SIMULATION_TYPE /= 3)");
  const size_t n = (size_t)(*NSPEC_AB) * NGLL2;
  if (*NSPEC_AB != mp->nspec) fail("transfer_kernels_el_to_host: NSPEC_AB mismatch");
  check_hip(hipStreamSynchronize(mp->stream), "transfer_kernels_el_to_host");
  std::vector<float> tmp(n);
  float* outs[3] = {h_rho_kl, h_mu_kl, h_kappa_kl};
  const float* devs[3] = {mp->d_rho_kl, mp->d_mu_kl, mp->d_kappa_kl};
  for (int q = 0; q < 3; q++) {
    check_hip(hipMemcpy(tmp.data(), devs[q], n * sizeof(float), hipMemcpyDeviceToHost), "kernels D2H");
    to_ispec_order(mp, tmp.data(), outs[q]);
  }
}
void transfer_compute_kernel_answers_from_device_(long* Mesh_pointer, realw* rho_kl, int* size_rho, realw* mu_kl,
                                                  int* size_mu, realw* kappa_kl,
                                                  int* size_kappa) {  // transfer_fields_cuda.cu:598
  Mesh* mp = handle(Mesh_pointer);
  if (!mp->d_rho_kl) fail("transfer_compute_kernel_answers_from_device: kernels not allocated");
  const size_t n = (size_t)mp->nspec * NGLL2;
  if ((size_t)*size_rho != n || (size_t)*size_mu != n || (size_t)*size_kappa != n)
    fail("transfer_compute_kernel_answers_from_device: sizes must be NGLL2*NSPEC_AB = %zu", n);
  int nspec = mp->nspec;
  transfer_kernels_el_to_host_(Mesh_pointer, rho_kl, mu_kl, kappa_kl, &nspec);
}
void transfer_compute_kernel_fields_from_device_(
    long* Mesh_pointer, realw* accel, int* size_accel, realw* b_displ, int* size_b_displ, realw* epsilondev_xx,
    realw* epsilondev_yy, realw* epsilondev_xy, realw* epsilondev_xz, realw* epsilondev_yz, int* size_epsilondev,
    realw* b_epsilondev_xx, realw* b_epsilondev_yy, realw* b_epsilondev_xy, realw* b_epsilondev_xz,
    realw* b_epsilondev_yz, int* size_b_epsilondev, realw* rho_kl, int* size_rho, realw* mu_kl, int* size_mu,
    realw* kappa_kl, int* size_kappa, realw* epsilon_trace_over_3, realw* b_epsilon_trace_over_3,
    int* size_epsilon_trace_over_3) {  // transfer_fields_cuda.cu:617 (3D leftover; strain arrays are not kept)
  (void)epsilondev_xx; (void)epsilondev_yy; (void)epsilondev_xy; (void)epsilondev_xz; (void)epsilondev_yz;
  (void)size_epsilondev; (void)b_epsilondev_xx; (void)b_epsilondev_yy; (void)b_epsilondev_xy;
  (void)b_epsilondev_xz; (void)b_epsilondev_yz; (void)size_b_epsilondev; (void)epsilon_trace_over_3;
  (void)b_epsilon_trace_over_3; (void)size_epsilon_trace_over_3;
  Mesh* mp = handle(Mesh_pointer);
  flush_correctors(mp);
  fetch_field(accel, mp->d_accel, *size_accel, mp, "transfer_compute_kernel_fields_from_device");
  fetch_field(b_displ, mp->d_b_displ, *size_b_displ, mp, "transfer_compute_kernel_fields_from_device");
  transfer_compute_kernel_answers_from_device_(Mesh_pointer, rho_kl, size_rho, mu_kl, size_mu, kappa_kl, size_kappa);
}

// ---------------------------------------------------------------------------
// time step
void update_displacement_cuda_(long* Mesh_pointer, realw* deltat_F, realw* deltatsqover2_F, realw* deltatover2_F,
                               realw* b_deltat_F, realw* b_deltatsqover2_F,
                               realw* b_deltatover2_F) {  // update_displacement_cuda.cu:52
  Mesh* mp = handle(Mesh_pointer);
  const int keep = mp->first_touch;  // first-touch assembly: the force kernels define accel
  // CPML: the previous-step field of the layer's points, from the fields before the predictor
  pml_elastic_before_predictor(mp, *deltatsqover2_F);
  if (mp->pending_corrector[0])
    launch_newmark_corrector_predictor(mp->d_displ, mp->d_veloc, mp->d_accel, mp->d_rmassx, rmass_z(mp), mp->nglob_el,
                                       mp->pending_dto2[0], *deltat_F, *deltatsqover2_F, *deltatover2_F, mp->stream,
                                       1, 0, !keep);
  else
    launch_newmark_predictor(mp->d_displ, mp->d_veloc, mp->d_accel, mp->nglob_el, *deltat_F, *deltatsqover2_F,
                             *deltatover2_F, mp->stream, 1, 0, !keep);
  if (keep) {
    launch_zero_points(mp, mp->d_accel);
    mp->accel_partial = 1;
  }
  if (mp->simulation_type == 3) {
    if (mp->pending_corrector[1])
      launch_newmark_corrector_predictor(mp->d_b_displ, mp->d_b_veloc, mp->d_b_accel, mp->d_rmassx, rmass_z(mp),
                                         mp->nglob_el, mp->pending_dto2[1], *b_deltat_F, *b_deltatsqover2_F,
                                         *b_deltatover2_F, mp->stream);
    else
      launch_newmark_predictor(mp->d_b_displ, mp->d_b_veloc, mp->d_b_accel, mp->nglob_el, *b_deltat_F, *b_deltatsqover2_F,
                               *b_deltatover2_F, mp->stream);
  } else if (mp->pending_corrector[1]) {
    flush_correctors(mp);
  }
  mp->pending_corrector[0] = mp->pending_corrector[1] = 0;
}

void compute_forces_viscoelastic_cuda_(long* Mesh_pointer, int* iphase, realw* deltat, int* nspec_outer_elastic,
                                       int* nspec_inner_elastic, int* ANISOTROPY, int* ATTENUATION_VISCOELASTIC,
                                       int* compute_wavefield_1,
                                       int* compute_wavefield_2) {  // compute_forces_viscoelastic_cuda.cu:327
  (void)deltat;
  Mesh* mp = handle(Mesh_pointer);
  flush_correctors(mp);
  if ((*ANISOTROPY != 0) != (mp->anisotropy != 0))
    fail("compute_forces_viscoelastic_cuda: ANISOTROPY differs from prepare_fields_elastic_device");
  if ((*ATTENUATION_VISCOELASTIC != 0) != (mp->attenuation != 0))
    fail("compute_forces_viscoelastic_cuda: ATTENUATION_VISCOELASTIC differs from prepare_fields_elastic_device");
  const int ph = (*iphase == 1) ? 0 : 1;
  const int nelem = (*iphase == 1) ? *nspec_outer_elastic : *nspec_inner_elastic;
  const int first_touch = mp->accel_partial && *compute_wavefield_1;
  if (ph == 1) mp->accel_partial = 0;  // the last force launches of the step complete accel
  if (nelem <= 0) return;
  if (!mp->sorted) fail("compute_forces: prepare_fields_elastic_device was not called");
  if (nelem != mp->phase_count[ph])
    fail("compute_forces: iphase %d has %d elements, the prepared phase list has %d", *iphase, nelem, mp->phase_count[ph]);
  const bool w1 = *compute_wavefield_1 != 0, w2 = *compute_wavefield_2 != 0;
  if (w2 && !mp->d_b_displ) fail("compute_forces: backward wavefield not allocated");
  if (w2 && mp->pml_state) fail("compute_forces: backward wavefields with CPML are not supported");
  if (w1) {
    apply_element_forces(mp, ph, mp->d_displ, mp->d_accel, 1, 0, first_touch);
    pml_elastic_forces(mp, ph);  // the CPML layer's elements, after the main launches of the phase
  }
  if (w2) apply_element_forces(mp, ph, mp->d_b_displ, mp->d_b_accel);
}

void compute_stacey_viscoelastic_cuda_(long* Mesh_pointer, int* iphasef, realw* h_b_absorb_elastic_left,
                                       realw* h_b_absorb_elastic_right, realw* h_b_absorb_elastic_top,
                                       realw* h_b_absorb_elastic_bottom, int* compute_wavefield_1,
                                       int* compute_wavefield_2,
                                       int* UNDO_ATTENUATION_AND_OR_PML) {  // compute_stacey_viscoelastic_cuda.cu:40
  Mesh* mp = handle(Mesh_pointer);
  flush_correctors(mp);
  if (mp->num_abs == 0 || *iphasef != 1) return;
  const int read_abs = (mp->simulation_type == 3 && !*UNDO_ATTENUATION_AND_OR_PML) ? 1 : 0;
  const int write_abs = (mp->simulation_type == 1 && mp->save_forward && !*UNDO_ATTENUATION_AND_OR_PML) ? 1 : 0;
  realw* host[5] = {nullptr, h_b_absorb_elastic_bottom, h_b_absorb_elastic_right, h_b_absorb_elastic_top,
                    h_b_absorb_elastic_left};
  const int counts[5] = {0, mp->nspec_bottom, mp->nspec_right, mp->nspec_top, mp->nspec_left};
  if (read_abs) {
    for (int s = 1; s <= 4; s++)
      if (counts[s] > 0)
        check_hip(hipMemcpyAsync(mp->d_b_absorb[s], host[s], sizeof(float) * 2 * NGLL * counts[s],
                                 hipMemcpyHostToDevice, mp->stream), "b_absorb H2D");
  }
  if (*compute_wavefield_1) launch_stacey(mp, mp->d_veloc, mp->d_accel, write_abs);
  if (*compute_wavefield_2) launch_stacey_backward(mp, mp->d_b_accel);
  if (write_abs) {
    for (int s = 1; s <= 4; s++)
      if (counts[s] > 0)
        check_hip(hipMemcpyAsync(host[s], mp->d_b_absorb[s], sizeof(float) * 2 * NGLL * counts[s],
                                 hipMemcpyDeviceToHost, mp->stream), "b_absorb D2H");
    check_hip(hipStreamSynchronize(mp->stream), "compute_stacey_viscoelastic_cuda");
  } else if (read_abs) {
    check_hip(hipStreamSynchronize(mp->stream), "compute_stacey_viscoelastic_cuda");  // host slice is reused
  }
}

void compute_add_sources_el_cuda_(long* Mesh_pointer, int* iphasef,
                                  int* itf) {  // compute_add_sources_viscoelastic_cuda.cu:46
  Mesh* mp = handle(Mesh_pointer);
  flush_correctors(mp);
  if (mp->nsources_local == 0 || *iphasef != 1) return;
  if (*itf < 1 || *itf > mp->nstep) fail("compute_add_sources_el_cuda: it=%d outside 1..%d", *itf, mp->nstep);
  launch_sources(mp, mp->d_accel, *itf - 1);
}

void compute_add_sources_el_s3_cuda_(long* Mesh_pointer, int* iphasef,
                                     int* itf) {  // compute_add_sources_viscoelastic_cuda.cu:86
  Mesh* mp = handle(Mesh_pointer);
  flush_correctors(mp);
  if (mp->nsources_local == 0 || *iphasef != 1) return;
  if (*itf < 1 || *itf > mp->nstep) fail("compute_add_sources_el_s3_cuda: it=%d outside 1..%d", *itf, mp->nstep);
  if (!mp->d_b_accel) fail("compute_add_sources_el_s3_cuda: backward wavefield not allocated");
  launch_sources(mp, mp->d_b_accel, *itf - 1);
}

void add_sources_el_sim_type_2_or_3_(long* Mesh_pointer, int* iphasef, int* itf, int* nadj_rec_local,
                                     int* NSTEP) {  // compute_add_sources_viscoelastic_cuda.cu:131
  Mesh* mp = handle(Mesh_pointer);
  flush_correctors(mp);
  if (*nadj_rec_local != mp->nadj_rec_local) fail("add_sources_el_sim_type_2_or_3: nadj_rec_local not equal");
  if (mp->nadj_rec_local == 0 || *iphasef != 1) return;
  if (*NSTEP != mp->nstep) fail("add_sources_el_sim_type_2_or_3: NSTEP mismatch");
  if (*itf < 1 || *itf > mp->nstep) fail("add_sources_el_sim_type_2_or_3: it=%d outside 1..%d", *itf, mp->nstep);
  launch_adjoint_sources(mp, mp->d_accel, *itf - 1);
}

void kernel_3_a_cuda_(long* Mesh_pointer, realw* deltatover2_F, realw* b_deltatover2_F, int* compute_wavefield_1,
                      int* compute_wavefield_2) {  // update_displacement_cuda.cu:221
  Mesh* mp = handle(Mesh_pointer);
  flush_correctors(mp);
  check_accel_complete(mp, "kernel_3_a_cuda");
  // recorded here, applied by the next predictor (or by the first entry point that reads veloc/accel)
  if (*compute_wavefield_1) {
    mp->pending_corrector[0] = 1;
    mp->pending_dto2[0] = *deltatover2_F;
  }
  if (*compute_wavefield_2) {
    if (!mp->d_b_veloc) fail("kernel_3_a_cuda: backward wavefield not allocated");
    mp->pending_corrector[1] = 1;
    mp->pending_dto2[1] = *b_deltatover2_F;
  }
}

void kernel_3_b_cuda_(long* Mesh_pointer, realw* deltatover2_F,
                      realw* b_deltatover2_F) {  // update_displacement_cuda.cu:273 (velocity-only corrector)
  Mesh* mp = handle(Mesh_pointer);
  flush_correctors(mp);
  launch_newmark_corrector(mp->d_veloc, mp->d_accel, nullptr, nullptr, mp->nglob_el, *deltatover2_F, 0, mp->stream);
  if (mp->simulation_type == 3)
    launch_newmark_corrector(mp->d_b_veloc, mp->d_b_accel, nullptr, nullptr, mp->nglob_el, *b_deltatover2_F, 0,
                             mp->stream);
}

void compute_seismograms_cuda_(long* Mesh_pointer_f, int* i_sigf, double* sisux, double* sisuz,
                               int* seismo_currentf, int* nlength_seismogramf, int* ELASTIC_SIMULATION,
                               int* ACOUSTIC_SIMULATION, int* USE_TRICK_FOR_BETTER_PRESSURE,
                               int* ATTENUATION_VISCOELASTIC, int* itf,
                               int* it_endf) {  // write_seismograms_cuda.cu:41
  (void)ELASTIC_SIMULATION; (void)ACOUSTIC_SIMULATION; (void)ATTENUATION_VISCOELASTIC;
  Mesh* mp = handle(Mesh_pointer_f);
  if (mp->nrec_local == 0) return;
  const int i_sig = *i_sigf - 1;
  const int cur = *seismo_currentf - 1;
  const int nlength = *nlength_seismogramf;
  if (i_sig < 0 || i_sig >= mp->nsigtype) fail("compute_seismograms_cuda: i_sig %d out of range", i_sig + 1);
  if (nlength != mp->nlength_seismogram) fail("compute_seismograms_cuda: nlength_seismogram mismatch");
  if (cur < 0 || cur >= nlength) fail("compute_seismograms_cuda: seismo_current %d outside 1..%d", cur + 1, nlength);
  const int stype = mp->h_seismotype[i_sig];
  const float* field = nullptr;
  if (stype == 2 || stype == 3) flush_correctors(mp);
  switch (stype) {
    case 1: field = mp->d_displ; break;
    case 2: field = mp->d_veloc; break;
    case 3:
      check_accel_complete(mp, "compute_seismograms_cuda");
      field = mp->d_accel;
      break;
    case 4: case 6: break;
    default: fail("seismotype %d (curl) is not supported by this library yet", stype);
  }
  float* d_dst = mp->d_seismo[i_sig];
  if (stype == 4 || stype == 6)
    launch_acoustic_pressure_seismograms(mp, d_dst, cur, *USE_TRICK_FOR_BETTER_PRESSURE, stype == 6);
  else {
    // receivers in solid elements interpolate the vector field; those in fluid elements then overwrite theirs with
    // the gradient of the matching scalar field over the density
    if (mp->nspec_elastic > 0) launch_seismograms(mp, field, d_dst, cur);
    if (acoustic_receivers(mp) > 0) launch_acoustic_vector_seismograms(mp, d_dst, cur, stype);
  }
  if (cur == nlength - 1 || *itf == *it_endf) {
    const size_t size = (size_t)mp->nrec_local * nlength;
    float* h = mp->h_seismo[i_sig];
    check_hip(hipMemcpyAsync(h, d_dst, sizeof(float) * 2 * size, hipMemcpyDeviceToHost, mp->stream), "seismo D2H");
    check_hip(hipStreamSynchronize(mp->stream), "compute_seismograms_cuda");
    for (size_t k = 0; k < size; k++) {
      sisux[k] = (double)h[k];
      sisuz[k] = (double)h[k + size];
    }
  }
}

void compute_kernels_elastic_cuda_(long* Mesh_pointer, realw* deltat) {  // compute_kernels_cuda.cu:47
  (void)deltat;
  Mesh* mp = handle(Mesh_pointer);
  flush_correctors(mp);
  if (!mp->d_rho_kl || !mp->d_b_displ) fail("compute_kernels_elastic_cuda: adjoint arrays not prepared");
  launch_kernels(mp);
}

// ---------------------------------------------------------------------------
// housekeeping
void get_norm_elastic_from_device_(realw* norm, long* Mesh_pointer,
                                   const int* FORWARD_OR_ADJOINT) {  // check_fields_cuda.cu:358
  Mesh* mp = handle(Mesh_pointer);
  const float* f = (*FORWARD_OR_ADJOINT == 3) ? mp->d_b_displ : mp->d_displ;
  if (!f) fail("get_norm_elastic_from_device: field not allocated");
  check_hip(hipMemsetAsync(mp->d_scratch_u32, 0, sizeof(unsigned int), mp->stream), "hipMemsetAsync");
  launch_max_norm(mp, f, mp->d_scratch_u32);
  unsigned int bits = 0;
  check_hip(hipMemcpyAsync(&bits, mp->d_scratch_u32, sizeof(bits), hipMemcpyDeviceToHost, mp->stream), "norm D2H");
  check_hip(hipStreamSynchronize(mp->stream), "get_norm_elastic_from_device");
  float m;
  memcpy(&m, &bits, sizeof(m));
  *norm = std::sqrt(m);
}

void get_free_device_memory_(realw* free_mb, realw* used_mb, realw* total_mb) {  // check_fields_cuda.cu:165
  size_t fr = 0, tot = 0;
  check_hip(hipMemGetInfo(&fr, &tot), "hipMemGetInfo");
  *free_mb = (realw)((double)fr / 1024.0 / 1024.0);
  *used_mb = (realw)((double)(tot - fr) / 1024.0 / 1024.0);
  *total_mb = (realw)((double)tot / 1024.0 / 1024.0);
}

void output_free_device_memory_(int* myrank_f) {  // check_fields_cuda.cu:149
  realw f, u, t;
  get_free_device_memory_(&f, &u, &t);
  printf("%d: GPU memory usage: used = %f MB, free = %f MB, total = %f MB\n", *myrank_f, u, f, t);
}

void get_max_accel_(int* itf, int* sizef, long* Mesh_pointer) {  // check_fields_cuda.cu:484
  Mesh* mp = handle(Mesh_pointer);
  flush_correctors(mp);
  check_accel_complete(mp, "get_max_accel");
  std::vector<float> a((size_t)*sizef);
  check_hip(hipStreamSynchronize(mp->stream), "get_max_accel");
  check_hip(hipMemcpy(a.data(), mp->d_accel, a.size() * sizeof(float), hipMemcpyDeviceToHost), "accel D2H");
  float m = 0.f;
  for (float v : a) m = std::fmax(m, std::fabs(v));
  printf("%d/%d: max=%e\n", *itf, *sizef, m);
}

static void print_max(const char* what, const float* v, int n, int id) {
  float m = 0.f;
  int at = 0;
  for (int k = 0; k < n; k++)
    if (std::fabs(v[k]) > m) {
      m = std::fabs(v[k]);
      at = k;
    }
  printf("%d: %s max = %e at %d\n", id, what, m, at);
}
static void print_device_max(Mesh* mp, const float* d, int n, int id, const char* what) {
  if (n != 2 * mp->nglob) fail("%s: size %d != 2*nglob = %d", what, n, 2 * mp->nglob);
  std::vector<float> h((size_t)n);
  check_hip(hipStreamSynchronize(mp->stream), what);
  points_from_device(mp, h.data(), d, 2, what);
  print_max(what, h.data(), n, id);
}
void check_max_norm_displ_gpu_(int* size, realw* displ, long* Mesh_pointer, int* announceID) {
  (void)displ;
  Mesh* mp = handle(Mesh_pointer);
  print_device_max(mp, mp->d_displ, *size, *announceID, "displ (gpu)");
}
void check_max_norm_vector_(int* size, realw* vector1, int* announceID) { print_max("vector", vector1, *size, *announceID); }
void check_max_norm_displ_(int* size, realw* displ, int* announceID) { print_max("displ", displ, *size, *announceID); }
void check_max_norm_b_displ_gpu_(int* size, realw* b_displ, long* Mesh_pointer, int* announceID) {
  (void)b_displ;
  Mesh* mp = handle(Mesh_pointer);
  print_device_max(mp, mp->d_b_displ, *size, *announceID, "b_displ (gpu)");
}
void check_max_norm_b_accel_gpu_(int* size, realw* b_accel, long* Mesh_pointer, int* announceID) {
  (void)b_accel;
  Mesh* mp = handle(Mesh_pointer);
  flush_correctors(mp);
  print_device_max(mp, mp->d_b_accel, *size, *announceID, "b_accel (gpu)");
}
void check_max_norm_b_veloc_gpu_(int* size, realw* b_veloc, long* Mesh_pointer, int* announceID) {
  (void)b_veloc;
  Mesh* mp = handle(Mesh_pointer);
  flush_correctors(mp);
  print_device_max(mp, mp->d_b_veloc, *size, *announceID, "b_veloc (gpu)");
}
void check_max_norm_b_displ_(int* size, realw* b_displ, int* announceID) { print_max("b_displ", b_displ, *size, *announceID); }
void check_max_norm_b_accel_(int* size, realw* b_accel, int* announceID) { print_max("b_accel", b_accel, *size, *announceID); }
void check_error_vectors_(int* sizef, realw* vector1, realw* vector2) {  // check_fields_cuda.cu:879
  double maxerr = 0.0;
  for (int k = 0; k < *sizef; k++) maxerr = std::fmax(maxerr, std::fabs((double)vector1[k] - (double)vector2[k]));
  printf("maximum error between vectors = %e\n", maxerr);
}
void pause_for_debug_(void) {}  // helper_functions.cu:281

}  // extern "C"

// ---------------------------------------------------------------------------
// extensions without a reference counterpart (specfem2d_gpu_ext.h)
extern "C" void* sf2d_mesh_stream(long* Mesh_pointer) { return (void*)handle(Mesh_pointer)->stream; }

// Builds the assembly plan for a mesh on the host and checks its invariants (no device involved):
// every listed element in exactly one unit, no two units of one colour share a point, patch slots address
// their patch's points.  stats = {colours phase 1, colours phase 2, patches, max points per patch}.
// Returns 0 when every check holds.
extern "C" int sf2d_plan_validate(int nspec, int nglob, const int* ibool_1based, int np, const int* phase_lists,
                                  int mode, int* stats) {
  std::vector<int> ib = to_zero_based(ibool_1based, (size_t)nspec * NGLL2, 1, nglob, "ibool");
  Plan P = build_plan(ib, nspec, nglob, phase_lists, np, mode);
  std::vector<int> seen(nspec, 0);
  for (int e : P.order) {
    if (e < 0 || e >= nspec) return 1;
    seen[e]++;
  }
  for (int e = 0; e < nspec; e++)
    if (seen[e] != 1) return 2;
  // point numbering of the patch mode: a permutation; owned ranges disjoint and consecutive
  std::vector<int> id_owner;
  if (mode == 0) {
    std::vector<char> hit(nglob, 0);
    for (int g = 0; g < nglob; g++) {
      const int id = P.newid[g];
      if (id < 0 || id >= nglob || hit[id]) return 6;
      hit[id] = 1;
    }
    id_owner.assign(nglob, -1);
    for (size_t u = 0; u < P.desc.size() / 4; u++) {
      const int o0 = P.desc[4 * u + 1], no = (P.desc[4 * u + 3] >> 8) & 0xfff;
      for (int id = o0; id < o0 + no; id++) {
        if (id < 0 || id >= nglob || id_owner[id] != -1) return 7;
        id_owner[id] = (int)u;
      }
    }
  }
  std::vector<int> owner(nglob, -1);
  int maxpts = 0;
  for (int ph = 0; ph < 2; ph++)
    for (int c = 0; c < P.ncolour[ph]; c++) {
      std::fill(owner.begin(), owner.end(), -1);
      for (int u = P.colour_start[ph][c]; u < P.colour_start[ph][c] + P.colour_count[ph][c]; u++) {
        int e0 = u, e1 = u + 1;  // element range of the unit
        if (mode == 0) {
          const int* d = P.desc.data() + 4 * (size_t)u;
          const int ne = d[3] & 0xf, no = (d[3] >> 8) & 0xfff, nh = (d[3] >> 20) & 0xfff;
          e0 = d[0];
          e1 = d[0] + ne;
          maxpts = std::max(maxpts, no + nh);
          if (ne < 1 || ne > Mesh::PATCH_ELEMS || no + nh > 256 || no < 1) return 3;
          for (int h = 0; h < nh; h++) {  // halo points: sorted, owned by another patch
            const int id = P.halo[(size_t)d[2] + h];
            if (id < 0 || id >= nglob || id_owner[id] == u || id_owner[id] < 0) return 8;
            if (h > 0 && P.halo[(size_t)d[2] + h - 1] >= id) return 8;
          }
          std::vector<char> used_slot(no + nh, 0);
          std::vector<int> rank_used(no + nh, 0);
          for (int pos = e0; pos < e1; pos++)
            for (int p = 0; p < NGLL2; p++) {
              const int sl = P.slot[(size_t)pos * NGLL2 + p] & 0xff;
              const int rank = P.slot[(size_t)pos * NGLL2 + p] >> 8;
              if (sl >= no + nh || rank >= Mesh::PATCH_SHARE) return 4;
              if (rank_used[sl] & (1 << rank)) return 10;  // two element points with one (point, rank) cell
              rank_used[sl] |= 1 << rank;
              const int id = sl < no ? d[1] + sl : P.halo[(size_t)d[2] + sl - no];
              if (id != P.newid[ib[(size_t)P.order[pos] * NGLL2 + p]]) return 4;
              used_slot[sl] = 1;
            }
          for (char x : used_slot)
            if (!x) return 9;  // every listed point belongs to an element of the patch
        }
        if (mode == 2) break;  // atomics may share points
        for (int pos = e0; pos < e1; pos++)
          for (int p = 0; p < NGLL2; p++) {
            const int g = ib[(size_t)P.order[pos] * NGLL2 + p];
            if (owner[g] != -1 && owner[g] != u) return 5;  // two units of one colour share a point
            owner[g] = u;
          }
      }
    }
  stats[0] = P.ncolour[0];
  stats[1] = P.ncolour[1];
  stats[2] = mode == 0 ? (int)P.desc.size() / 4 : 0;
  stats[3] = maxpts;
  stats[4] = mode == 0 ? (int)P.halo.size() - 1 : 0;  // halo entries of all patches (spare entry excluded)
  long long owned = 0;
  for (size_t u = 0; mode == 0 && u < P.desc.size() / 4; u++) owned += (P.desc[4 * u + 3] >> 8) & 0xfff;
  stats[5] = (int)owned;
  return 0;
}

extern "C" void sf2d_mesh_synchronize(long* Mesh_pointer) {
  Mesh* mp = handle(Mesh_pointer);
  flush_correctors(mp);
  check_hip(hipStreamSynchronize(mp->stream), "sf2d_mesh_synchronize");
}

extern "C" void sf2d_mesh_flush(long* Mesh_pointer) { flush_correctors(handle(Mesh_pointer)); }

extern "C" void sf2d_attenuation_reset(long* Mesh_pointer) {
  Mesh* mp = handle(Mesh_pointer);
  if (!mp->attenuation) return;
  const size_t ne = (size_t)mp->nspec * NGLL2;
  check_hip(hipMemsetAsync(mp->d_att_e, 0, 12 * ne * sizeof(float), mp->stream), "attenuation reset");
}

extern "C" void sf2d_pml_reset(long* Mesh_pointer) { pml_reset(handle(Mesh_pointer)); }

extern "C" int sf2d_mesh_affine_elements(long* Mesh_pointer) {
  Mesh* mp = handle(Mesh_pointer);
  return mp->no_affine ? 0 : mp->n_affine;
}

// whether every patch of the plan for this mesh is a chain of elements (host only)
extern "C" int sf2d_plan_chains(int nspec, int nglob, const int* ibool_1based, int np, const int* phase_lists,
                                int* first_writes) {
  std::vector<int> ib = to_zero_based(ibool_1based, (size_t)nspec * NGLL2, 1, nglob, "ibool");
  Plan P = build_plan(ib, nspec, nglob, phase_lists, np, 0, Mesh::CHAIN_ELEMS);  // as the device plan is chosen
  if (!P.chains) P = build_plan(ib, nspec, nglob, phase_lists, np, 0);
  if (first_writes) {
    *first_writes = 0;
    for (uint8_t f : P.first) *first_writes += f;
  }
  return P.chains ? 1 : 0;
}

// host-only: whether the column-lane kernel's row records hold for the chain plan of this mesh (every chain row
// numbered consecutively in the plan's numbering), and the number of records; -1 when the plan is not all chains
extern "C" int sf2d_plan_row_records(int nspec, int nglob, const int* ibool_1based, int np, const int* phase_lists,
                                     int* nrecords) {
  std::vector<int> ib = to_zero_based(ibool_1based, (size_t)nspec * NGLL2, 1, nglob, "ibool");
  Plan P = build_plan(ib, nspec, nglob, phase_lists, np, 0, Mesh::CHAIN_ELEMS);  // as the device plan is chosen
  if (!P.chains) P = build_plan(ib, nspec, nglob, phase_lists, np, 0);
  if (!P.chains) return -1;
  std::vector<int> ids((size_t)nspec * NGLL2);
  for (int pos = 0; pos < nspec; pos++)
    for (int p = 0; p < NGLL2; p++) ids[(size_t)pos * NGLL2 + p] = P.newid[ib[(size_t)P.order[pos] * NGLL2 + p]];
  std::vector<int> flagged(ids);
  for (size_t l = 0; l < flagged.size() && l < P.first.size(); l++)
    if (P.first[l]) flagged[l] |= Mesh::FIRST_TOUCH;
  std::vector<int4> rows;
  const bool ok = chain_row_records(P.desc, ids, flagged, rows);
  if (nrecords) *nrecords = ok ? (int)rows.size() : 0;
  if (!ok) return 0;
  // the records reproduce every writer's id and flag
  for (int q = 0; q < (int)P.desc.size() / 4; q++) {
    const int e0 = P.desc[4 * q], ne = P.desc[4 * q + 3] & 0xf;
    for (int j = 0; j < NGLL; j++)
      for (int k = 0; k < ne; k++)
        for (int i = 0; i < NGLL; i++) {
          const int x = 4 * k + i;
          const int4 r = rows[(size_t)q * NGLL + j];
          const int id = x == 0 ? r.y : r.x + x;
          const size_t l = (size_t)(e0 + k) * NGLL2 + j * NGLL + i;
          if (id != ids[l]) return 0;
          const bool writer = i < NGLL - 1 || k == ne - 1;
          const unsigned bits = x < 32 ? ((unsigned)r.z >> x) : ((unsigned)r.w >> (x - 32));
          if (writer && ((bits & 1u) != 0) != ((flagged[l] & Mesh::FIRST_TOUCH) != 0)) return 0;
        }
  }
  return 1;
}

extern "C" int sf2d_mesh_chain(long* Mesh_pointer) { return handle(Mesh_pointer)->chain; }

extern "C" int sf2d_mesh_rect(long* Mesh_pointer) { return handle(Mesh_pointer)->rect; }

extern "C" int sf2d_mesh_col_rows(long* Mesh_pointer) { return handle(Mesh_pointer)->d_col_rows != nullptr; }

// tuning knobs of a handle: elements per chain patch (fixed at set-up) and shots per wave of batched launches
extern "C" int sf2d_mesh_chain_elems(long* Mesh_pointer) { return handle(Mesh_pointer)->chain_elems; }

// elements per chain patch of this handle's plan (1..12), before prepare_fields_elastic_device_ builds it; takes
// precedence over SF2D_CHAIN_ELEMS
extern "C" void sf2d_mesh_set_chain_elems(long* Mesh_pointer, int n) {
  Mesh* mp = handle(Mesh_pointer);
  if (mp->sorted) fail("sf2d_mesh_set_chain_elems: the element plan is already built");
  if (n < 0 || n > Mesh::CHAIN_ELEMS) fail("sf2d_mesh_set_chain_elems: %d outside 0..%d", n, Mesh::CHAIN_ELEMS);
  mp->chain_request = n;
}

extern "C" void sf2d_mesh_set_shots_per_wave(long* Mesh_pointer, int shots) {
  Mesh* mp = handle(Mesh_pointer);
  if (shots < 0 || shots > 64) fail("sf2d_mesh_set_shots_per_wave: %d outside 0..64", shots);
  mp->shots_per_wave = shots;
}

extern "C" int sf2d_mesh_colours(long* Mesh_pointer, int iphase) {
  Mesh* mp = handle(Mesh_pointer);
  return (iphase == 1 || iphase == 2) ? mp->ncolour[iphase - 1] : -1;
}

extern "C" void* sf2d_mesh_field(long* Mesh_pointer, int which) {
  Mesh* mp = handle(Mesh_pointer);
  switch (which) {
    case 0: return mp->d_displ;
    case 1: return mp->d_veloc;
    case 2: return mp->d_accel;
    case 3: return mp->d_b_displ;
    case 4: return mp->d_b_veloc;
    case 5: return mp->d_b_accel;
    case 6: return mp->d_rho_kl;
    case 7: return mp->d_mu_kl;
    case 8: return mp->d_kappa_kl;
    default: return nullptr;
  }
}
