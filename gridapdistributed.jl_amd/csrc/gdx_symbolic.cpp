// Host symbolic phase of libgdx: partition, conforming dof numbering, global
// ids, matrix row/col PRanges, local CSR pattern and the position patterns the
// numeric kernels scatter through.  Runs once per problem; its outputs are
// uploaded to HBM by gdx_capi.hip.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <unordered_map>

#include "gdx_internal.h"

namespace gdx {

namespace {

// Host threads for the per-part loops: GDX_HOST_THREADS, else the hardware
// threads capped at 16 (parts are independent; every id and order below is
// the sequential one, whatever the thread count).
int host_threads() {
  const char *e = std::getenv("GDX_HOST_THREADS");
  if (e && std::atoi(e) > 0) return std::atoi(e);
  const unsigned hw = std::thread::hardware_concurrency();
  return (int)std::max(1u, std::min(hw, 16u));
}

// f(r) for every part r in [0, n), parts spread over the host threads; the
// first exception thrown by any part is rethrown here
template <class F>
void for_parts(int n, F &&f) {
  const int nt = std::min(n, host_threads());
  if (nt <= 1) {
    for (int r = 0; r < n; ++r) f(r);
    return;
  }
  std::atomic<int> next{0};
  std::exception_ptr err;
  std::mutex m;
  std::vector<std::thread> th;
  th.reserve(nt);
  for (int t = 0; t < nt; ++t)
    th.emplace_back([&]() {
      for (int r; (r = next.fetch_add(1)) < n;) {
        try {
          f(r);
        } catch (...) {
          std::lock_guard<std::mutex> g(m);
          if (!err) err = std::current_exception();
        }
      }
    });
  for (std::thread &t : th) t.join();
  if (err) std::rethrow_exception(err);
}

// src/PArraysExtras.jl:72-84 (_local_range), 0-based and half-open: the last
// `rem` parts get one extra cell, `ghost` adds one layer on interior sides.
void local_range(int pi, int np, int64_t n, int ghost, int64_t &lo, int64_t &hi) {
  int64_t l = n / np, rem = n % np, p = pi + 1, offset = l * (p - 1);
  if (rem >= np - p + 1) {
    l += 1;
    offset += p - (np - rem) - 1;
  }
  lo = std::max<int64_t>(offset - ghost, 0);
  hi = std::min<int64_t>(l + offset + ghost, n);
}

// Local faces of Gridap's QUAD/HEX polytopes as points of the doubled lattice
// of a cell (vertices, then edges, then faces, in the polytope's order).
const int kHexV[8][3] = {{0, 0, 0}, {2, 0, 0}, {0, 2, 0}, {2, 2, 0}, {0, 0, 2}, {2, 0, 2}, {0, 2, 2}, {2, 2, 2}};
const int kHexE[12][3] = {{1, 0, 0}, {1, 2, 0}, {1, 0, 2}, {1, 2, 2}, {0, 1, 0}, {2, 1, 0},
                          {0, 1, 2}, {2, 1, 2}, {0, 0, 1}, {2, 0, 1}, {0, 2, 1}, {2, 2, 1}};
const int kHexF[6][3] = {{1, 1, 0}, {1, 1, 2}, {1, 0, 1}, {1, 2, 1}, {0, 1, 1}, {2, 1, 1}};
const int kQuadV[4][3] = {{0, 0, 0}, {2, 0, 0}, {0, 2, 0}, {2, 2, 0}};
const int kQuadE[4][3] = {{1, 0, 0}, {1, 2, 0}, {0, 1, 0}, {2, 1, 0}};

// Cell dofs of LagrangianRefFE(order K) in Gridap's local order, as offsets in
// the order-K node lattice of the cell.
int cell_dof_offsets(int D, int K, int off[27][3]) {
  int n = 0;
  const int(*V)[3] = D == 3 ? kHexV : kQuadV;
  for (int v = 0; v < (D == 3 ? 8 : 4); ++v, ++n)
    for (int d = 0; d < 3; ++d) off[n][d] = V[v][d] * K / 2;
  if (K == 1) return n;
  const int(*E)[3] = D == 3 ? kHexE : kQuadE;
  for (int e = 0; e < (D == 3 ? 12 : 4); ++e, ++n)
    for (int d = 0; d < 3; ++d) off[n][d] = E[e][d];
  if (D == 3)
    for (int f = 0; f < 6; ++f, ++n)
      for (int d = 0; d < 3; ++d) off[n][d] = kHexF[f][d];
  off[n][0] = 1;
  off[n][1] = 1;
  off[n][2] = D == 3 ? 1 : 0;
  return n + 1;
}

struct Ctx {
  gdx_problem *p;
  int D, K;
  int64_t n[3], nn[3];
  std::vector<int> node_owner_dim[3];  // part index along d owning global node N
  std::vector<int64_t> own_lo_dim[3], own_hi_dim[3];
};

inline int64_t lin(const int64_t e[3], int64_t x, int64_t y, int64_t z) { return x + e[0] * (y + e[1] * z); }

bool is_dirichlet(const Ctx &c, const int64_t g[3]) {
  if (c.p->desc.dir_kind == 0) return false;
  if (c.p->desc.dir_kind == GDX_DIR_CLAMPED_X0) return g[0] == 0;
  for (int d = 0; d < c.D; ++d)
    if (g[d] == 0 || g[d] == c.nn[d] - 1) return true;
  return false;
}

// Conforming numbering of the part's local model (Gridap: faces dimension-major,
// each face's own dofs consecutive; edges and faces numbered at their first
// appearance when cells are visited in order and local faces in polytope
// order; Dirichlet dofs skipped).  The first cell of a doubled-lattice point
// is the lexicographically smallest local cell containing it.
void number_part(const Ctx &c, Part &P) {
  const int D = c.D, K = c.K;
  const int64_t *lnn = P.lnn;
  const int64_t nl = lnn[0] * lnn[1] * lnn[2];
  int64_t lc[3] = {P.loc_hi[0] - P.loc_lo[0], P.loc_hi[1] - P.loc_lo[1], D == 3 ? P.loc_hi[2] - P.loc_lo[2] : 1};
  std::vector<int64_t> order;
  order.reserve(nl);
  const int64_t nz = D == 3 ? lnn[2] : 1;
  for (int64_t z = 0; z < nz; ++z)
    for (int64_t y = 0; y < lnn[1]; ++y)
      for (int64_t x = 0; x < lnn[0]; ++x)
        if (K == 1 || !((x | y | z) & 1)) order.push_back(lin(lnn, x, y, z));
  if (K == 2) {
    auto first_cell = [](int64_t pt) { return (pt & 1) ? (pt - 1) / 2 : std::max<int64_t>(pt / 2 - 1, 0); };
    for (int fdim = 1; fdim < D; ++fdim) {
      const int(*F)[3] = D == 3 ? (fdim == 1 ? kHexE : kHexF) : kQuadE;
      const int nf = D == 3 ? (fdim == 1 ? 12 : 6) : 4;
      for (int64_t cz = 0; cz < lc[2]; ++cz)
        for (int64_t cy = 0; cy < lc[1]; ++cy)
          for (int64_t cx = 0; cx < lc[0]; ++cx)
            for (int f = 0; f < nf; ++f) {
              int64_t pt[3] = {2 * cx + F[f][0], 2 * cy + F[f][1], D == 3 ? 2 * cz + F[f][2] : 0};
              bool first = first_cell(pt[0]) == cx && first_cell(pt[1]) == cy && (D == 2 || first_cell(pt[2]) == cz);
              if (first) order.push_back(lin(lnn, pt[0], pt[1], pt[2]));
            }
    }
    for (int64_t cz = 0; cz < lc[2]; ++cz)
      for (int64_t cy = 0; cy < lc[1]; ++cy)
        for (int64_t cx = 0; cx < lc[0]; ++cx) order.push_back(lin(lnn, 2 * cx + 1, 2 * cy + 1, D == 3 ? 2 * cz + 1 : 0));
  }
  if ((int64_t)order.size() != nl) throw std::runtime_error("internal: local numbering does not cover all nodes");
  P.lnode_dof.assign(nl, -1);
  P.ldof_lnode.clear();
  for (int64_t l : order) {
    int64_t g[3] = {K * P.loc_lo[0] + l % lnn[0], K * P.loc_lo[1] + (l / lnn[0]) % lnn[1],
                    K * P.loc_lo[2] + l / (lnn[0] * lnn[1])};
    if (!is_dirichlet(c, g)) {
      P.lnode_dof[l] = (int64_t)P.ldof_lnode.size();
      P.ldof_lnode.push_back(l);
    }
  }
  P.n_ldofs = (int64_t)P.ldof_lnode.size();
}

inline void gnode_of(const Ctx &c, const Part &P, int64_t l, int64_t g[3]) {
  g[0] = c.K * P.loc_lo[0] + l % P.lnn[0];
  g[1] = c.K * P.loc_lo[1] + (l / P.lnn[0]) % P.lnn[1];
  g[2] = c.K * P.loc_lo[2] + l / (P.lnn[0] * P.lnn[1]);
}

inline int64_t lnode_in(const Ctx &c, const Part &P, const int64_t g[3]) {
  int64_t l[3];
  for (int d = 0; d < 3; ++d) {
    l[d] = g[d] - c.K * P.loc_lo[d];
    if (l[d] < 0 || l[d] >= P.lnn[d]) return -1;
  }
  return lin(P.lnn, l[0], l[1], l[2]);
}

inline int node_owner(const Ctx &c, const int64_t g[3]) {
  const int *np = c.p->desc.parts;
  int r = c.node_owner_dim[2][g[2]];
  r = r * np[1] + c.node_owner_dim[1][g[1]];
  return r * np[0] + c.node_owner_dim[0][g[0]];
}

// Global nodes sharing a cell with node i along one dimension, restricted to
// cells [clo, chi): [jlo, jhi] (empty when jlo > jhi).
inline void shared_range(int K, int64_t i, int64_t clo, int64_t chi, int64_t &jlo, int64_t &jhi) {
  int64_t cmin = std::max<int64_t>((i + K - 1) / K - 1, clo), cmax = std::min<int64_t>(i / K, chi - 1);
  jlo = K * cmin;
  jhi = K * cmax + K;
  if (cmin > cmax) jlo = 1, jhi = 0;
}

// Local CSR of a vector-valued space (nc dofs per node) and the position
// patterns of its owned rows.  The nc rows of a node share one column set: for
// every stencil node j, its nc component columns.  Own columns (local id
// nc * own index + component) sort like their nodes; ghost columns follow the
// col PRange's first-encounter order, which interleaves components by cell.
// A pattern maps row position -> stencil slot * nc + column component (16 bit:
// a Q2 vertex row holds 125 * 3 entries).
void vector_csr(gdx_problem *p, const Ctx &c) {
  const int K = c.K, nc = p->ncomp, W = 2 * K + 1;
  const int S = c.D == 3 ? W * W * W : W * W;
  const int SN = S * nc;
  const int SP = (SN + 8) & ~7;  // padded stride, at least one spare entry
  constexpr uint16_t kAbsent16 = 0xFFFF, kDir16 = 0xFFFE;
  const int nparts = p->nparts;
  // per part: its CSR and its distinct patterns in first-occurrence order,
  // merged below in part order (ids as one sequential pass would give them)
  struct LocalPats {
    std::vector<std::string> keys;
    std::vector<uint16_t> inv;  // [npat][SP]
    std::vector<int> len, dir;
    std::vector<int> own_pid;  // per owned row: local id, then global id
  };
  std::vector<LocalPats> lp(nparts);
  for_parts(nparts, [&](int r) {
    Part &P = p->P[r];
    LocalPats &L = lp[r];
    std::unordered_map<std::string, int> ids;
    std::vector<uint16_t> key(SN);
    std::vector<std::pair<int64_t, int>> nodes;  // (comp-0 col, node slot) of own stencil nodes
    std::vector<std::pair<int64_t, int>> ents;   // (col, slot * nc + comp)
    P.rowptr.assign(P.nrows + 1, 0);
    P.colidx.clear();
    L.own_pid.assign(P.n_own, -1);
    const int64_t own_cols = P.n_own;  // local col ids below this are own
    for (int64_t row = 0; row < P.nrows; row += nc) {
      if (P.row_ldof[row] % nc != 0) throw std::runtime_error("internal: rows of a node are not grouped");
      int64_t g[3];
      gnode_of(c, P, P.ldof_lnode[P.row_ldof[row]], g);
      const bool own = row < P.n_own;
      int64_t jlo[3] = {0, 0, 0}, jhi[3] = {0, 0, 0};
      for (int k = 0; k < c.D; ++k) {
        if (own) shared_range(K, g[k], 0, c.n[k], jlo[k], jhi[k]);
        else shared_range(K, g[k], P.own_lo[k], P.own_hi[k], jlo[k], jhi[k]);
      }
      nodes.clear();
      ents.clear();
      std::fill(key.begin(), key.end(), kAbsent16);
      bool has_dir = false;
      for (int64_t jz = jlo[2]; jz <= jhi[2]; ++jz)
        for (int64_t jy = jlo[1]; jy <= jhi[1]; ++jy)
          for (int64_t jx = jlo[0]; jx <= jhi[0]; ++jx) {
            int64_t j[3] = {jx, jy, jz};
            int slot = (int)(jx - g[0] + K) + W * (int)(jy - g[1] + K);
            if (c.D == 3) slot += W * W * (int)(jz - g[2] + K);
            if (is_dirichlet(c, j)) {
              for (int k = 0; k < nc; ++k) key[slot * nc + k] = kDir16;
              has_dir = true;
              continue;
            }
            int64_t lj = lnode_in(c, P, j);
            int64_t j0 = lj >= 0 ? P.lnode_dof[lj] : -1;
            int64_t col0 = j0 >= 0 ? P.ldof_col[j0] : -1;
            if (col0 < 0) throw std::runtime_error("internal: stencil column missing from the col PRange");
            if (col0 < own_cols) {
              nodes.emplace_back(col0, slot);
            } else {
              for (int k = 0; k < nc; ++k) ents.emplace_back(P.ldof_col[j0 + k], slot * nc + k);
            }
          }
      std::sort(nodes.begin(), nodes.end());
      std::sort(ents.begin(), ents.end());
      const int64_t base = (int64_t)P.colidx.size();
      int e = 0;
      for (const auto &nd : nodes)
        for (int k = 0; k < nc; ++k, ++e) {
          P.colidx.push_back((int32_t)(nd.first + k));
          key[nd.second * nc + k] = (uint16_t)e;
        }
      for (const auto &en : ents) {
        P.colidx.push_back((int32_t)en.first);
        key[en.second] = (uint16_t)e++;
      }
      // the node's other rows repeat the column set
      for (int k = 1; k < nc; ++k) P.colidx.insert(P.colidx.end(), P.colidx.begin() + base, P.colidx.begin() + base + e);
      for (int k = 0; k < nc; ++k) P.rowptr[row + k + 1] = base + (int64_t)(k + 1) * e;
      if (own) {
        std::string skey(reinterpret_cast<const char *>(key.data()), key.size() * sizeof(uint16_t));
        auto it = ids.find(skey);
        int pid;
        if (it == ids.end()) {
          pid = (int)L.len.size();
          ids.emplace(skey, pid);
          L.keys.push_back(std::move(skey));
          std::vector<uint16_t> inv(SP, 0);
          (void)vector_pattern_inv(key.data(), SN, SP, inv.data());
          L.inv.insert(L.inv.end(), inv.begin(), inv.end());
          L.len.push_back(e);
          L.dir.push_back(has_dir ? 1 : 0);
        } else {
          pid = it->second;
        }
        for (int k = 0; k < nc; ++k) L.own_pid[row + k] = pid;
      }
    }
  });
  std::unordered_map<std::string, int> pat_ids;
  p->pat_fwd.clear();
  p->pat_inv.clear();
  p->pat_invr.clear();
  p->pat_len.clear();
  p->pat_dir.clear();
  p->pat16_inv.clear();
  p->pat16_stride = SP;
  int64_t val_base = 0, row_base = 0, owned_nnz = 0, rinfo_n = 0;
  for (int r = 0; r < nparts; ++r) {
    Part &P = p->P[r];
    LocalPats &L = lp[r];
    std::vector<int> to_global(L.len.size());
    for (size_t l = 0; l < L.len.size(); ++l) {
      auto it = pat_ids.find(L.keys[l]);
      if (it != pat_ids.end()) {
        to_global[l] = it->second;
        continue;
      }
      const int pid = (int)p->pat_len.size();
      pat_ids.emplace(L.keys[l], pid);
      to_global[l] = pid;
      p->pat16_inv.insert(p->pat16_inv.end(), L.inv.begin() + l * SP, L.inv.begin() + (l + 1) * SP);
      p->pat_len.push_back(L.len[l]);
      p->pat_dir.push_back(L.dir[l]);
    }
    for (int &pid : L.own_pid) pid = to_global[pid];
    P.val_base = val_base;
    P.row_base = row_base;
    owned_nnz += P.rowptr[P.n_own];
    P.nnz = P.rowptr[P.nrows];
    val_base += P.rowptr[P.nrows];
    row_base += P.nrows;
    P.rowinfo_base = rinfo_n;
    rinfo_n += P.onb_n[0] * P.onb_n[1] * P.onb_n[2];
  }
  // row table over the owned node box: the node's first row
  p->rowinfo.assign(rinfo_n, RowInfo{0, 0, -1, 0});
  for_parts(nparts, [&](int r) {
    const Part &P = p->P[r];
    const std::vector<int> &own_pid = lp[r].own_pid;
    std::vector<int64_t> ldof_own(P.n_ldofs, -1);
    for (int64_t o = 0; o < P.n_own; ++o) ldof_own[P.own_ldofs[o]] = o;
    RowInfo *out = p->rowinfo.data() + P.rowinfo_base;
    for (int64_t z = 0; z < P.onb_n[2]; ++z)
      for (int64_t y = 0; y < P.onb_n[1]; ++y)
        for (int64_t x = 0; x < P.onb_n[0]; ++x, ++out) {
          int64_t g[3] = {P.onb_lo[0] + x, P.onb_lo[1] + y, P.onb_lo[2] + z};
          int64_t l = lnode_in(c, P, g);
          int64_t ld = l >= 0 ? P.lnode_dof[l] : -1;
          if (ld >= 0) {
            int64_t o = ldof_own[ld];
            if (o < 0) throw std::runtime_error("internal: owned node box holds a non-owned dof");
            out->rowstart = P.val_base + P.rowptr[o];
            out->bidx = P.row_base + o;
            out->pid = own_pid[o];
            out->lenf = p->pat_len[out->pid] | (p->pat_dir[out->pid] << 16);
          }
        }
  });
  p->npat = (int)p->pat_len.size();
  p->nnz_total = val_base;
  p->rows_total = row_base;
  p->owned_nnz = owned_nnz;
}

}  // namespace

bool scalar_pattern_tables(int D, int K, const uint8_t *key, int S, int SP, const int64_t g[3], int len,
                           uint8_t *fwd, uint8_t *inv, uint8_t *invr) {
  bool has_dir = false;
  for (int s = 0; s < SP; ++s) fwd[s] = kSlotAbsent, inv[s] = 0, invr[s] = 0;
  for (int s = 0; s < S; ++s) {
    const uint8_t v = key[s];
    if (v == kSlotDirichlet) has_dir = true;
    else if (v != kSlotAbsent) inv[v] = (uint8_t)s;
    fwd[s] = v;
  }
  // reduced numbering (the pattern fixes the node parity of its rows); the
  // Dirichlet slots follow the row's entries, their count in the last byte
  const int Wd = 2 * K + 1;
  auto reduced = [&](int s) {
    const int sx = s % Wd - K, sy = (s / Wd) % Wd - K, sz = D == 3 ? s / (Wd * Wd) - K : 0;
    int r = sx - stencil_lo(K, (int)g[0]);
    r += stencil_width(K, (int)g[0]) * (sy - stencil_lo(K, (int)g[1]));
    if (D == 3) r += stencil_width(K, (int)g[0]) * stencil_width(K, (int)g[1]) * (sz - stencil_lo(K, (int)g[2]));
    return r;
  };
  int ndir = 0;
  for (int s = 0; s < S; ++s) {
    const uint8_t v = fwd[s];
    if (v != kSlotDirichlet && v != kSlotAbsent) invr[v] = (uint8_t)reduced(s);
  }
  for (int s = 0; s < S; ++s)
    if (fwd[s] == kSlotDirichlet) invr[len + ndir++] = (uint8_t)reduced(s);
  invr[SP - 1] = (uint8_t)ndir;
  return has_dir;
}

bool vector_pattern_inv(const uint16_t *key, int SN, int SP, uint16_t *inv) {
  bool has_dir = false;
  for (int s = 0; s < SP; ++s) inv[s] = 0;
  for (int s3 = 0; s3 < SN; ++s3) {
    if (key[s3] < 0xFFFE) inv[key[s3]] = (uint16_t)s3;
    else if (key[s3] == 0xFFFE) has_dir = true;
  }
  return has_dir;
}

void symbolic_mark(const char *stage, bool start) {
  static thread_local std::chrono::steady_clock::time_point last;
  const auto now = std::chrono::steady_clock::now();
  const char *e = std::getenv("GDX_TRACE_SYMBOLIC");
  if (!start && e && std::atoi(e) > 0)
    std::fprintf(stderr, "gdx symbolic: %-16s %8.2f ms\n", stage,
                 std::chrono::duration<double, std::milli>(now - last).count());
  last = now;
}

void tile_shape(int dim, int K, int R[3]) {
  // must match the template instances launched by launch_matrix
  if (dim == 3) {
    // z-marching columns: R[2] rows along z per workgroup
    R[0] = K == 2 ? 4 : 8;
    R[1] = 8;
    R[2] = 32;
  } else {
    R[0] = 16; R[1] = 16; R[2] = 1;
  }
}

// Host numbering, owners, global ids and row/col PRanges of every part.
static void host_fe_stage(gdx_problem *p, const Ctx &c) {
  const int nparts = p->nparts;
  const gdx_problem_desc &d = p->desc;
  const bool FA = d.strategy == GDX_FULLY_ASSEMBLED_ROWS;
  (void)FA;
  (void)d;
  for_parts(nparts, [&](int r) { number_part(c, p->P[r]); });
  symbolic_mark("numbering");
  // owners and rank-blocked global ids (src/FESpaces.jl:139-261): owned dofs of
  // part r get consecutive ids after those of parts < r
  for_parts(nparts, [&](int r) {
    Part &P = p->P[r];
    P.ldof_owner.resize(P.n_ldofs);
    P.ldof_gid.assign(P.n_ldofs, -1);
    P.own_ldofs.clear();
    for (int64_t i = 0; i < P.n_ldofs; ++i) {
      int64_t g[3];
      gnode_of(c, P, P.ldof_lnode[i], g);
      P.ldof_owner[i] = node_owner(c, g);
      if (P.ldof_owner[i] == r) P.own_ldofs.push_back(i);
    }
    P.n_own = (int64_t)P.own_ldofs.size();
  });
  std::vector<int64_t> gid_first(nparts + 1, 0);
  for (int r = 0; r < nparts; ++r) gid_first[r + 1] = gid_first[r] + p->P[r].n_own;
  for_parts(nparts, [&](int r) {
    Part &P = p->P[r];
    for (int64_t o = 0; o < P.n_own; ++o) P.ldof_gid[P.own_ldofs[o]] = gid_first[r] + o;
  });
  p->ngdofs = gid_first[nparts];
  for_parts(nparts, [&](int r) {
    Part &P = p->P[r];
    for (int64_t i = 0; i < P.n_ldofs; ++i) {
      if (P.ldof_gid[i] >= 0) continue;
      int64_t g[3];
      gnode_of(c, P, P.ldof_lnode[i], g);
      const Part &Q = p->P[P.ldof_owner[i]];
      int64_t lq = lnode_in(c, Q, g);
      if (lq < 0 || Q.lnode_dof[lq] < 0) throw std::runtime_error("internal: ghost dof missing in its owner");
      P.ldof_gid[i] = Q.ldof_gid[Q.lnode_dof[lq]];
    }
  });
  // vector-valued space: each free node carries ncomp consecutive dofs (Gridap
  // numbers a face's own dofs together; Q1/Q2 faces own one node), so local
  // and global ids become ncomp * (node id) + component, owners per node
  const int nc = p->ncomp;
  if (nc > 1) {
    for_parts(nparts, [&](int r) {
      Part &P = p->P[r];
      std::vector<int64_t> lnode(P.n_ldofs * nc), gid(P.n_ldofs * nc), own;
      std::vector<int32_t> owner(P.n_ldofs * nc);
      for (int64_t i = 0; i < P.n_ldofs; ++i)
        for (int k = 0; k < nc; ++k) {
          lnode[nc * i + k] = P.ldof_lnode[i];
          gid[nc * i + k] = nc * P.ldof_gid[i] + k;
          owner[nc * i + k] = P.ldof_owner[i];
        }
      own.reserve(P.own_ldofs.size() * nc);
      for (int64_t o : P.own_ldofs)
        for (int k = 0; k < nc; ++k) own.push_back(nc * o + k);
      for (int64_t &l : P.lnode_dof)
        if (l >= 0) l *= nc;
      P.ldof_lnode.swap(lnode);
      P.ldof_gid.swap(gid);
      P.ldof_owner.swap(owner);
      P.own_ldofs.swap(own);
      P.n_ldofs *= nc;
      P.n_own *= nc;
    });
    p->ngdofs *= nc;
  }

  symbolic_mark("global ids");
  int off[27][3];
  const int ncd = cell_dof_offsets(c.D, c.K, off);
  const int K = c.K;
  auto cell_ldof = [&](const Part &P, int64_t cx, int64_t cy, int64_t cz, int a) -> int64_t {
    int64_t l = lin(P.lnn, K * (cx - P.loc_lo[0]) + off[a][0], K * (cy - P.loc_lo[1]) + off[a][1],
                    K * (cz - P.loc_lo[2]) + off[a][2]);
    return P.lnode_dof[l];
  };

  // row and col PRanges (src/Algebra.jl:593-646, 1136-1257)
  for_parts(nparts, [&](int r) {
    Part &P = p->P[r];
    P.row_gid.clear();
    P.row_owner.clear();
    P.row_ldof.clear();
    for (int64_t o = 0; o < P.n_own; ++o) {
      P.row_gid.push_back(P.ldof_gid[P.own_ldofs[o]]);
      P.row_owner.push_back(r);
      P.row_ldof.push_back(P.own_ldofs[o]);
    }
    if (!FA) {
      // ghost rows touched by owned cells, in FE local id order
      for (int64_t i = 0; i < P.n_ldofs; ++i) {
        if (P.ldof_owner[i] == r) continue;
        int64_t g[3];
        gnode_of(c, P, P.ldof_lnode[i], g);
        bool in = true;
        for (int k = 0; k < c.D; ++k) in = in && g[k] >= K * P.own_lo[k] && g[k] <= K * P.own_hi[k];
        if (in) {
          P.row_gid.push_back(P.ldof_gid[i]);
          P.row_owner.push_back(P.ldof_owner[i]);
          P.row_ldof.push_back(i);
        }
      }
    }
    P.nrows = (int64_t)P.row_gid.size();
    P.n_ghost_rows = P.nrows - P.n_own;

    P.ldof_col.assign(P.n_ldofs, -1);
    P.col_gid.clear();
    P.col_owner.clear();
    for (int64_t o = 0; o < P.n_own; ++o) {
      P.ldof_col[P.own_ldofs[o]] = o;
      P.col_gid.push_back(P.ldof_gid[P.own_ldofs[o]]);
      P.col_owner.push_back(r);
    }
    std::vector<int64_t> ghosts;  // ldofs of ghost cols in PRange order
    // a cell's dofs in the reference FE's order: component-major over its nodes
    auto visit_cell_cols = [&](int64_t cx, int64_t cy, int64_t cz) {
      for (int k = 0; k < nc; ++k)
        for (int b = 0; b < ncd; ++b) {
          int64_t j = cell_ldof(P, cx, cy, cz, b);
          if (j < 0) continue;
          j += k;
          if (P.ldof_owner[j] != r && P.ldof_col[j] == -1) {
            P.ldof_col[j] = -2;  // seen
            ghosts.push_back(j);
          }
        }
    };
    auto cell_has_row = [&](const Part &Q, int64_t cx, int64_t cy, int64_t cz, int want_owner) {
      for (int a = 0; a < ncd; ++a) {
        int64_t i = cell_ldof(Q, cx, cy, cz, a);
        if (i >= 0 && (want_owner < 0 || Q.ldof_owner[i] == want_owner)) return true;
      }
      return false;
    };
    if (!FA) {
      // first encounter over the local COO: own cells (columns outer), then the
      // triplets received from each sender in ascending rank order
      for (int64_t cz = P.own_lo[2]; cz < P.own_hi[2]; ++cz)
        for (int64_t cy = P.own_lo[1]; cy < P.own_hi[1]; ++cy)
          for (int64_t cx = P.own_lo[0]; cx < P.own_hi[0]; ++cx)
            if (cell_has_row(P, cx, cy, cz, -1)) visit_cell_cols(cx, cy, cz);
      for (int s = 0; s < nparts; ++s) {
        if (s == r) continue;
        const Part &S = p->P[s];
        int64_t lo[3], hi[3];
        bool empty = false;
        for (int k = 0; k < 3; ++k) {
          // sender cells whose closure meets this part's owned node box
          int64_t a = std::max<int64_t>(S.own_lo[k], (P.onb_lo[k] + K - 1) / K - 1);
          int64_t b = std::min<int64_t>(S.own_hi[k], (P.onb_lo[k] + P.onb_n[k] - 1) / K + 1);
          if (k >= c.D) a = 0, b = 1;
          lo[k] = a;
          hi[k] = b;
          if (a >= b) empty = true;
        }
        if (empty) continue;
        for (int64_t cz = lo[2]; cz < hi[2]; ++cz)
          for (int64_t cy = lo[1]; cy < hi[1]; ++cy)
            for (int64_t cx = lo[0]; cx < hi[0]; ++cx)
              if (cell_has_row(S, cx, cy, cz, r)) visit_cell_cols(cx, cy, cz);
      }
    } else {
      // touched ghost columns of cells holding an owned row, sorted by FE local id
      for (int64_t cz = P.loc_lo[2]; cz < P.loc_hi[2]; ++cz)
        for (int64_t cy = P.loc_lo[1]; cy < P.loc_hi[1]; ++cy)
          for (int64_t cx = P.loc_lo[0]; cx < P.loc_hi[0]; ++cx)
            if (cell_has_row(P, cx, cy, cz, r)) visit_cell_cols(cx, cy, cz);
      std::sort(ghosts.begin(), ghosts.end());
    }
    for (size_t g = 0; g < ghosts.size(); ++g) {
      P.ldof_col[ghosts[g]] = P.n_own + (int64_t)g;
      P.col_gid.push_back(P.ldof_gid[ghosts[g]]);
      P.col_owner.push_back(P.ldof_owner[ghosts[g]]);
    }
    P.ncols = (int64_t)P.col_gid.size();
    P.n_ghost_cols = P.ncols - P.n_own;
  });

  symbolic_mark("PRanges");
}

// Host CSR, position patterns, row table and tiles of every part.
static int host_csr_stage(gdx_problem *p, const Ctx &c) {
  const int nparts = p->nparts;
  const gdx_problem_desc &d = p->desc;
  const bool FA = d.strategy == GDX_FULLY_ASSEMBLED_ROWS;
  (void)FA;
  (void)d;
  const int K = c.K, nc = p->ncomp;
  if (nc > 1) {
    vector_csr(p, c);
    int R[3];
    tile_shape(c.D, K, R);
    build_tiles(p, R);
    return GDX_OK;
  }

  // local CSR patterns and the position patterns of owned rows.  Each part
  // builds its CSR and a table of its distinct patterns in first-occurrence
  // order; merging the tables in part order gives every pattern the id of its
  // first occurrence over (part, row), as one sequential pass would.
  const int S = (c.D == 3) ? (2 * K + 1) * (2 * K + 1) * (2 * K + 1) : (2 * K + 1) * (2 * K + 1);
  const int SP = pattern_stride(S);
  struct LocalPats {
    std::vector<std::string> keys;
    std::vector<uint8_t> fwd, inv, invr;  // [npat][SP]
    std::vector<int> len, dir;
    std::vector<int> own_pid;  // per owned row: local id, then global id
  };
  std::vector<LocalPats> lp(nparts);
  for_parts(nparts, [&](int r) {
    Part &P = p->P[r];
    LocalPats &L = lp[r];
    std::unordered_map<std::string, int> ids;
    P.rowptr.assign(P.nrows + 1, 0);
    P.colidx.clear();
    std::vector<std::pair<int32_t, int>> ents;
    L.own_pid.assign(P.n_own, -1);
    std::string key(S, (char)kSlotAbsent);
    for (int64_t row = 0; row < P.nrows; ++row) {
      int64_t g[3];
      gnode_of(c, P, P.ldof_lnode[P.row_ldof[row]], g);
      const bool own = row < P.n_own;
      int64_t jlo[3] = {0, 0, 0}, jhi[3] = {0, 0, 0};
      for (int k = 0; k < c.D; ++k) {
        if (own) shared_range(K, g[k], 0, c.n[k], jlo[k], jhi[k]);
        else shared_range(K, g[k], P.own_lo[k], P.own_hi[k], jlo[k], jhi[k]);
      }
      ents.clear();
      std::fill(key.begin(), key.end(), (char)kSlotAbsent);
      for (int64_t jz = jlo[2]; jz <= jhi[2]; ++jz)
        for (int64_t jy = jlo[1]; jy <= jhi[1]; ++jy)
          for (int64_t jx = jlo[0]; jx <= jhi[0]; ++jx) {
            int64_t j[3] = {jx, jy, jz};
            int slot = (int)(jx - g[0] + K);
            slot += (2 * K + 1) * (int)(jy - g[1] + K);
            if (c.D == 3) slot += (2 * K + 1) * (2 * K + 1) * (int)(jz - g[2] + K);
            if (is_dirichlet(c, j)) {
              key[slot] = (char)kSlotDirichlet;
              continue;
            }
            int64_t lj = lnode_in(c, P, j);
            int64_t col = lj >= 0 ? P.ldof_col[P.lnode_dof[lj]] : -1;
            if (col < 0) throw std::runtime_error("internal: stencil column missing from the col PRange");
            ents.emplace_back((int32_t)col, slot);
          }
      std::sort(ents.begin(), ents.end());
      for (size_t e = 0; e < ents.size(); ++e) {
        P.colidx.push_back(ents[e].first);
        key[ents[e].second] = (char)e;
      }
      P.rowptr[row + 1] = (int64_t)P.colidx.size();
      if (own) {
        auto it = ids.find(key);
        int pid;
        if (it == ids.end()) {
          pid = (int)L.len.size();
          ids.emplace(key, pid);
          L.keys.push_back(key);
          std::vector<uint8_t> inv(SP, 0), fwd(SP, kSlotAbsent), invr(SP, 0);
          const bool has_dir = scalar_pattern_tables(c.D, K, reinterpret_cast<const uint8_t *>(key.data()), S, SP, g,
                                                     (int)ents.size(), fwd.data(), inv.data(), invr.data());
          L.fwd.insert(L.fwd.end(), fwd.begin(), fwd.end());
          L.inv.insert(L.inv.end(), inv.begin(), inv.end());
          L.invr.insert(L.invr.end(), invr.begin(), invr.end());
          L.len.push_back((int)ents.size());
          L.dir.push_back(has_dir ? 1 : 0);
        } else {
          pid = it->second;
        }
        L.own_pid[row] = pid;
      }
    }
  });
  std::unordered_map<std::string, int> pat_ids;
  p->pat_fwd.clear();
  p->pat_inv.clear();
  p->pat_invr.clear();
  p->pat_len.clear();
  p->pat_dir.clear();
  int64_t val_base = 0, row_base = 0, owned_nnz = 0, rinfo_n = 0;
  for (int r = 0; r < nparts; ++r) {
    Part &P = p->P[r];
    LocalPats &L = lp[r];
    std::vector<int> to_global(L.len.size());
    for (size_t l = 0; l < L.len.size(); ++l) {
      auto it = pat_ids.find(L.keys[l]);
      if (it != pat_ids.end()) {
        to_global[l] = it->second;
        continue;
      }
      const int pid = (int)p->pat_len.size();
      pat_ids.emplace(L.keys[l], pid);
      to_global[l] = pid;
      p->pat_fwd.insert(p->pat_fwd.end(), L.fwd.begin() + l * SP, L.fwd.begin() + (l + 1) * SP);
      p->pat_inv.insert(p->pat_inv.end(), L.inv.begin() + l * SP, L.inv.begin() + (l + 1) * SP);
      p->pat_invr.insert(p->pat_invr.end(), L.invr.begin() + l * SP, L.invr.begin() + (l + 1) * SP);
      p->pat_len.push_back(L.len[l]);
      p->pat_dir.push_back(L.dir[l]);
    }
    for (int &pid : L.own_pid) pid = to_global[pid];
    P.val_base = val_base;
    P.row_base = row_base;
    owned_nnz += P.rowptr[P.n_own];
    P.nnz = P.rowptr[P.nrows];
    val_base += P.rowptr[P.nrows];
    row_base += P.nrows;
    P.rowinfo_base = rinfo_n;
    rinfo_n += P.onb_n[0] * P.onb_n[1] * P.onb_n[2];
  }
  // row table over each part's owned node box
  p->rowinfo.assign(rinfo_n, RowInfo{0, 0, -1, 0});
  for_parts(nparts, [&](int r) {
    const Part &P = p->P[r];
    const std::vector<int> &own_pid = lp[r].own_pid;
    std::vector<int64_t> ldof_own(P.n_ldofs, -1);
    for (int64_t o = 0; o < P.n_own; ++o) ldof_own[P.own_ldofs[o]] = o;
    RowInfo *out = p->rowinfo.data() + P.rowinfo_base;
    for (int64_t z = 0; z < P.onb_n[2]; ++z)
      for (int64_t y = 0; y < P.onb_n[1]; ++y)
        for (int64_t x = 0; x < P.onb_n[0]; ++x, ++out) {
          int64_t g[3] = {P.onb_lo[0] + x, P.onb_lo[1] + y, P.onb_lo[2] + z};
          int64_t l = lnode_in(c, P, g);
          int64_t ld = l >= 0 ? P.lnode_dof[l] : -1;
          if (ld >= 0) {
            int64_t o = ldof_own[ld];
            if (o < 0) throw std::runtime_error("internal: owned node box holds a non-owned dof");
            out->rowstart = P.val_base + P.rowptr[o];
            out->bidx = P.row_base + o;
            out->pid = own_pid[o];
            out->lenf = p->pat_len[out->pid] | (p->pat_dir[out->pid] << 16);
          }
        }
  });
  p->npat = (int)p->pat_len.size();
  p->nnz_total = val_base;
  p->rows_total = row_base;
  p->owned_nnz = owned_nnz;

  // tiles of owned rows (the flux operator's fused kernel takes its own shape)
  int R[3];
  tile_shape(c.D, K, R);
  if (p->desc.op_kind == GDX_OP_FLUX && c.D == 3) R[1] = kFluxR2, R[2] = kFluxR3;
  build_tiles(p, R);
  return GDX_OK;
}

int symbolic(gdx_problem *p) {
  p->dev_structure = p->dev_numbering = false;
  p->host_csr_stale = p->host_rowinfo_stale = p->host_fe_stale = false;
  symbolic_mark("", true);
  Ctx c;
  c.p = p;
  c.D = p->D;
  c.K = p->K;
  const gdx_problem_desc &d = p->desc;
  for (int k = 0; k < 3; ++k) {
    c.n[k] = k < c.D ? d.cells[k] : 1;
    c.nn[k] = k < c.D ? c.K * c.n[k] + 1 : 1;
    int np = k < c.D ? d.parts[k] : 1;
    c.own_lo_dim[k].resize(np);
    c.own_hi_dim[k].resize(np);
    for (int q = 0; q < np; ++q) local_range(q, np, c.n[k], 0, c.own_lo_dim[k][q], c.own_hi_dim[k][q]);
    // owner of node N along k: the highest part whose first node is <= N
    // (generate_gids: owner = max owner of the cells touching the dof)
    c.node_owner_dim[k].assign(c.nn[k], 0);
    for (int64_t N = 0; N < c.nn[k]; ++N) {
      int o = 0;
      for (int q = 0; q < np; ++q)
        if (c.K * c.own_lo_dim[k][q] <= N) o = q;
      c.node_owner_dim[k][N] = o;
    }
  }
  const int nparts = p->nparts;
  p->P.assign(nparts, Part{});
  for_parts(nparts, [&](int r) {
    Part &P = p->P[r];
    P.rank = r;
    P.pc[0] = r % d.parts[0];
    P.pc[1] = (r / d.parts[0]) % d.parts[1];
    P.pc[2] = c.D == 3 ? r / (d.parts[0] * d.parts[1]) : 0;
    for (int k = 0; k < 3; ++k) {
      if (k < c.D) {
        local_range(P.pc[k], d.parts[k], c.n[k], 0, P.own_lo[k], P.own_hi[k]);
        local_range(P.pc[k], d.parts[k], c.n[k], 1, P.loc_lo[k], P.loc_hi[k]);
        P.lnn[k] = c.K * (P.loc_hi[k] - P.loc_lo[k]) + 1;
        P.onb_lo[k] = c.K * P.own_lo[k];
        P.onb_n[k] = c.K * (P.own_hi[k] - P.own_lo[k]) + (P.pc[k] == d.parts[k] - 1 ? 1 : 0);
      } else {
        P.own_lo[k] = P.loc_lo[k] = 0;
        P.own_hi[k] = P.loc_hi[k] = 1;
        P.lnn[k] = 1;
        P.onb_lo[k] = 0;
        P.onb_n[k] = 1;
      }
    }
  });
  // numbering, ids and PRanges, then CSR, position patterns, row table and
  // tiles: on the device unless the problem has none (host-only problems) or
  // GDX_HOST_SYMBOLIC=1 asks for the host stages; the device CSR stage hands
  // back to the host one when its pattern table overflows or two keys collide
  // in a hash
  const char *hs = std::getenv("GDX_HOST_SYMBOLIC");
  const bool device = p->ctx && !(hs && std::atoi(hs) > 0);
  bool numbered = false;
  if (device) {
    int off[27][3];
    const int ncd = cell_dof_offsets(c.D, c.K, off);
    const int rc = numbering_device(p, c.node_owner_dim, off, ncd, c.D == 3 ? kHexE : kQuadE, c.D == 3 ? 12 : 4,
                                    kHexF, c.D == 3 ? 6 : 0);
    if (rc != kUseHost && rc != GDX_OK) return rc;
    numbered = rc == GDX_OK;
    symbolic_mark("numbering+PRanges (device)");
  }
  if (!numbered) host_fe_stage(p, c);
  if (device) {
    const int rc = structure_device(p);
    if (rc != kUseHost) {
      if (rc) return rc;
      int R[3];
      tile_shape(c.D, c.K, R);
      if (p->desc.op_kind == GDX_OP_FLUX && c.D == 3) R[1] = kFluxR2, R[2] = kFluxR3;
      const int rt = build_tiles_device(p, R);
      symbolic_mark("tiles (device)");
      return rt;
    }
    if (numbered) {
      const int rf = host_fe(p);  // the host stage reads the numbering from the host copies
      if (rf) return rf;
    }
  }
  return host_csr_stage(p, c);
}

// Tiles of owned rows of extent R (rows per direction; multiples of the order),
// each with its distinct position patterns; the rows' tile-local pattern index
// goes into bits 17.. of RowInfo::lenf.
void build_tiles(gdx_problem *p, const int R[3]) {
  const int nparts = p->nparts;
  for (int k = 0; k < 3; ++k) p->tileR[k] = R[k];
  // per part (its rows' RowInfo entries are its own), then concatenated in part order
  std::vector<std::vector<Tile>> ptiles(nparts);
  std::vector<std::vector<int>> ppats(nparts);
  for_parts(nparts, [&](int r) {
    const Part &P = p->P[r];
    std::vector<Tile> &tiles = ptiles[r];
    std::vector<int> &tpats = ppats[r];
    int64_t nt[3];
    for (int k = 0; k < 3; ++k) nt[k] = (P.onb_n[k] + p->tileR[k] - 1) / p->tileR[k];
    std::vector<int> pids;
    for (int64_t tz = 0; tz < nt[2]; ++tz)
      for (int64_t ty = 0; ty < nt[1]; ++ty)
        for (int64_t tx = 0; tx < nt[0]; ++tx) {
          Tile t{};
          t.part = r;
          int64_t tt[3] = {tx, ty, tz};
          bool any = false;
          for (int k = 0; k < 3; ++k) {
            t.org[k] = (int)(P.onb_lo[k] + tt[k] * p->tileR[k]);
            t.ext[k] = (int)std::min<int64_t>(p->tileR[k], P.onb_n[k] - tt[k] * p->tileR[k]);
            t.onb_lo[k] = (int)P.onb_lo[k];
            t.onb_n[k] = (int)P.onb_n[k];
          }
          t.rowinfo_base = P.rowinfo_base;
          // distinct patterns of the tile; rows get their tile-local index
          pids.clear();
          t.pat_off = (long long)tpats.size();  // part-relative until concatenated
          for (int z = 0; z < t.ext[2]; ++z)
            for (int y = 0; y < t.ext[1]; ++y)
              for (int x = 0; x < t.ext[0]; ++x) {
                int64_t li = ((int64_t)(t.org[2] - t.onb_lo[2] + z) * P.onb_n[1] + (t.org[1] - t.onb_lo[1] + y)) *
                                 P.onb_n[0] +
                             (t.org[0] - t.onb_lo[0] + x);
                RowInfo &ri = p->rowinfo[P.rowinfo_base + li];
                if (ri.pid < 0) continue;
                any = true;
                int l = (int)(std::find(pids.begin(), pids.end(), ri.pid) - pids.begin());
                if (l == (int)pids.size()) pids.push_back(ri.pid);
                if (l >= (1 << 14)) throw std::runtime_error("internal: too many patterns in one tile");
                ri.lenf = (ri.lenf & 0x1FFFF) | (l << 17);
              }
          t.npat = (int)pids.size();
          if (any) {
            tpats.insert(tpats.end(), pids.begin(), pids.end());
            tiles.push_back(t);
          }
        }
  });
  p->tiles.clear();
  p->tile_pats.clear();
  for (int r = 0; r < nparts; ++r) {
    const long long off = (long long)p->tile_pats.size();
    for (Tile &t : ptiles[r]) {
      t.pat_off += off;
      p->tiles.push_back(t);
    }
    p->tile_pats.insert(p->tile_pats.end(), ppats[r].begin(), ppats[r].end());
  }
}

// Reference-cell tables: Gauss-Legendre points/weights on [0,1] and the 1D
// Lagrange basis on equispaced nodes j/K (tensor index order).
void fill_kparams(gdx_problem *p) {
  KParams &kp = p->kp;
  std::memset(&kp, 0, sizeof(kp));
  const gdx_problem_desc &d = p->desc;
  kp.dim = p->D;
  kp.K = p->K;
  kp.NQ = p->NQ;
  kp.S = p->D == 3 ? (2 * p->K + 1) * (2 * p->K + 1) * (2 * p->K + 1) : (2 * p->K + 1) * (2 * p->K + 1);
  for (int k = 0; k < 3; ++k) {
    kp.ncells[k] = k < p->D ? (int)d.cells[k] : 1;
    kp.qn[k] = k < p->D ? p->NQ * kp.ncells[k] : 1;
    kp.nn[k] = k < p->D ? p->K * kp.ncells[k] + 1 : 1;
    kp.x0[k] = k < p->D ? d.domain[2 * k] : 0.0;
    kp.h[k] = k < p->D ? (d.domain[2 * k + 1] - d.domain[2 * k]) / (double)kp.ncells[k] : 1.0;
    kp.hK[k] = kp.h[k] / p->K;
  }
  // Gauss-Legendre on [-1,1] by Newton on P_n, mapped to [0,1]
  const int n = p->NQ;
  for (int i = 0; i < n; ++i) {
    double z = std::cos(M_PI * (i + 0.75) / (n + 0.5)), dp = 0.0;
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = z;
      for (int j = 2; j <= n; ++j) {
        double p2 = ((2.0 * j - 1.0) * z * p1 - (j - 1.0) * p0) / j;
        p0 = p1;
        p1 = p2;
      }
      if (n == 1) p1 = z, p0 = 1.0;
      dp = n * (z * p1 - p0) / (z * z - 1.0);
      double dz = p1 / dp;
      z -= dz;
      if (std::fabs(dz) < 1e-17) break;
    }
    kp.qx[n - 1 - i] = 0.5 * (1.0 + z);
    kp.qw[n - 1 - i] = 1.0 / ((1.0 - z * z) * dp * dp);
  }
  const int K = p->K;
  for (int a = 0; a <= K; ++a)
    for (int q = 0; q < n; ++q) {
      double x = kp.qx[q], v = 1.0, dv = 0.0;
      for (int j = 0; j <= K; ++j)
        if (j != a) v *= (x - (double)j / K) / ((double)(a - j) / K);
      for (int m = 0; m <= K; ++m) {
        if (m == a) continue;
        double t = 1.0 / ((double)(a - m) / K);
        for (int j = 0; j <= K; ++j)
          if (j != a && j != m) t *= (x - (double)j / K) / ((double)(a - j) / K);
        dv += t;
      }
      kp.phi[a][q] = v;
      kp.dphi[a][q] = dv;
    }
  for (int d = 0; d < 3; ++d)
    for (int a = 0; a <= K; ++a)
      for (int b = 0; b <= K; ++b)
        for (int q = 0; q < n; ++q) {
          kp.fG[d][a][b][q] = kp.qw[q] * kp.dphi[a][q] * kp.dphi[b][q] / kp.h[d];
          kp.fM[d][a][b][q] = kp.qw[q] * kp.h[d] * kp.phi[a][q] * kp.phi[b][q];
        }
  kp.coef_const = d.coef_const;
  kp.nmodes = d.nmodes;
  kp.rhs_kind = d.rhs_kind;
  kp.rhs_const = d.rhs_const;
  kp.dir_kind = d.dir_kind;
  for (int k = 0; k < 4; ++k) kp.dir_coef[k] = d.dir_coef[k];
}

}  // namespace gdx
