// Device stage of libgdx's symbolic phase.  From the host's conforming dof
// numbering and matrix PRanges (gdx_symbolic.cpp) it builds in HBM:
//   - every part's local CSR: row lengths counted over each row's stencil,
//     offsets by a scan, and the columns of a row placed by rank counting
//     (a column's position is the number of smaller columns in its row, so the
//     row comes out sorted, as the reference's CSR from sorted COO triplets,
//     src/Algebra.jl:1136-1257);
//   - the position pattern of every owned row (stencil slot -> position),
//     deduplicated through a hash table whose slots keep the first owned row
//     (over parts, then rows) that has the key; pattern ids follow those first
//     occurrences, which is the host stage's sequential first-encounter order,
//     and every row's key is compared with its representative's, so that a
//     hash collision is found (the host stage then builds the structure);
//   - the row table over each part's owned node box, and the tiles.
// Everything equals the host stage's output bit for bit.  Host copies of the
// CSRs and the row table are made only when host code asks for them.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "gdx_internal.h"

namespace gdx {
namespace {

constexpr uint16_t kAbs16 = 0xFFFF, kDir16 = 0xFFFE;
constexpr int kTableBits = 18;
constexpr unsigned kTableSize = 1u << kTableBits;
constexpr int kBlock = 256;

struct MeshG {
  long long n[3], nn[3];
  int D, K, nc, dir_kind;
};

// one part, by value; offsets index the concatenation over parts in part order
struct PartG {
  long long loc_lo[3], lnn[3], own_lo[3], own_hi[3], onb_lo[3], onb_n[3];
  long long n_own, nrows;
  long long row0;    // first row of the part among all parts' rows
  long long ogrp0;   // first owned row group (nc consecutive rows of one node)
  long long rinfo0;  // first row-table entry
  long long lnode0, ldof0, own0;
};

__device__ inline bool dirichlet(const MeshG &m, long long g0, long long g1, long long g2) {
  if (m.dir_kind == 0) return false;
  if (m.dir_kind == GDX_DIR_CLAMPED_X0) return g0 == 0;
  if (g0 == 0 || g0 == m.nn[0] - 1 || g1 == 0 || g1 == m.nn[1] - 1) return true;
  return m.D == 3 && (g2 == 0 || g2 == m.nn[2] - 1);
}

// global nodes sharing a cell of [clo, chi) with node i along one direction
__device__ inline void shared_range(int K, long long i, long long clo, long long chi, long long &jlo, long long &jhi) {
  const long long cmin = max((i + K - 1) / K - 1, clo), cmax = min(i / K, chi - 1);
  jlo = K * cmin;
  jhi = K * cmax + K;
  if (cmin > cmax) jlo = 1, jhi = 0;
}

__device__ inline void node_of(const MeshG &m, const PartG &P, long long l, long long g[3]) {
  g[0] = m.K * P.loc_lo[0] + l % P.lnn[0];
  g[1] = m.K * P.loc_lo[1] + (l / P.lnn[0]) % P.lnn[1];
  g[2] = m.K * P.loc_lo[2] + l / (P.lnn[0] * P.lnn[1]);
}

__device__ inline long long lnode_in(const MeshG &m, const PartG &P, long long g0, long long g1, long long g2) {
  const long long l0 = g0 - m.K * P.loc_lo[0], l1 = g1 - m.K * P.loc_lo[1], l2 = g2 - m.K * P.loc_lo[2];
  if (l0 < 0 || l0 >= P.lnn[0] || l1 < 0 || l1 >= P.lnn[1] || l2 < 0 || l2 >= P.lnn[2]) return -1;
  return l0 + P.lnn[0] * (l1 + P.lnn[1] * l2);
}

// the row's stencil range per direction: owned rows couple over the whole
// mesh, ghost rows only through the part's owned cells
__device__ inline void row_ranges(const MeshG &m, const PartG &P, const long long g[3], bool own, long long jlo[3],
                                  long long jhi[3]) {
  for (int d = 0; d < 3; ++d) jlo[d] = jhi[d] = 0;
  for (int d = 0; d < m.D; ++d) {
    if (own) shared_range(m.K, g[d], 0, m.n[d], jlo[d], jhi[d]);
    else shared_range(m.K, g[d], P.own_lo[d], P.own_hi[d], jlo[d], jhi[d]);
  }
}

// SplitMix64 finaliser (Steele, Lea, Flood 2014): spreads an entry's
// (index, key) pair over 64 bits before the order-free sum of a row's hash
__device__ inline unsigned long long mix64(unsigned long long x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

__global__ void k_fill_i64(int64_t *x, long long n, int64_t v) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < n) x[i] = v;
}

__global__ void k_fill_u64(unsigned long long *x, long long n, unsigned long long v) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < n) x[i] = v;
}

// owned row of every owned dof (-1 elsewhere) and the node of every row group
__global__ void k_ldof_orow(PartG P, const int64_t *own_ldofs, int64_t *ldof_orow) {
  const long long o = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (o < P.n_own) ldof_orow[own_ldofs[o]] = o;
}

__global__ void k_row_lnode(PartG P, int nc, const int64_t *row_ldof, const int64_t *ldof_lnode, int64_t *row_lnode,
                            int *err) {
  const long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (q >= P.nrows / nc) return;
  const long long ld = row_ldof[q * nc];
  if (ld % nc != 0) atomicOr(err, 1);  // the rows of a node must be grouped
  row_lnode[q] = ldof_lnode[ld];
}

// row lengths: nc columns per free stencil node of the row's range
__global__ void k_rowlen(MeshG m, PartG P, const int64_t *row_lnode, const int64_t *lnode_dof, const int64_t *ldof_col,
                         long long *rowlen, int *err) {
  const long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (q >= P.nrows / m.nc) return;
  long long g[3], jlo[3], jhi[3];
  node_of(m, P, row_lnode[q], g);
  row_ranges(m, P, g, q * m.nc < P.n_own, jlo, jhi);
  long long cnt = 0;
  for (long long jz = jlo[2]; jz <= jhi[2]; ++jz)
    for (long long jy = jlo[1]; jy <= jhi[1]; ++jy)
      for (long long jx = jlo[0]; jx <= jhi[0]; ++jx) {
        if (dirichlet(m, jx, jy, jz)) continue;
        const long long lj = lnode_in(m, P, jx, jy, jz);
        const long long j0 = lj >= 0 ? lnode_dof[lj] : -1;
        if (j0 < 0 || ldof_col[j0] < 0) {
          atomicOr(err, 1);  // a stencil column missing from the col PRange
          continue;
        }
        ++cnt;
      }
  for (int k = 0; k < m.nc; ++k) rowlen[P.row0 + q * m.nc + k] = cnt * m.nc;
}

// Exclusive scan of n int64 values into out[0..n] (out[n] = total): blocks of
// kBlock threads x kItems values scan in LDS; the block totals are scanned the
// same way and added back.
constexpr int kItems = 8;
constexpr long long kScanTile = (long long)kBlock * kItems;

__global__ __launch_bounds__(kBlock) void k_scan_tile(const long long *in, long long *out, long long n, long long *sums) {
  __shared__ long long part[kBlock];
  const long long base = blockIdx.x * kScanTile + (long long)threadIdx.x * kItems;
  long long v[kItems], acc = 0;
  for (int i = 0; i < kItems; ++i) {
    v[i] = base + i < n ? in[base + i] : 0;
    acc += v[i];
  }
  part[threadIdx.x] = acc;
  __syncthreads();
  for (int off = 1; off < kBlock; off <<= 1) {  // inclusive Hillis-Steele over the thread totals
    const long long t = threadIdx.x >= off ? part[threadIdx.x - off] : 0;
    __syncthreads();
    part[threadIdx.x] += t;
    __syncthreads();
  }
  long long run = part[threadIdx.x] - acc;
  for (int i = 0; i < kItems; ++i) {
    if (base + i < n) out[base + i] = run;
    run += v[i];
  }
  if (threadIdx.x == kBlock - 1) sums[blockIdx.x] = part[kBlock - 1];
}

__global__ void k_scan_add(long long *out, long long n, const long long *offs) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < n) out[i] += offs[i / kScanTile];
}

__global__ void k_copy1(long long *dst, const long long *src) { *dst = *src; }

// Launches of the scan; ws holds the block sums and their scans of every
// level (scan_workspace(n) values).  Stream-ordered, no synchronisation.
long long scan_workspace(long long n) {
  const long long nb = std::max<long long>((n + kScanTile - 1) / kScanTile, 1);
  return nb > 1 ? 2 * nb + 1 + scan_workspace(nb) : 2;
}

void scan_launch(const long long *in, long long *out, long long n, long long *ws, hipStream_t st) {
  const long long nb = std::max<long long>((n + kScanTile - 1) / kScanTile, 1);
  long long *sums = ws, *offs = ws + nb;
  hipLaunchKernelGGL(k_scan_tile, dim3((unsigned)nb), dim3(kBlock), 0, st, in, out, n, sums);
  if (nb > 1) {
    scan_launch(sums, offs, nb, ws + 2 * nb + 1, st);
    hipLaunchKernelGGL(k_scan_add, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, st, out, n, offs);
  }
  // total: the single block's sum, or the scan of the block sums' total
  hipLaunchKernelGGL(k_copy1, dim3(1), dim3(1), 0, st, out + n, nb > 1 ? offs + nb : sums);
}

// workspaces of the scans of one symbolic phase, grown as needed
struct ScanWs {
  hipStream_t st;
  long long *ws = nullptr;
  long long cap = 0;
  explicit ScanWs(hipStream_t s) : st(s) {}
  ~ScanWs() {
    if (!ws) return;
    (void)hipStreamSynchronize(st);  // queued scans may still read it
    (void)hipFree(ws);
  }
};

int scan_exclusive(const long long *in, long long *out, long long n, hipStream_t st, ScanWs &W) {
  const long long need = scan_workspace(n);
  if (need > W.cap) {
    // the previous workspace may still be read by queued scans
    if (W.ws && (hipStreamSynchronize(st) != hipSuccess || hipFree(W.ws) != hipSuccess)) return GDX_ERR_HIP;
    W.ws = nullptr;
    W.cap = 0;
    if (dev_malloc(&W.ws, need * sizeof(long long)) != hipSuccess) return GDX_ERR_ALLOC;
    W.cap = need;
  }
  scan_launch(in, out, n, W.ws, st);
  return hipGetLastError() == hipSuccess ? GDX_OK : GDX_ERR_HIP;
}

// One row group (the nc rows of a node) per LP lanes, one lane per stencil
// node: its columns, their positions in the row, the column ids of the group's
// rows and, for owned rows, the position key and its order-free hash.  The
// own columns come first, in the order of their nodes (an own node's nc
// columns are consecutive), then the ghost columns by id; a column's position
// counts the smaller columns of its kind in the row's compact lists.
constexpr int kMaxComp = 3;

template <int LP, int BS>
__global__ __launch_bounds__(BS) void k_fill_rows(MeshG m, PartG P, const int64_t *row_lnode,
                                                  const int64_t *lnode_dof, const int64_t *ldof_col,
                                                  const long long *rowptr, int *colidx, uint16_t *keys, int KS,
                                                  unsigned long long *hash, int *err) {
  constexpr int GPB = BS / LP;
  __shared__ __align__(16) int ownl[BS];            // per group: first columns of the own nodes
  __shared__ __align__(16) int ghl[BS * kMaxComp];  // per group: ghost columns
  __shared__ int nown[GPB], ngh[GPB];
  __shared__ unsigned long long hs[GPB];
  __shared__ long long geo[GPB][9];  // the group's node and stencil ranges (one lane computes them)
  const int grp = threadIdx.x / LP, s = threadIdx.x % LP, nc = m.nc;
  const long long q = (long long)blockIdx.x * GPB + grp;
  const bool live = q < P.nrows / nc;
  const int W = 2 * m.K + 1, S = m.D == 3 ? W * W * W : W * W, SN = S * nc;
  const bool own_row = live && q * nc < P.n_own;
  if (threadIdx.x < GPB) {
    hs[threadIdx.x] = 0;
    nown[threadIdx.x] = 0;
    ngh[threadIdx.x] = 0;
  }
  // unused list entries read as larger than every column (the rank loops read four at a time)
  ownl[threadIdx.x] = INT_MAX;
  for (int k = 0; k < kMaxComp; ++k) ghl[threadIdx.x * kMaxComp + k] = INT_MAX;
  if (live && s == 0) {
    long long g[3], jlo[3], jhi[3];
    node_of(m, P, row_lnode[q], g);
    row_ranges(m, P, g, own_row, jlo, jhi);
    for (int d = 0; d < 3; ++d) {
      geo[grp][d] = g[d];
      geo[grp][3 + d] = jlo[d];
      geo[grp][6 + d] = jhi[d];
    }
  }
  __syncthreads();
  int kind = 0;  // 0 absent, 1 Dirichlet, 2 own node, 3 ghost node
  int c0 = 0, ck[kMaxComp] = {0, 0, 0};
  if (live && s < S) {
    const long long g[3] = {geo[grp][0], geo[grp][1], geo[grp][2]};
    const long long jlo[3] = {geo[grp][3], geo[grp][4], geo[grp][5]}, jhi[3] = {geo[grp][6], geo[grp][7], geo[grp][8]};
    const long long j0 = g[0] + s % W - m.K, j1 = g[1] + (s / W) % W - m.K,
                    j2 = g[2] + (m.D == 3 ? s / (W * W) - m.K : 0);
    if (j0 >= jlo[0] && j0 <= jhi[0] && j1 >= jlo[1] && j1 <= jhi[1] && j2 >= jlo[2] && j2 <= jhi[2]) {
      if (dirichlet(m, j0, j1, j2)) {
        kind = 1;
      } else {
        // the same test as the row length's (the node's first column), so a
        // row never holds more columns than were counted for it
        const long long lj = lnode_in(m, P, j0, j1, j2);
        const long long d0 = lj >= 0 ? lnode_dof[lj] : -1;
        const long long cc = d0 >= 0 ? ldof_col[d0] : -1;
        if (cc < 0 || cc > INT_MAX - kMaxComp) {
          atomicOr(err, 1);
        } else if (cc < P.n_own) {
          kind = 2;
          c0 = (int)cc;
          for (int k = 1; k < nc; ++k)
            if (ldof_col[d0 + k] != cc + k) atomicOr(err, 1);
          ownl[grp * LP + atomicAdd(&nown[grp], 1)] = c0;
        } else {
          kind = 3;
          for (int k = 0; k < nc; ++k) {
            const long long c = ldof_col[d0 + k];
            if (c < P.n_own || c > INT_MAX - 1) atomicOr(err, 1);
            ck[k] = (int)c;
            ghl[grp * LP * kMaxComp + atomicAdd(&ngh[grp], 1)] = ck[k];
          }
        }
      }
    }
  }
  __syncthreads();
  uint16_t key[kMaxComp] = {kAbs16, kAbs16, kAbs16};
  const long long r0 = P.row0 + q * nc;
  if (kind == 1) {
    for (int k = 0; k < nc; ++k) key[k] = kDir16;
  } else if (kind == 2) {
    const int4 *ol = reinterpret_cast<const int4 *>(ownl + grp * LP);
    int r = 0;
    for (int i = 0; i < (nown[grp] + 3) / 4; ++i) {
      const int4 v = ol[i];
      r += (v.x < c0) + (v.y < c0) + (v.z < c0) + (v.w < c0);
    }
    for (int k = 0; k < nc; ++k) {
      key[k] = (uint16_t)(nc * r + k);
      for (int kk = 0; kk < nc; ++kk) colidx[rowptr[r0 + kk] + nc * r + k] = c0 + k;
    }
  } else if (kind == 3) {
    const int4 *gl = reinterpret_cast<const int4 *>(ghl + grp * LP * kMaxComp);
    const int ng = ngh[grp], base = nc * nown[grp];
    for (int k = 0; k < nc; ++k) {
      int r = 0;
      for (int i = 0; i < (ng + 3) / 4; ++i) {
        const int4 v = gl[i];
        r += (v.x < ck[k]) + (v.y < ck[k]) + (v.z < ck[k]) + (v.w < ck[k]);
      }
      key[k] = (uint16_t)(base + r);
      for (int kk = 0; kk < nc; ++kk) colidx[rowptr[r0 + kk] + base + r] = ck[k];
    }
  }
  if (own_row && s < S) {
    unsigned long long h = 0;
    for (int k = 0; k < nc; ++k) {
      const int e = s * nc + k;
      keys[(P.ogrp0 + q) * KS + e] = key[k];
      h += mix64(((unsigned long long)e << 16) | key[k]);
    }
    atomicAdd(&hs[grp], h);
    if (s == 0)
      for (int e = SN; e < KS; ++e) keys[(P.ogrp0 + q) * KS + e] = kAbs16;  // padding reads as absent
  }
  __syncthreads();
  if (own_row && s == 0) hash[P.ogrp0 + q] = hs[grp] | 1ull;  // 0 marks an empty table slot
}

// first owned row group of every distinct hash (open addressing, linear
// probing).  Neighbouring rows mostly share a pattern: the lanes of a wave with
// the same hash elect their first lane, which alone probes the table and takes
// the minimum (the wave's smallest row of that hash), and passes the slot on.
__global__ __launch_bounds__(256) void k_insert(const unsigned long long *hash, long long n, unsigned long long *tkey,
                                                unsigned long long *tmin, unsigned *slot, int *err) {
  const long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  const bool live = q < n;
  const unsigned long long h = live ? hash[q] : 0ull;
  const int lane = (int)(threadIdx.x & 63);
  int leader = lane;
  for (int l = 63; l >= 0; --l)
    if (__shfl(h, l) == h) leader = l;  // the lowest lane with this hash
  unsigned s = (unsigned)(h >> 1) & (kTableSize - 1);
  if (live && leader == lane) {
    bool placed = false;
    for (unsigned t = 0; t < kTableSize && !placed; ++t) {
      const unsigned long long old = atomicCAS(&tkey[s], 0ull, h);
      if (old == 0ull || old == h) {
        if (tmin[s] > (unsigned long long)q) atomicMin(&tmin[s], (unsigned long long)q);
        placed = true;
      } else {
        s = (s + 1) & (kTableSize - 1);
      }
    }
    if (!placed) atomicOr(err, 2);  // table full
  }
  s = (unsigned)__shfl((int)s, leader);
  if (live) slot[q] = s;
}

__global__ void k_collect(const unsigned long long *tkey, const unsigned long long *tmin, unsigned *cnt,
                          unsigned long long *out_min, unsigned *out_slot) {
  const unsigned s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= kTableSize || tkey[s] == 0ull) return;
  const unsigned i = atomicAdd(cnt, 1u);
  out_min[i] = tmin[s];
  out_slot[i] = s;
}

__global__ void k_gather_i64(const int64_t *src, const long long *idx, long long n, long long *out) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < n) out[i] = src[idx[i]];
}

__global__ void k_gather_keys(const uint16_t *keys, int KS, const long long *rep, int npat, uint16_t *out) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= (long long)npat * KS) return;
  out[i] = keys[rep[i / KS] * KS + i % KS];
}

// pattern id of every owned row group; its key must equal its representative's
__global__ void k_pattern_ids(const uint16_t *keys, int KS, long long n, const unsigned *slot,
                              const unsigned long long *tmin, const int *tpid, int *pid, int *err) {
  const long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (q >= n) return;
  const unsigned s = slot[q];
  const long long rep = (long long)tmin[s];
  const uint64_t *a = reinterpret_cast<const uint64_t *>(keys + q * KS);
  const uint64_t *b = reinterpret_cast<const uint64_t *>(keys + rep * KS);
  bool same = true;
  for (int w = 0; w < KS / 4; ++w) same = same && a[w] == b[w];
  if (!same) atomicOr(err, 4);  // hash collision
  pid[q] = tpid[s];
}

__global__ void k_rowinfo(MeshG m, PartG P, const int64_t *lnode_dof, const int64_t *ldof_orow,
                          const long long *rowptr, const int *pid_grp, const int *pat_dir, RowInfo *out, int *err) {
  const long long li = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (li >= P.onb_n[0] * P.onb_n[1] * P.onb_n[2]) return;
  const long long g0 = P.onb_lo[0] + li % P.onb_n[0], g1 = P.onb_lo[1] + (li / P.onb_n[0]) % P.onb_n[1],
                  g2 = P.onb_lo[2] + li / (P.onb_n[0] * P.onb_n[1]);
  const long long l = lnode_in(m, P, g0, g1, g2);
  const long long ld = l >= 0 ? lnode_dof[l] : -1;
  RowInfo ri{0, 0, -1, 0};
  if (ld >= 0) {
    const long long o = ldof_orow[ld];
    if (o < 0) {
      atomicOr(err, 8);  // the owned node box holds a non-owned dof
    } else {
      const long long grow = P.row0 + o;
      const int pid = pid_grp[P.ogrp0 + o / m.nc];
      ri.rowstart = rowptr[grow];
      ri.bidx = grow;
      ri.pid = pid;
      ri.lenf = (int)(rowptr[grow + 1] - rowptr[grow]) | (pat_dir[pid] << 16);
    }
  }
  out[P.rinfo0 + li] = ri;
}

// distinct patterns of each candidate tile in first-occurrence order (rows z,
// y, x); the rows get their tile-local index in bits 17.. of lenf
__global__ void k_tile_patterns(const Tile *cand, int ncand, int cap, RowInfo *rowinfo, int *lists, int *npat,
                                int *any, int *err) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ncand) return;
  const Tile T = cand[t];
  int *list = lists + (long long)t * cap;
  int n = 0, has = 0;
  for (int z = 0; z < T.ext[2]; ++z)
    for (int y = 0; y < T.ext[1]; ++y)
      for (int x = 0; x < T.ext[0]; ++x) {
        const long long li = ((long long)(T.org[2] - T.onb_lo[2] + z) * T.onb_n[1] + (T.org[1] - T.onb_lo[1] + y)) *
                                 T.onb_n[0] +
                             (T.org[0] - T.onb_lo[0] + x);
        RowInfo &ri = rowinfo[T.rowinfo_base + li];
        if (ri.pid < 0) continue;
        has = 1;
        int l = 0;
        while (l < n && list[l] != ri.pid) ++l;
        if (l == n) list[n++] = ri.pid;
        if (l >= (1 << 14)) atomicOr(err, 16);
        ri.lenf = (ri.lenf & 0x1FFFF) | (l << 17);
      }
  npat[t] = n;
  any[t] = has;
}

// The same with one workgroup per candidate tile, for pattern counts that fit
// in LDS: every pattern's first row in the tile (atomic minimum), then each
// distinct pattern's rank among those first rows is its tile-local index.
constexpr int kTileThreads = 256;

__global__ __launch_bounds__(kTileThreads) void k_tile_patterns_wg(const Tile *cand, int ncand, int cap, int npat_all,
                                                                   RowInfo *rowinfo, int *lists, int *npat, int *any,
                                                                   int *err) {
  extern __shared__ int lds[];
  int *first = lds, *dl = lds + npat_all, *rk = lds + 2 * npat_all;
  __shared__ int cnt;
  const int t = blockIdx.x;
  if (t >= ncand) return;
  const Tile T = cand[t];
  const int nrow = T.ext[0] * T.ext[1] * T.ext[2];
  for (int p = threadIdx.x; p < npat_all; p += kTileThreads) first[p] = INT_MAX;
  if (threadIdx.x == 0) cnt = 0;
  __syncthreads();
  auto entry = [&](int i) -> RowInfo & {
    const int x = i % T.ext[0], y = (i / T.ext[0]) % T.ext[1], z = i / (T.ext[0] * T.ext[1]);
    const long long li = ((long long)(T.org[2] - T.onb_lo[2] + z) * T.onb_n[1] + (T.org[1] - T.onb_lo[1] + y)) *
                             T.onb_n[0] +
                         (T.org[0] - T.onb_lo[0] + x);
    return rowinfo[T.rowinfo_base + li];
  };
  for (int i = threadIdx.x; i < nrow; i += kTileThreads) {
    const int pid = entry(i).pid;
    if (pid >= 0) atomicMin(&first[pid], i);
  }
  __syncthreads();
  for (int p = threadIdx.x; p < npat_all; p += kTileThreads)
    if (first[p] != INT_MAX) dl[atomicAdd(&cnt, 1)] = p;
  __syncthreads();
  const int n = cnt;
  for (int u = threadIdx.x; u < n; u += kTileThreads) {
    const int pu = dl[u], fu = first[pu];
    int r = 0;
    for (int v = 0; v < n; ++v) r += first[dl[v]] < fu;
    rk[pu] = r;
    lists[(long long)t * cap + r] = pu;
  }
  __syncthreads();
  if (n > (1 << 14)) atomicOr(err, 16);
  for (int i = threadIdx.x; i < nrow; i += kTileThreads) {
    RowInfo &ri = entry(i);
    if (ri.pid >= 0) ri.lenf = (ri.lenf & 0x1FFFF) | (rk[ri.pid] << 17);
  }
  if (threadIdx.x == 0) {
    npat[t] = n;
    any[t] = n > 0;
  }
}

__global__ void k_tile_write(const int *lists, int cap, const int *npat, const long long *off, int ncand,
                             int *tile_pats) {
  const int t = blockIdx.x;
  if (t >= ncand || off[t] < 0) return;
  for (int i = threadIdx.x; i < npat[t]; i += blockDim.x) tile_pats[off[t] + i] = lists[(long long)t * cap + i];
}

unsigned grid_of(long long n) { return (unsigned)std::max<long long>((n + kBlock - 1) / kBlock, 1); }

// device allocations of one call, freed together once the stream's queued
// work (which may still read them, after an early error return) is done
struct Scratch {
  hipStream_t st;
  std::vector<void *> ptrs;
  explicit Scratch(hipStream_t s) : st(s) {}
  ~Scratch() {
    (void)hipStreamSynchronize(st);
    for (void *q : ptrs) (void)hipFree(q);
  }
  template <class T>
  T *get(long long n) {
    void *q = nullptr;
    if (dev_malloc(&q, std::max<long long>(n, 1) * sizeof(T)) != hipSuccess) return nullptr;
    ptrs.push_back(q);
    return static_cast<T *>(q);
  }
};

template <class T>
bool put(T *dst, const std::vector<T> &src) {
  return src.empty() || hipMemcpy(dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice) == hipSuccess;
}

template <int LP>
void launch_fill(const MeshG &m, const PartG &G, hipStream_t st, const int64_t *row_lnode, const int64_t *lnode_dof,
                 const int64_t *ldof_col, const long long *rowptr, int *colidx, uint16_t *keys, int KS,
                 unsigned long long *hash, int *err) {
  constexpr int BS = LP > kBlock ? LP : kBlock;
  constexpr int GPB = BS / LP;
  const long long ng = G.nrows / m.nc;
  if (ng == 0) return;
  hipLaunchKernelGGL((k_fill_rows<LP, BS>), dim3((unsigned)((ng + GPB - 1) / GPB)), dim3(BS), 0, st, m, G, row_lnode,
                     lnode_dof, ldof_col, rowptr, colidx, keys, KS, hash, err);
}

}  // namespace

namespace {

// ---- numbering, owners, global ids and PRanges on the device ----

struct PartN {
  long long loc_lo[3], lnn[3], lc[3], own_lo[3], own_hi[3];
  long long nlnode, lnode0;  // local nodes; offset among all parts' nodes
  long long ncl, ef0;        // local cells; offset of the edge/face flags
  long long ldof0, nldof, own0, n_own, gid0;
  long long key0, nkey;      // ghost-column key space
  long long row0, col0;
  int rank;
};

// the reference cell: its dofs as order-K lattice offsets; its edges and faces
// as doubled-lattice offsets in polytope order; a doubled offset's index in
// its dimension's list
struct NumTabs {
  int ncd, nE, nF;
  int off[27][3];
  int E[12][3], F[6][3];
  int fidx[27];
};

struct OwnerTabs {
  const int *dim[3];  // part index along d owning lattice node N
  int np[3];
};

// the first local cell holding doubled-lattice coordinate pt along a direction
__device__ inline long long first_cell(long long pt) { return (pt & 1) ? (pt - 1) / 2 : max(pt / 2 - 1, 0LL); }

__device__ inline long long pn_lnode_in(const MeshG &m, const PartN &P, long long g0, long long g1, long long g2) {
  const long long l0 = g0 - m.K * P.loc_lo[0], l1 = g1 - m.K * P.loc_lo[1], l2 = g2 - m.K * P.loc_lo[2];
  if (l0 < 0 || l0 >= P.lnn[0] || l1 < 0 || l1 >= P.lnn[1] || l2 < 0 || l2 >= P.lnn[2]) return -1;
  return l0 + P.lnn[0] * (l1 + P.lnn[1] * l2);
}

__device__ inline void pn_node(const MeshG &m, const PartN &P, long long l, long long g[3]) {
  g[0] = m.K * P.loc_lo[0] + l % P.lnn[0];
  g[1] = m.K * P.loc_lo[1] + (l / P.lnn[0]) % P.lnn[1];
  g[2] = m.K * P.loc_lo[2] + l / (P.lnn[0] * P.lnn[1]);
}

// whether (cell c, edge or face t) is where the edge or face is numbered: the
// first local cell holding it (edges block, then faces block)
__global__ void k_ef_flags(MeshG m, PartN P, NumTabs T, long long *flag) {
  const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  const long long nE = P.ncl * T.nE;
  if (t >= nE + P.ncl * T.nF) return;
  long long c;
  const int *o;
  if (t < nE) {
    c = t / T.nE;
    o = T.E[t % T.nE];
  } else {
    c = (t - nE) / T.nF;
    o = T.F[(t - nE) % T.nF];
  }
  const long long cc[3] = {c % P.lc[0], (c / P.lc[0]) % P.lc[1], c / (P.lc[0] * P.lc[1])};
  bool first = true;
  for (int d = 0; d < m.D; ++d) first = first && first_cell(2 * cc[d] + o[d]) == cc[d];
  flag[P.ef0 + t] = first ? 1 : 0;
}

// a local node's position in the numbering order (vertices lexicographic, then
// first edges, first faces, cell interiors; every node for order 1) and
// whether it is free; free flags are stored by position
__global__ void k_node_pos(MeshG m, PartN P, NumTabs T, const long long *efscan, long long nef, long long *pos,
                           long long *freef) {
  const long long l = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (l >= P.nlnode) return;
  const long long x = l % P.lnn[0], y = (l / P.lnn[0]) % P.lnn[1], z = l / (P.lnn[0] * P.lnn[1]);
  long long ps = l;
  if (m.K == 2) {
    const int odd = (int)(x & 1) + (int)(y & 1) + (m.D == 3 ? (int)(z & 1) : 0);
    const long long nv = (P.lc[0] + 1) * (P.lc[1] + 1) * (m.D == 3 ? P.lc[2] + 1 : 1);
    if (odd == 0) {
      ps = x / 2 + (P.lc[0] + 1) * (y / 2 + (P.lc[1] + 1) * (z / 2));
    } else {
      const long long c0 = first_cell(x), c1 = first_cell(y), c2 = m.D == 3 ? first_cell(z) : 0;
      const long long cell = c0 + P.lc[0] * (c1 + P.lc[1] * c2);
      if (odd == m.D) {
        ps = nv + nef + cell;
      } else {
        const int o = (int)(x - 2 * c0) + 3 * (int)(y - 2 * c1) + 9 * (int)(m.D == 3 ? z - 2 * c2 : 0);
        const int f = T.fidx[o];
        const long long t = odd == 1 ? cell * T.nE + f : P.ncl * T.nE + cell * T.nF + f;
        ps = nv + efscan[P.ef0 + t] - efscan[P.ef0];
      }
    }
  }
  pos[P.lnode0 + l] = ps;
  long long g[3];
  pn_node(m, P, l, g);
  freef[P.lnode0 + ps] = dirichlet(m, g[0], g[1], g[2]) ? 0 : 1;
}

__global__ void k_lnode_dof(MeshG m, PartN P, const long long *pos, const long long *freef, const long long *fscan,
                            int64_t *lnode_dof, int64_t *ldof_lnode) {
  const long long l = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (l >= P.nlnode) return;
  const long long ps = P.lnode0 + pos[P.lnode0 + l];
  if (!freef[ps]) {
    lnode_dof[P.lnode0 + l] = -1;
    return;
  }
  const long long dd = (long long)m.nc * (fscan[ps] - fscan[P.lnode0]);
  lnode_dof[P.lnode0 + l] = dd;
  for (int k = 0; k < m.nc; ++k) ldof_lnode[P.ldof0 + dd + k] = l;
}

// owner of every dof (the owner of its node: generate_gids' max-owner rule)
// every dof has one node of the part (the order covered every node once)
__global__ void k_check_ldofs(PartN P, const int64_t *ldof_lnode, int *err) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= P.nldof) return;
  const long long l = ldof_lnode[P.ldof0 + i];
  if (l < 0 || l >= P.nlnode) atomicOr(err, 4);
}

__global__ void k_owners(MeshG m, PartN P, OwnerTabs O, const int64_t *ldof_lnode, int32_t *ldof_owner,
                         long long *ownf) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= P.nldof) return;
  long long g[3];
  pn_node(m, P, ldof_lnode[P.ldof0 + i], g);
  const int o = (O.dim[2][g[2]] * O.np[1] + O.dim[1][g[1]]) * O.np[0] + O.dim[0][g[0]];
  ldof_owner[P.ldof0 + i] = o;
  ownf[P.ldof0 + i] = o == P.rank ? 1 : 0;
}

// owned dofs: own index, global id, column id
__global__ void k_own_ids(PartN P, const long long *ownf, const long long *oscan, int64_t *own_ldofs,
                          int64_t *ldof_gid, int64_t *ldof_col) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= P.nldof) return;
  if (!ownf[P.ldof0 + i]) {
    ldof_col[P.ldof0 + i] = -1;
    return;
  }
  const long long o = oscan[P.ldof0 + i] - oscan[P.ldof0];
  own_ldofs[P.own0 + o] = i;
  ldof_gid[P.ldof0 + i] = P.gid0 + o;
  ldof_col[P.ldof0 + i] = o;
}

// ghost dofs: global id from the owner's numbering of the node; ghost rows
// (SubAssembledRows) are the ghosts on the closure of the owned cells
__global__ void k_ghost_ids(MeshG m, const PartN *parts, int r, int fa, const int32_t *ldof_owner,
                            const int64_t *ldof_lnode, const int64_t *lnode_dof, int64_t *ldof_gid, long long *grf,
                            int *err) {
  const PartN P = parts[r];
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= P.nldof) return;
  const int ow = ldof_owner[P.ldof0 + i];
  if (ow == P.rank) {
    grf[P.ldof0 + i] = 0;
    return;
  }
  long long g[3];
  pn_node(m, P, ldof_lnode[P.ldof0 + i], g);
  const PartN &Q = parts[ow];
  const long long lq = pn_lnode_in(m, Q, g[0], g[1], g[2]);
  const long long dq = lq >= 0 ? lnode_dof[Q.lnode0 + lq] : -1;
  if (dq < 0) atomicOr(err, 1);  // a ghost dof missing in its owner
  else ldof_gid[P.ldof0 + i] = ldof_gid[Q.ldof0 + dq + i % m.nc];
  bool in = !fa;
  for (int d = 0; d < m.D; ++d) in = in && g[d] >= m.K * P.own_lo[d] && g[d] <= m.K * P.own_hi[d];
  grf[P.ldof0 + i] = in ? 1 : 0;
}

__global__ void k_rows(PartN P, const long long *ownf, const long long *oscan, const long long *grf,
                       const long long *grscan, const int64_t *ldof_gid, const int32_t *ldof_owner, int64_t *row_ldof,
                       int64_t *row_gid, int32_t *row_owner) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= P.nldof) return;
  long long row;
  if (ownf[P.ldof0 + i]) row = oscan[P.ldof0 + i] - oscan[P.ldof0];
  else if (grf[P.ldof0 + i]) row = P.n_own + grscan[P.ldof0 + i] - grscan[P.ldof0];
  else return;
  row_ldof[P.row0 + row] = i;
  row_gid[P.row0 + row] = ldof_gid[P.ldof0 + i];
  row_owner[P.row0 + row] = ldof_owner[P.ldof0 + i];
}

// One pass of the ghost-column scan: the cells of a box (of part s), those
// holding a row of interest (any free dof when want < 0, else a dof owned by
// want) visited in order; each of their non-owned dofs (in part r's
// numbering) keeps its first visit, as a key ordered like the visits
// (SubAssembledRows: cells, then components, then cell dofs), or its own id
// (FullyAssembledRows: ordered by local id).
__global__ void k_visit(MeshG m, NumTabs T, const PartN *parts, int r, int s, int want, int fa, long long lo0,
                        long long lo1, long long lo2, long long n0, long long n1, long long n2, long long koff,
                        const int64_t *lnode_dof, const int32_t *ldof_owner, unsigned long long *minkey, int *err) {
  const long long ci = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (ci >= n0 * n1 * n2) return;
  const long long c[3] = {lo0 + ci % n0, lo1 + (ci / n0) % n1, lo2 + ci / (n0 * n1)};
  const PartN &S = parts[s];
  bool has = false;
  for (int a = 0; a < T.ncd && !has; ++a) {
    const long long ls = pn_lnode_in(m, S, m.K * c[0] + T.off[a][0], m.K * c[1] + T.off[a][1], m.K * c[2] + T.off[a][2]);
    if (ls < 0) {
      atomicOr(err, 2);
      return;
    }
    const long long i = lnode_dof[S.lnode0 + ls];
    has = i >= 0 && (want < 0 || ldof_owner[S.ldof0 + i] == want);
  }
  if (!has) return;
  const PartN &P = parts[r];
  for (int k = 0; k < m.nc; ++k)
    for (int b = 0; b < T.ncd; ++b) {
      const long long lp = pn_lnode_in(m, P, m.K * c[0] + T.off[b][0], m.K * c[1] + T.off[b][1], m.K * c[2] + T.off[b][2]);
      if (lp < 0) {
        atomicOr(err, 2);  // a visited cell outside the part's local cells
        continue;
      }
      long long j = lnode_dof[P.lnode0 + lp];
      if (j < 0) continue;
      j += k;
      if (j >= P.nldof) {
        atomicOr(err, 2);
        continue;
      }
      if (ldof_owner[P.ldof0 + j] == P.rank) continue;
      const unsigned long long key = fa ? (unsigned long long)j : (unsigned long long)(koff + (ci * m.nc + k) * T.ncd + b);
      atomicMin(&minkey[P.ldof0 + j], key);
    }
}

__global__ void k_key_flags(PartN P, const unsigned long long *minkey, long long *kf) {
  const long long j = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (j >= P.nldof) return;
  const unsigned long long k = minkey[P.ldof0 + j];
  if (k != ~0ull) kf[P.key0 + (long long)k] = 1;
}

__global__ void k_cols(PartN P, const long long *ownf, const long long *oscan, const unsigned long long *minkey,
                       const long long *kscan, const int64_t *ldof_gid, const int32_t *ldof_owner, int64_t *ldof_col,
                       int64_t *col_gid, int32_t *col_owner) {
  const long long j = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (j >= P.nldof) return;
  long long col;
  if (ownf[P.ldof0 + j]) {
    col = oscan[P.ldof0 + j] - oscan[P.ldof0];
  } else {
    const unsigned long long k = minkey[P.ldof0 + j];
    if (k == ~0ull) return;
    col = P.n_own + kscan[P.key0 + (long long)k] - kscan[P.key0];
    ldof_col[P.ldof0 + j] = col;
  }
  col_gid[P.col0 + col] = ldof_gid[P.ldof0 + j];
  col_owner[P.col0 + col] = ldof_owner[P.ldof0 + j];
}

__global__ void k_gid_node(MeshG m, PartN P, const int64_t *ldof_lnode, const int64_t *ldof_gid, long long ngdofs,
                           long long *gid_node) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= P.nldof) return;
  long long g[3];
  pn_node(m, P, ldof_lnode[P.ldof0 + i], g);
  const long long gid = ldof_gid[P.ldof0 + i];
  if (gid >= 0 && gid < ngdofs) gid_node[gid] = g[0] + m.nn[0] * (g[1] + m.nn[1] * g[2]);
}

__global__ void k_gather_ll(const long long *src, const long long *idx, int n, long long *out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = src[idx[i]];
}

// src[idx[i]] for every i, read back with one copy (stream-ordered)
int read_at(const long long *src, const std::vector<long long> &idx, std::vector<long long> &out, hipStream_t st) {
  const int n = (int)idx.size();
  out.assign(n, 0);
  if (n == 0) return GDX_OK;
  long long *d = nullptr;
  if (dev_malloc(&d, 2 * n * sizeof(long long)) != hipSuccess) return GDX_ERR_ALLOC;
  int rc = GDX_OK;
  if (hipMemcpyAsync(d, idx.data(), n * sizeof(long long), hipMemcpyHostToDevice, st) != hipSuccess) rc = GDX_ERR_HIP;
  if (!rc) hipLaunchKernelGGL(k_gather_ll, dim3((n + 255) / 256), dim3(256), 0, st, src, d, n, d + n);
  if (!rc && (hipGetLastError() != hipSuccess ||
              hipMemcpyAsync(out.data(), d + n, n * sizeof(long long), hipMemcpyDeviceToHost, st) != hipSuccess ||
              hipStreamSynchronize(st) != hipSuccess))
    rc = GDX_ERR_HIP;
  (void)hipFree(d);
  return rc;
}

MeshG mesh_of(const gdx_problem *p) {
  MeshG m{};
  for (int k = 0; k < 3; ++k) {
    m.n[k] = k < p->D ? p->desc.cells[k] : 1;
    m.nn[k] = k < p->D ? p->K * m.n[k] + 1 : 1;
  }
  m.D = p->D;
  m.K = p->K;
  m.nc = p->ncomp;
  m.dir_kind = p->desc.dir_kind;
  return m;
}

// the per-part view of the device numbering
PartN part_view(const gdx_problem *p, int r) {
  const Part &P = p->P[r];
  PartN n{};
  for (int k = 0; k < 3; ++k) {
    n.loc_lo[k] = P.loc_lo[k];
    n.lnn[k] = P.lnn[k];
    n.lc[k] = k < p->D ? P.loc_hi[k] - P.loc_lo[k] : 1;
    n.own_lo[k] = P.own_lo[k];
    n.own_hi[k] = P.own_hi[k];
  }
  n.nlnode = P.lnn[0] * P.lnn[1] * P.lnn[2];
  n.lnode0 = P.lnode_base;
  n.ncl = n.lc[0] * n.lc[1] * n.lc[2];
  n.ldof0 = P.ldof_base;
  n.nldof = P.n_ldofs;
  n.own0 = P.own_base;
  n.n_own = P.n_own;
  n.row0 = P.row_base;
  n.col0 = P.col_base;
  n.rank = r;
  return n;
}

}  // namespace

void free_numbering(gdx_problem *p) {
  DevNumbering &d = p->dn;
  void *ptrs[] = {d.lnode_dof, d.ldof_lnode, d.ldof_gid, d.own_ldofs, d.ldof_col, d.row_ldof,
                  d.row_gid,   d.col_gid,    d.ldof_owner, d.row_owner, d.col_owner};
  for (void *q : ptrs)
    if (q) (void)hipFree(q);
  d = DevNumbering{};
  p->dev_numbering = false;
  p->host_fe_stale = false;
}

int numbering_device(gdx_problem *p, const std::vector<int> owner_dim[3], const int off[27][3], int ncd,
                     const int (*E)[3], int nE, const int (*F)[3], int nF) {
  hipStream_t st = p->ctx->stream;
  free_numbering(p);
  const int nparts = p->nparts, nc = p->ncomp, K = p->K, D = p->D;
  const bool fa = p->desc.strategy == GDX_FULLY_ASSEMBLED_ROWS;
  const MeshG m = mesh_of(p);
  NumTabs T{};
  T.ncd = ncd;
  T.nE = K == 2 ? nE : 0;
  T.nF = K == 2 ? nF : 0;
  for (int a = 0; a < ncd; ++a)
    for (int d = 0; d < 3; ++d) T.off[a][d] = off[a][d];
  for (int f = 0; f < 27; ++f) T.fidx[f] = -1;
  for (int f = 0; f < nE; ++f) {
    for (int d = 0; d < 3; ++d) T.E[f][d] = E[f][d];
    T.fidx[E[f][0] + 3 * E[f][1] + 9 * E[f][2]] = f;
  }
  for (int f = 0; f < nF; ++f) {
    for (int d = 0; d < 3; ++d) T.F[f][d] = F[f][d];
    T.fidx[F[f][0] + 3 * F[f][1] + 9 * F[f][2]] = f;
  }
  Scratch X(st);
  ScanWs SW(st);
  int *d_err = X.get<int>(1);
  std::vector<int> odim;
  long long odoff[3];
  for (int d = 0; d < 3; ++d) {
    odoff[d] = (long long)odim.size();
    odim.insert(odim.end(), owner_dim[d].begin(), owner_dim[d].end());
  }
  int *d_odim = X.get<int>((long long)odim.size());
  if (!d_err || !d_odim || !put(d_odim, odim) || hipMemsetAsync(d_err, 0, sizeof(int), st) != hipSuccess)
    return GDX_ERR_ALLOC;
  OwnerTabs O{};
  for (int d = 0; d < 3; ++d) {
    O.dim[d] = d_odim + odoff[d];
    O.np[d] = d < D ? p->desc.parts[d] : 1;
  }
  // 1. node positions in the numbering order and free flags
  std::vector<PartN> V(nparts);
  long long nodes = 0, efs = 0;
  for (int r = 0; r < nparts; ++r) {
    Part &P = p->P[r];
    P.lnode_base = nodes;
    V[r] = part_view(p, r);
    V[r].ef0 = efs;
    nodes += V[r].nlnode;
    efs += V[r].ncl * (T.nE + T.nF);
  }
  long long *efl = X.get<long long>(efs), *efscan = X.get<long long>(efs + 1), *pos = X.get<long long>(nodes),
            *freef = X.get<long long>(nodes), *fscan = X.get<long long>(nodes + 1);
  if (!efl || !efscan || !pos || !freef || !fscan) return GDX_ERR_ALLOC;
  int rc;
  if (efs > 0) {
    for (int r = 0; r < nparts; ++r)
      hipLaunchKernelGGL(k_ef_flags, dim3(grid_of(V[r].ncl * (T.nE + T.nF))), dim3(kBlock), 0, st, m, V[r], T, efl);
    if ((rc = scan_exclusive(efl, efscan, efs, st, SW))) return rc;
  }
  std::vector<long long> nef(nparts, 0), at, val;
  for (int r = 0; r < nparts && efs > 0; ++r) at.push_back(V[r].ef0);
  if (efs > 0) {
    at.push_back(efs);
    if ((rc = read_at(efscan, at, val, st))) return rc;
  }
  for (int r = 0; r < nparts && efs > 0; ++r) {
    nef[r] = val[r + 1] - val[r];
    const long long nv = (V[r].lc[0] + 1) * (V[r].lc[1] + 1) * (D == 3 ? V[r].lc[2] + 1 : 1);
    if (nv + nef[r] + V[r].ncl != V[r].nlnode) return kUseHost;  // the order must cover every node once
  }
  // positions not hit stay non-free (a numbering that missed a node is caught below)
  hipLaunchKernelGGL(k_fill_i64, dim3(grid_of(nodes)), dim3(kBlock), 0, st, reinterpret_cast<int64_t *>(freef), nodes,
                     (int64_t)0);
  for (int r = 0; r < nparts; ++r)
    hipLaunchKernelGGL(k_node_pos, dim3(grid_of(V[r].nlnode)), dim3(kBlock), 0, st, m, V[r], T, efscan, nef[r], pos,
                       freef);
  if (hipGetLastError() != hipSuccess) return GDX_ERR_HIP;
  if ((rc = scan_exclusive(freef, fscan, nodes, st, SW))) return rc;
  // 2. dofs of the free nodes (nc consecutive per node)
  long long ldofs = 0;
  std::vector<long long> fb;
  at.clear();
  for (int r = 0; r <= nparts; ++r) at.push_back(r < nparts ? V[r].lnode0 : nodes);
  if ((rc = read_at(fscan, at, fb, st))) return rc;
  for (int r = 0; r < nparts; ++r) {
    Part &P = p->P[r];
    P.n_ldofs = nc * (fb[r + 1] - fb[r]);
    P.ldof_base = ldofs;
    V[r].ldof0 = ldofs;
    V[r].nldof = P.n_ldofs;
    ldofs += P.n_ldofs;
  }
  DevNumbering &dn = p->dn;
  auto alloc = [](auto **q, long long n) {
    return dev_malloc(reinterpret_cast<void **>(q), std::max<long long>(n, 1) * sizeof(**q)) == hipSuccess;
  };
  if (!alloc(&dn.lnode_dof, nodes) || !alloc(&dn.ldof_lnode, ldofs) || !alloc(&dn.ldof_owner, ldofs) ||
      !alloc(&dn.ldof_gid, ldofs) || !alloc(&dn.ldof_col, ldofs)) {
    free_numbering(p);
    return GDX_ERR_ALLOC;
  }
  auto fail = [&](int code) {
    (void)hipStreamSynchronize(st);
    free_numbering(p);
    return code;
  };
  hipLaunchKernelGGL(k_fill_i64, dim3(grid_of(ldofs)), dim3(kBlock), 0, st, dn.ldof_lnode, ldofs, (int64_t)-1);
  for (int r = 0; r < nparts; ++r)
    hipLaunchKernelGGL(k_lnode_dof, dim3(grid_of(V[r].nlnode)), dim3(kBlock), 0, st, m, V[r], pos, freef, fscan,
                       dn.lnode_dof, dn.ldof_lnode);
  for (int r = 0; r < nparts; ++r)
    hipLaunchKernelGGL(k_check_ldofs, dim3(grid_of(V[r].nldof)), dim3(kBlock), 0, st, V[r], dn.ldof_lnode, d_err);
  {
    int err = 0;
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess ||
        hipMemcpy(&err, d_err, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)
      return fail(GDX_ERR_HIP);
    if (err) return fail(kUseHost);
  }
  // 3. owners, owned dofs, global ids (rank-blocked: part r's owned dofs after
  // those of parts < r, src/FESpaces.jl:139-261)
  long long *ownf = X.get<long long>(ldofs), *oscan = X.get<long long>(ldofs + 1), *grf = X.get<long long>(ldofs),
            *grscan = X.get<long long>(ldofs + 1);
  if (!ownf || !oscan || !grf || !grscan) return fail(GDX_ERR_ALLOC);
  for (int r = 0; r < nparts; ++r)
    hipLaunchKernelGGL(k_owners, dim3(grid_of(V[r].nldof)), dim3(kBlock), 0, st, m, V[r], O, dn.ldof_lnode,
                       dn.ldof_owner, ownf);
  if (hipGetLastError() != hipSuccess) return fail(GDX_ERR_HIP);
  if ((rc = scan_exclusive(ownf, oscan, ldofs, st, SW))) return fail(rc);
  long long owns = 0;
  at.clear();
  for (int r = 0; r <= nparts; ++r) at.push_back(r < nparts ? V[r].ldof0 : ldofs);
  if ((rc = read_at(oscan, at, val, st))) return fail(rc);
  for (int r = 0; r < nparts; ++r) {
    Part &P = p->P[r];
    P.n_own = val[r + 1] - val[r];
    P.own_base = owns;
    V[r].n_own = P.n_own;
    V[r].own0 = owns;
    V[r].gid0 = owns;  // owned dofs of all parts in part order: the global ids
    owns += P.n_own;
  }
  p->ngdofs = owns;
  if (!alloc(&dn.own_ldofs, owns)) return fail(GDX_ERR_ALLOC);
  for (int r = 0; r < nparts; ++r)
    hipLaunchKernelGGL(k_own_ids, dim3(grid_of(V[r].nldof)), dim3(kBlock), 0, st, V[r], ownf, oscan, dn.own_ldofs,
                       dn.ldof_gid, dn.ldof_col);
  PartN *d_parts = X.get<PartN>(nparts);
  if (!d_parts || !put(d_parts, V)) return fail(GDX_ERR_ALLOC);
  for (int r = 0; r < nparts; ++r)
    hipLaunchKernelGGL(k_ghost_ids, dim3(grid_of(V[r].nldof)), dim3(kBlock), 0, st, m, d_parts, r, fa ? 1 : 0,
                       dn.ldof_owner, dn.ldof_lnode, dn.lnode_dof, dn.ldof_gid, grf, d_err);
  if (hipGetLastError() != hipSuccess) return fail(GDX_ERR_HIP);
  if ((rc = scan_exclusive(grf, grscan, ldofs, st, SW))) return fail(rc);
  // 4. row PRange: owned dofs, then the ghost rows, each in local id order
  long long rows = 0;
  if ((rc = read_at(grscan, at, val, st))) return fail(rc);
  for (int r = 0; r < nparts; ++r) {
    Part &P = p->P[r];
    P.n_ghost_rows = val[r + 1] - val[r];
    P.nrows = P.n_own + P.n_ghost_rows;
    P.row_base = rows;
    V[r].row0 = rows;
    rows += P.nrows;
  }
  if (!alloc(&dn.row_ldof, rows) || !alloc(&dn.row_gid, rows) || !alloc(&dn.row_owner, rows))
    return fail(GDX_ERR_ALLOC);
  for (int r = 0; r < nparts; ++r)
    hipLaunchKernelGGL(k_rows, dim3(grid_of(V[r].nldof)), dim3(kBlock), 0, st, V[r], ownf, oscan, grf, grscan,
                       dn.ldof_gid, dn.ldof_owner, dn.row_ldof, dn.row_gid, dn.row_owner);
  // 5. col PRange: owned dofs, then the ghost columns in first-encounter order
  // over the local COO (src/Algebra.jl:1136-1257): the own cells, then the
  // cells whose triplets each sender ships, senders in ascending rank order
  struct Pass {
    int r, s, want;
    long long lo[3], n[3], koff;
  };
  std::vector<Pass> passes;
  long long keys = 0;
  for (int r = 0; r < nparts; ++r) {
    const Part &P = p->P[r];
    V[r].key0 = keys;
    long long nk = 0;
    auto add = [&](int s, int want, const long long lo[3], const long long hi[3]) {
      Pass q{r, s, want, {lo[0], lo[1], lo[2]}, {hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]}, nk};
      const long long cells = q.n[0] * q.n[1] * q.n[2];
      if (cells <= 0) return;
      passes.push_back(q);
      nk += cells * nc * ncd;
    };
    if (fa) {
      passes.push_back(Pass{r, r, r, {P.loc_lo[0], P.loc_lo[1], P.loc_lo[2]},
                            {P.loc_hi[0] - P.loc_lo[0], P.loc_hi[1] - P.loc_lo[1], P.loc_hi[2] - P.loc_lo[2]}, 0});
      nk = P.n_ldofs;
    } else {
      const long long lo[3] = {P.own_lo[0], P.own_lo[1], P.own_lo[2]}, hi[3] = {P.own_hi[0], P.own_hi[1], P.own_hi[2]};
      add(r, -1, lo, hi);
      for (int s = 0; s < nparts; ++s) {
        if (s == r) continue;
        const Part &S = p->P[s];
        long long a[3], b[3];
        bool empty = false;
        for (int k = 0; k < 3; ++k) {
          // sender cells whose closure meets this part's owned node box
          a[k] = std::max<long long>(S.own_lo[k], (P.onb_lo[k] + K - 1) / K - 1);
          b[k] = std::min<long long>(S.own_hi[k], (P.onb_lo[k] + P.onb_n[k] - 1) / K + 1);
          if (k >= D) a[k] = 0, b[k] = 1;
          if (a[k] >= b[k]) empty = true;
        }
        if (!empty) add(s, r, a, b);
      }
    }
    V[r].nkey = nk;
    keys += nk;
  }
  unsigned long long *minkey = X.get<unsigned long long>(ldofs);
  long long *kf = X.get<long long>(keys), *kscan = X.get<long long>(keys + 1);
  if (!minkey || !kf || !kscan || !put(d_parts, V)) return fail(GDX_ERR_ALLOC);
  hipLaunchKernelGGL(k_fill_u64, dim3(grid_of(ldofs)), dim3(kBlock), 0, st, minkey, ldofs, ~0ull);
  hipLaunchKernelGGL(k_fill_i64, dim3(grid_of(keys)), dim3(kBlock), 0, st, reinterpret_cast<int64_t *>(kf), keys,
                     (int64_t)0);
  for (const Pass &q : passes)
    hipLaunchKernelGGL(k_visit, dim3(grid_of(q.n[0] * q.n[1] * q.n[2])), dim3(kBlock), 0, st, m, T, d_parts, q.r, q.s,
                       q.want, fa ? 1 : 0, q.lo[0], q.lo[1], q.lo[2], q.n[0], q.n[1], q.n[2], q.koff, dn.lnode_dof,
                       dn.ldof_owner, minkey, d_err);
  for (int r = 0; r < nparts; ++r)
    hipLaunchKernelGGL(k_key_flags, dim3(grid_of(V[r].nldof)), dim3(kBlock), 0, st, V[r], minkey, kf);
  if (hipGetLastError() != hipSuccess) return fail(GDX_ERR_HIP);
  if ((rc = scan_exclusive(kf, kscan, keys, st, SW))) return fail(rc);
  long long cols = 0;
  at.clear();
  for (int r = 0; r <= nparts; ++r) at.push_back(r < nparts ? V[r].key0 : keys);
  if ((rc = read_at(kscan, at, val, st))) return fail(rc);
  for (int r = 0; r < nparts; ++r) {
    Part &P = p->P[r];
    P.n_ghost_cols = val[r + 1] - val[r];
    P.ncols = P.n_own + P.n_ghost_cols;
    P.col_base = cols;
    V[r].col0 = cols;
    cols += P.ncols;
  }
  if (!alloc(&dn.col_gid, cols) || !alloc(&dn.col_owner, cols)) return fail(GDX_ERR_ALLOC);
  for (int r = 0; r < nparts; ++r)
    hipLaunchKernelGGL(k_cols, dim3(grid_of(V[r].nldof)), dim3(kBlock), 0, st, V[r], ownf, oscan, minkey, kscan,
                       dn.ldof_gid, dn.ldof_owner, dn.ldof_col, dn.col_gid, dn.col_owner);
  int err = 0;
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess ||
      hipMemcpy(&err, d_err, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)
    return fail(GDX_ERR_HIP);
  if (err) return fail(kUseHost);
  // host copies of the numbering are made on first host use
  for (Part &P : p->P) {
    P.lnode_dof.clear();
    P.ldof_lnode.clear();
    P.ldof_owner.clear();
    P.ldof_gid.clear();
    P.own_ldofs.clear();
    P.ldof_col.clear();
    P.row_ldof.clear();
    P.row_gid.clear();
    P.row_owner.clear();
    P.col_gid.clear();
    P.col_owner.clear();
  }
  p->dev_numbering = true;
  p->host_fe_stale = true;
  return GDX_OK;
}

int host_fe(gdx_problem *p) {
  if (!p->host_fe_stale) return GDX_OK;
  if (hipSetDevice(p->ctx->device) != hipSuccess || hipStreamSynchronize(p->ctx->stream) != hipSuccess)
    return GDX_ERR_HIP;
  const DevNumbering &d = p->dn;
  auto get = [](auto &v, const auto *src, long long n) {
    v.resize(n);
    return n == 0 || hipMemcpy(v.data(), src, n * sizeof(v[0]), hipMemcpyDeviceToHost) == hipSuccess;
  };
  for (Part &P : p->P) {
    const long long nl = P.lnn[0] * P.lnn[1] * P.lnn[2];
    if (!get(P.lnode_dof, d.lnode_dof + P.lnode_base, nl) || !get(P.ldof_lnode, d.ldof_lnode + P.ldof_base, P.n_ldofs) ||
        !get(P.ldof_owner, d.ldof_owner + P.ldof_base, P.n_ldofs) ||
        !get(P.ldof_gid, d.ldof_gid + P.ldof_base, P.n_ldofs) || !get(P.own_ldofs, d.own_ldofs + P.own_base, P.n_own) ||
        !get(P.ldof_col, d.ldof_col + P.ldof_base, P.n_ldofs) || !get(P.row_ldof, d.row_ldof + P.row_base, P.nrows) ||
        !get(P.row_gid, d.row_gid + P.row_base, P.nrows) || !get(P.row_owner, d.row_owner + P.row_base, P.nrows) ||
        !get(P.col_gid, d.col_gid + P.col_base, P.ncols) || !get(P.col_owner, d.col_owner + P.col_base, P.ncols))
      return GDX_ERR_HIP;
  }
  p->host_fe_stale = false;
  return GDX_OK;
}

int gid_node_device(gdx_problem *p, long long *gid_node) {
  const MeshG m = mesh_of(p);
  for (int r = 0; r < p->nparts; ++r) {
    const PartN V = part_view(p, r);
    if (V.nldof > 0)
      hipLaunchKernelGGL(k_gid_node, dim3(grid_of(V.nldof)), dim3(kBlock), 0, p->ctx->stream, m, V, p->dn.ldof_lnode,
                         p->dn.ldof_gid, (long long)p->ngdofs, gid_node);
  }
  return hipGetLastError() == hipSuccess && hipStreamSynchronize(p->ctx->stream) == hipSuccess ? GDX_OK : GDX_ERR_HIP;
}

int structure_device(gdx_problem *p) {
  hipStream_t st = p->ctx->stream;
  const int nparts = p->nparts, nc = p->ncomp, D = p->D, K = p->K;
  MeshG m{};
  for (int k = 0; k < 3; ++k) {
    m.n[k] = k < D ? p->desc.cells[k] : 1;
    m.nn[k] = k < D ? K * m.n[k] + 1 : 1;
  }
  m.D = D;
  m.K = K;
  m.nc = nc;
  m.dir_kind = p->desc.dir_kind;
  const int W = 2 * K + 1, S = D == 3 ? W * W * W : W * W, SN = S * nc, KS = (SN + 3) & ~3;
  int LP = 32;  // lanes per row group: one per stencil node
  while (LP < S) LP *= 2;
  if (LP > 128 || nc > kMaxComp) return kUseHost;
  std::vector<PartG> G(nparts);
  long long rows = 0, ogrps = 0, rinfo = 0, lnodes = 0, ldofs = 0, owns = 0;
  for (int r = 0; r < nparts; ++r) {
    const Part &P = p->P[r];
    PartG &g = G[r];
    for (int k = 0; k < 3; ++k) {
      g.loc_lo[k] = P.loc_lo[k];
      g.lnn[k] = P.lnn[k];
      g.own_lo[k] = P.own_lo[k];
      g.own_hi[k] = P.own_hi[k];
      g.onb_lo[k] = P.onb_lo[k];
      g.onb_n[k] = P.onb_n[k];
    }
    if (P.nrows % nc || P.n_own % nc) return kUseHost;
    g.n_own = P.n_own;
    g.nrows = P.nrows;
    g.row0 = rows;
    g.ogrp0 = ogrps;
    g.rinfo0 = rinfo;
    g.lnode0 = lnodes;
    g.ldof0 = ldofs;
    g.own0 = owns;
    rows += P.nrows;
    ogrps += P.n_own / nc;
    rinfo += P.onb_n[0] * P.onb_n[1] * P.onb_n[2];
    lnodes += P.lnn[0] * P.lnn[1] * P.lnn[2];
    ldofs += P.n_ldofs;
    owns += P.n_own;
  }
  Scratch X(st);
  // the numbering and PRanges: made on the device already, or uploaded
  const bool dn = p->dev_numbering;
  int64_t *d_lnode_dof = dn ? p->dn.lnode_dof : X.get<int64_t>(lnodes),
          *d_ldof_col = dn ? p->dn.ldof_col : X.get<int64_t>(ldofs),
          *d_ldof_lnode = dn ? p->dn.ldof_lnode : X.get<int64_t>(ldofs),
          *d_row_ldof = dn ? p->dn.row_ldof : X.get<int64_t>(rows), *d_own = dn ? p->dn.own_ldofs : X.get<int64_t>(owns),
          *d_ldof_orow = X.get<int64_t>(ldofs), *d_row_lnode = X.get<int64_t>(rows / nc);
  long long *d_rowlen = X.get<long long>(rows);
  int *d_err = X.get<int>(1);
  if (!d_lnode_dof || !d_ldof_col || !d_ldof_lnode || !d_ldof_orow || !d_row_ldof || !d_own || !d_row_lnode ||
      !d_rowlen || !d_err)
    return GDX_ERR_ALLOC;
  for (int r = 0; r < nparts && !dn; ++r) {
    const Part &P = p->P[r];
    if (!put(d_lnode_dof + G[r].lnode0, P.lnode_dof) || !put(d_ldof_col + G[r].ldof0, P.ldof_col) ||
        !put(d_ldof_lnode + G[r].ldof0, P.ldof_lnode) || !put(d_row_ldof + G[r].row0, P.row_ldof) ||
        !put(d_own + G[r].own0, P.own_ldofs))
      return GDX_ERR_HIP;
  }
  symbolic_mark("uploads");
  if (hipMemsetAsync(d_err, 0, sizeof(int), st) != hipSuccess) return GDX_ERR_HIP;
  hipLaunchKernelGGL(k_fill_i64, dim3(grid_of(ldofs)), dim3(kBlock), 0, st, d_ldof_orow, ldofs, (int64_t)-1);
  for (int r = 0; r < nparts; ++r) {
    const PartG &g = G[r];
    hipLaunchKernelGGL(k_ldof_orow, dim3(grid_of(g.n_own)), dim3(kBlock), 0, st, g, d_own + g.own0,
                       d_ldof_orow + g.ldof0);
    hipLaunchKernelGGL(k_row_lnode, dim3(grid_of(g.nrows / nc)), dim3(kBlock), 0, st, g, nc, d_row_ldof + g.row0,
                       d_ldof_lnode + g.ldof0, d_row_lnode + g.row0 / nc, d_err);
    hipLaunchKernelGGL(k_rowlen, dim3(grid_of(g.nrows / nc)), dim3(kBlock), 0, st, m, g, d_row_lnode + g.row0 / nc,
                       d_lnode_dof + g.lnode0, d_ldof_col + g.ldof0, d_rowlen, d_err);
  }
  if (hipGetLastError() != hipSuccess) return GDX_ERR_HIP;
  // row offsets: the parts' CSRs back to back, so part r's local rowptr is the
  // global one minus its first value
  long long *rowptr = nullptr;
  if (dev_malloc(&rowptr, (rows + 1) * sizeof(long long)) != hipSuccess) return GDX_ERR_ALLOC;
  ScanWs SW(st);
  int rc = scan_exclusive(d_rowlen, rowptr, rows, st, SW);
  if (rc) {
    (void)hipFree(rowptr);
    return rc;
  }
  symbolic_mark("row lengths+scan");
  // the part boundaries of the offsets: first row, first ghost row, end
  std::vector<long long> bound, at;
  for (int r = 0; r < nparts; ++r)
    for (long long a : {G[r].row0, G[r].row0 + G[r].n_own, G[r].row0 + G[r].nrows}) at.push_back(a);
  at.push_back(rows);
  rc = read_at(rowptr, at, bound, st);
  const long long nnz = rc ? 0 : bound[3 * nparts];
  int *colidx = nullptr;
  if (!rc && dev_malloc(&colidx, std::max<long long>(nnz, 1) * sizeof(int)) != hipSuccess) rc = GDX_ERR_ALLOC;
  uint16_t *keys = rc ? nullptr : X.get<uint16_t>(ogrps * KS);
  unsigned long long *hash = rc ? nullptr : X.get<unsigned long long>(ogrps);
  unsigned long long *tkey = rc ? nullptr : X.get<unsigned long long>(kTableSize);
  unsigned long long *tmin = rc ? nullptr : X.get<unsigned long long>(kTableSize);
  unsigned *slot = rc ? nullptr : X.get<unsigned>(ogrps);
  if (!rc && (!keys || !hash || !tkey || !tmin || !slot)) rc = GDX_ERR_ALLOC;
  auto fail = [&](int code) {
    (void)hipStreamSynchronize(st);
    if (colidx) (void)hipFree(colidx);
    (void)hipFree(rowptr);
    return code;
  };
  if (rc) return fail(rc);
  for (int r = 0; r < nparts; ++r) {
    const PartG &g = G[r];
    const int64_t *rl = d_row_lnode + g.row0 / nc, *ldn = d_lnode_dof + g.lnode0, *ldc = d_ldof_col + g.ldof0;
    switch (LP) {
      case 32: launch_fill<32>(m, g, st, rl, ldn, ldc, rowptr, colidx, keys, KS, hash, d_err); break;
      case 64: launch_fill<64>(m, g, st, rl, ldn, ldc, rowptr, colidx, keys, KS, hash, d_err); break;
      default: launch_fill<128>(m, g, st, rl, ldn, ldc, rowptr, colidx, keys, KS, hash, d_err); break;
    }
  }
  // distinct keys: first occurrences through the hash table
  hipLaunchKernelGGL(k_fill_u64, dim3(grid_of(kTableSize)), dim3(kBlock), 0, st, tkey, (long long)kTableSize, 0ull);
  hipLaunchKernelGGL(k_fill_u64, dim3(grid_of(kTableSize)), dim3(kBlock), 0, st, tmin, (long long)kTableSize, ~0ull);
  if (ogrps > 0)
    hipLaunchKernelGGL(k_insert, dim3(grid_of(ogrps)), dim3(kBlock), 0, st, hash, ogrps, tkey, tmin, slot, d_err);
  unsigned *d_cnt = X.get<unsigned>(1);
  unsigned long long *l_min = X.get<unsigned long long>(kTableSize);
  unsigned *l_slot = X.get<unsigned>(kTableSize);
  if (!d_cnt || !l_min || !l_slot) return fail(GDX_ERR_ALLOC);
  if (hipMemsetAsync(d_cnt, 0, sizeof(unsigned), st) != hipSuccess) return fail(GDX_ERR_HIP);
  hipLaunchKernelGGL(k_collect, dim3(grid_of(kTableSize)), dim3(kBlock), 0, st, tkey, tmin, d_cnt, l_min, l_slot);
  int err = 0;
  unsigned npat = 0;
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess ||
      hipMemcpy(&err, d_err, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(&npat, d_cnt, sizeof(unsigned), hipMemcpyDeviceToHost) != hipSuccess)
    return fail(GDX_ERR_HIP);
  if (err) return fail(kUseHost);
  symbolic_mark("CSR + hash");
  std::vector<unsigned long long> hmin(npat);
  std::vector<unsigned> hslot(npat);
  if (npat && (hipMemcpy(hmin.data(), l_min, npat * sizeof(unsigned long long), hipMemcpyDeviceToHost) != hipSuccess ||
               hipMemcpy(hslot.data(), l_slot, npat * sizeof(unsigned), hipMemcpyDeviceToHost) != hipSuccess))
    return fail(GDX_ERR_HIP);
  // pattern ids in order of first occurrence
  std::vector<unsigned> order(npat);
  for (unsigned i = 0; i < npat; ++i) order[i] = i;
  std::sort(order.begin(), order.end(), [&](unsigned a, unsigned b) { return hmin[a] < hmin[b]; });
  std::vector<int> tpid(kTableSize, -1);
  std::vector<long long> rep(npat);
  for (unsigned i = 0; i < npat; ++i) {
    tpid[hslot[order[i]]] = (int)i;
    rep[i] = (long long)hmin[order[i]];
  }
  int *d_tpid = X.get<int>(kTableSize), *d_pid = X.get<int>(ogrps);
  long long *d_rep = X.get<long long>(npat);
  uint16_t *d_repkeys = X.get<uint16_t>((long long)npat * KS);
  if (!d_tpid || !d_pid || !d_rep || !d_repkeys) return fail(GDX_ERR_ALLOC);
  if (!put(d_tpid, tpid) || !put(d_rep, rep)) return fail(GDX_ERR_HIP);
  if (ogrps > 0)
    hipLaunchKernelGGL(k_pattern_ids, dim3(grid_of(ogrps)), dim3(kBlock), 0, st, keys, KS, ogrps, slot, tmin, d_tpid,
                       d_pid, d_err);
  // GDX_FAULT_HASH=1: report a hash collision, so that the host stage's
  // take-over after a collision can be tested
  {
    const char *fh = std::getenv("GDX_FAULT_HASH");
    if (fh && std::atoi(fh) > 0 && hipMemsetAsync(d_err, 4, 1, st) != hipSuccess) return fail(GDX_ERR_HIP);
  }
  if (npat > 0)
    hipLaunchKernelGGL(k_gather_keys, dim3(grid_of((long long)npat * KS)), dim3(kBlock), 0, st, keys, KS, d_rep,
                       (int)npat, d_repkeys);
  std::vector<uint16_t> repkeys((size_t)npat * KS);
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess ||
      hipMemcpy(&err, d_err, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess ||
      (npat && hipMemcpy(repkeys.data(), d_repkeys, repkeys.size() * sizeof(uint16_t), hipMemcpyDeviceToHost) !=
                   hipSuccess))
    return fail(GDX_ERR_HIP);
  if (err) return fail(kUseHost);
  // pattern tables from the representatives' keys, as the host stage makes them
  p->pat_fwd.clear();
  p->pat_inv.clear();
  p->pat_invr.clear();
  p->pat_len.assign(npat, 0);
  p->pat_dir.assign(npat, 0);
  p->pat16_inv.clear();
  std::vector<long long> ogrp0(nparts);
  for (int r = 0; r < nparts; ++r) ogrp0[r] = G[r].ogrp0;
  if (nc == 1) {
    // the representatives' nodes (their parity fixes the reduced numbering)
    std::vector<long long> rgrp(npat), rnode(npat);
    std::vector<int> rpart(npat);
    for (unsigned i = 0; i < npat; ++i) {
      const int r = (int)(std::upper_bound(ogrp0.begin(), ogrp0.end(), rep[i]) - ogrp0.begin()) - 1;
      rpart[i] = r;
      rgrp[i] = G[r].row0 / nc + (rep[i] - ogrp0[r]);
    }
    long long *d_rg = X.get<long long>(npat), *d_rn = X.get<long long>(npat);
    if (!d_rg || !d_rn) return fail(GDX_ERR_ALLOC);
    if (!put(d_rg, rgrp)) return fail(GDX_ERR_HIP);
    if (npat) hipLaunchKernelGGL(k_gather_i64, dim3(grid_of(npat)), dim3(kBlock), 0, st, d_row_lnode, d_rg, (long long)npat, d_rn);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess ||
        (npat && hipMemcpy(rnode.data(), d_rn, npat * sizeof(long long), hipMemcpyDeviceToHost) != hipSuccess))
      return fail(GDX_ERR_HIP);
    const int SP = pattern_stride(S);
    p->pat_fwd.resize((size_t)npat * SP);
    p->pat_inv.resize((size_t)npat * SP);
    p->pat_invr.resize((size_t)npat * SP);
    std::vector<uint8_t> key8(S);
    for (unsigned i = 0; i < npat; ++i) {
      const uint16_t *k16 = &repkeys[(size_t)i * KS];
      int len = 0;
      for (int s = 0; s < S; ++s) {
        key8[s] = k16[s] == kDir16 ? kSlotDirichlet : k16[s] == kAbs16 ? kSlotAbsent : (uint8_t)k16[s];
        len += k16[s] < kDir16;
      }
      const Part &P = p->P[rpart[i]];
      const int64_t l = rnode[i];
      const int64_t g[3] = {K * P.loc_lo[0] + l % P.lnn[0], K * P.loc_lo[1] + (l / P.lnn[0]) % P.lnn[1],
                            K * P.loc_lo[2] + l / (P.lnn[0] * P.lnn[1])};
      p->pat_dir[i] = scalar_pattern_tables(D, K, key8.data(), S, SP, g, len, &p->pat_fwd[(size_t)i * SP],
                                            &p->pat_inv[(size_t)i * SP], &p->pat_invr[(size_t)i * SP])
                          ? 1
                          : 0;
      p->pat_len[i] = len;
    }
  } else {
    const int SP = (SN + 8) & ~7;
    p->pat16_stride = SP;
    p->pat16_inv.resize((size_t)npat * SP);
    for (unsigned i = 0; i < npat; ++i) {
      const uint16_t *k16 = &repkeys[(size_t)i * KS];
      int len = 0;
      for (int s = 0; s < SN; ++s) len += k16[s] < kDir16;
      p->pat_dir[i] = vector_pattern_inv(k16, SN, SP, &p->pat16_inv[(size_t)i * SP]) ? 1 : 0;
      p->pat_len[i] = len;
    }
  }
  symbolic_mark("pattern ids");
  // row table
  int *d_dir = X.get<int>(npat);
  RowInfo *rinfo_d = nullptr;
  if (!d_dir || !put(d_dir, p->pat_dir)) return fail(GDX_ERR_ALLOC);
  if (dev_malloc(&rinfo_d, std::max<long long>(rinfo, 1) * sizeof(RowInfo)) != hipSuccess) return fail(GDX_ERR_ALLOC);
  for (int r = 0; r < nparts; ++r) {
    const PartG &g = G[r];
    hipLaunchKernelGGL(k_rowinfo, dim3(grid_of(g.onb_n[0] * g.onb_n[1] * g.onb_n[2])), dim3(kBlock), 0, st, m, g,
                       d_lnode_dof + g.lnode0, d_ldof_orow + g.ldof0, rowptr, d_pid, d_dir, rinfo_d, d_err);
  }
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess ||
      hipMemcpy(&err, d_err, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) {
    (void)hipFree(rinfo_d);
    return fail(GDX_ERR_HIP);
  }
  if (err) {
    (void)hipFree(rinfo_d);
    return fail(kUseHost);
  }
  symbolic_mark("row table");
  // sizes and offsets of the parts; host copies of the CSRs and the row table
  // are made on demand
  int64_t owned_nnz = 0;
  for (int r = 0; r < nparts; ++r) {
    Part &P = p->P[r];
    P.val_base = bound[3 * r];
    P.row_base = G[r].row0;
    P.nnz = bound[3 * r + 2] - bound[3 * r];
    P.rowinfo_base = G[r].rinfo0;
    P.rowptr.clear();
    P.colidx.clear();
    owned_nnz += bound[3 * r + 1] - bound[3 * r];
  }
  p->rowinfo.clear();
  p->npat = (int)npat;
  p->nnz_total = nnz;
  p->rows_total = rows;
  p->owned_nnz = owned_nnz;
  Device &d = p->dev;
  if (d.csr_rowptr) (void)hipFree(d.csr_rowptr);
  if (d.csr_colidx) (void)hipFree(d.csr_colidx);
  if (d.rowinfo) (void)hipFree(d.rowinfo);
  d.csr_rowptr = rowptr;
  d.csr_colidx = colidx;
  d.rowinfo = rinfo_d;
  p->dev_structure = true;
  p->host_csr_stale = p->host_rowinfo_stale = true;
  return GDX_OK;
}

int build_tiles_device(gdx_problem *p, const int R[3]) {
  hipStream_t st = p->ctx->stream;
  for (int k = 0; k < 3; ++k) p->tileR[k] = R[k];
  // candidate tiles in the host stage's order: parts, then tz, ty, tx
  std::vector<Tile> cand;
  for (int r = 0; r < p->nparts; ++r) {
    const Part &P = p->P[r];
    int64_t nt[3];
    for (int k = 0; k < 3; ++k) nt[k] = (P.onb_n[k] + R[k] - 1) / R[k];
    for (int64_t tz = 0; tz < nt[2]; ++tz)
      for (int64_t ty = 0; ty < nt[1]; ++ty)
        for (int64_t tx = 0; tx < nt[0]; ++tx) {
          Tile t{};
          t.part = r;
          const int64_t tt[3] = {tx, ty, tz};
          for (int k = 0; k < 3; ++k) {
            t.org[k] = (int)(P.onb_lo[k] + tt[k] * R[k]);
            t.ext[k] = (int)std::min<int64_t>(R[k], P.onb_n[k] - tt[k] * R[k]);
            t.onb_lo[k] = (int)P.onb_lo[k];
            t.onb_n[k] = (int)P.onb_n[k];
          }
          t.rowinfo_base = P.rowinfo_base;
          cand.push_back(t);
        }
  }
  const int nc_ = (int)cand.size(), cap = R[0] * R[1] * R[2];
  Scratch X(st);
  Tile *d_cand = X.get<Tile>(nc_);
  int *lists = X.get<int>((long long)nc_ * cap), *d_npat = X.get<int>(nc_), *d_any = X.get<int>(nc_),
      *d_err = X.get<int>(1);
  long long *d_off = X.get<long long>(nc_);
  if (!d_cand || !lists || !d_npat || !d_any || !d_err || !d_off) return GDX_ERR_ALLOC;
  if (!put(d_cand, cand) || hipMemsetAsync(d_err, 0, sizeof(int), st) != hipSuccess) return GDX_ERR_HIP;
  // the workgroup form when three ints per pattern fit in 64 KB of LDS
  const size_t lds = (size_t)3 * p->npat * sizeof(int);
  if (nc_ > 0 && lds <= 65536)
    hipLaunchKernelGGL(k_tile_patterns_wg, dim3(nc_), dim3(kTileThreads), lds, st, d_cand, nc_, cap, p->npat,
                       p->dev.rowinfo, lists, d_npat, d_any, d_err);
  else if (nc_ > 0)
    hipLaunchKernelGGL(k_tile_patterns, dim3((nc_ + 63) / 64), dim3(64), 0, st, d_cand, nc_, cap, p->dev.rowinfo,
                       lists, d_npat, d_any, d_err);
  std::vector<int> npat(nc_), any(nc_);
  int err = 0;
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess ||
      hipMemcpy(&err, d_err, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess ||
      (nc_ && (hipMemcpy(npat.data(), d_npat, nc_ * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess ||
               hipMemcpy(any.data(), d_any, nc_ * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)))
    return GDX_ERR_HIP;
  if (err) {
    set_error("internal: too many patterns in one tile");
    return GDX_ERR_ARG;
  }
  std::vector<long long> off(nc_, -1);
  p->tiles.clear();
  long long total = 0;
  for (int t = 0; t < nc_; ++t) {
    if (!any[t]) continue;
    Tile T = cand[t];
    T.npat = npat[t];
    T.pat_off = total;
    off[t] = total;
    total += npat[t];
    p->tiles.push_back(T);
  }
  int *tp = nullptr;
  if (dev_malloc(&tp, std::max<long long>(total, 1) * sizeof(int)) != hipSuccess) return GDX_ERR_ALLOC;
  p->tile_pats.assign(total, 0);
  if (!put(d_off, off)) {
    (void)hipFree(tp);
    return GDX_ERR_HIP;
  }
  if (nc_ > 0) hipLaunchKernelGGL(k_tile_write, dim3(nc_), dim3(64), 0, st, lists, cap, d_npat, d_off, nc_, tp);
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess ||
      (total && hipMemcpy(p->tile_pats.data(), tp, total * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)) {
    (void)hipFree(tp);
    return GDX_ERR_HIP;
  }
  if (p->dev.tile_pats) (void)hipFree(p->dev.tile_pats);
  p->dev.tile_pats = tp;
  p->host_rowinfo_stale = true;
  return GDX_OK;
}

int host_csr(gdx_problem *p) {
  if (!p->host_csr_stale) return GDX_OK;
  if (hipSetDevice(p->ctx->device) != hipSuccess || hipStreamSynchronize(p->ctx->stream) != hipSuccess)
    return GDX_ERR_HIP;
  for (Part &P : p->P) {
    P.rowptr.resize(P.nrows + 1);
    P.colidx.resize(P.nnz);
    if (hipMemcpy(P.rowptr.data(), p->dev.csr_rowptr + P.row_base, (P.nrows + 1) * sizeof(long long),
                  hipMemcpyDeviceToHost) != hipSuccess ||
        (P.nnz && hipMemcpy(P.colidx.data(), p->dev.csr_colidx + P.val_base, P.nnz * sizeof(int),
                            hipMemcpyDeviceToHost) != hipSuccess))
      return GDX_ERR_HIP;
    for (int64_t &v : P.rowptr) v -= P.val_base;
  }
  p->host_csr_stale = false;
  return GDX_OK;
}

int host_rowinfo(gdx_problem *p) {
  if (!p->host_rowinfo_stale) return GDX_OK;
  int64_t n = 0;
  for (const Part &P : p->P) n += P.onb_n[0] * P.onb_n[1] * P.onb_n[2];
  p->rowinfo.resize(n);
  if (hipSetDevice(p->ctx->device) != hipSuccess || hipStreamSynchronize(p->ctx->stream) != hipSuccess ||
      (n && hipMemcpy(p->rowinfo.data(), p->dev.rowinfo, n * sizeof(RowInfo), hipMemcpyDeviceToHost) != hipSuccess))
    return GDX_ERR_HIP;
  p->host_rowinfo_stale = false;
  return GDX_OK;
}

}  // namespace gdx
