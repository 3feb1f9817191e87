// fv_regrid.hip -- device-resident mesh adaptation (meshadaptation.c:3-44).
//
// One fv_dev_regrid call runs, without leaving the device:
//  1. the indicator (interface layers and per-leaf targets, fv_dev_indicator);
//  2. criterionsmth (criterionsmth.c:3-58): LEVEL+1 passes of sibling consensus
//     and 2:1 smoothing.  The reference sweeps the leaves in order and updates
//     in place.  The sibling step only reads the leaf's own family, so each
//     family is decided at once (all eight want to coarsen, or none coarsens).
//     The smoothing step new_i = max(old_i, max_j cur_j - 1) reads earlier
//     leaves' new values and later leaves' old ones.  Its result is the unique
//     solution of y_i = max(c_i, max_{j<i} y_j - 1), where c_i holds the later
//     neighbours' contribution.  Jacobi sweeps from y = c rise monotonically to
//     that solution in at most a few sweeps (each hop lowers a value by one);
//     the result equals the sequential sweep exactly;
//  3. merge and split with the state remap (merge.c:20-55: mean of the eight
//     kids, then relaxation; split.c:425-434: kids copy the parent) at offsets
//     from a prefix sum, which keeps the DFS leaf order;
//  4. connectivity from 64-bit DFS keys by binary search (same level, one
//     level coarser, or four finer leaves as elconnfin.c:14-22 orders them),
//     and the face kinds, 1:4 rows and neighbour table of fv_mesh_pack.
// The result equals fv_criterionsmth + fv_mesh_adapt + fv_mesh_pack on the host
// (tests/test_regrid.py) at a fraction of the time.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <stdint.h>
#include <string.h>
#include <vector>
#include "fv_internal.h"

namespace {

using fvk::Params;

constexpr int BLK = 256;
__constant__ int D_FAX[6] = {1, 0, 1, 0, 2, 2};
__constant__ int D_FSG[6] = {-1, 1, 1, -1, -1, 1};
__constant__ int D_FCOPP[6] = {2, 3, 0, 1, 5, 4};            // operatetp.c:268-279
__constant__ int D_BCSIDE[6] = {2, 1, 3, 0, 4, 5};           // face -> -x,+x,-y,+y,-z,+z side
__constant__ int D_FCREF[6][4] = {{0, 4, 5, 1}, {3, 1, 5, 7}, {2, 3, 7, 6}, {2, 6, 4, 0}, {2, 0, 1, 3}, {6, 7, 5, 4}};

struct Forest {
  int nroot[3];
  int bcside[6];
};

__device__ __forceinline__ int kid_of(const int *ci, int i) {
  return (ci[3 * i] & 1) + 2 * (ci[3 * i + 1] & 1) + 4 * (ci[3 * i + 2] & 1);
}

// 8 consecutive leaves from `first` are the complete family of leaf i at level l
__device__ bool family_at(int first, int i, int n, const int *__restrict__ root, const int *__restrict__ level,
                          const int *__restrict__ ci) {
  if (first < 0 || first + 7 >= n) return false;
  const int l = level[i], r = root[i];
  const int px = ci[3 * i] >> 1, py = ci[3 * i + 1] >> 1, pz = ci[3 * i + 2] >> 1;
  for (int k = 0; k < 8; k++) {
    const int s = first + k;
    if (level[s] != l || root[s] != r || kid_of(ci, s) != k || (ci[3 * s] >> 1) != px ||
        (ci[3 * s + 1] >> 1) != py || (ci[3 * s + 2] >> 1) != pz)
      return false;
  }
  return true;
}

// ---- criterionsmth ----

// sibling consensus and the one-level cap (criterionsmth.c:13-37)
__global__ __launch_bounds__(BLK) void k_smooth_family(int n, const int *__restrict__ root,
                                                       const int *__restrict__ level, const int *__restrict__ ci,
                                                       const int *__restrict__ sin, int *__restrict__ sout) {
  const int i = blockIdx.x * BLK + threadIdx.x;
  if (i >= n) return;
  const int l = level[i];
  int v = sin[i];
  if (l > 1) {
    const int first = i - kid_of(ci, i);
    bool all8 = false;
    if (family_at(first, i, n, root, level, ci)) {
      all8 = true;
      for (int k = 0; k < 8; k++) all8 = all8 && (sin[first + k] < l);
    }
    if (all8) v = l - 1;
    else if (l > v) v = l;
  }
  if (l < v) v = l + 1;
  sout[i] = v;
}

template <class Fn>
__device__ __forceinline__ void for_face_neighbours(int i, int n, const int *__restrict__ nbr,
                                                    const uint32_t *__restrict__ info,
                                                    const int *__restrict__ fine4, Fn fn) {
  const uint32_t inf = info[i];
  for (int f = 0; f < 6; f++) {
    const int kind = fvk::face_kind(inf, f);
    if (kind == FV_FK_BOUNDARY) continue;
    const int nb = nbr[(size_t)f * n + i];
    if (kind == FV_FK_FINE) {
      for (int q = 0; q < 4; q++) fn(fine4[4 * nb + q]);
    } else {
      fn(nb);
    }
  }
}

// c_i = max(x_i, 0, max over later neighbours x_j - 1)  (criterionsmth.c:39-55)
__global__ __launch_bounds__(BLK) void k_smooth_later(int n, const int *__restrict__ nbr,
                                                      const uint32_t *__restrict__ info,
                                                      const int *__restrict__ fine4, const int *__restrict__ x,
                                                      int *__restrict__ c) {
  const int i = blockIdx.x * BLK + threadIdx.x;
  if (i >= n) return;
  int v = x[i] > 0 ? x[i] : 0;
  for_face_neighbours(i, n, nbr, info, fine4, [&](int j) {
    if (j > i && x[j] - 1 > v) v = x[j] - 1;
  });
  c[i] = v;
}

// one Jacobi sweep of y_i = max(c_i, max over earlier neighbours y_j - 1)
__global__ __launch_bounds__(BLK) void k_smooth_earlier(int n, const int *__restrict__ nbr,
                                                        const uint32_t *__restrict__ info,
                                                        const int *__restrict__ fine4, const int *__restrict__ c,
                                                        const int *__restrict__ yin, int *__restrict__ yout,
                                                        int *__restrict__ changed) {
  const int i = blockIdx.x * BLK + threadIdx.x;
  if (i >= n) return;
  int v = c[i];
  for_face_neighbours(i, n, nbr, info, fine4, [&](int j) {
    if (j < i && yin[j] - 1 > v) v = yin[j] - 1;
  });
  yout[i] = v;
  if (v != yin[i]) atomicOr(changed, 1);
}

// ---- merge / split ----

// leaves written by leaf i: 1 (kept, or first of a merged family), 0 (later
// member of a merged family) or 8 (split); fv_mesh_adapt's count loop
__global__ __launch_bounds__(BLK) void k_adapt_count(int n, const int *__restrict__ root,
                                                     const int *__restrict__ level, const int *__restrict__ ci,
                                                     const int *__restrict__ split, int *__restrict__ cnt,
                                                     uint8_t *__restrict__ kind, int *__restrict__ err) {
  const int i = blockIdx.x * BLK + threadIdx.x;
  if (i >= n) return;
  const int l = level[i];
  if (l > 1) {
    const int first = i - kid_of(ci, i);
    if (family_at(first, i, n, root, level, ci)) {
      bool all = true;
      for (int k = 0; k < 8; k++) all = all && (split[first + k] < l);
      if (all) { cnt[i] = (i == first) ? 1 : 0; kind[i] = 1; return; }
    }
  }
  if (split[i] > l) {
    if (l + 1 > FV_MAXLEVEL) atomicOr(err, 1);
    cnt[i] = 8;
    kind[i] = 2;
  } else {
    cnt[i] = 1;
    kind[i] = 0;
  }
}

template <int NM>
__global__ __launch_bounds__(BLK) void k_adapt_emit(int n, int nn, Params P, const int *__restrict__ root,
                                                    const int *__restrict__ level, const int *__restrict__ ci,
                                                    const int *__restrict__ SIold, const uint8_t *__restrict__ kind,
                                                    const int *__restrict__ off, const double *__restrict__ S,
                                                    int *__restrict__ root2, int *__restrict__ level2,
                                                    int *__restrict__ ci2, int *__restrict__ SI2,
                                                    double *__restrict__ S2) {
  using L = fvk::Lay<NM>;
  const int i = blockIdx.x * BLK + threadIdx.x;
  if (i >= n) return;
  const int l = level[i], o = off[i];
  if (kind[i] == 1) {
    if (kid_of(ci, i) != 0) return;               // the first kid writes the parent
    root2[o] = root[i];
    level2[o] = l - 1;
    for (int a = 0; a < 3; a++) ci2[3 * o + a] = ci[3 * i + a] >> 1;
    double d[L::NEQ];
#pragma unroll
    for (int iv = 0; iv < L::NEQ; iv++) {
      double s = 0.0;
      for (int k = 0; k < 8; k++) s += S[(size_t)iv * n + i + k] / ((double)8);
      d[iv] = s;
    }
    fvk::relax<NM>(P, d);                         // merge.c:47-52
#pragma unroll
    for (int iv = 0; iv < L::NEQ; iv++) S2[(size_t)iv * nn + o] = d[iv];
    SI2[2 * o] = 0;
    SI2[2 * o + 1] = 0;
  } else if (kind[i] == 2) {
    for (int k = 0; k < 8; k++) {
      const int q = o + k;
      root2[q] = root[i];
      level2[q] = l + 1;
      ci2[3 * q] = 2 * ci[3 * i] + (k & 1);
      ci2[3 * q + 1] = 2 * ci[3 * i + 1] + ((k >> 1) & 1);
      ci2[3 * q + 2] = 2 * ci[3 * i + 2] + ((k >> 2) & 1);
#pragma unroll
      for (int iv = 0; iv < L::NEQ; iv++) S2[(size_t)iv * nn + q] = S[(size_t)iv * n + i];
      SI2[2 * q] = SIold[2 * i];
      SI2[2 * q + 1] = 0;
    }
  } else {
    root2[o] = root[i];
    level2[o] = l;
    for (int a = 0; a < 3; a++) ci2[3 * o + a] = ci[3 * i + a];
#pragma unroll
    for (int iv = 0; iv < L::NEQ; iv++) S2[(size_t)iv * nn + o] = S[(size_t)iv * n + i];
    SI2[2 * o] = SIold[2 * i];
    SI2[2 * o + 1] = SIold[2 * i + 1];
  }
}

// the forest side of k_adapt_emit alone (a regrid whose state is remapped by a
// fv_state_hook): addresses and interface flags of the new leaves
__global__ __launch_bounds__(BLK) void k_adapt_forest(int n, const int *__restrict__ root,
                                                      const int *__restrict__ level, const int *__restrict__ ci,
                                                      const int *__restrict__ SIold, const uint8_t *__restrict__ kind,
                                                      const int *__restrict__ off, int *__restrict__ root2,
                                                      int *__restrict__ level2, int *__restrict__ ci2,
                                                      int *__restrict__ SI2) {
  const int i = blockIdx.x * BLK + threadIdx.x;
  if (i >= n) return;
  const int l = level[i], o = off[i];
  if (kind[i] == 1) {
    if (kid_of(ci, i) != 0) return;
    root2[o] = root[i];
    level2[o] = l - 1;
    for (int a = 0; a < 3; a++) ci2[3 * o + a] = ci[3 * i + a] >> 1;
    SI2[2 * o] = 0;
    SI2[2 * o + 1] = 0;
  } else if (kind[i] == 2) {
    for (int k = 0; k < 8; k++) {
      const int q = o + k;
      root2[q] = root[i];
      level2[q] = l + 1;
      ci2[3 * q] = 2 * ci[3 * i] + (k & 1);
      ci2[3 * q + 1] = 2 * ci[3 * i + 1] + ((k >> 1) & 1);
      ci2[3 * q + 2] = 2 * ci[3 * i + 2] + ((k >> 2) & 1);
      SI2[2 * q] = SIold[2 * i];
      SI2[2 * q + 1] = 0;
    }
  } else {
    root2[o] = root[i];
    level2[o] = l;
    for (int a = 0; a < 3; a++) ci2[3 * o + a] = ci[3 * i + a];
    SI2[2 * o] = SIold[2 * i];
    SI2[2 * o + 1] = SIold[2 * i + 1];
  }
}

// ---- connectivity ----

__device__ __forceinline__ uint64_t spread3(uint32_t x) {
  uint64_t v = x & 0x3ff;
  v = (v | (v << 16)) & 0x30000ffULL;
  v = (v | (v << 8)) & 0x300f00fULL;
  v = (v | (v << 4)) & 0x30c30c3ULL;
  v = (v | (v << 2)) & 0x9249249ULL;
  return v;
}

// DFS order key: root, then the Morton code of the octant's first finest cell
__device__ __forceinline__ uint64_t dfs_key(int r, int l, int x, int y, int z) {
  const int sh = FV_MAXLEVEL - l;
  return ((uint64_t)r << 32) | spread3((uint32_t)x << sh) | (spread3((uint32_t)y << sh) << 1) |
         (spread3((uint32_t)z << sh) << 2);
}

__global__ __launch_bounds__(BLK) void k_keys(int n, const int *__restrict__ root, const int *__restrict__ level,
                                              const int *__restrict__ ci, uint64_t *__restrict__ key) {
  const int i = blockIdx.x * BLK + threadIdx.x;
  if (i >= n) return;
  key[i] = dfs_key(root[i], level[i], ci[3 * i], ci[3 * i + 1], ci[3 * i + 2]);
}

// the leaf at exactly (r, l, x, y, z), or -1
__device__ int find_leaf(int n, const uint64_t *__restrict__ key, const int *__restrict__ level, int r, int l, int x,
                         int y, int z) {
  if (l < 1 || l > FV_MAXLEVEL) return -1;
  const int side = 1 << (l - 1);
  if (x < 0 || y < 0 || z < 0 || x >= side || y >= side || z >= side) return -1;
  const uint64_t K = dfs_key(r, l, x, y, z);
  int lo = 0, hi = n;                               // first index with key > K
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (key[mid] <= K) lo = mid + 1;
    else hi = mid;
  }
  const int j = lo - 1;
  return (j >= 0 && key[j] == K && level[j] == l) ? j : -1;
}

struct Nb {
  int lfc;      // 0 boundary, 1 same or coarser, 4 finer
  int bc;       // boundary code
  int j[4];
};

// connect() of fv_mesh.c for leaf i and face f
__device__ Nb neighbour(int i, int f, int n, Forest Fo, const uint64_t *__restrict__ key,
                        const int *__restrict__ root, const int *__restrict__ level, const int *__restrict__ ci,
                        int *__restrict__ err) {
  Nb out;
  out.lfc = 0; out.bc = 0; out.j[0] = out.j[1] = out.j[2] = out.j[3] = -1;
  const int r = root[i], l = level[i], side = 1 << (l - 1);
  int c[3] = {ci[3 * i], ci[3 * i + 1], ci[3 * i + 2]};
  const int a = D_FAX[f];
  int rr = r;
  c[a] += D_FSG[f];
  if (c[a] < 0 || c[a] >= side) {
    int p[3] = {r % Fo.nroot[0], (r / Fo.nroot[0]) % Fo.nroot[1], r / (Fo.nroot[0] * Fo.nroot[1])};
    p[a] += D_FSG[f];
    if (p[a] < 0 || p[a] >= Fo.nroot[a]) rr = -1;
    else rr = p[0] + Fo.nroot[0] * (p[1] + Fo.nroot[1] * p[2]);
    c[a] = (c[a] < 0) ? side - 1 : 0;
  }
  if (rr < 0) {
    out.bc = Fo.bcside[D_BCSIDE[f]];
    if (out.bc == 0) atomicOr(err, 4);
    return out;
  }
  int j = find_leaf(n, key, level, rr, l, c[0], c[1], c[2]);
  if (j >= 0) { out.lfc = 1; out.j[0] = j; return out; }
  if (l > 1) {
    j = find_leaf(n, key, level, rr, l - 1, c[0] >> 1, c[1] >> 1, c[2] >> 1);
    if (j >= 0) { out.lfc = 1; out.j[0] = j; return out; }
  }
  const int fo = D_FCOPP[f];
  for (int q = 0; q < 4; q++) {
    const int nd = D_FCREF[fo][q];
    const int k = find_leaf(n, key, level, rr, l + 1, 2 * c[0] + (nd & 1), 2 * c[1] + ((nd >> 1) & 1),
                            2 * c[2] + ((nd >> 2) & 1));
    if (k < 0) atomicOr(err, 2);
    out.j[q] = k < 0 ? 0 : k;
  }
  out.lfc = 4;
  return out;
}

__global__ __launch_bounds__(BLK) void k_conn_count(int n, Forest Fo, const uint64_t *__restrict__ key,
                                                    const int *__restrict__ root, const int *__restrict__ level,
                                                    const int *__restrict__ ci, int *__restrict__ nfc,
                                                    int *__restrict__ err) {
  const int i = blockIdx.x * BLK + threadIdx.x;
  if (i >= n) return;
  int cnt = 0;
  for (int f = 0; f < 6; f++) cnt += neighbour(i, f, n, Fo, key, root, level, ci, err).lfc == 4;
  nfc[i] = cnt;
}

// fv_mesh_pack: neighbour table, info word with the face kinds, 1:4 rows in (leaf, face) order
__global__ __launch_bounds__(BLK) void k_conn_emit(int n, Forest Fo, const uint64_t *__restrict__ key,
                                                   const int *__restrict__ root, const int *__restrict__ level,
                                                   const int *__restrict__ ci, const int *__restrict__ SI,
                                                   const int *__restrict__ frow, int *__restrict__ nbr,
                                                   uint32_t *__restrict__ info, int *__restrict__ fine4,
                                                   int *__restrict__ err) {
  const int i = blockIdx.x * BLK + threadIdx.x;
  if (i >= n) return;
  uint32_t inf = (uint32_t)level[i] | (SI[2 * i] != 0 ? 16u : 0u);
  int kf = frow[i];
  for (int f = 0; f < 6; f++) {
    const Nb b = neighbour(i, f, n, Fo, key, root, level, ci, err);
    int kind, nb;
    if (b.lfc == 0) { kind = FV_FK_BOUNDARY; nb = -b.bc; }
    else if (b.lfc == 4) {
      kind = FV_FK_FINE;
      nb = kf;
      for (int q = 0; q < 4; q++) fine4[4 * kf + q] = b.j[q];
      kf++;
    } else {
      const int j = b.j[0];
      nb = j;
      if (level[j] < level[i]) kind = FV_FK_COARSE;
      else if (SI[2 * i] == 0 && SI[2 * j] == 0 && j < i) kind = FV_FK_MIRROR;
      else kind = FV_FK_SAME;
    }
    nbr[(size_t)f * n + i] = nb;
    inf |= (uint32_t)kind << (8 + 3 * f);
  }
  info[i] = inf;
}

// ---- host side ----

// temporaries of one call, taken from and given back to the context's pool of
// device blocks, so that a regrid allocates only when the mesh outgrows them
struct Tmp {
  fv_dev *c;
  std::vector<std::pair<void *, size_t>> p;
  explicit Tmp(fv_dev *ctx) : c(ctx) {}
  ~Tmp() {
    for (auto &b : p)
      if (b.first) c->pool.push_back(b);
    // a mesh that keeps growing leaves smaller generations behind: free the smallest
    while (c->pool.size() > 48) {
      size_t k = 0;
      for (size_t j = 1; j < c->pool.size(); j++)
        if (c->pool[j].second < c->pool[k].second) k = j;
      hipFree(c->pool[k].first);
      c->pool.erase(c->pool.begin() + k);
    }
  }
  template <class T>
  hipError_t get(T **out, size_t count) {
    const size_t bytes = (count > 0 ? count : 1) * sizeof(T);
    size_t best = (size_t)-1;
    int pick = -1;
    for (size_t k = 0; k < c->pool.size(); k++)
      if (c->pool[k].second >= bytes && c->pool[k].second < best) { best = c->pool[k].second; pick = (int)k; }
    if (pick >= 0) {
      p.push_back(c->pool[pick]);
      c->pool.erase(c->pool.begin() + pick);
      *out = (T *)p.back().first;
      return hipSuccess;
    }
    void *q = nullptr;
    const size_t alloc = bytes + bytes / 4;            // headroom for a slightly larger mesh next time
    hipError_t e = hipMalloc(&q, alloc);
    if (e == hipSuccess) p.push_back({q, alloc});
    *out = (T *)q;
    return e;
  }
  // hand a block over to the caller (it leaves the pool for good)
  void keep(const void *q) {
    for (auto &b : p)
      if (b.first == q) b.first = nullptr;
  }
};

inline dim3 grid_of(int n) { return dim3((unsigned)((n + BLK - 1) / BLK)); }

// exclusive prefix sum of cnt [n] into off [n]; returns the total in *total
int exclusive_sum(Tmp &T, hipStream_t st, const int *cnt, int *off, int n, long *total) {
  size_t bytes = 0;
  HIPCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, cnt, off, n, st));
  void *tmp = nullptr;
  HIPCHK(T.get((char **)&tmp, bytes));
  HIPCHK(hipcub::DeviceScan::ExclusiveSum(tmp, bytes, cnt, off, n, st));
  int last_off = 0, last_cnt = 0;
  HIPCHK(hipMemcpyAsync(&last_off, off + n - 1, sizeof(int), hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(&last_cnt, cnt + n - 1, sizeof(int), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  *total = (long)last_off + last_cnt;
  return FV_OK;
}

// criterionsmth on the device mesh of c: split (device, [n]) updated in place
int smooth(fv_dev *c, Tmp &T, int LEVEL, int *split) {
  const int n = c->n;
  int *x, *cv, *y2, *flag;
  HIPCHK(T.get(&x, n)); HIPCHK(T.get(&cv, n)); HIPCHK(T.get(&y2, n)); HIPCHK(T.get(&flag, 1));
  for (int pass = 0; pass < LEVEL + 1; pass++) {
    hipLaunchKernelGGL(k_smooth_family, grid_of(n), dim3(BLK), 0, c->stream, n, c->lroot, c->llevel, c->lci, split, x);
    hipLaunchKernelGGL(k_smooth_later, grid_of(n), dim3(BLK), 0, c->stream, n, c->nbr, c->info, c->fine4, x, cv);
    HIPCHK(hipMemcpyAsync(split, cv, sizeof(int) * n, hipMemcpyDeviceToDevice, c->stream));
    int *yin = split, *yout = y2;
    // sweeps past the fixpoint change nothing, so the host checks only every
    // second sweep whether the last one still changed a value
    for (int it = 0;; it++) {
      if (it > 64) return fv_set_err(FV_ERR_MESH, "criterionsmth: smoothing sweeps did not settle");
      const bool check = (it % 2) == 1;
      if (check) HIPCHK(hipMemsetAsync(flag, 0, sizeof(int), c->stream));
      hipLaunchKernelGGL(k_smooth_earlier, grid_of(n), dim3(BLK), 0, c->stream, n, c->nbr, c->info, c->fine4, cv,
                         yin, yout, flag);
      int *t = yin; yin = yout; yout = t;
      if (check) {
        int h = 0;
        HIPCHK(hipMemcpyAsync(&h, flag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        if (!h) break;
      }
    }
    if (yin != split) HIPCHK(hipMemcpyAsync(split, yin, sizeof(int) * n, hipMemcpyDeviceToDevice, c->stream));
  }
  HIPCHK(hipGetLastError());
  return FV_OK;
}

// merge/split, connectivity and pack of the mesh of c with targets split (device [n])
// and mesh interface flags SIsrc (device [n][2]); installs the new mesh and state in c
int adapt_and_connect(fv_dev *c, Tmp &T, const int *split, const int *SIsrc, const fv_state_hook *h = nullptr) {
  const int n = c->n, neq = c->neq;
  int *cnt, *off, *err;
  uint8_t *kind;
  HIPCHK(T.get(&cnt, n)); HIPCHK(T.get(&off, n)); HIPCHK(T.get(&kind, n)); HIPCHK(T.get(&err, 1));
  HIPCHK(hipMemsetAsync(err, 0, sizeof(int), c->stream));
  hipLaunchKernelGGL(k_adapt_count, grid_of(n), dim3(BLK), 0, c->stream, n, c->lroot, c->llevel, c->lci, split, cnt,
                     kind, err);
  long total = 0;
  int rc = exclusive_sum(T, c->stream, cnt, off, n, &total);
  if (rc != FV_OK) return rc;
  int herr = 0;
  HIPCHK(hipMemcpy(&herr, err, sizeof(int), hipMemcpyDeviceToHost));
  if (herr & 1) return fv_set_err(FV_ERR_MESH, "regrid: a leaf would exceed FV_MAXLEVEL");
  if (total <= 0 || total > 0x7fffffffL) return fv_set_err(FV_ERR_MESH, "regrid: leaf count out of range");
  const int nn = (int)total;
  int *root2, *level2, *ci2, *SI2;
  double *S2;
  HIPCHK(T.get(&root2, nn)); HIPCHK(T.get(&level2, nn));
  HIPCHK(T.get(&ci2, 3 * (size_t)nn)); HIPCHK(T.get(&SI2, 2 * (size_t)nn));
  if (h) {
    // the state is the hook's: the forest side here, the state remap there
    S2 = nullptr;
    hipLaunchKernelGGL(k_adapt_forest, grid_of(n), dim3(BLK), 0, c->stream, n, c->lroot, c->llevel, c->lci, SIsrc,
                       kind, off, root2, level2, ci2, SI2);
    HIPCHK(hipGetLastError());
    if ((rc = h->remap(h->ctx, c->stream, n, nn, kind, off, c->lci)) != FV_OK) return rc;
  } else {
    HIPCHK(T.get(&S2, (size_t)nn * neq));
    if (c->nmat == 2)
      hipLaunchKernelGGL(k_adapt_emit<2>, grid_of(n), dim3(BLK), 0, c->stream, n, nn, c->P, c->lroot, c->llevel,
                         c->lci, SIsrc, kind, off, c->S, root2, level2, ci2, SI2, S2);
    else
      hipLaunchKernelGGL(k_adapt_emit<3>, grid_of(n), dim3(BLK), 0, c->stream, n, nn, c->P, c->lroot, c->llevel,
                         c->lci, SIsrc, kind, off, c->S, root2, level2, ci2, SI2, S2);
  }
  HIPCHK(hipGetLastError());
  // connectivity of the new leaves
  Forest Fo;
  for (int k = 0; k < 3; k++) Fo.nroot[k] = c->nroot[k];
  for (int k = 0; k < 6; k++) Fo.bcside[k] = c->bcside[k];
  uint64_t *key;
  int *nfc, *frow, *nbr2, *fine2;
  uint32_t *info2;
  HIPCHK(T.get(&key, nn)); HIPCHK(T.get(&nfc, nn)); HIPCHK(T.get(&frow, nn));
  HIPCHK(hipMemsetAsync(err, 0, sizeof(int), c->stream));
  hipLaunchKernelGGL(k_keys, grid_of(nn), dim3(BLK), 0, c->stream, nn, root2, level2, ci2, key);
  hipLaunchKernelGGL(k_conn_count, grid_of(nn), dim3(BLK), 0, c->stream, nn, Fo, key, root2, level2, ci2, nfc, err);
  long nfine = 0;
  if ((rc = exclusive_sum(T, c->stream, nfc, frow, nn, &nfine)) != FV_OK) return rc;
  HIPCHK(hipMemcpy(&herr, err, sizeof(int), hipMemcpyDeviceToHost));
  if (herr & 2) return fv_set_err(FV_ERR_MESH, "regrid: the new forest is not 2:1 face balanced");
  if (herr & 4) return fv_set_err(FV_ERR_ARG, "regrid: a boundary face needs a non-zero BC code");
  HIPCHK(T.get(&nbr2, (size_t)nn * 6)); HIPCHK(T.get(&info2, nn)); HIPCHK(T.get(&fine2, 4 * (size_t)(nfine > 0 ? nfine : 1)));
  hipLaunchKernelGGL(k_conn_emit, grid_of(nn), dim3(BLK), 0, c->stream, nn, Fo, key, root2, level2, ci2, SI2, frow,
                     nbr2, info2, fine2, err);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  // install: the device buffers may be reallocated (with headroom, so that a
  // slowly growing mesh does not reallocate every time); the new state waits in S2
  if (nn > c->cap_n || nfine > c->cap_fine) {
    if ((rc = fv_dev_reserve(c, nn + nn / 4, (int)(nfine + nfine / 4))) != FV_OK) return rc;
  }
  c->broken = 1;   // until the new mesh and state are installed
  c->brick_ok = 0;
  c->n = nn;
  c->nfine = (int)nfine;
  if (!h) HIPCHK(hipMemcpyAsync(c->S, S2, sizeof(double) * (size_t)nn * neq, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->nbr, nbr2, sizeof(int) * (size_t)nn * 6, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->info, info2, sizeof(uint32_t) * nn, hipMemcpyDeviceToDevice, c->stream));
  if (nfine > 0)
    HIPCHK(hipMemcpyAsync(c->fine4, fine2, sizeof(int) * 4 * (size_t)nfine, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(hipMemsetAsync(c->nanflag, 0, sizeof(int), c->stream));
  HIPCHK(hipMemsetAsync(c->lim, 0, (size_t)nn, c->stream));
  // the forest arrays: the new ones belong to the context from here on, the old
  // ones go to the pool once the copies above are done
  HIPCHK(hipStreamSynchronize(c->stream));
  const size_t old_n = (size_t)(c->cap_leaf > 0 ? c->cap_leaf : 1);
  if (c->lroot) c->pool.push_back({c->lroot, old_n * sizeof(int)});
  if (c->llevel) c->pool.push_back({c->llevel, old_n * sizeof(int)});
  if (c->lci) c->pool.push_back({c->lci, 3 * old_n * sizeof(int)});
  if (c->lSI) c->pool.push_back({c->lSI, 2 * old_n * sizeof(int)});
  c->lroot = root2; c->llevel = level2; c->lci = ci2; c->lSI = SI2;
  c->cap_leaf = nn;
  T.keep(root2); T.keep(level2); T.keep(ci2); T.keep(SI2);
  c->broken = 0;
  if (h && (rc = h->install(h->ctx, c)) != FV_OK) return rc;
  return FV_OK;
}

int need_forest(fv_dev *c) {
  if (!c) return fv_set_err(FV_ERR_ARG, "null argument");
  if (!c->forest) return fv_set_err(FV_ERR_ARG, "no forest: call fv_dev_set_forest after attach/reattach");
  return FV_OK;
}

}  // namespace

extern "C" int fv_dev_set_forest(fv_dev *c, const fv_mesh *m) {
  FV_GUARD(c);
  if (!c || !m) return fv_set_err(FV_ERR_ARG, "null argument");
  if (m->n != c->n) return fv_set_err(FV_ERR_ARG, "fv_dev_set_forest: mesh and context differ in leaf count");
  const size_t n = (size_t)m->n;
  if (m->n > c->cap_leaf) {
    hipFree(c->lroot); hipFree(c->llevel); hipFree(c->lci); hipFree(c->lSI);
    c->lroot = c->llevel = c->lci = c->lSI = nullptr;
    c->cap_leaf = 0;
    HIPCHK(hipMalloc(&c->lroot, sizeof(int) * n));
    HIPCHK(hipMalloc(&c->llevel, sizeof(int) * n));
    HIPCHK(hipMalloc(&c->lci, sizeof(int) * 3 * n));
    HIPCHK(hipMalloc(&c->lSI, sizeof(int) * 2 * n));
    c->cap_leaf = m->n;
  }
  HIPCHK(hipMemcpy(c->lroot, m->root, sizeof(int) * n, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->llevel, m->level, sizeof(int) * n, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->lci, m->ci, sizeof(int) * 3 * n, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->lSI, m->SI, sizeof(int) * 2 * n, hipMemcpyHostToDevice));
  for (int k = 0; k < 3; k++) c->nroot[k] = m->nroot[k];
  for (int k = 0; k < 6; k++) c->bcside[k] = m->bcside[k];
  c->forest = 1;
  return FV_OK;
}

extern "C" int fv_dev_get_forest(fv_dev *c, int *root, int *level, int *ci, int *SI) {
  FV_GUARD(c);
  int rc = need_forest(c);
  if (rc != FV_OK) return rc;
  const size_t n = (size_t)c->n;
  if (root && (rc = fv_copy_to_host(c, root, c->lroot, sizeof(int) * n)) != FV_OK) return rc;
  if (level && (rc = fv_copy_to_host(c, level, c->llevel, sizeof(int) * n)) != FV_OK) return rc;
  if (ci && (rc = fv_copy_to_host(c, ci, c->lci, sizeof(int) * 3 * n)) != FV_OK) return rc;
  if (SI && (rc = fv_copy_to_host(c, SI, c->lSI, sizeof(int) * 2 * n)) != FV_OK) return rc;
  return FV_OK;
}

extern "C" int fv_dev_get_ntrees(fv_dev *c, int *ntrees) {
  FV_GUARD(c);
  int rc = need_forest(c);
  if (rc != FV_OK) return rc;
  if (!ntrees) return fv_set_err(FV_ERR_ARG, "null argument");
  *ntrees = c->nroot[0] * c->nroot[1] * c->nroot[2];
  return FV_OK;
}

extern "C" int fv_dev_get_pack(fv_dev *c, int *nbr, uint32_t *info, int *nfine, int *fine4) {
  FV_GUARD(c);
  if (!c) return fv_set_err(FV_ERR_ARG, "null argument");
  const size_t n = (size_t)c->n;
  int rc;
  if (nfine) *nfine = c->nfine;
  if (nbr && (rc = fv_copy_to_host(c, nbr, c->nbr, sizeof(int) * 6 * n)) != FV_OK) return rc;
  if (info && (rc = fv_copy_to_host(c, info, c->info, sizeof(uint32_t) * n)) != FV_OK) return rc;
  if (fine4 && c->nfine > 0 && (rc = fv_copy_to_host(c, fine4, c->fine4, sizeof(int) * 4 * (size_t)c->nfine)) != FV_OK)
    return rc;
  return FV_OK;
}

extern "C" int fv_dev_criterionsmth(fv_dev *c, int LEVEL, int *split) {
  FV_GUARD(c);
  int rc = need_forest(c);
  if (rc != FV_OK) return rc;
  if (!split) return fv_set_err(FV_ERR_ARG, "null argument");
  Tmp T(c);
  int *d;
  HIPCHK(T.get(&d, c->n));
  HIPCHK(hipMemcpy(d, split, sizeof(int) * c->n, hipMemcpyHostToDevice));
  if ((rc = smooth(c, T, LEVEL, d)) != FV_OK) return rc;
  return fv_copy_to_host(c, split, d, sizeof(int) * c->n);
}

extern "C" int fv_dev_regrid(fv_dev *c, const fv_crit *cr, int maxlevel, int *n_out) {
  FV_GUARD(c);
  int rc = need_forest(c);
  if (rc != FV_OK) return rc;
  if (!cr) return fv_set_err(FV_ERR_ARG, "null argument");
  Tmp T(c);
  int *SIdev = nullptr;
  if ((rc = fv_dev_indicator_raw(c, cr, &SIdev)) != FV_OK) return rc;
  // the new interface flags become the mesh's (interface_location before meshadaptation)
  HIPCHK(hipMemcpyAsync(c->lSI, SIdev, sizeof(int) * 2 * (size_t)c->n, hipMemcpyDeviceToDevice, c->stream));
  int *split;
  HIPCHK(T.get(&split, c->n));
  HIPCHK(hipMemcpyAsync(split, c->split, sizeof(int) * c->n, hipMemcpyDeviceToDevice, c->stream));
  if ((rc = smooth(c, T, maxlevel, split)) != FV_OK) return rc;
  if ((rc = adapt_and_connect(c, T, split, c->lSI)) != FV_OK) return rc;
  if (n_out) *n_out = c->n;
  return FV_OK;
}

int fv_dev_regrid_hook(fv_dev *c, int *split, const int *SI, int maxlevel, const fv_state_hook *h, int *n_out) {
  FV_GUARD(c);
  int rc = need_forest(c);
  if (rc != FV_OK) return rc;
  if (!split || !SI || !h) return fv_set_err(FV_ERR_ARG, "null argument");
  Tmp T(c);
  if (SI != c->lSI)
    HIPCHK(hipMemcpyAsync(c->lSI, SI, sizeof(int) * 2 * (size_t)c->n, hipMemcpyDeviceToDevice, c->stream));
  if ((rc = smooth(c, T, maxlevel, split)) != FV_OK) return rc;
  if ((rc = adapt_and_connect(c, T, split, c->lSI, h)) != FV_OK) return rc;
  if (n_out) *n_out = c->n;
  return FV_OK;
}
