// fv_mlp.hip -- learned refinement / limiter indicator (new; the reference
// has only the plain gradient indicator of criterionsplt.c:40-140).
//
// Input: a 3x3x3 stencil of 4 normalised fields per leaf (rho/rho_c - 1,
// relative pressure jump (p - p_c)/max(|p|,|p_c|), volume fraction, |u|/c_c)
// = 108 features, zero-padded to K = 112.  Network 112 -> 64 -> 64 -> 4 with
// SiLU; outputs are logits of refine / keep / coarsen and of the smooth flag
// (MC limiter instead of minmod).  Labels come from the plain indicator on the
// same state: its target level against the leaf level, and a smooth flag set
// where the largest relative pressure jump to a face neighbour is under a
// threshold (no shock or acoustic front: smooth flow and contacts, where the
// compressive MC limiter pays, against minmod across pressure jumps).
//
// Training is FP32 with hand-written forward, backward and Adam kernels (no
// atomics: every weight gradient is one thread's ordered sum over the batch).
// Inference is a BF16 MFMA GEMM chain (v_mfma_f32_16x16x16_bf16, FP32
// accumulation) with the hidden activations staged through LDS; any leaf the
// plain indicator refines is refined whatever the network says.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <vector>
#include "fv_internal.h"

using namespace fvk;

#define MLP_NIN 108
#define MLP_K 112
#define MLP_H 64
#define MLP_NO 4
#define MLP_HP (MLP_H + 2)   // row pitch of the BF16 activations in LDS
#define MLP_NW (MLP_K * MLP_H + MLP_H + MLP_H * MLP_H + MLP_H + MLP_H * MLP_NO + MLP_NO)
#define MLP_OFF_B1 (MLP_K * MLP_H)
#define MLP_OFF_W2 (MLP_OFF_B1 + MLP_H)
#define MLP_OFF_B2 (MLP_OFF_W2 + MLP_H * MLP_H)
#define MLP_OFF_W3 (MLP_OFF_B2 + MLP_H)
#define MLP_OFF_B3 (MLP_OFF_W3 + MLP_H * MLP_NO)

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

struct fv_mlp {
  float *w = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;   // FP32 master weights, grad, Adam moments
  uint16_t *w1t = nullptr, *w2t = nullptr, *w3t = nullptr;        // BF16 transposed copies for inference
  int bf16_valid = 0;
  long adam_t = 0;
  // dataset (device)
  float *X = nullptr, *smooth = nullptr;
  int *cls = nullptr;
  long nd = 0, cap = 0;
  // batch scratch
  int bcap = 0;
  float *bx = nullptr, *bs = nullptr, *z1 = nullptr, *h1 = nullptr, *z2 = nullptr, *h2 = nullptr, *o = nullptr;
  float *d3 = nullptr, *d2 = nullptr, *d1 = nullptr, *lossv = nullptr, *wpart = nullptr;
  int *bc = nullptr, *bidx = nullptr;
  // per-leaf stencil fields
  long fcap = 0;
  float4 *fld = nullptr;
  float *snd = nullptr;
  // inference scratch
  long icap = 0, xcap = 0;
  float *ix = nullptr, *iout = nullptr;
  uint8_t *ismooth = nullptr;
  float cw[3] = {1.f, 1.f, 1.f};
  float smooth_thr = 0.1f;
};

// ---------------------------------------------------------------------
// features and labels
// ---------------------------------------------------------------------
struct CellF { double rho, p, alpha, umag, c; };

template <int NM>
FV_DEV CellF cell_fields(const Params &P, const double *__restrict__ S, int n, int j) {
  using L = Lay<NM>;
  double X[L::NEQ];
#pragma unroll
  for (int v = 0; v < L::NEQ; v++) X[v] = S[(size_t)v * n + j];
  Sol<NM, double> s;
  fill_sol<NM, double>(P, X, s);
  CellF o;
  o.rho = s.r;
  double pm = 0.0, c2 = 0.0;
#pragma unroll
  for (int k = 0; k < NM; k++) { pm += s.avf[k] * s.p[k]; c2 += s.c2Y[k]; }
  o.p = pm;
  o.alpha = s.avf[0];
  o.umag = sqrt(s.u * s.u + s.v * s.v + s.w * s.w);
  o.c = sqrt(c2);
  return o;
}

FV_DEV int step_face(int j, int f, int n, const int *__restrict__ nbr, const uint32_t *__restrict__ info,
                     const int *__restrict__ fine4) {
  const int kind = face_kind(info[j], f);
  const int nb = nbr[(size_t)f * n + j];
  if (kind == FV_FK_BOUNDARY) return j;
  if (kind == FV_FK_FINE) return fine4[4 * nb];
  return nb;
}

FV_DEV float clampf(float x) { return fminf(8.f, fmaxf(-8.f, x)); }

// per-leaf fields for the stencil features: (rho, p, alpha_0, |u|) and c
template <int NM>
__global__ __launch_bounds__(256) void k_mlp_fields(int n, Params P, const double *__restrict__ S,
                                                    float4 *__restrict__ fld, float *__restrict__ snd) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const CellF c = cell_fields<NM>(P, S, n, i);
  fld[i] = make_float4((float)c.rho, (float)c.p, (float)c.alpha, (float)c.umag);
  snd[i] = (float)c.c;
}

__global__ __launch_bounds__(256) void k_mlp_features(int n, const float4 *__restrict__ fld,
                                                      const float *__restrict__ snd, const int *__restrict__ nbr,
                                                      const uint32_t *__restrict__ info,
                                                      const int *__restrict__ fine4, float *__restrict__ X, long row0) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 c0 = fld[i];
  const float ic = 1.f / fmaxf(snd[i], 1e-30f);
  float *x = X + (size_t)(row0 + i) * MLP_K;
  const int fx[2] = {3, 1}, fy[2] = {0, 2}, fz[2] = {4, 5};
  int s = 0;
  for (int dz = -1; dz <= 1; dz++)
    for (int dy = -1; dy <= 1; dy++)
      for (int dx = -1; dx <= 1; dx++, s++) {
        int j = i;
        if (dx) j = step_face(j, fx[dx > 0], n, nbr, info, fine4);
        if (dy) j = step_face(j, fy[dy > 0], n, nbr, info, fine4);
        if (dz) j = step_face(j, fz[dz > 0], n, nbr, info, fine4);
        const float4 cj = (j == i) ? c0 : fld[j];
        const float pden = fmaxf(fmaxf(fabsf(cj.y), fabsf(c0.y)), 1e-30f);
        x[4 * s + 0] = clampf(cj.x / c0.x - 1.f);
        x[4 * s + 1] = clampf((cj.y - c0.y) / pden);
        x[4 * s + 2] = clampf(cj.z);
        x[4 * s + 3] = clampf(cj.w * ic);
      }
  for (int k = MLP_NIN; k < MLP_K; k++) x[k] = 0.f;
}

// class: 0 refine (target above the level), 1 keep, 2 coarsen; smooth: largest
// relative pressure jump to a face neighbour below thr
template <int NM>
__global__ __launch_bounds__(256) void k_mlp_labels(int n, Params P, const double *__restrict__ S,
                                                    const int *__restrict__ nbr, const uint32_t *__restrict__ info,
                                                    const int *__restrict__ fine4, const int *__restrict__ split,
                                                    float thr, int *__restrict__ cls, float *__restrict__ smooth,
                                                    long row0) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t inf = info[i];
  const int lvl = cell_level(inf);
  cls[row0 + i] = split[i] > lvl ? 0 : (split[i] < lvl ? 2 : 1);
  const CellF c0 = cell_fields<NM>(P, S, n, i);
  double jmax = 0.0;
  for (int f = 0; f < 6; f++) {
    const int kind = face_kind(inf, f);
    if (kind == FV_FK_BOUNDARY) continue;
    const int nb = nbr[(size_t)f * n + i];
    const int ns = (kind == FV_FK_FINE) ? 4 : 1;
    for (int k = 0; k < ns; k++) {
      const int j = (kind == FV_FK_FINE) ? fine4[4 * nb + k] : nb;
      const CellF cj = cell_fields<NM>(P, S, n, j);
      jmax = fmax(jmax, fabs(cj.p - c0.p) / fmax(fmax(fabs(cj.p), fabs(c0.p)), 1e-30));
    }
  }
  smooth[row0 + i] = (jmax < (double)thr) ? 1.f : 0.f;
}

// ---------------------------------------------------------------------
// FP32 training kernels
// ---------------------------------------------------------------------
FV_DEV float sigm(float z) { return 1.f / (1.f + expf(-z)); }

__global__ void k_gather_rows(int B, const int *__restrict__ idx, const float *__restrict__ X, const int *__restrict__ cls,
                              const float *__restrict__ sm, float *__restrict__ bx, int *__restrict__ bc,
                              float *__restrict__ bs) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= B * MLP_K) return;
  const int b = t / MLP_K, k = t % MLP_K;
  const long r = idx[b];
  bx[t] = X[r * MLP_K + k];
  if (k == 0) { bc[b] = cls[r]; bs[b] = sm[r]; }
}

// Z = X W + bias (X: [B][K] row stride ldx, W: [K][N]); H = silu(Z) when act
__global__ void k_fwd(int B, int K, int N, const float *__restrict__ X, int ldx, const float *__restrict__ W,
                      const float *__restrict__ bias, float *__restrict__ Z, float *__restrict__ H, int act) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= B * N) return;
  const int b = t / N, nn = t % N;
  float z = bias[nn];
  for (int k = 0; k < K; k++) z += X[(size_t)b * ldx + k] * W[k * N + nn];
  Z[t] = z;
  H[t] = act ? z * sigm(z) : z;
}

// softmax cross entropy on logits 0..2 (class weights cw) + BCE on logit 3
__global__ void k_loss(int B, const float *__restrict__ O, const int *__restrict__ cls, const float *__restrict__ sm,
                       float w0, float w1, float w2, float *__restrict__ dO, float *__restrict__ lossv) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  const float *o = O + 4 * b;
  const float mx = fmaxf(o[0], fmaxf(o[1], o[2]));
  const float e0 = expf(o[0] - mx), e1 = expf(o[1] - mx), e2 = expf(o[2] - mx);
  const float se = e0 + e1 + e2;
  const int y = cls[b];
  const float wy = (y == 0) ? w0 : (y == 1 ? w1 : w2);
  const float p[3] = {e0 / se, e1 / se, e2 / se};
  const float s = sigm(o[3]);
  const float t = sm[b];
  const float iB = 1.f / (float)B;
  for (int c = 0; c < 3; c++) dO[4 * b + c] = wy * (p[c] - (c == y ? 1.f : 0.f)) * iB;
  dO[4 * b + 3] = (s - t) * iB;
  const float ce = -wy * logf(fmaxf(p[y], 1e-30f));
  const float bce = -(t * logf(fmaxf(s, 1e-30f)) + (1.f - t) * logf(fmaxf(1.f - s, 1e-30f)));
  lossv[b] = ce + bce;
}

// dZprev = (dZ W^T) * silu'(Zprev)   (dZ: [B][N], W: [K][N], Zprev: [B][K])
__global__ void k_bwd_data(int B, int K, int N, const float *__restrict__ dZ, const float *__restrict__ W,
                           const float *__restrict__ Zp, float *__restrict__ dZp) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= B * K) return;
  const int b = t / K, k = t % K;
  float dh = 0.f;
  for (int nn = 0; nn < N; nn++) dh += dZ[b * N + nn] * W[k * N + nn];
  const float z = Zp[t], s = sigm(z);
  dZp[t] = dh * (s + z * s * (1.f - s));
}

// weight gradient in two deterministic steps: partial sums over chunks of
// MLP_WCHUNK batch rows, then the chunks summed in order.
// part[c][k][n] = sum_{b in chunk c} X[b][k] dZ[b][n] (rows k < K) and
// sum_{b in chunk c} dZ[b][n] (row K, the bias)
#define MLP_WCHUNK 64
__global__ void k_wgrad_part(int B, int K, int N, const float *__restrict__ X, int ldx, const float *__restrict__ dZ,
                             float *__restrict__ part) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int per = (K + 1) * N;
  const int nch = (B + MLP_WCHUNK - 1) / MLP_WCHUNK;
  if (t >= (long)per * nch) return;
  const int c = (int)(t / per), r = (int)(t % per);
  const int k = r / N, nn = r % N;
  const int b0 = c * MLP_WCHUNK, b1 = min(B, b0 + MLP_WCHUNK);
  float acc = 0.f;
  if (k < K) {
    for (int b = b0; b < b1; b++) acc += X[(size_t)b * ldx + k] * dZ[b * N + nn];
  } else {
    for (int b = b0; b < b1; b++) acc += dZ[b * N + nn];
  }
  part[t] = acc;
}

__global__ void k_wgrad_sum(int B, int K, int N, const float *__restrict__ part, float *__restrict__ dW,
                            float *__restrict__ db) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int per = (K + 1) * N;
  if (t >= per) return;
  const int nch = (B + MLP_WCHUNK - 1) / MLP_WCHUNK;
  float acc = 0.f;
  for (int c = 0; c < nch; c++) acc += part[(size_t)c * per + t];
  if (t < K * N) dW[t] = acc;
  else db[t - K * N] = acc;
}

__global__ void k_adam(int np, float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= np) return;
  const float gt = g[t];
  const float mt = b1 * m[t] + (1.f - b1) * gt;
  const float vt = b2 * v[t] + (1.f - b2) * gt * gt;
  m[t] = mt;
  v[t] = vt;
  w[t] -= lr * (mt / c1) / (sqrtf(vt / c2) + eps);
}

// ---------------------------------------------------------------------
// BF16 MFMA inference
// ---------------------------------------------------------------------
FV_DEV uint16_t to_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);   // round to nearest even
  return (uint16_t)(u >> 16);
}

__global__ void k_to_bf16_t(int K, int N, int Npad, const float *__restrict__ W, uint16_t *__restrict__ Wt) {
  // Wt[n][k] = bf16(W[k][n]) for n < N, zero for N <= n < Npad
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= K * Npad) return;
  const int nn = t / K, k = t % K;
  Wt[t] = (nn < N) ? to_bf16(W[k * N + nn]) : (uint16_t)0;
}

FV_DEV bf16x4 pack4(uint16_t a, uint16_t b, uint16_t c, uint16_t d) {
  bf16x4 r;
  r[0] = (short)a; r[1] = (short)b; r[2] = (short)c; r[3] = (short)d;
  return r;
}

// one wave = 16 samples; 4 waves per block.  A operand lane l: row l%16,
// k = 4*(l/16)+0..3; B operand lane l: column l%16, same k; D lane l: rows
// 4*(l/16)+0..3, column l%16.  xa(row, k) returns the FP32 input feature.
template <typename XA>
FV_DEV void mlp_infer_wave(long n, long r0, int wv, int lane, XA xa, const uint16_t *__restrict__ w1t,
                           const uint16_t *__restrict__ w2t, const uint16_t *__restrict__ w3t,
                           const float *__restrict__ wf, uint16_t (*hs)[16][MLP_HP], float *__restrict__ out) {
  const int row = lane & 15, kq = 4 * (lane >> 4);
  const float *b1 = wf + MLP_OFF_B1, *b2 = wf + MLP_OFF_B2, *b3 = wf + MLP_OFF_B3;
  // layer 1: [16 x 112] x [112 x 64]
  for (int nt = 0; nt < 4; nt++) {
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int kt = 0; kt < MLP_K / 16; kt++) {
      const int k0 = kt * 16 + kq;
      const bf16x4 a = pack4(xa(row, k0), xa(row, k0 + 1), xa(row, k0 + 2), xa(row, k0 + 3));
      const uint16_t *wb = w1t + (size_t)(nt * 16 + row) * MLP_K + k0;
      const bf16x4 b = pack4(wb[0], wb[1], wb[2], wb[3]);
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
    for (int q = 0; q < 4; q++) {
      const float z = acc[q] + b1[nt * 16 + row];
      hs[wv][kq + q][nt * 16 + row] = to_bf16(z / (1.f + expf(-z)));
    }
  }
  __syncthreads();
  // layer 2: [16 x 64] x [64 x 64]
  f32x4 acc2[4];
  for (int nt = 0; nt < 4; nt++) {
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int kt = 0; kt < MLP_H / 16; kt++) {
      const uint16_t *ha = &hs[wv][row][kt * 16 + kq];
      const bf16x4 a = pack4(ha[0], ha[1], ha[2], ha[3]);
      const uint16_t *wb = w2t + (size_t)(nt * 16 + row) * MLP_H + kt * 16 + kq;
      const bf16x4 b = pack4(wb[0], wb[1], wb[2], wb[3]);
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
    acc2[nt] = acc;
  }
  __syncthreads();
  for (int nt = 0; nt < 4; nt++)
    for (int q = 0; q < 4; q++) {
      const float z = acc2[nt][q] + b2[nt * 16 + row];
      hs[wv][kq + q][nt * 16 + row] = to_bf16(z / (1.f + expf(-z)));
    }
  __syncthreads();
  // layer 3: [16 x 64] x [64 x 16] (4 outputs used)
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  for (int kt = 0; kt < MLP_H / 16; kt++) {
    const uint16_t *ha = &hs[wv][row][kt * 16 + kq];
    const bf16x4 a = pack4(ha[0], ha[1], ha[2], ha[3]);
    const uint16_t *wb = w3t + (size_t)row * MLP_H + kt * 16 + kq;
    const bf16x4 b = pack4(wb[0], wb[1], wb[2], wb[3]);
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
  }
  if (row < MLP_NO)
    for (int q = 0; q < 4; q++) {
      const long r = r0 + kq + q;
      if (r < n) out[r * MLP_NO + row] = acc[q] + b3[row];
    }
}

// inference of feature rows stored in HBM (host batches)
__global__ __launch_bounds__(256) void k_mlp_infer(long n, const float *__restrict__ X, const uint16_t *__restrict__ w1t,
                                                   const uint16_t *__restrict__ w2t, const uint16_t *__restrict__ w3t,
                                                   const float *__restrict__ wf, float *__restrict__ out) {
  __shared__ uint16_t hs[4][16][MLP_HP];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const long r0 = ((long)blockIdx.x * 4 + wv) * 16;
  auto xa = [&](int row, int k) {
    const long rr = (r0 + row < n) ? r0 + row : n - 1;
    return to_bf16(X[rr * MLP_K + k]);
  };
  mlp_infer_wave(n, r0, wv, lane, xa, w1t, w2t, w3t, wf, hs, out);
}

// inference of every leaf with the stencil features gathered straight into LDS
// (no feature matrix in HBM); same features as k_mlp_features, same MFMA order
// as k_mlp_infer, so the outputs equal k_mlp_infer on k_mlp_features' rows
__global__ __launch_bounds__(256) void k_mlp_infer_stencil(int n, const float4 *__restrict__ fld,
                                                           const float *__restrict__ snd, const int *__restrict__ nbr,
                                                           const uint32_t *__restrict__ info,
                                                           const int *__restrict__ fine4,
                                                           const uint16_t *__restrict__ w1t,
                                                           const uint16_t *__restrict__ w2t,
                                                           const uint16_t *__restrict__ w3t,
                                                           const float *__restrict__ wf, float *__restrict__ out) {
  // the features and hidden activations are kept as the BF16 bits the MFMAs
  // read, and the neighbour table shares its memory with the activations
  // (it is dead once the features are written): 23 KB per block
  __shared__ uint16_t xs[4][16][MLP_K + 8];
  __shared__ __align__(16) unsigned char nh[(sizeof(int) * 4 * 16 * 28 > sizeof(uint16_t) * 4 * 16 * MLP_HP)
                                                ? sizeof(int) * 4 * 16 * 28 : sizeof(uint16_t) * 4 * 16 * MLP_HP];
  int (*nix)[16][28] = reinterpret_cast<int (*)[16][28]>(nh);
  uint16_t (*hs)[16][MLP_HP] = reinterpret_cast<uint16_t (*)[16][MLP_HP]>(nh);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const long r0 = ((long)blockIdx.x * 4 + wv) * 16;
  // four lanes per row; the 26 stencil neighbours are found level by level
  // (faces, then edges from the x/y faces, then corners from the x-y edges),
  // the same face chains x -> y -> z as k_mlp_features
  const int row = lane >> 2, q = lane & 3;
  const int i = (int)((r0 + row < n) ? r0 + row : n - 1);
  int *nb = nix[wv][row];
  const int fx[2] = {3, 1}, fy[2] = {0, 2}, fz[2] = {4, 5};
  auto sidx = [](int dx, int dy, int dz) { return (dz + 1) * 9 + (dy + 1) * 3 + (dx + 1); };
  if (q == 0) nb[13] = i;
  // fixed trip counts, so that a lane's independent neighbour chains overlap
#pragma unroll
  for (int t = 0; t < 2; t++) {
    const int e = q + 4 * t;
    if (e >= 6) break;
    const int ax = e >> 1, pm = e & 1, d = pm ? 1 : -1;
    const int f = (ax == 0) ? fx[pm] : (ax == 1 ? fy[pm] : fz[pm]);
    nb[sidx(ax == 0 ? d : 0, ax == 1 ? d : 0, ax == 2 ? d : 0)] = step_face(i, f, n, nbr, info, fine4);
  }
  __syncthreads();
  {
    // the table reads first, then the chains, then the writes
    int src[3], fc[3], dst[3], res[3];
#pragma unroll
    for (int t = 0; t < 3; t++) {
      const int e = q + 4 * t;
      const int kind = e >> 2, a = (e & 1) ? 1 : -1, b = (e & 2) ? 1 : -1;
      src[t] = (kind == 2) ? nb[sidx(0, a, 0)] : nb[sidx(a, 0, 0)];
      fc[t] = (kind == 0) ? fy[b > 0] : fz[b > 0];
      dst[t] = (kind == 0) ? sidx(a, b, 0) : (kind == 1 ? sidx(a, 0, b) : sidx(0, a, b));
    }
#pragma unroll
    for (int t = 0; t < 3; t++) res[t] = step_face(src[t], fc[t], n, nbr, info, fine4);
#pragma unroll
    for (int t = 0; t < 3; t++) nb[dst[t]] = res[t];
  }
  __syncthreads();
  {
    int src[2], fc[2], dst[2], res[2];
#pragma unroll
    for (int t = 0; t < 2; t++) {
      const int e = q + 4 * t;
      const int a = (e & 1) ? 1 : -1, b = (e & 2) ? 1 : -1, c = (e & 4) ? 1 : -1;
      src[t] = nb[sidx(a, b, 0)];
      fc[t] = fz[c > 0];
      dst[t] = sidx(a, b, c);
    }
#pragma unroll
    for (int t = 0; t < 2; t++) res[t] = step_face(src[t], fc[t], n, nbr, info, fine4);
#pragma unroll
    for (int t = 0; t < 2; t++) nb[dst[t]] = res[t];
  }
  __syncthreads();
  {
    const float4 c0 = fld[i];
    const float ic = 1.f / fmaxf(snd[i], 1e-30f);
#pragma unroll
    for (int t = 0; t < 7; t++) {
      const int s = q + 4 * t;
      if (s >= 27) break;
      const int j = nb[s];
      const float4 cj = (j == i) ? c0 : fld[j];
      const float pden = fmaxf(fmaxf(fabsf(cj.y), fabsf(c0.y)), 1e-30f);
      xs[wv][row][4 * s + 0] = to_bf16(clampf(cj.x / c0.x - 1.f));
      xs[wv][row][4 * s + 1] = to_bf16(clampf((cj.y - c0.y) / pden));
      xs[wv][row][4 * s + 2] = to_bf16(clampf(cj.z));
      xs[wv][row][4 * s + 3] = to_bf16(clampf(cj.w * ic));
    }
    xs[wv][row][MLP_NIN + q] = to_bf16(0.f);
  }
  __syncthreads();
  auto xa = [&](int row, int k) { return xs[wv][row][k]; };
  mlp_infer_wave((long)n, r0, wv, lane, xa, w1t, w2t, w3t, wf, hs, out);
}

// target level from the network, with the plain indicator's refinement forced:
// where the plain target is at or above the leaf's level (it refines or keeps
// the leaf, as in its interface bands) it is a floor, so the network never
// coarsens a leaf that the plain indicator would refine back at the next regrid
__global__ void k_mlp_decide(int n, const float *__restrict__ out, const uint32_t *__restrict__ info,
                             int *__restrict__ split, uint8_t *__restrict__ lim, uint8_t *__restrict__ smooth,
                             int maxlevel) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float *o = out + (size_t)i * MLP_NO;
  const int lvl = cell_level(info[i]);
  const int c = (o[0] >= o[1] && o[0] >= o[2]) ? 0 : (o[1] >= o[2] ? 1 : 2);
  int t = (c == 0) ? lvl + 1 : (c == 2 ? lvl - 1 : lvl);
  if (t < 1) t = 1;
  if (t > maxlevel) t = maxlevel;   // the network refines no further than the run's finest level
  const int plain = split[i];
  if (plain >= lvl && plain > t) t = plain;
  split[i] = t;
  const uint8_t s = o[3] > 0.f ? 1 : 0;
  lim[i] = s;
  smooth[i] = s;
}

// ---------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------
static inline dim3 g1(long n, int b = 256) { return dim3((unsigned)((n + b - 1) / b)); }

static int ensure_batch(fv_mlp *m, int B) {
  if (B <= m->bcap) return FV_OK;
  float **fl[] = {&m->bx, &m->bs, &m->z1, &m->h1, &m->z2, &m->h2, &m->o, &m->d3, &m->d2, &m->d1, &m->lossv, &m->wpart};
  for (auto p : fl) { HIPCHK(hipFree(*p)); *p = nullptr; }
  HIPCHK(hipFree(m->bc)); HIPCHK(hipFree(m->bidx));
  HIPCHK(hipMalloc(&m->bx, sizeof(float) * B * MLP_K));
  HIPCHK(hipMalloc(&m->bs, sizeof(float) * B));
  HIPCHK(hipMalloc(&m->z1, sizeof(float) * B * MLP_H));
  HIPCHK(hipMalloc(&m->h1, sizeof(float) * B * MLP_H));
  HIPCHK(hipMalloc(&m->z2, sizeof(float) * B * MLP_H));
  HIPCHK(hipMalloc(&m->h2, sizeof(float) * B * MLP_H));
  HIPCHK(hipMalloc(&m->o, sizeof(float) * B * MLP_NO));
  HIPCHK(hipMalloc(&m->d3, sizeof(float) * B * MLP_NO));
  HIPCHK(hipMalloc(&m->d2, sizeof(float) * B * MLP_H));
  HIPCHK(hipMalloc(&m->d1, sizeof(float) * B * MLP_H));
  HIPCHK(hipMalloc(&m->lossv, sizeof(float) * B));
  HIPCHK(hipMalloc(&m->wpart, sizeof(float) * (size_t)((B + MLP_WCHUNK - 1) / MLP_WCHUNK) * (MLP_K + 1) * MLP_H));
  HIPCHK(hipMalloc(&m->bc, sizeof(int) * B));
  HIPCHK(hipMalloc(&m->bidx, sizeof(int) * B));
  m->bcap = B;
  return FV_OK;
}

static void wgrad(fv_mlp *m, int B, int K, int N, const float *X, int ldx, const float *dZ, float *dW, float *db) {
  const long per = (long)(K + 1) * N;
  const long nch = (B + MLP_WCHUNK - 1) / MLP_WCHUNK;
  hipLaunchKernelGGL(k_wgrad_part, g1(per * nch), dim3(256), 0, 0, B, K, N, X, ldx, dZ, m->wpart);
  hipLaunchKernelGGL(k_wgrad_sum, g1(per), dim3(256), 0, 0, B, K, N, m->wpart, dW, db);
}

// forward + backward on the batch already in bx/bc/bs; gradient in m->g, loss summed to *loss
static int fwd_bwd(fv_mlp *m, int B, float *loss, bool backward) {
  float *W1 = m->w, *B1 = m->w + MLP_OFF_B1, *W2 = m->w + MLP_OFF_W2, *B2 = m->w + MLP_OFF_B2;
  float *W3 = m->w + MLP_OFF_W3, *B3 = m->w + MLP_OFF_B3;
  hipLaunchKernelGGL(k_fwd, g1((long)B * MLP_H), dim3(256), 0, 0, B, MLP_K, MLP_H, m->bx, MLP_K, W1, B1, m->z1, m->h1, 1);
  hipLaunchKernelGGL(k_fwd, g1((long)B * MLP_H), dim3(256), 0, 0, B, MLP_H, MLP_H, m->h1, MLP_H, W2, B2, m->z2, m->h2, 1);
  hipLaunchKernelGGL(k_fwd, g1((long)B * MLP_NO), dim3(256), 0, 0, B, MLP_H, MLP_NO, m->h2, MLP_H, W3, B3, m->o, m->o, 0);
  hipLaunchKernelGGL(k_loss, g1(B), dim3(256), 0, 0, B, m->o, m->bc, m->bs, m->cw[0], m->cw[1], m->cw[2], m->d3, m->lossv);
  if (backward) {
    float *G = m->g;
    wgrad(m, B, MLP_H, MLP_NO, m->h2, MLP_H, m->d3, G + MLP_OFF_W3, G + MLP_OFF_B3);
    hipLaunchKernelGGL(k_bwd_data, g1((long)B * MLP_H), dim3(256), 0, 0, B, MLP_H, MLP_NO, m->d3, W3, m->z2, m->d2);
    wgrad(m, B, MLP_H, MLP_H, m->h1, MLP_H, m->d2, G + MLP_OFF_W2, G + MLP_OFF_B2);
    hipLaunchKernelGGL(k_bwd_data, g1((long)B * MLP_H), dim3(256), 0, 0, B, MLP_H, MLP_H, m->d2, W2, m->z1, m->d1);
    wgrad(m, B, MLP_K, MLP_H, m->bx, MLP_K, m->d1, G, G + MLP_OFF_B1);
  }
  HIPCHK(hipGetLastError());
  if (loss) {
    std::vector<float> lv(B);
    HIPCHK(hipMemcpy(lv.data(), m->lossv, sizeof(float) * B, hipMemcpyDeviceToHost));
    double s = 0.0;
    for (int b = 0; b < B; b++) s += lv[b];
    *loss = (float)(s / B);
  }
  return FV_OK;
}

static uint64_t sm64(uint64_t *s) {
  uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

extern "C" int fv_mlp_create(fv_mlp **out, unsigned seed) {
  if (!out) return fv_set_err(FV_ERR_ARG, "null argument");
  fv_mlp *m = new fv_mlp();
  std::vector<float> w(MLP_NW, 0.f);
  uint64_t st = seed ? seed : 1;
  auto uni = [&]() { return (float)((double)(sm64(&st) >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0); };
  const float a1 = sqrtf(6.f / (MLP_NIN + MLP_H)), a2 = sqrtf(6.f / (2 * MLP_H)), a3 = sqrtf(6.f / (MLP_H + MLP_NO));
  for (int k = 0; k < MLP_NIN; k++)
    for (int j = 0; j < MLP_H; j++) w[k * MLP_H + j] = a1 * uni();
  for (int k = 0; k < MLP_H * MLP_H; k++) w[MLP_OFF_W2 + k] = a2 * uni();
  for (int k = 0; k < MLP_H * MLP_NO; k++) w[MLP_OFF_W3 + k] = a3 * uni();
  HIPCHK(hipMalloc(&m->w, sizeof(float) * MLP_NW));
  HIPCHK(hipMalloc(&m->g, sizeof(float) * MLP_NW));
  HIPCHK(hipMalloc(&m->m, sizeof(float) * MLP_NW));
  HIPCHK(hipMalloc(&m->v, sizeof(float) * MLP_NW));
  HIPCHK(hipMalloc(&m->w1t, sizeof(uint16_t) * MLP_H * MLP_K));
  HIPCHK(hipMalloc(&m->w2t, sizeof(uint16_t) * MLP_H * MLP_H));
  HIPCHK(hipMalloc(&m->w3t, sizeof(uint16_t) * 16 * MLP_H));
  HIPCHK(hipMemcpy(m->w, w.data(), sizeof(float) * MLP_NW, hipMemcpyHostToDevice));
  HIPCHK(hipMemset(m->g, 0, sizeof(float) * MLP_NW));
  HIPCHK(hipMemset(m->m, 0, sizeof(float) * MLP_NW));
  HIPCHK(hipMemset(m->v, 0, sizeof(float) * MLP_NW));
  *out = m;
  return FV_OK;
}

extern "C" int fv_mlp_destroy(fv_mlp *m) {
  if (!m) return FV_OK;
  void *ps[] = {m->w, m->g, m->m, m->v, m->w1t, m->w2t, m->w3t, m->X, m->smooth, m->cls, m->bx, m->bs, m->z1, m->h1,
                m->z2, m->h2, m->o, m->d3, m->d2, m->d1, m->lossv, m->bc, m->bidx, m->ix, m->iout, m->ismooth, m->fld, m->snd,
                m->wpart};
  for (void *p : ps) (void)hipFree(p);
  delete m;
  return FV_OK;
}

extern "C" int fv_mlp_num_weights(void) { return MLP_NW; }

extern "C" int fv_mlp_get_weights(fv_mlp *m, float *w) {
  if (!m || !w) return fv_set_err(FV_ERR_ARG, "null argument");
  HIPCHK(hipMemcpy(w, m->w, sizeof(float) * MLP_NW, hipMemcpyDeviceToHost));
  return FV_OK;
}

extern "C" int fv_mlp_set_weights(fv_mlp *m, const float *w) {
  if (!m || !w) return fv_set_err(FV_ERR_ARG, "null argument");
  HIPCHK(hipMemcpy(m->w, w, sizeof(float) * MLP_NW, hipMemcpyHostToDevice));
  m->bf16_valid = 0;
  return FV_OK;
}

extern "C" int fv_mlp_set_options(fv_mlp *m, const float *class_weights, float smooth_thr) {
  if (!m) return fv_set_err(FV_ERR_ARG, "null argument");
  if (class_weights) for (int c = 0; c < 3; c++) m->cw[c] = class_weights[c];
  if (smooth_thr > 0.f) m->smooth_thr = smooth_thr;
  return FV_OK;
}

static int ensure_data(fv_mlp *m, long need) {
  if (need <= m->cap) return FV_OK;
  long nc = need + need / 2 + 1024;
  float *X = nullptr, *sm = nullptr;
  int *cl = nullptr;
  HIPCHK(hipMalloc(&X, sizeof(float) * nc * MLP_K));
  HIPCHK(hipMalloc(&sm, sizeof(float) * nc));
  HIPCHK(hipMalloc(&cl, sizeof(int) * nc));
  if (m->nd > 0) {
    HIPCHK(hipMemcpy(X, m->X, sizeof(float) * m->nd * MLP_K, hipMemcpyDeviceToDevice));
    HIPCHK(hipMemcpy(sm, m->smooth, sizeof(float) * m->nd, hipMemcpyDeviceToDevice));
    HIPCHK(hipMemcpy(cl, m->cls, sizeof(int) * m->nd, hipMemcpyDeviceToDevice));
  }
  HIPCHK(hipFree(m->X)); HIPCHK(hipFree(m->smooth)); HIPCHK(hipFree(m->cls));
  m->X = X; m->smooth = sm; m->cls = cl; m->cap = nc;
  return FV_OK;
}

static int ensure_fields(fv_mlp *m, long n) {
  if (n <= m->fcap) return FV_OK;
  HIPCHK(hipFree(m->fld)); HIPCHK(hipFree(m->snd));
  m->fld = nullptr; m->snd = nullptr; m->fcap = 0;
  HIPCHK(hipMalloc(&m->fld, sizeof(float4) * n));
  HIPCHK(hipMalloc(&m->snd, sizeof(float) * n));
  m->fcap = n;
  return FV_OK;
}

// stencil features of every leaf of ctx into rows row0.. of X, on stream st
static int launch_features(fv_dev *c, fv_mlp *m, float *X, long row0, hipStream_t st) {
  int rc = ensure_fields(m, c->n);
  if (rc != FV_OK) return rc;
  if (c->nmat == 2)
    hipLaunchKernelGGL(k_mlp_fields<2>, g1(c->n), dim3(256), 0, st, c->n, c->P, c->S, m->fld, m->snd);
  else
    hipLaunchKernelGGL(k_mlp_fields<3>, g1(c->n), dim3(256), 0, st, c->n, c->P, c->S, m->fld, m->snd);
  hipLaunchKernelGGL(k_mlp_features, g1(c->n), dim3(256), 0, st, c->n, m->fld, m->snd, c->nbr, c->info, c->fine4, X, row0);
  HIPCHK(hipGetLastError());
  return FV_OK;
}

// append the (features, labels) of every leaf of ctx, labels from the plain indicator
extern "C" int fv_mlp_sample(fv_dev *c, fv_mlp *m, const fv_crit *cr) {
  FV_GUARD(c);
  if (!c || !m || !cr) return fv_set_err(FV_ERR_ARG, "null argument");
  int *SIdev = nullptr;
  int rc = fv_dev_indicator_raw(c, cr, &SIdev);
  if (rc != FV_OK) return rc;
  HIPCHK(hipStreamSynchronize(c->stream));
  rc = ensure_data(m, m->nd + c->n);
  if (rc != FV_OK) return rc;
  rc = launch_features(c, m, m->X, m->nd, 0);
  if (rc != FV_OK) return rc;
  if (c->nmat == 2)
    hipLaunchKernelGGL(k_mlp_labels<2>, g1(c->n), dim3(256), 0, 0, c->n, c->P, c->S, c->nbr, c->info, c->fine4, c->split,
                       m->smooth_thr, m->cls, m->smooth, m->nd);
  else
    hipLaunchKernelGGL(k_mlp_labels<3>, g1(c->n), dim3(256), 0, 0, c->n, c->P, c->S, c->nbr, c->info, c->fine4, c->split,
                       m->smooth_thr, m->cls, m->smooth, m->nd);
  HIPCHK(hipGetLastError());
  HIPCHK(hipDeviceSynchronize());
  m->nd += c->n;
  return FV_OK;
}

extern "C" int fv_mlp_dataset(fv_mlp *m, long *n, float *X, int *cls, float *smooth) {
  if (!m) return fv_set_err(FV_ERR_ARG, "null argument");
  if (n) *n = m->nd;
  if (X) HIPCHK(hipMemcpy(X, m->X, sizeof(float) * m->nd * MLP_K, hipMemcpyDeviceToHost));
  if (cls) HIPCHK(hipMemcpy(cls, m->cls, sizeof(int) * m->nd, hipMemcpyDeviceToHost));
  if (smooth) HIPCHK(hipMemcpy(smooth, m->smooth, sizeof(float) * m->nd, hipMemcpyDeviceToHost));
  return FV_OK;
}

extern "C" int fv_mlp_clear_dataset(fv_mlp *m) {
  if (!m) return fv_set_err(FV_ERR_ARG, "null argument");
  m->nd = 0;
  return FV_OK;
}

// Adam on the device dataset, deterministic order from seed; mean batch loss of the last epoch
extern "C" int fv_mlp_train(fv_mlp *m, int epochs, int batch, float lr, unsigned seed, float *loss) {
  if (!m || epochs < 1 || batch < 1 || m->nd < batch) return fv_set_err(FV_ERR_ARG, "bad training arguments");
  int rc = ensure_batch(m, batch);
  if (rc != FV_OK) return rc;
  std::vector<int> perm(m->nd);
  uint64_t st = seed ? seed : 7;
  double tot = 0.0;
  long nb = 0;
  for (int e = 0; e < epochs; e++) {
    for (long i = 0; i < m->nd; i++) perm[i] = (int)i;
    for (long i = m->nd - 1; i > 0; i--) {
      long j = (long)(sm64(&st) % (uint64_t)(i + 1));
      int t = perm[i]; perm[i] = perm[j]; perm[j] = t;
    }
    tot = 0.0;
    nb = 0;
    for (long s0 = 0; s0 + batch <= m->nd; s0 += batch) {
      HIPCHK(hipMemcpy(m->bidx, perm.data() + s0, sizeof(int) * batch, hipMemcpyHostToDevice));
      hipLaunchKernelGGL(k_gather_rows, g1((long)batch * MLP_K), dim3(256), 0, 0, batch, m->bidx, m->X, m->cls, m->smooth,
                         m->bx, m->bc, m->bs);
      float l = 0.f;
      rc = fwd_bwd(m, batch, (e == epochs - 1) ? &l : nullptr, true);
      if (rc != FV_OK) return rc;
      m->adam_t++;
      const float b1 = 0.9f, b2 = 0.999f;
      const float c1 = 1.f - powf(b1, (float)m->adam_t), c2 = 1.f - powf(b2, (float)m->adam_t);
      hipLaunchKernelGGL(k_adam, g1(MLP_NW), dim3(256), 0, 0, MLP_NW, m->w, m->g, m->m, m->v, lr, b1, b2, 1e-8f, c1, c2);
      tot += l;
      nb++;
    }
  }
  HIPCHK(hipDeviceSynchronize());
  m->bf16_valid = 0;
  if (loss) *loss = (float)(tot / (double)(nb > 0 ? nb : 1));
  return FV_OK;
}

// gradient and mean loss of a host batch (for checks against the CPU oracle)
extern "C" int fv_mlp_grad(fv_mlp *m, const float *X, const int *cls, const float *smooth, int B, float *grad,
                           float *loss) {
  if (!m || !X || !cls || !smooth || B < 1) return fv_set_err(FV_ERR_ARG, "bad argument");
  int rc = ensure_batch(m, B);
  if (rc != FV_OK) return rc;
  HIPCHK(hipMemcpy(m->bx, X, sizeof(float) * B * MLP_K, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(m->bc, cls, sizeof(int) * B, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(m->bs, smooth, sizeof(float) * B, hipMemcpyHostToDevice));
  rc = fwd_bwd(m, B, loss, true);
  if (rc != FV_OK) return rc;
  if (grad) HIPCHK(hipMemcpy(grad, m->g, sizeof(float) * MLP_NW, hipMemcpyDeviceToHost));
  return FV_OK;
}

// FP32 forward of a host batch: logits [B][4]
extern "C" int fv_mlp_forward(fv_mlp *m, const float *X, int B, float *out) {
  if (!m || !X || !out || B < 1) return fv_set_err(FV_ERR_ARG, "bad argument");
  int rc = ensure_batch(m, B);
  if (rc != FV_OK) return rc;
  HIPCHK(hipMemcpy(m->bx, X, sizeof(float) * B * MLP_K, hipMemcpyHostToDevice));
  HIPCHK(hipMemset(m->bc, 0, sizeof(int) * B));
  HIPCHK(hipMemset(m->bs, 0, sizeof(float) * B));
  rc = fwd_bwd(m, B, nullptr, false);
  if (rc != FV_OK) return rc;
  HIPCHK(hipMemcpy(out, m->o, sizeof(float) * B * MLP_NO, hipMemcpyDeviceToHost));
  return FV_OK;
}

static int refresh_bf16(fv_mlp *m) {
  if (m->bf16_valid) return FV_OK;
  hipLaunchKernelGGL(k_to_bf16_t, g1(MLP_K * MLP_H), dim3(256), 0, 0, MLP_K, MLP_H, MLP_H, m->w, m->w1t);
  hipLaunchKernelGGL(k_to_bf16_t, g1(MLP_H * MLP_H), dim3(256), 0, 0, MLP_H, MLP_H, MLP_H, m->w + MLP_OFF_W2, m->w2t);
  hipLaunchKernelGGL(k_to_bf16_t, g1(MLP_H * 16), dim3(256), 0, 0, MLP_H, MLP_NO, 16, m->w + MLP_OFF_W3, m->w3t);
  HIPCHK(hipGetLastError());
  HIPCHK(hipDeviceSynchronize());   // the copies are read from the solver's own (non-blocking) stream
  m->bf16_valid = 1;
  return FV_OK;
}

// outputs of n rows; with_x also the n x 112 input rows of a host batch
static int ensure_infer(fv_mlp *m, long n, bool with_x) {
  if (n > m->icap) {
    HIPCHK(hipFree(m->iout)); HIPCHK(hipFree(m->ismooth));
    m->iout = nullptr; m->ismooth = nullptr; m->icap = 0;
    HIPCHK(hipMalloc(&m->iout, sizeof(float) * n * MLP_NO));
    HIPCHK(hipMalloc(&m->ismooth, n));
    m->icap = n;
  }
  if (with_x && n > m->xcap) {
    HIPCHK(hipFree(m->ix));
    m->ix = nullptr; m->xcap = 0;
    HIPCHK(hipMalloc(&m->ix, sizeof(float) * n * MLP_K));
    m->xcap = n;
  }
  return FV_OK;
}

// BF16 MFMA forward of a host batch: logits [B][4]
extern "C" int fv_mlp_infer(fv_mlp *m, const float *X, int B, float *out) {
  if (!m || !X || !out || B < 1) return fv_set_err(FV_ERR_ARG, "bad argument");
  int rc = refresh_bf16(m);
  if (rc == FV_OK) rc = ensure_infer(m, B, true);
  if (rc != FV_OK) return rc;
  HIPCHK(hipMemcpy(m->ix, X, sizeof(float) * B * MLP_K, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_mlp_infer, dim3((unsigned)((B + 63) / 64)), dim3(256), 0, 0, (long)B, m->ix, m->w1t, m->w2t, m->w3t,
                     m->w, m->iout);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpy(out, m->iout, sizeof(float) * B * MLP_NO, hipMemcpyDeviceToHost));
  return FV_OK;
}

// Learned indicator on the current state of ctx: interface layers and plain
// targets as fv_dev_indicator, then the network's targets with the plain
// refinement forced; the smooth flag becomes the per-leaf limiter.
extern "C" int fv_dev_indicator_mlp(fv_dev *c, fv_mlp *m, const fv_crit *cr, int *SI, int *split, uint8_t *smooth) {
  FV_GUARD(c);
  if (!c || !m || !cr) return fv_set_err(FV_ERR_ARG, "null argument");
  int *SIdev = nullptr;
  int rc = fv_dev_indicator_raw(c, cr, &SIdev);
  if (rc == FV_OK) rc = refresh_bf16(m);
  if (rc == FV_OK) rc = ensure_infer(m, c->n, false);
  if (rc != FV_OK) return rc;
  rc = ensure_fields(m, c->n);
  if (rc != FV_OK) return rc;
  if (c->nmat == 2)
    hipLaunchKernelGGL(k_mlp_fields<2>, g1(c->n), dim3(256), 0, c->stream, c->n, c->P, c->S, m->fld, m->snd);
  else
    hipLaunchKernelGGL(k_mlp_fields<3>, g1(c->n), dim3(256), 0, c->stream, c->n, c->P, c->S, m->fld, m->snd);
  hipLaunchKernelGGL(k_mlp_infer_stencil, dim3((unsigned)((c->n + 63) / 64)), dim3(256), 0, c->stream, c->n, m->fld,
                     m->snd, c->nbr, c->info, c->fine4, m->w1t, m->w2t, m->w3t, m->w, m->iout);
  hipLaunchKernelGGL(k_mlp_decide, g1(c->n), dim3(256), 0, c->stream, c->n, m->iout, c->info, c->split, c->lim, m->ismooth,
                     FV_MAXLEVEL);
  HIPCHK(hipGetLastError());
  if (SI && (rc = fv_copy_to_host(c, SI, SIdev, sizeof(int) * 2 * (size_t)c->n)) != FV_OK) return rc;
  if (split && (rc = fv_copy_to_host(c, split, c->split, sizeof(int) * (size_t)c->n)) != FV_OK) return rc;
  if (smooth && (rc = fv_copy_to_host(c, smooth, m->ismooth, (size_t)c->n)) != FV_OK) return rc;
  HIPCHK(hipStreamSynchronize(c->stream));
  return FV_OK;
}

// ---------------------------------------------------------------------
// model-5 contexts (config 2): the same network on the model-5 state; the
// fields and labels come from fv_m5.hip, the indicator kernels run on the
// stream of the context's topology (fv_m5_dev_set_forest)
// ---------------------------------------------------------------------

// the smooth flags alone (the targets were used already): the per-leaf limiter
__global__ void k_mlp_flags(int n, const float *__restrict__ out, uint8_t *__restrict__ lim,
                            uint8_t *__restrict__ smooth) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint8_t s = out[(size_t)i * MLP_NO + 3] > 0.f ? 1 : 0;
  lim[i] = s;
  smooth[i] = s;
}

// network outputs of every leaf of the model-5 context into m->iout
static int m5_infer(fv_m5_dev *c, fv_mlp *m, fv_m5_view &v) {
  int rc = fv_m5_view_of(c, &v);
  if (rc == FV_OK) rc = refresh_bf16(m);
  if (rc == FV_OK) rc = ensure_infer(m, v.n, false);
  if (rc == FV_OK) rc = ensure_fields(m, v.n);
  if (rc == FV_OK) rc = fv_m5_mlp_fields(c, m->fld, m->snd);
  if (rc != FV_OK) return rc;
  hipLaunchKernelGGL(k_mlp_infer_stencil, dim3((unsigned)((v.n + 63) / 64)), dim3(256), 0, v.stream, v.n, m->fld,
                     m->snd, v.nbr, v.info, v.fine4, m->w1t, m->w2t, m->w3t, m->w, m->iout);
  HIPCHK(hipGetLastError());
  return FV_OK;
}

extern "C" int fv_mlp_sample_m5(fv_m5_dev *c, fv_mlp *m, const fv_crit *cr) {
  if (!c || !m || !cr) return fv_set_err(FV_ERR_ARG, "null argument");
  int *SI = nullptr, *split = nullptr;
  int rc = fv_m5_indicator_raw(c, cr, &SI, &split);
  fv_m5_view v;
  if (rc == FV_OK) rc = fv_m5_view_of(c, &v);
  if (rc != FV_OK) return rc;
  HIPCHK(hipStreamSynchronize(v.stream));
  if ((rc = ensure_data(m, m->nd + v.n)) != FV_OK) return rc;
  if ((rc = ensure_fields(m, v.n)) != FV_OK) return rc;
  if ((rc = fv_m5_mlp_fields(c, m->fld, m->snd)) != FV_OK) return rc;
  hipLaunchKernelGGL(k_mlp_features, g1(v.n), dim3(256), 0, v.stream, v.n, m->fld, m->snd, v.nbr, v.info, v.fine4,
                     m->X, m->nd);
  HIPCHK(hipGetLastError());
  if ((rc = fv_m5_mlp_labels(c, split, m->smooth_thr, m->cls, m->smooth, m->nd)) != FV_OK) return rc;
  HIPCHK(hipStreamSynchronize(v.stream));
  m->nd += v.n;
  return FV_OK;
}

// the learned indicator on a model-5 context: interface layers and plain
// targets, then the network's targets with the plain refinement forced (left in
// the topology context, where fv_m5_dev_regrid_mlp reads them) and the smooth
// flag as the per-leaf limiter
extern "C" int fv_m5_dev_indicator_mlp(fv_m5_dev *c, fv_mlp *m, const fv_crit *cr, int *SI, int *split,
                                       uint8_t *smooth) {
  if (!c || !m || !cr) return fv_set_err(FV_ERR_ARG, "null argument");
  int *SIdev = nullptr, *sp = nullptr;
  int rc = fv_m5_indicator_raw(c, cr, &SIdev, &sp);
  fv_m5_view v;
  if (rc == FV_OK) rc = m5_infer(c, m, v);
  if (rc != FV_OK) return rc;
  hipLaunchKernelGGL(k_mlp_decide, g1(v.n), dim3(256), 0, v.stream, v.n, m->iout, v.info, sp, v.lim, m->ismooth,
                     FV_MAXLEVEL);
  HIPCHK(hipGetLastError());
  if (SI) HIPCHK(hipMemcpyAsync(SI, SIdev, sizeof(int) * 2 * (size_t)v.n, hipMemcpyDeviceToHost, v.stream));
  if (split) HIPCHK(hipMemcpyAsync(split, sp, sizeof(int) * (size_t)v.n, hipMemcpyDeviceToHost, v.stream));
  if (smooth) HIPCHK(hipMemcpyAsync(smooth, m->ismooth, (size_t)v.n, hipMemcpyDeviceToHost, v.stream));
  HIPCHK(hipStreamSynchronize(v.stream));
  return FV_OK;
}

// meshadaptation with the learned targets: the learned indicator, the regrid
// (criterionsmth, merge/split with the model-5 remap, connectivity), then the
// network's smooth flags of the new leaves as their limiter
extern "C" int fv_m5_dev_regrid_mlp(fv_m5_dev *c, fv_mlp *m, const fv_crit *cr, int maxlevel, int *n_out) {
  if (!c || !m || !cr) return fv_set_err(FV_ERR_ARG, "null argument");
  int *SIdev = nullptr, *sp = nullptr;
  int rc = fv_m5_indicator_raw(c, cr, &SIdev, &sp);
  fv_m5_view v;
  if (rc == FV_OK) rc = m5_infer(c, m, v);
  if (rc != FV_OK) return rc;
  hipLaunchKernelGGL(k_mlp_decide, g1(v.n), dim3(256), 0, v.stream, v.n, m->iout, v.info, sp, v.lim, m->ismooth,
                     maxlevel);
  HIPCHK(hipGetLastError());
  if ((rc = fv_m5_regrid_targets(c, sp, SIdev, maxlevel, n_out)) != FV_OK) return rc;
  if ((rc = m5_infer(c, m, v)) != FV_OK) return rc;
  hipLaunchKernelGGL(k_mlp_flags, g1(v.n), dim3(256), 0, v.stream, v.n, m->iout, v.lim, m->ismooth);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(v.stream));
  return FV_OK;
}
