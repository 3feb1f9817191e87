// fv_internal.h -- the device context shared by the solver and MLP units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <vector>
#include "forestfv_dev.h"
#include "fv_phys.h"

extern thread_local char fv_g_err[256];
int fv_set_err(int code, const char *msg);

#define HIPCHK(x)                                                                                      \
  do {                                                                                                 \
    hipError_t e_ = (x);                                                                               \
    if (e_ != hipSuccess) {                                                                            \
      snprintf(fv_g_err, sizeof fv_g_err, "%s:%d %s: %s", __FILE__, __LINE__, #x, hipGetErrorString(e_)); \
      return FV_ERR_HIP;                                                                               \
    }                                                                                                  \
  } while (0)

struct fv_dev {
  int n = 0, neq = 0, nmat = 0, nfine = 0, prec = FV_PREC_FP64;
  int fused = 0;          // FP64 / MIXED stages: fused kernel (1) or recon + flux with stored face states (0)
  fvk::Params P{};
  fvk::Geo G{};
  fvk::Geo *Gdev = nullptr;   // device copy of G for kernels that index it per leaf
  hipStream_t stream = nullptr;
  double *S = nullptr, *S2 = nullptr, *SN = nullptr, *SF = nullptr, *part = nullptr, *dtmp = nullptr;
  float *SFm = nullptr;   // mixed-precision face records, allocated on first use
  int cap_m = 0;
  int *nbr = nullptr, *fine4 = nullptr, *nanflag = nullptr, *SIa = nullptr, *SIb = nullptr, *split = nullptr;
  int *iflag = nullptr;   // [4] queue counters of the interface-layer search
  int *bfsq = nullptr;    // [2][n] current and next queue of the interface-layer search
  double *tpart = nullptr; // per-block partial sums of fv_dev_totals
  size_t cap_tpart = 0;
  // device blocks kept between regrids for their temporaries: (pointer, bytes)
  std::vector<std::pair<void *, size_t>> pool;
  uint32_t *info = nullptr;
  uint8_t *lim = nullptr;  // per-leaf limiter (0 minmod, 1 MC) read when limiter == 2
  int cap_n = 0, cap_fine = 0;
  void *pinned = nullptr;  // host staging for device -> host copies of per-leaf outputs
  size_t pinned_cap = 0;
  // leaf addresses of the forest for device regrids (fv_dev_set_forest): root [n],
  // level [n], ci [n][3] and the mesh's interface flags SI [n][2]
  int forest = 0, cap_leaf = 0;
  int nroot[3] = {1, 1, 1}, bcside[6] = {0, 0, 0, 0, 0, 0};
  int *lroot = nullptr, *llevel = nullptr, *lci = nullptr, *lSI = nullptr;
  // set while a mesh is being installed (attach, reattach, regrid) and left set
  // when the install fails part way: the buffers and n may then disagree, so
  // every later entry point refuses the context instead of launching on it
  int broken = 0;
  // 4x4x4 same-level blocks of the mesh for the block stage kernel (FP64 and
  // mixed stages): their first leaves, the leaves outside them, a per-leaf flag
  int use_bricks = 2, brick_ok = 0, brick_cap = 0, nbricks = 0, nplain = 0, ngraded = 0, nrest = 0;
  int *bricks = nullptr, *rest = nullptr, *bmark = nullptr, *bcnt = nullptr;
  uint8_t *inbrick = nullptr, *notbrick = nullptr;
  void *bcub = nullptr;
  size_t bcub_bytes = 0;
};

// entry-point guard: a null or broken context is an error status, never a launch
#define FV_GUARD(c)                                                                                    \
  do {                                                                                                 \
    if (!(c)) return fv_set_err(FV_ERR_ARG, "null context");                                          \
    if ((c)->broken)                                                                                   \
      return fv_set_err(FV_ERR_HIP, "context unusable: an earlier mesh install failed part way");     \
  } while (0)

// capacity for n leaves and nfine 1:4 faces (reallocates, losing contents, only
// when too small); defined in fv_device.hip
int fv_dev_reserve(fv_dev *c, int n, int nfine);

// device -> host copy through the context's pinned staging buffer (on its stream)
int fv_copy_to_host(fv_dev *c, void *dst, const void *src, size_t bytes);

// launches the indicator kernels (interface layers, target level) and leaves
// SI in *SIdev and the targets in c->split; defined in fv_device.hip
int fv_dev_indicator_raw(fv_dev *c, const fv_crit *cr, int **SIdev);

// layers 2 .. ADAPT_INTER_MAX - 1 of the interface flags from the first layer in
// c->SIa; *out: the device buffer ([n][2]) that holds them all (fv_device.hip)
int fv_dev_si_layers(fv_dev *c, const fv_crit *cr, int **out);

// the face kinds of info rewritten for the interface flags SI (the mirror rule of
// m4_residual.c:62-85 and m5_residual.c:62-103), on stream st
void fv_dev_refresh_kinds_launch(hipStream_t st, int n, const int *nbr, const int *SI, uint32_t *info);

// a context with the topology, forest and indicator buffers of pk and no state
// (neq = 0); freed with fv_dev_detach
int fv_dev_attach_topology(fv_dev **out, const fv_pack *pk, int nmat);

// the state side of a regrid run by a topology context (fv_dev_regrid_hook):
// remap launches, on stream st, the state of the new leaves from the merge/split
// plan of the n old leaves (kind: 0 kept, 1 merged -- the family's first kid
// writes the parent --, 2 split into eight; off: the first new leaf; ci: the old
// leaves' octant addresses); install takes the new state once the new topology
// is in the topology context c
struct fv_state_hook {
  void *ctx;
  int (*remap)(void *ctx, hipStream_t st, int n, int nn, const uint8_t *kind, const int *off, const int *ci);
  int (*install)(void *ctx, fv_dev *c);
};

// criterionsmth of the targets split (device [n], changed in place), merge/split
// with the hook's state remap, connectivity and install; SI (device [n][2]) are
// the interface flags that become the mesh's (fv_regrid.hip)
int fv_dev_regrid_hook(fv_dev *c, int *split, const int *SI, int maxlevel, const fv_state_hook *h, int *n_out);

// ---- model-5 contexts seen by the learned indicator (fv_mlp.hip) ----
struct fv_m5_dev;
struct fv_m5_view {
  int n;
  const int *nbr;
  const uint32_t *info;
  const int *fine4;
  uint8_t *lim;        // per-leaf slope flags of limiter 2
  hipStream_t stream;  // the topology context's stream, where the indicator kernels run
};
int fv_m5_view_of(fv_m5_dev *c, fv_m5_view *v);
// the model-5 indicator of the device state: interface flags (*SI) and plain
// targets (*split), device buffers of the topology context (fv_m5_dev_set_forest)
int fv_m5_indicator_raw(fv_m5_dev *c, const fv_crit *cr, int **SI, int **split);
// per-leaf fields of the learned indicator from the model-5 state: (rho, mixture
// pressure, liquid fraction alpha_1, |u|) and the longitudinal sound speed
int fv_m5_mlp_fields(fv_m5_dev *c, float4 *fld, float *snd);
// labels of rows row0.. from the targets split: class (0 refine, 1 keep, 2
// coarsen) and the smooth flag (largest relative pressure jump below thr)
int fv_m5_mlp_labels(fv_m5_dev *c, const int *split, float thr, int *cls, float *smooth, long row0);
// regrid with the targets split (device [n], changed in place) and interface flags SI
int fv_m5_regrid_targets(fv_m5_dev *c, int *split, const int *SI, int maxlevel, int *n_out);
