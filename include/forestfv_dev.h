/*
 * forestfv_dev.h -- C ABI of the MI355X device path.
 *
 * The reference has no plugin or operator registry: its explicit update is
 * reached through plain C calls, loop() -> advancetimeexplicit(run)
 * (src/c/loop.c:21), and inside it grad_scheme(run), RHS(run), RK(irk,run),
 * m4_relaxation_fluid(run) and communicate_S(run,0) per stage
 * (src/c/advancetimeexplicit.c:9-24).  These entry points take their place:
 * a caller packs its leaves once per mesh change (fv_dev_attach), then runs
 * whole steps on the device (fv_dev_advance) and copies the state back only
 * before export, restart or regrid (fv_dev_get_state).  Plain pointers and
 * sizes only; every call returns an FV_* status instead of exit(0).
 *
 * One context per RUN, one HIP stream per context, no concurrent calls on a
 * context (the reference is single-threaded per rank and non-reentrant,
 * strdata.h:185-217).
 */
#ifndef FORESTFV_DEV_H
#define FORESTFV_DEV_H

#include "forestfv.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fv_dev fv_dev;

/* Precision of the reconstruction / flux / ghost-fill kernels.  The state,
 * its RK accumulation and pressure recovery from energy stay FP64 in all.
 *  FV_PREC_FP64  (default) plain FP64, reciprocal-multiply form of the
 *                reference arithmetic; within 1e-10 relative L1 of the CPU
 *                path after 100 steps;
 *  FV_PREC_MIXED FP32 slopes, face records and HLLC;
 *  FV_PREC_EXACT FP64 in the reference's operation order, bit-identical to
 *                the CPU restatement. */
#define FV_PREC_FP64  0
#define FV_PREC_MIXED 1
#define FV_PREC_EXACT 2

/* Upload the packed mesh and parameters; allocates every device buffer of
 * the context (the device mirror of leafallocation.c / createutility.c). */
int  fv_dev_attach(fv_dev **ctx, const fv_params *p, const fv_pack *pk);
/* Release the context (leafdeallocation.c equivalent). */
int  fv_dev_detach(fv_dev *ctx);
/* Re-upload topology after a regrid, keeping parameters; state is resized. */
int  fv_dev_reattach(fv_dev *ctx, const fv_pack *pk);

/* State transfer, S as [n][NEQ] in leaf order (LEAF.S). */
int  fv_dev_set_state(fv_dev *ctx, const double *S);
int  fv_dev_get_state(fv_dev *ctx, double *S);

/* advancetimeexplicit (advancetimeexplicit.c:3-29) repeated nsteps times. */
int  fv_dev_advance(fv_dev *ctx, int nsteps);
/* RK(0,run): SN = S (RK.c:17-21). */
int  fv_dev_begin_step(fv_dev *ctx);
/* One stage irk: grad_scheme + RHS + RK(irk) + m4_relaxation_fluid. */
int  fv_dev_stage(fv_dev *ctx, int irk);

/* DT used by the stages (the reference's fixed DT, inputset.c:69). */
int  fv_dev_set_dt(fv_dev *ctx, double dt);
/* Largest stable step cfl * min h / (max|u_i| + c) (new: no CFL in the reference). */
int  fv_dev_cfl_dt(fv_dev *ctx, double cfl, double *dt);
/* FV_PREC_FP64 (default), FV_PREC_MIXED or FV_PREC_EXACT. */
int  fv_dev_set_precision(fv_dev *ctx, int prec);
/* FV_PREC_FP64 / FV_PREC_MIXED stages: 0 (default) reconstruction and flux
 * kernels with the face states stored in between; 1 one fused kernel per stage
 * that keeps no face states and recomputes each neighbour's face state.  Both
 * give the same state up to FMA contraction.  FV_PREC_EXACT always stores the
 * face states. */
int  fv_dev_set_fused(fv_dev *ctx, int on);

/* FV_PREC_FP64 / FV_PREC_MIXED stages with stored face states (fused off):
 * with on = 1 every 4x4x4 block of 64 consecutive same-level leaves whose
 * surface faces have same-level neighbours or boundaries runs the LDS block
 * kernel (each face evaluated once, no face states in HBM), the other leaves
 * the per-leaf kernels; on = 2 (the default) also takes the blocks with 2:1
 * or 1:4 faces on their surface whose neighbours across those faces are
 * outside blocks ("graded" blocks, evaluated as the per-leaf flux evaluates
 * those faces); 0 every leaf runs the per-leaf kernels.  All give the same
 * state up to rounding (within 1e-13 relative per stage).
 * fv_dev_brick_count: the blocks of the current mesh and the leaves outside
 * them; fv_dev_brick_kinds: how many of the blocks are plain and graded (no
 * counterpart in the reference). */
int  fv_dev_set_bricks(fv_dev *ctx, int on);
int  fv_dev_brick_count(fv_dev *ctx, int *nbricks, int *nrest);
int  fv_dev_brick_kinds(fv_dev *ctx, int *nplain, int *ngraded);

/* Face states of the last stage, SF as [n][6][NEQ] (LEAF.SF); written by the
 * stages of FV_PREC_EXACT and of the unfused throughput path with the block
 * kernel off (fv_dev_set_bricks(ctx, 0)). */
int  fv_dev_get_faces(fv_dev *ctx, double *SF);

/* interface_location (interface_location.c:7-86) and the per-leaf target
 * level of criterionsplt (criterionsplt.c:3-140) on the current state.
 * SI: [n][2] out, split: [n] out (either may be NULL). */
int  fv_dev_indicator(fv_dev *ctx, const fv_crit *c, int *SI, int *split);

/* interface_location as the time loop calls it after every step
 * (loop.c:23, interface_location.c:7-86): recomputes SI from ctx's state and
 * makes the following stages use it (the interface bit and the computed-once
 * choice of equal-level faces, as fv_mesh_pack derives them from the mesh's
 * SI).  SI: [n][2] out, may be NULL. */
int  fv_dev_interface_location(fv_dev *ctx, const fv_crit *c, int *SI);

/* ---- device-resident mesh adaptation ----
 * meshadaptation (meshadaptation.c:3-44) without a host round trip: the
 * indicator, criterionsmth (criterionsmth.c:3-58), merge and split with the
 * state remap (merge.c:20-55, split.c:425-434), the new DFS order and the
 * connectivity and face lists of fv_mesh_pack, all on the device.  The result
 * equals fv_dev_indicator + fv_criterionsmth + fv_mesh_adapt + fv_mesh_pack.
 * fv_dev_set_forest uploads the leaf addresses (root, level, ci) and SI of the
 * mesh ctx is attached to; it is needed once after attach or reattach. */
int  fv_dev_set_forest(fv_dev *ctx, const fv_mesh *m);
/* *n_out: the new leaf count; the state stays on the device, remapped. */
int  fv_dev_regrid(fv_dev *ctx, const fv_crit *c, int maxlevel, int *n_out);
/* criterionsmth alone on the device mesh; split [n] host array, in place. */
int  fv_dev_criterionsmth(fv_dev *ctx, int LEVEL, int *split);
/* Leaf addresses of the device mesh: root [n], level [n], ci [n][3], SI [n][2]
 * (any may be NULL); fv_mesh_from_leaves rebuilds the host forest from them. */
int  fv_dev_get_forest(fv_dev *ctx, int *root, int *level, int *ci, int *SI);
/* Number of root trees of the forest given to fv_dev_set_forest. */
int  fv_dev_get_ntrees(fv_dev *ctx, int *ntrees);
/* restart_save / export_ultra_fast of the device mesh and state (forest set
 * with fv_dev_set_forest): text (binary = 0) or lossless binary (binary = 1),
 * the formats of fv_restart_write. */
int  fv_dev_write_restart(fv_dev *ctx, const char *path, int istep, double time, int binary);
/* The device mesh's fv_pack arrays: nbr [6][n], info [n], fine4 [nfine][4]
 * (any may be NULL). */
int  fv_dev_get_pack(fv_dev *ctx, int *nbr, uint32_t *info, int *nfine, int *fine4);

/* Per-leaf limiter when fv_params.limiter == 2: 0 minmod, 1 MC; written by
 * fv_dev_indicator_mlp from the learned smooth flag, or by the caller. */
int  fv_dev_set_limiter_flags(fv_dev *ctx, const uint8_t *lim);
int  fv_dev_get_limiter_flags(fv_dev *ctx, uint8_t *lim);

/* Timing of the stage kernels with hipEvents on the context's stream:
 * runs nrep full stages (cycling irk) and returns the mean milliseconds per
 * stage and per launch of the dominant kernel: the 4x4x4 block kernel where
 * the FP64 or mixed stage has blocks (fv_dev_brick_count), else the flux or
 * fused stage kernel. */
int  fv_dev_time_stages(fv_dev *ctx, int nrep, double *ms_stage, double *ms_flux_kernel);

/* Measurement utility with no counterpart in the reference: copies `bytes`
 * between two device buffers `reps` times and returns read+write GB/s (the
 * measured HBM stream bandwidth; one dispatch also calibrates byte counters). */
int  fv_dev_stream_copy(size_t bytes, int reps, double *gbps);

/* ---- learned refinement / limiter indicator (new; no reference counterpart) ----
 * A 108 -> 64 -> 64 -> 4 SiLU network on a 3x3x3 stencil of normalised
 * (rho, p, alpha, |u|); outputs refine / keep / coarsen and a smooth flag.
 * Trained on the device with FP32 forward, backward and Adam kernels against
 * the plain indicator's labels; inference is a BF16 MFMA GEMM chain.  The
 * plain indicator's refinement always wins. */
typedef struct fv_mlp fv_mlp;
int  fv_mlp_create(fv_mlp **m, unsigned seed);
int  fv_mlp_destroy(fv_mlp *m);
int  fv_mlp_num_weights(void);
/* weights as [W1 112x64 | b1 64 | W2 64x64 | b2 64 | W3 64x4 | b3 4], row-major [in][out] */
int  fv_mlp_get_weights(fv_mlp *m, float *w);
int  fv_mlp_set_weights(fv_mlp *m, const float *w);
/* class weights of refine/keep/coarsen in the loss; smooth-flag threshold on the
 * largest relative pressure jump to a face neighbour, default 0.1 (<= 0 keeps the current value) */
int  fv_mlp_set_options(fv_mlp *m, const float *class_weights, float smooth_thr);
/* append (features, labels) of every leaf of ctx; labels from the plain indicator */
int  fv_mlp_sample(fv_dev *ctx, fv_mlp *m, const fv_crit *c);
int  fv_mlp_dataset(fv_mlp *m, long *n, float *X, int *cls, float *smooth);
int  fv_mlp_clear_dataset(fv_mlp *m);
int  fv_mlp_train(fv_mlp *m, int epochs, int batch, float lr, unsigned seed, float *loss);
/* host-batch helpers: FP32 gradient and loss, FP32 forward, BF16 MFMA forward */
int  fv_mlp_grad(fv_mlp *m, const float *X, const int *cls, const float *smooth, int B, float *grad, float *loss);
int  fv_mlp_forward(fv_mlp *m, const float *X, int B, float *out);
int  fv_mlp_infer(fv_mlp *m, const float *X, int B, float *out);
/* learned indicator on ctx's state: SI [n][2], split [n], smooth [n] (any may be NULL);
 * also sets the per-leaf limiter used when fv_params.limiter == 2 */
int  fv_dev_indicator_mlp(fv_dev *ctx, fv_mlp *m, const fv_crit *c, int *SI, int *split, uint8_t *smooth);

/* Volume integral of every state field, tot [NEQ] (the conservation monitor of
 * conservationcalc.c:3-36): sums per block, then over the blocks in order, so
 * the result is the same on every run. */
int  fv_dev_totals(fv_dev *ctx, double *tot);

/* Number of leaves and NEQ of the attached mesh. */
int  fv_dev_size(const fv_dev *ctx, int *n, int *neq);

/* Allocation-free CFL step, with no counterpart in the reference: for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  It is fv_dev_cfl_dt's
 * computation, cfl * min h / (max|u_i| + c), on caller-owned device arrays:
 * S [NEQ][n] (structure of arrays, as a context holds it) and info [n] (the
 * fv_pack info words), with the geometry tables g of fv_pack.  The result is
 * written to the device double *dt, in stream order on `stream` (a
 * hipStream_t; NULL is the null stream); the call does not wait for it.  The
 * per-block partial minima live in a fixed device array of
 * FV_CFL_NOALLOC_MAX_LEAVES / 256 doubles (2 MiB per device, held until the
 * process ends), so n above FV_CFL_NOALLOC_MAX_LEAVES is rejected with
 * FV_ERR_ARG.  Calls that overlap on one device share that array and race. */
#define FV_CFL_NOALLOC_MAX_LEAVES (1 << 26)
int  fv_cfl_dt_noalloc(const fv_params *p, const fv_geo *g, int n, const double *S, const uint32_t *info,
                       double cfl, double *dt, void *stream);

/* Host-only self-check with no counterpart in the reference: a context left
 * unusable by a failed mesh install is refused by the entry points (status
 * FV_ERR_HIP, no kernel launched), and a device reservation that cannot be
 * satisfied leaves the context's buffers and size unchanged.  FV_OK when both
 * hold. */
int  fv_dev_selftest_guard(void);

/* Message of the last failing call on this thread. */
const char *fv_dev_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
