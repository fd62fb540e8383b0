/*
 * mi3hip.h -- C ABI of libmi3hip.so, the MI355X (gfx950) engine behind the
 * Mi3 inference round.
 *
 * The reference reaches its device code through pyopencl kernel enqueues
 * issued by the per-device wrapper class MCMCGPU (mi3gpu/mcmcGPU.py:110-1018)
 * on named device buffers.  This header is that boundary as a C ABI: plain
 * pointers, sizes and NUL-terminated buffer names, no OpenCL, HIP or torch
 * types.  Each entry point cites the reference call it replaces.
 *
 * Conventions
 *  - Every function returns 0 on success or a negative MI3_E* code; the
 *    message of the last failure on a context is mi3_last_error(ctx).
 *  - A context owns one device, one in-order HIP stream and a table of named
 *    device buffers; all compute entry points are asynchronous on that
 *    stream (the reference's event chaining, mcmcGPU.py:211-222, becomes
 *    stream order).  Contexts are not thread-safe.
 *  - Buffer names follow the reference (mcmcGPU.py:173-184, 271-330):
 *    "J", "dJ", "bi", "bi target", "bicount", "seq main", "seq large",
 *    "E main", "E large", "E tmp", "weights", "weights large",
 *    "weightstats", "minout", "cprob", ...  The device layout of "seq *"
 *    buffers is site-major: L rows of `ld` bytes, row i holding residue i of
 *    every sequence (host code converts from/to the reference's (N, L) uint8).
 */
#ifndef MI3HIP_H
#define MI3HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MI3_OK 0
#define MI3_EINVAL (-1)   /* bad argument: ValueError in the reference wrapper */
#define MI3_ENOBUF (-2)   /* unknown buffer name */
#define MI3_ESTATE (-3)   /* component not initialised / already initialised */
#define MI3_EHIP (-4)     /* HIP runtime error */
#define MI3_ENOMEM (-5)   /* allocation failed */
#define MI3_EUNSUP (-6)   /* size outside what this build supports (e.g. q > 25) */

typedef struct mi3_ctx mi3_ctx;

/* Library version string. */
const char *mi3_version(void);

/* Number of visible HIP devices (printGPUs, mcmcGPU.py:58-63). */
int mi3_device_count(int *count);

/* Device name and properties as text (printDevice, mcmcGPU.py:45-56). */
int mi3_device_info(int device, char *buf, size_t buflen);

/* Create a context for an L-site, q-letter Potts model with `nwalkers`
 * walkers (multiple of 512) on `device`.  Replaces MCMCGPU.__init__
 * (mcmcGPU.py:111-193) together with setup_GPU_context / initGPU
 * (mcmcGPU.py:1038-1119): no program is compiled at run time, L and q are
 * kernel arguments (2 <= L <= 1024, 2 <= q <= 25: walker residues are held in
 * registers; MI3_EUNSUP beyond).  `seed` keys every counter-based random
 * stream.  `profile` != 0 records one event pair per launch (the reference's
 * PROFILING_ENABLE queue, mcmcGPU.py:153-157). */
int mi3_create(int device, uint32_t L, uint32_t q, uint32_t nwalkers, uint64_t seed, int profile,
               mi3_ctx **out);

/* Release every device buffer and the stream. */
int mi3_destroy(mi3_ctx *ctx);

/* Message of the last failure on `ctx` (or of the last mi3_create failure
 * when ctx is NULL). */
const char *mi3_last_error(mi3_ctx *ctx);

/* ---- named buffers ------------------------------------------------------ */

/* Allocate a named device buffer of `nbytes` (zero-filled).  Replaces
 * MCMCGPU._setupBuffer (mcmcGPU.py:239-259). */
int mi3_buf_alloc(mi3_ctx *ctx, const char *name, size_t nbytes);

/* 1 if `name` exists, else 0. */
int mi3_buf_exists(mi3_ctx *ctx, const char *name);

/* Size in bytes of a named buffer. */
int mi3_buf_size(mi3_ctx *ctx, const char *name, size_t *nbytes);

/* Copy host memory into a buffer at byte `offset`, ordered on the stream;
 * returns when the host memory may be reused.  Replaces MCMCGPU.setBuf with a
 * host array (mcmcGPU.py:869-907). */
int mi3_buf_write(mi3_ctx *ctx, const char *name, const void *src, size_t nbytes, size_t offset);

/* Device-to-device copy (setBuf with a device buffer, mcmcGPU.py:873-879). */
int mi3_buf_copy(mi3_ctx *ctx, const char *dst, const char *src, size_t nbytes);

/* Copy `nbytes` (a multiple of 16) between two buffers with a plain
 * streaming kernel: the reference for the HBM bandwidth a kernel reaches
 * (bench.py reports the sweep against it).  No counterpart in the reference. */
int mi3_stream_copy(mi3_ctx *ctx, const char *dst, const char *src, size_t nbytes);

/* Read `nbytes` (a multiple of 16) of a buffer with a plain streaming kernel
 * (the read-only bandwidth reference); `out` (4 bytes) is written only in the
 * unlikely case the data's xor equals an internal key.  No counterpart in
 * the reference. */
int mi3_stream_read(mi3_ctx *ctx, const char *src, size_t nbytes, const char *out);

/* Enqueue a read of `nbytes` at `offset` into library-owned pinned staging;
 * returns a ticket.  The data is a snapshot taken in stream order, as
 * getBuf's enqueue_copy (mcmcGPU.py:845-867). */
int mi3_buf_read_async(mi3_ctx *ctx, const char *name, size_t nbytes, size_t offset, int *ticket);

/* Wait for a ticket and copy its data to `dst` (FutureBuf.read,
 * mcmcGPU.py:95-108).  The ticket is released. */
int mi3_ticket_wait(mi3_ctx *ctx, int ticket, void *dst, size_t nbytes);

/* Fill a buffer with a 32-bit pattern (fillBuf, mcmcGPU.py:909-917).  Uses a
 * fill kernel, so it is safe inside captured graphs. */
int mi3_buf_fill_u32(mi3_ctx *ctx, const char *name, uint32_t pattern);

/* Device address of a named buffer and the context's HIP stream, for
 * callers that run their own kernels on the engine's memory (the learned
 * proposal's training step).  After writing a buffer through its address,
 * call mi3_buf_mark_written so derived caches (the chunked coupling copy)
 * are rebuilt.  No counterpart in the reference. */
int mi3_buf_device_ptr(mi3_ctx *ctx, const char *name, void **ptr);
int mi3_buf_mark_written(mi3_ctx *ctx, const char *name);
int mi3_stream(mi3_ctx *ctx, void **stream);

/* Wait for all work on the context's stream (MCMCGPU.wait, mcmcGPU.py:1016). */
int mi3_synchronize(mi3_ctx *ctx);

/* ---- sequences ---------------------------------------------------------- */

/* Fill every sequence of a seq buffer with one sequence of L residues
 * (fillSeqs, mcmcGPU.py:927-932). */
int mi3_fill_seqs(mi3_ctx *ctx, const char *seqbuf, const uint8_t *seq);

/* Copy sequences [src_offset, src_offset+n) of one seq buffer to
 * [dst_offset, dst_offset+n) of another (storeSeqs / restoreSeqs,
 * mcmcGPU.py:934-998; mcmc.cl:67-96). */
int mi3_copy_seqs(mi3_ctx *ctx, const char *src, uint32_t src_offset, const char *dst,
                  uint32_t dst_offset, uint32_t n);

/* At every site i with marks[i] != 0 set residue i of every sequence of dst
 * to residue i of sequence `seqind` of src (copySubseq, mcmcGPU.py:1002-1014;
 * mcmc.cl:98-122).  `marks` names a buffer of at least L bytes. */
int mi3_copy_subseq(mi3_ctx *ctx, const char *src, uint32_t seqind, const char *dst,
                    const char *marks);

/* ---- MCMC ---------------------------------------------------------------- */

/* Launch shape of the sweep: walkers per work-group (128, 256 or 512; other
 * powers of two are rounded into that range).  Takes the role of the
 * reference's work-group size option (WGSIZE compile define,
 * mcmcGPU.py:1080-1093; wgsize_heuristic mcmcGPU.py:1121-1132; --wgsize
 * Mi3.py:115).  A request of 512 runs tiles of 256 when L > 384 and q <= 8. */
int mi3_set_wgsize(mi3_ctx *ctx, uint32_t wgsize);

/* Set nsteps per runMCMC call and the walker-stream offset (initMCMC,
 * mcmcGPU.py:271-283, and _initMCMC_RNG 410-445).  `rng_offset` /
 * `rng_span` select a disjoint walker-index range per device, as the
 * reference's per-GPU stream spans (node_manager.py:73-85). */
int mi3_init_mcmc(mi3_ctx *ctx, uint32_t nsteps, uint64_t rng_offset, uint64_t rng_span);

/* nsteps Metropolis site updates of every walker of "seq main" under the
 * couplings "J" at inverse temperature beta.  Replaces MCMCGPU.runMCMC
 * (mcmcGPU.py:460-482) and the metropolis kernel (mcmc.cl:415-458). */
int mi3_run_mcmc(mi3_ctx *ctx, float beta);

/* Measurement probe with no counterpart in the reference: the coupling-row
 * staging of one mi3_run_mcmc call alone (the same chunk copies into LDS, by
 * the same launch shape, for the sites the next call would visit, with one
 * barrier per chunk and no gathers, reduction or accept).  inflight = 1
 * awaits each chunk's copy before the following barrier, as the sweep does;
 * inflight = 2 keeps one copy in flight across it.  Recorded in the profile
 * as "stageProbe" / "stageProbe2"; the walker state is not touched. */
int mi3_stage_probe(mi3_ctx *ctx, int inflight);

/* Measurement probe with no counterpart in the reference: the 64-bit LDS
 * atomic adds of mi3_weighted_marg alone, in its work-group size (1024
 * threads), histogram (q*q rows of 32 uint64) and column pattern, to
 * pseudo-random rows, with no global loads.  work_groups (1..65536) groups
 * of 1024 threads make iters (1..2^20) adds each.  Recorded in the profile
 * as "wmAtomicProbe". */
int mi3_wm_atomic_probe(mi3_ctx *ctx, uint32_t work_groups, uint32_t iters);

/* Number of MCMC steps executed so far on this context (the global step
 * counter of the random streams). */
int mi3_mcmc_step_count(mi3_ctx *ctx, uint64_t *steps);

/* Independent-site reseed of a seq buffer from "cprob" (gen_indep,
 * mcmcGPU.py:380-391, mcmc.cl:170-244). */
int mi3_gen_indep(mi3_ctx *ctx, const char *seqbuf);

/* ---- energies and marginals ---------------------------------------------- */

/* Ebuf[n] = sum_{i<j} Jbuf[pair(i,j), q*s_i + s_j] for the first nseq
 * sequences of seqbuf.  Replaces calcEnergies (mcmcGPU.py:536-553) and
 * getEnergies (mcmc.cl:249-345). */
int mi3_calc_energies(mi3_ctx *ctx, const char *seqbuf, const char *Jbuf, const char *Ebuf,
                      uint32_t nseq);

/* "bicount"[pair][a*q+b] = #{n < nseq : s_i = a, s_j = b}, exact uint32.
 * Replaces calcBicounts (mcmcGPU.py:501-518) and countBivariate
 * (mcmc.cl:464-516). */
int mi3_calc_bicounts(mi3_ctx *ctx, const char *seqbuf, uint32_t nseq);

/* Number of parts mi3_calc_bicounts splits nseq sequences into (partial
 * counts of the parts are added with integer atomics when it is above 1).
 * A query of the launch plan behind calcBicounts (mcmcGPU.py:501-518); the
 * reference's countBivariate has no split (mcmc.cl:464-516). */
int mi3_bicount_split(mi3_ctx *ctx, uint32_t nseq, uint32_t *ksplit);

/* "bi" = float("bicount") / float(nseq) (bicounts_to_bimarg,
 * mcmcGPU.py:520-534, mcmc.cl:518-528). */
int mi3_bicounts_to_bimarg(mi3_ctx *ctx, uint32_t nseq);

/* "bi"[pair][a*q+b] = sum_n wbuf[n] [s_i=a][s_j=b] over the first nseq
 * sequences (weightedMarg, mcmcGPU.py:634-661, mcmc.cl:646-718).  Summed
 * exactly in 64-bit fixed point of sum(w) * 2^-62, then rounded to float
 * once, so the result does not depend on summation order. */
int mi3_weighted_marg(mi3_ctx *ctx, const char *seqbuf, const char *wbuf, uint32_t nseq);

/* Allocation-free variant of mi3_weighted_marg.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The caller passes device pointers: `seq` site-major with row
 * stride `ld` bytes (a multiple of 16, nseq <= ld), `w` nseq floats, and the
 * output `bi` of (L(L-1)/2) * q * q floats; the result equals
 * mi3_weighted_marg's bit for bit.  Scratch is a library-owned static device
 * array of MI3_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends:
 * the call allocates nothing and rejects nseq > MI3_NOALLOC_MAX_SEQ with
 * MI3_EINVAL.  Every call on one GPU shares that GPU's array, so calls that
 * overlap on one GPU (from several contexts or streams) race; calls on one
 * context's stream are ordered and safe. */
#define MI3_NOALLOC_MAX_SEQ (1u << 20)
#define MI3_NOALLOC_SCRATCH_BYTES \
    ((size_t)MI3_NOALLOC_MAX_SEQ * 8 + 256 + 512 * 16 + 512 + ((size_t)8 << 21))
int mi3_weighted_marg_noalloc(mi3_ctx *ctx, const uint8_t *seq, uint32_t ld, const float *w,
                              uint32_t nseq, float *bi);

/* Divide each pair block of "bi" by its sum (renormalize_bimarg,
 * mcmcGPU.py:663-668; merge_bimarg on one device, node_manager.py:283-287). */
int mi3_renormalize_bimarg(mi3_ctx *ctx);

/* ---- quasi-Newton step ---------------------------------------------------- */

/* "minout"[0] = min(0, min buf[0..n)) (min_buf, mcmcGPU.py:555-567;
 * minFloats, mcmc.cl:558-584). */
int mi3_min_buf(mi3_ctx *ctx, const char *buf, uint32_t n);

/* "weightstats" = (sum w, sum w^2) over wbuf[0..n) (weight_statistics,
 * mcmcGPU.py:569-589; weight_stats, mcmc.cl:586-622). */
int mi3_weight_stats(mi3_ctx *ctx, const char *wbuf, uint32_t n);

/* wbuf = exp(-Ebuf + ref) (dE_to_weights, mcmcGPU.py:591-610,
 * mcmc.cl:624-633). */
int mi3_dE_to_weights(mi3_ctx *ctx, const char *Ebuf, const char *wbuf, uint32_t n, float ref);

/* wbuf = exp((beta-1)*(Ebuf - ref)) (fixed_beta_weights, mcmcGPU.py:612-631,
 * mcmc.cl:635-644). */
int mi3_fixed_beta_weights(mi3_ctx *ctx, const char *Ebuf, const char *wbuf, uint32_t n,
                           float ref, float beta);

/* dst += src over the first n floats (addFloatBuf, mcmcGPU.py:670-685). */
int mi3_add_float_bufs(mi3_ctx *ctx, const char *dst, const char *src, uint32_t n);

/* Jbuf <- Jbuf - gamma*("bi target" - "bi")/("bi" + pc) (updateJ,
 * mcmcGPU.py:704-718; updatedJ, mcmc.cl:829-843). */
int mi3_update_J(mi3_ctx *ctx, float gamma, float pc, const char *Jbuf);

/* Zero-gauge L1 / L2 step on "dJ" given "J" and "bi" (reg_l1z / reg_l2z,
 * mcmcGPU.py:720-745; mcmc.cl:845-902). */
int mi3_reg_l1z(mi3_ctx *ctx, float gamma, float pc, float lJ);
int mi3_reg_l2z(mi3_ctx *ctx, float gamma, float pc, float lJ);

/* updateJ on "dJ" fused with the following reg_l1z (reg = 1), reg_l2z
 * (reg = 2) or nothing (reg = 0) in one pass over the pair blocks: the
 * sequence NewtonSteps.iterNewton issues every iteration
 * (NewtonSteps.py:218-220). */
int mi3_update_J_reg(mi3_ctx *ctx, float gamma, float pc, int reg, float lJ);

/* The remaining regularisers of the coupling step, each one pass over the
 * pair blocks of "dJ" given "J" and "bi" (one q x q block per wave).
 *   SCADJ   (r, a): SCAD penalty on the zero-gauge couplings
 *           (reg_SCADJ, mcmcGPU.py:747-759; mcmc.cl:904-932)
 *   X       (lX): lambda_X |X|, X = sum_ab J(a,b) C(a,b)
 *           (reg_X, mcmcGPU.py:775-787; kernel mcmc.cl:1004-1028)
 *   Xij     one lambda per pair read from "Creg" (first nPairs floats,
 *           as the kernel's lambdas[group]) (reg_Xij, mcmcGPU.py:761-773;
 *           kernel mcmc.cl:972-1002).  The reference's wrappers call each
 *           other's kernel with mismatched arguments; here each name runs
 *           its own kernel's arithmetic.
 *   SCADX   (s, r, a) (reg_SCADX, mcmcGPU.py:789-801; mcmc.cl:1077-1124)
 *   expX    (lam) (reg_expX, mcmcGPU.py:803-815; mcmc.cl:1126-1150)
 *   ddE     (lam) (reg_ddE, mcmcGPU.py:817-829; mcmc.cl:1152-1187)
 *   SCADddE (lam, r) (reg_SCADddE, mcmcGPU.py:831-843; mcmc.cl:1191-1243) */
int mi3_reg_SCADJ(mi3_ctx *ctx, float gamma, float pc, float r, float a);
int mi3_reg_X(mi3_ctx *ctx, float gamma, float pc, float lX);
int mi3_reg_Xij(mi3_ctx *ctx, float gamma, float pc);
int mi3_reg_SCADX(mi3_ctx *ctx, float gamma, float pc, float s, float r, float a);
int mi3_reg_expX(mi3_ctx *ctx, float gamma, float pc, float lam);
int mi3_reg_ddE(mi3_ctx *ctx, float gamma, float pc, float lam);
int mi3_reg_SCADddE(mi3_ctx *ctx, float gamma, float pc, float lam, float r);

/* ---- batched models -------------------------------------------------------
 * Many independent Potts models of uneven L and q, each with its own
 * couplings, walkers and random streams (model m is keyed by seed + m and
 * runs exactly the chains a single-model context created with that seed
 * would), swept together: one persistent kernel per class of models sharing
 * the register layout (ceil(L/128) state words, the chunk class of q), the
 * classes on parallel branches of one HIP graph per round.  No counterpart
 * in the reference, which runs one model per process. */
typedef struct mi3_batch mi3_batch;

/* nmodels models, model m with L[m] sites, q[m] letters, nwalkers[m]
 * walkers (a multiple of 512).  Couplings start at zero, walkers at
 * residue 0. */
int mi3_batch_create(int device, uint32_t nmodels, const uint32_t *L, const uint32_t *q,
                     const uint32_t *nwalkers, uint64_t seed, mi3_batch **out);
int mi3_batch_destroy(mi3_batch *b);
const char *mi3_batch_last_error(mi3_batch *b);

/* Packed (L(L-1)/2, q*q) float32 couplings of one model. */
int mi3_batch_set_couplings(mi3_batch *b, uint32_t model, const float *J);

/* Walkers of one model, site-major: L rows of nwalkers residues. */
int mi3_batch_set_seqs(mi3_batch *b, uint32_t model, const uint8_t *seqs);
int mi3_batch_get_seqs(mi3_batch *b, uint32_t model, uint8_t *out);

/* nsweeps sweeps (nsweeps*L site updates per walker) of every model: one
 * graph launch; the graph is captured on the first call for a given
 * nsweeps and after any tile change. */
int mi3_batch_run(mi3_batch *b, uint32_t nsweeps);
int mi3_batch_synchronize(mi3_batch *b);

/* Autotuner: for every class, time each allowed walkers-per-tile choice
 * (128, 256, 512) on nsweeps sweeps of its models and
 * keep the fastest.  The timed sweeps are real MCMC steps of the walkers. */
int mi3_batch_tune(mi3_batch *b, uint32_t nsweeps);
int mi3_batch_tile(mi3_batch *b, uint32_t model, uint32_t *wpb);
int mi3_batch_set_tile(mi3_batch *b, uint32_t model, uint32_t wpb);

/* Global step counter of a model's random streams. */
int mi3_batch_step_count(mi3_batch *b, uint32_t model, uint64_t *steps);

/* ---- learned proposal (masked autoregressive MLP) -------------------------
 * Kernels of a MADE network (Germain et al. 2015) over sequence positions,
 * used as an independence Metropolis-Hastings proposal beside the sweep:
 *   h1 = relu(b1 + sum_i W1[:, q*i + s_i]), h2 = relu(W2 h1 + b2),
 *   logits = W3 h2 + b3, p(s_j | s_<j) = softmax of site j's q logits,
 * with the MADE masks applied to the weights by the caller.  Launchers take
 * raw device pointers and a HIP stream (NULL = default stream); sequences
 * are site-major (L rows of ld bytes); bf16 operands are raw 16-bit arrays.
 * No counterpart in the reference.  Errors: mi3_made_last_error(). */
const char *mi3_made_last_error(void);

/* C = alpha*op(A)op(B) + beta*C + bias[col], relu if set; fp32 C and/or
 * bf16 Cb.  op(A) is M x K (A is K x M row-major if transA), op(B) is K x N
 * (B is N x K if transB).  M, N, K, lda, ldb multiples of 8. */
int mi3_gemm_bf16(void *stream, int transA, int transB, int M, int N, int K, float alpha,
                  const void *A, int lda, const void *B, int ldb, float beta, float *C, int ldc,
                  void *Cb, int ldcb, const float *bias, int relu);

/* h1 = b1 + sum_i W1t[q*i + s_i][:] (W1t: (L*q, H) float32), h1b = bf16(relu(h1)). */
int mi3_made_embed(void *stream, const uint8_t *seq, uint32_t ld, uint32_t B, uint32_t L,
                   uint32_t q, uint32_t H, const float *W1t, const float *b1, float *h1,
                   void *h1b);

/* dW1t[q*i + a][:] = sum of dh1[b][:] over the b with s_i(b) == a (every
 * entry written). */
int mi3_made_embed_grad(void *stream, const uint8_t *seq, uint32_t ld, uint32_t B, uint32_t L,
                        uint32_t q, uint32_t H, const float *dh1, float *dW1t);

/* nll[b] = -sum_j log softmax(logits[b, site j])[s_j]; optional gradient
 * gscale*(softmax - onehot) as bf16 and/or float32 (B x L*q). */
int mi3_made_xent(void *stream, const float *logits, const uint8_t *seq, uint32_t ld, uint32_t B,
                  uint32_t L, uint32_t q, float gscale, float *nll, void *dlogits_bf16,
                  float *dlogits);

/* out = dh * (hpre > 0), float32 and/or bf16. */
int mi3_relu_grad(void *stream, uint64_t n, const float *dh, const float *hpre, float *out,
                  void *outb);

/* Autoregressive pass over all L sites for B walkers (tiles of 16 per
 * work-group): with draw != 0 it samples s_j (u from Philox((j,
 * walker_base+b, counter, 'MADE')) keyed by seed, by inversion) into row j
 * of the site-major `seq`; with draw == 0 it reads s_j from seq.  Either way
 * nll[b] = -sum_j log p(s_j).  Each hidden unit is formed once, at its
 * degree (the masks make it final then).  W2b, W3b are bf16 copies of the
 * masked W2 (H, H) and W3 (L*q, H); W1t (L*q, H) float32.  Needs q <= 25 and
 * H a multiple of 32 up to 512 (MI3_EUNSUP otherwise). */
int mi3_made_fused(void *stream, uint32_t L, uint32_t q, uint32_t H, uint32_t B, const float *W1t,
                   const float *b1, const void *W2b, const float *b2, const void *W3b,
                   const float *b3, uint64_t seed, uint32_t counter, uint32_t walker_base,
                   uint8_t *seq, uint32_t ld, float *nll, int draw);

/* Independence Metropolis-Hastings step: walker w takes its proposal with
 * probability min(1, exp(-(Eprop - Ecur) - nll_cur + nll_prop)), where
 * nll = -log q under the proposal, u from Philox((counter, walker_base+w,
 * 'MHAC')); accepted columns, energies and nll are copied; naccept counts
 * the moves. */
int mi3_mh_accept(void *stream, uint32_t N, uint32_t L, uint8_t *cur, uint32_t ldc,
                  const uint8_t *prop, uint32_t ldp, float *Ecur, const float *Eprop,
                  float *nll_cur, const float *nll_prop, uint64_t seed, uint64_t counter,
                  uint32_t walker_base, uint32_t *naccept);

/* Training helpers: one Adam step (moments m, v; gradient times the 0/1
 * mask if given; the updated parameters masked again and copied to bf16 pb
 * if given), column sums of a rows x cols float32 matrix, float32 -> bf16. */
int mi3_adam_step(void *stream, uint64_t n, float *p, const float *g, float *m, float *v,
                  const float *mask, float lr, float beta1, float beta2, float eps, uint32_t t,
                  void *pb);
int mi3_colsum(void *stream, uint32_t rows, uint32_t cols, const float *x, float *out);
int mi3_to_bf16(void *stream, uint64_t n, const float *p, void *pb);

/* ---- profiling ------------------------------------------------------------ */

/* Append "EVT name start_ns end_ns bytes" lines for every completed launch to
 * `path` (logProfile, mcmcGPU.py:224-237; the format read by
 * utils/analyze_gpu_profiling.py:4-36).  Returns the number written. */
int mi3_profile_flush(mi3_ctx *ctx, const char *path);

/* Average duration in ms of the launches of kernel `name` recorded since the
 * last reset, and their count. */
int mi3_profile_kernel_ms(mi3_ctx *ctx, const char *name, double *avg_ms, int *count);
int mi3_profile_reset(mi3_ctx *ctx);

#ifdef __cplusplus
}
#endif

#endif /* MI3HIP_H */
