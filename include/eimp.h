/*
 * eimp.h -- C ABI of the MI355X (gfx950) engine for EagleImp's phasing
 * pre-stage: top-K conditioning haplotypes, split/call sites, condensed
 * reference and compressed forward/backward PBWT.
 *
 * Library: eagleimp_amd/libeimp.so.  Plain pointers and sizes only.  Every
 * function returns EIMP_OK (0) or a negative status; eimp_last_error() gives
 * the message of the last failure on the calling thread.
 *
 * Threading: an engine handle is used by one host thread at a time, exactly
 * as the reference's GPUEngine is used by one producer thread
 * (src/GPUHandler.cpp:148).  Calls are synchronous (when a call returns, its
 * results are complete in the caller's buffers) except eimp_stream_submit,
 * whose results eimp_stream_wait collects.
 *
 * Output layout (the reference's drop-in GPU layout, src/GPUHandler.cpp:259,
 * src/GPUEngine.cu:160,182-211): per target 2*Mp rows of 4*Kwords uint32,
 * Kwords = roundToMultiple(K,512)/64, forward rows r = 0..Mp-1 first, then the
 * backward rows in processing order (row Mp + i holds site Mp-1-i).  Each row
 * holds ceil(K/32) (data, cumulative zero count) pairs; the remaining pairs of
 * a row are zero (uninitialised in the reference).
 */
#ifndef EIMP_H_
#define EIMP_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EIMP_OK 0
#define EIMP_ERR_INVALID (-1)     /* bad argument or inconsistent sizes */
#define EIMP_ERR_HIP (-2)         /* HIP runtime error */
#define EIMP_ERR_NOMEM (-3)       /* device or host allocation failed */
#define EIMP_ERR_STATE (-4)       /* call not valid in the engine's state */
#define EIMP_ERR_UNSUPPORTED (-5) /* size beyond what the kernels support */

typedef struct eimp_engine eimp_engine;

/* Message of the last failure on this thread ("" if none). */
const char *eimp_last_error(void);

/* Number of visible HIP devices.
 * Replaces hybridsys GPU discovery (src/hybridsys/Hybridsys.cpp:55-60,116-153). */
int eimp_device_count(int *count);

/* ---------------------------------------------------------------------- */
/* Reference GPUEngine interface (src/GPUEngine.h:46-66)                    */
/* ---------------------------------------------------------------------- */

/* GPUEngine::GPUEngine(int index)  (src/GPUEngine.cu:13-42) */
int eimp_engine_create(int device_index, eimp_engine **engine);

/* ~GPUEngine; also releases what eimp_engine_free would. */
int eimp_engine_destroy(eimp_engine *engine);

/* GPUEngine::initialize(K, Kwords, M, Nrefhapscorr, NrefhapsAD, refTCapacity,
 * refTData)  (src/GPUEngine.h:50, src/GPUEngine.cu:44-96).
 * refTData: M site-major rows of refTCapacity bytes (refTCapacity a multiple of
 * 64, bit h of row m = allele of haplotype h at site m, padding bits zero),
 * uploaded once and kept resident.  Kwords must equal roundToMultiple(K,512)/64. */
int eimp_engine_initialize(eimp_engine *engine, size_t K, size_t Kwords, size_t M,
                           size_t Nrefhapscorr, size_t NrefhapsAD, size_t refTCapacity,
                           const uint64_t *refTData);

/* GPUEngine::getBlockSize()  (src/GPUEngine.h:53) -- targets per call. */
int eimp_engine_block_size(eimp_engine *engine, size_t *block_size);

/* GPUEngine::runCrefPBWTKernel(tgtData, splitdata, besthapdata, crefTOffsets,
 * cPBWT, count0s)  (src/GPUEngine.h:55-61, src/GPUEngine.cu:98-217).
 * For each of the n targets: tgtData[t] = GenotypeVector words over all M
 * sites (roundToMultiple(2M,512)/8 bytes are read), splitdata[t] = split-site
 * bitmap (roundToMultiple(M,512)/8 bytes), besthapdata[t] = best-haps bitmap
 * (refTCapacity bytes, exactly K bits set).  crefTOffsets has n+1 entries, the
 * prefix sums of Mp = 2*#splits+1.  cPBWT[t] receives 8*Mp*Kwords uint32
 * (layout above) and count0s[t] receives Mp int32. */
int eimp_engine_run_cref_pbwt(eimp_engine *engine, size_t n, const uint64_t *const *tgtData,
                              const uint64_t *const *splitdata,
                              const uint64_t *const *besthapdata, const uint32_t *crefTOffsets,
                              uint32_t *const *cPBWT, int32_t *const *count0s);

/* GPUEngine::free()  (src/GPUEngine.cu:219-225). */
int eimp_engine_free(eimp_engine *engine);

/* GPUEngine::getCUDAIndex()  (src/GPUEngine.h:66). */
int eimp_engine_device_index(eimp_engine *engine, int *index);

/* ---------------------------------------------------------------------- */
/* Device-side Target::prepPhase (src/Target.cpp:95-163): new work -- the   */
/* reference runs findBestHaps/findMonos/findCallLocs on the host           */
/* (src/GPUHandler.cpp:119-136).                                            */
/* ---------------------------------------------------------------------- */

/* Genetic positions (float cM per site, src/VCFData.cpp:1189) and, optionally,
 * the hap-major panel (src/VCFData.h:203; NrefhapsAD rows of hapStride uint64
 * words, stride a multiple of 8, bit m = site m).  Without hapData the engine
 * transposes the resident site-major panel on the device.  haploidRef: NULL
 * or NrefhapsAD/2 bytes, nonzero = haploid reference sample whose second
 * haplotype is skipped (src/Target.cpp:804-806,835). */
int eimp_engine_set_panel_extras(eimp_engine *engine, const float *cM, const uint64_t *hapData,
                                 size_t hapStride, const uint8_t *haploidRef);

/* One target of a batch: window [s0, s0+W) of the panel's sites (the
 * reference's chunk semantics), genotypes over the window as GenotypeVector
 * words (2*roundToMultiple(W,512)/64 uint64, src/Datatypes.cpp:27-56), K
 * (clamped to Ncorr as src/PBWTPhaser.cpp:683), ignoreSample (-1 none; the
 * sample index skipped in top-K, src/Target.cpp:791-795). */
typedef struct eimp_target {
    const uint64_t *gt;
    uint64_t s0;
    uint64_t W;
    uint64_t K;
    int64_t ignoreSample;
} eimp_target;

/* Batch flags */
#define EIMP_BATCH_DIGEST 1u      /* compute per-target output digests */
#define EIMP_BATCH_RAW 2u         /* rows of roundToMultiple(K,512)/32 data words
                                     only, bits past K and words past the row one:
                                     the FPGA handler's raw buffer, which
                                     PBWT(cpbwt, K, Kwords, M, ntarget) turns into
                                     data/offset pairs (src/PBWT.cpp:72-110,
                                     src/FPGAHandler.cpp:896-900) */
#define EIMP_BATCH_COMPACT 4u     /* rows of 2*ceil(K/32) uint32 (CPU layout,
                                     src/PBWT.cpp:36) instead of 4*Kwords */

/* Per-target summary of a batch */
typedef struct eimp_target_result {
    uint64_t S;          /* split sites */
    uint64_t Mp;         /* 2*S+1 PBWT rows per direction */
    uint64_t ncalls;     /* call (het split) sites */
    uint64_t K;          /* K after clamping */
    uint64_t digest_pbwt;   /* sum of rotl32(v ^ g*0x9E3779B1, g mod 32) over logical words */
    uint64_t digest_count0; /* sum of fmix32(c0 + r*0x7FEB352D + 0x68E31DA4) */
} eimp_target_result;

/* Whole pre-stage for n targets on the device: top-K by (nerr, index),
 * monos, split/call sites, condensed reference, forward and backward PBWT.
 * results[t] is filled for every target.  cMmaxSplit as src/Args.cpp:290.
 * Outputs stay in the engine's device pool (the sink) until the next call. */
int eimp_engine_run_batch(eimp_engine *engine, size_t n, const eimp_target *targets,
                          float cMmaxSplit, uint32_t flags, eimp_target_result *results);

/* eimp_engine_run_batch without any device (or pinned host) allocation.  No
 * counterpart in the reference: for callers that cannot allocate device
 * memory -- code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc.  The caller passes inputs and outputs in device memory:
 * targets[t].gt points to the target's GenotypeVector words ON THE DEVICE,
 * rows[t] to 2*Mp*rowstride uint32 and count0[t] to Mp int32 of device
 * memory (rowstride as the layout flag chooses; Mp <= max_Mp, which the
 * caller sizes the buffers by).  Everything else (distances, best haps,
 * split data, condensed reference, tables) lives in a fixed 64 MiB device
 * scratch array that every device running the library holds until the
 * process ends; a batch needing more, more than 256 targets, or an Mp above
 * max_Mp fails with EIMP_ERR_UNSUPPORTED.  The panel parts a batch reads must
 * be resident already: the mono bits are built by
 * eimp_engine_set_panel_extras, the hap-major rows (targets with K < Ncorr or
 * windows) by passing hapData there or by one earlier eimp_engine_run_batch;
 * otherwise EIMP_ERR_STATE.  All calls on one device share that device's
 * scratch: calls that overlap on one device (two engines, two threads) race.
 * Results, digests and eimp_engine_copy_target behave as after
 * eimp_engine_run_batch (until the next call on the device). */
int eimp_engine_run_batch_noalloc(eimp_engine *engine, size_t n, const eimp_target *targets,
                                  float cMmaxSplit, uint32_t flags, size_t max_Mp,
                                  uint32_t *const *rows, int32_t *const *count0,
                                  eimp_target_result *results);

/* Copies outputs of target t of the last batch (they stay on the device
 * until the next batch or drop-in call).  Any pointer may be NULL.  Sizes: besthaps K uint32, bhflags
 * roundToMultiple(NrefhapsAD,512)/64 uint64, splitsites S uint32, splits
 * roundToMultiple(W,512)/64 uint64, cMdiffs S+1 float (0 if S==0), count0 Mp
 * int32, pbwt 2*Mp*rowstride uint32, cref_rows Mp*roundToMultiple(K,512)/64
 * uint64 (condensed reference rows, src/Target.cpp:253-262). */
int eimp_engine_copy_target(eimp_engine *engine, size_t t, uint32_t *besthaps, uint64_t *bhflags,
                            uint32_t *splitsites, uint64_t *splits, float *cMdiffs,
                            int32_t *count0, uint32_t *pbwt, uint64_t *cref_rows);

/* Mono bits of target t of the last batch (Target::findMonos,
 * src/Target.cpp:878-897: bit m set when all K chosen haplotypes carry the
 * same allele at window site m), roundToMultiple(W,512)/64 uint64 words, bits
 * past W zero. */
int eimp_engine_copy_monos(eimp_engine *engine, size_t t, uint64_t *monos);

/* Panel PBWT for imputation: the reference's PBWT(refT, true)
 * (src/PBWT.cpp:53-71,131-153,207-285, built by src/PBWTImputer.cpp:110,119)
 * over sites [m0, m0+nsites) of the resident panel with all NrefhapsAD
 * haplotypes, forward only: rows receives nsites rows of 2*ceil(N/32) uint32
 * (data / cumulative zero pairs, 1-padding), count0 nsites int32, and absperm,
 * if not NULL, (nsites+1) x N int32 absolute orders (row m = order before site
 * m0+m; getSortOrder, src/PBWT.h:78).  device_ms (NULL or 1 double): device
 * time.  N up to 24,576 (one workgroup carries the order). */
int eimp_engine_ref_pbwt(eimp_engine *engine, size_t m0, size_t nsites, uint32_t *rows,
                         int32_t *count0, int32_t *absperm, double *device_ms);

/* ---------------------------------------------------------------------- */
/* Stream contexts: asynchronous, device-driven batches.  No counterpart in  */
/* the reference, whose GPU calls are synchronous on the default stream     */
/* (src/GPUEngine.cu:98-217); a context moves the producer/consumer         */
/* pipeline of src/GPUHandler.cpp:104-186 onto the device.                  */
/* ---------------------------------------------------------------------- */

typedef struct eimp_stream eimp_stream;

#define EIMP_BUSY 1                /* eimp_stream_query: submission still running */
#define EIMP_STREAM_NO_GRAPH 8u    /* launch kernels one by one instead of replaying
                                      the context's captured hipGraph */
#define EIMP_STREAM_TIMED 16u      /* direct launches with HIP events between the
                                      kernel classes (eimp_stream_kernel_ms) */

/* Creates a context on an initialized engine (cM set): its own HIP stream,
 * fixed device buffers for up to max_targets (<= 1024) targets with windows of
 * up to max_window sites and K up to max_K (after clamping), an output pool of
 * pool_bytes, and the captured launch sequence.  flags: EIMP_BATCH_DIGEST,
 * EIMP_BATCH_COMPACT or EIMP_BATCH_RAW, EIMP_STREAM_NO_GRAPH,
 * EIMP_STREAM_TIMED.  The panel is shared with the engine and its other
 * contexts; each context is used by one host thread at a time. */
int eimp_stream_create(eimp_engine *engine, size_t max_targets, size_t max_window, size_t max_K,
                       size_t pool_bytes, uint32_t flags, eimp_stream **stream);

/* As eimp_stream_create, with the output pool of `other` (same engine)
 * instead of a pool of its own.  The contexts sharing a pool take turns:
 * each submission's build (planner, cref, PBWT) starts after the pool's
 * previous build, whichever context submitted it, while its preparation
 * (distances, top-K, monos, call locations) runs beside that build.  A
 * context's outputs in the pool (count0, rows, cref) stay valid until any
 * context sharing the pool submits again; copying them afterwards returns
 * EIMP_ERR_STATE.  The pool lives until its last context is destroyed. */
int eimp_stream_create_shared(eimp_stream *other, size_t max_targets, size_t max_window,
                              size_t max_K, uint32_t flags, eimp_stream **stream);

/* Queues a batch and returns without waiting.  On the device: the prep
 * kernels, a planner that sizes each target's outputs from its split count
 * and places them in the pool, then the cref and PBWT kernels pulling work
 * from device queues (largest first).  Targets whose outputs do not fit the
 * pool are deferred.  The caller may reuse targets[].gt when the call
 * returns.  One submission per context until eimp_stream_wait collects it. */
int eimp_stream_submit(eimp_stream *stream, size_t n, const eimp_target *targets,
                       float cMmaxSplit);

/* EIMP_OK when the last submission has completed, EIMP_BUSY while it runs. */
int eimp_stream_query(eimp_stream *stream);

/* Waits for the last submission and fills results[0..n-1].  A deferred target
 * has results[t].Mp == 0 (its S is set); *ndeferred counts them.  times: NULL
 * or 2 doubles, device milliseconds of [0] the kernels and [1] the whole
 * submission including its copies. */
int eimp_stream_wait(eimp_stream *stream, eimp_target_result *results, size_t *ndeferred,
                     double *times);

/* Device milliseconds per kernel class (EIMP_KCLASS_*, EIMP_NKCLASS values)
 * of the last collected submission of a context created with
 * EIMP_STREAM_TIMED.  The planner counts as cref, the count0 digest as PBWT;
 * EIMP_KCLASS_COPY holds the submission's host<->device copies. */
int eimp_stream_kernel_ms(eimp_stream *stream, double *ms);

/* As eimp_engine_copy_target, for target t of the last collected submission
 * (outputs stay in the context's pool until its next submission, or the next
 * submission of a context sharing the pool). */
int eimp_stream_copy_target(eimp_stream *stream, size_t t, uint32_t *besthaps, uint64_t *bhflags,
                            uint32_t *splitsites, uint64_t *splits, float *cMdiffs,
                            int32_t *count0, uint32_t *pbwt, uint64_t *cref_rows);

/* As eimp_engine_copy_monos, for target t of the last collected submission. */
int eimp_stream_copy_monos(eimp_stream *stream, size_t t, uint64_t *monos);

/* Digests of the outputs the last collected submission left in the context's
 * pool, recomputed from its stored rows and count0: the same two sums as
 * EIMP_BATCH_DIGEST (results[].digest_pbwt / digest_count0), for contexts
 * created without it -- a check of those very outputs after the fact.
 * digests: 2 uint64 per target of the submission, 0 for deferred targets.
 * The raw layout holds no zero counts: EIMP_ERR_UNSUPPORTED. */
int eimp_stream_digest_outputs(eimp_stream *stream, uint64_t *digests);

/* Waits for outstanding work and releases the context. */
int eimp_stream_destroy(eimp_stream *stream);

/* Free and total device memory of the engine's device (admission by free
 * HBM). */
int eimp_engine_mem_info(eimp_engine *engine, size_t *free_bytes, size_t *total_bytes);

/* ---------------------------------------------------------------------- */
/* Measurement                                                              */
/* ---------------------------------------------------------------------- */

/* Kernel classes timed with HIP events on the engine's stream. */
#define EIMP_KCLASS_DISTANCE 0
#define EIMP_KCLASS_SELECT 1
#define EIMP_KCLASS_MONOS 2
#define EIMP_KCLASS_SPLITS 3
#define EIMP_KCLASS_CREF 4
#define EIMP_KCLASS_PBWT 5
#define EIMP_KCLASS_COPY 6
#define EIMP_NKCLASS 7

/* Enables (1) or disables (0) per-launch event timing. */
int eimp_engine_set_timing(eimp_engine *engine, int enable);
/* Accumulated milliseconds and launches per kernel class since the last reset. */
int eimp_engine_get_timing(eimp_engine *engine, double *ms, uint64_t *launches);
int eimp_engine_reset_timing(eimp_engine *engine);

/* PBWT kernel variant: -1 automatic, otherwise an index into the engine's
 * variant table (threads per workgroup x chunks per wave); the autotuner in
 * the Python package sweeps these. */
int eimp_engine_set_pbwt_variant(eimp_engine *engine, int variant);
int eimp_pbwt_variant_count(void);
int eimp_pbwt_variant_info(int variant, int *threads, int *chunks_per_wave, int *max_k);

/* ------------------------------------------------------------------------ */
/* Beam-search phasing over a target's compressed PBWT (host code): the     */
/* consumer of the engine's output.  The reference's GPU handler wraps each */
/* target's cPBWT as PBWT(cpbwt, count0, K, 4*Kwords, Mp) and calls          */
/* Target::phaseExtPBWT (src/GPUHandler.cpp:259-266, src/Target.cpp:291-777,*/
/* src/Beam.cpp, src/BeamPath.cpp, src/HaplotypePath.cpp, src/PBWT.cpp:      */
/* 288-382).  Results are identical to the reference's (same float          */
/* arithmetic, same pruning order).                                         */
/* ------------------------------------------------------------------------ */

/* Parameters; eimp_phase_default_params gives the reference's defaults
 * (src/Target.cpp:41-51; src/Args.cpp:276-278, src/main.cpp:305-306,326:
 * hist 1, pErr 0.003, pLimit = pErr, expectIBDcM 2; reverse phasing on,
 * pre-phasing off, last iteration, missing genotypes not imputed). */
typedef struct eimp_phase_params {
    uint32_t history;          /* HistorySize (100); times hist, at most 128 */
    uint32_t history_fast;     /* HistorySizeFast (30), pre-phasing */
    uint32_t beam_width;       /* BeamWidth (50) */
    uint32_t beam_width_fast;  /* BeamWidthFast (30) */
    uint32_t delta;            /* Delta (20) */
    uint32_t delta_fast;       /* DeltaFast (10) */
    float hist;                /* history length multiplier */
    float p_err;               /* genotype error probability */
    float p_limit;             /* beam probability deviation limit */
    float expect_ibd_cM;       /* expected IBD length (cM) */
    float min_fix_size, max_fix_size, min_fix_thresh, max_fix_thresh; /* 0.5 0.9 0.99 1.0 */
    int32_t pre_phasing;       /* doPrePhasing */
    int32_t rev_phasing;       /* doRevPhasing */
    int32_t last_iter;         /* lastIter */
    int32_t imp_missing;       /* impMissing (needs af) */
} eimp_phase_params;

void eimp_phase_default_params(eimp_phase_params *params);

/* One target.  pbwt: 2*Mp rows of row_stride uint32 holding (data, cumulative
 * zero count) pairs, forward rows first, then the backward rows in processing
 * order (the drop-in layout has row_stride 4*Kwords, the compact CPU layout
 * 2*ceil(K/32)); with raw_rows set, rows of data words only (the FPGA
 * handler's raw buffer, EIMP_BATCH_RAW, bits past K one), turned into pairs
 * as PBWT(cpbwt, K, Kwords, M) does (src/PBWT.cpp:72-110, consumed by
 * src/FPGAHandler.cpp:861-905), count0 then optional.  gt: GenotypeVector words over the window's W sites.
 * splitsites: the S = (Mp-1)/2 window-relative split sites, cMdiffs S+1
 * floats (both as eimp_engine_copy_target returns them).  af: NULL or W
 * allele frequencies (used with imp_missing).  start_phase: haplotype of the
 * maternal path at the first call site, 0 Ref / 1 Alt
 * (VCFData::getStartPhase). */
typedef struct eimp_phase_input {
    const uint32_t *pbwt;
    const int32_t *count0;
    uint32_t K, row_stride, Mp, W;
    const uint64_t *gt;
    const uint32_t *splitsites;
    const float *cMdiffs;
    const float *af;
    int32_t start_phase;
    int32_t raw_rows;
} eimp_phase_input;

/* mat/pat: ceil(W/64) words each (bit m = allele at window site m),
 * dosage_mat/pat: W floats each, call_conf: NULL or one float per call site
 * (written in the last iteration).  total_conf, ncallsites, nmissing,
 * nmonhets: the target's totals (Target::getTotalConfidence,
 * getNCallSites, getMissings, getMonHets). */
typedef struct eimp_phase_output {
    uint64_t *mat, *pat;
    float *dosage_mat, *dosage_pat;
    float *call_conf;
    float total_conf;
    uint32_t ncalls;
    uint64_t ncallsites;
    int32_t nmissing, nmonhets;
} eimp_phase_output;

int eimp_phase_target(const eimp_phase_params *params, const eimp_phase_input *in,
                      eimp_phase_output *out);

/* n independent targets over `threads` host threads (<= 0: hardware
 * concurrency), as the reference's phasing threads take targets from their
 * queue (src/PBWTPhaser.cpp:490-499).  Returns the first failure. */
int eimp_phase_batch(const eimp_phase_params *params, size_t n, const eimp_phase_input *in,
                     eimp_phase_output *out, int threads);

/* ------------------------------------------------------------------------ */
/* Qref reference panel reader (host code): loads the reference's binary     */
/* panel format (qref_spec.md; VCFData::qRefOpenReadMeta and                 */
/* qRefLoadNextVariant, src/VCFData.cpp:2107-2432; run-length code           */
/* src/utils.h:133-315) into the site-major rows eimp_engine_initialize      */
/* takes, with the reference's sample masking of participant withdrawals.    */
/* Target matching (VCF/HTSlib) is not part of it.                           */
/* ------------------------------------------------------------------------ */
typedef struct eimp_qref eimp_qref;

typedef struct eimp_qref_info {
    int32_t v_major, v_minor;   /* format version (1.<=1 accepted, src/VCFData.h:413-414) */
    int32_t chrom;              /* 1..24, 23 = X */
    int32_t has_haploid_flags;  /* per-sample flags stored in the file */
    uint64_t Nref;              /* samples (haplotypes: 2*Nref, haploids encoded homozygous) */
    uint64_t NrefHap;           /* haploid samples */
    uint64_t Mref;              /* variants */
    uint64_t MrefMA;            /* variants split from multi-allelics */
    uint64_t hap_words;         /* uint64 words per stored haplotype vector (2*Nref rounded to 512 bits) */
} eimp_qref_info;

/* Maps the file, checks header and sizes, indexes the variants. */
int eimp_qref_open(const char *path, eimp_qref **qref, eimp_qref_info *info);
int eimp_qref_close(eimp_qref *qref);

/* Any pointer may be NULL.  positions: Mref int64 (0-based), af: Mref float
 * as stored in the file (the masked frequencies come from
 * eimp_qref_read_sites_af), haploid: Nref bytes of the kept samples (1 =
 * haploid; all equal when the file stores no flags), ma_flags: ceil(Mref/64)
 * words (bit m = split from a multi-allelic). */
int eimp_qref_metadata(eimp_qref *qref, int64_t *positions, float *af, uint8_t *haploid,
                       uint64_t *ma_flags);

/* Alleles (reference/major first) and ID of variant m: zero-terminated
 * strings valid until eimp_qref_close. */
int eimp_qref_variant(eimp_qref *qref, size_t m, const char **ref, const char **alt,
                      const char **id);

/* Haplotypes of variants [m0, m0+n) as n site-major rows of row_words uint64
 * (bit h = allele of haplotype h; row_words >= ceil(2*Nref/64), bits past
 * 2*Nref zero), decoded over `threads` host threads (<= 0: all). */
int eimp_qref_read_sites(eimp_qref *qref, size_t m0, size_t n, uint64_t *rows, size_t row_words,
                         int threads);

/* Masks samples (participant withdrawals; replaces the mask the reference
 * builds in VCFData::parseSampleList, src/VCFData.cpp:3698-3746, and applies
 * in qRefOpenReadMeta / qRefLoadNextVariant, src/VCFData.cpp:2139-2208 and
 * 2418-2530): `samples` strictly ascending sample indices of the file (n = 0
 * clears the mask).  Afterwards Nref, NrefHap and hap_words describe the kept
 * samples (*info, may be NULL, receives them), rows hold only their
 * haplotypes and eimp_qref_metadata's haploid flags only theirs. */
int eimp_qref_set_mask(eimp_qref *qref, const uint64_t *samples, size_t n, eimp_qref_info *info);

/* VCFData::parseSampleList (src/VCFData.cpp:3698-3746): one sample per line
 * (first token; the second one after two header lines in a file ending in
 * ".sample"), a withdrawal is an ID of 'W' or '-' followed by digits only.
 * Writes up to cap withdrawal indices (masked may be NULL to count),
 * *nmasked = withdrawals, *nsamples = samples listed. */
int eimp_qref_parse_sample_list(const char *path, uint64_t *masked, size_t cap, size_t *nmasked,
                                size_t *nsamples);

/* Masks the withdrawals of a sample list; the list must name as many samples
 * as the file holds when it has any withdrawal (src/VCFData.cpp:2139-2142). */
int eimp_qref_mask_sample_list(eimp_qref *qref, const char *path, eimp_qref_info *info);

/* eimp_qref_read_sites plus, per variant, the allele frequency the reference
 * holds after masking (af, src/VCFData.cpp:2426-2432 and 2524-2529: the
 * file's frequency without a mask) and the variant alleles the masked
 * samples carried (masked).  rows, af and masked may each be NULL. */
int eimp_qref_read_sites_af(eimp_qref *qref, size_t m0, size_t n, uint64_t *rows,
                            size_t row_words, float *af, uint64_t *masked, int threads);

#ifdef __cplusplus
}
#endif

#endif /* EIMP_H_ */
