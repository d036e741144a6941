#pragma once

#include <stddef.h>

/*
 * MaxMulSK for AMD Instinct MI355X (gfx950).
 *
 * The first two entry points keep the reference library's C ABI byte for byte
 * (capi/maxmulsk.h:24-37): C = A*B for row-major, tightly packed FP32 matrices,
 * A is M x K, B is K x N, C is M x N, C is overwritten (alpha = 1, beta = 0) and
 * its previous contents are never read. Dimensions are in (M, K, N) order, not
 * BLAS's (M, N, K).
 *
 * Pointers may be host memory (staged through device buffers, a true drop-in for
 * the CPU library) or device memory from hipMalloc (zero-copy). Both entry points
 * are synchronous: they return with C complete. When no usable gfx950 device is
 * present they return MAXMULSK_ERR_UNSUPPORTED and leave C untouched instead of
 * crashing (the reference dies with SIGILL without SME, capi/maxmulsk.h:16-17).
 *
 * Everything after maxmulsk_sgemm_v4 is an additive extension with no
 * counterpart in the reference's C ABI unless its comment names one.
 */

#ifdef __cplusplus
extern "C" {
#endif

/* Same names and values as capi/maxmulsk.h:24-28. MAXMULSK_ERR_DEVICE is
 * additive: a HIP runtime call failed after the arguments were accepted. */
typedef enum {
    MAXMULSK_OK = 0,
    MAXMULSK_ERR_UNSUPPORTED,   /* shape, plan, mode or device rejected; C untouched */
    MAXMULSK_ERR_ALLOC,         /* a device workspace or staging buffer could not be allocated */
    MAXMULSK_ERR_DEVICE,        /* a HIP call failed (launch or copy) */
} maxmulsk_status;

/* Replaces capi/maxmulsk.h:32-33 (maxmulsk_sgemm_v3 -> SMEKernels1x4AccKcOut
 * with the v3 tuning table, sme/support/gemm_tuning.cpp:81-100). Here: plain
 * FP32 MFMA GEMM, plan from the fixed shape rule only (plus the MAXMULSK_PLAN
 * environment override, the analogue of MAXMULSK_BLOCKING). */
maxmulsk_status maxmulsk_sgemm_v3(const float* A, const float* B, float* C,
                                  size_t M, size_t K, size_t N);

/* Replaces capi/maxmulsk.h:36-37 (maxmulsk_sgemm_v4 -> the v4c shape rule,
 * sme/v4/sme-1x4-v4c-dispatch.cpp:17-31). Here: plain FP32 MFMA GEMM, plan from
 * the autotuner cache, then the learned selector, then the shape rule. */
maxmulsk_status maxmulsk_sgemm_v4(const float* A, const float* B, float* C,
                                  size_t M, size_t K, size_t N);

/* ------------------------------------------------------------------------- */
/* Extensions                                                                */
/* ------------------------------------------------------------------------- */

typedef enum {
    MAXMULSK_MODE_PLAIN  = 0,  /* FP32 operands, v_mfma_f32_32x32x2_f32, FP32 accumulate */
    MAXMULSK_MODE_SPLIT3 = 1,  /* a = hi + lo (BF16), hi*hi + hi*lo + lo*hi on BF16 MFMA; error
                                * within (K * 2^-24 + 3 * 2^-16) * |A||B|: the FP32 bound for long
                                * K, not for short K (about K < 64 on uniform data) */
    MAXMULSK_MODE_SPLIT6 = 2,  /* a = hi + mid + lo (BF16), six cross products on BF16 MFMA */
} maxmulsk_mode;

/* A launch plan: output tile edge (64, 128 or 256; the M edge shrinks to 64 or
 * 128 for narrow M) and the number of K slices (split-K, 1..32). {0, 0} asks
 * the planner. Analogue of the reference's Blocking {Mc, Nc, Kc}
 * (sme/v4/sme-1x4-kcout-ncblock-apack4za.hpp:30). */
typedef struct {
    int tile;
    int split_k;
} maxmulsk_plan;

typedef struct {
    maxmulsk_mode mode;
    maxmulsk_plan plan;        /* {0,0}: planner */
    void* stream;              /* hipStream_t; NULL: the legacy default stream */
    int async;                 /* 0: return with C complete; 1: stream-ordered (device pointers only) */
    void* workspace;           /* optional caller-owned device workspace (split-K partials) */
    size_t workspace_bytes;
} maxmulsk_options;

/* C = A*B with an explicit mode, plan, stream and workspace. opt may be NULL
 * (plain mode, planner, default stream, synchronous). Without a workspace, a
 * plan that needs one (split-K partials, packed operands) takes the library's
 * per-device pooled workspace: a call on another stream than the previous user
 * waits for it on the device (an event), while consecutive calls on the legacy
 * default stream are ordered by that stream alone. */
maxmulsk_status maxmulsk_sgemm_ex(const float* A, const float* B, float* C,
                                  size_t M, size_t K, size_t N,
                                  const maxmulsk_options* opt);

/* Bytes of device workspace a call with this shape, mode and plan needs
 * (0 when the plan needs none). Counterpart of
 * SMEKernels1x4FusedPackV51::workspace_bytes (sme/v5/sme-1x4-fusedpack-v51.hpp:72).
 * plan may be NULL (planner). Returns 0 for an invalid plan. */
size_t maxmulsk_workspace_bytes(size_t M, size_t K, size_t N, maxmulsk_mode mode,
                                const maxmulsk_plan* plan);

/* Caller-owned workspace variant: no allocation inside the call when A, B and C
 * are device pointers. ws must be 256-byte aligned device memory of at least
 * maxmulsk_workspace_bytes() bytes (NULL allowed when that is 0); its contents
 * need no initialisation. Counterpart of run_multiplication_ws
 * (sme/v5/sme-1x4-fusedpack-v51.hpp:76-78); a null, short or misaligned
 * workspace is MAXMULSK_ERR_ALLOC, as AllocationFailed is there (v51.cpp:734). */
maxmulsk_status maxmulsk_sgemm_ws(const float* A, const float* B, float* C,
                                  size_t M, size_t K, size_t N,
                                  maxmulsk_mode mode, const maxmulsk_plan* plan,
                                  void* ws, size_t ws_bytes, void* stream);

/* No counterpart in the reference. For callers that cannot allocate device
 * memory (code under stream capture, tight serving loops, harnesses that forbid
 * hipMalloc): maxmulsk_sgemm_ws with the workspace taken from a static device
 * array of MAXMULSK_NOALLOC_SCRATCH_BYTES that the library holds on every GPU
 * it runs on until the process ends. The call allocates nothing (no hipMalloc,
 * hipFree or allocating library routine); its planner uses the
 * MAXMULSK_PLAN override, the autotuner cache and the shape rule, not the
 * learned selector. A, B and C must be device memory; the call is synchronous
 * on `stream`. A plan whose workspace (maxmulsk_workspace_bytes) exceeds the
 * scratch returns MAXMULSK_ERR_ALLOC with C untouched. All calls on one GPU
 * share its scratch: calls that overlap on one GPU (from several threads or
 * streams) race and must be serialised by the caller. */
#define MAXMULSK_NOALLOC_SCRATCH_BYTES ((size_t)32 << 20)
maxmulsk_status maxmulsk_sgemm_noalloc(const float* A, const float* B, float* C,
                                       size_t M, size_t K, size_t N,
                                       maxmulsk_mode mode, const maxmulsk_plan* plan,
                                       void* stream);

/* Pack once, multiply many: an operand rewritten into the MFMA fragment
 * layout (split into its BF16 terms in the split modes), owned by the library.
 * Operand 0 is an A (rows = M, row-major M x K), operand 1 a B (rows = N,
 * row-major K x N). Packed operands feed the packed 256 x 256 GEMM kernel,
 * which maxmulsk_sgemm_ex runs for 256-wide plans on more than 128 rows in
 * plain mode and on ragged split-mode shapes, packing both operands on every
 * call (split-mode calls on whole 256 x 256 tiles split on the fly instead); a
 * caller that reuses an operand (weights) packs it once. Results are the same
 * bit for bit on every path. Counterpart of the reference's B panel packed once per
 * (Kc, Nc) block (sme/v4/sme-1x4-kcout-ncblock-apack4za.cpp:498-503). Device
 * pointers only. */
typedef struct maxmulsk_packed_s* maxmulsk_packed;
maxmulsk_status maxmulsk_pack(const float* X, size_t rows, size_t K, int operand, maxmulsk_mode mode,
                              void* stream, maxmulsk_packed* out);
/* C = A*B from two packed operands of the same mode and K (M = A's rows,
 * N = B's rows). opt->mode and opt->plan.tile are ignored; opt->plan.split_k
 * (0 = 1) splits K; workspace, stream and async as in maxmulsk_sgemm_ex. */
maxmulsk_status maxmulsk_sgemm_packed(maxmulsk_packed A, maxmulsk_packed B, float* C,
                                      const maxmulsk_options* opt);
void maxmulsk_packed_destroy(maxmulsk_packed p);

/* The plan the planner would run for this shape and mode, without running
 * anything (counterpart of SMEKernels1x4V4C::choose,
 * sme/v4/sme-1x4-v4c-dispatch.hpp:61). source receives 0 = environment
 * override, 1 = autotuner cache, 2 = learned selector, 3 = shape rule
 * (may be NULL). */
maxmulsk_status maxmulsk_choose_plan(size_t M, size_t K, size_t N, maxmulsk_mode mode,
                                     maxmulsk_plan* out, int* source);

/* Shape rule only (tier 3 of maxmulsk_choose_plan), pure. */
maxmulsk_plan maxmulsk_rule_plan(size_t M, size_t K, size_t N, maxmulsk_mode mode);

/* One GEMM of a batch: counterpart of struct Gemm
 * (sme/v4/sme-1x4-v4c-onestream.hpp:62). */
typedef struct {
    const float* A;
    const float* B;
    float* C;
    size_t M, K, N;
} maxmulsk_gemm;

/* Runs every problem of the batch in ONE persistent launch (plus one split-K
 * reduce launch when any problem is split). Device pointers only. Every problem
 * is validated before anything runs; the first failure is returned and no C is
 * touched (the validate-all-first contract of run_batch,
 * sme/v4/sme-1x4-v4c-onestream.cpp:683-694). opt->plan is ignored: each problem
 * takes the grouped default plan (see maxmulsk_grouped_default_plan). Returns
 * with every C complete; opt->async is ignored. */
maxmulsk_status maxmulsk_sgemm_grouped(const maxmulsk_gemm* problems, size_t n,
                                       const maxmulsk_options* opt);

/* A prepared batch: plans, work list and workspace are built once and the
 * launch sequence (counter reset, persistent kernel, split-K reduce) is captured
 * in a hipGraph; each launch replays the graph. The pointers inside the
 * problems are baked in. plans may be NULL (grouped policy per problem:
 * grouped autotuner cache, else the grouped rule; see
 * maxmulsk_grouped_default_plan), or hold one plan
 * per problem. In a grouped launch a 256-wide tile on more than 128 rows runs
 * on the 128 x 256 geometry, and in the split modes a 64-wide tile runs on the
 * 64 x 128 geometry (same K slices). A problem on 128-row tiles whose last tile
 * holds at most 64 rows runs that tail on a 64-row tile (plain mode, and the
 * split modes on the decoupled feed); the products per element are the same.
 * The split modes run on the decoupled feed unless the feed setting is
 * MAXMULSK_FEED_DIRECT when the plan is created. A plan owns its work counter and split-K
 * workspace: launches of one plan must not overlap on the device (launch it on
 * one stream, or order the streams); different plans are independent. */
typedef struct maxmulsk_grouped_plan_s* maxmulsk_grouped_plan;
maxmulsk_status maxmulsk_grouped_create(const maxmulsk_gemm* problems, size_t n,
                                        maxmulsk_mode mode, const maxmulsk_plan* plans,
                                        void* stream, maxmulsk_grouped_plan* out);
maxmulsk_status maxmulsk_grouped_launch(maxmulsk_grouped_plan plan, void* stream);
/* Work-list statistics: number of work items (tile x k-slice), algorithmic
 * FLOPs (2MNK summed over problems) and whether the launch replays a graph. */
maxmulsk_status maxmulsk_grouped_info(maxmulsk_grouped_plan plan, size_t* work_items,
                                      double* flops, int* used_graph);
/* 1 when the plan's launch runs on the decoupled feed (the persistent kernel
 * with producer and consumer waves per workgroup, as in the reference's Trio,
 * sme/trio/trio_sme.cpp:274-334), else 0. Split-mode plans take it unless they
 * were built under MAXMULSK_FEED_DIRECT; the work items, tiles and K slices are
 * those of the direct kernel, and so is the result, bit for bit. */
int maxmulsk_grouped_uses_feed(maxmulsk_grouped_plan plan);
void maxmulsk_grouped_destroy(maxmulsk_grouped_plan plan);

/* Plan a problem gets inside a grouped launch when none is given: the grouped
 * autotuner's choice for its shape class if there is one, else the grouped
 * rule (tile edge from N, K sliced into pieces of at most 4096). This is the
 * plan of a problem alone in its launch: in a launch of at least 1024 tiles
 * (one K slice each) the rule's problems run unsliced, unless one of their
 * tiles holds more work than the launch's total divided by 256 CUs. */
maxmulsk_plan maxmulsk_grouped_default_plan(size_t M, size_t K, size_t N, maxmulsk_mode mode);
/* The plans a grouped launch of these problems takes when none are given, one
 * per problem (out: n entries; {0, 0} for problems with M, N or K = 0). Pure:
 * reads only the sizes, runs nothing. */
maxmulsk_status maxmulsk_grouped_plans(const maxmulsk_gemm* problems, size_t n, maxmulsk_mode mode,
                                       maxmulsk_plan* out);

/* Grouped autotuner: shape class = (M rounded up to 64/128/256/512+, N, K).
 * Coordinate search over the candidate set, one class at a time (heaviest
 * classes first, at most max_classes of them; <= 0 means all), timing the whole
 * grouped launch (median of reps replays). The per-class winners are stored in
 * the process-wide grouped cache and returned per problem in plans_out (n
 * entries, may be NULL); *best_ms receives the final launch time (may be NULL). */
maxmulsk_status maxmulsk_grouped_autotune(const maxmulsk_gemm* problems, size_t n,
                                          maxmulsk_mode mode, int max_classes, int reps,
                                          maxmulsk_plan* plans_out, float* best_ms);

/* ---- autotuner (two-stage, correctness-gated; counterpart of the blocking
 * sweep bench/results/2026-09-13/source/bench_nb_sweep.cpp:4-18) ---------- */

/* Number of candidates in the fixed candidate set and the i-th candidate. */
int maxmulsk_num_candidates(void);
maxmulsk_plan maxmulsk_candidate(int i);

/* Times every candidate on device buffers A, B, C (device pointers), writes
 * per-candidate median milliseconds to times_ms[maxmulsk_num_candidates()]
 * (negative = failed the correctness gate or could not run), stores the best
 * plan in *best and in the in-process cache. stage1_reps / stage2_reps <= 0
 * select defaults. */
maxmulsk_status maxmulsk_autotune(const float* A, const float* B, float* C,
                                  size_t M, size_t K, size_t N, maxmulsk_mode mode,
                                  int stage1_reps, int stage2_reps,
                                  float* times_ms, maxmulsk_plan* best);

/* Cache of autotuned plans, keyed by (M, K, N, mode). */
maxmulsk_status maxmulsk_cache_put(size_t M, size_t K, size_t N, maxmulsk_mode mode,
                                   maxmulsk_plan plan);
void maxmulsk_cache_clear(void);
size_t maxmulsk_cache_size(void);
/* Versioned little-endian file with the device name in its header; a file
 * written for another device or format is refused (MAXMULSK_ERR_UNSUPPORTED). */
maxmulsk_status maxmulsk_cache_save(const char* path);
maxmulsk_status maxmulsk_cache_load(const char* path);

/* ---- learned tile selector (3-layer MLP cost model) -------------------- */

/* Network geometry: 4 inputs (log2 M, log2 N, log2 K, mode), two hidden layers
 * of 64 ReLU units, one output per candidate (predicted log runtime). Weights
 * are FP32, row-major W[out][in], laid out W1,b1,W2,b2,W3,b3 in one buffer. */
#define MAXMULSK_MLP_IN 4
#define MAXMULSK_MLP_HIDDEN 64
size_t maxmulsk_mlp_num_params(void);

/* Forward pass over a batch on the device (device pointers): X[batch][4] ->
 * Y[batch][ncand]; H1, H2 (batch x 64) keep the post-ReLU activations for the
 * backward pass (may be NULL for inference). */
maxmulsk_status maxmulsk_mlp_forward(const float* params, const float* X, int batch,
                                     float* H1, float* H2, float* Y, void* stream);
/* Masked mean-squared-error loss gradient and backward pass: T[batch][ncand]
 * targets, Mask[batch][ncand] in {0,1}; writes dParams (same layout as params,
 * overwritten) and the loss into *loss_dev (device float, may be NULL). */
maxmulsk_status maxmulsk_mlp_backward(const float* params, const float* X, int batch,
                                      const float* H1, const float* H2, const float* Y,
                                      const float* T, const float* Mask,
                                      float* dparams, float* loss_dev, void* stream);
/* One Adam step (device pointers): m, v moment buffers; step is 1-based. */
maxmulsk_status maxmulsk_mlp_adam(float* params, const float* dparams, float* m, float* v,
                                  int step, float lr, float beta1, float beta2, float eps,
                                  void* stream);
/* Network input of a shape: (log2 M, log2 N, log2 K) / 13 and mode / 2. */
void maxmulsk_selector_features(size_t M, size_t K, size_t N, maxmulsk_mode mode, float* out4);
/* Installs host-side weights as the process-wide selector used by the planner
 * (tier 2); NULL uninstalls it. A plan is taken from the network only where it
 * is valid for the shape. Planner picks are memoised per shape. */
maxmulsk_status maxmulsk_selector_set(const float* host_params);
/* Loads weights from a file written by maxmulsk_selector_save (magic, geometry,
 * then the FP32 parameters, little endian) and installs them. Without an
 * explicit call the planner loads $MAXMULSK_SELECTOR, else
 * maxmulsk_selector.bin next to the library, on first use (if present). */
maxmulsk_status maxmulsk_selector_load(const char* path);
maxmulsk_status maxmulsk_selector_save(const char* path, const float* host_params);
/* 1 when a selector is installed. */
int maxmulsk_selector_active(void);
/* Selector inference on the device through the dedicated one-wave kernel:
 * writes the predicted log-runtime of every candidate into pred (host array of
 * maxmulsk_num_candidates() floats) and the arg-min into *best. */
maxmulsk_status maxmulsk_selector_predict(size_t M, size_t K, size_t N, maxmulsk_mode mode,
                                          float* pred, maxmulsk_plan* best);

/* ---- introspection --------------------------------------------------------- */

/* Phase instrumentation. Event counters of the host runtime (allocation,
 * stage_in, pack, compute, reduce, writeback, zero_fill), counted per process
 * since the last reset; with MAXMULSK_ROCTX=1 the host phases are also roctx
 * ranges ("maxmulsk/plan", "maxmulsk/allocation", "maxmulsk/stage_in",
 * "maxmulsk/launch", "maxmulsk/writeback", "maxmulsk/sync"). Counterpart of the
 * instrumentation scopes of sme/support/lab_hooks.hpp:10-22 and of the call
 * counts of NB::counts (sme/v4/sme-1x4-kcout-ncblock-apack4za.cpp:448-466). */
int maxmulsk_num_phase_counters(void);
const char* maxmulsk_phase_counter_name(int i);
unsigned long long maxmulsk_phase_counter(int i);
void maxmulsk_phase_counters_reset(void);

/* Matrix-core throughput probe, no counterpart in the reference: every wave
 * of two 4-wave workgroups per CU issues `iterations` x 8 back-to-back MFMAs
 * (mode 0: v_mfma_f32_32x32x2_f32, 1: v_mfma_f32_32x32x16_bf16) into eight
 * accumulators, with register operands: eight different random operand pairs
 * (random_operands = 1, the bits change on every instruction) or one constant
 * pair (0). *tflops receives the issued rate on the current device: the
 * ceiling the GEMM kernels meet on random data, and the cost of the data. */
maxmulsk_status maxmulsk_probe_mfma(int mode, int random_operands, int iterations, double* tflops);

/* Operand feed of the 64 x 64 tile (plans with tile 64, and narrow-M plans
 * whose tile shrinks to 64 x 64) and of the split modes' 128 x 128 tile.
 * MAXMULSK_FEED_DIRECT: each wave stages, splits and multiplies (the direct
 * tile). MAXMULSK_FEED_DECOUPLED: every such launch runs producer waves that
 * stage and split k-blocks into an LDS ring and consumer waves that only
 * multiply and write C, handing slots over through ready/consumed counters.
 * MAXMULSK_FEED_AUTO (the default): the decoupled feed where it measured
 * faster, on launches of at most one workgroup per CU: 64 x 64 split6 with K
 * slices of at least 256, 64 x 64 split3 with slices of at least 1024,
 * 128 x 128 split3/split6 whose tiles and k-blocks are all full (M, N
 * multiples of 128, K and slices multiples of 64, slices of at least 1024),
 * and 128 x 128 split3 with M, N >= 256 and slices of at least 2048;
 * the direct tile elsewhere. Both feeds give the same C (same K
 * slices, products and order), except that an exact -0 may come out as +0.
 * Split-mode grouped launches run on the feed under DECOUPLED and AUTO (it
 * measured 1.03-1.37x faster on every batch tried), on the direct grouped
 * kernel under DIRECT; the setting is read when the grouped plan is created.
 * Plain grouped launches keep their own kernel. Counterpart of the reference's Trio
 * engine, one compute thread fed by separate packing threads through rings
 * (sme/trio/trio.hpp:10-24, sme/trio/trio_sme.cpp:274-334). The setting is
 * process-wide and read at each launch; the first use reads
 * MAXMULSK_FEED=direct|decoupled (anything else: auto). maxmulsk_set_feed
 * returns the previous setting, or -1 for an unknown value (nothing
 * changes). */
#define MAXMULSK_FEED_DIRECT 0
#define MAXMULSK_FEED_DECOUPLED 1
#define MAXMULSK_FEED_AUTO 2
int maxmulsk_set_feed(int feed);
int maxmulsk_get_feed(void);
/* Kernel a call with this shape, mode and plan launches under the current
 * feed setting: 0 = the direct tile (or the packed path), 1..6 = a 64 x 64
 * decoupled-feed variant, 7..8 = a 128 x 128 one (k-block depth, ring slots,
 * waves and loads in flight differ per variant), -1 = invalid plan. */
int maxmulsk_feed_variant(size_t M, size_t K, size_t N, maxmulsk_mode mode, const maxmulsk_plan* plan);

/* Streaming-floor probe, no counterpart in the reference: for a direct-tile
 * plan (not the packed 256 x 256 path) of an M x K x N GEMM, times on
 * `stream` (NULL: the legacy default stream) of the current device, per call
 * over `calls` back-to-back launches: us[0] an empty kernel on the GEMM's grid
 * (the launch floor), us[1] a kernel on the same grid that reads exactly each
 * workgroup's A rows and B columns over its K slice and writes its block of
 * split-K partials, with no arithmetic but the loads' sum, and us[2] that
 * kernel followed by the library's split-K reduce (as us[1] when the plan has
 * no split). Buffers are the probe's own; the call is synchronous. K and N
 * must be multiples of 4. What a call with that plan's launches and bytes
 * costs without staging, splitting or MFMA work. */
maxmulsk_status maxmulsk_probe_stream(size_t M, size_t K, size_t N, const maxmulsk_plan* plan, int calls,
                                      void* stream, double* us);

/* Number of usable gfx950 devices (0 when none). */
int maxmulsk_device_count(void);
/* Text of the last error on this thread (never NULL). */
const char* maxmulsk_last_error(void);
/* Host-side view of the LDS staging layout, for known-answer tests of the
 * swizzle: byte offset inside one pipeline stage of the 16-byte chunk that
 * holds (row, chunk c, term) of operand op (0 = A tile, 1 = B tile), for a
 * tile of `rows` rows (BM or BN) and depth bk, in the given mode; -1 when out
 * of range. */
int maxmulsk_lds_chunk_offset(int mode, int op, int rows, int bk, int row, int c, int term);
/* Host-side view of the epilogue staging of the 256 x 256 tile (a test hook
 * with no counterpart in the reference): each wave passes 32 rows of its wave
 * tile, wtn = 64 columns wide (the only width the tile has), through a
 * row-major block of 32 x (wtn + 4) floats. _write: float offset in that block of lane `lane`'s
 * 16-byte store of accumulator values 4q..4q+3 of fragment column tj (row
 * lane % 32). _read: float offset of the four floats lane `lane` loads in read
 * step i (0..wtn/8-1), whose row and first column of the wave tile go to
 * *row and *col. Both return -1 when out of range. */
int maxmulsk_lds_epilogue_write(int wtn, int lane, int tj, int q);
int maxmulsk_lds_epilogue_read(int wtn, int lane, int i, int* row, int* col);
/* Staging geometry of tile geometry cfg (0..maxmulsk_num_tile_geometries()-1)
 * in a mode, as one-call launches run it: out[0..5] = BM, BN, BK, threads per
 * workgroup, columns per B staging unit, waves (eight for the 128-row tiles of
 * the split modes). Launches of at most 512 workgroups may run the 64 x 64 and
 * split3 128 x 128 tiles with deeper k-blocks than BK (same K slices). Returns
 * 0, or -1 for an unknown cfg or mode. */
int maxmulsk_num_tile_geometries(void);
int maxmulsk_tile_info(int cfg, int mode, int* out);

#ifdef __cplusplus
}
#endif
