/*
 * zkgpu.h — C ABI of libzkgpu.so, the MI355X (gfx950) BN254 MSM + Fr NTT library.
 *
 * This is the drop-in boundary for zerotruss/zk-gpu's public Rust API (src/lib.rs:1-6). Each
 * entry point names the reference item it replaces (file:line in the reference repository).
 * The reference has no C ABI of its own; a Rust facade binds these symbols one-to-one through
 * extern "C" declarations of the same names and argument types.
 *
 * Conventions
 *  - Field elements are 8 little-endian u32 words, plain integers below the modulus
 *    (README.md:77-78; src/host.rs:104-110). G1 affine points are 16 words (x, y); G2 affine
 *    points are 32 words (x.a, x.b, y.a, y.b) with Fq2 = Fq[u]/(u^2+1); infinity is all zeros
 *    (src/group.rs:57-69).
 *  - Every function returns ZKG_OK (0) or an error code; zkg_last_error() then holds a message
 *    for the calling thread. Size and range violations use the reference's panic strings
 *    (src/msm.rs:97-98,275-281; src/ntt.rs:27,54-59), so a facade can re-raise them verbatim.
 *  - All work on a device is queued in FIFO order on one HIP stream. zkg_msm_run,
 *    zkg_ntt_forward and zkg_ntt_inverse return before the GPU finishes; zkg_buffer_read and
 *    zkg_wait synchronise (src/gpu.rs:151-199).
 *  - Plans are not re-entrant: a plan owns its workspace and its result buffer
 *    (src/msm.rs:25-64,286-298; CONTRIBUTING.md:83-84).
 */
#ifndef ZKGPU_H
#define ZKGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct zkg_device zkg_device;
typedef struct zkg_buffer zkg_buffer;
typedef struct zkg_msm zkg_msm;
typedef struct zkg_ntt zkg_ntt;
typedef struct zkg_zkey zkg_zkey; /* a mapped snarkjs proving key (see "snarkjs .zkey ingestion") */

enum {
    ZKG_OK = 0,
    ZKG_ERR_INVALID = 1,  /* argument out of range */
    ZKG_ERR_SIZE = 2,     /* buffer smaller than the plan needs */
    ZKG_ERR_DEVICE = 3,   /* no GPU / HIP runtime error */
    ZKG_ERR_NOMEM = 4,    /* device allocation failed or over budget */
    ZKG_ERR_INTERNAL = 5
};

enum { ZKG_G1 = 1, ZKG_G2 = 2 }; /* Group::{G1, G2}  (src/group.rs:10-14) */

/* Message of the last failed call on this thread ("" if none). */
const char* zkg_last_error(void);
/* ABI version (major * 10000 + minor * 100 + patch). */
int zkg_version(void);

/* ------------------------------------------------------------------ device and buffers */

/* Gpu::new (src/gpu.rs:51-55). Fails with "GPU init failed" when no device exists. */
int zkg_device_open(int ordinal, zkg_device** out);
int zkg_device_close(zkg_device* dev);
/* Compute units, shader clock (kHz), HBM bytes and gfx arch name of the device. */
int zkg_device_info(zkg_device* dev, int* cu_count, int* clock_khz, size_t* total_mem, char* name,
                    size_t name_len);
/* No counterpart in the reference (wgpu hides its queue): the HIP stream (hipStream_t) on which
 * the library queues all work for this device, so callers can record events or order their
 * own work against it. */
void* zkg_device_stream(zkg_device* dev);
/* Gpu::report_errors (src/gpu.rs:40-49): `fn` receives "GPU error: <message>" for every error the
 * HIP runtime returns on this device -- in a call (which also returns the error code) or where no
 * call waits for it (a proof-job server's scheduler thread). It may run on any thread and must
 * not call back into the library for the same device. fn = NULL removes the callback. */
typedef void (*zkg_error_fn)(const char* message, void* user);
int zkg_device_report_errors(zkg_device* dev, zkg_error_fn fn, void* user);
/* Gpu::buffer (src/gpu.rs:57-64): device buffer initialised with n_words u32. */
int zkg_buffer_create(zkg_device* dev, const uint32_t* words, size_t n_words, zkg_buffer** out);
/* Gpu::empty_buffer (src/gpu.rs:66-69): zero-filled buffer of n_words u32. */
int zkg_buffer_empty(zkg_device* dev, size_t n_words, zkg_buffer** out);
/* Gpu::read (src/gpu.rs:192-199): blocking read of the first n_words u32 (synchronises). */
int zkg_buffer_read(zkg_buffer* buf, size_t n_words, uint32_t* out);
/* Gpu::try_read_async / read_async (src/gpu.rs:157-190): queue a read of the first n_words into
 * `out` and return at once with a ticket. zkg_read_done polls it; zkg_read_wait blocks until the
 * words are in `out` (which must stay valid until then) and frees the ticket. A failure returns
 * "GPU read failed: ..." instead of aborting. */
typedef struct zkg_read zkg_read;
int zkg_buffer_read_async(zkg_buffer* buf, size_t n_words, uint32_t* out, zkg_read** ticket);
int zkg_read_done(zkg_read* ticket, int* done);
int zkg_read_wait(zkg_read* ticket);
/* wgpu Queue::write_buffer as used by callers that refill a buffer: host->device copy of n_words
 * at offset_words (returns once `words` may be reused). */
int zkg_buffer_write(zkg_buffer* buf, size_t offset_words, const uint32_t* words, size_t n_words);
/* wgpu::Buffer::size: capacity in bytes. */
size_t zkg_buffer_size(const zkg_buffer* buf);
/* No counterpart in the reference: wrap caller-owned device memory (e.g. a torch tensor) as a
 * buffer without copying; the memory must outlive the handle and is never freed by the library. */
int zkg_buffer_wrap(zkg_device* dev, void* device_ptr, size_t bytes, zkg_buffer** out);
void* zkg_buffer_device_ptr(const zkg_buffer* buf);
int zkg_buffer_free(zkg_buffer* buf);
/* Gpu::wait (src/gpu.rs:151-155): block until all queued work on the device is done. */
int zkg_wait(zkg_device* dev);

/* ------------------------------------------------------------------ MSM */

/* Msm::default_window (src/msm.rs:76-80): clamp(ilog2(max(n,1)) - 5, 11, 16). */
uint32_t zkg_msm_default_window(size_t n);
/* Msm::for_group (src/msm.rs:95-180); Msm::new (:86) is group = ZKG_G1, item_points = 0;
 * Msm::with_item_points (:91) passes item_points. item_points = 0 picks the size automatically.
 * Errors: n = 0 -> "an MSM needs at least one point"; c outside 9..=20 ->
 * "window size c must be in 9..=20, got {c}". */
int zkg_msm_create(zkg_device* dev, int group, size_t n, uint32_t c, uint32_t item_points, zkg_msm** out);
/* The checks Msm::for_group makes (src/msm.rs:96-98) plus this library's size limits, without a
 * device and without allocating: ZKG_OK and the plan's device workspace in bytes when
 * zkg_msm_create (precomputed = 0) or zkg_msm_create_precomputed (precomputed = 1) would accept
 * (group, n, c, item_points); otherwise the error those calls return. The library indexes the
 * sorted (point, digit) entries with 32 bits, so plans needing 2^32 or more of them (W * nv, with
 * W windows over nv virtual points) are rejected: "an MSM of {n} points at c = {c} needs {k}
 * sorted entries; at most 2^32 - 1 fit one plan". */
int zkg_msm_plan_check(int group, size_t n, uint32_t c, uint32_t item_points, int precomputed,
                       size_t* workspace_bytes);
/* No counterpart in the reference: a plan whose window, work-item size and sort-block size come
 * from the tuning table (zkg_tune_msm / zkg_tune_load) for this GPU and size; untuned sizes use
 * default_window(n) with automatic item and sort sizes. Results equal any other plan's. */
int zkg_msm_create_tuned(zkg_device* dev, int group, size_t n, zkg_msm** out);
/* No counterpart in the reference (fixed-base precomputation for bases that stay resident, such as
 * a proving key's): *out receives a new buffer of 2n points of `group` holding the n points of
 * `bases` (library form, as zkg_msm_upload_bases* leave them) followed by 2^64 times each (64
 * doublings and one inversion per point, on the device; infinity stays (0, 0)). Queued on the
 * device's stream. Errors: "bases buffer holds ..." when `bases` has fewer than n points. */
int zkg_msm_precompute_bases(zkg_device* dev, int group, const zkg_buffer* bases, size_t n, zkg_buffer** out);
/* No counterpart in the reference: a plan whose zkg_msm_run takes a zkg_msm_precompute_bases
 * buffer (2n points) as `bases`. Each GLV half-scalar splits at bit 64, so the stages run over 4n
 * virtual points with ceil(64 / c) windows: the same bucket additions, half the buckets to reduce
 * and about half the doublings of the window combine. Same result as zkg_msm_create's plan on the
 * original bases. c in 9..=16 ("a plan over precomputed bases takes c in 9..=16, got {c}"),
 * n < 2^29. */
int zkg_msm_create_precomputed(zkg_device* dev, int group, size_t n, uint32_t c, uint32_t item_points,
                               zkg_msm** out);
/* Points a plan's bases buffer must hold: n, or 2n for a plan over precomputed bases. */
size_t zkg_msm_base_points(const zkg_msm* plan);
/* Points per work item the plan uses (the automatic choice when created with item_points = 0). */
uint32_t zkg_msm_item_points(const zkg_msm* plan);
int zkg_msm_destroy(zkg_msm* plan);
/* Msm::group (src/msm.rs:267-269). */
int zkg_msm_group(const zkg_msm* plan);
/* Window size and point count of a plan. */
uint32_t zkg_msm_window(const zkg_msm* plan);
size_t zkg_msm_points(const zkg_msm* plan);
/* Msm::upload_bases (src/msm.rs:182-188): plain affine coordinates -> device buffer in the
 * library's Montgomery form. Group-agnostic: n_words is 8 per coordinate. */
int zkg_msm_upload_bases(zkg_device* dev, const uint32_t* plain_xy, size_t n_words, zkg_buffer** out);
/* Msm::upload_bases_montgomery_256 (src/msm.rs:190-197): coordinates already in Montgomery form
 * with R = 2^256 (the snarkjs/ffjavascript form). The library's R is 2^256, so this is a copy
 * followed by a reduction of every word below p on the device (any 256-bit word is accepted, as the
 * reference's scaling product accepts it). */
int zkg_msm_upload_bases_montgomery_256(zkg_device* dev, const uint32_t* xy, size_t n_words, zkg_buffer** out);
/* Convert an existing device buffer of plain coordinates to the library's form in place
 * (what upload_bases does after its copy; for bases produced on the device). */
int zkg_msm_convert_bases(zkg_buffer* buf, size_t n_words);
/* No counterpart in the reference (it checks nothing about its inputs, README.md:218-219): counts
 * the first n points of `bases` (library form, as the upload calls leave them) that are neither
 * infinity nor on the curve of `group` (G1: y^2 = x^3 + 3; G2: the twist y^2 = x^3 + 3/(9+u)) and,
 * for G2, in the order-r subgroup (r P = O, by double-and-add on the device); *first_bad = the
 * smallest such index (0 when *bad = 0). Waits for the device.
 *
 * Why it matters: plans split scalars with the GLV endomorphism phi(x, y) = (beta x, y), which
 * acts as multiplication by lambda only on the order-r subgroup, and G2 plans up to 2^18 points
 * also with psi (the Frobenius carried to the twist), which acts as p mod r only there. G1 has
 * cofactor 1 (on the curve is enough), but the G2 twist has cofactor 2p - r: a G2 base on the
 * twist but outside the subgroup gives a different sum than the reference's plain Pippenger. Bases of a Groth16 proving
 * key are subgroup points; zkg_zkey_upload and zkg_server_add_zkey run this check themselves. */
int zkg_msm_check_bases(zkg_device* dev, int group, const zkg_buffer* bases, size_t n, uint64_t* bad,
                        uint64_t* first_bad);
/* Msm::run (src/msm.rs:284-298): queue sum(scalars[i] * bases[i]). *result (unless `result` is
 * NULL; zkg_msm_result gives the same handle at any time) receives the plan-owned buffer holding one XYZZ point (32 words G1, 64 words G2, Montgomery form), valid
 * until the next run on this plan. Errors: "scalars buffer holds ..." / "bases buffer holds ...".
 * Precondition (not checked here, see zkg_msm_check_bases): G2 bases lie in the order-r subgroup.
 * ZKGPU_MSM_GLV=0 in the environment at plan creation selects plans without the GLV (and GLS)
 * split, which sum any points of the twist exactly as the reference does. */
int zkg_msm_run(zkg_msm* plan, const zkg_buffer* bases, const zkg_buffer* scalars, zkg_buffer** result);
/* No counterpart in the reference. For callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc): the MSM that
 * zkg_msm_create(dev, group, n, c, 0) + zkg_msm_run compute, with the same kernels and the same
 * result words written to the caller's `result` (32 words G1, 64 words G2: one XYZZ point in the
 * library's form, as zkg_msm_run's result buffer), but no plan and no allocation: the workspace
 * is carved from a static device array of 64 MiB per GPU (held from zkg_device_open until the
 * process ends), which covers e.g. G1 up to n = 2^16 at the default window, G1 2^17 at c = 12,
 * G2 2^16 at c = 11. Larger workspaces are rejected ("an MSM of {n} points at c = {c} needs a
 * workspace of {k} bytes; the static scratch holds 67108864"); plan and buffer errors as
 * zkg_msm_create / zkg_msm_run. Queued on the device's stream (returns before the GPU finishes).
 * Every call on one GPU shares that GPU's array: calls that overlap on one GPU (from several host
 * threads, or this call next to a batch or server using it) race. Every stage writes the scratch
 * it reads within the same call, so nothing is carried over from the previous call. */
int zkg_msm_run_noalloc(zkg_device* dev, int group, size_t n, uint32_t c, const zkg_buffer* bases,
                        const zkg_buffer* scalars, zkg_buffer* result);
/* Msm::run_yielding (src/msm.rs:300-318): as zkg_msm_run, but waits for the GPU after the
 * heavy stages (accumulate, finalize, reduce chunks, combine), so a display sharing the GPU gets it
 * in between. Same result. */
int zkg_msm_run_yielding(zkg_msm* plan, const zkg_buffer* bases, const zkg_buffer* scalars, zkg_buffer** result);
/* The plan-owned result buffer (what zkg_msm_run returns), e.g. after a batch ran the plan. */
int zkg_msm_result(zkg_msm* plan, zkg_buffer** result);
/* Msm::profile (src/msm.rs:321-342): run once, synchronising around every stage; writes up to
 * `cap` (stage name, milliseconds) pairs, names as NUL-separated text in `names`. */
int zkg_msm_profile(zkg_msm* plan, const zkg_buffer* bases, const zkg_buffer* scalars, char* names,
                    size_t names_len, double* ms, size_t cap, size_t* count);
/* xyzz_to_affine (src/msm.rs:345-354): 32 result words -> plain affine x, y (16 words).
 * *is_finite = 0 for the point at infinity (the reference's None); out is then zeroed. */
int zkg_xyzz_to_affine(const uint32_t* words32, uint32_t* out_xy16, int* is_finite);
/* xyzz_to_affine_g2 (src/msm.rs:356-373): 64 result words -> (x.a, x.b, y.a, y.b). */
int zkg_xyzz_to_affine_g2(const uint32_t* words64, uint32_t* out_xy32, int* is_finite);

/* ------------------------------------------------------------------ NTT */

/* Ntt::new (src/ntt.rs:24-40): plan for 2^log_n; "log_n must be in 10..=27" otherwise. */
int zkg_ntt_create(zkg_device* dev, uint32_t log_n, zkg_ntt** out);
int zkg_ntt_destroy(zkg_ntt* plan);
/* Ntt::log_n (src/ntt.rs:21). */
uint32_t zkg_ntt_log_n(const zkg_ntt* plan);
/* Kernel launches (LDS-tiled passes) per transform of this plan. */
uint32_t zkg_ntt_passes(const zkg_ntt* plan);
/* Ntt::forward (src/ntt.rs:42-45): in place, natural order in and out. Errors:
 * "data buffer holds {} bytes; a 2^{} transform needs {}". */
int zkg_ntt_forward(zkg_ntt* plan, zkg_buffer* data);
/* Ntt::inverse (src/ntt.rs:47-50): in place, includes the 1/n scaling. */
int zkg_ntt_inverse(zkg_ntt* plan, zkg_buffer* data);

/* ------------------------------------------------------------------ batches
 * No counterpart in the reference (one call at a time). A batch queues many MSMs and NTTs of
 * uneven size as one unit: all NTT jobs run as grouped launches (one launch per pass phase, each
 * workgroup finding its transform in a table), every MSM job runs on its own HIP stream, and the
 * device stream waits for the whole batch, so it is ordered like one call. Jobs must use distinct
 * data buffers; an MSM plan may appear only once per batch (plans are not re-entrant). */
typedef struct zkg_batch zkg_batch;
int zkg_batch_create(zkg_device* dev, zkg_batch** out);
int zkg_batch_destroy(zkg_batch* batch);
/* In-place transform of the first 2^log_n elements of `data` (10 <= log_n <= 27), as Ntt::forward
 * / Ntt::inverse. Twiddle tables are built once per size and kept by the batch. */
int zkg_batch_add_ntt(zkg_batch* batch, zkg_buffer* data, uint32_t log_n, int inverse);
/* sum(scalars[i] * bases[i]) into `plan`'s result buffer, as zkg_msm_run (same size checks). */
int zkg_batch_add_msm(zkg_batch* batch, zkg_msm* plan, const zkg_buffer* bases, const zkg_buffer* scalars);
/* Queue every job (asynchronous); the batch keeps its jobs and can be run again. */
int zkg_batch_run(zkg_batch* batch);
int zkg_batch_clear(zkg_batch* batch);
/* Capture the whole batch (every NTT phase launch and every MSM stage, on all of the batch's
 * streams) into one hipGraph; zkg_batch_run then replays it with a single launch until the batch
 * is changed (add / clear drop the graph). The graph refers to the jobs' buffers and plans: they
 * must outlive it, and their contents are read at replay time. */
int zkg_batch_capture(zkg_batch* batch);
int zkg_batch_captured(const zkg_batch* batch);  /* 1 if a current graph exists */

/* ------------------------------------------------------------------ Groth16 quotient polynomial
 * No counterpart in the reference (README.md:229-230 names it as the next step around the NTT).
 * a, b, c hold the evaluations of A, B, C on the domain {w^i} of size n = 2^log_n (natural
 * order, plain words); run leaves in a the n coefficients of H = (A B - C) / (x^n - 1), computed
 * on the coset 5 H (3 inverse + 3 forward NTTs with the coset shift fused into the inverse's
 * scaling, a pointwise step, 1 inverse NTT with the unshift fused); b and c are clobbered.
 * When A B - C vanishes on the domain, H is the exact quotient (its top coefficient is 0). */
typedef struct zkg_groth16_h zkg_groth16_h;
int zkg_groth16_h_create(zkg_device* dev, uint32_t log_n, zkg_groth16_h** out);
/* The quotient in the form a proving key's H bases expect: ZKG_H_COEFFICIENTS (the above: H's
 * coefficients, for bases tau^i Z(tau)/delta) or ZKG_H_ODD_COSET (snarkjs keys: A B - C on the odd
 * coset w_2n H, a primitive 2n-th root times the domain, with no division; six transforms). */
enum { ZKG_H_COEFFICIENTS = 0, ZKG_H_ODD_COSET = 1 };
int zkg_groth16_h_create_form(zkg_device* dev, uint32_t log_n, int form, zkg_groth16_h** out);
int zkg_groth16_h_destroy(zkg_groth16_h* plan);
int zkg_groth16_h_run(zkg_groth16_h* plan, zkg_buffer* a, zkg_buffer* b, zkg_buffer* c);

/* ------------------------------------------------------------------ autotuning
 * No counterpart in the reference. A process-wide table, keyed by GPU architecture and size
 * (MSM: group and ceil(log2 n); NTT: log_n), holds the measured best MSM window c, work-item
 * size and counting-sort block size, and the NTT tile; zkg_msm_create_tuned, zkg_ntt_create and
 * the server read it. ZKGPU_TUNE_FILE names a table file to load at first use. The tuned choices
 * only change how work is split: every result is bit-identical to an untuned plan's. */
/* Times candidates on synthetic inputs of n points (coordinate search: c in default_window(n)
 * +-2 within 9..=16, then item size x0.5 / x2, then sort block x2 / x4), records and returns the
 * best (0 = automatic for item_points / sort_points) and its time in milliseconds. */
int zkg_tune_msm(zkg_device* dev, int group, size_t n, uint32_t* c, uint32_t* item_points, uint32_t* sort_points,
                 double* ms);
/* Times the LDS tiles 2^10, 2^11, 2^12 for a forward NTT of 2^log_n; records and returns the best. */
int zkg_tune_ntt(zkg_device* dev, uint32_t log_n, uint32_t* tile_log, double* ms);
int zkg_tune_lookup_msm(const char* arch, int group, size_t n, int* found, uint32_t* c, uint32_t* item_points,
                        uint32_t* sort_points);
int zkg_tune_lookup_ntt(const char* arch, uint32_t log_n, int* found, uint32_t* tile_log);
/* Text file: "msm <arch> <group> <ceil log2 n> <c> <item points> <sort points> <ms>" and
 * "ntt <arch> <log_n> <tile log2> <ms>" lines; '#' starts a comment line. load merges. */
int zkg_tune_load(const char* path);
int zkg_tune_save(const char* path);
int zkg_tune_clear(void);

/* ------------------------------------------------------------------ proof-job server
 * No counterpart in the reference. A device-resident request loop with continuous batching:
 * proving keys stay in HBM; jobs are admitted in arrival order as workspace slots and the HBM
 * budget allow (footprint per job: zkg_server_slot_bytes / zkg_server_key_slot_bytes), and
 * admitted jobs overlap on the device, each on its own streams. A job = 7 in-place NTTs of
 * 2^log_n elements (bit k of inverse_mask: polynomial k inverse) + MSMs A, B1 and B2 (G2) over
 * the witness, C over the witness from element c_offset, and H over the first outputs of
 * polynomial 6. Results: 192 words = XYZZ of A, B1, C, H (32 each), then B2 (64), Montgomery form. */
typedef struct zkg_server zkg_server;
typedef struct {
    uint64_t submitted, completed;
    double p50_ms, p99_ms, mean_ms, jobs_per_s; /* latency = submit -> results in host memory */
    size_t bytes_in_use, peak_bytes;
    uint32_t inflight, peak_inflight;
} zkg_server_stats;
/* A key's MSM sizes: points[] = A, B1, C, H (G1) and B2 (G2); H holds at most 2^log_n points. */
typedef struct {
    uint32_t log_n;
    size_t points[5];
    size_t c_offset; /* C's scalars start at this witness element */
    int h_form;      /* quotient jobs: ZKG_H_COEFFICIENTS or ZKG_H_ODD_COSET (H bases' convention) */
} zkg_key_shape;
int zkg_server_create(zkg_device* dev, size_t mem_budget_bytes, uint32_t max_inflight, zkg_server** out);
int zkg_server_destroy(zkg_server* server);
/* g1[0..3]: A, B1, C (n points each) and H (n - 1 points); g2: B2 (n points); all in the
 * library's form (zkg_msm_upload_bases*). The buffers must outlive the server. */
int zkg_server_add_key(zkg_server* server, uint32_t log_n, zkg_buffer* const g1[4], zkg_buffer* g2, int* key_id);
/* As zkg_server_add_key with explicit MSM sizes (bases buffers hold shape->points[] points). */
int zkg_server_add_key_shaped(zkg_server* server, const zkg_key_shape* shape, zkg_buffer* const g1[4],
                              zkg_buffer* g2, int* key_id);
/* Upload a snarkjs key's A, B1, C, H and B2 sections into server-owned HBM (freed with the
 * server) as a key of shape (log2 domainSize; nVars, nVars, nVars - nPublic - 1, domainSize,
 * nVars; c_offset = nPublic + 1): the witness is snarkjs's full assignment and polynomial 6's
 * outputs are the H scalars. Fails unless log2(domainSize) is in 10..=27, and (as zkg_zkey_upload)
 * when a section holds a point off the curve, or a B2 point outside the order-r subgroup. */
int zkg_server_add_zkey(zkg_server* server, const zkg_zkey* zkey, int* key_id);
int zkg_server_key_shape(zkg_server* server, int key_id, zkg_key_shape* out);
int zkg_server_slot_bytes(zkg_server* server, uint32_t log_n, size_t* bytes);
int zkg_server_key_slot_bytes(zkg_server* server, int key_id, size_t* bytes);
/* A quotient job: a, b, c hold A, B, C's evaluations on the domain (2^log_n elements each,
 * clobbered; a receives the H MSM's scalars in the key's H form, see zkg_groth16_h_create_form);
 * then the five MSMs as for zkg_server_submit. Keys from zkg_server_add_zkey use the odd-coset
 * form, zkg_server_add_key the coefficient form. */
int zkg_server_submit_groth16(zkg_server* server, int key_id, zkg_buffer* a, zkg_buffer* b, zkg_buffer* c,
                              const zkg_buffer* witness, uint64_t* job_id);
/* witness must hold max(points A, B1, B2, c_offset + points C) scalars. */
int zkg_server_submit(zkg_server* server, int key_id, zkg_buffer* const polys[7], uint32_t inverse_mask,
                      const zkg_buffer* witness, uint64_t* job_id);
int zkg_server_done(zkg_server* server, uint64_t job_id, int* done);
/* Blocks until the job finishes; results: 192 words (may be NULL); frees the job's record. */
int zkg_server_wait(zkg_server* server, uint64_t job_id, uint32_t* results, double* latency_ms);
int zkg_server_get_stats(zkg_server* server, zkg_server_stats* out);
int zkg_server_reset_stats(zkg_server* server);

/* ------------------------------------------------------------------ snarkjs .zkey ingestion
 * No counterpart in the reference (its README.md:224-225 leaves the snarkjs front end to another
 * project; src/msm.rs:190-197 accepts the Montgomery-2^256 bases such a key holds). Reads the
 * Groth16 proving-key layout snarkjs writes (version 1): the file is memory-mapped and checked
 * (magic, version, Groth16, BN254 q and r, section sizes). Point sections: 3 IC (nPublic + 1),
 * 5 A, 6 B1 (nVars G1), 7 B2 (nVars G2), 8 C (nVars - nPublic - 1), 9 H (domainSize). */
typedef struct {
    uint32_t n_vars, n_public, domain_size;
    /* verification key, plain affine words (infinity = zeros) */
    uint32_t alpha1[16], beta1[16], beta2[32], gamma2[32], delta1[16], delta2[32];
} zkg_zkey_info;
int zkg_zkey_open(const char* path, zkg_zkey** out);
int zkg_zkey_close(zkg_zkey* zkey);
int zkg_zkey_get_info(const zkg_zkey* zkey, zkg_zkey_info* out);
int zkg_zkey_points(const zkg_zkey* zkey, int section, size_t* count);
/* A point section as MSM bases (the stored words are the library's form: a copy reduced below p,
 * as zkg_msm_upload_bases_montgomery_256), then checked on the device as zkg_msm_check_bases does:
 * a section holding a point that is neither infinity nor on the curve (G1) / on the twist and in
 * the order-r subgroup (B2) is rejected ("the .zkey section {s} holds {k} point(s) off the curve
 * ..." with the first index) and no buffer is returned. Swapped coordinates or plain
 * (non-Montgomery) words fail this check. */
int zkg_zkey_upload(zkg_device* dev, const zkg_zkey* zkey, int section, zkg_buffer** out);

/* ------------------------------------------------------------------ test and benchmark hooks
 * No counterpart in the reference's public API. They play the role of the reference's WGSL test
 * harness (tests/curve.rs:8-30) and field benchmark (examples/field_bench.rs:10-97). */

/* Element-wise field op on `count` plain elements: field 0 = Fq, 1 = Fr, 2 = Fq2 (16 words);
 * op 0 = mul, 1 = add, 2 = sub, 3 = square. Synchronous; host arrays in and out. */
int zkg_test_field_op(zkg_device* dev, int field, int op, const uint32_t* a, const uint32_t* b, uint32_t* out,
                      size_t count);
/* Point ops on `count` pairs of plain affine points (infinity = zeros), results as plain affine:
 * op 0 = P+Q (XYZZ add), 1 = 2P (XYZZ double), 2 = P+Q (mixed add), 3 = (P+Q)+2P. */
int zkg_test_point_op(zkg_device* dev, int group, int op, const uint32_t* p, const uint32_t* q, uint32_t* out,
                      size_t count);
/* Field-multiply throughput: `threads` lanes each run `chain` dependent Montgomery products,
 * `ways` independent chains per lane. field 0 = Fq and 1 = Fr in 8 x 32-bit limbs (CIOS), 2 = Fq in
 * the MSM's 9 x 29-bit limbs (product scanning, R = 2^261); field 3 = Fq inversions in those limbs
 * (Fermat, 362 products each; `chain` dependent inversions per lane, `ways` ignored, each checked:
 * a failed x * x^-1 = 1 returns an error). Reports kernel milliseconds (HIP events). */
int zkg_bench_field_mul(zkg_device* dev, int field, uint32_t threads, uint32_t chain, uint32_t ways, double* ms);
/* A plain device-to-device copy kernel of `bytes` (16 per lane), run `reps` times; reports the
 * mean milliseconds per launch. The known-size reference for calibrating HBM byte counters. */
int zkg_bench_copy(zkg_device* dev, size_t bytes, uint32_t reps, double* ms);
/* The MSM's GLV scalar split on the device: out[0..4n) = k1 of each scalar, out[4n..8n) = k2,
 * each as |k| (127 bits, 4 words) with the sign in bit 127, such that s = k1 + k2 * lambda mod r
 * (lambda = 0xb3c4d79d41a917585bfc41088d8daaa78b17ea66b99c90dd). Synchronous; host arrays. */
int zkg_test_glv_split(zkg_device* dev, const uint32_t* scalars, size_t n, uint32_t* out);
/* The G2 MSM's GLS scalar split on the device (no counterpart in the reference, which has no such
 * split): out[12 i + 3 t], out[12 i + 3 t + 1] = |k_t| (low, high word), out[12 i + 3 t + 2] = the
 * sign of k_t (1 = negative), t = 0..3, with s_i = k0 + k1 lambda + k2 mu + k3 lambda mu mod r,
 * mu = p mod r, and |k_t| < 2^64. Synchronous; host arrays. */
int zkg_test_gls_split(zkg_device* dev, const uint32_t* scalars, size_t n, uint32_t* out);
/* Fixed-base multiples out[i] = ks[i] * generator (plain affine, infinity = zeros), for building
 * large synthetic inputs on the device. ks: n scalars (8 words each). */
int zkg_util_generator_multiples(zkg_device* dev, int group, const zkg_buffer* ks, size_t n, zkg_buffer* out);

#ifdef __cplusplus
}
#endif

#endif /* ZKGPU_H */
