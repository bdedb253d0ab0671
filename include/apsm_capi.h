/*
 * apsm_capi.h - C ABI of the MI355X (gfx950) APSM / LLS multi-user detector library
 * (libapsm_amd.so). Plain pointers and sizes only; no C++ or torch types.
 *
 * Every entry point stands in for a piece of the reference library's C++ API
 * (fraunhoferhhi/libapsm); the citation on each declaration names the reference
 * function it replaces. The drop-in C++ classes (NomaDetector, ApsmNomaDetector,
 * LlsNomaDetector, ...) are thin wrappers over these calls.
 *
 * Data layouts (the reference's host layouts, cpp/lib/noma/types.cuh:69-100):
 *   rx  : float [num_antennas][num_samples][2]   complex, antenna-major ([antenna][sample])
 *   tx  : float [num_users][num_samples][2]      complex pilots, one row per detected user
 *   est : float [num_users][num_samples][2]      complex soft estimates
 * The *_dev entry points take the same layouts in device memory and run on the
 * handle's stream without host synchronisation.
 *
 * Errors: every call returns APSM_OK (0) or a negative status; apsm_last_error()
 * returns the message of the last failure on the calling thread. Handles are not
 * thread-safe; distinct handles may be used from different threads.
 */
#ifndef APSM_CAPI_H
#define APSM_CAPI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define APSM_OK 0
#define APSM_ERR_ARG -1 /* invalid argument or configuration (reference: std::invalid_argument) */
#define APSM_ERR_DEVICE -2 /* HIP runtime failure (reference: CUDA_CHECK -> exit) */
#define APSM_ERR_NUMERIC -3 /* least-squares system not positive definite */
#define APSM_ERR_UNSUPPORTED -4 /* configuration the device path does not implement */
#define APSM_ERR_BUSY -5 /* admission refused: the HBM budget cannot hold it */

/* detectVersion values (cpp/lib/apsm/apsm_versions.cuh:75-89). All four select the same
 * dictionary-streaming kernel here; the names are accepted for drop-in compatibility. */
#define APSM_VERSION_DETECT_ORIGINAL 0x10001
#define APSM_VERSION_DETECT_OLDFAST 0x10002
#define APSM_VERSION_DETECT_SHMEM 0x10006
#define APSM_VERSION_DETECT_BALANCED 0x10007

/* detect_precision values: arithmetic of the detect cross terms x.z (no counterpart in the
 * reference, whose detect is FP32 on the CUDA cores). Squared norms, distances, exponentials,
 * the coefficient contraction and the linear term are FP32 in every mode; the 16-bit modes
 * round only the cross-term inputs and accumulate in FP32. */
#define APSM_PRECISION_FP32 0 /* FP32; the library picks the vector or matrix-core kernel by antenna count */
#define APSM_PRECISION_FP32_VECTOR 1 /* FP32 on the vector ALUs */
#define APSM_PRECISION_FP32_MATRIX 2 /* FP32 on the matrix cores */
#define APSM_PRECISION_BF16 3 /* BF16 inputs on the matrix cores, FP32 accumulation */
#define APSM_PRECISION_FP16 4 /* FP16 inputs on the matrix cores, FP32 accumulation */

/* Parameter set. Field meaning, units and defaults follow apsm_parameters
 * (cpp/lib/apsm/apsm_parameters.cuh:68-86) and the factory defaults
 * (cpp/lib/apsm/apsm_noma_detector.cu:361-372). window_size and dictionary_size count
 * complex samples; sample_step counts real rows. gaussian_kernel_variance is the
 * unscaled configuration value: the library applies var = v * N^2 when v != 0
 * (apsm_noma_detector.cu:79-80) and the pairwise-distance heuristic when v == 0
 * (apsm_noma_detector.cu:194-218). */
typedef struct apsm_params
{
    float gaussian_kernel_weight;
    float gaussian_kernel_variance;
    float eB;
    float norm_constraint;
    uint32_t window_size;
    uint32_t sample_step;
    uint32_t train_passes;
    uint32_t dictionary_size;
    uint8_t start_with_full_window;
    uint8_t lls_initialization;
    uint8_t training_shuffle;
    uint8_t reserved;
    int32_t detect_version;
    float linear_kernel_weight; /* w_L; the reference factory sets 1 - gaussian_kernel_weight */
    int32_t detect_precision; /* APSM_PRECISION_* (default APSM_PRECISION_FP32) */
} apsm_params;

typedef struct apsm_handle apsm_handle;
typedef struct lls_handle lls_handle;

/* Fills *p with the reference factory defaults (apsm_noma_detector.cu:361-372). */
void apsm_params_default(apsm_params* p);

/* Message of the last failed call on this thread ("" if none). */
const char* apsm_last_error(void);

/* Replaces ApsmNomaDetector<version>::ApsmNomaDetector(par, numAntennas)
 * (cpp/lib/apsm/apsm_noma_detector.cu:72-81). num_users > 1 trains that many detectors
 * on the same received signal in one handle (one pilot row per user in tx); they share
 * the Gaussian dictionary, which is made of the received pilots. */
int apsm_create(const apsm_params* p, uint32_t num_antennas, uint32_t num_users, apsm_handle** out);

/* Replaces ApsmNomaDetector::~ApsmNomaDetector (apsm_noma_detector.cu:83-87). */
void apsm_destroy(apsm_handle* h);

/* Replaces ApsmNomaDetector::resetState (apsm_noma_detector.cu:89-93). */
int apsm_reset(apsm_handle* h);

/* Replaces ApsmNomaDetector::train (apsm_noma_detector.cu:96-128 and the host loop
 * apsm_train.cu:80-166). Host rx [N][T][2], tx [U][T][2]. *ms (may be NULL) receives the
 * device time of the train loop, like the reference's event timer (apsm_train.cu:99,163-165).
 * A later call without apsm_reset continues the same pilot stream. */
int apsm_train(apsm_handle* h, const float* rx, const float* tx, uint32_t T, float* ms);

/* Replaces ApsmNomaDetector::detect (apsm_noma_detector.cu:131-143). Host rx [N][M][2],
 * est [U][M][2]. *ms receives the detect kernel time (apsm_detect.cu:972-974). */
int apsm_detect(apsm_handle* h, const float* rx, uint32_t M, float* est, float* ms);

/* Device-memory variants of apsm_train / apsm_detect (no counterpart in the reference,
 * which always uploads from host vectors, cuda_matrix_host.cu:105-126). Same layouts,
 * device pointers, enqueued on the handle's stream; no host synchronisation. */
int apsm_train_dev(apsm_handle* h, const float* d_rx, const float* d_tx, uint32_t T);
int apsm_detect_dev(apsm_handle* h, const float* d_rx, uint32_t M, float* d_est);

/* Device time (HIP events recorded on the handle's stream) of the most recent train
 * sequencer launch and the most recent detect launch, in ms; waits for those events.
 * Stands in for the event timer the reference returns from train/detect
 * (cpp/lib/cuda/cuda_eventtimer.cuh:73-133); covers the *_dev calls as well. */
int apsm_kernel_times(apsm_handle* h, float* train_ms, float* detect_ms);

/* Per-kernel device time of the handle's launches (no counterpart in the reference, whose
 * timer covers the whole train loop or detect kernel). While enabled, pairs of HIP events on
 * the handle's stream bracket every launch of the train sequencer (APSM_PROFILE_SEQUENCER,
 * apsm_train_seq2, one per train block), every block precompute (APSM_PROFILE_BLOCK_PRE: the
 * entry/Gram/eviction launch and the entry combine of a block) and every detect
 * (APSM_PROFILE_DETECT) of the blocked train path and the detect entry points.
 * apsm_profile_read waits for the recorded events and returns, per kind, the summed device
 * time (ms[APSM_PROFILE_KINDS]) and the number of bracketed launches (launches[...]) since the
 * previous read, then clears them. */
#define APSM_PROFILE_SEQUENCER 0
#define APSM_PROFILE_BLOCK_PRE 1
#define APSM_PROFILE_DETECT 2
#define APSM_PROFILE_KINDS 3
int apsm_profile_enable(apsm_handle* h, int enable);
int apsm_profile_read(apsm_handle* h, double* ms, uint64_t* launches);

/* Waits for all work enqueued on the handle's stream. */
int apsm_synchronize(apsm_handle* h);

/* The handle's stream as an opaque hipStream_t. */
void* apsm_stream(apsm_handle* h);

/* Training state readback, in the spirit of getTrainingState()
 * (apsm_noma_detector.cuh:93): stream length in real rows, dictionary size in complex
 * entries, the scaled kernel variance, theta [U][2N], and the real-row coefficients of the
 * dictionary [U][2*dict] in stream order (oldest first). Any output pointer may be NULL. */
int apsm_get_state(apsm_handle* h, uint64_t* used_rows, uint32_t* dict_entries, float* variance,
                   float* theta, float* coeffs, uint32_t coeffs_cap);

/* The dictionary's complex basis vectors (the basis of HostTrainingState,
 * cpp/lib/cuda/cuda_matrix_host.cuh:476-529) in the rx layout [N][dict_entries][2], oldest
 * first. dict may be NULL to query *dict_entries only. */
int apsm_get_dictionary(apsm_handle* h, float* dict, uint32_t cap_entries, uint32_t* dict_entries);

/* Switches the detect arithmetic of an existing handle (APSM_PRECISION_*); the trained state
 * is kept. No counterpart in the reference (its detect arithmetic is fixed). */
int apsm_set_detect_precision(apsm_handle* h, int32_t precision);

/* Replaces ApsmNomaDetector::computeNormRatio (cpp/lib/apsm/apsm_noma_detector.cu:220-319):
 * (w_L |theta|^2) / (w_G c^T K c) over the live dictionary, for one user of the handle. */
int apsm_norm_ratio(apsm_handle* h, uint32_t user, float* ratio);

/* Replaces ApsmNomaDetector::train with a validation set (NomaValSet,
 * cpp/lib/noma/noma_train_history.cuh:99; per-step metrics in apsm_train.cu:147-158): trains
 * on rx/tx like apsm_train, and after every train step detects the whole training set and the
 * validation set (rx_val [N][T_val][2], tx_val [U][T_val][2]) and records, per step and user,
 * history[step][user][0..5] = train EVM, SER, BER, validation EVM, SER, BER with the rules of
 * cpp/lib/noma/noma_train_history.cu:105-236 for a constellation of bits_per_symbol bits and
 * the given modulation scale. *num_steps receives the number of steps (<= history_cap).
 * Needs train_passes == 1 (APSM_ERR_UNSUPPORTED otherwise). *ms: device time of the loop. */
int apsm_train_history(apsm_handle* h, const float* rx, const float* tx, uint32_t T, const float* rx_val,
                       const float* tx_val, uint32_t T_val, uint32_t bits_per_symbol, float modulation_scale,
                       float* history, uint32_t history_cap, uint32_t* num_steps, float* ms);

/* LLS detector: replaces LlsNomaDetector (cpp/lib/noma/lls_detector.cu:97-218).
 * Normal equations accumulated on the device in FP32, solved by an FP64 Cholesky on the
 * device; detect is est = A w. lls_get_weights replaces getWeightMat() (2N floats). */
int lls_create(uint32_t num_antennas, lls_handle** out);
void lls_destroy(lls_handle* h);
int lls_train(lls_handle* h, const float* rx, const float* tx, uint32_t T, float* ms);
int lls_detect(lls_handle* h, const float* rx, uint32_t M, float* est, float* ms);
int lls_train_dev(lls_handle* h, const float* d_rx, const float* d_tx, uint32_t T);
int lls_detect_dev(lls_handle* h, const float* d_rx, uint32_t M, float* d_est);
int lls_get_weights(lls_handle* h, float* w);
int lls_synchronize(lls_handle* h);

/* LlsNomaDetector(numAntennas, std::shared_ptr<CudaStream> extStream)
 * (cpp/lib/noma/lls_detector.cu:97-107): the handle enqueues on the caller's stream and
 * does not destroy it. */
int lls_create_on_stream(uint32_t num_antennas, void* stream, lls_handle** out);

/* trainLowLevel (cpp/lib/noma/lls_detector.cu:157-179): least squares over a general
 * column-major device matrix A (num_samples x num_input_dims, leading dimension
 * num_samples) and device target y (num_samples). num_output_dims must be 1 (the only value
 * the reference passes); num_input_dims <= 128. A and y are not modified. */
int lls_train_lowlevel(lls_handle* h, const float* d_A, const float* d_y, int32_t num_input_dims,
                       int32_t num_output_dims, int32_t num_samples);

/* lls_train_lowlevel without any device allocation. No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc). Same arithmetic and kernels as lls_train_lowlevel (the same
 * weights for the same inputs), but no handle: the solved weights go to d_w (num_input_dims
 * floats) and the status to *d_status (0, or -1 when the normal matrix is not positive
 * definite), both device memory, enqueued on `stream` (a hipStream_t; NULL is the null stream)
 * with no host synchronisation. Its scratch is a fixed 32 MiB per device, held until the
 * process ends and shared by every call on that device (calls that overlap on one device
 * race); a problem whose partial sums need more (about nPart * n * (n+1) * 4 bytes with
 * nPart = ceil(num_samples / floor(65536 / (4 * (n+1)))) plus n * (n+1) * 8) is rejected with
 * APSM_ERR_ARG before anything is enqueued: about 3.9M rows at 32 unknowns, 64k rows at 128. */
int lls_train_lowlevel_noalloc(const float* d_A, const float* d_y, int32_t num_input_dims, int32_t num_output_dims,
                               int32_t num_samples, float* d_w, int32_t* d_status, void* stream);

/* detectLowLevel (cpp/lib/noma/lls_detector.cu:207-218): est = alpha * A w + beta * est on
 * the handle stream (est is not read when beta == 0). */
int lls_detect_lowlevel(lls_handle* h, const float* d_A, float* d_est, int32_t num_input_dims,
                        int32_t num_output_dims, int32_t num_samples, float alpha, float beta);

/* computeSignalPower (cpp/lib/noma/lls_detector.cu:236-242): sqrt(sum d^2 / n) of a device
 * array; synchronises the handle stream. */
int lls_signal_power(lls_handle* h, const float* d_data, uint32_t num_values, float* out);

/* getWeightMat() (cpp/lib/noma/lls_detector.cuh:92): device pointer of the solved weights. */
float* lls_weights_dev(lls_handle* h);

/* The handle's stream (hipStream_t), for callers that order their own work after it. */
void* lls_stream(lls_handle* h);

/* Batched instances (no counterpart in the reference, which runs one detector per host
 * thread, cpp/cli/detect/noma_detect.cpp:100-121): n independent detectors, each with its own
 * parameters, antenna count and users, run as one launch sequence. A run is, for every
 * instance: reset, train on its pilots, detect its data samples -- the same results as
 * apsm_reset + apsm_train_dev + apsm_detect_dev on a handle of its own.
 * apsm_batch_plan fixes the inputs (device pointers in the *_dev layouts, one per instance),
 * builds the launch plan, picks the detect launch shape per (antennas, users) bucket by timing
 * candidates on these inputs, and (use_graph != 0) captures the run in a hipGraph; it leaves
 * one completed run behind. apsm_batch_run enqueues a run (graph replay) on the batch stream.
 * Instances need the blocked train path (trainPasses 1, windowSize <= 160, sampleStep 2 or 4,
 * no startWithFullWindow, no normConstraint), a non-zero gaussianKernelVariance, no LLS
 * initialisation or shuffle, and at most 16 users (APSM_ERR_UNSUPPORTED otherwise). */
typedef struct apsm_batch apsm_batch;
int apsm_batch_create(uint32_t n, const apsm_params* params, const uint32_t* num_antennas, const uint32_t* num_users,
                      apsm_batch** out);
void apsm_batch_destroy(apsm_batch* b);
int apsm_batch_plan(apsm_batch* b, const float* const* d_rx_train, const float* const* d_tx_train, const uint32_t* T,
                    const float* const* d_rx_data, const uint32_t* M, float* const* d_est, int use_graph);
int apsm_batch_run(apsm_batch* b);
int apsm_batch_synchronize(apsm_batch* b);
/* Device time of the most recent run (HIP events around it on the batch stream), in ms. */
int apsm_batch_last_ms(apsm_batch* b, float* ms);
/* The same run split at the end of the train launches: train and detect device time, in ms. */
int apsm_batch_last_split(apsm_batch* b, float* train_ms, float* detect_ms);
/* Instance i's handle (state readback with apsm_get_state; owned by the batch). */
apsm_handle* apsm_batch_instance(apsm_batch* b, uint32_t i);
/* Autotuner choices, one row of 8 ints per detect bucket: (NMAX, UMAX, threads per workgroup,
 * workgroups per CU, work items, kernel: -1 vector ALU, else the matrix-core input precision
 * 0 FP32, 1 BF16, 2 FP16, dictionary entries per LDS tile, samples per workgroup tile). */
int apsm_batch_tuning(apsm_batch* b, int32_t* table, uint32_t cap_rows, uint32_t* rows);

/* Serving engine (no counterpart in the reference, whose tool detects one data set per run).
 * A server keeps links -- detectors whose state stays on the device between frames -- and
 * serves frames of them: detect the frame's data samples with the link's current state, then
 * train the link incrementally on the frame's pilots (the next frame of the link sees the
 * update). Frames of different links are served by continuous batching: every round detects
 * the frames admitted since the previous round in batched launches and advances every link
 * train in progress by one block, so frames that arrive while others are on the GPU join
 * the next round. Each result equals what apsm_detect followed by apsm_train on a handle of
 * its own computes (same kernels, same order). Links are admitted against hbm_budget (their
 * buffers are reserved for max_pilots pilots per frame; APSM_ERR_BUSY when it is spent) and
 * frames against the frame arena (a frame waits for room, in arrival order). Link
 * requirements are those of apsm_batch. Frame buffers are host memory (pinned for overlap);
 * est is written when the frame's result is reported. */
typedef struct apsm_server apsm_server;
typedef struct apsm_server_config
{
    uint64_t hbm_budget; /* device bytes for link states and the frame arena together */
    uint64_t frame_arena; /* device bytes of the frame arena (part of hbm_budget) */
    /* Detect batching (0 selects the default): a batch of newly admitted frames closes at
     * batch_max_frames frames (default 32) or batch_max_bytes bytes of frame data (default
     * 64 MiB); at most detects_in_flight batches (1..8, default 4) are on the device at once. */
    uint32_t batch_max_frames;
    uint32_t detects_in_flight;
    uint64_t batch_max_bytes;
} apsm_server_config;
typedef struct apsm_frame
{
    double arrival_s; /* apsm_server_replay: arrival time relative to the start of the replay */
    uint32_t link;
    uint32_t n_data; /* samples to detect */
    uint32_t n_pilots; /* pilots trained on after the detect (0: no train) */
    uint32_t reserved;
    const float* rx_data; /* [N][n_data][2] */
    const float* rx_pilots; /* [N][n_pilots][2] */
    const float* tx_pilots; /* [U][n_pilots][2] */
    float* est; /* [U][n_data][2] */
} apsm_frame;
typedef struct apsm_frame_result
{
    int32_t status;
    uint32_t reserved;
    uint64_t ticket;
    double arrival_s, admit_s; /* seconds on the server clock (replay: from its start) */
    double output_s; /* estimates in est (the frame's latency is output_s - arrival_s) */
    double trained_s; /* the link's train on the frame's pilots finished */
    int64_t detect_round, last_round;
} apsm_frame_result;
int apsm_server_create(const apsm_server_config* cfg, apsm_server** out);
void apsm_server_destroy(apsm_server* s);
int apsm_server_add_link(apsm_server* s, const apsm_params* p, uint32_t num_antennas, uint32_t num_users, uint32_t max_pilots,
                         uint32_t* link);
/* The link's detector handle (owned by the server) for state readback. */
apsm_handle* apsm_server_link(apsm_server* s, uint32_t link);
/* Queues a frame (arrival = now); *ticket identifies its result. */
int apsm_server_submit(apsm_server* s, const apsm_frame* frame, uint64_t* ticket);
/* Makes progress (retires finished rounds, launches the next) and reports finished frames,
 * up to cap; wait != 0 keeps going until every submitted frame has finished. A reported
 * frame's est is complete in host memory (pinned or pageable) when the call returns. On a
 * failure every frame still held is dropped (not reported) and the server is idle again. */
int apsm_server_poll(apsm_server* s, int wait, apsm_frame_result* out, uint32_t cap, uint32_t* n);
/* Serves n frames arriving at frames[i].arrival_s (non-decreasing) after the call starts, on
 * the host's clock, and returns when all have finished; results[i] belongs to frames[i].
 * Every frame is checked before any is enqueued (APSM_ERR_ARG: nothing ran); on a failure
 * during the run the frames are dropped and the server is idle again. */
int apsm_server_replay(apsm_server* s, const apsm_frame* frames, uint32_t n, apsm_frame_result* results);
/* Rounds launched so far, the largest number of frames one round carried, device bytes of
 * the links. */
int apsm_server_stats(apsm_server* s, uint64_t* rounds, uint64_t* max_frames_per_round, uint64_t* link_bytes);

/* Pinned host memory (for frame buffers whose copies overlap the device work). */
int apsm_host_malloc(void** ptr, uint64_t bytes);
int apsm_host_free(void* ptr);

/* CudaStream (cpp/lib/cuda/cuda_stream.cuh:67-96): a HIP stream of default priority. */
int apsm_stream_create(void** out);
int apsm_stream_destroy(void* stream);
int apsm_stream_synchronize(void* stream);

/* Number of visible HIP devices (0 when no GPU is present). Never fails. */
int apsm_device_count(void);

/* Device memory plumbing for callers without a device runtime of their own (tests, the
 * Python mirror, the benchmark); the reference allocates inside its matrix classes
 * (cpp/lib/cuda/cuda_matrix_host.cuh). Copies are synchronous with the whole device: a
 * download waits for the work of every stream, an upload completes before it returns. */
int apsm_device_malloc(void** ptr, uint64_t bytes);
int apsm_device_free(void* ptr);
int apsm_memcpy_h2d(void* dst, const void* src, uint64_t bytes);
int apsm_memcpy_d2h(void* dst, const void* src, uint64_t bytes);
int apsm_device_synchronize(void);

#ifdef __cplusplus
}
#endif

#endif /* APSM_CAPI_H */
