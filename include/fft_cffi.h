/* fft_cffi.h -- C ABI of the MI355X (gfx950) batched FFT/STFT library, libfft_cffi.so.
 *
 * Drop-in for the reference library's C API (fft_cffi.hpp): the entry points marked
 * "reference" keep the reference's names, signatures, argument meaning and error
 * behaviour; each cites the declaration it replaces. The entry points marked "additive"
 * have no reference counterpart; they take device pointers (for roofline timing and
 * torch callers) or expose data the reference only offers through its C++ API.
 *
 * Pointers in reference entry points are HOST pointers, exactly as in the reference, and
 * those calls return when the results are in the caller's buffers. "_device" entry points
 * take device pointers and a hipStream_t passed as void* (NULL = the default stream); they
 * only enqueue work on that stream and return without waiting for it.
 *
 * All functions are safe to call with a NULL handle (they return 0 / do nothing).
 * A context is not re-entrant: concurrent calls on one context race, separate
 * contexts are independent (reference: fft_cffi.hpp:99-176, eigen_fft.hpp:1418).
 */
#ifndef FFTFREE_AMD_FFT_CFFI_H
#define FFTFREE_AMD_FFT_CFFI_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Kernel, radix and transform constants (fft_cffi.hpp:62-81; values follow the
 * eigfft::KernelKind / ButterflyRadix enum order, eigen_fft.hpp:1892,
 * butterfly_kernel.hpp:16-21). At the C level the kernel argument is read as
 * 2 = Stockham and anything else = Cooley-Tukey (fft_cffi.cpp:799-800). */
enum { FFT_KERNEL_COOLEYTUKEY = 0, FFT_KERNEL_STOCKHAM = 1, FFT_KERNEL_EXTERNAL = 2 };
enum { FFT_BUTTERFLY_RADIX_2 = 0, FFT_BUTTERFLY_RADIX_4 = 1, FFT_BUTTERFLY_RADIX_8 = 2, FFT_BUTTERFLY_RADIX_16 = 3 };
enum { FFT_TRANSFORM_C2C = 0, FFT_TRANSFORM_R2C = 1, FFT_TRANSFORM_C2R = 2, FFT_TRANSFORM_R2R = 3 };
/* Window kinds, normalisation and COLA modes (fft_cffi.hpp:292-313). */
enum { FFT_WINDOW_RECT = 0, FFT_WINDOW_HANN = 1, FFT_WINDOW_HAMMING = 2, FFT_WINDOW_BLACKMAN = 3,
       FFT_WINDOW_TUKEY = 4, FFT_WINDOW_KAISER = 5 };
enum { FFT_WINDOW_NORM_NONE = 0, FFT_WINDOW_NORM_L2 = 1, FFT_WINDOW_NORM_AREA = 2 };
enum { FFT_COLA_OFF = 0, FFT_COLA_NORMALIZE = 1 };
enum { FFT_STREAM_FLUSH_NONE = 0, FFT_STREAM_FLUSH_PARTIAL = 1, FFT_STREAM_FLUSH_FINAL = 2 };
/* Metadata kinds of fft_plan_metadata (eigfft::MetadataKind order, eigen_fft.hpp:1477-1491). */
enum { FFT_META_TWIDDLE = 0, FFT_META_LAYOUT_MAP = 1, FFT_META_STAGE_PARAMS = 3, FFT_META_BUTTERFLY_PAIRS = 4,
       FFT_META_TWIDDLE_INDEX_MAP = 6, FFT_META_OMEGA = 7 };

/* ---- reference entry points -------------------------------------------------------- */

/* reference: fft_cffi.hpp:31 (fft_cffi.cpp:431-514). One forward FFT of n samples padded
 * to the next power of two; writes the first n bins. */
void fft_pcm_to_channels(const float* in_pcm, float* out_real, float* out_imag, float* out_mag, size_t n,
                         int threads);

/* reference: fft_cffi.hpp:99 (fft_cffi.cpp:535-538). */
void* fft_init(size_t n, int threads, int lanes, int inverse, int kernel, int radix, const int* radix_pattern,
               size_t radix_pattern_len, int pad_mode);

/* reference: fft_cffi.hpp:105 (fft_cffi.cpp:540-658). Pads N to a power of two only when
 * pad_mode == 1; honours FFTFREE_DEFAULT_KERNEL when kernel == 0. */
void* fft_init_ex(size_t n, int threads, int lanes, int inverse, int kernel, int radix, const int* radix_pattern,
                  size_t radix_pattern_len, int pad_mode, int window, int hop);

/* reference: fft_cffi.hpp:108-127 (fft_cffi.cpp:1792-1822). Never windows. */
void* fft_init_full(size_t n, int threads, int lanes, int inverse, int kernel, int radix, const int* radix_pattern,
                    size_t radix_pattern_len, int pad_mode, int window, int hop, int transform,
                    int reduce_magnitude, int store_polar, int half_spectrum, int allow_outer_parallel,
                    int allow_inner_parallel, int inner_threads, int save_crash_logs, int silent_crash_reports);

/* reference: fft_cffi.hpp:130-150 (fft_cffi.cpp:1824-1978). NULL unless n is a power of
 * two. Crash-handler flags are accepted and ignored. */
void* fft_init_full_v2(size_t n, int threads, int lanes, int inverse, int kernel, int radix,
                       const int* radix_pattern, size_t radix_pattern_len, int pad_mode, int window, int hop,
                       int transform, int reduce_magnitude, int store_polar, int half_spectrum,
                       int allow_outer_parallel, int allow_inner_parallel, int inner_threads, int save_crash_logs,
                       int silent_crash_reports, int apply_windows, int apply_ola, int analysis_window_kind,
                       float analysis_param1, float analysis_param2, int synthesis_window_kind,
                       float synthesis_param1, float synthesis_param2, int window_norm_policy, int cola_mode);

/* reference: fft_cffi.hpp:153-158 (fft_cffi.cpp:660-699). One frame of n <= N samples;
 * returns 1 on success, 0 on error. */
int fft_execute(void* handle, const float* in_pcm, float* out_real, float* out_imag, float* out_mag, size_t n);

/* reference: fft_cffi.hpp:168-176 (fft_cffi.cpp:701-1034). Frames pcm with the context's
 * window/hop, pads per pad_mode (0 = context setting), writes max_frames*bins floats per
 * plane at most; returns frames produced, 0 on error. enable_backup is accepted and has
 * no effect on results (the reference's recovery path computes the same values). */
size_t fft_execute_batched(void* handle, const float* pcm, size_t pcm_len, float* out_real, float* out_imag,
                           float* out_mag, int pad_mode, int enable_backup, size_t max_frames);

/* reference: fft_cffi.hpp:181-187 (fft_cffi.cpp:1458-1466) = batched with pad_mode 1. */
size_t fft_stft_execute(void* handle, const float* pcm, size_t pcm_len, float* out_real, float* out_imag,
                        float* out_mag, size_t max_frames);

/* reference: fft_cffi.hpp:228-235 (fft_cffi.cpp:1042-1340). Complex frame-major bins in,
 * N real samples per frame out (times the synthesis window when windows are enabled),
 * optional internal overlap-add when the context was created with apply_ola. */
size_t fft_execute_complex_batched(void* handle, const float* in_real, const float* in_imag, size_t frames,
                                   float* out_pcm, int pad_mode, int enable_backup, size_t max_frames);

/* reference: fft_cffi.hpp:237-241 (fft_cffi.cpp:1744-1773). */
size_t fft_ctx_size(void* handle);
size_t fft_ctx_worker_threads(void* handle);
size_t fft_ctx_effective_threads(void* handle);

/* reference: fft_cffi.hpp:280 (fft_cffi.cpp:1775-1779). fft_free(NULL) is a no-op. */
void fft_free(void* handle);

/* reference: fft_cffi.hpp:329-345 (fft_cffi.cpp:1983-2032). Rebuild (or replace) the
 * analysis/synthesis windows; like the reference this leaves internal windowing off. */
int fft_config_window(void* handle, int analysis_kind, int synthesis_kind, float param1, float param2,
                      int norm_policy, int cola_mode);
int fft_config_window_custom(void* handle, const float* analysis_window, size_t analysis_len,
                             const float* synthesis_window, size_t synthesis_len, int norm_policy, int cola_mode);

/* reference: fft_cffi.hpp:246-266, 286-287. The reference's CPU worker-pool fault
 * injection hooks; there is no worker pool here, so these record nothing and the
 * observer is never called. Present so existing callers link. */
void fft_set_workerpool_debug_dropout(int pct);
void fft_set_workerpool_dropout_pattern(int period, int on_len, int phase);
void fft_set_workerpool_dropout_single(long long index);
void fft_set_workerpool_dropout_persistent(int on);
void fft_clear_workerpool_dropout_history(void);
typedef void (*fft_restore_observer_t)(size_t start, size_t end);
void fft_register_restore_observer(fft_restore_observer_t cb);

/* ---- additive entry points --------------------------------------------------------- */

/* additive. Frames fft_execute_batched would produce for pcm_len samples (the
 * reference's rule, fft_cffi.cpp:715-750); 0 = refused. Host-only, no GPU work. */
size_t fft_frame_count(void* handle, size_t pcm_len, int pad_mode, size_t max_frames);

/* additive. fft_execute_batched on device-resident buffers: d_pcm holds pcm_len floats,
 * the planes hold frames*bins floats each; any plane may be NULL to skip writing it. */
size_t fft_execute_batched_device(void* handle, const float* d_pcm, size_t pcm_len, float* d_out_real,
                                  float* d_out_imag, float* d_out_mag, int pad_mode, size_t max_frames,
                                  void* stream);

/* additive. Batched complex transform of `batch` columns of N interleaved complex64
 * values (the C++ fft_inplace_batched path, eigen_fft.hpp:3152-3175, including the
 * plan's pre/post hooks). in and out are the same buffer or do not overlap. N above
 * 4096 runs as two passes through a device scratch of batch*N complex values owned by
 * the context (grown on first use). Host pointers / device pointers. */
size_t fft_execute_c2c_batched(void* handle, const float* in, float* out, size_t batch);
size_t fft_execute_c2c_batched_device(void* handle, const float* d_in, float* d_out, size_t batch, void* stream);

/* additive. fft_execute_complex_batched (fft_cffi.cpp:1042-1157) on device-resident
 * buffers: d_in_real / d_in_imag hold frames*bins floats (frame-major), d_out_pcm
 * frames*N floats, written in full as the reference writes them (overlap-added output
 * followed by zeros when the context has apply_ola). pad_mode and enable_backup of the
 * host entry have no effect there either and are not taken. Returns frames, 0 on error.
 * Enqueued on `stream`; a context whose flags the fused kernel does not cover uses a
 * device scratch of frames*N complex values owned by the context. */
size_t fft_execute_complex_batched_device(void* handle, const float* d_in_real, const float* d_in_imag,
                                          size_t frames, float* d_out_pcm, size_t max_frames, void* stream);

/* reference: fft_cffi.hpp:269-277 (fft_cffi.cpp:2058-2174, phase_infer.hpp:23-66). Phase
 * inference from a frame-major magnitude spectrogram (frames x bins, bins = N/2+1 with
 * half_spectrum, else N). mode 0: the linear phase advance phi[f][k] = f * 2 pi k hop / N
 * (accumulated in double, as the reference does); mode 1: the minimum-phase spectrum of
 * each frame through the real cepstrum (N a power of two). phase_infer_execute returns
 * frames, or 0 (unknown mode, bad arguments). Host pointers; phase_free(NULL) is a no-op. */
void* phase_init(int N, int hop, int half_spectrum, int mode, int iterations);
size_t phase_infer_execute(void* handle, const float* in_mag, size_t frames, float* out_real, float* out_imag);
void phase_free(void* handle);

/* additive. phase_infer_execute on device buffers, enqueued on `stream`. */
size_t phase_infer_execute_device(void* handle, const float* d_in_mag, size_t frames, float* d_out_real,
                                  float* d_out_imag, void* stream);

/* reference: fft_cffi.hpp:188-211 (fft_cffi.cpp:1342-1742). Streaming STFT on a context.
 * fft_stream_push_pcm appends `samples` to the context's buffer and writes the frames that
 * are now complete (at most max_frames, 0 = all; hop-aligned from the stream start,
 * partial frames kept for later) to the planes (host pointers, frames*bins floats each;
 * NULL planes are skipped); returns the frame count. flush_mode is accepted and, as in the
 * reference, not used. fft_stream_push_spectrum (inverse contexts) transforms `frames`
 * frames of bins, overlap-adds them into the context's accumulator (COLA-normalised when
 * the context's cola_mode is FFT_COLA_NORMALIZE), queues `hop` finished samples per frame
 * (all N of the last frame, or the whole accumulator when frames == 0, with
 * FFT_STREAM_FLUSH_FINAL), and writes up to max_samples (0 = all) queued samples to
 * out_pcm; returns the sample count written. The reference prints a line to stderr on
 * every fft_stream_push_pcm call; here that line needs FFTFREE_VERBOSE. */
void fft_stream_reset(void* handle);
size_t fft_stream_push_pcm(void* handle, const float* pcm, size_t samples, float* out_real, float* out_imag,
                           float* out_mag, size_t max_frames, int flush_mode);
size_t fft_stream_pending_frames(void* handle);
size_t fft_stream_backlog_samples(void* handle);
size_t fft_stream_push_spectrum(void* handle, const float* in_real, const float* in_imag, size_t frames,
                                float* out_pcm, size_t max_samples, int flush_mode);
size_t fft_stream_pending_pcm(void* handle);

/* reference: fft_cffi.hpp:213-223 (fft_cffi.cpp:2176-2248, phase_infer.hpp:69-204).
 * Griffin-Lim phase reconstruction of a frame-major magnitude spectrogram (frames x
 * bins): a random start (mt19937 seeded with `seed`, 0x47d7bfad when 0), then
 * `iterations` rounds (32 when <= 0) of inverse transform with ctx_inverse, overlap-add
 * averaged over the frames per sample (hop: the argument, else ctx_inverse's, else
 * ctx_forward's, else N), forward STFT with ctx_forward and the phase update
 * X = mag * (cos, sin)(arg X). Writes the final spectrum to out_real / out_imag (frames x
 * bins). Returns 1, or 0 on bad arguments (NULL or aliased buffers, half_spectrum
 * differing from either context, contexts of different N). Host pointers. */
int fft_griffin_lim(void* ctx_forward, void* ctx_inverse, const float* in_mag, size_t frames, int hop,
                    int half_spectrum, int iterations, int pad_mode, unsigned int seed, float* out_real,
                    float* out_imag);

/* additive. fft_griffin_lim on device buffers, enqueued on `stream` (the mt19937 words
 * of the start are drawn on the host and copied in before the call returns). Scratch
 * lives in ctx_inverse. */
int fft_griffin_lim_device(void* ctx_forward, void* ctx_inverse, const float* d_in_mag, size_t frames, int hop,
                           int half_spectrum, int iterations, int pad_mode, unsigned int seed, float* d_out_real,
                           float* d_out_imag, void* stream);

/* additive. Batched STFT of `signals` signals of pcm_len samples each (signal s at
 * d_pcm + s*pcm_len), for a forward R2C half-spectrum context with N in {512, 1024,
 * 2048}: every signal is framed with the context's window/hop/pad rule (frames per
 * signal = fft_frame_count) and its planes, as fft_execute_batched writes them, go to
 * d_out_* + s*frames*(N/2+1). Any plane may be NULL. Returns frames per signal, 0 on
 * error. Device pointers; enqueued on `stream`. */
size_t fft_stft_signals_device(void* handle, const float* d_pcm, size_t signals, size_t pcm_len, float* d_out_real,
                               float* d_out_imag, float* d_out_mag, int pad_mode, void* stream);

/* additive. Adjoint of fft_stft_signals_device: given the gradients of a loss with
 * respect to the first two output planes (re and im; |X| and arg X when store_polar;
 * |X| when reduce_magnitude, where d_grad_imag is ignored), laid out like those planes,
 * writes dL/dpcm (signals*pcm_len floats) to d_grad_pcm. Each frame's spectrum is
 * recomputed from d_pcm. Either gradient may be NULL (zero). Needs window + k*hop <= 8N
 * (k = 2 for N = 2048, 4 otherwise).
 * Deterministic (fixed summation order). Returns frames per signal, 0 on error. */
size_t fft_stft_adjoint_device(void* handle, const float* d_pcm, size_t signals, size_t pcm_len,
                               const float* d_grad_real, const float* d_grad_imag, float* d_grad_pcm, int pad_mode,
                               void* stream);

/* additive. The same adjoint from the forward's own output planes instead of the samples:
 * d_plane_real / d_plane_imag are the first two planes fft_stft_signals_device wrote for
 * this context and signals (|X| and arg X when store_polar; not read for complex planes,
 * may be NULL there). V = dL/d|X| e^{i arg X} + i dL/d(arg X) e^{i arg X} / |X| (0 where
 * |X| = 0) or dL/dre + i dL/dim, then one inverse real DFT per frame, the analysis window
 * and the overlap-add: no forward transform is recomputed. A magnitude-only context is
 * refused (its planes do not hold the direction of X). The same window/hop limits as
 * fft_stft_adjoint_device apply. Either gradient may be NULL (zero). Deterministic (fixed
 * summation order). Returns frames per signal, 0 on error. */
size_t fft_stft_adjoint_saved_device(void* handle, const float* d_plane_real, const float* d_plane_imag,
                                     size_t signals, size_t pcm_len, const float* d_grad_real,
                                     const float* d_grad_imag, float* d_grad_pcm, int pad_mode, void* stream);

/* No counterpart in the reference: an allocation-free variant of fft_stft_adjoint_device,
 * for callers that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc). Same arguments and results; it calls
 * no hipMalloc, hipFree or allocating routine. Its halo scratch is a static 16 MiB device
 * array per GPU (4M floats), shared by every call on that GPU and kept until the process
 * ends: calls that overlap on one GPU race. A shape that needs more halo scratch is
 * refused (0, with a message on stderr). The context must already hold its device state:
 * make one allocating device call on it first (e.g. fft_stft_signals_device, which a
 * forward/backward pair makes anyway); otherwise it is refused. */
size_t fft_stft_adjoint_device_noalloc(void* handle, const float* d_pcm, size_t signals, size_t pcm_len,
                                       const float* d_grad_real, const float* d_grad_imag, float* d_grad_pcm,
                                       int pad_mode, void* stream);

/* additive. One request of a ragged batch (device pointers): pcm_len samples, framed with
 * window = N (a power of two from 256 to 4096) and hop; output planes of frames*(N/2+1)
 * floats each, as fft_execute_batched writes them for a R2C half-spectrum context (any
 * may be NULL). fft_ragged_create writes `frames`. */
typedef struct fft_ragged_request {
  const float* pcm;
  size_t pcm_len;
  int window;
  int hop;
  float* out_real;
  float* out_imag;
  float* out_mag;
  size_t frames;
} fft_ragged_request;

/* additive. A ragged batch: every request gets the same flags (pad_mode as in
 * fft_execute_batched, reduce_magnitude / store_polar output, analysis window kind when
 * apply_windows). The buffers are fixed: fft_ragged_execute reads and writes the ones
 * given here, so refill them between executions. Returns NULL on bad requests. */
void* fft_ragged_create(fft_ragged_request* requests, size_t count, int pad_mode, int reduce_magnitude,
                        int store_polar, int apply_windows, int window_kind);
/* additive. Times every engine of each N of the batch (the tile shapes: radix pattern and
 * frames per tile; the line-aligned engine where every request of N qualifies) at 1, 2, 4
 * and 8 blocks per work item on the batch itself (writing its outputs), by the median of
 * `repeats` runs, keeps a new choice only when it is more than 3% faster than the current
 * one, and rebuilds the work lists and graph. Returns 1 on success. */
int fft_ragged_autotune(void* ragged, int repeats);
/* additive. Sets the engine of window size n (an index below fft_ragged_engine_count: a
 * tile shape, or a line-aligned engine) and the blocks per work item (>= 1), and rebuilds
 * the graph. Returns 0 when the engine does not run size n or, for a line-aligned engine,
 * some request of size n does not qualify (window != N, odd hop, buffers not aligned). The
 * line-aligned engine of N = 2048 runs one block per workgroup and ignores
 * blocks_per_item. */
int fft_ragged_set_engine(void* ragged, int n, int engine, int blocks_per_item);
int fft_ragged_engine_count(void);
/* additive. The whole batch as one graph launch (one launch per window size: persistent
 * workgroups walking work items, or at N = 2048 one workgroup per block) on `stream`. */
int fft_ragged_execute(void* ragged, void* stream);
/* additive. The same outputs with one plain launch per request (the unbatched baseline). */
int fft_ragged_execute_plain(void* ragged, void* stream);
/* additive. Text summary of the grid, item count and per-N choices. */
int fft_ragged_describe(void* ragged, char* buf, size_t buflen);
void fft_ragged_free(void* ragged);

/* additive. Index metadata of the context's plan (MetadataRequest kinds of the C++ API,
 * eigen_fft.hpp:1496-1501), written by a GPU index kernel as complex<float> pairs into
 * the host buffer `out` (capacity in complex entries). Returns entries written; 0 if the
 * kind is not recorded by the plan's kernel or capacity is too small. */
size_t fft_plan_metadata(void* handle, int kind, float* out, size_t capacity);

/* additive. Selects the kernel variant (radix pattern / frames per workgroup) for the
 * context's size; -1 restores the default. Returns 1 if the variant exists. The count
 * and a description ("8x8x8 fpw=4") are queryable for autotuners. R2C and C2R contexts
 * select among the fused STFT kernels; C2C contexts among the complex-transform kernels
 * that their batched complex entries run. */
int fft_set_variant(void* handle, int variant);
int fft_variant_count(void* handle);
int fft_variant_name(void* handle, int variant, char* buf, size_t buflen);
/* additive. The variant index a launch of `frames` frames (STFT contexts) or of a batch
 * (complex contexts) runs on this context: the one fft_set_variant chose, else the
 * default for that shape on the current device; -1 for a NULL handle. The batched STFT
 * of a C2C context always runs its default STFT variant: fft_set_variant on a C2C
 * context tunes the complex-transform kernels. */
int fft_variant_default(void* handle, size_t frames);

#ifdef __cplusplus
}
#endif

#endif /* FFTFREE_AMD_FFT_CFFI_H */
