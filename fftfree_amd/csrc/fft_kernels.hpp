// fft_kernels.hpp -- host-side view of the HIP kernels: argument blocks and launchers.
// Everything here takes plain device pointers and a hipStream_t; no torch types.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace fftfree {

// Real-input batched STFT (fft_execute_batched semantics, fft_cffi.cpp:701-825):
// frame f covers pcm[f*hop, f*hop + window), times the analysis window, zero-padded to
// N; output planes are frame-major with `bins` floats per frame.
struct StftArgs {
  const float* pcm = nullptr;
  size_t pcm_len = 0;
  const float* win = nullptr;  // `window` taps, or nullptr when windows are off
  int window = 0;              // W <= N
  int hop = 0;
  size_t frames = 0;
  const float2* tw = nullptr;  // forward table exp(-2 pi i m / N), m < N
  int n = 0;                   // N (power of two)
  int bins = 0;                // N/2+1 or N
  int out_mode = 0;            // dev::OutMode
  float scale = 1.0f;          // 1 or 1/N
  float* o_re = nullptr;       // any plane may be nullptr (device entry only)
  float* o_im = nullptr;
  float* o_mag = nullptr;
  size_t frame_offset = 0;     // first frame index (ragged / chunked callers)
  float* sink = nullptr;       // N floats of scratch that skipped planes are written to
  size_t lead = 0;             // multi-frame kernel: frames covered by whole tiles
  int rest_rank = 0;           // persistent multi-frame kernel: the rank that runs the rest
};

// Complex batched transform (fft_inplace_batched semantics, eigen_fft.hpp:3152-3175):
// `batch` transforms of N complex values, transform-major, pre hook on load and post
// hook on store (eigen_fft.hpp:2484-2596). in may equal out.
struct C2CArgs {
  const float2* in = nullptr;
  float2* out = nullptr;
  size_t batch = 0;
  const float2* tw = nullptr;  // forward table exp(-2 pi i m / N), m < N
  int n = 0;
  int pre_mode = 0;   // 0 none, 1 clear imag (R2C), 2 Hermitian rebuild (C2R half)
  int pre_polar = 0;  // with pre_mode 2: bins hold (mag, phase)
  int post_half = 0;  // R2C half: DC/Nyquist imag = 0, bins > N/2 zero
  int post_mode = 0;  // 0 none, 1 C2R imag clear, 2 magnitude, 3 polar
  int post_upper = 0; // bins the magnitude/polar reduction covers
  float scale = 1.0f;
};

// Batched real-input STFT of `signals` equal-length signals (signal s at pcm + s*len)
// and its adjoint; planes and upstream gradients are [signal][frame][N/2+1] floats.
struct SignalArgs {
  const float* pcm = nullptr;
  size_t len = 0;              // samples per signal
  size_t signals = 0;
  size_t frames = 0;           // frames per signal
  const float* win = nullptr;  // N taps (taps >= window are 1), or nullptr when windows are off
  int window = 0;
  int hop = 0;
  int n = 0;
  const float2* tw = nullptr;  // forward table exp(-2 pi i m / N), m < N
  float* o_re = nullptr;       // forward planes (any may be nullptr)
  float* o_im = nullptr;
  float* o_mag = nullptr;
  const float* g_mag = nullptr;    // adjoint: gradient of plane re (|X| when polar/magnitude)
  const float* g_phase = nullptr;  // adjoint: gradient of plane im (arg X when polar), may be nullptr
  // adjoint from the forward's saved polar planes (launch_stft_adjoint with `saved`):
  // V = e^{i p_phase} (g_mag + i g_phase / p_mag), 0 where p_mag = 0; pcm is not read
  const float* p_mag = nullptr;
  const float* p_phase = nullptr;
  float* grad = nullptr;           // adjoint output: dL/dpcm, signals*len floats
  int tiles = 0;                   // set by the launchers: workgroups per signal
  int steps = 0;                   // adjoint: steps per workgroup
  float* halo = nullptr;           // adjoint: boundary partial sums (2 slots of W-H per run)
  size_t halo_capacity = 0;        // floats
};
// Device scratch (floats) the adjoint needs for `signals` signals of pcm_len samples.
size_t adjoint_scratch(int n, int window, int hop, size_t signals, size_t pcm_len);

// Transform of frame-major complex bins with synthesis window and overlap-add
// (fft_execute_complex_batched semantics, fft_cffi.cpp:1042-1157).
struct IstftArgs {
  const float* in_re = nullptr;  // frames x bins
  const float* in_im = nullptr;
  size_t frames = 0;
  int n = 0;
  int bins = 0;
  int hop = 0;                   // overlap-add hop
  const float2* tw = nullptr;    // forward table exp(-2 pi i m / N), m < N
  const float* swin = nullptr;   // N synthesis taps, or nullptr when windows are off
  const float* awin = nullptr;   // N analysis taps, or nullptr; COLA weight = awin*swin
  int ola = 0;                   // 0: N samples per frame, frame-major; 1: overlap-add
  int cola = 0;                  // with ola: divide by the summed COLA weights
  int unit_cola = 0;             // with cola: every weight is 1 (the sum counts frames)
  float* out = nullptr;          // frames*N floats (out_len with overlap-add)
  size_t out_len = 0;            // overlap-add: samples written (the rest up to it zero); 0 = frames*N
  // Griffin-Lim: frames below gl_frames take X = gl_mag * dir(gl_re + i gl_im) (the
  // phase update, dir(z) = (cos, sin)(atan2(Im z, Re z))) instead of in_re / in_im
  const float* gl_mag = nullptr;
  const float* gl_re = nullptr;
  const float* gl_im = nullptr;
  size_t gl_frames = 0;
  size_t ola_len = 0;            // set by the launcher: samples the overlap-add reaches
  int kd = 0;                    // set by the launcher: ceil(N / hop)
  int warm = 0;                  // set by the launcher: warm-up steps per run
  int tiles = 0;                 // set by the launcher: workgroups
  int steps = 0;                 // set by the launcher: steps per workgroup
};
// Fused kernel: an inverse C2R half-spectrum plan, N in {512, 1024, 2048}, and (with
// overlap-add) 1 <= hop <= N.
bool istft_fused_supported(int n, int hop, int ola);
size_t istft_ola_len(size_t frames, int n, int hop);
hipError_t launch_istft(IstftArgs a, hipStream_t s);
// General path: column rebuild into `cols` (frames x N complex), the plan's transform,
// then launch_overlap for the output.
hipError_t launch_complex_columns(const IstftArgs& a, int hermitian, float2* cols, hipStream_t s);
hipError_t launch_overlap(IstftArgs a, const float2* cols, hipStream_t s);

// Griffin-Lim (phase_infer.hpp:69-204) pieces. Start: bins i < count get mag[i] *
// (cos, sin)(a) with a the float uniform on [-pi, pi) that mt19937 word words[i] gives
// (generate_canonical<float, 24>, then a + (b - a) u as a fused multiply-add).
hipError_t launch_gl_start(const float* mag, const uint32_t* words, size_t count, float* o_re, float* o_im,
                           hipStream_t s);
// Phase update of frames [0, frames): o = mag * dir(t); with `half`, imag of bins 0 and
// bins-1 set to 0 (when N is even).
hipError_t launch_gl_update(const float* mag, const float* t_re, const float* t_im, size_t frames, int bins,
                            int half, int n_even, float* o_re, float* o_im, hipStream_t s);
// Overlap-add of frames*N frame-major samples with hop H into out[0, ola_len), divided by
// the number of frames per sample.
// The streaming inverse's overlap-add of `frames` frames of n samples at offsets f*e
// (e <= n) onto the n carried samples (and, when w is given, the COLA weight sums onto
// the carried norm): len = (frames - 1)*e + n samples out.
hipError_t launch_stream_ola(const float* frames_pcm, size_t frames, int n, int e, const float* carry,
                             const float* carry_norm, const float* w, float* out, float* out_norm, size_t len,
                             hipStream_t s);
hipError_t launch_gl_overlap(const float* frames_pcm, size_t frames, int n, int hop, float* out, size_t ola_len,
                             hipStream_t s);

// Phase inference (phase_infer_execute, fft_cffi.cpp:2072-2162). Mode 0: the linear phase
// of phase_infer.hpp:36-66, one launch, no scratch. Mode 1:
// log spectrum into frames x N complex columns, then (after the inverse DFT) the causal
// cepstrum fold, then (after the forward DFT) exp into the planes.
hipError_t launch_phase_linear(const float* mag, size_t frames, int bins, int n, int hop, float* o_re, float* o_im,
                               hipStream_t s);
// Minimum phase in one launch, one wave per frame (N = 512, 1024: phase_minimum_fused);
// tw is the N-point forward table.
bool phase_minimum_fused(int n);
hipError_t launch_phase_minimum(const float* mag, size_t frames, int bins, int n, int half, const float2* tw,
                                float* o_re, float* o_im, hipStream_t s);
hipError_t launch_phase_log(const float* mag, size_t frames, int bins, int n, int half, float2* cols, hipStream_t s);
hipError_t launch_phase_fold(size_t frames, int n, float2* cols, hipStream_t s);
hipError_t launch_phase_exp(const float2* cols, size_t frames, int bins, int n, float* o_re, float* o_im,
                            hipStream_t s);

// Metadata kinds (eigfft::MetadataKind order, eigen_fft.hpp:1477-1491).
enum MetaKind : int {
  kMetaTwiddle = 0,
  kMetaLayoutMap = 1,
  kMetaStageParams = 3,
  kMetaButterflyPairs = 4,
  kMetaTwiddleIndexMap = 6,
  kMetaOmega = 7,
};

// Launchers; return hipSuccess or the launch error. `variant` < 0 picks the default
// configuration for N; otherwise an index into the variant table (autotuner).
hipError_t launch_stft(const StftArgs& a, int inverse, int variant, hipStream_t s);
// The variant launch_stft runs for `variant` < 0 (by N and frame count, on the current device).
int default_stft_variant(const StftArgs& a);
// The variant for planes the next launch reads straight back (ordinary stores), or -1.
int readback_stft_variant(const StftArgs& a);
hipError_t launch_c2c(const C2CArgs& a, int inverse, int variant, hipStream_t s);
// Generic real-input path: frames as complex columns (N per frame), and the output
// copy of transformed columns into planes (mag plane = re when polar).
hipError_t launch_frame_complex(const StftArgs& a, float2* out, hipStream_t s);
hipError_t launch_planes(const float2* in, size_t frames, int n, int bins, int polar, float* o_re, float* o_im,
                         float* o_mag, hipStream_t s);
// Index metadata of one plan (kernel: 2 Stockham, else Cooley-Tukey), written as
// complex<float> pairs; returns the entry count in *count (0 if not recorded).
hipError_t launch_metadata(int n, int kernel, int kind, int r2c_half, const float2* tw, int inverse,
                           float2* out, size_t capacity, size_t* count, hipStream_t s);

// Complex DFT for N above c2c_max_lds_n(): two LDS-resident passes (columns, then rows)
// through `tmp` (batch*N complex values), with every pre and post hook.
bool large_supported(size_t n);
void large_split(size_t n, int* n1, int* n2, int* s);
hipError_t launch_c2c_large(const C2CArgs& a, int inverse, int variant, float2* tmp, hipStream_t s);
// Line-kernel choices of the two passes (autotuner); variant -1 = default.
int large_variant_count(size_t n);
bool large_variant_desc(size_t n, int variant, char* buf, size_t len);

// Signals STFT and adjoint: N in {512, 1024, 2048}, forward R2C half spectrum, output
// mode kOutComplex / kOutMagnitude / kOutPolar; the adjoint needs window + k*hop <= 8N
// (k = 2 for N = 2048, else 4).
bool signals_supported(int n, int window, int hop, bool adjoint);
hipError_t launch_stft_signals(SignalArgs a, int mode, hipStream_t s);
// With halo_need, launches nothing and reports the halo floats the launch would use. With
// `saved`, V comes from the gradients and the forward's saved planes (complex: V = g_mag +
// i g_phase; polar: p_mag, p_phase) and no frame is transformed forward; magnitude mode is
// refused there.
hipError_t launch_stft_adjoint(SignalArgs a, int mode, hipStream_t s, size_t* halo_need = nullptr,
                               bool saved = false);
// Device address (on the current device) and size in floats of the static halo scratch
// of the allocation-free adjoint entry; nullptr if the lookup fails.
float* noalloc_halo_scratch(size_t* capacity);

// Variant tables for the autotuner.
struct VariantInfo {
  int l;            // complex length the kernel transforms
  int e;            // values per thread
  int fpw;          // transforms per workgroup
  const char* radices;
};

// Ragged batches: one work item = `ntiles` consecutive blocks of request `req`, run by
// engine `cfg` (ragged_cfg): a tile engine (blocks = tiles of FPW frames) or a
// line-aligned engine (blocks as in ragged_lines_layout).
struct RaggedItem {
  unsigned req;
  unsigned cfg;
  unsigned tile;
  unsigned ntiles;
};
int ragged_cfg_count();
const VariantInfo* ragged_cfg(int id);  // l = N/2; fpw = frames per block
bool ragged_cfg_lines(int id);
// Engines that run one block per workgroup (no item loop): items of one block each, grid
// = 8 x the longest chunk.
bool ragged_cfg_flat(int id);          // a line-aligned engine
// Blocks of request `a` under line-aligned engine `cfg`: frames in whole tiles (`lead`,
// to be set in the request's StftArgs), blocks, frames covered (`take`); false when the
// request does not qualify (window != N, odd hop, misaligned buffers).
bool ragged_lines_layout(int cfg, const StftArgs& a, size_t* lead, size_t* blocks, size_t* take);
// Resident workgroups of engine `cfg`'s kernel on the current device (CUs x occupancy).
hipError_t ragged_grid(int mode, int cfg, int* grid);
// One persistent launch of engine `cfg` over its items: workgroups of XCD x
// (blockIdx % 8 == x) stride through items[chunks[x] .. chunks[x+1]). Forward R2C, half
// spectrum, output mode kOutComplex / kOutMagnitude / kOutPolar.
hipError_t launch_stft_ragged(const StftArgs* reqs, const RaggedItem* items, const unsigned* chunks, int mode,
                              int cfg, int grid, hipStream_t s);
int stft_variant_count(int n);
const VariantInfo* stft_variant(int n, int idx);
int c2c_variant_count(int n);
const VariantInfo* c2c_variant(int n, int idx);
// Largest N the single-launch LDS kernels cover.
int stft_max_lds_n();
int c2c_max_lds_n();

}  // namespace fftfree
