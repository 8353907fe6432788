// A C++ caller of include/fft_cffi.h, compiled with g++ against the header and linked to
// libfft_cffi.so: the reference's own C-ABI callers are C++ programs like this one, and
// their call sites must compile and run unchanged against this library.
//
//   caller roundtrip N B out.bin
//       the roundtrip mode of the reference's tests/restore_test_cffi.cpp:120-210:
//       B*N samples from std::mt19937_64(1337) + std::normal_distribution<double>(0, 1),
//       a forward R2C half-spectrum context and a C2R inverse context from fft_init_full
//       (kernel/radix auto, pad_mode 1, window = hop = N), fft_execute_batched, then
//       fft_execute_complex_batched; exit 0 when max|x - x'| / max|x| <= 1e-3, 6 when
//       not (the reference test's tolerance and codes). Writes N, B, frames and the
//       input, the three planes and the restored samples to out.bin.
//   caller stft N hop in.bin out.bin
//       fft_init_full(N, ..., kernel 2, pad_mode 0, window N, hop, R2C, reduce_magnitude,
//       half spectrum), the settings of the N = 1024 STFT workload (fft_init_full never
//       windows), on the float32 samples of in.bin; fft_execute_batched with max_frames
//       0, then fft_stft_execute on the same samples; writes frames, bins and both
//       calls' three planes to out.bin.
// Also calls the worker-pool test hooks, which exist so that such callers link.
#include "fft_cffi.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <string>
#include <vector>

namespace {

bool write_all(std::FILE* f, const void* p, size_t bytes) { return std::fwrite(p, 1, bytes, f) == bytes; }

int roundtrip(int N, int B, const char* out_path) {
  const int threads = 4;
  const size_t pcm_len = static_cast<size_t>(N) * static_cast<size_t>(B);
  std::vector<float> pcm(pcm_len);
  std::mt19937_64 rng(1337);
  std::normal_distribution<double> dist(0.0, 1.0);
  for (size_t i = 0; i < pcm_len; ++i) pcm[i] = static_cast<float>(dist(rng));

  fft_clear_workerpool_dropout_history();
  fft_set_workerpool_dropout_pattern(4, 1, 0);
  fft_set_workerpool_dropout_persistent(0);
  fft_register_restore_observer(nullptr);

  auto make_ctx = [&](int inverse, int transform) -> void* {
    return fft_init_full(static_cast<size_t>(N), threads, 1, inverse, 0, 0, nullptr, 0, 1, N, N, transform, 0, 0, 1,
                         1, 0, 0, 0, 1);
  };
  void* ctx = make_ctx(0, 1);
  if (!ctx) {
    std::fprintf(stderr, "fft_init_full failed\n");
    return 2;
  }
  const size_t bins = static_cast<size_t>(N) / 2 + 1;
  const size_t max_frames = static_cast<size_t>(B);
  std::vector<float> re(max_frames * bins), im(max_frames * bins), mag(max_frames * bins);
  const size_t produced =
      fft_execute_batched(ctx, pcm.data(), pcm_len, re.data(), im.data(), mag.data(), 1, 0, max_frames);
  if (produced == 0) {
    std::fprintf(stderr, "fft_execute_batched returned 0 frames\n");
    fft_free(ctx);
    return 3;
  }
  void* inv = make_ctx(1, 2);
  if (!inv) {
    std::fprintf(stderr, "fft_init_full (inverse) failed\n");
    fft_free(ctx);
    return 4;
  }
  std::vector<float> restored(pcm_len);
  const size_t recovered =
      fft_execute_complex_batched(inv, re.data(), im.data(), produced, restored.data(), 1, 0, produced);
  if (recovered == 0) {
    std::fprintf(stderr, "fft_execute_complex_batched failed\n");
    fft_free(ctx);
    fft_free(inv);
    return 5;
  }
  double max_abs = 0.0, max_err = 0.0;
  for (size_t i = 0; i < pcm_len; ++i) {
    max_abs = std::fmax(max_abs, std::fabs((double)pcm[i]));
    max_err = std::fmax(max_err, std::fabs((double)pcm[i] - (double)restored[i]));
  }
  const double rel = max_err / (max_abs + 1e-30);
  std::printf("ROUND_TRIP: frames=%zu max_err=%g rel=%g\n", produced, max_err, rel);
  fft_free(ctx);
  fft_free(inv);
  fft_free(nullptr);  // a no-op, as in the reference

  std::FILE* f = std::fopen(out_path, "wb");
  if (!f) return 7;
  const long long hdr[3] = {N, B, static_cast<long long>(produced)};
  bool ok = write_all(f, hdr, sizeof hdr) && write_all(f, pcm.data(), pcm_len * 4) &&
            write_all(f, re.data(), produced * bins * 4) && write_all(f, im.data(), produced * bins * 4) &&
            write_all(f, mag.data(), produced * bins * 4) && write_all(f, restored.data(), pcm_len * 4);
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) return 7;
  return rel <= 1e-3 ? 0 : 6;
}

int stft(int N, int hop, const char* in_path, const char* out_path) {
  std::FILE* f = std::fopen(in_path, "rb");
  if (!f) return 7;
  std::vector<float> pcm;
  float buf[4096];
  size_t got;
  while ((got = std::fread(buf, 4, 4096, f)) > 0) pcm.insert(pcm.end(), buf, buf + got);
  std::fclose(f);

  const int pattern[5] = {4, 4, 4, 4, 4};  // radix hints: copied, no numeric effect
  void* ctx = fft_init_full(static_cast<size_t>(N), 16, 1, 0, 2, FFT_BUTTERFLY_RADIX_4, pattern, 5, 0, N, hop,
                            FFT_TRANSFORM_R2C, 1, 0, 1, 1, 0, 0, 0, 1);
  if (!ctx) return 2;
  if (fft_ctx_size(ctx) != static_cast<size_t>(N)) return 8;
  const size_t bins = static_cast<size_t>(N) / 2 + 1;
  const size_t cap = pcm.size() / static_cast<size_t>(hop) + 4;
  std::vector<float> re(cap * bins), im(cap * bins), mag(cap * bins);
  const size_t frames = fft_execute_batched(ctx, pcm.data(), pcm.size(), re.data(), im.data(), mag.data(), 0, 0, 0);
  std::vector<float> re2(cap * bins), im2(cap * bins), mag2(cap * bins);
  const size_t frames2 = fft_stft_execute(ctx, pcm.data(), pcm.size(), re2.data(), im2.data(), mag2.data(), 0);
  fft_free(ctx);
  if (frames == 0 || frames2 == 0) return 3;

  f = std::fopen(out_path, "wb");
  if (!f) return 7;
  const long long hdr[3] = {static_cast<long long>(frames), static_cast<long long>(frames2),
                            static_cast<long long>(bins)};
  bool ok = write_all(f, hdr, sizeof hdr);
  for (const std::vector<float>* v : {&re, &im, &mag}) ok = ok && write_all(f, v->data(), frames * bins * 4);
  for (const std::vector<float>* v : {&re2, &im2, &mag2}) ok = ok && write_all(f, v->data(), frames2 * bins * 4);
  ok = (std::fclose(f) == 0) && ok;
  return ok ? 0 : 7;
}

}  // namespace

int main(int argc, char** argv) {
  const std::string mode = argc > 1 ? argv[1] : "";
  if (mode == "roundtrip" && argc == 5) return roundtrip(std::atoi(argv[2]), std::atoi(argv[3]), argv[4]);
  if (mode == "stft" && argc == 6) return stft(std::atoi(argv[2]), std::atoi(argv[3]), argv[4], argv[5]);
  std::fprintf(stderr, "usage: caller roundtrip N B out.bin | caller stft N hop in.bin out.bin\n");
  return 1;
}
