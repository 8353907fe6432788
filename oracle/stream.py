"""Restatement of the reference's streaming STFT (fft_cffi.cpp:1342-1742) over the float
restatement of its batched paths (oracle/ref.py). TEST INFRASTRUCTURE ONLY.

Each method follows the reference function of the same name: a PCM buffer with a read
offset for the forward direction; for the inverse direction an N-sample overlap-add
accumulator (with COLA weight sums when normalising) and a queue of finished samples.
"""
from __future__ import annotations

import numpy as np

from . import ref

FLUSH_NONE, FLUSH_PARTIAL, FLUSH_FINAL = 0, 1, 2


class Stream:
    def __init__(self, N, window=0, hop=0, inverse=False, kernel=0, transform=1, half_spectrum=True,
                 reduce_magnitude=False, store_polar=False, taps=None, analysis=None, synthesis=None,
                 windows_enabled=False, cola_mode=0):
        self.N, self.inverse, self.kernel, self.transform = N, inverse, kernel, transform
        self.half, self.reduce, self.polar = half_spectrum, reduce_magnitude, store_polar
        self.W = window if window > 0 else N                       # stream_window, 1344-1350
        self.H = hop if hop > 0 else self.W                         # stream_hop, 1352-1358
        self.window, self.hop = window, hop
        self.taps, self.analysis, self.synthesis = taps, analysis, synthesis
        self.windows_enabled, self.cola = windows_enabled, cola_mode
        self.bins = N // 2 + 1 if half_spectrum else N
        self.reset()

    def reset(self):                                                # fft_stream_reset, 1460-1471
        self.buf = np.zeros(0, np.float32)
        self.offset = 0
        self.queue = []
        self.qden = []    # the COLA weight sum each queued sample was divided by (1 without)
        self.last_den = np.zeros(0)
        self.accum = np.zeros(0, np.float32)
        self.norm = np.zeros(0, np.float32)

    def backlog(self):                                              # stream_backlog, 1360-1365
        return max(0, self.buf.size - self.offset)

    def pending_frames(self):                                       # stream_available_frames, 1367-1375
        b = self.backlog()
        return 0 if b < self.W else 1 + (b - self.W) // self.H

    def pending_pcm(self):
        return len(self.queue) if self.inverse else 0

    def push_pcm(self, pcm, max_frames=0):                          # fft_stream_push_pcm, 1508-1600
        if pcm is not None and len(pcm):
            self.buf = np.concatenate([self.buf, np.asarray(pcm, np.float32)])
        frames = self.pending_frames()
        if frames == 0:
            return 0, None
        if max_frames and frames > max_frames:
            frames = max_frames
        length = min(self.W + (frames - 1) * self.H, self.backlog())
        chunk = self.buf[self.offset:self.offset + length]
        got, re, im, mg = ref.execute_batched(chunk, self.N, self.window, self.hop, ctx_pad=0, pad_mode=2,
                                              max_frames=frames, inverse=self.inverse, kernel=self.kernel,
                                              transform=self.transform, reduce_magnitude=self.reduce,
                                              store_polar=self.polar, half_spectrum=self.half,
                                              window_taps=self.taps)
        if got == 0:
            return 0, None
        self.offset += got * self.H
        self.buf = self.buf[self.offset:]
        self.offset = 0
        return got, (re, im, mg)

    def _weight(self, i):                                           # cola_window_weight, 210-226
        if not self.windows_enabled:
            return np.float32(1.0)
        a = self.analysis[i] if self.analysis is not None else np.float32(1.0)
        b = self.synthesis[i] if self.synthesis is not None else np.float32(1.0)
        return np.float32(a * b)

    def _emit(self, max_samples):                                   # stream_inverse_emit, 1420-1442
        take = len(self.queue)
        if max_samples and take > max_samples:
            take = max_samples
        out = np.array(self.queue[:take], np.float32)
        self.last_den = np.array(self.qden[:take], np.float64)
        self.queue = self.queue[take:]
        self.qden = self.qden[take:]
        return out

    def push_spectrum(self, re, im, max_samples=0, flush=FLUSH_NONE):  # fft_stream_push_spectrum, 1602-1742
        if not self.inverse:
            return np.zeros(0, np.float32)
        N = self.N
        if self.accum.size != N:
            self.accum = np.zeros(N, np.float32)
        use_norm = self.cola == 1
        if use_norm:
            if self.norm.size != N:
                self.norm = np.zeros(N, np.float32)
        else:
            self.norm = np.zeros(0, np.float32)
        frames = 0 if re is None else re.shape[0]
        if frames > 0:
            syn = self.synthesis if self.windows_enabled else None
            got, fr = ref.execute_complex_batched(re, im, N, inverse=True, kernel=self.kernel,
                                                  transform=self.transform, reduce_magnitude=self.reduce,
                                                  store_polar=self.polar, half_spectrum=self.half, synth_taps=syn,
                                                  max_frames=frames)
            if got == 0:
                return self._emit(max_samples)
            w = np.array([self._weight(i) for i in range(N)], np.float32) if use_norm else None
            for f in range(min(got, frames)):
                self.accum = (self.accum + fr[f]).astype(np.float32)
                if use_norm:
                    self.norm = (self.norm + w).astype(np.float32)
                emit = N if (flush == FLUSH_FINAL and f + 1 == got) else self.H
                emit = min(emit, N)
                v = self.accum[:emit].copy()
                if use_norm:
                    d = self.norm[:emit]
                    m = np.abs(d) > 1e-12
                    v[m] = v[m] / d[m]
                    self.qden.extend(d.astype(np.float64).tolist())
                else:
                    self.qden.extend([1.0] * emit)
                self.queue.extend(v.tolist())
                self.accum = np.concatenate([self.accum[emit:], np.zeros(emit, np.float32)])
                if use_norm:
                    self.norm = np.concatenate([self.norm[emit:], np.zeros(emit, np.float32)])
        if flush == FLUSH_FINAL and frames == 0 and self.accum.size:
            v = self.accum.copy()
            if self.norm.size:
                m = np.abs(self.norm) > 1e-12
                v[m] = v[m] / self.norm[m]
                self.qden.extend(self.norm.astype(np.float64).tolist())
            else:
                self.qden.extend([1.0] * v.size)
            self.queue.extend(v.tolist())
            self.accum = np.zeros_like(self.accum)
            if self.norm.size:
                self.norm = np.zeros_like(self.norm)
        return self._emit(max_samples)
