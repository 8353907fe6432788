"""The oracle: CPU restatements of the reference fftfree path, used as the checker.

TEST INFRASTRUCTURE ONLY. Only tests/, ``__graft_entry__.smoke()`` and ``bench.py``'s
CPU-baseline leg import this package. The product path (``fftfree_amd``) never does,
and fails loudly when its HIP library is missing instead of falling back to anything here.

Modules:
  fp64       FP64 DFT oracle (numpy float64 FFT, cross-checked by a direct O(N^2) DFT),
             float32-exact window builder, frame-count rule, post-transform rules.
  index_maps closed forms of the reference's integer metadata (bit reversal, LayoutMap,
             TwiddleIndexMap, ButterflyPairs, StageParams, Twiddle indices).
  ref        ctypes binding of cpu_ref.cpp, the C++ restatement of the reference CPU
             path (radix-2 Cooley-Tukey and Stockham in float, framing, output planes).
  signals    seeded synthetic inputs of the benchmark configurations.
"""
