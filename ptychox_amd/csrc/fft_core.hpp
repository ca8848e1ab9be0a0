// LDS-staged Stockham FFT core for CDNA4 (gfx950), complex FP32.
//
// A length-N sequence is held by T = N/E threads, E complex values each, in
// "natural" layout: thread t holds x[t + q*T] in v[q].  Each stage does E/R
// radix-R butterflies per thread in registers; between stages the values go
// through LDS once (ping-pong buffers, one barrier per exchange).  The first
// stage reads, and the last stage leaves, the natural layout, so the caller's
// global loads and stores are coalesced across the T threads of a sequence
// (consecutive t -> consecutive addresses) and any prologue / epilogue
// (gather, probe multiply, fftshift-by-sign, amplitude, loss gradient) is
// applied in registers with no extra HBM round trip.
//
// Stage s uses radix E (or the remainder radix for the last stage) and
// Ns = E^s; butterfly j reads x[j + m*N/R], twiddles by w_{Ns R}^{(j mod Ns) m}
// and writes y[(j - j mod Ns) R + j mod Ns + m Ns] (Stockham autosort).
//
// Twiddles come from a compile-time table of exp(-2 pi i q / TW_MAX) computed
// in double precision (Taylor series on an exact octant reduction) and rounded
// once to float, so no fast-math sincos enters the transform.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ptx {

// ---------------------------------------------------------------- complex helpers

__device__ __forceinline__ float2 c_add(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 c_sub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 c_mul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
__device__ __forceinline__ float2 c_mulc(float2 a, float2 b) {
  return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ float2 c_scale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__device__ __forceinline__ float2 c_neg(float2 a) { return make_float2(-a.x, -a.y); }
// |a| with the hardware square root (v_sqrt_f32, 1 ulp): the correctly
// rounded sqrtf expands to ~16 instructions, a tenth of a column pass's ALU work
__device__ __forceinline__ float c_abs(float2 a) { return __builtin_amdgcn_sqrtf(a.x * a.x + a.y * a.y); }
__device__ __forceinline__ float2 c_zero() { return make_float2(0.0f, 0.0f); }
// 1/x by the hardware reciprocal (v_rcp_f32, ~1 ulp) for the normalising
// divides of the loss / phase epilogues: the IEEE divide sequence costs ~10
// instructions and several registers per element.  x >= FLT_EPSILON there.
__device__ __forceinline__ float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
// |a| and 1/max(|a|, floor) from one hardware reciprocal square root
// (v_rsq_f32) instead of a square root and a reciprocal (two quarter-rate
// instructions): with s = |a|^2 and r = rsq(max(s, FLT_MIN)), |a| = s r and
// 1/max(|a|, floor) = min(r, 1/floor).  Above |a| = 1.1e-19 (s = FLT_MIN) the
// magnitude is within ~2 ulp; below, it is off by less than 1.1e-19 absolute.
struct AbsInv {
  float a, inv;
};
__device__ __forceinline__ AbsInv c_abs_inv(float2 v, float inv_floor) {
  const float s = v.x * v.x + v.y * v.y;
  const float r = __builtin_amdgcn_rsqf(fmaxf(s, 1.17549435e-38f));
  return AbsInv{s * r, fminf(r, inv_floor)};
}

// ---------------------------------------------------------------- compile-time twiddle table

constexpr int TW_MAX = 2048;  // largest supported transform length

namespace detail {
constexpr double kPi = 3.14159265358979323846264338327950288;

constexpr double sin_small(double x) {  // |x| <= pi/4, error < 1e-17
  double x2 = x * x, term = x, sum = x;
  for (int n = 1; n < 14; ++n) {
    term *= -x2 / double((2 * n) * (2 * n + 1));
    sum += term;
  }
  return sum;
}
constexpr double cos_small(double x) {
  double x2 = x * x, term = 1.0, sum = 1.0;
  for (int n = 1; n < 14; ++n) {
    term *= -x2 / double((2 * n - 1) * (2 * n));
    sum += term;
  }
  return sum;
}
// cos and sin of 2 pi q / M, exact octant reduction on the integer q
constexpr void sincos_turn(int q, int M, double& c, double& s) {
  const int quarter = M / 4, eighth = M / 8;
  const int quad = q / quarter;
  int r = q % quarter;
  double cr = 0, sr = 0;
  if (r <= eighth) {
    const double a = 2.0 * kPi * double(r) / double(M);
    cr = cos_small(a);
    sr = sin_small(a);
  } else {
    const double a = 2.0 * kPi * double(quarter - r) / double(M);
    cr = sin_small(a);
    sr = cos_small(a);
  }
  switch (quad) {
    case 0: c = cr; s = sr; break;
    case 1: c = -sr; s = cr; break;
    case 2: c = -cr; s = -sr; break;
    default: c = sr; s = -cr; break;
  }
}
struct TwiddleTable {
  float v[2 * TW_MAX];  // interleaved (re, im) of exp(-2 pi i q / TW_MAX)
};
constexpr TwiddleTable make_twiddles() {
  TwiddleTable t{};
  for (int q = 0; q < TW_MAX; ++q) {
    double c = 0, s = 0;
    sincos_turn(q, TW_MAX, c, s);
    t.v[2 * q] = float(c);
    t.v[2 * q + 1] = float(-s);
  }
  return t;
}
// (w^q, w^(4q)) pairs, w = exp(-2 pi i / TW_MAX), interleaved (re, im, re, im):
// one 16-byte load gives a radix-8 butterfly both of its table twiddles
struct TwiddlePairTable {
  float v[4 * TW_MAX];
};
constexpr TwiddlePairTable make_twiddle_pairs() {
  TwiddlePairTable t{};
  for (int q = 0; q < TW_MAX; ++q) {
    double c = 0, s = 0;
    sincos_turn(q, TW_MAX, c, s);
    t.v[4 * q] = float(c);
    t.v[4 * q + 1] = float(-s);
    sincos_turn((4 * q) % TW_MAX, TW_MAX, c, s);
    t.v[4 * q + 2] = float(c);
    t.v[4 * q + 3] = float(-s);
  }
  return t;
}
// exp(-2 pi i k / 16), k = 0..7, as exact-as-float constants for the in-register DFTs
constexpr float kC16[8] = {1.0f, 0.92387953251128674f, 0.70710678118654752f, 0.38268343236508977f,
                           0.0f, -0.38268343236508977f, -0.70710678118654752f, -0.92387953251128674f};
constexpr float kS16[8] = {0.0f, 0.38268343236508977f, 0.70710678118654752f, 0.92387953251128674f,
                           1.0f, 0.92387953251128674f, 0.70710678118654752f, 0.38268343236508977f};
}  // namespace detail

__constant__ const detail::TwiddleTable c_twiddle = detail::make_twiddles();
__constant__ const detail::TwiddlePairTable c_twiddle_pair = detail::make_twiddle_pairs();

// w^q with w = exp(-2 pi i / TW_MAX); DIR > 0 returns the conjugate (inverse transform)
template <int DIR>
__device__ __forceinline__ float2 twiddle(int q) {
  const float2 w = reinterpret_cast<const float2*>(c_twiddle.v)[q];
  return DIR > 0 ? make_float2(w.x, -w.y) : w;
}

// ---------------------------------------------------------------- in-register DFTs

// multiply by exp(DIR * 2 pi i * k16 / 16) with k16 a compile-time constant after unrolling
template <int DIR>
__device__ __forceinline__ float2 rot16(float2 a, int k16) {
  if (k16 == 0) return a;
  if (k16 == 4) return DIR < 0 ? make_float2(a.y, -a.x) : make_float2(-a.y, a.x);
  const float c = detail::kC16[k16];
  const float s = DIR < 0 ? -detail::kS16[k16] : detail::kS16[k16];
  return make_float2(a.x * c - a.y * s, a.x * s + a.y * c);
}

// length-R DFT (R = 1, 2, 4, 8, 16) in natural order, sign DIR (-1 forward, +1 inverse), unnormalised
template <int R, int DIR>
__device__ __forceinline__ void dft(float2 (&v)[R]) {
  if constexpr (R > 1) {
    float2 e[R / 2], o[R / 2];
#pragma unroll
    for (int m = 0; m < R / 2; ++m) {
      e[m] = v[2 * m];
      o[m] = v[2 * m + 1];
    }
    dft<R / 2, DIR>(e);
    dft<R / 2, DIR>(o);
#pragma unroll
    for (int k = 0; k < R / 2; ++k) {
      const float2 t = rot16<DIR>(o[k], k * (16 / R));
      v[k] = c_add(e[k], t);
      v[k + R / 2] = c_sub(e[k], t);
    }
  }
}

// ---------------------------------------------------------------- stage machinery

constexpr int ilog2(int n) { return n <= 1 ? 0 : 1 + ilog2(n / 2); }
constexpr int ipow(int b, int e) { return e == 0 ? 1 : b * ipow(b, e - 1); }

// LDS index with one pad slot per 16 entries (breaks power-of-two bank strides)
__device__ __forceinline__ int lds_pad(int i) { return i + (i >> 4); }
// per-sequence LDS length (float2 units); the +1 keeps neighbouring sequences in different banks
template <int N>
constexpr int lds_len() { return N == 1 ? 1 : N + N / 16 + 1; }

template <int N, int E>
struct FftShape {
  static_assert((N & (N - 1)) == 0 && N >= 2 && N <= TW_MAX, "N must be a power of two <= TW_MAX");
  static_assert(E == 2 || E == 4 || E == 8 || E == 16, "E must be 2, 4, 8 or 16");
  static constexpr int T = N / E;  // threads per sequence
  static constexpr int FULL = ilog2(N) / ilog2(E);
  static constexpr int REM = 1 << (ilog2(N) % ilog2(E));
  static constexpr int STAGES = FULL + (REM > 1 ? 1 : 0);
  static constexpr bool NEEDS_LDS = STAGES > 1;
};

// Work-group barrier that orders LDS accesses only: waits for this wave's LDS
// operations (lgkmcnt) but not for its global loads and stores, so global
// loads issued ahead of the transform (prefetch of the next tile) stay in
// flight across the exchanges.  __syncthreads() would also drain vmcnt.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Ordering of LDS accesses among the lanes of one wave only: the LDS performs
// a wave's operations in issue order, so a sequence held entirely by one wave
// (a row of T <= 64 threads; rows never straddle waves since T divides 64)
// needs no work-group barrier between an exchange's writes and reads -- only
// that the compiler keeps them in program order.  The other waves of the
// work-group then run their transforms without waiting for each other.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
}

// The synchronisation an exchange needs: the wave's own (WAVE: the sequence
// is held by one wave) or the work-group's.
template <bool WAVE>
__device__ __forceinline__ void seq_sync() {
  if constexpr (WAVE)
    wave_lds_sync();
  else
    lds_barrier();
}

// LDS_TW: the stage's base twiddles come from `lds_tw`, an LDS copy of
// w_N^k (k < N, forward direction; see load_lds_twiddles) instead of the
// global table.  REG_TW: every twiddle of the stage comes from `reg_tw`
// (RegTwiddles below, forward direction, conjugated here for the inverse).
// PAIR_TW (global-table twiddles): w and w^4 from one load of the pair table
// rather than two loads of the single table (same values).
template <int N, int E, int DIR, int R, int NS, bool LAST, bool SYNC_FIRST, bool LDS_TW = false, bool REG_TW = false,
          bool PAIR_TW = true, bool WAVE = false>
__device__ __forceinline__ void fft_stage(float2 (&v)[E], int t, float2* __restrict__ lds,
                                          const float2* __restrict__ lds_tw = nullptr,
                                          const float2* reg_tw = nullptr) {
  constexpr int T = N / E;
  constexpr int NB = E / R;               // butterflies per thread
  constexpr int TWS = TW_MAX / (NS * R);  // twiddle stride into the table
  // single-buffer mode: everyone must have finished reading the previous exchange
  if constexpr (SYNC_FIRST) seq_sync<WAVE>();
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    const int j = t + i * T;
    const int k = j & (NS - 1);
    float2 a[R];
#pragma unroll
    for (int m = 0; m < R; ++m) a[m] = v[i + m * NB];
    if constexpr (NS > 1 && REG_TW) {
#pragma unroll
      for (int m = 1; m < R; ++m) a[m] = DIR > 0 ? c_mulc(a[m], reg_tw[i + m * NB]) : c_mul(a[m], reg_tw[i + m * NB]);
    } else if constexpr (NS > 1) {
      // w^(k m): m = 1, 4 and 8 from the tables, w^2 = w w, every other power
      // the product of the power of two below it and an earlier power
      // (m = 7: w^4 w^3, w^3 = w^2 w).  A recurrence from w alone
      // (w^(m+1) = w^m w) multiplies the table's rounding by m and adds one
      // rounding per step: twiddles 2x less accurate (rms 1.2e-7 vs 5.7e-8 at
      // radix 8, 2.0e-7 vs 6.1e-8 at radix 16) and transforms ~1.5x less
      // accurate than rocFFT at 512 points and above.  From the global
      // tables, w and w^4 are one 16-byte load of the pair table (as many
      // load instructions as the recurrence needs: the one-row passes that
      // read the global table are sensitive to the count).
      float2 w[R];
      constexpr bool PAIR = PAIR_TW && !LDS_TW && R > 4;
      if constexpr (PAIR) {
        const float4 pr = reinterpret_cast<const float4*>(c_twiddle_pair.v)[k * TWS];
        w[1] = DIR > 0 ? make_float2(pr.x, -pr.y) : make_float2(pr.x, pr.y);
        w[4] = DIR > 0 ? make_float2(pr.z, -pr.w) : make_float2(pr.z, pr.w);
      }
#pragma unroll
      for (int m = 1; m < R; ++m) {
        if (m == 2) {
          w[m] = c_mul(w[1], w[1]);
        } else if ((m & (m - 1)) == 0) {
          if constexpr (LDS_TW) {
            const float2 t0 = lds_tw[k * m * (N / (NS * R))];
            w[m] = DIR > 0 ? make_float2(t0.x, -t0.y) : t0;
          } else if (!PAIR || m == 8) {  // (m = 1, 4 came from the pair table)
            w[m] = twiddle<DIR>(k * m * TWS);
          }
        } else {
          const int hb = 1 << ilog2(m);
          w[m] = c_mul(w[hb], w[m - hb]);
        }
        a[m] = c_mul(a[m], w[m]);
      }
    }
    dft<R, DIR>(a);
    if constexpr (LAST) {
#pragma unroll
      for (int m = 0; m < R; ++m) v[i + m * NB] = a[m];
    } else {
      // pad(base + m NS) = pad(base) + pad(m NS): the R outputs lie in one
      // NS R-aligned block (base = (j - k) R + k, k < NS), which either sits
      // inside one 16-entry pad group (NS R <= 16) or steps m NS by whole
      // groups from an offset below 16 -- so one address per butterfly and
      // compile-time offsets (checked exhaustively for N <= 4096, R <= 16)
      float2* wb = lds + lds_pad((j - k) * R + k);
#pragma unroll
      for (int m = 0; m < R; ++m) wb[lds_pad(m * NS)] = a[m];
    }
  }
  if constexpr (!LAST) {
    seq_sync<WAVE>();
    // pad(t + q T) = pad(t) + pad(q T) likewise (t < T, T a power of two)
    const float2* rb = lds + lds_pad(t);
#pragma unroll
    for (int q = 0; q < E; ++q) v[q] = rb[lds_pad(q * T)];
  }
}

// LDS copy of the forward twiddles w_N^k, k < N (N float2), written by the
// whole work-group; a kernel that transforms many sequences reads its stage
// twiddles from here, so its sequence loop issues no global twiddle loads --
// which matters when the loop also keeps global loads of the next sequence in
// flight (vector-memory counters retire in order, so waiting for a later
// twiddle load would also wait for those).  Ends with a barrier.
template <int N>
__device__ __forceinline__ void load_lds_twiddles(float2* lds_tw) {
  for (int k = threadIdx.x; k < N; k += blockDim.x) lds_tw[k] = twiddle<-1>(k * (TW_MAX / N));
  lds_barrier();
}

// PINGPONG: alternate two LDS regions (one barrier per exchange); otherwise a
// single region and two barriers per exchange (half the LDS footprint).
// WAVE: the sequence is held by one wave (wave_lds_sync instead of barriers).
// Stages S .. SEND - 1 only: a transform can be run in parts, with other work
// (the next tile's loads) issued between its stages.
template <int N, int E, int DIR, bool PINGPONG, int S, bool LDS_TW = false, bool PAIR_TW = true, bool WAVE = false,
          int SEND = 64>
__device__ __forceinline__ void fft_stages(float2 (&v)[E], int t, float2* lds0, float2* lds1,
                                           const float2* lds_tw = nullptr) {
  using Sh = FftShape<N, E>;
  if constexpr (S < Sh::STAGES && S < SEND) {
    constexpr int R = S < Sh::FULL ? E : Sh::REM;
    constexpr int NS = ipow(E, S);
    constexpr bool LAST = S == Sh::STAGES - 1;
    constexpr bool SYNC_FIRST = !PINGPONG && S > 0 && !LAST;
    fft_stage<N, E, DIR, R, NS, LAST, SYNC_FIRST, LDS_TW, false, PAIR_TW, WAVE>(
        v, t, (PINGPONG && (S & 1)) ? lds1 : lds0, lds_tw);
    fft_stages<N, E, DIR, PINGPONG, S + 1, LDS_TW, PAIR_TW, WAVE, SEND>(v, t, lds0, lds1, lds_tw);
  }
}

// Stage twiddles held in registers, for passes whose threads keep the same
// element indices from one sequence to the next (a column pass: thread t holds
// rows t + q T of every tile).  Stage S >= 1 multiplies input m of butterfly i
// by w^(k m), k = (t + i T) mod NS; the table values are loaded once per
// thread and serve every sequence, forward and (conjugated) inverse, in place
// of the per-stage recurrence.  (E - E/R) float2 per stage.
template <int N, int E>
struct RegTwiddles {
  using Sh = FftShape<N, E>;
  static constexpr int NST = Sh::STAGES > 1 ? Sh::STAGES - 1 : 1;
  float2 w[NST][E];  // [S - 1][i + m NB], forward direction; entries m = 0 unused
  template <int S = 1>
  __device__ __forceinline__ void init(int t) {
    if constexpr (S < Sh::STAGES) {
      constexpr int T = N / E;
      constexpr int R = S < Sh::FULL ? E : Sh::REM;
      constexpr int NS = ipow(E, S);
      constexpr int NB = E / R;
      constexpr int TWS = TW_MAX / (NS * R);
#pragma unroll
      for (int i = 0; i < NB; ++i) {
        const int k = (t + i * T) & (NS - 1);
        w[S - 1][i] = make_float2(1.0f, 0.0f);
#pragma unroll
        for (int m = 1; m < R; ++m) w[S - 1][i + m * NB] = twiddle<-1>(k * m * TWS);
      }
      init<S + 1>(t);
    }
  }
};

template <int N, int E, int DIR, bool PINGPONG, int S = 0, int SEND = 64>
__device__ __forceinline__ void fft_stages_reg(float2 (&v)[E], int t, float2* lds0, float2* lds1,
                                               const RegTwiddles<N, E>& tw) {
  using Sh = FftShape<N, E>;
  if constexpr (S < Sh::STAGES && S < SEND) {
    constexpr int R = S < Sh::FULL ? E : Sh::REM;
    constexpr int NS = ipow(E, S);
    constexpr bool LAST = S == Sh::STAGES - 1;
    constexpr bool SYNC_FIRST = !PINGPONG && S > 0 && !LAST;
    fft_stage<N, E, DIR, R, NS, LAST, SYNC_FIRST, false, true>(v, t, (PINGPONG && (S & 1)) ? lds1 : lds0, nullptr,
                                                               tw.w[S > 0 ? S - 1 : 0]);
    fft_stages_reg<N, E, DIR, PINGPONG, S + 1, SEND>(v, t, lds0, lds1, tw);
  }
}

// Stage twiddles of a kernel that transforms many sequences with a fixed
// thread-to-index map: in registers at up to 8 elements per thread (up to 14
// float2 per stage) unless the kernel opts out (ALLOW_REG = false: its own
// state already fills the register budget), otherwise from an LDS table.  The
// kernel declares `__shared__ float2 s_tw[LoopTwiddles<...>::TABLE]` and calls
// init uniformly before its first transform.
template <int N, int E, bool ALLOW_REG = true>
struct LoopTwiddles {
  static constexpr bool REG = ALLOW_REG && E <= 8;
  static constexpr int TABLE = REG ? 1 : N;
  RegTwiddles<N, E> reg;
  const float2* table;
  __device__ __forceinline__ void init(int t, float2* s_tw) {
    if constexpr (REG) {
      reg.init(t);
    } else {
      load_lds_twiddles<N>(s_tw);
      table = s_tw;
    }
  }
  // stages S0 .. S1 - 1 of the transform (all by default)
  template <int DIR, bool PINGPONG = true, int S0 = 0, int S1 = 64>
  __device__ __forceinline__ void stages(float2 (&v)[E], int t, float2* lds0, float2* lds1) const {
    if constexpr (REG)
      fft_stages_reg<N, E, DIR, PINGPONG, S0, S1>(v, t, lds0, lds1, reg);
    else
      fft_stages<N, E, DIR, PINGPONG, S0, true, true, false, S1>(v, t, lds0, lds1, table);
  }
};

// Transform one sequence held in natural layout.  Every thread of the work-group
// must call this uniformly (it contains work-group barriers when N > E).
// lds0 / lds1: per-sequence regions of lds_len<N>() float2 each (lds1 unused
// unless PINGPONG).
template <int N, int E, int DIR, bool PINGPONG = false, bool PAIR_TW = true>
__device__ __forceinline__ void fft_seq(float2 (&v)[E], int t, float2* lds0, float2* lds1 = nullptr) {
  fft_stages<N, E, DIR, PINGPONG, 0, false, PAIR_TW>(v, t, lds0, lds1);
}

// The same transform with the stage twiddles read from an LDS table
// (load_lds_twiddles<N>).
template <int N, int E, int DIR, bool WAVE = false>
__device__ __forceinline__ void fft_seq_tw(float2 (&v)[E], int t, float2* lds, const float2* lds_tw) {
  fft_stages<N, E, DIR, false, 0, true, true, WAVE>(v, t, lds, nullptr, lds_tw);
}

// A row of a row pass: thread t < T = N/E of the row's T consecutive threads
// (row s = threadIdx.x / T).  With T <= 64 the row lies within one wave and
// its exchanges synchronise that wave only.
template <int N>
constexpr bool row_in_wave(int E) { return N / E <= 64; }
template <int N, int E, int DIR>
__device__ __forceinline__ void fft_row(float2 (&v)[E], int t, float2* lds) {
  fft_stages<N, E, DIR, false, 0, false, true, row_in_wave<N>(E)>(v, t, lds, nullptr);
}
template <int N, int E, int DIR>
__device__ __forceinline__ void fft_row_tw(float2 (&v)[E], int t, float2* lds, const float2* lds_tw) {
  fft_stages<N, E, DIR, false, 0, true, true, row_in_wave<N>(E)>(v, t, lds, nullptr, lds_tw);
}

// Elements per thread used for a transform of length N.
template <int N>
constexpr int elems_for() { return N <= 16 ? N : 16; }

}  // namespace ptx
