// Device FFT core for gfx950: radix-2/4/8 Stockham passes over a batch of
// lines held in LDS.
//
// Layout in LDS: a batch of L lines of length N, element (line l, position n)
// at buf[n * L + l]; consecutive threads take consecutive lines, so every
// pass reads and writes LDS at unit stride across a wavefront.
//
// Pass p (radix R, stride Ns = product of earlier radices) takes butterfly j
// (0 <= j < N/R) from positions j + r*N/R, multiplies input r by
// exp(DIR*2*pi*i * r*(j % Ns)/(Ns*R)), runs a length-R DFT and stores output r
// at (j/Ns)*Ns*R + (j % Ns) + r*Ns (Stockham autosort: output in natural
// order, no bit reversal).  Every thread loads all of its butterflies'
// inputs before a barrier and then stores, so the passes run in place.
// The last pass can leave its outputs in registers: output r of butterfly j
// is then position j + r*N/R.
//
// Twiddles come from a table tw[m] = exp(-2*pi*i*m/N) computed on the host in
// FP64 (rounded to T for the FP32 path); the inverse uses conj(tw).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace ph {

// Barrier for LDS hand-offs only: waits for this wave's LDS traffic but not
// for its outstanding global loads, so prefetches stay in flight across it.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

template <typename T>
struct alignas(2 * sizeof(T)) cx {
  T x, y;
};

// Two FP32 lanes handled as one element: the column passes transform two
// adjacent FP32 columns per slot (cx<f2> = re0 re1 im0 im1, 16 bytes), which
// gives FP32 tiles the FP64 shape (128-byte rows at 8 slots, the same
// registers per thread).  Scalars (twiddles, sigma, k) are splatted.
struct alignas(8) f2 {
  float a, b;
  f2() = default;
  __host__ __device__ constexpr f2(float a_, float b_) : a(a_), b(b_) {}
  __host__ __device__ constexpr f2(double v) : a(float(v)), b(float(v)) {}
  __host__ __device__ constexpr f2(float v) : a(v), b(v) {}
  __host__ __device__ constexpr f2(int v) : a(float(v)), b(float(v)) {}
};
__device__ __forceinline__ f2 operator+(f2 u, f2 v) { return f2(u.a + v.a, u.b + v.b); }
__device__ __forceinline__ f2 operator-(f2 u, f2 v) { return f2(u.a - v.a, u.b - v.b); }
__device__ __forceinline__ f2 operator*(f2 u, f2 v) { return f2(u.a * v.a, u.b * v.b); }
__device__ __forceinline__ f2 operator-(f2 u) { return f2(-u.a, -u.b); }

// Streaming global accesses.  Every pipeline buffer (states, spectra) is read
// or written once per kernel and is far larger than the caches, so the x and
// column kernels move it with non-temporal loads and stores (measured on the
// k_xmid access pattern at 512^3: 3.97 -> 4.86 TB/s); PH_NT_LOAD /
// PH_NT_STORE 0 give plain accesses.  The host pass and the CPU emulation
// take plain accesses.
#ifndef PH_NT_LOAD
#define PH_NT_LOAD 1
#endif
#ifndef PH_NT_STORE
#define PH_NT_STORE 1
#endif
template <int BYTES> struct StreamWord;
template <> struct StreamWord<4> { using type = unsigned int; };
template <> struct StreamWord<8> { using type = unsigned int __attribute__((ext_vector_type(2))); };
template <> struct StreamWord<16> { using type = unsigned int __attribute__((ext_vector_type(4))); };
template <typename V>
__device__ __forceinline__ V ld_stream(const V* p) {
#if defined(__HIP_DEVICE_COMPILE__) && PH_NT_LOAD
  using W = typename StreamWord<sizeof(V)>::type;
  return __builtin_bit_cast(V, __builtin_nontemporal_load(reinterpret_cast<const W*>(p)));
#else
  return *p;
#endif
}
template <typename V>
__device__ __forceinline__ void st_stream(V* p, const V& v) {
#if defined(__HIP_DEVICE_COMPILE__) && PH_NT_STORE
  using W = typename StreamWord<sizeof(V)>::type;
  __builtin_nontemporal_store(__builtin_bit_cast(W, v), reinterpret_cast<W*>(p));
#else
  *p = v;
#endif
}

template <typename T>
__device__ __forceinline__ cx<T> mk(T a, T b) { cx<T> r; r.x = a; r.y = b; return r; }
template <typename T>
__device__ __forceinline__ cx<T> operator+(cx<T> a, cx<T> b) { return mk(a.x + b.x, a.y + b.y); }
template <typename T>
__device__ __forceinline__ cx<T> operator-(cx<T> a, cx<T> b) { return mk(a.x - b.x, a.y - b.y); }
template <typename T>
__device__ __forceinline__ cx<T> cmul(cx<T> a, cx<T> b) {
  return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename T>
__device__ __forceinline__ cx<T> cscale(cx<T> a, T s) { return mk(a.x * s, a.y * s); }
// multiply by i*k (k real)
template <typename T>
__device__ __forceinline__ cx<T> mul_ik(cx<T> a, T k) { return mk(-a.y * k, a.x * k); }
// multiply by DIR*i
template <int DIR, typename T>
__device__ __forceinline__ cx<T> rot(cx<T> a) {
  return DIR < 0 ? mk(a.y, -a.x) : mk(-a.y, a.x);
}

template <int DIR, typename T>
__device__ __forceinline__ void dft2(cx<T>* v) {
  cx<T> a = v[0], b = v[1];
  v[0] = a + b;
  v[1] = a - b;
}

template <int DIR, typename T>
__device__ __forceinline__ void dft4(cx<T>* v) {
  cx<T> s02 = v[0] + v[2], d02 = v[0] - v[2];
  cx<T> s13 = v[1] + v[3], d13 = rot<DIR>(v[1] - v[3]);
  v[0] = s02 + s13;
  v[2] = s02 - s13;
  v[1] = d02 + d13;
  v[3] = d02 - d13;
}

template <int DIR, typename T>
__device__ __forceinline__ void dft8(cx<T>* v) {
  cx<T> e[4] = {v[0], v[2], v[4], v[6]};
  cx<T> o[4] = {v[1], v[3], v[5], v[7]};
  dft4<DIR>(e);
  dft4<DIR>(o);
  const T c = T(0.70710678118654752440084436210485);
  // o1 *= exp(DIR*i*pi/4), o2 *= DIR*i, o3 *= exp(DIR*3i*pi/4)
  cx<T> o1 = mk(c * (o[1].x - T(DIR) * o[1].y), c * (o[1].y + T(DIR) * o[1].x));
  cx<T> o2 = rot<DIR>(o[2]);
  cx<T> o3 = mk(-c * (o[3].x + T(DIR) * o[3].y), c * (T(DIR) * o[3].x - o[3].y));
  v[0] = e[0] + o[0];
  v[4] = e[0] - o[0];
  v[1] = e[1] + o1;
  v[5] = e[1] - o1;
  v[2] = e[2] + o2;
  v[6] = e[2] - o2;
  v[3] = e[3] + o3;
  v[7] = e[3] - o3;
}

// length-16 DFT as 4 x 4: n = j + 4m, k = k1 + 4k2; DFT-4 over m, twiddle
// exp(DIR*2*pi*i*j*k1/16), DFT-4 over j
template <int DIR, typename T>
__device__ __forceinline__ void dft16(cx<T>* v) {
  cx<T> t[4][4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    t[j][0] = v[j];
    t[j][1] = v[j + 4];
    t[j][2] = v[j + 8];
    t[j][3] = v[j + 12];
    dft4<DIR>(t[j]);
  }
  const T c1 = T(0.92387953251128675613), s1 = T(0.38268343236508977173);  // cos, sin pi/8
  const T c2 = T(0.70710678118654752440);                                   // cos pi/4
  // (cos, sin) of 2*pi*q/16 for q = j*k1 in 0..9
  const T cq[10] = {T(1), c1, c2, s1, T(0), -s1, -c2, -c1, T(-1), -c1};
  const T sq[10] = {T(0), s1, c2, c1, T(1), c1, c2, s1, T(0), -s1};
#pragma unroll
  for (int j = 1; j < 4; ++j)
#pragma unroll
    for (int k1 = 1; k1 < 4; ++k1) {
      const int q = j * k1;
      t[j][k1] = cmul(t[j][k1], mk(cq[q], T(DIR) * sq[q]));
    }
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1) {
    cx<T> u[4] = {t[0][k1], t[1][k1], t[2][k1], t[3][k1]};
    dft4<DIR>(u);
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2) v[k1 + 4 * k2] = u[k2];
  }
}

template <int R, int DIR, typename T>
__device__ __forceinline__ void dftR(cx<T>* v) {
  if constexpr (R == 2) dft2<DIR>(v);
  else if constexpr (R == 4) dft4<DIR>(v);
  else if constexpr (R == 8) dft8<DIR>(v);
  else dft16<DIR>(v);
}

constexpr int ilog2c(int n) { int k = 0; while ((1 << k) < n) ++k; return k; }

// Radix plan: radix-8 passes, preceded by one radix-4 pass (log2 N = 3q+2) or
// two radix-4 passes (log2 N = 3q+1, q >= 1); N = 2 is a single radix-2 pass.
template <int N>
struct Plan {
  static_assert(N >= 2 && (N & (N - 1)) == 0, "power-of-two lengths only");
  static constexpr int K = ilog2c(N);
  static constexpr int n8 = K / 3;
  static constexpr int rem = K % 3;
  static constexpr int npass = rem == 0 ? n8 : (rem == 1 ? (n8 >= 1 ? n8 + 1 : 1) : n8 + 1);
  static constexpr int radix(int p) {
    return rem == 0 ? 8 : (rem == 1 ? (n8 >= 1 ? (p < 2 ? 4 : 8) : 2) : (p == 0 ? 4 : 8));
  }
};

// Plan for transforms that run entirely in LDS (the x kernels): as Plan, but
// N = 1024 takes one radix-16 pass instead of two radix-4 passes, which saves
// one LDS round trip and barrier per transform (measured 4 % faster 2D steps
// at 1024^2; at 128 the radix-16 pass leaves too many threads idle).
template <int N>
struct PlanX {
  static constexpr int K = ilog2c(N);
  static constexpr bool R16 = K % 3 == 1 && K >= 10;
  static constexpr int npass = R16 ? K / 3 : Plan<N>::npass;
  static constexpr int radix(int p) { return R16 ? (p == 0 ? 16 : 8) : Plan<N>::radix(p); }
};

template <int N, int P, class PL = Plan<N>>
struct PassNs {  // product of the radices of passes 0..P-1
  static constexpr int value = PassNs<N, P - 1, PL>::value * PL::radix(P - 1);
};
template <int N, class PL>
struct PassNs<N, 0, PL> { static constexpr int value = 1; };

template <int N, int L, int NT>
struct LastRegs {
  static constexpr int R = Plan<N>::radix(Plan<N>::npass - 1);
  static constexpr int NB = L * (N / R);
  static constexpr int PER = (NB + NT - 1) / NT;
};

// load the R inputs of butterfly b of pass (R, Ns) and apply the twiddles
// LDS slot of (position, line) for line stride LS; PADB > 0 inserts one
// empty position after every 2^PADB positions (breaks power-of-two strides
// across LDS banks).
// PADB < 0 (LS a power of two): no padding; line l of position pos is
// rotated by pos LS / W, W = 16 (PADB = -1, 16-byte elements) or 32 (-2,
// 8-byte elements) slots of a 256-byte bank window, so that W consecutive
// positions of one line, and W consecutive slots of whole positions, each
// fall in distinct banks
template <int LS, int PADB>
__device__ __forceinline__ int lds_slot(int pos, int l) {
  if constexpr (PADB < 0) return pos * LS + ((l + ((pos * LS) >> (PADB == -1 ? 4 : 5))) & (LS - 1));
  else return (PADB > 0 ? pos + (pos >> PADB) : pos) * LS + l;
}
template <int N, int LS, int PADB>
struct LdsSize {
  static constexpr int value = (PADB > 0 ? N + (N >> PADB) : N) * LS;
};

// TWS: stride of the twiddle table (a table of exp(-2 pi i m / (TWS N))
// entries serves a length-N transform with TWS = 2: the single-row x kernels
// share one table between their N/2-point transforms and the real-to-complex
// split)
// Slot of twiddle m in a table stored with TSW = 1: entries swizzled within
// aligned blocks of 16 (m ^ (m / 16 mod 16)), so reads at strides that are
// multiples of 16 entries (the early passes of a transform) fall in distinct
// LDS banks
template <int TSW>
__device__ __forceinline__ int twslot(int m) { return TSW ? (m ^ ((m >> 4) & 15)) : m; }

template <typename T, int N, int L, int DIR, int R, int NS, int LS = L, int PADB = 0, int TWS = 1, int TSW = 0>
__device__ __forceinline__ void bfly_load(const cx<T>* buf, const cx<T>* tw, int b, cx<T>* v) {
  const int l = b % L, j = b / L;
#pragma unroll
  for (int r = 0; r < R; ++r) v[r] = buf[lds_slot<LS, PADB>(j + r * (N / R), l)];
  if constexpr (NS > 1) {
    const int k = j % NS;
#pragma unroll
    for (int r = 1; r < R; ++r) {
      cx<T> w = tw[twslot<TSW>(TWS * (r * k * (N / (NS * R))))];
      if constexpr (DIR > 0) w.y = -w.y;
      v[r] = cmul(v[r], w);
    }
  }
  dftR<R, DIR>(v);
}

// one Stockham pass, LDS -> LDS (in place; ends with a barrier).  Lines are
// interleaved with stride LS >= L (padding keeps 16-byte accesses free of
// bank conflicts when L is not a multiple of 8).
template <typename T, int N, int L, int DIR, int NT, int P, int LS = L, int PADB = 0, class PL = Plan<N>, int TWS = 1,
          int TSW = 0>
__device__ __forceinline__ void pass_lds(cx<T>* buf, const cx<T>* tw, int tid) {
  constexpr int R = PL::radix(P);
  constexpr int NS = PassNs<N, P, PL>::value;
  constexpr int NB = L * (N / R);
  constexpr int PER = (NB + NT - 1) / NT;
  cx<T> v[PER][R];
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int b = tid + q * NT;
    if (NB % NT == 0 || b < NB) bfly_load<T, N, L, DIR, R, NS, LS, PADB, TWS, TSW>(buf, tw, b, v[q]);
  }
  lds_barrier();
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int b = tid + q * NT;
    if (NB % NT == 0 || b < NB) {
      const int l = b % L, j = b / L;
      const int k = j % NS;
      const int d = (j / NS) * NS * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) buf[lds_slot<LS, PADB>(d + r * NS, l)] = v[q][r];
    }
  }
  lds_barrier();
}

template <typename T, int N, int L, int DIR, int NT, int P, int PEND, int LS = L, int PADB = 0,
          class PL = Plan<N>, int TWS = 1, int TSW = 0>
__device__ __forceinline__ void passes_lds_range(cx<T>* buf, const cx<T>* tw, int tid) {
  if constexpr (P < PEND) {
    pass_lds<T, N, L, DIR, NT, P, LS, PADB, PL, TWS, TSW>(buf, tw, tid);
    passes_lds_range<T, N, L, DIR, NT, P + 1, PEND, LS, PADB, PL, TWS, TSW>(buf, tw, tid);
  }
}

// Full transform of the L lines in buf (line stride LS), result in buf
// (natural order).  Caller must have synchronised after filling buf; ends
// with a barrier.
template <typename T, int N, int L, int DIR, int NT, int LS = L, int PADB = 0, class PL = Plan<N>, int TWS = 1,
          int TSW = 0>
__device__ __forceinline__ void fft_lds(cx<T>* buf, const cx<T>* tw, int tid) {
  passes_lds_range<T, N, L, DIR, NT, 0, PL::npass, LS, PADB, PL, TWS, TSW>(buf, tw, tid);
}

// Full transform with the last pass left in registers v[q][r]; output r of
// butterfly b = tid + q*NT is line b % L, position b / L + r*N/R (last_pos).
// A barrier separates the last pass's loads from the return, so the caller
// may overwrite buf at once.
template <typename T, int N, int L, int DIR, int NT>
__device__ __forceinline__ void fft_regs(cx<T>* buf, const cx<T>* tw, int tid,
                                         cx<T> (&v)[LastRegs<N, L, NT>::PER][LastRegs<N, L, NT>::R]) {
  constexpr int LP = Plan<N>::npass - 1;
  constexpr int R = Plan<N>::radix(LP);
  constexpr int NS = PassNs<N, LP>::value;
  constexpr int NB = L * (N / R);
  constexpr int PER = LastRegs<N, L, NT>::PER;
  passes_lds_range<T, N, L, DIR, NT, 0, LP>(buf, tw, tid);
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int b = tid + q * NT;
    if (NB % NT == 0 || b < NB) bfly_load<T, N, L, DIR, R, NS>(buf, tw, b, v[q]);
  }
  lds_barrier();
}

// As fft_regs, but the first pass takes its inputs from ld(line, position)
// (global memory, zero-fill, ...) instead of LDS, which saves the staging
// write/read and one barrier per transform.  The caller's previous use of buf
// must have ended with a barrier (fft_regs and fft_regs_g end with one).
template <typename T, int N, int L, int DIR, int NT, class Ld>
__device__ __forceinline__ void fft_regs_g(cx<T>* buf, const cx<T>* tw, int tid,
                                           cx<T> (&v)[LastRegs<N, L, NT>::PER][LastRegs<N, L, NT>::R],
                                           Ld ld) {
  constexpr int NP = Plan<N>::npass;
  constexpr int R0 = Plan<N>::radix(0);
  constexpr int NB0 = L * (N / R0);
  if constexpr (NP == 1) {
    constexpr int PER = LastRegs<N, L, NT>::PER;
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int b = tid + q * NT;
      if (NB0 % NT == 0 || b < NB0) {
        const int l = b % L, j = b / L;
#pragma unroll
        for (int r = 0; r < R0; ++r) v[q][r] = ld(l, j + r * (N / R0));
        dftR<R0, DIR>(v[q]);
      }
    }
  } else {
    constexpr int PER0 = (NB0 + NT - 1) / NT;
    cx<T> w[PER0][R0];
#pragma unroll
    for (int q = 0; q < PER0; ++q) {
      const int b = tid + q * NT;
      if (NB0 % NT == 0 || b < NB0) {
        const int l = b % L, j = b / L;
#pragma unroll
        for (int r = 0; r < R0; ++r) w[q][r] = ld(l, j + r * (N / R0));
        dftR<R0, DIR>(w[q]);
      }
    }
#pragma unroll
    for (int q = 0; q < PER0; ++q) {
      const int b = tid + q * NT;
      if (NB0 % NT == 0 || b < NB0) {
        const int l = b % L, j = b / L;
#pragma unroll
        for (int r = 0; r < R0; ++r) buf[(j * R0 + r) * L + l] = w[q][r];
      }
    }
    lds_barrier();
    constexpr int LP = NP - 1;
    constexpr int R = Plan<N>::radix(LP);
    constexpr int NS = PassNs<N, LP>::value;
    constexpr int NB = L * (N / R);
    constexpr int PER = LastRegs<N, L, NT>::PER;
    passes_lds_range<T, N, L, DIR, NT, 1, LP>(buf, tw, tid);
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int b = tid + q * NT;
      if (NB % NT == 0 || b < NB) bfly_load<T, N, L, DIR, R, NS>(buf, tw, b, v[q]);
    }
    lds_barrier();
  }
}

// Pass-0 inputs of a transform: w[q][r] = ld(line, j + r*N/R0) for the
// butterflies b = tid + q*NT of pass 0.
template <int N, int L, int NT>
struct P0Shape {
  static constexpr int R = Plan<N>::radix(0);
  static constexpr int NB = L * (N / R);
  static constexpr int PER = (NB + NT - 1) / NT;
};

template <typename T, int N, int L, int NT, class Ld>
__device__ __forceinline__ void load_pass0(int tid, Ld ld,
                                           cx<T> (&w)[P0Shape<N, L, NT>::PER][P0Shape<N, L, NT>::R]) {
  using P0 = P0Shape<N, L, NT>;
#pragma unroll
  for (int q = 0; q < P0::PER; ++q) {
    const int b = tid + q * NT;
    if (P0::NB % NT == 0 || b < P0::NB) {
      const int l = b % L, j = b / L;
#pragma unroll
      for (int r = 0; r < P0::R; ++r) w[q][r] = ld(l, j + r * (N / P0::R));
    }
  }
}

// Transform whose pass-0 inputs were loaded into w beforehand; as soon as w
// has been consumed, the pass-0 inputs of the next transform are loaded into
// it (when `more`), so those global loads overlap this transform's LDS
// passes.  Output as fft_regs (last pass in registers, ends with a barrier).
template <typename T, int N, int L, int DIR, int NT, class Ld>
__device__ __forceinline__ void fft_regs_pipe(cx<T>* buf, const cx<T>* tw, int tid,
                                              cx<T> (&v)[LastRegs<N, L, NT>::PER][LastRegs<N, L, NT>::R],
                                              cx<T> (&w)[P0Shape<N, L, NT>::PER][P0Shape<N, L, NT>::R],
                                              bool more, Ld next) {
  using P0 = P0Shape<N, L, NT>;
  constexpr int NP = Plan<N>::npass;
  if constexpr (NP == 1) {
    static_assert(P0::PER == LastRegs<N, L, NT>::PER, "single pass layout");
#pragma unroll
    for (int q = 0; q < P0::PER; ++q) {
      const int b = tid + q * NT;
      if (P0::NB % NT == 0 || b < P0::NB) {
#pragma unroll
        for (int r = 0; r < P0::R; ++r) v[q][r] = w[q][r];
        dftR<P0::R, DIR>(v[q]);
      }
    }
    if (more) load_pass0<T, N, L, NT>(tid, next, w);
  } else {
#pragma unroll
    for (int q = 0; q < P0::PER; ++q) {
      const int b = tid + q * NT;
      if (P0::NB % NT == 0 || b < P0::NB) {
        const int l = b % L, j = b / L;
        dftR<P0::R, DIR>(w[q]);
#pragma unroll
        for (int r = 0; r < P0::R; ++r) buf[(j * P0::R + r) * L + l] = w[q][r];
      }
    }
    lds_barrier();
    if (more) load_pass0<T, N, L, NT>(tid, next, w);
    constexpr int LP = NP - 1;
    constexpr int R = Plan<N>::radix(LP);
    constexpr int NS = PassNs<N, LP>::value;
    constexpr int NB = L * (N / R);
    constexpr int PER = LastRegs<N, L, NT>::PER;
    passes_lds_range<T, N, L, DIR, NT, 1, LP>(buf, tw, tid);
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int b = tid + q * NT;
      if (NB % NT == 0 || b < NB) bfly_load<T, N, L, DIR, R, NS>(buf, tw, b, v[q]);
    }
    lds_barrier();
  }
}

// Is register slot q valid for this thread, and which (line, position) is
// output r of it?
template <int N, int L, int NT>
__device__ __forceinline__ bool last_valid(int tid, int q) {
  constexpr int NB = LastRegs<N, L, NT>::NB;
  return (NB % NT == 0) || (tid + q * NT < NB);
}
template <int N, int L, int NT>
__device__ __forceinline__ void last_pos(int tid, int q, int r, int& l, int& n) {
  constexpr int R = LastRegs<N, L, NT>::R;
  const int b = tid + q * NT;
  l = b % L;
  n = b / L + r * (N / R);
}

}  // namespace ph
