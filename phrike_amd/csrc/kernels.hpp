// Pencil kernels of the pseudo-spectral Euler RHS and RK4 step (gfx950).
//
// One RHS evaluation in 3D is five launches (the 2D form skips the z pass):
//
//   k_xfwd   fluxes (in registers) + real-to-complex FFT along x of one
//            y-row (a half-length complex FFT), keeps |mx| <= Nx/3, scales by
//            sigma_x/(Nx*Ny*Nz).  3D Euler: the 12 distinct fluxes (the
//            momentum flux is symmetric), without i kx -> A[12][z][y][mx];
//            3D MHD: the 15 distinct ones; 2D Euler: the 7 distinct ones
//   k_col    y pass: forward FFT along y of column tiles, keeps |my| <= Ny/3,
//            forms G = sigma_y (i kx Fx^ + i ky Fy^) and sigma_y Fz^
//                                                   -> B[10][z][my][mx]
//   k_col    z pass: forward FFT along z of G and Fz^, D = sigma_z (G + i kz
//            Fz^) on |mz| <= Nz/3, zero elsewhere (3D Euler steps: also the
//            RK4 sum in spectral space, CM_FWDINV_DIVACC), inverse FFT along
//            z                                      -> [5][z][my][mx] (in B)
//   k_col    y inverse: zero-filled inverse FFT along y -> [5][z][y][mx] (in A)
//   k_xinv   complex-to-real inverse along x of one row, then the output (or
//            a stage combine) in registers.
//
// In a step, k_xmid joins the x inverse of one stage to the next stage's
// fluxes and x forward (the stage states never reach HBM), and k_xend writes
// the new state and forms the next step's fluxes and x forward.
//
// Because differentiation, the 2/3 mask and the filter are all diagonal in
// Fourier space, dx Fx + dy Fy + dz Fz is formed once in spectral space and
// only 5 fields are transformed back: the same operator as three separate
// masked derivatives, with 2/3 of the modes never written to HBM.
//
// All spectral buffers hold complex cx<T>; row pitch mxp >= mx keeps rows
// aligned.  The x kernels take NX/2 (FP64) or NX/4 (FP32) threads per row
// (at least 64); the column kernels N * TM / 8.
#pragma once
#include <type_traits>

#include "fft_core.hpp"

namespace ph {

constexpr int NT = 256;  // reductions and small helpers

constexpr int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Element type of the column passes: FP64 complex, or (PH_COL_PAIR_F32) two
// adjacent FP32 columns per slot.  FP32 spectral rows are then stored
// pair-planar: columns 2m, 2m+1 as re(2m) re(2m+1) im(2m) im(2m+1).
#ifndef PH_COL_PAIR_F32
#define PH_COL_PAIR_F32 1
#endif
template <typename T>
struct ColElem {
  using type = T;
  static constexpr int PW = 1;
};
#if PH_COL_PAIR_F32
template <>
struct ColElem<float> {
  using type = f2;
  static constexpr int PW = 2;
};
#endif

// 1/v, or 1 where v == 0 (per lane for a column pair)
__device__ __forceinline__ double rcp_or_one(double v) { return v == 0.0 ? 1.0 : 1.0 / v; }
__device__ __forceinline__ float rcp_or_one(float v) { return v == 0.f ? 1.f : 1.f / v; }
__device__ __forceinline__ f2 rcp_or_one(f2 v) { return f2(rcp_or_one(v.a), rcp_or_one(v.b)); }

// The y pass's order of the 12 distinct 3D Euler fluxes of the symmetric
// layout (euler3_flux_sym's u), each transformed once, 4 bits per step:
// mass (x, y, z), rho ux uy, xx, yy, xz, yz, zz, energy (x, y, z)
constexpr unsigned long long kEulerSymSeq =
    (0ull << 0) | (1ull << 4) | (2ull << 8) |                     // u0 u1 u2
    (6ull << 12) | (3ull << 16) | (4ull << 20) |                  // u6 u3 u4
    (7ull << 24) | (8ull << 28) | (5ull << 32) |                  // u7 u8 u5
    (9ull << 36) | (10ull << 40) | (11ull << 44);                 // u9 u10 u11
__device__ __forceinline__ int sym_seq(int k) { return int((kEulerSymSeq >> (4 * k)) & 15ull); }
// per step of that order: the output field of sigma F(Fz[c]) (5 + c) and of
// G[c] (c), 15 for none; whether the step's spectrum is kept for the next G,
// and whether the kept spectrum is that G's x flux (else its y flux)
constexpr unsigned long long kEulerSymZ =
    (15ull << 0) | (15ull << 4) | (5ull << 8) | (15ull << 12) | (15ull << 16) | (15ull << 20) |
    (6ull << 24) | (7ull << 28) | (8ull << 32) | (15ull << 36) | (15ull << 40) | (9ull << 44);
constexpr unsigned long long kEulerSymG =
    (15ull << 0) | (0ull << 4) | (15ull << 8) | (15ull << 12) | (1ull << 16) | (2ull << 20) |
    (15ull << 24) | (3ull << 28) | (15ull << 32) | (15ull << 36) | (4ull << 40) | (15ull << 44);
constexpr unsigned kEulerSymKeep = (1u << 0) | (1u << 3) | (1u << 6) | (1u << 9);
constexpr unsigned kEulerSymFirstX = (1u << 1) | (1u << 5) | (1u << 7) | (1u << 10);
__device__ __forceinline__ int sym_z(int k) { return int((kEulerSymZ >> (4 * k)) & 15ull); }
__device__ __forceinline__ int sym_g(int k) { return int((kEulerSymG >> (4 * k)) & 15ull); }
__device__ __forceinline__ bool sym_keep(int k) { return (kEulerSymKeep >> k) & 1u; }
__device__ __forceinline__ bool sym_first_x(int k) { return (kEulerSymFirstX >> k) & 1u; }

// The y pass's order of the 15 distinct ideal-MHD fluxes (mhd3_flux_sym's u)
// and, per step: the component c of a sigma F(Fz[c]) output (field 8 + c; 15
// none) and its sign; up to two G outputs (component, 15 none) with a form
// (1: i kx kept + i ky X, 2: i kx X + i ky kept, 3: i kx X, 4: i ky X) and a
// sign; whether the step's spectrum is kept.  Fz[Bz] = 0 is written as zeros.
constexpr unsigned long long kMhdSymSeq =
    (0ull << 0) | (1ull << 4) | (2ull << 8) | (6ull << 12) | (3ull << 16) | (4ull << 20) | (7ull << 24) |
    (8ull << 28) | (5ull << 32) | (9ull << 36) | (10ull << 40) | (11ull << 44) | (12ull << 48) | (13ull << 52) |
    (14ull << 56);
constexpr unsigned long long kMhdSymZ =
    (15ull << 0) | (15ull << 4) | (0ull << 8) | (15ull << 12) | (15ull << 16) | (15ull << 20) | (1ull << 24) |
    (2ull << 28) | (3ull << 32) | (15ull << 36) | (15ull << 40) | (4ull << 44) | (15ull << 48) | (5ull << 52) |
    (6ull << 56);
constexpr unsigned kMhdSymZNeg = (1u << 13) | (1u << 14);
constexpr unsigned long long kMhdSymG =
    (15ull << 0) | (0ull << 4) | (15ull << 8) | (15ull << 12) | (1ull << 16) | (2ull << 20) | (15ull << 24) |
    (3ull << 28) | (15ull << 32) | (15ull << 36) | (4ull << 40) | (15ull << 44) | (5ull << 48) | (15ull << 52) |
    (7ull << 56);
constexpr unsigned long long kMhdSymGForm =
    (1ull << 4) | (2ull << 16) | (1ull << 20) | (1ull << 28) | (1ull << 40) | (4ull << 48) | (1ull << 56);
constexpr unsigned kMhdSymGNeg = 1u << 12;  // G[Bx] = i ky (-(ux By - uy Bx))
constexpr int kMhdSymG2Step = 12, kMhdSymG2 = 6, kMhdSymG2Form = 3;  // G[By] = i kx (ux By - uy Bx)
constexpr unsigned kMhdSymKeep = (1u << 0) | (1u << 3) | (1u << 6) | (1u << 9) | (1u << 13);
__device__ __forceinline__ int mhd_seq(int k) { return int((kMhdSymSeq >> (4 * k)) & 15ull); }
__device__ __forceinline__ int mhd_z(int k) { return int((kMhdSymZ >> (4 * k)) & 15ull); }
__device__ __forceinline__ int mhd_g(int k) { return int((kMhdSymG >> (4 * k)) & 15ull); }
__device__ __forceinline__ int mhd_gform(int k) { return int((kMhdSymGForm >> (4 * k)) & 15ull); }

// ----------------------------------------------------------------------------
// retained-mode bookkeeping along one axis: positions n <= kp and
// n >= N - kn are kept; compact index j = n (n <= kp) or n - (N - nr).
// ----------------------------------------------------------------------------
struct Keep {
  int kp, kn, nr;
  __device__ __forceinline__ bool kept(int n, int N) const { return n <= kp || n >= N - kn; }
  __device__ __forceinline__ int compact(int n, int N) const { return n <= kp ? n : n - (N - nr); }
  __device__ __forceinline__ int natural(int j, int N) const { return j <= kp ? j : j + (N - nr); }
};

// ----------------------------------------------------------------------------
// x kernels: forward (fluxes or copies -> kept x modes), inverse (kept x modes
// -> physical rows, optionally with the RK4 stage combine) and the fused
// stage boundary k_xmid.
//
// The kept-mode count MX (NX/3 + 1 with the 2/3 rule, NX/2 + 1 without) is a
// compile-time constant; spectral rows move in 16-byte chunks: one FP64
// complex, or one pair-planar FP32 pair (re(2m) re(2m+1) im(2m) im(2m+1)).
// Only the chunks that hold kept modes are read or written (the column passes
// skip the row padding).
// ----------------------------------------------------------------------------

// radix plan of the x kernels' in-LDS transforms: PlanX (a radix-16 pass for
// log2 N = 3q+1) or the column kernels' Plan
#ifndef PH_X_RADIX16
#define PH_X_RADIX16 1
#endif
template <int N>
using XPlan = std::conditional_t<PH_X_RADIX16 != 0, PlanX<N>, Plan<N>>;

#ifndef PH_XT_MAX_F64
#define PH_XT_MAX_F64 1024
#endif
#ifndef PH_XT_MAX_F32
#define PH_XT_MAX_F32 512
#endif
// threads per block of the hybrid solver's x kernel (two rows per block)
template <typename T, int NX>
struct XThreads {
  static constexpr int value = sizeof(T) == 8 ? clampi(NX, 64, PH_XT_MAX_F64) : clampi(NX / 2, 64, PH_XT_MAX_F32);
};

// LDS line stride (complex values per x position) of the x kernels: by
// default the line count itself (dense, smallest LDS footprint, which sets
// how many blocks share a CU); PH_XLS > 0 pads every position to PH_XLS.
#ifndef PH_XLS
#define PH_XLS 0
#endif
// PH_XODD: an odd stride (lines | 1), so 16-byte accesses of consecutive
// positions by consecutive lanes fall in distinct banks
#ifndef PH_XODD
#define PH_XODD 1
#endif
// PH_XROT: power-of-two line counts keep their stride and rotate the lines
// of each position instead (lds_slot with PADB < 0), for 16-byte elements
// (bit 0: FP64) and 8-byte elements (bit 1: FP32)
#ifndef PH_XROT
#define PH_XROT 1
#endif
// optional position padding of the x kernels' LDS lines (see lds_slot)
#ifndef PH_XPADB
#define PH_XPADB 0
#endif
constexpr bool xrot(int lines, int es) {
  return PH_XLS == 0 && lines > 1 && (lines & (lines - 1)) == 0 &&
         ((es >= 16 && (PH_XROT & 1)) || (es < 16 && (PH_XROT & 2)));
}
constexpr int xstride(int lines, int es) {
  return PH_XLS > 0 ? PH_XLS : (xrot(lines, es) ? lines : (PH_XODD ? (lines | 1) : lines));
}
constexpr int xpadb(int lines, int es) { return xrot(lines, es) ? (es >= 16 ? -1 : -2) : PH_XPADB; }
template <int LINES, typename T>
__device__ __forceinline__ int xslot(int pos, int l) {
  return lds_slot<xstride(LINES, sizeof(cx<T>)), xpadb(LINES, sizeof(cx<T>))>(pos, l);
}
template <int NX, int LINES, typename T>
constexpr int xbuf_size() { return LdsSize<NX, xstride(LINES, sizeof(cx<T>)), xpadb(LINES, sizeof(cx<T>))>::value; }

// Twiddles of the x kernels: an LDS copy per block (PH_XTW_GLOBAL 0), or read
// where they lie in global memory through the L1 cache (1).
#ifndef PH_XTW_GLOBAL
#define PH_XTW_GLOBAL 0
#endif
// The LDS copy is stored swizzled (twslot<1>; PH_XTW_SWZ 0: plain).
#if PH_XTW_GLOBAL
#define PH_XTW_SWZ 0
#define PH_XTW_DECL(T, NX) const cx<T>* tw = nullptr
#define PH_XTW_FILL(src, NX, NT, tid) tw = (src)
#else
#ifndef PH_XTW_SWZ
#define PH_XTW_SWZ 1
#endif
#define PH_XTW_DECL(T, NX) __shared__ cx<T> tw[NX]
#define PH_XTW_FILL(src, NX, NT, tid) \
  for (int i_ = (tid); i_ < (NX); i_ += (NT)) tw[twslot<PH_XTW_SWZ>(i_)] = (src)[i_]
#endif
// k_xmid loads the RK combine's inputs before its inverse transform (1: U and
// the accumulator, 2: U only, 0: none) so their latency overlaps it
#ifndef PH_XMID_PREFETCH
#define PH_XMID_PREFETCH 1
#endif
// the same for the 8 ideal-MHD fields (twice the registers per prefetched field set)
#ifndef PH_XMID_PREFETCH_MHD
#define PH_XMID_PREFETCH_MHD 2
#endif
// rows of one slab per x-kernel block (1 or 2; every axis length is a
// multiple of 2)
#ifndef PH_X_ROWS
#define PH_X_ROWS 1
#endif
// minimum waves per SIMD requested for the x kernels (1: compiler's choice)
#ifndef PH_X_WAVES_PER_EU
#define PH_X_WAVES_PER_EU 1
#endif

// the 16-byte unit of a spectral row
struct alignas(16) pair4f {
  float a, b, c, d;
};
template <typename T, int K>
struct ChunkT { using type = cx<T>; };
template <>
struct ChunkT<float, 2> { using type = pair4f; };
template <typename T>
__device__ __forceinline__ cx<T> ch_get(const cx<T>& c, int) { return c; }
__device__ __forceinline__ cx<float> ch_get(const pair4f& c, int i) { return i == 0 ? mk(c.a, c.c) : mk(c.b, c.d); }
template <typename T>
__device__ __forceinline__ void ch_set(cx<T>& c, int, cx<T> v) { c = v; }
__device__ __forceinline__ void ch_set(pair4f& c, int i, cx<float> v) {
  if (i == 0) {
    c.a = v.x;
    c.c = v.y;
  } else {
    c.b = v.x;
    c.d = v.y;
  }
}

// quotient of the flux primitives: IEEE division in FP64 (the reference's
// operation); FP32 multiplies by the hardware reciprocal (1 ulp)
__device__ __forceinline__ double xrcp(double v) { return 1.0 / v; }
__device__ __forceinline__ float xrcp(float v) { return __builtin_amdgcn_rcpf(v); }
template <typename T>
__device__ __forceinline__ T xquot(T a, T b, T rb) {
  if constexpr (sizeof(T) == 8) return a / b;
  else return a * rb;
}


// ----------------------------------------------------------------------------
// x forward
// ----------------------------------------------------------------------------
enum XFwdMode { XF_FLUX3 = 0, XF_FLUX2 = 1, XF_COPY = 2, XF_MHD3 = 3, XF_COPY8 = 4 };
// copy modes transform the conserved values themselves (no fluxes)
constexpr bool xf_copy(int mode) { return mode == XF_COPY || mode == XF_COPY8; }

template <typename T>
struct XFwdArgs {
  const T* in;
  long long in_fs, in_ss;  // field / slab strides (elements); rows of nx
  cx<T>* out;
  long long out_fs, out_ss;  // rows of mxp complex
  int nf;                    // XF_COPY: number of fields (<= 5; XF_COPY8: <= 8)
  int mx, mxp;
  unsigned ikmask;  // bit f: multiply field f by i*kx (uniform over a group of lines)
  const cx<T>* tw;
  const T* sig;  // sigma_x(m) * scale, m < mx
  const T* kk;   // kx(m), zero at the Nyquist mode
  T gamma;
  const int* active;  // per-slab flag (batched 2D) or null
  int act_all;        // 1: active[0] gates every slab (a device-resident 3D run loop)
};

// is this block's slab switched off?  active[slab], or active[0] for all
__device__ __forceinline__ bool gated_off(const int* active, int act_all, int slab) {
  return active != nullptr && active[act_all ? 0 : slab] == 0;
}

template <int MODE> struct XFwdShape;
// 3D Euler: the 12 distinct fluxes (the momentum flux tensor is symmetric),
// 3 groups of 4 lines; see euler3_flux_sym
template <> struct XFwdShape<XF_FLUX3> { static constexpr int NC = 5, LX = 4, NG = 3; };
// 2D Euler: the 7 distinct fluxes (euler2_flux_sym) in groups of 4 and 3
// lines, without i kx (the y pass CM_FWDINV_DIV2S applies it); PH_2D_SYM 0:
// the 8 fluxes, x-fluxes times i kx
#ifndef PH_2D_SYM
#define PH_2D_SYM 1
#endif
template <> struct XFwdShape<XF_FLUX2> { static constexpr int NC = 4, LX = 4, NG = 2; };
template <> struct XFwdShape<XF_COPY>  { static constexpr int NC = 5, LX = 5, NG = 1; };
// ideal MHD: the 15 distinct fluxes in three groups of 5 (mhd3_flux_sym; the
// y pass CM_FWD_DIV3M applies i kx; PH_MHD_SYM 0: the 24 fluxes of the three
// directions, i kx in the x pass)
#ifndef PH_MHD_SYM
#define PH_MHD_SYM 1
#endif
template <> struct XFwdShape<XF_MHD3>  { static constexpr int NC = 8, LX = PH_MHD_SYM ? 5 : 8, NG = 3; };
template <> struct XFwdShape<XF_COPY8> { static constexpr int NC = 8, LX = 8, NG = 1; };

// The Euler flux transforms of a row in one FFT over all NG*LX lines instead
// of NG groups of LX (FP32 rows up to PH_XF_MERGE_F32_MAXNX points, FP64 up to
// PH_XF_MERGE_F64_MAXNX): the radix passes keep ~94 % of the lanes busy
// instead of 62-83 % and take a third of the barriers, for NG times the LDS.
// Off by default: at 256^3 FP32 the kernels then need 139 VGPRs and the step
// took 4.47 instead of 4.19 ms.
#ifndef PH_XF_MERGE_F32_MAXNX
#define PH_XF_MERGE_F32_MAXNX 0
#endif
#ifndef PH_XF_MERGE_F64_MAXNX
#define PH_XF_MERGE_F64_MAXNX 0
#endif
template <typename T, int NX, int MODE>
struct XMerge {
  static constexpr bool value = (MODE == XF_FLUX3 || MODE == XF_FLUX2) &&
                                NX <= (sizeof(T) == 4 ? PH_XF_MERGE_F32_MAXNX : PH_XF_MERGE_F64_MAXNX);
  // LDS lines of the forward transform
  static constexpr int lines = value ? XFwdShape<MODE>::LX * XFwdShape<MODE>::NG : XFwdShape<MODE>::LX;
};


// ----------------------------------------------------------------------------
// column passes (y and z): tiles of TM consecutive mx columns, lines along the
// transform axis with element stride ls.
// ----------------------------------------------------------------------------
enum ColMode {
  CM_FWD_DIV3 = 0,    // 3D y forward: A(12 or 15) -> B(10), compact out
  CM_FWD_COPY = 1,    // forward, sigma (and i*k), compact out, nc fields
  CM_FWDINV_DIV = 2,  // forward both halves, D = sigma (X + i k Z), inverse, full out
  CM_FWDINV_COPY = 3, // forward, sigma (and i*k) on kept modes, inverse, full out
  CM_INV = 4,         // compact in, zero-filled inverse, full out
  CM_FWDINV_CORR = 5, // forward, (sx sy sz - 1) on kept modes, inverse, full out
  CM_FWDINV_PROJ = 6, // three fields (B) together: forward, sigma and the Leray
                      // projection on kept modes, inverse, full out
  CM_FWDINV_DIVACC = 7, // CM_FWDINV_DIV and the RK4 sum of the step's RHS spectra
                        // (below)
  CM_FWD_DIV3M = 8,     // 3D ideal-MHD y forward, symmetric flux layout: A(15) -> B(16)
  CM_FWDINV_DIV2S = 9   // 2D y pass, symmetric flux layout: the 7 fluxes -> 4 fields, in place
};

template <typename T>
struct ColArgs {
  const cx<T>* in;
  long long in_fs, in_bs, in_ls;
  cx<T>* out;
  long long out_fs, out_bs, out_ls;
  int ncol;   // valid columns
  int ncolw;  // columns a store may cover (the row pitch; tiles then write
              // whole 128-byte lines, zeros past ncol); 0: ncol
  int nc;     // output fields
  int cpb;    // fields per block along blockIdx.z (0: all in every block; not CM_FWD_DIV3)
  Keep keep;  // retained modes along this axis
  int ikflag; // COPY modes: multiply by i*k
  const cx<T>* tw;
  const T* sig;  // by compact index
  const T* kk;   // by compact index
  const int* active;  // per blockIdx.y flag (batched 2D runs) or null
  int act_all;        // 1: active[0] gates every block
  // CM_FWDINV_CORR: sigma - 1 of this axis (by compact index), of the column
  // axis (by column) and of the axis of blockIdx.y (by blockIdx.y)
  const T* sm1;
  const T* colsm1;
  const T* rowsm1;
  // CM_FWDINV_PROJ: k of the column axis (by column) and of the axis of
  // blockIdx.y (by blockIdx.y); this axis's k is kk
  const T* colk;
  const T* rowk;
  // CM_FWDINV_DIVACC: the step's RK4 sum in 3D spectral space, S (kept
  // modes of all three axes; element (field c, kept z j, blockIdx.y, column)
  // at acc + c acc_fs + j acc_ls + blockIdx.y acc_bs + column).  With D this
  // stage's RHS spectrum (-D the transform of k), stages 1-3 leave
  // S - w D in S (stage 1 starts from S = F U, the input state's masked
  // spectrum) and transform D back; stage 4 leaves U' = sigma_x sigma_y
  // sigma_z (S - dt/6 D), the filtered new state's spectrum, in S and
  // transforms U' back.  w = dt/6, dt/3, dt/3, dt/6; dt = dt_dev[0] or dt.
  cx<T>* acc;
  long long acc_fs, acc_bs, acc_ls;
  int stage;
  double dt;
  const double* dt_dev;
  const T* colsig;  // sigma of the column axis (by column)
  const T* rowsig;  // sigma of the axis of blockIdx.y (by blockIdx.y)
  // CM_FWD_DIV3 with the 12-field symmetric 3D Euler flux layout (nc = 5;
  // order kEulerSymSeq): F_x[c] is read without i kx, which is applied here
  // (colk)
  int symflux;
};

#ifndef PH_COL_ELEMS_F64
#define PH_COL_ELEMS_F64 4096
#endif
#ifndef PH_COL_ELEMS_F32
#define PH_COL_ELEMS_F32 4096
#endif
// FP32 modes with one transform per field (no second spectrum held in LDS)
// take twice the columns, so a row segment is 128 bytes
#ifndef PH_COL_ELEMS_F32_SINGLE
#define PH_COL_ELEMS_F32_SINGLE 8192
#endif

// the projection mode holds three spectra of a tile in LDS (two besides the
// transform buffer): 2048 elements per spectrum keep it at ~100 KB per block
#ifndef PH_COL_ELEMS_PROJ
#define PH_COL_ELEMS_PROJ 2048
#endif

// Two-input modes hold the first spectrum in registers (YREG) or in a second
// LDS array (default: measured equal in FP32 at 8-column tiles, and the
// registers spill in FP64 or at 16 FP32 columns).
#ifndef PH_COL_TM_MAX
#define PH_COL_TM_MAX 8
#endif
#ifndef PH_COL_PER_THREAD_F32
#define PH_COL_PER_THREAD_F32 8
#endif
#ifndef PH_COL_YREG_F32
#define PH_COL_YREG_F32 0
#endif
#ifndef PH_COL_YREG_F64
#define PH_COL_YREG_F64 0
#endif

template <typename T, int N, int MODE>
struct ColTile {
  static constexpr bool PAIR =
      (MODE == CM_FWD_DIV3 || MODE == CM_FWDINV_DIV || MODE == CM_FWDINV_DIVACC || MODE == CM_FWD_DIV3M ||
       MODE == CM_FWDINV_DIV2S);
  static constexpr bool YREG = PAIR && ((sizeof(T) == 8) ? PH_COL_YREG_F64 : PH_COL_YREG_F32);
  static constexpr int ELEMS = MODE == CM_FWDINV_PROJ
                                   ? PH_COL_ELEMS_PROJ
                                   : ((sizeof(T) == 8) ? PH_COL_ELEMS_F64
                                                       : (PAIR ? PH_COL_ELEMS_F32 : PH_COL_ELEMS_F32_SINGLE));
  // columns per tile: PH_COL_TM_MAX 16-byte elements at most (128-byte rows),
  // so short lines make more, smaller blocks rather than wider rows
  static constexpr int TM = clampi(ELEMS / N < PH_COL_TM_MAX ? ELEMS / N : PH_COL_TM_MAX, 1, 64);
  // complex values of each register array per thread
  static constexpr int PERT = (sizeof(T) == 8) ? 8 : PH_COL_PER_THREAD_F32;
  static constexpr int NT = clampi((N * TM) / PERT, 64, 1024);
};

// loads of the spectral RK4 sum issued together per thread (CM_FWDINV_DIVACC)
#ifndef PH_ZACC_BATCH
#define PH_ZACC_BATCH 4
#endif
#ifndef PH_COL_WAVES_PER_EU
#define PH_COL_WAVES_PER_EU 2
#endif

// Block -> tile.  Tiles narrower than a 128-byte line (TM < 8 at long lines)
// share each line with the next tile, and blocks go round-robin over the 8
// XCDs, each with its own L2: in x order the two tiles of a line sit on two
// XCDs and the line is fetched from HBM twice.  Within each run of 16 blocks
// of a row, blocks i and i + 8 (one XCD, dispatched together) take tiles 2i
// and 2i + 1; a shorter last run pairs what it can and keeps the rest in
// order (a permutation of the row's tiles).  PH_COL_XCD_PAIR_ALL applies it
// at every tile width (tests of the mapping).
#ifndef PH_COL_XCD_PAIR
#define PH_COL_XCD_PAIR 1
#endif
#ifndef PH_COL_XCD_PAIR_ALL
#define PH_COL_XCD_PAIR_ALL 0
#endif
__device__ __forceinline__ int xcd_pair_tile(int b, int g) {
  const int base = b & ~15, t = g - base < 16 ? g - base : 16, i = b - base;
  if (t <= 8) return b;
  if (i < t - 8) return base + 2 * i;
  if (i >= 8) return base + 2 * (i - 8) + 1;
  return base + 2 * (t - 8) + (i - (t - 8));
}

// Per block, the global-source transforms run in a fixed order (field
// sequence below); the pass-0 inputs of transform k+1 are loaded while
// transform k runs its LDS passes.  For the two-input modes the first
// spectrum waits in registers (Y) or in a second LDS array (ybuf) while the
// second is transformed; each thread reads back only the slots it wrote, and
// every output is written to HBM once.
template <typename T, int N, int MODE>
__global__ __launch_bounds__((ColTile<T, N, MODE>::NT), PH_COL_WAVES_PER_EU) void k_col(ColArgs<T> a) {
  constexpr int TM = ColTile<T, N, MODE>::TM;
  constexpr int NT = ColTile<T, N, MODE>::NT;
  using LR = LastRegs<N, TM, NT>;
  using P0 = P0Shape<N, TM, NT>;
  constexpr int PER = LR::PER, R = LR::R;
  __shared__ cx<T> tw[N];
  __shared__ cx<T> buf[N * TM];
  if (gated_off(a.active, a.act_all, blockIdx.y)) return;
  const int tid = threadIdx.x;
  constexpr bool XPAIR = PH_COL_XCD_PAIR_ALL || (PH_COL_XCD_PAIR && TM * sizeof(cx<T>) < 128);
  const int c0 = (XPAIR ? xcd_pair_tile(int(blockIdx.x), int(gridDim.x)) : int(blockIdx.x)) * TM;
  for (int i = tid; i < N; i += NT) tw[i] = a.tw[i];
  const cx<T>* in = a.in + (long long)blockIdx.y * a.in_bs;
  cx<T>* out = a.out + (long long)blockIdx.y * a.out_bs;
  const Keep kp = a.keep;
  const long long ls = a.in_ls, ols = a.out_ls;
  const int ncol = a.ncol;
  const int ncolw = a.ncolw > 0 ? a.ncolw : ncol;
  const int nc = a.nc;
  // tile rows of a whole 128-byte line stream (non-temporal); narrower rows
  // (long lines: TM < 8) keep cached accesses, so the two halves of a line,
  // read and written by the tiles of one XCD (xcd_pair_tile), meet in L2
  // rather than in HBM
  constexpr bool LINE = TM * sizeof(cx<T>) >= 128;
  auto ld = [](const cx<T>* p) {
    if constexpr (LINE) return ld_stream(p);
    else return *p;
  };
  auto st = [](cx<T>* p, const cx<T>& v) {
    if constexpr (LINE) st_stream(p, v);
    else *p = v;
  };
  cx<T> X[PER][R];
  cx<T> W[P0::PER][P0::R];
  // first spectrum of a pair, at the kept positions this thread owns
  constexpr bool PAIR = ColTile<T, N, MODE>::PAIR;
  constexpr bool YREG = ColTile<T, N, MODE>::YREG;
  __shared__ cx<T> ybuf[(PAIR && !YREG) ? N * TM : 1];
  // CM_FWDINV_PROJ: the x and y components' spectra while the z one is transformed
  __shared__ cx<T> pbuf[MODE == CM_FWDINV_PROJ ? 2 * N * TM : 1];
  cx<T> Y[YREG ? PER : 1][R];
  auto first = [&](int q, int r, int l, int n) {
    if constexpr (YREG) return Y[q][r];
    else return ybuf[n * TM + l];
  };

  // CM_FWD_DIV3, symmetric layout: kx of the tile's columns
  constexpr bool SYMY = MODE == CM_FWD_DIV3 || MODE == CM_FWD_DIV3M || MODE == CM_FWDINV_DIV2S;
  __shared__ T s_colk[SYMY ? TM : 1];
  if constexpr (SYMY) {
    if ((MODE != CM_FWD_DIV3 || a.symflux) && tid < TM) s_colk[tid] = c0 + tid < ncolw ? a.colk[c0 + tid] : T(0);  // twiddle barrier below
  }
  constexpr int PERC = (MODE == CM_FWD_DIV3) ? 3 : ((MODE == CM_FWDINV_DIV || MODE == CM_FWDINV_DIVACC) ? 2 : 1);
  // this block's fields cb..ce-1 (all, or a.cpb of them by blockIdx.z:
  // launches of few tiles then have more, shorter blocks)
  const int cb = (!SYMY && a.cpb > 0) ? int(blockIdx.z) * a.cpb : 0;
  const int ce = (!SYMY && a.cpb > 0) ? (cb + a.cpb < nc ? cb + a.cpb : nc) : nc;
  // field read by the k-th global-source transform of this block
  auto field_of = [&](int k) {
    if constexpr (MODE == CM_FWD_DIV3M) return mhd_seq(k);
    else if constexpr (MODE == CM_FWDINV_DIV2S) return int((0x6532410ull >> (4 * k)) & 15ull);  // 0 1 4 2 3 5 6
    else if constexpr (MODE == CM_FWD_DIV3) return a.symflux ? sym_seq(k) : (k % 3) * nc + k / 3;
    else if constexpr (MODE == CM_FWDINV_DIV || MODE == CM_FWDINV_DIVACC) return (k % 2) * nc + cb + k / 2;
    else return cb + k;
  };
  const int nsrc = MODE == CM_FWD_DIV3M ? 15
                   : (MODE == CM_FWDINV_DIV2S ? 7 : ((MODE == CM_FWD_DIV3 && a.symflux) ? 12 : PERC * (ce - cb)));
  // element (line l, position n) of field f's tile, zero outside the grid
  auto loader = [&](int k) {
    const cx<T>* src = in + (long long)field_of(k) * a.in_fs + c0;
    return [=](int l, int n) {
      if constexpr (MODE == CM_INV)
        return (c0 + l < ncol && kp.kept(n, N)) ? ld(src + (long long)kp.compact(n, N) * ls + l)
                                                : mk(T(0), T(0));
      else
        return (c0 + l < ncol) ? ld(src + (long long)n * ls + l) : mk(T(0), T(0));
    };
  };
  // visit every (q, r) output slot of this thread: f(q, r, l, n, col_ok)
  auto each = [&](auto&& f) {
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      if (!last_valid<N, TM, NT>(tid, q)) continue;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        int l, n;
        last_pos<N, TM, NT>(tid, q, r, l, n);
        f(q, r, l, n, c0 + l < ncolw);
      }
    }
  };
  auto keep_first = [&]() {
    each([&](int q, int r, int l, int n, bool ok) {
      if constexpr (YREG) Y[q][r] = X[q][r];
      else if (kp.kept(n, N)) ybuf[n * TM + l] = X[q][r];
    });
  };
  int k = 0;
  // next global-source transform into X; with pf its last passes overlap the
  // loads of the following transform's first pass (into W)
  auto fwd = [&](auto dir, bool pf = true) {
    constexpr int D = decltype(dir)::value;
    const bool more = pf && k + 1 < nsrc;
    fft_regs_pipe<T, N, TM, D, NT>(buf, tw, tid, X, W, more, loader(more ? k + 1 : k));
    ++k;
  };
  using FWD = std::integral_constant<int, -1>;
  using INV = std::integral_constant<int, +1>;
  load_pass0<T, N, TM, NT>(tid, loader(0), W);
  lds_barrier();  // twiddles
#pragma unroll 1
  for (int c = cb; c < ce; ++c) {
    cx<T>* oc = out + (long long)c * a.out_fs + c0;
    if constexpr (MODE == CM_INV) {
      fwd(INV{});
      each([&](int q, int r, int l, int n, bool ok) {
        if (ok) st(oc + (long long)n * ols + l, X[q][r]);
      });
    } else if constexpr (MODE == CM_FWD_COPY) {
      fwd(FWD{});
      each([&](int q, int r, int l, int n, bool ok) {
        if (ok && kp.kept(n, N)) {
          const int j = kp.compact(n, N);
          cx<T> v = cscale(X[q][r], a.sig[j]);
          if (a.ikflag) v = mul_ik(v, a.kk[j]);
          st(oc + (long long)j * ols + l, v);
        }
      });
    } else if constexpr (MODE == CM_FWD_DIV3M) {
      {
        // ideal MHD, symmetric layout (nc = 8): the 15 distinct fluxes, each
        // transformed once (order mhd_seq), feed G[c] -> field c and
        // sigma F(Fz[c]) -> field 8 + c by the step tables above
#pragma unroll 1
        for (int s = 0; s < 15; ++s) {
          fwd(FWD{});
          const int zc = mhd_z(s), g1 = mhd_g(s), f1 = mhd_gform(s);
          const bool zneg = (kMhdSymZNeg >> s) & 1u, g1neg = (kMhdSymGNeg >> s) & 1u;
          const bool keep = (kMhdSymKeep >> s) & 1u, two = s == kMhdSymG2Step;
          cx<T>* oz = out + (long long)(8 + zc) * a.out_fs + c0;
          cx<T>* og1 = out + (long long)g1 * a.out_fs + c0;
          cx<T>* og2 = out + (long long)kMhdSymG2 * a.out_fs + c0;
          each([&](int q, int r, int l, int n, bool ok) {
            const cx<T> x = X[q][r];
            if (ok && kp.kept(n, N)) {
              const int j = kp.compact(n, N);
              const T sg = a.sig[j], kx = s_colk[l], ky = a.kk[j];
              auto g_of = [&](int form) {
                if (form == 3) return mul_ik(x, kx);
                if (form == 4) return mul_ik(x, ky);
                const cx<T> y0 = first(q, r, l, n);
                return form == 1 ? mul_ik(y0, kx) + mul_ik(x, ky) : mul_ik(x, kx) + mul_ik(y0, ky);
              };
              if (zc < 15) st(oz + (long long)j * ols + l, cscale(zneg ? mk(-x.x, -x.y) : x, sg));
              if (g1 < 15) {
                const cx<T> v = g_of(f1);
                st(og1 + (long long)j * ols + l, cscale(g1neg ? mk(-v.x, -v.y) : v, sg));
              }
              if (two) st(og2 + (long long)j * ols + l, cscale(g_of(kMhdSymG2Form), sg));
            }
            if (keep) {
              if constexpr (YREG) Y[q][r] = x;
              else if (kp.kept(n, N)) ybuf[n * TM + l] = x;
            }
          });
        }
        cx<T>* oz7 = out + (long long)15 * a.out_fs + c0;  // Fz[Bz] = 0
        each([&](int q, int r, int l, int n, bool ok) {
          if (ok && kp.kept(n, N)) st(oz7 + (long long)kp.compact(n, N) * ols + l, mk(T(0), T(0)));
        });
        break;
      }
    } else if constexpr (MODE == CM_FWD_DIV3) {
      if (a.symflux) {
        // 3D Euler, symmetric layout (nc = 5; no i kx in the x pass): the 12
        // distinct fluxes are transformed once each, in the order sym_seq,
        // and every spectrum serves all the outputs that use it:
        // G[c] = sigma (i kx F(Fx[c]) + i ky F(Fy[c])) -> field c (sym_g),
        // sigma F(Fz[c]) -> field 5 + c (sym_z).  G is formed with the kept
        // spectrum (first: the x flux where sym_first_x, else the y flux).
#pragma unroll 1
        for (int s = 0; s < 12; ++s) {
          fwd(FWD{});
          const int zf = sym_z(s), gf = sym_g(s);
          const bool fix = sym_first_x(s), keep = sym_keep(s);
          cx<T>* oz = out + (long long)zf * a.out_fs + c0;
          cx<T>* og = out + (long long)gf * a.out_fs + c0;
          each([&](int q, int r, int l, int n, bool ok) {
            const cx<T> x = X[q][r];
            if (ok && kp.kept(n, N)) {
              const int j = kp.compact(n, N);
              const T sg = a.sig[j];
              if (zf < 15) st(oz + (long long)j * ols + l, cscale(x, sg));
              if (gf < 15) {
                const cx<T> y0 = first(q, r, l, n);
                const cx<T> fx = fix ? y0 : x;
                const cx<T> fy = fix ? x : y0;
                st(og + (long long)j * ols + l, cscale(mul_ik(fx, s_colk[l]) + mul_ik(fy, a.kk[j]), sg));
              }
            }
            if (keep) {
              if constexpr (YREG) Y[q][r] = x;
              else if (kp.kept(n, N)) ybuf[n * TM + l] = x;
            }
          });
        }
        break;
      }
      // G = sigma (X + i k Z) with X = F(i kx Fx^), Z = F(Fy^); then sigma
      // F(Fz^)
      fwd(FWD{});
      keep_first();
      fwd(FWD{});
      each([&](int q, int r, int l, int n, bool ok) {
        if (ok && kp.kept(n, N)) {
          const int j = kp.compact(n, N);
          st(oc + (long long)j * ols + l, cscale(first(q, r, l, n) + mul_ik(X[q][r], a.kk[j]), a.sig[j]));
        }
      });
      cx<T>* o2 = out + (long long)(nc + c) * a.out_fs + c0;
      fwd(FWD{});
      each([&](int q, int r, int l, int n, bool ok) {
        if (ok && kp.kept(n, N)) {
          const int j = kp.compact(n, N);
          st(o2 + (long long)j * ols + l, cscale(X[q][r], a.sig[j]));
        }
      });
    } else if constexpr (MODE == CM_FWDINV_PROJ) {
      // nc = 3: this iteration does all three fields.  Forward transforms
      // of the x and y components wait in pbuf (each thread reads back only
      // the slots it wrote), the z component stays in X; on kept modes
      // sigma, then B^ -= k (k.B^)/|k|^2 (|k|^2 = 1 at k = 0); the inverse
      // transforms run in place in buf, pbuf and pbuf + N TM.
      cx<T>* pb0 = pbuf;
      cx<T>* pb1 = pbuf + N * TM;
      auto keep_to = [&](cx<T>* dst) {
        each([&](int q, int r, int l, int n, bool ok) { dst[n * TM + l] = X[q][r]; });
      };
      fwd(FWD{});
      keep_to(pb0);
      fwd(FWD{});
      keep_to(pb1);
      fwd(FWD{});
      const T ky = a.rowk[blockIdx.y];
      each([&](int q, int r, int l, int n, bool ok) {
        cx<T> bx = mk(T(0), T(0)), by = bx, bz = bx;
        if (kp.kept(n, N)) {
          const int j = kp.compact(n, N);
          const T sg = a.sig[j], kx = a.colk[c0 + l], kz = a.kk[j];
          bx = cscale(pb0[n * TM + l], sg);
          by = cscale(pb1[n * TM + l], sg);
          bz = cscale(X[q][r], sg);
          const T k2 = (kx * kx + ky * ky) + kz * kz;
          const cx<T> kb = mk((kx * bx.x + ky * by.x) + kz * bz.x, (kx * bx.y + ky * by.y) + kz * bz.y);
          const T inv = rcp_or_one(k2);
          const cx<T> cf = mk(kb.x * inv, kb.y * inv);
          bx = bx - cscale(cf, kx);
          by = by - cscale(cf, ky);
          bz = bz - cscale(cf, kz);
        }
        pb0[n * TM + l] = bx;
        pb1[n * TM + l] = by;
        buf[n * TM + l] = bz;
      });
      lds_barrier();
#pragma unroll
      for (int f = 0; f < 3; ++f) {
        fft_regs<T, N, TM, +1, NT>(f == 0 ? pb0 : (f == 1 ? pb1 : buf), tw, tid, X);
        cx<T>* of = out + (long long)f * a.out_fs + c0;
        each([&](int q, int r, int l, int n, bool ok) {
          if (ok) st(of + (long long)n * ols + l, X[q][r]);
        });
      }
      break;
    } else {
      // CM_FWDINV_DIV / CM_FWDINV_DIVACC / CM_FWDINV_COPY / CM_FWDINV_CORR:
      // forward, combine on kept modes, inverse
      constexpr bool DIV = MODE == CM_FWDINV_DIV || MODE == CM_FWDINV_DIVACC;
      constexpr bool DIV2S = MODE == CM_FWDINV_DIV2S;
      if constexpr (DIV2S) {
        // fields 0 1 | 4 2 | 3 | 5 6 for components 0..3: each component but
        // 2 keeps its first spectrum (rho ux, rho ux uy, (E+p) ux); component
        // 2 transforms its y flux only and reuses rho ux uy, still kept, as
        // its x flux.  The kept spectrum is component 1's y flux and the x
        // flux of the others.
        if (c != 2) {
          fwd(FWD{});
          keep_first();
        }
        fwd(FWD{});
      } else {
        fwd(FWD{});
      }
      if constexpr (DIV) {
        keep_first();
        // CM_FWDINV_DIVACC loads the next field's first pass after the RK4
        // sum instead, so W is free during it
        fwd(FWD{}, MODE != CM_FWDINV_DIVACC);
      }
      each([&](int q, int r, int l, int n, bool ok) {
        cx<T> v = mk(T(0), T(0));
        if (kp.kept(n, N)) {
          const int j = kp.compact(n, N);
          if constexpr (DIV2S) {
            const cx<T> kf = first(q, r, l, n), x = X[q][r];
            const T kx = s_colk[l], ky = a.kk[j];
            v = cscale(c == 1 ? mul_ik(x, kx) + mul_ik(kf, ky) : mul_ik(kf, kx) + mul_ik(x, ky), a.sig[j]);
          } else if constexpr (DIV) {
            v = cscale(first(q, r, l, n) + mul_ik(X[q][r], a.kk[j]), a.sig[j]);
          } else if constexpr (MODE == CM_FWDINV_CORR) {
            // sx sy sz - 1 from the three (sigma - 1): no cancellation
            const T sa = a.colsm1[c0 + l], sb = a.rowsm1[blockIdx.y], sc = a.sm1[j];
            const T f = ((sa + sb) + sc) + ((sa * sb + sa * sc) + sb * sc) + (sa * sb) * sc;
            v = cscale(X[q][r], f);
          } else {
            v = cscale(X[q][r], a.sig[j]);
            if (a.ikflag) v = mul_ik(v, a.kk[j]);
          }
        }
        buf[n * TM + l] = v;
      });
      lds_barrier();
      if constexpr (MODE == CM_FWDINV_DIVACC) {
        // the RK4 sum on the tile's kept modes, D read back from buf in
        // whole tile rows (X is free here)
        const double dtd = a.dt_dev != nullptr ? a.dt_dev[0] : a.dt;
        const bool last = a.stage >= 4;
        const T w = T(dtd * ((a.stage == 1 || last) ? 1.0 / 6.0 : 1.0 / 3.0));
        const T sy = a.rowsig[blockIdx.y];
        cx<T>* sc = a.acc + (long long)c * a.acc_fs + (long long)blockIdx.y * a.acc_bs + c0;
        constexpr int NTASK = ((N / 3) * 2 + 1) * TM;  // kept positions at most (2/3 rule)
        constexpr int NIT = (NTASK + NT - 1) / NT;
        // the thread's loads of S in batches of SB issued together (a store
        // between two loads would order them: the compiler cannot tell the
        // slots apart), SB bounded so the batch fits next to the prefetch
        // registers of the next transform
        constexpr int SB = PH_ZACC_BATCH < NIT ? PH_ZACC_BATCH : NIT;
#pragma unroll
        for (int i0 = 0; i0 < NIT; i0 += SB) {
          cx<T> s0[SB];
#pragma unroll
          for (int ii = 0; ii < SB; ++ii) {
            const int t = tid + (i0 + ii) * NT;
            const int j = t / TM, l = t - j * TM;
            s0[ii] = (i0 + ii < NIT && j < kp.nr && c0 + l < ncolw) ? ld(sc + (long long)j * a.acc_ls + l)
                                                                    : mk(T(0), T(0));
          }
#pragma unroll
          for (int ii = 0; ii < SB; ++ii) {
            const int i = i0 + ii;
            if (i >= NIT) break;
            const int t = tid + i * NT;
            const int j = t / TM, l = t - j * TM;
            if (j < kp.nr && c0 + l < ncolw) {
              const int n = kp.natural(j, N);
              cx<T>* sp = sc + (long long)j * a.acc_ls + l;
              const cx<T> s1 = s0[ii] - cscale(buf[n * TM + l], w);
              if (last) {
                const cx<T> u = cscale(s1, (a.colsig[c0 + l] * sy) * a.sig[j]);
                st(sp, u);
                buf[n * TM + l] = u;
              } else {
                st(sp, s1);
              }
            }
          }
        }
        if (last) lds_barrier();
        if (k < nsrc) load_pass0<T, N, TM, NT>(tid, loader(k), W);
      }
      fft_regs<T, N, TM, +1, NT>(buf, tw, tid, X);
      each([&](int q, int r, int l, int n, bool ok) {
        if (ok) st(oc + (long long)n * ols + l, X[q][r]);
      });
    }
  }
}

// ----------------------------------------------------------------------------
// x inverse + stage combine
// ----------------------------------------------------------------------------
enum XInvMode {
  XI_OUT = 0,      // out = value
  XI_RHS = 1,      // out = -value
  XI_STAGE3 = 2,   // 3D RK4 stage combine
  XI_STAGE2 = 3,   // 2D RK4 stage combine with positivity clamp
  XI_ADD = 4,      // out += value
  XI_PROJE = 5,    // MHD: value = projected B (3 fields) -> state fields 5-7,
                   // E recomputed at the thermal pressure of the unprojected state
  XI_MHDSTAGE = 6  // MHD: k = -value; RK4 stage combine in the reference's order
                   // and the clamp, next state to us, the k sum in out
};

template <typename T>
struct XInvArgs {
  const cx<T>* in;
  long long in_fs, in_ss;  // rows of mxp
  int mx, mxp;
  int nf;                  // XI_OUT / XI_RHS: fields (<= NC)
  T* out;                  // output (and the RK accumulator in stage modes)
  const T* un;             // state at the start of the step
  T* us;                   // next stage state
  long long fs, ss;        // physical field / slab strides (elements); rows of nx
  int stage;               // 1..4
  double dt;               // used when dt_dev is null
  const double* dt_dev;    // device dt: dt_dev[slab] (dt_per_slab) or dt_dev[0]; or null
  int dt_per_slab;         // 1: one dt per slab (batched 2D runs)
  int keep_state;          // k_xmid stage 4: also write the new state to out
  int specacc;             // k_xmid stages 1-3: no accumulator (the RK4 sum is
                           // kept in spectral space by the z pass)
  const int* active;       // per-slab flag (batched 2D) or null
  int act_all;             // 1: active[0] gates every slab
  T gamma, rho_min, p_min;
  const cx<T>* tw;
  // XI_OUT / XI_ADD / XI_STAGE2 stage 4 with the 4 or 5 Euler fields: the
  // max wave speed of the row just written -> cfl_rows[slab * rows + y]
  // (null: none)
  double* cfl_rows;
};

// Wave speed of one point from its conserved values u[0..DIM+1].  No FMA
// contraction, so the standalone reduction and the fused one of the step
// (which takes these values from registers) give the same bits.
template <typename T, int DIM>
__device__ __forceinline__ double wave_speed_v(const T* u, T gamma) {
#pragma clang fp contract(off)
  // 3D: max(|ux|,|uy|,|uz|) + a (phrike/equations.py:199-211);
  // 2D: |ux| + a (phrike/equations.py:137-147)
  const T rho = u[0];
  if constexpr (DIM == 3) {
    const T ux = u[1] / rho, uy = u[2] / rho, uz = u[3] / rho;
    const T kin = (T(0.5) * rho) * ((ux * ux + uy * uy) + uz * uz);
    const T p = (gamma - T(1)) * (u[4] - kin);
    const T a = sqrt(gamma * p / rho);
    T m = fabs(ux) > fabs(uy) ? fabs(ux) : fabs(uy);
    m = m > fabs(uz) ? m : fabs(uz);
    return double(m + a);
  } else {
    const T ux = u[1] / rho, uy = u[2] / rho;
    const T kin = (T(0.5) * rho) * (ux * ux + uy * uy);
    const T p = (gamma - T(1)) * (u[3] - kin);
    const T a = sqrt(gamma * p / rho);
    return double(fabs(ux) + a);
  }
}
template <typename T, int DIM>
__device__ __forceinline__ double wave_speed(const T* u, long long fs, long long i, T gamma) {
  T v[DIM + 2];
#pragma unroll
  for (int c = 0; c < DIM + 2; ++c) v[c] = u[c * fs + i];
  return wave_speed_v<T, DIM>(v, gamma);
}

__device__ __forceinline__ double dmax_nan(double a, double b) {
  // NaN wins (a NaN speed must not be hidden)
  if (a != a) return a;
  if (b != b) return b;
  return a > b ? a : b;
}

// Max of m over the block -> *dst: xor shuffles within each wavefront, then
// the block's wavefronts through LDS.  Every thread of the block calls it.
template <int NTB>
__device__ __forceinline__ void block_max_store(double m, double* dst, int tid) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = dmax_nan(m, __shfl_xor(m, o, 64));
  constexpr int NW = NTB / 64;
  if constexpr (NW == 1) {
    if (tid == 0) *dst = m;
  } else {
    __shared__ double wred[NW];
    if ((tid & 63) == 0) wred[tid >> 6] = m;
    __syncthreads();
    if (tid == 0) {
      double r = wred[0];
#pragma unroll
      for (int w = 1; w < NW; ++w) r = dmax_nan(r, wred[w]);
      *dst = r;
    }
  }
}

template <typename T>
__device__ __forceinline__ void clamp2d(T* u, T gamma, T rho_min, T p_min) {
  // phrike/solver.py:467-488: (rho,ux,uy,p) -> clamp rho, p -> conservative
  const T rho = u[0];
  const T ux = u[1] / rho, uy = u[2] / rho;
  const T kin = (T(0.5) * rho) * (ux * ux + uy * uy);
  const T p = (gamma - T(1)) * (u[3] - kin);
  const T rc = rho < rho_min ? rho_min : rho;  // NaN propagates like np.maximum
  const T pc = p < p_min ? p_min : p;
  u[0] = rc;
  u[1] = rc * ux;
  u[2] = rc * uy;
  u[3] = pc / (gamma - T(1)) + (T(0.5) * rc) * (ux * ux + uy * uy);
}

// RK4 stage combine at one point (phrike/solver.py:1220-1230 3D, 502-513 2D)
// with k_s = -v.  The accumulator a.out carries the step's sum so far,
// acc = U + dt/6 k1 (stage 1), + dt/3 k2, + dt/3 k3 (stages 2, 3); stages 1-3
// return the next stage state U + c k_s in u, stage 4 returns acc + dt/6 k4:
// the reference's U + dt/6 (k1 + 2 k2 + 2 k3 + k4), summed in another order
// so that stage 4 needs no U.  The 2D form clamps u (phrike/solver.py:467-488).
template <typename T, int NC, bool CLAMP>
__device__ __forceinline__ void rk_point(const XInvArgs<T>& a, const T (&v)[NC], long long o, double dtd,
                                         T (&u)[NC], const T (&un)[NC], const T (&acc)[NC]) {
  const T half = T(0.5 * dtd), full = T(dtd), sixth = T(dtd / 6.0), third = T(dtd / 3.0);
  if (a.stage < 4) {
    const T coef = (a.stage == 3) ? full : half;
    const T w = (a.stage == 1) ? sixth : third;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const T k = -v[c];
      const long long idx = c * a.fs + o;
      a.out[idx] = (a.stage == 1 ? un[c] : acc[c]) + w * k;
      u[c] = un[c] + coef * k;
    }
  } else {
#pragma unroll
    for (int c = 0; c < NC; ++c) u[c] = acc[c] + sixth * (-v[c]);
  }
  if constexpr (CLAMP) clamp2d(u, a.gamma, a.rho_min, a.p_min);
}

// MHD positivity clamp through prim -> cons (phrike/solver.py:956-964):
// rho >= rho_min, thermal p >= p_min, velocities and B kept
template <typename T>
__device__ __forceinline__ void clamp_mhd(T (&u)[8], T gamma, T rho_min, T p_min) {
  const T rho = u[0];
  const T ux = u[1] / rho, uy = u[2] / rho, uz = u[3] / rho;
  const T Bx = u[5], By = u[6], Bz = u[7];
  const T kin = (T(0.5) * rho) * ((ux * ux + uy * uy) + uz * uz);
  const T mag = T(0.5) * ((Bx * Bx + By * By) + Bz * Bz);
  const T p = (gamma - T(1)) * ((u[4] - kin) - mag);
  const T rc = rho < rho_min ? rho_min : rho;  // NaN propagates like np.maximum
  const T pc = p < p_min ? p_min : p;
  u[0] = rc;
  u[1] = rc * ux;
  u[2] = rc * uy;
  u[3] = rc * uz;
  u[4] = (pc / (gamma - T(1)) + (T(0.5) * rc) * ((ux * ux + uy * uy) + uz * uz)) + mag;
}

// MHD stage combine at one point (phrike/solver.py:988-1006) with k = -v:
// stage 1: acc = k; stages 2, 3: acc = acc + 2k; next state U + c k (c = dt/2,
// dt/2, dt); stage 4: U + dt/6 (acc + k).  Then the clamp rho >= rho_min,
// p >= p_min through prim -> cons (phrike/solver.py:956-964) into u.
template <typename T>
__device__ __forceinline__ void mhd_stage_vals(const XInvArgs<T>& a, const T (&v)[8], const T (&un)[8],
                                               const T (&acc)[8], double dtd, T (&u)[8], T (&acc_new)[8]) {
  const T c = T(a.stage == 3 ? dtd : 0.5 * dtd), sixth = T(dtd / 6.0);
#pragma unroll
  for (int f = 0; f < 8; ++f) {
    const T k = -v[f];
    acc_new[f] = a.stage == 1 ? k : acc[f] + T(2) * k;
    u[f] = a.stage < 4 ? un[f] + c * k : un[f] + sixth * (acc[f] + k);
  }
  clamp_mhd(u, a.gamma, a.rho_min, a.p_min);
}

// Fused stage boundary: x inverse + RK combine of stage s, then (stages 1-3)
// the fluxes of the stage s+1 state and their forward x transform, or (stage
// 4) the new state's forward x transform for the post-step filter, which
// always follows and writes the state itself (with keep_state the new state
// is written here too: the filter then adds its correction to it).  The
// intermediate stage states
// never go to HBM.  The spectra are read from and written to the same rows of
// the same buffer; the block loads its rows completely before it writes them.
template <typename T>
struct XMidArgs {
  XInvArgs<T> inv;
  XFwdArgs<T> fwd;       // stages 1-3
  XFwdArgs<T> fwd_last;  // stage 4 (copy mode, 5 or 4 fields)
};


// ----------------------------------------------------------------------------
// Single-row x kernels.  A block owns row y of one slab for every field.  The
// real row x[0..NX) is packed as z[p] = x[2p] + i x[2p+1] (p < M = NX/2) and
// transformed with an M-point complex FFT in LDS.  With E, O the spectra of
// the even and odd samples, Z = E + i O, so the half spectrum is
//   X[k] = (Z[k] + Z*[M-k])/2 - i W^k (Z[k] - Z*[M-k])/2,  W = exp(-2 pi i/NX),
// k = 0..M (indices mod M), and the inverse packs the kept modes back as
//   Z'[k] = (X[k] + X*[M-k]) + i W^-k (X[k] - X*[M-k]),
//   z[p] = sum_k Z'[k] exp(+2 pi i k p/M)   (x[2p] = Re z, x[2p+1] = Im z),
// taking X[0] and X[M] real (the real part of the full inverse, as the
// reference's ifft(...).real does).  Thread t owns the packed points
// p = t + j*NT (j < PJ), i.e. the x pairs (2p, 2p+1), which it loads and
// stores as one complex value (16 bytes in FP64).  Blocks are small (NX/2
// FP64, NX/4 FP32 threads), so several share a CU and the load, transform
// and store phases of different blocks overlap.
// ----------------------------------------------------------------------------
// threads per row: NX / PH_XR_DIV_F64 (FP64) or NX / PH_XR_DIV_F32 (FP32), at
// least one wave; a thread owns (NX/2) / threads packed points (at the
// defaults one in FP64, two adjacent ones in FP32: one 16-byte access)
#ifndef PH_XR_DIV_F64
#define PH_XR_DIV_F64 2
#endif
// unroll the 3D Euler flux groups of the x forward transforms (0: a loop)
#ifndef PH_XF_UNROLL_G
#define PH_XF_UNROLL_G 1
#endif
#ifndef PH_XR_DIV_F32
#define PH_XR_DIV_F32 4
#endif
template <typename T, int NX>
struct XRThreads {
  static constexpr int value =
      sizeof(T) == 8 ? clampi(NX / PH_XR_DIV_F64, 64, 512) : clampi(NX / PH_XR_DIV_F32, 64, 256);
};

template <typename T, int NX, bool DA>
struct XRGeo {
  static constexpr int NT = XRThreads<T, NX>::value;
  static constexpr int M = NX / 2;                         // packed points of a row
  static constexpr int MX = DA ? NX / 3 + 1 : NX / 2 + 1;  // kept x modes
  static constexpr int K = ColElem<T>::PW;                 // modes per 16-byte chunk
  static constexpr int MC = (MX + K - 1) / K;              // chunks holding kept modes
  // chunks per row pitch (mxp: 128-byte lines, 8 chunks).  Stores cover the
  // whole pitch, zeros past the kept modes: a partly written cache line
  // costs more than a whole one (15 rows of 171 FP64 chunks 2.73 ms, of 176
  // chunks 2.01 ms at 512^3)
  static constexpr int MCP = (MC + 7) / 8 * 8;
  static constexpr bool PALL = M >= NT;                    // every thread owns points
  static constexpr int PJ = PALL ? M / NT : 1;             // packed points per thread
  // FP32 threads own two adjacent points (x = 4t..4t+3, one 16-byte access);
  // otherwise point p = t + j*NT
  static constexpr bool CONSEC = sizeof(T) == 4 && PJ == 2;
  static __device__ __forceinline__ int point(int tid, int j) { return CONSEC ? 2 * tid + j : tid + j * NT; }
};

// the thread's x pairs of one row (row: element 0 of the row): loaded /
// stored as PJ complex values, one 16-byte access per thread in FP32
template <typename G, typename T>
__device__ __forceinline__ void xr_load_pairs(const T* row, int tid, cx<T> (&v)[G::PJ]) {
  if constexpr (G::CONSEC) {
    const pair4f q = ld_stream(reinterpret_cast<const pair4f*>(row) + tid);
    v[0] = mk(q.a, q.b);
    v[1] = mk(q.c, q.d);
  } else {
#pragma unroll
    for (int j = 0; j < G::PJ; ++j) {
      const int p = G::point(tid, j);
      v[j] = (G::PALL || p < G::M) ? ld_stream(reinterpret_cast<const cx<T>*>(row) + p) : mk(T(0), T(0));
    }
  }
}
template <typename G, typename T>
__device__ __forceinline__ void xr_store_pairs(T* row, int tid, const cx<T> (&v)[G::PJ]) {
  if constexpr (G::CONSEC) {
    pair4f q;
    q.a = v[0].x;
    q.b = v[0].y;
    q.c = v[1].x;
    q.d = v[1].y;
    st_stream(reinterpret_cast<pair4f*>(row) + tid, q);
  } else {
#pragma unroll
    for (int j = 0; j < G::PJ; ++j) {
      const int p = G::point(tid, j);
      if (G::PALL || p < G::M) st_stream(reinterpret_cast<cx<T>*>(row) + p, v[j]);
    }
  }
}
// LDS of a packed row: M transform points and the Nyquist slot M
template <int NX, int LINES, typename T>
constexpr int xrbuf_size() {
  return LdsSize<NX / 2 + 1, xstride(LINES, sizeof(cx<T>)), xpadb(LINES, sizeof(cx<T>))>::value;
}

template <typename T, int NX, bool DA>
__device__ __forceinline__ void xr_tables(const T* sig, const T* kk, T* s_sig, T* s_k, int tid) {
  using G = XRGeo<T, NX, DA>;
  for (int i = tid; i < G::MX; i += G::NT) {
    s_sig[i] = sig[i];
    s_k[i] = kk[i];
  }
}

// Kept modes of one row of NCI fields, loaded into registers in 16-byte
// chunks; place() writes them to LDS at their natural positions (line l =
// field l) with zeros in the band MX..M.
template <typename T, int NX, bool DA, int NCI>
struct XRSpec {
  using G = XRGeo<T, NX, DA>;
  using C = typename ChunkT<T, G::K>::type;
  static constexpr int NTASK = NCI * G::MC;
  static constexpr int PER = (NTASK + G::NT - 1) / G::NT;
  C a[PER];

  __device__ __forceinline__ void load(const cx<T>* src, long long fs, int nfl, int tid) {
    const C* s = reinterpret_cast<const C*>(src);
    const long long fsc = fs / G::K;
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int t = tid + q * G::NT;
      a[q] = C{};
      if (NTASK % G::NT == 0 || t < NTASK) {
        const int l = t / G::MC, m = t - l * G::MC;
        if (l < nfl) a[q] = ld_stream(s + l * fsc + m);
      }
    }
  }

  __device__ __forceinline__ void place(cx<T>* buf, int tid) const {
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int t = tid + q * G::NT;
      if (NTASK % G::NT == 0 || t < NTASK) {
        const int l = t / G::MC, m = t - l * G::MC;
#pragma unroll
        for (int i = 0; i < G::K; ++i) {
          const int k = m * G::K + i;
          if (k < G::MX) buf[xslot<NCI, T>(k, l)] = ch_get(a[q], i);
        }
      }
    }
    constexpr int ZB = G::M + 1 - G::MX;
    if constexpr (ZB > 0) {
#pragma unroll
      for (int i = 0; i < (NCI * ZB + G::NT - 1) / G::NT; ++i) {
        const int t = tid + i * G::NT;
        if ((NCI * ZB) % G::NT != 0 && t >= NCI * ZB) break;
        const int l = t / ZB, k = G::MX + (t - l * ZB);
        buf[xslot<NCI, T>(k, l)] = mk(T(0), T(0));
      }
    }
  }
};

// natural-order half spectra in LDS (positions 0..M) -> packed spectrum Z'
// (positions 0..M-1), in place: the task of pair (k, M-k) reads and writes
// only those two slots
template <typename T, int NX, int LINES, int NT>
__device__ __forceinline__ void xr_pack(cx<T>* buf, const cx<T>* tw, int tid) {
  constexpr int M = NX / 2, H = M / 2 + 1;
#pragma unroll
  for (int i = 0; i < (LINES * H + NT - 1) / NT; ++i) {
    const int t = tid + i * NT;
    if ((LINES * H) % NT != 0 && t >= LINES * H) break;
    const int l = t / H, k = t - l * H, k2 = M - k;
    cx<T> a = buf[xslot<LINES, T>(k, l)], b = buf[xslot<LINES, T>(k2, l)];
    if (k == 0) {  // X[0], X[M]: real parts
      a.y = T(0);
      b.y = T(0);
    }
    const cx<T> bc = mk(b.x, -b.y);
    const cx<T> w1 = tw[twslot<PH_XTW_SWZ>(k)], w2 = tw[twslot<PH_XTW_SWZ>(k2)];
    const cx<T> d1 = cmul(mk(w1.x, -w1.y), a - bc);  // W^-k (X[k] - X*[M-k])
    buf[xslot<LINES, T>(k, l)] = (a + bc) + mk(-d1.y, d1.x);
    if (k != 0 && k2 != k) {
      const cx<T> ac = mk(a.x, -a.y);
      const cx<T> d2 = cmul(mk(w2.x, -w2.y), b - ac);
      buf[xslot<LINES, T>(k2, l)] = (b + ac) + mk(-d2.y, d2.x);
    }
  }
}

// Z (M-point forward transforms of LX packed lines, in LDS) -> kept modes
// X[k] = (E + W^k O) * sig[k] (* i k[k] where bit (l / GL) * GL of ikmask is
// set) of lines l < nl, stored as 16-byte chunks to dst (one row of the
// line's field; field stride fs)
template <typename T, int NX, bool DA, int LX, int GL = LX, bool IK = true>
__device__ __forceinline__ void xr_split_store(const cx<T>* buf, const cx<T>* tw, cx<T>* dst, long long fs, int nl,
                                               const T* sig, const T* kk, unsigned ikmask, int tid) {
  using G = XRGeo<T, NX, DA>;
  using C = typename ChunkT<T, G::K>::type;
  constexpr int NTASK = LX * G::MCP, M = G::M;
  C* d = reinterpret_cast<C*>(dst);
  const long long fsc = fs / G::K;
#pragma unroll
  for (int i = 0; i < (NTASK + G::NT - 1) / G::NT; ++i) {
    const int t = tid + i * G::NT;
    const int l = t / G::MCP, m = t - l * G::MCP;
    if ((NTASK % G::NT == 0 || t < NTASK) && l < nl) {
      C c{};
#pragma unroll
      for (int i = 0; i < G::K; ++i) {
        const int k = m * G::K + i;
        if (k < G::MX) {
          const cx<T> a = buf[xslot<LX, T>(k & (M - 1), l)];
          const cx<T> z2 = buf[xslot<LX, T>((M - k) & (M - 1), l)];
          const cx<T> e = mk(T(0.5) * (a.x + z2.x), T(0.5) * (a.y - z2.y));
          const cx<T> h = mk(T(0.5) * (a.x - z2.x), T(0.5) * (a.y + z2.y));  // (Z[k] - Z*[M-k]) / 2
          cx<T> X = e + cmul(tw[twslot<PH_XTW_SWZ>(k)], mk(h.y, -h.x));       // E + W^k (-i h)
          X = cscale(X, sig[k]);
          if (IK && ((ikmask >> ((l / GL) * GL)) & 1u)) X = mul_ik(X, kk[k]);  // i k of line l's group
          ch_set(c, i, X);
        }
      }
      st_stream(d + l * fsc + m, c);
    }
  }
}

// Euler flux of direction g, component l from the primitives w = (rho, ux,
// uy, uz, p, E + p) (phrike/equations.py:119-134, 167-196); the direction is
// picked by selects (a register index by a run-time g would put w in scratch
// memory)
template <typename T, int NC>
__device__ __forceinline__ T euler_flux(const T* w, int g, int l) {
  const T ud = g == 0 ? w[1] : (g == 1 ? w[2] : w[3]);
  const T md = w[0] * ud;
  if (l == 0) return md;
  if (l == NC - 1) return w[5] * ud;
  T v = md * w[l];
  if (l == g + 1) v = v + w[4];
  return v;
}

// The 12 distinct 3D Euler fluxes from the primitives w = (rho, ux, uy, uz, p,
// E + p) (phrike/equations.py:167-196).  Of the 15 fluxes F_d[c], the mass
// fluxes are rho u_d and the momentum fluxes rho u_i u_j + p delta_ij are
// symmetric in (i, j), so F_x[rho v] = F_y[rho u] etc.:
//   u = 0..2   rho ux, rho uy, rho uz              (F_x,y,z[rho])
//   u = 3..5   (rho ux) ux + p, (rho uy) uy + p, (rho uz) uz + p
//   u = 6..8   (rho ux) uy, (rho ux) uz, (rho uy) uz
//   u = 9..11  (E + p) ux, (E + p) uy, (E + p) uz   (F_x,y,z[E])
// Each is formed as the reference forms F_x[.] (mx' = rho ux times u_j), so
// the shared off-diagonal entries differ from the reference's F_y[rho u] =
// (rho uy) ux by one rounding at most.
template <typename T>
__device__ __forceinline__ T euler3_flux_sym(const T* w, int u) {
  const T rux = w[0] * w[1], ruy = w[0] * w[2], ruz = w[0] * w[3];
  const T f1 = u < 3 ? w[0]
                     : ((u == 3 || u == 6 || u == 7) ? rux : ((u == 4 || u == 8) ? ruy : (u == 5 ? ruz : w[5])));
  const T f2 = (u == 0 || u == 3 || u == 9) ? w[1] : ((u == 1 || u == 4 || u == 6 || u == 10) ? w[2] : w[3]);
  const T v = f1 * f2;
  return (u >= 3 && u <= 5) ? v + w[4] : v;
}

// The 15 distinct ideal-MHD fluxes from the primitives w = (rho, ux, uy, uz,
// p*, Bx, By, Bz, E, u.B, E + p*) (phrike/equations.py:949-1000).  Of the 24
// fluxes F_d[c], the momentum fluxes rho u_i u_j + p* delta_ij - B_i B_j are
// symmetric in (i, j) and the induction fluxes u_d B_i - B_d u_i are
// antisymmetric (zero for i = d):
//   u = 0..2    rho ux, rho uy, rho uz                  (F_x,y,z[rho])
//   u = 3..5    (rho ud) ud + p* - Bd Bd, d = x, y, z   (F_d[rho ud])
//   u = 6..8    (rho ux) uy - Bx By, (rho ux) uz - Bx Bz, (rho uy) uz - By Bz
//   u = 9..11   (E + p*) ud - Bd (u.B)                  (F_d[E])
//   u = 12..14  ux By - uy Bx, ux Bz - uz Bx, uy Bz - uz By
//               (F_x[By] = -F_y[Bx], F_x[Bz] = -F_z[Bx], F_y[Bz] = -F_z[By])
// each in the reference's operation order for the first direction it names.
template <typename T>
__device__ __forceinline__ T mhd3_flux_sym(const T* w, int u) {
  if (u < 3) return w[0] * w[1 + u];
  if (u < 6) {
    const int d = u - 3;
    return (w[0] * w[1 + d] * w[1 + d] + w[4]) - w[5 + d] * w[5 + d];
  }
  if (u < 9) {
    const int a = u == 8 ? 1 : 0, b = u == 6 ? 1 : 2;  // (x, y), (x, z), (y, z)
    return (w[0] * w[1 + a]) * w[1 + b] - w[5 + a] * w[5 + b];
  }
  if (u < 12) return w[10] * w[1 + (u - 9)] - w[5 + (u - 9)] * w[9];
  const int a = u == 14 ? 1 : 0, b = u == 12 ? 1 : 2;
  return w[1 + a] * w[5 + b] - w[1 + b] * w[5 + a];
}

// The 7 distinct 2D Euler fluxes from the primitives w = (rho, ux, uy, 0, p,
// E + p) (phrike/equations.py:119-134): u = 0, 1: rho ux, rho uy; 2, 3:
// (rho ux) ux + p, (rho uy) uy + p; 4: (rho ux) uy (F_x[rho v] = F_y[rho u];
// the reference forms the latter as (rho uy) ux, one rounding apart at most);
// 5, 6: (E + p) ux, (E + p) uy; 7: none (zero)
template <typename T>
__device__ __forceinline__ T euler2_flux_sym(const T* w, int u) {
  if (u < 2) return w[0] * w[1 + u];
  if (u < 4) return (w[0] * w[u - 1]) * w[u - 1] + w[4];
  if (u == 4) return (w[0] * w[1]) * w[2];
  if (u < 7) return w[5] * w[u - 4];
  return T(0);
}

// Fluxes (or copies) of the conserved values q at the thread's x pairs ->
// forward transform of the row -> kept modes to a.out.  Shared by k_xfwd and
// k_xmid.
template <typename T, int NX, bool DA, int MODE>
__device__ __forceinline__ void xr_forward(const XFwdArgs<T>& a, int ncomp, const cx<T>* tw, cx<T>* buf,
                                           const T* sig, const T* kk, int tid, int slab, int y,
                                           const T (&q)[XRGeo<T, NX, DA>::PJ][2][XFwdShape<MODE>::NC]) {
  using G = XRGeo<T, NX, DA>;
  using S = XFwdShape<MODE>;
  constexpr int LX = S::LX, NC = S::NC, NT = G::NT, PJ = G::PJ, M = G::M;
  // primitive recovery (phrike/equations.py:104-117, 150-164; MHD 922-933)
  constexpr int NPR = MODE == XF_MHD3 ? 11 : 6;
  T pr[PJ][2][NPR];
  if constexpr (MODE == XF_MHD3) {
    // rho ux uy uz p* Bx By Bz E uB (E + p*)  (phrike/equations.py:922-933, 949-1000)
#pragma unroll
    for (int j = 0; j < PJ; ++j)
#pragma unroll
      for (int r = 0; r < 2; ++r) {
        const T* u = q[j][r];
        const T rho = u[0];
        const T ri = sizeof(T) == 8 ? T(0) : xrcp(rho);
        const T ux = xquot(u[1], rho, ri), uy = xquot(u[2], rho, ri), uz = xquot(u[3], rho, ri);
        const T Bx = u[5], By = u[6], Bz = u[7];
        const T kin = (T(0.5) * rho) * ((ux * ux + uy * uy) + uz * uz);
        const T mag = T(0.5) * ((Bx * Bx + By * By) + Bz * Bz);
        const T p = (a.gamma - T(1)) * ((u[4] - kin) - mag);
        const T ps = p + mag;
        T* w = pr[j][r];
        w[0] = rho; w[1] = ux; w[2] = uy; w[3] = uz; w[4] = ps;
        w[5] = Bx; w[6] = By; w[7] = Bz; w[8] = u[4];
        w[9] = (ux * Bx + uy * By) + uz * Bz;
        w[10] = u[4] + ps;
      }
  } else if constexpr (!xf_copy(MODE)) {
#pragma unroll
    for (int j = 0; j < PJ; ++j)
#pragma unroll
      for (int r = 0; r < 2; ++r) {
        const T* u = q[j][r];
        const T rho = u[0];
        const T ri = sizeof(T) == 8 ? T(0) : xrcp(rho);
        if constexpr (MODE == XF_FLUX3) {
          const T ux = xquot(u[1], rho, ri), uy = xquot(u[2], rho, ri), uz = xquot(u[3], rho, ri);
          const T kin = (T(0.5) * rho) * ((ux * ux + uy * uy) + uz * uz);
          const T p = (a.gamma - T(1)) * (u[4] - kin);
          pr[j][r][0] = rho; pr[j][r][1] = ux; pr[j][r][2] = uy; pr[j][r][3] = uz;
          pr[j][r][4] = p;
          pr[j][r][5] = u[4] + p;  // E + p
        } else {
          const T ux = xquot(u[1], rho, ri), uy = xquot(u[2], rho, ri);
          const T kin = (T(0.5) * rho) * (ux * ux + uy * uy);
          const T p = (a.gamma - T(1)) * (u[3] - kin);
          pr[j][r][0] = rho; pr[j][r][1] = ux; pr[j][r][2] = uy; pr[j][r][3] = T(0);
          pr[j][r][4] = p;
          pr[j][r][5] = u[3] + p;
        }
      }
  }
  cx<T>* dst = a.out + (long long)slab * a.out_ss + (long long)y * a.mxp;
  if constexpr (XMerge<T, NX, MODE>::value) {
    static_assert(!(MODE == XF_FLUX2 && PH_2D_SYM), "the merged transform takes the general 2D layout");
    // all NG directions' fluxes as one batch of lines (group g = lines g*LX..)
    constexpr int LG = XMerge<T, NX, MODE>::lines;
#pragma unroll
    for (int g = 0; g < S::NG; ++g)
#pragma unroll
      for (int j = 0; j < PJ; ++j) {
        const int p = G::point(tid, j);
        if (G::PALL || p < M) {
#pragma unroll
          for (int l = 0; l < LX; ++l)
            buf[xslot<LG, T>(p, g * LX + l)] =
                MODE == XF_FLUX3 ? mk(euler3_flux_sym<T>(pr[j][0], g * LX + l), euler3_flux_sym<T>(pr[j][1], g * LX + l))
                                 : mk(euler_flux<T, NC>(pr[j][0], g, l), euler_flux<T, NC>(pr[j][1], g, l));
        }
      }
    lds_barrier();
    fft_lds<T, M, LG, -1, NT, xstride(LG, sizeof(cx<T>)), xpadb(LG, sizeof(cx<T>)), XPlan<M>, 2, PH_XTW_SWZ>(buf, tw, tid);
    xr_split_store<T, NX, DA, LG, LX, MODE != XF_FLUX3>(buf, tw, dst, a.out_fs, LG, sig, kk, a.ikmask, tid);
    lds_barrier();
    return;
  }
  const int nl0 = xf_copy(MODE) ? ncomp : LX;
  // the 3D Euler groups unrolled: each line's flux is then fixed at compile
  // time (a run-time group index selects it through chains of selects)
  constexpr bool SYM2 = MODE == XF_FLUX2 && PH_2D_SYM;
  constexpr bool SYM = MODE == XF_FLUX3 || (MODE == XF_MHD3 && PH_MHD_SYM) || SYM2;
  constexpr int GU = (SYM && PH_XF_UNROLL_G) ? S::NG : 1;
  // the symmetric layouts' fluxes are stored without i k (the y pass applies
  // i kx)
  constexpr bool IK = !SYM;
#pragma unroll GU
  for (int g = 0; g < S::NG; ++g) {
#pragma unroll
    for (int j = 0; j < PJ; ++j) {
      const int p = G::point(tid, j);
      if (G::PALL || p < M) {
#pragma unroll
        for (int l = 0; l < LX; ++l) {
          T v[2];
#pragma unroll
          for (int r = 0; r < 2; ++r) {
            if constexpr (xf_copy(MODE)) {
              v[r] = q[j][r][l];
            } else if constexpr (MODE == XF_MHD3 && PH_MHD_SYM) {
              v[r] = mhd3_flux_sym<T>(pr[j][r], g * LX + l);
            } else if constexpr (MODE == XF_MHD3) {
              // ideal-MHD flux of direction g, component l (phrike/equations.py:949-1000);
              // the direction is picked by selects (a register index by the
              // run-time g would put pr in scratch memory)
              const T* w = pr[j][r];
              const T ug = g == 0 ? w[1] : (g == 1 ? w[2] : w[3]);
              const T Bg = g == 0 ? w[5] : (g == 1 ? w[6] : w[7]);
              const T mg = w[0] * ug;
              if (l == 0) v[r] = mg;
              else if (l <= 3) {
                v[r] = mg * w[l];
                if (l == g + 1) v[r] = v[r] + w[4];
                v[r] = v[r] - Bg * w[4 + l];
              } else if (l == 4) v[r] = w[10] * ug - Bg * w[9];
              else v[r] = (l - 5 == g) ? T(0) : ug * w[l] - w[l - 4] * Bg;
            } else if constexpr (MODE == XF_FLUX3) {
              v[r] = euler3_flux_sym<T>(pr[j][r], g * LX + l);
            } else {
              v[r] = SYM2 ? euler2_flux_sym<T>(pr[j][r], g * LX + l) : euler_flux<T, NC>(pr[j][r], g, l);
            }
          }
          buf[xslot<LX, T>(p, l)] = mk(v[0], v[1]);
        }
      }
    }
    lds_barrier();
    fft_lds<T, M, LX, -1, NT, xstride(LX, sizeof(cx<T>)), xpadb(LX, sizeof(cx<T>)), XPlan<M>, 2, PH_XTW_SWZ>(buf, tw, tid);
    // the 2D symmetric layout's second group holds 3 fluxes
    const int nl = (SYM2 && g == S::NG - 1) ? 3 : nl0;
    xr_split_store<T, NX, DA, LX, LX, IK>(buf, tw, dst + (long long)(g * LX) * a.out_fs, a.out_fs, nl, sig, kk,
                                          (a.ikmask >> (g * LX)) & 1u, tid);
    lds_barrier();
  }
}

template <typename T, int NX, bool DA, int MODE>
__global__ __launch_bounds__((XRThreads<T, NX>::value), PH_X_WAVES_PER_EU) void k_xfwd(XFwdArgs<T> a) {
  using G = XRGeo<T, NX, DA>;
  using S = XFwdShape<MODE>;
  constexpr int NT = G::NT, NC = S::NC, PJ = G::PJ;
  PH_XTW_DECL(T, NX);
  __shared__ cx<T> buf[xrbuf_size<NX, XMerge<T, NX, MODE>::lines, T>()];
  __shared__ T s_sig[G::MX], s_k[G::MX];
  const int tid = threadIdx.x;
  const int slab = blockIdx.y;
  if (gated_off(a.active, a.act_all, slab)) return;
  // PH_X_ROWS consecutive rows per block; the twiddle and mode tables are
  // filled once
#pragma unroll 1
  for (int rr = 0; rr < PH_X_ROWS; ++rr) {
    const int y = blockIdx.x * PH_X_ROWS + rr;
    if (rr > 0) lds_barrier();  // the previous row is done with buf
    if (rr == 0) PH_XTW_FILL(a.tw, NX, NT, tid);
    if (rr == 0) xr_tables<T, NX, DA>(a.sig, a.kk, s_sig, s_k, tid);
    const T* src = a.in + (long long)slab * a.in_ss + (long long)y * NX;
    const int ncomp = xf_copy(MODE) ? a.nf : NC;
    T q[PJ][2][NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      cx<T> v[PJ];
#pragma unroll
      for (int j = 0; j < PJ; ++j) v[j] = mk(T(0), T(0));
      if (c < ncomp) xr_load_pairs<G>(src + (long long)c * a.in_fs, tid, v);
#pragma unroll
      for (int j = 0; j < PJ; ++j) {
        q[j][0][c] = v[j].x;
        q[j][1][c] = v[j].y;
      }
    }
    xr_forward<T, NX, DA, MODE>(a, ncomp, tw, buf, s_sig, s_k, tid, slab, y, q);
  }
}

// RK4 stage combine of the x pair (2p, 2p+1) with k_s = -v (phrike/solver.py:
// 1220-1230 3D, 502-513 2D); as rk_point: stages 1-3 give the next stage
// state in u and the updated accumulator in acc_new, stage 4 gives
// acc + dt/6 k4 in u and acc_new; the 2D form clamps u (phrike/solver.py:
// 467-488), the 8-field ideal-MHD form through its own prim -> cons
// (phrike/solver.py:956-964).  acc_new at stage 4 is the unclamped sum.
template <typename T, int NC, bool CLAMP>
__device__ __forceinline__ void rk_pair(const XInvArgs<T>& a, const T (&v)[2][NC], double dtd, const cx<T> (&un)[NC],
                                        const cx<T> (&acc)[NC], T (&u)[2][NC], cx<T> (&acc_new)[NC]) {
  // one assignment per output (selects, no branches), which keeps the
  // arrays in registers
  const T half = T(0.5 * dtd), full = T(dtd), sixth = T(dtd / 6.0), third = T(dtd / 3.0);
  const bool last = a.stage >= 4;
  const T coef = last ? sixth : (a.stage == 3 ? full : half);
  const T w = (a.stage == 1) ? sixth : third;
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const T k0 = -v[0][c], k1 = -v[1][c];
    const cx<T> bu = last ? acc[c] : un[c];
    u[0][c] = bu.x + coef * k0;
    u[1][c] = bu.y + coef * k1;
    const cx<T> ba = a.stage == 1 ? un[c] : acc[c];
    acc_new[c] = last ? mk(u[0][c], u[1][c]) : mk(ba.x + w * k0, ba.y + w * k1);
  }
  if constexpr (CLAMP) {
    if constexpr (NC == 8) {
      clamp_mhd(u[0], a.gamma, a.rho_min, a.p_min);
      clamp_mhd(u[1], a.gamma, a.rho_min, a.p_min);
    } else {
      clamp2d(u[0], a.gamma, a.rho_min, a.p_min);
      clamp2d(u[1], a.gamma, a.rho_min, a.p_min);
    }
  }
}

// the row's inverse transform of NC fields into buf (natural packed order)
template <typename T, int NX, bool DA, int NC>
__device__ __forceinline__ void xr_inverse(const XRSpec<T, NX, DA, NC>& sp, cx<T>* buf, const cx<T>* tw, int tid) {
  using G = XRGeo<T, NX, DA>;
  sp.place(buf, tid);
  lds_barrier();  // also covers the block's twiddle fill
  xr_pack<T, NX, NC, G::NT>(buf, tw, tid);
  lds_barrier();
  fft_lds<T, G::M, NC, +1, G::NT, xstride(NC, sizeof(cx<T>)), xpadb(NC, sizeof(cx<T>)), XPlan<G::M>, 2, PH_XTW_SWZ>(buf, tw, tid);
}

template <typename T, int NX, bool DA, int NC, int MODE>
__global__ __launch_bounds__((XRThreads<T, NX>::value), PH_X_WAVES_PER_EU) void k_xinv(XInvArgs<T> a) {
  using G = XRGeo<T, NX, DA>;
  constexpr int NT = G::NT, PJ = G::PJ;
  PH_XTW_DECL(T, NX);
  __shared__ cx<T> buf[xrbuf_size<NX, NC, T>()];
  const int tid = threadIdx.x;
  const int slab = blockIdx.y;
  if (gated_off(a.active, a.act_all, slab)) return;
  // PH_X_ROWS consecutive rows per block; the twiddle and mode tables are
  // filled once
#pragma unroll 1
  for (int rr = 0; rr < PH_X_ROWS; ++rr) {
    const int y = blockIdx.x * PH_X_ROWS + rr;
    if (rr > 0) lds_barrier();  // the previous row is done with buf
    if (rr == 0) PH_XTW_FILL(a.tw, NX, NT, tid);
    const int nfl = (MODE == XI_OUT || MODE == XI_RHS || MODE == XI_ADD) ? a.nf : NC;
    static_assert(MODE != XI_PROJE || NC == 3, "projection write-back takes the 3 B fields");
    static_assert(MODE != XI_MHDSTAGE || NC == 8, "the MHD stage combine takes the 8 MHD fields");
    XRSpec<T, NX, DA, NC> sp;
    sp.load(a.in + (long long)slab * a.in_ss + (long long)y * a.mxp, a.in_fs, nfl, tid);
    xr_inverse<T, NX, DA, NC>(sp, buf, tw, tid);
    const long long base = (long long)slab * a.ss + (long long)y * NX;
    const double dtd = a.dt_dev != nullptr ? a.dt_dev[a.dt_per_slab ? slab : 0] : a.dt;
    // the transformed values at the thread's x pairs: v[j][r][c] at x = 2p + r
    T v[PJ][2][NC];
#pragma unroll
    for (int j = 0; j < PJ; ++j) {
      const int p = G::point(tid, j);
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        const cx<T> z = (G::PALL || p < G::M) ? buf[xslot<NC, T>(p, c)] : mk(T(0), T(0));
        v[j][0][c] = z.x;
        v[j][1][c] = z.y;
      }
    }
    auto row = [&](T* f, int c) { return f + c * a.fs + base; };
    // the fused CFL reduction of a final state: max wave speed of the values
    // this thread wrote (u[j][r][c] at x = 2p + r), one partial per row
    constexpr bool EULER = NC == 4 || NC == 5;
    auto cfl_store = [&](const T (&u)[PJ][2][NC]) {
      if constexpr (EULER) {
        double m = 0.0;
#pragma unroll
        for (int j = 0; j < PJ; ++j) {
          if (!G::PALL && G::point(tid, j) >= G::M) continue;
#pragma unroll
          for (int r = 0; r < 2; ++r) m = dmax_nan(m, wave_speed_v<T, NC - 2>(u[j][r], a.gamma));
        }
        block_max_store<NT>(m, a.cfl_rows + (long long)slab * (gridDim.x * PH_X_ROWS) + y, tid);
      }
    };
    if constexpr (MODE == XI_OUT || MODE == XI_RHS || MODE == XI_ADD) {
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        if (c >= nfl) continue;
        cx<T> o[PJ];
        if constexpr (MODE == XI_ADD) xr_load_pairs<G>(row(a.out, c), tid, o);
#pragma unroll
        for (int j = 0; j < PJ; ++j) {
          const T s0 = MODE == XI_RHS ? -v[j][0][c] : v[j][0][c];
          const T s1 = MODE == XI_RHS ? -v[j][1][c] : v[j][1][c];
          o[j] = MODE == XI_ADD ? mk(o[j].x + s0, o[j].y + s1) : mk(s0, s1);
          v[j][0][c] = o[j].x;  // the values written
          v[j][1][c] = o[j].y;
        }
        xr_store_pairs<G>(row(a.out, c), tid, o);
      }
      if (MODE != XI_RHS && a.cfl_rows != nullptr) cfl_store(v);
    } else if constexpr (MODE == XI_PROJE) {
      // phrike/solver.py:966-986 with the primitives of phrike/equations.py:922-933:
      // B <- the projected B, E at the thermal pressure of the unprojected state
      cx<T> w[8][PJ];
#pragma unroll
      for (int f = 0; f < 8; ++f) xr_load_pairs<G>(row(a.out, f), tid, w[f]);
#pragma unroll
      for (int j = 0; j < PJ; ++j) {
        T e[2], b[2][3];
#pragma unroll
        for (int r = 0; r < 2; ++r) {
          auto at = [&](int f) { return r == 0 ? w[f][j].x : w[f][j].y; };
          const T rho = at(0);
          const T ux = at(1) / rho, uy = at(2) / rho, uz = at(3) / rho;
          const T b0 = at(5), b1 = at(6), b2 = at(7);
          const T kin = (T(0.5) * rho) * ((ux * ux + uy * uy) + uz * uz);
          const T mag = T(0.5) * ((b0 * b0 + b1 * b1) + b2 * b2);
          const T pp = (a.gamma - T(1)) * ((at(4) - kin) - mag);
          const T Bx = v[j][r][0], By = v[j][r][1], Bz = v[j][r][2];
          e[r] = (pp / (a.gamma - T(1)) + kin) + T(0.5) * ((Bx * Bx + By * By) + Bz * Bz);
          b[r][0] = Bx;
          b[r][1] = By;
          b[r][2] = Bz;
        }
        w[4][j] = mk(e[0], e[1]);
#pragma unroll
        for (int f = 0; f < 3; ++f) w[5 + f][j] = mk(b[0][f], b[1][f]);
      }
#pragma unroll
      for (int f = 4; f < 8; ++f) xr_store_pairs<G>(row(a.out, f), tid, w[f]);
    } else if constexpr (MODE == XI_MHDSTAGE) {
      cx<T> un[8][PJ], acc[8][PJ];
#pragma unroll
      for (int f = 0; f < 8; ++f) {
        xr_load_pairs<G>(row(const_cast<T*>(a.un), f), tid, un[f]);
#pragma unroll
        for (int j = 0; j < PJ; ++j) acc[f][j] = mk(T(0), T(0));
        if (a.stage > 1) xr_load_pairs<G>(row(a.out, f), tid, acc[f]);
      }
#pragma unroll
      for (int j = 0; j < PJ; ++j) {
        T u[2][8], an[2][8];
#pragma unroll
        for (int r = 0; r < 2; ++r) {
          T unr[8], accr[8];
#pragma unroll
          for (int f = 0; f < 8; ++f) {
            unr[f] = r == 0 ? un[f][j].x : un[f][j].y;
            accr[f] = r == 0 ? acc[f][j].x : acc[f][j].y;
          }
          mhd_stage_vals<T>(a, v[j][r], unr, accr, dtd, u[r], an[r]);
        }
#pragma unroll
        for (int f = 0; f < 8; ++f) {
          un[f][j] = mk(u[0][f], u[1][f]);  // the next state
          acc[f][j] = mk(an[0][f], an[1][f]);
        }
      }
#pragma unroll
      for (int f = 0; f < 8; ++f) {
        if (a.stage < 4) xr_store_pairs<G>(row(a.out, f), tid, acc[f]);
        xr_store_pairs<G>(row(a.us, f), tid, un[f]);
      }
    } else {
      cx<T> un[NC][PJ], acc[NC][PJ];
#pragma unroll
      for (int c = 0; c < NC; ++c) {
#pragma unroll
        for (int j = 0; j < PJ; ++j) {
          un[c][j] = mk(T(0), T(0));
          acc[c][j] = mk(T(0), T(0));
        }
        if (a.stage < 4) xr_load_pairs<G>(row(const_cast<T*>(a.un), c), tid, un[c]);
        if (a.stage > 1) xr_load_pairs<G>(row(a.out, c), tid, acc[c]);
      }
      cx<T> an[NC][PJ], us[NC][PJ];
#pragma unroll
      for (int j = 0; j < PJ; ++j) {
        cx<T> un_j[NC], acc_j[NC], an_j[NC];
        T u[2][NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          un_j[c] = un[c][j];
          acc_j[c] = acc[c][j];
        }
        rk_pair<T, NC, MODE == XI_STAGE2>(a, v[j], dtd, un_j, acc_j, u, an_j);
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          an[c][j] = an_j[c];
          us[c][j] = mk(u[0][c], u[1][c]);
        }
      }
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        if (a.stage < 4) {
          xr_store_pairs<G>(row(a.out, c), tid, an[c]);
          xr_store_pairs<G>(row(a.us, c), tid, us[c]);
        } else {
          xr_store_pairs<G>(row(a.out, c), tid, us[c]);
        }
      }
      if (a.stage == 4 && a.cfl_rows != nullptr) {
        T u[PJ][2][NC];
#pragma unroll
        for (int j = 0; j < PJ; ++j)
#pragma unroll
          for (int c = 0; c < NC; ++c) {
            u[j][0][c] = us[c][j].x;
            u[j][1][c] = us[c][j].y;
          }
        cfl_store(u);
      }
    }
  }
}

// Fused stage boundary: x inverse + RK combine
// of stage s, then the fluxes of the stage s+1 state and their forward x
// transform (stages 1-3), or the new state's forward x transform for the
// post-step filter (stage 4).  The block loads its spectral row completely
// before it writes that row.
template <typename T, int NX, bool DA, int NC, int FMODE>
__global__ __launch_bounds__((XRThreads<T, NX>::value), PH_X_WAVES_PER_EU) void k_xmid(XMidArgs<T> m) {
  using G = XRGeo<T, NX, DA>;
  constexpr int NT = G::NT, PJ = G::PJ;
  static_assert(XFwdShape<FMODE>::NC == NC || FMODE == XF_COPY, "field counts");
  // stage 4 transforms the new state itself: 5 fields (Euler; 4 in 2D) or 8 (MHD)
  constexpr int LMODE = NC == 8 ? XF_COPY8 : XF_COPY;
  constexpr int QN = XFwdShape<LMODE>::NC;
  PH_XTW_DECL(T, NX);
  constexpr int LB0 = NC > XMerge<T, NX, FMODE>::lines ? NC : XMerge<T, NX, FMODE>::lines;
  constexpr int LB = LB0 > XFwdShape<LMODE>::LX ? LB0 : XFwdShape<LMODE>::LX;
  __shared__ cx<T> buf[xrbuf_size<NX, LB, T>()];
  __shared__ T s_sig[G::MX], s_k[G::MX];
  const XInvArgs<T>& a = m.inv;
  const int tid = threadIdx.x;
  const int slab = blockIdx.y;
  if (gated_off(a.active, a.act_all, slab)) return;
  // PH_X_ROWS consecutive rows per block; the twiddle and mode tables are
  // filled once
#pragma unroll 1
  for (int rr = 0; rr < PH_X_ROWS; ++rr) {
    const int y = blockIdx.x * PH_X_ROWS + rr;
    if (rr > 0) lds_barrier();  // the previous row is done with buf
    XRSpec<T, NX, DA, NC> sp;
    sp.load(a.in + (long long)slab * a.in_ss + (long long)y * a.mxp, a.in_fs, NC, tid);
    if (rr == 0) PH_XTW_FILL(a.tw, NX, NT, tid);
    const XFwdArgs<T>& fa = a.stage < 4 ? m.fwd : m.fwd_last;
    if (rr == 0) xr_tables<T, NX, DA>(fa.sig, fa.kk, s_sig, s_k, tid);
    const long long base = (long long)slab * a.ss + (long long)y * NX;
    // the RK combine's inputs (x pairs), per field: U (PH_XMID_PREFETCH 2), or
    // U and the accumulator (1), loaded before the inverse transform so their
    // latency overlaps it; the rest after it
    constexpr int PF = NC == 8 ? PH_XMID_PREFETCH_MHD : PH_XMID_PREFETCH;
    constexpr bool PRE_UN = PF != 0, PRE_ACC = PF == 1;
    auto row = [&](T* f, int c) { return f + c * a.fs + base; };
    cx<T> un[NC][PJ], acc[NC][PJ];
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
      for (int j = 0; j < PJ; ++j) {
        un[c][j] = mk(T(0), T(0));
        acc[c][j] = mk(T(0), T(0));
      }
    if constexpr (PRE_UN) {
      if (a.stage < 4) {
#pragma unroll
        for (int c = 0; c < NC; ++c) xr_load_pairs<G>(row(const_cast<T*>(a.un), c), tid, un[c]);
      }
    }
    if constexpr (PRE_ACC) {
      if (a.stage > 1 && !a.specacc) {
#pragma unroll
        for (int c = 0; c < NC; ++c) xr_load_pairs<G>(row(a.out, c), tid, acc[c]);
      }
    }
    xr_inverse<T, NX, DA, NC>(sp, buf, tw, tid);
    const double dtd = a.dt_dev != nullptr ? a.dt_dev[a.dt_per_slab ? slab : 0] : a.dt;
    if constexpr (!PRE_UN) {
      if (a.stage < 4) {
#pragma unroll
        for (int c = 0; c < NC; ++c) xr_load_pairs<G>(row(const_cast<T*>(a.un), c), tid, un[c]);
      }
    }
    if constexpr (!PRE_ACC) {
      if (a.stage > 1 && !a.specacc) {
#pragma unroll
        for (int c = 0; c < NC; ++c) xr_load_pairs<G>(row(a.out, c), tid, acc[c]);
      }
    }
    T q[PJ][2][QN];
    cx<T> an[NC][PJ];
#pragma unroll
    for (int j = 0; j < PJ; ++j) {
      const int p = G::point(tid, j);
      T v[2][NC];
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        const cx<T> z = (G::PALL || p < G::M) ? buf[xslot<NC, T>(p, c)] : mk(T(0), T(0));
        v[0][c] = z.x;
        v[1][c] = z.y;
      }
      T u[2][NC];
      cx<T> un_j[NC], acc_j[NC], an_j[NC];
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        un_j[c] = un[c][j];
        acc_j[c] = acc[c][j];
      }
      rk_pair<T, NC, NC == 4 || NC == 8>(a, v, dtd, un_j, acc_j, u, an_j);
#pragma unroll
      for (int c = 0; c < NC; ++c) an[c][j] = an_j[c];
#pragma unroll
      for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int c = 0; c < QN; ++c) q[j][r][c] = c < NC ? u[r][c] : T(0);
    }
    // keep_state (stage 4 also writes the new state) is an option of the
    // unclamped 3D Euler form only, where the stage-4 accumulator is the state
    if (!a.specacc && (a.stage < 4 || (NC == 5 && a.keep_state))) {
#pragma unroll
      for (int c = 0; c < NC; ++c) xr_store_pairs<G>(row(a.out, c), tid, an[c]);
    }
    lds_barrier();  // buf is reused by the forward transform
    if (a.stage < 4) {
      T qf[PJ][2][XFwdShape<FMODE>::NC];
#pragma unroll
      for (int j = 0; j < PJ; ++j)
#pragma unroll
        for (int r = 0; r < 2; ++r)
#pragma unroll
          for (int c = 0; c < XFwdShape<FMODE>::NC; ++c) qf[j][r][c] = q[j][r][c];
      xr_forward<T, NX, DA, FMODE>(m.fwd, NC, tw, buf, s_sig, s_k, tid, slab, y, qf);
    } else {
      xr_forward<T, NX, DA, LMODE>(m.fwd_last, NC, tw, buf, s_sig, s_k, tid, slab, y, q);
    }
  }
}

// Fused step boundary between two steps: the post-step filter's x inverse
// (XI_OUT: the filtered state; XI_ADD: the state plus the FP32 filter
// correction) is written as the new state, with its per-row CFL maxima, and
// then the fluxes of that state and their x transform (k_xfwd's work for the
// next step's first stage) go to the spectral buffer, so the next step does
// not read the state again.  The block loads its spectral row completely
// before it writes that row.
template <typename T, int NX, bool DA, int NC, int MODE>
__global__ __launch_bounds__((XRThreads<T, NX>::value), PH_X_WAVES_PER_EU) void k_xend(XMidArgs<T> m) {
  using G = XRGeo<T, NX, DA>;
  constexpr int NT = G::NT, PJ = G::PJ;
  constexpr int FMODE = NC == 5 ? XF_FLUX3 : XF_FLUX2;  // 3D or batched-2D Euler state
  static_assert(NC == 4 || NC == 5, "Euler states");
  static_assert(MODE == XI_OUT || MODE == XI_ADD, "the filter's last pass writes or adds");
  constexpr int LB = NC > XMerge<T, NX, FMODE>::lines ? NC : XMerge<T, NX, FMODE>::lines;
  PH_XTW_DECL(T, NX);
  __shared__ cx<T> buf[xrbuf_size<NX, LB, T>()];
  __shared__ T s_sig[G::MX], s_k[G::MX];
  const XInvArgs<T>& a = m.inv;
  const int tid = threadIdx.x;
  const int slab = blockIdx.y;
  if (gated_off(a.active, a.act_all, slab)) return;
#pragma unroll 1
  for (int rr = 0; rr < PH_X_ROWS; ++rr) {
    const int y = blockIdx.x * PH_X_ROWS + rr;
    if (rr > 0) lds_barrier();  // the previous row is done with buf
    XRSpec<T, NX, DA, NC> sp;
    sp.load(a.in + (long long)slab * a.in_ss + (long long)y * a.mxp, a.in_fs, NC, tid);
    if (rr == 0) PH_XTW_FILL(a.tw, NX, NT, tid);
    if (rr == 0) xr_tables<T, NX, DA>(m.fwd.sig, m.fwd.kk, s_sig, s_k, tid);
    const long long base = (long long)slab * a.ss + (long long)y * NX;
    auto row = [&](T* f, int c) { return f + c * a.fs + base; };
    cx<T> o[NC][PJ];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
#pragma unroll
      for (int j = 0; j < PJ; ++j) o[c][j] = mk(T(0), T(0));
      if constexpr (MODE == XI_ADD) xr_load_pairs<G>(row(a.out, c), tid, o[c]);
    }
    xr_inverse<T, NX, DA, NC>(sp, buf, tw, tid);
    T q[PJ][2][NC];
#pragma unroll
    for (int j = 0; j < PJ; ++j) {
      const int p = G::point(tid, j);
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        const cx<T> z = (G::PALL || p < G::M) ? buf[xslot<NC, T>(p, c)] : mk(T(0), T(0));
        o[c][j] = MODE == XI_ADD ? mk(o[c][j].x + z.x, o[c][j].y + z.y) : z;
        q[j][0][c] = o[c][j].x;
        q[j][1][c] = o[c][j].y;
      }
    }
#pragma unroll
    for (int c = 0; c < NC; ++c) xr_store_pairs<G>(row(a.out, c), tid, o[c]);
    if (a.cfl_rows != nullptr) {
      double mm = 0.0;
#pragma unroll
      for (int j = 0; j < PJ; ++j) {
        if (!G::PALL && G::point(tid, j) >= G::M) continue;
#pragma unroll
        for (int r = 0; r < 2; ++r) mm = dmax_nan(mm, wave_speed_v<T, NC - 2>(q[j][r], a.gamma));
      }
      block_max_store<NT>(mm, a.cfl_rows + (long long)slab * (gridDim.x * PH_X_ROWS) + y, tid);
    }
    lds_barrier();  // buf is reused by the forward transform
    xr_forward<T, NX, DA, FMODE>(m.fwd, NC, tw, buf, s_sig, s_k, tid, slab, y, q);
  }
}

// launch shape of the x kernels: threads per block and rows per block
template <typename T, int NX>
struct XLaunch {
  static constexpr int threads = XRThreads<T, NX>::value, rows = PH_X_ROWS;
};

// ----------------------------------------------------------------------------
// Half spectra of a plan that keeps every mode, between its spectral buffer B
// ([field][kz][ky][mxp], FP32 rows pair-planar) and a plain complex array
// (field, kz, ky, kx < mx) interleaved re/im; the copy in scales by `scale`
// and zeroes B's padding modes mx..mxp-1 (the column passes read them).
// ----------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ long long spec_index(long long row, int m, int mxp, bool im) {
  // position (in T elements) of mode m of a spectral row in B
  if constexpr (sizeof(T) == 8) return (row * mxp + m) * 2 + (im ? 1 : 0);
  else return row * mxp * 2 + 4 * (m >> 1) + (im ? 2 : 0) + (m & 1);
}

template <typename T>
__global__ __launch_bounds__(NT) void k_spec_out(const T* B, T* out, long long rows, int mx, int mxp) {
  const long long i = (long long)blockIdx.x * NT + threadIdx.x;
  if (i >= rows * mx) return;
  const long long row = i / mx;
  const int m = int(i - row * mx);
  out[2 * i] = B[spec_index<T>(row, m, mxp, false)];
  out[2 * i + 1] = B[spec_index<T>(row, m, mxp, true)];
}

template <typename T>
__global__ __launch_bounds__(NT) void k_spec_in(const T* in, T* B, long long rows, int mx, int mxp, T scale) {
  const long long i = (long long)blockIdx.x * NT + threadIdx.x;
  if (i >= rows * mxp) return;
  const long long row = i / mxp;
  const int m = int(i - row * mxp);
  const long long o = row * mx + m;
  B[spec_index<T>(row, m, mxp, false)] = m < mx ? in[2 * o] * scale : T(0);
  B[spec_index<T>(row, m, mxp, true)] = m < mx ? in[2 * o + 1] * scale : T(0);
}

// ----------------------------------------------------------------------------
// reductions: per-block partials, then one block reduces them in a fixed
// order (deterministic; no atomics, no memset).
// ----------------------------------------------------------------------------
template <typename T, int DIM>
__global__ __launch_bounds__(NT) void k_maxspeed_partial(const T* u, long long fs, long long ss,
                                                         long long n, T gamma, double* part,
                                                         const int* active) {
  // grid: (blocks per slab, slabs); part[slab * gridDim.x + blockIdx.x]
  __shared__ double red[NT];
  const T* us = u + (long long)blockIdx.y * ss;
  double m = 0.0;
  const bool on = active == nullptr || active[blockIdx.y] != 0;
  for (long long i = (long long)blockIdx.x * NT + threadIdx.x; on && i < n; i += (long long)gridDim.x * NT)
    m = dmax_nan(m, wave_speed<T, DIM>(us, fs, i, gamma));
  red[threadIdx.x] = m;
  lds_barrier();
  for (int s = NT / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] = dmax_nan(red[threadIdx.x], red[threadIdx.x + s]);
    lds_barrier();
  }
  if (threadIdx.x == 0) part[(long long)blockIdx.y * gridDim.x + blockIdx.x] = red[0];
}

// first level of a large partial reduction: n values (per slab, slab stride
// n) -> gridDim.x maxima per slab, part[slab * gridDim.x + blockIdx.x]
static __global__ __launch_bounds__(NT) void k_max_mid(const double* in, long long n, double* part) {
  __shared__ double red[NT];
  const double* ps = in + (long long)blockIdx.y * n;
  double m = 0.0;
  for (long long i = (long long)blockIdx.x * NT + threadIdx.x; i < n; i += (long long)gridDim.x * NT)
    m = dmax_nan(m, ps[i]);
  red[threadIdx.x] = m;
  lds_barrier();
  for (int s = NT / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] = dmax_nan(red[threadIdx.x], red[threadIdx.x + s]);
    lds_barrier();
  }
  if (threadIdx.x == 0) part[(long long)blockIdx.y * gridDim.x + blockIdx.x] = red[0];
}

// one block per slab: reduce nb partials -> out[slab]; optionally dt = cfl*h/max
template <int DIM>
__global__ __launch_bounds__(NT) void k_max_final(const double* part, int nb, double* out,
                                                  double* dt_out, double cfl_h, const int* active) {
  // an inactive run keeps its previous max and dt
  if (active != nullptr && active[blockIdx.x] == 0) return;
  __shared__ double red[NT];
  const double* ps = part + (long long)blockIdx.x * nb;
  double m = 0.0;
  for (int i = threadIdx.x; i < nb; i += NT) m = dmax_nan(m, ps[i]);
  red[threadIdx.x] = m;
  lds_barrier();
  for (int s = NT / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] = dmax_nan(red[threadIdx.x], red[threadIdx.x + s]);
    lds_barrier();
  }
  if (threadIdx.x == 0) {
    const double m = red[0];
    if (out) out[blockIdx.x] = m;
    // the reference's compute_dt: 1e-6 when the speed is <= 0, else cfl*h/max,
    // which is NaN for a NaN speed (a diverged run must not keep stepping)
    if (dt_out) dt_out[blockIdx.x] = m != m ? m : (m > 0.0 ? cfl_h / m : 1e-6);
  }
}

// conserved sums: 3D mass, mx, my, mz, energy from primitives
// (phrike/solver.py:1439-1468); 2D mass, mx, my, energy (solver.py:734-757)
template <typename T, int DIM>
__global__ __launch_bounds__(NT) void k_sums_partial(const T* u, long long fs, long long ss,
                                                     long long n, T gamma, double* part) {
  constexpr int NS = DIM == 3 ? 5 : 4;
  __shared__ double red[NS][NT];
  const T* us = u + (long long)blockIdx.y * ss;
  double acc[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) acc[s] = 0.0;
  for (long long i = (long long)blockIdx.x * NT + threadIdx.x; i < n; i += (long long)gridDim.x * NT) {
    const T rho = us[i];
    if constexpr (DIM == 3) {
      const T ux = us[fs + i] / rho, uy = us[2 * fs + i] / rho, uz = us[3 * fs + i] / rho;
      const T q2 = (ux * ux + uy * uy) + uz * uz;
      const T p = (gamma - T(1)) * (us[4 * fs + i] - (T(0.5) * rho) * q2);
      acc[0] += double(rho);
      acc[1] += double(rho * ux);
      acc[2] += double(rho * uy);
      acc[3] += double(rho * uz);
      acc[4] += double(p / (gamma - T(1)) + (T(0.5) * rho) * q2);
    } else {
      const T ux = us[fs + i] / rho, uy = us[2 * fs + i] / rho;
      const T q2 = ux * ux + uy * uy;
      const T p = (gamma - T(1)) * (us[3 * fs + i] - (T(0.5) * rho) * q2);
      acc[0] += double(rho);
      acc[1] += double(rho * ux);
      acc[2] += double(rho * uy);
      acc[3] += double(p / (gamma - T(1)) + (T(0.5) * rho) * q2);
    }
  }
#pragma unroll
  for (int s = 0; s < NS; ++s) red[s][threadIdx.x] = acc[s];
  lds_barrier();
  for (int w = NT / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
#pragma unroll
      for (int s = 0; s < NS; ++s) red[s][threadIdx.x] += red[s][threadIdx.x + w];
    }
    lds_barrier();
  }
  if (threadIdx.x < NS)
    part[((long long)blockIdx.y * gridDim.x + blockIdx.x) * NS + threadIdx.x] = red[threadIdx.x][0];
}

template <int NS>
__global__ __launch_bounds__(NT) void k_sums_final(const double* part, int nb, double* out) {
  __shared__ double red[NS][NT];
  const double* ps = part + (long long)blockIdx.x * nb * NS;
  double acc[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) acc[s] = 0.0;
  for (int i = threadIdx.x; i < nb; i += NT) {
#pragma unroll
    for (int s = 0; s < NS; ++s) acc[s] += ps[(long long)i * NS + s];
  }
#pragma unroll
  for (int s = 0; s < NS; ++s) red[s][threadIdx.x] = acc[s];
  lds_barrier();
  for (int w = NT / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
#pragma unroll
      for (int s = 0; s < NS; ++s) red[s][threadIdx.x] += red[s][threadIdx.x + w];
    }
    lds_barrier();
  }
  if (threadIdx.x < NS) out[(long long)blockIdx.x * NS + threadIdx.x] = red[threadIdx.x][0];
}

// ----------------------------------------------------------------------------
// monitoring reductions (phrike/problems/base.py:490-822: velocity, density
// and pressure statistics): per run, in FP64,
//   MON_MASS   sum rho              MON_RV2   sum rho |u|^2
//   MON_VMAX   max |u|              MON_VMIN  min |u|
//   MON_RMAX   max rho              MON_RMIN  min rho
//   MON_RS1    sum (rho - c)        MON_RS2   sum (rho - c)^2   (c = rho of cell 0,
//              a shift that keeps the variance free of cancellation)
//   MON_PMAX   max p                MON_PMIN  min p
//   MON_PSUM   sum p                MON_B2    sum |B|^2 (MHD; 0 otherwise)
//   MON_SHIFT  c
// NaN in any cell propagates into the max / min entries.
// ----------------------------------------------------------------------------
enum MonField {
  MON_MASS = 0, MON_RV2, MON_VMAX, MON_VMIN, MON_RMAX, MON_RMIN, MON_RS1, MON_RS2, MON_PMAX, MON_PMIN,
  MON_PSUM, MON_B2, MON_NRED, MON_SHIFT = MON_NRED, MON_N
};

__device__ __forceinline__ double dmin_nan(double a, double b) {
  if (a != a) return a;
  if (b != b) return b;
  return a < b ? a : b;
}
__device__ __forceinline__ bool mon_is_max(int f) { return f == MON_VMAX || f == MON_RMAX || f == MON_PMAX; }
__device__ __forceinline__ bool mon_is_min(int f) { return f == MON_VMIN || f == MON_RMIN || f == MON_PMIN; }
__device__ __forceinline__ double mon_init(int f) {
  return mon_is_max(f) ? -__builtin_inf() : (mon_is_min(f) ? __builtin_inf() : 0.0);
}
__device__ __forceinline__ double mon_join(int f, double a, double b) {
  return mon_is_max(f) ? dmax_nan(a, b) : (mon_is_min(f) ? dmin_nan(a, b) : a + b);
}

// NCOMP: 4 (2D Euler), 5 (3D Euler), 8 (3D ideal MHD)
template <typename T, int NCOMP>
__global__ __launch_bounds__(NT) void k_monitor_partial(const T* u, long long fs, long long ss, long long n,
                                                        T gamma, double* part) {
  __shared__ double red[MON_NRED][NT];
  const T* us = u + (long long)blockIdx.y * ss;
  const double c = double(us[0]);
  double acc[MON_NRED];
#pragma unroll
  for (int f = 0; f < MON_NRED; ++f) acc[f] = mon_init(f);
  for (long long i = (long long)blockIdx.x * NT + threadIdx.x; i < n; i += (long long)gridDim.x * NT) {
    const T rho = us[i];
    const T ux = us[fs + i] / rho, uy = us[2 * fs + i] / rho;
    const T uz = NCOMP >= 5 ? us[3 * fs + i] / rho : T(0);
    const T q2 = NCOMP >= 5 ? (ux * ux + uy * uy) + uz * uz : ux * ux + uy * uy;
    const T E = us[(NCOMP == 4 ? 3 : 4) * fs + i];
    T b2 = T(0);
    if constexpr (NCOMP == 8) {
      const T bx = us[5 * fs + i], by = us[6 * fs + i], bz = us[7 * fs + i];
      b2 = (bx * bx + by * by) + bz * bz;
    }
    const T p = NCOMP == 8 ? (gamma - T(1)) * ((E - (T(0.5) * rho) * q2) - T(0.5) * b2)
                           : (gamma - T(1)) * (E - (T(0.5) * rho) * q2);
    const double v = sqrt(double(q2)), r = double(rho), dr = r - c;
    const double val[MON_NRED] = {r, r * double(q2), v, v, r, r, dr, dr * dr, double(p), double(p), double(p),
                                  double(b2)};
#pragma unroll
    for (int f = 0; f < MON_NRED; ++f) acc[f] = mon_join(f, acc[f], val[f]);
  }
#pragma unroll
  for (int f = 0; f < MON_NRED; ++f) red[f][threadIdx.x] = acc[f];
  lds_barrier();
  for (int w = NT / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
#pragma unroll
      for (int f = 0; f < MON_NRED; ++f) red[f][threadIdx.x] = mon_join(f, red[f][threadIdx.x], red[f][threadIdx.x + w]);
    }
    lds_barrier();
  }
  if (threadIdx.x < MON_NRED)
    part[((long long)blockIdx.y * gridDim.x + blockIdx.x) * MON_NRED + threadIdx.x] = red[threadIdx.x][0];
}

// one block per run: reduce nb partials -> out[run * MON_N + f]
template <typename T>
__global__ __launch_bounds__(NT) void k_monitor_final(const double* part, int nb, const T* u, long long ss,
                                                      double* out) {
  __shared__ double red[MON_NRED][NT];
  const double* ps = part + (long long)blockIdx.x * nb * MON_NRED;
  double acc[MON_NRED];
#pragma unroll
  for (int f = 0; f < MON_NRED; ++f) acc[f] = mon_init(f);
  for (int i = threadIdx.x; i < nb; i += NT) {
#pragma unroll
    for (int f = 0; f < MON_NRED; ++f) acc[f] = mon_join(f, acc[f], ps[(long long)i * MON_NRED + f]);
  }
#pragma unroll
  for (int f = 0; f < MON_NRED; ++f) red[f][threadIdx.x] = acc[f];
  lds_barrier();
  for (int w = NT / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
#pragma unroll
      for (int f = 0; f < MON_NRED; ++f) red[f][threadIdx.x] = mon_join(f, red[f][threadIdx.x], red[f][threadIdx.x + w]);
    }
    lds_barrier();
  }
  if (threadIdx.x < MON_NRED) out[(long long)blockIdx.x * MON_N + threadIdx.x] = red[threadIdx.x][0];
  if (threadIdx.x == 0) out[(long long)blockIdx.x * MON_N + MON_SHIFT] = double(u[(long long)blockIdx.x * ss]);
}

// ----------------------------------------------------------------------------
// per-run time bookkeeping of the run loop (phrike/solver.py:764-797):
//   dt = min(dt, t_end - t) before a step; t += dt and
//   active = t < t_end - 1e-12 after it.  FP64 throughout.
// ----------------------------------------------------------------------------
static __global__ __launch_bounds__(NT) void k_runs_clip(int nb, const double* t, const double* t_end, double* dt,
                                                  const int* active) {
  const int r = blockIdx.x * NT + threadIdx.x;
  if (r < nb && active[r]) {
    const double rem = t_end[r] - t[r];
    if (rem < dt[r]) dt[r] = rem;
  }
}

// A run whose dt is not finite (a NaN max speed) retires with t = NaN: the
// caller reports it as failed (the reference's while t < t_end loop ends the
// same way, phrike/solver.py:1475).
static __global__ __launch_bounds__(NT) void k_runs_advance(int nb, double* t, const double* t_end,
                                                     const double* dt, int* active, int* steps) {
  const int r = blockIdx.x * NT + threadIdx.x;
  if (r < nb && active[r]) {
    const double d = dt[r];
    const double tn = (d - d == 0.0) ? t[r] + d : __builtin_nan("");
    t[r] = tn;
    steps[r] += 1;
    active[r] = tn < t_end[r] - 1e-12 ? 1 : 0;
  }
}

// ----------------------------------------------------------------------------
// Device-resident run loop (phrike/solver.py:1475-1484 for 3D, 764-779 for
// 2D): per-run scalars in rs[field * nb + run] (RS_* fields, FP64) and
// rf[field * nb + run] (RF_* fields, int32).  Per step: clip dt to t_end,
// step with dt, CFL dt of the new state into dt_next, then advance: t += dt,
// dt_used = dt, dt = dt_next, steps += 1, and the run stays active while
// t < t_end - 1e-12 and t + 1e-12 < t_stop (t_stop: the caller's next output
// or checkpoint time, where it wants the state back).  A non-finite dt
// retires the run with t = NaN.
// ----------------------------------------------------------------------------
enum RunScalar { RS_T = 0, RS_DT = 1, RS_TEND = 2, RS_TSTOP = 3, RS_DTNEXT = 4, RS_DTUSED = 5, RS_N = 6 };
enum RunFlag { RF_ACTIVE = 0, RF_STEPS = 1, RF_N = 2 };

static __global__ __launch_bounds__(NT) void k_run_clip(int nb, double* rs, const int* rf) {
  const int r = blockIdx.x * NT + threadIdx.x;
  if (r < nb && rf[RF_ACTIVE * nb + r]) {
    const double rem = rs[RS_TEND * nb + r] - rs[RS_T * nb + r];
    if (rem < rs[RS_DT * nb + r]) rs[RS_DT * nb + r] = rem;
  }
}

static __global__ __launch_bounds__(NT) void k_run_advance(int nb, double* rs, int* rf) {
  const int r = blockIdx.x * NT + threadIdx.x;
  if (r < nb && rf[RF_ACTIVE * nb + r]) {
    const double d = rs[RS_DT * nb + r];
    const double tn = (d - d == 0.0) ? rs[RS_T * nb + r] + d : __builtin_nan("");
    rs[RS_T * nb + r] = tn;
    rs[RS_DTUSED * nb + r] = d;
    rs[RS_DT * nb + r] = rs[RS_DTNEXT * nb + r];
    rf[RF_STEPS * nb + r] += 1;
    rf[RF_ACTIVE * nb + r] = (tn < rs[RS_TEND * nb + r] - 1e-12 && tn + 1e-12 < rs[RS_TSTOP * nb + r]) ? 1 : 0;
  }
}

}  // namespace ph
