// fg_kernels.hip -- gfx950 kernels and C ABI of the raster-scan generalised
// geodesic distance transform (see include/fastgeodis_hip.h).
//
// Semantics (reference FastGeodis v1.0.5):
//   candidate  cur = fl(fl(d_q + fl(l_eucl*LD)) + fl(l_grad*l)),  l = sum_c |I_c(p) - I_c(q)|
//              (fastgeodis_cpu.cpp:88-110, 257-277)
//   update     d(p) = min(d(p), cur) keeping d(p) unless cur is strictly smaller
//   passes     2D: rows then columns (fastgeodis_cpu.cpp:177-198)
//              3D: D, H, W pass pairs with per-pass step tables (fastgeodis_cpu.cpp:220-231, 359-404)
//
// Layout in HBM.  Every sweep runs along a non-contiguous axis A with lanes on
// the contiguous axis E and wave rows on the remaining axis B:
//   D pass: A=D, B=H, E=W on the native (D,H,W) array
//   H pass: A=H, B=D, E=W on the native array
//   W pass: A=W, B=H, E=D on a transposed (W,H,D) copy (image transposed once
//           per call; the distance enters that layout through the H pass's
//           backward sweep, which writes it directly, and leaves it through
//           the W pass's backward sweep, which writes the native layout; a
//           transpose instead where those forms do not apply (3-4 channels,
//           other shapes and halo depths) and after a signed or GSF
//           transform's last W pass, where it applies the epilogue)
//   2D: rows A=H, E=W on (H,W); columns A=W, E=H on (W,H), the backward row
//       and column sweeps on the 30-slice ring writing the other layout
// B and E are the reference's in-plane axes in its neighbour order (B outer,
// E inner), so the exact-order path visits candidates in the reference order.
//
// Engines.
//   tiled (default): one persistent kernel per sweep (all workgroups resident).
//     A workgroup owns a tile of the slice plane and walks along A.  Its
//     front (the previous slice) lives in VGPRs: lane = E position, register
//     = B row; E neighbours arrive through DPP wave shifts folded into the
//     arithmetic, B neighbours are other registers or, across waves, LDS.
//     The window is the tile grown by K on every tiled side; every K slices a
//     tile publishes its own values as (value, tag) pairs with device-scope
//     relaxed 64-bit stores and reloads its K-wide halo from its neighbours'
//     publications, spinning on the tag (bounded; expiry sets a status word).
//     Between exchanges the halo is recomputed redundantly: values at window
//     positions more than j slices from the window edge are exact at step j,
//     so the tile's own positions are exact for all K steps.
//   per-slice: one small launch per slice (any channel count); kept as the
//     fallback and as an independent second path for tests.
//
// Arithmetic.  Compiled with -ffp-contract=off (no FMA) and without fast-math.
// With l_grad and every l_eucl*LD non-negative (sign bit clear), no candidate
// can be -0, so min over candidates with IEEE minNum (v_min3_f32, NaN
// candidates ignored) followed by "(m < d) ? m : d" equals the reference's
// sequential std::min fold bit for bit; otherwise the exact path folds the
// candidates one by one in the reference order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "fastgeodis_hip.h"

namespace {

constexpr float kInf = __builtin_huge_valf();
// Deepest halo: 2D rows use up to 24 of a 64-lane window (one-wave tiles are
// bound by the exchange, and a lone image leaves most CUs idle); 3D windows
// of 32 rows stop at 15 by themselves.
constexpr int kMaxK = 30;
// s_sleep units (64 clocks each) between publishing a halo and first polling
// the neighbours' publications.
constexpr int kHaloPollDelay = 20;
// The same for 16-row windows (measured: 12 on 128 x 256 x 256 volumes, 3 %
// faster than 24; 12 made the 32-row windows of 512^2 planes 1 % slower).
constexpr int kHaloPollDelay16 = 12;
// The same for one-wave row tiles (two neighbours each): 512 clocks.
constexpr int kRowPollDelay = 8;
// Multi-wave tiles skip the slab loads of window rows whose values at that
// step are not exact (see load_rows).
constexpr bool kSkipEdgeLoads = true;
// Waves that perform the halo exchanges in kernels with exchange waves.
constexpr int kCommWaves = 4;
// A-contiguous output (the W pass writing the native layout): slices staged
// per run (16 x 4 B = one 64-byte segment per position; measured in the
// sweep's streaming pattern: runs of 16 cost ~10 % more than plain row
// stores, runs of 8 ~50 %) and the padding between run slots in LDS (words).
constexpr int kRunLen = 16;
constexpr int kRunPad = 4;
// A completed run is stored at the next halo exchange, so the LDS ring also
// holds the slices staged until then (K - 1 at most, the host using this
// form with K == 5).
constexpr int kRunSlots = kRunLen + 4;
// Buckets of the exchange-wait histogram: bucket b counts waits of
// [2^(b-1), 2^b) ticks of the constant-rate device clock (bucket 0: 0 ticks).
constexpr int kXHistBuckets = 32;

// ---------------------------------------------------------------------------
// Descriptors
// ---------------------------------------------------------------------------

struct SweepDesc {
  const float *img;  // image channel 0 in this sweep's layout
  float *d;          // distance in this sweep's layout
  long long cs;      // image channel stride (elements)
  long long sA, sB;  // slice stride and row stride; lanes are contiguous
  int nA, nB, nE;
  int dir;  // +1 forward, -1 backward
  int C;
  float lg;
  float e1[9];  // fl(l_eucl * LD[i*3+j]), i = row offset + 1, j = lane offset + 1
  int nTB, nTE;  // tile grid
  int TB, TE;    // own tile extent
  int tile0;     // first tile of this descriptor within its launch
  long long halo0;  // first halo entry of this descriptor
  float *dxo;       // other-layout output: element (a, b, e) at
  long long sXe, sXa;  //   dxo[e * sXe + a * sXa + b] (XST, b contiguous)
  long long sXb;       //   dxo[e * sXe + b * sXb + a] (XSA, a contiguous)
  // First sweep of a transform: the input distance is the initialisation,
  // computed from the mask as it is loaded (init_mode 1: fl(v*m), 2:
  // fl(v*fl(1-m))), and the sweep's untouched first slice is written too.
  const float *init_src;  // mask channel 0 in this sweep's layout (init_mode != 0)
  float init_v;
  int init_mode;
};

// Transposes, optionally combining several sources elementwise (the epilogues
// of signed and GSF transforms, applied while their distances leave the W
// pass layout).
enum XposeKind : int {
  XK_COPY = 0,       // dst = src
  XK_SIGNED = 1,     // dst = src - src2                      (D(M) - D(1-M))
  XK_GSF_MID = 2,    // ds = src - src2; dst..dst4 = the inits of the four second-stage transforms
  XK_GSF_FINAL = 3,  // dst = -(src2 - src3) + (src4 - src)
};

struct XposeDesc {  // out[c][w][h][d] = in[c][d][h][w]
  const float *src;
  float *dst;
  int C, D, H, W;
  int tiles_dw;  // tiles per (c, h) plane
  int tile0;
  int kind;                          // XposeKind (C == 1 unless XK_COPY)
  const float *src2, *src3, *src4;   // further sources, same layout as src
  float *dst2, *dst3, *dst4;         // further outputs (XK_GSF_MID), same layout as dst
  float v, theta;
};

// ---------------------------------------------------------------------------
// Device helpers
// ---------------------------------------------------------------------------

// Lane l receives lane l-1 (J == 0) or lane l+1 (J == 2).  The wave's edge
// lane receives 0 (wave shift) or the opposite edge lane's value (ALT: wave
// rotate); that lane is always a halo lane, whose value stays outside the
// part of the window that is exact, or a pad lane, forced to +inf each step,
// so both forms give the same results.  A shift that several candidates use
// is computed once by the compiler and cannot fold into its users; shifting
// one front row for two new rows with the two different forms keeps each
// shift single-use, so it folds into the VALU instruction that consumes it
// (v_add_f32_dpp / v_sub_f32_dpp).  TAG keeps other call sites distinct.
template <int J, int TAG, bool ALT = false>
__device__ __forceinline__ float shiftE(float x) {
  if constexpr (J == 1) return x;
  else {
    constexpr int ctrl = (J == 0 ? 0x138 : 0x130) | (ALT ? 0x4 : 0x0);  // wave_shr/ror:1, wave_shl/rol:1
    return __int_as_float(__builtin_amdgcn_update_dpp(TAG, __float_as_int(x), ctrl, 0xF, 0xF, true));
  }
}

// Workgroups are dispatched round-robin over the 8 XCDs; give each XCD a
// contiguous range of tiles so that most neighbours share an L2.
__device__ __forceinline__ int xcd_remap(int g, int G) {
  constexpr int nx = 8;
  const int x = g % nx, i = g / nx;
  const int q = G / nx, rem = G % nx;
  return x * q + min(x, rem) + i;
}

__device__ __forceinline__ unsigned long long pack(float v, unsigned tag) {
  return ((unsigned long long)tag << 32) | (unsigned long long)__float_as_uint(v);
}

// sum over channels C0..CT-1 of |I_c(p) - I_c(q)|, accumulated left to right
template <int I, int J, int C0, int CT>
__device__ __forceinline__ float absdiff_sum(float acc, const float (&Iq)[CT], const float (&In)[CT]) {
  if constexpr (C0 >= CT) {
    return acc;
  } else {
    const float q = shiftE<J, 1 + C0 + I * 16 + J * 4, (I & 1) != 0>(Iq[C0]);
    return absdiff_sum<I, J, C0 + 1, CT>(acc + fabsf(q - In[C0]), Iq, In);
  }
}

// Candidate for the element at this lane from front row I (B offset I-1) and
// lane offset J-1:  (d_q + e1) + l_grad * sum_c |I_c(p) - I_c(q)|.  The
// channel sum starts from |I_0(p) - I_0(q)|, which equals the reference's
// 0.0f + |..| bit for bit.
template <int I, int J, int CT, bool LAMB1>
__device__ __forceinline__ float candidate(float dq, const float (&Iq)[CT], const float (&In)[CT], float e1,
                                           float lg) {
  // |q - p| == |p - q| exactly; the shifted operand goes first so that the
  // shift folds into v_sub_f32_dpp without commuting
  const float l =
      absdiff_sum<I, J, 1, CT>(fabsf(shiftE<J, 1 + I * 16 + J * 4, (I & 1) != 0>(Iq[0]) - In[0]), Iq, In);
  const float ds = shiftE<J, 64 + I * 4 + J, (I & 1) != 0>(dq);
  if constexpr (LAMB1) {
    // l_grad == 1 and l_eucl*LD == +0: fl(1*l) == l and dropping "+ 0" is exact here
    return ds + l;
  } else {
    return (ds + e1) + lg * l;
  }
}

template <int I, int CT, bool LAMB1>
__device__ __forceinline__ void candidates_row(float (&c)[9], float dq, const float (&Iq)[CT],
                                               const float (&In)[CT], const float (&e1)[9], float lg) {
  c[I * 3 + 0] = candidate<I, 0, CT, LAMB1>(dq, Iq, In, e1[I * 3 + 0], lg);
  c[I * 3 + 1] = candidate<I, 1, CT, LAMB1>(dq, Iq, In, e1[I * 3 + 1], lg);
  c[I * 3 + 2] = candidate<I, 2, CT, LAMB1>(dq, Iq, In, e1[I * 3 + 2], lg);
}

// New value of one row element.  dq[i] / Iq[i][c]: front rows at B offset
// i-1 (unshifted); In[c] / dinit: the element's own new-slice values.
template <int CT, bool LAMB1, bool FAST>
__device__ __forceinline__ float relax(const float (&dq)[3], const float (&Iq)[3][CT],
                                       const float (&In)[CT], float dinit, const float (&e1)[9],
                                       float lg) {
  float c[9];
  candidates_row<0, CT, LAMB1>(c, dq[0], Iq[0], In, e1, lg);
  candidates_row<1, CT, LAMB1>(c, dq[1], Iq[1], In, e1, lg);
  candidates_row<2, CT, LAMB1>(c, dq[2], Iq[2], In, e1, lg);
  if constexpr (FAST) {
    float m = fminf(fminf(c[0], c[1]), c[2]);
    m = fminf(fminf(m, c[3]), c[4]);
    m = fminf(fminf(m, c[5]), c[6]);
    m = fminf(fminf(m, c[7]), c[8]);
    return (m < dinit) ? m : dinit;
  } else {
    float nd = dinit;
#pragma unroll
    for (int k = 0; k < 9; ++k) nd = (c[k] < nd) ? c[k] : nd;
    return nd;
  }
}

// New value of an element of a one-row plane: the three candidates of the
// front row (lanes -1, 0, +1; table entries 3..5).  The six candidates of
// rows -1 and +1, which lie outside such a plane, are skipped as the
// reference skips out-of-range neighbours.
template <int CT, bool LAMB1, bool FAST>
__device__ __forceinline__ float relax_row(float dq, const float (&Iq)[CT], const float (&In)[CT], float dinit,
                                           const float (&e1)[9], float lg) {
  const float c3 = candidate<1, 0, CT, LAMB1>(dq, Iq, In, e1[3], lg);
  const float c4 = candidate<1, 1, CT, LAMB1>(dq, Iq, In, e1[4], lg);
  const float c5 = candidate<1, 2, CT, LAMB1>(dq, Iq, In, e1[5], lg);
  if constexpr (FAST) {
    const float m = fminf(fminf(c3, c4), c5);
    return (m < dinit) ? m : dinit;
  } else {
    float nd = dinit;
    nd = (c3 < nd) ? c3 : nd;
    nd = (c4 < nd) ? c4 : nd;
    nd = (c5 < nd) ? c5 : nd;
    return nd;
  }
}

// ---------------------------------------------------------------------------
// Tiled persistent sweep kernel
// ---------------------------------------------------------------------------

// Global-address-space view of the halo publications (64-bit atomics use
// SGPR base + VGPR offset addressing instead of flat addressing).
typedef __attribute__((address_space(1))) unsigned long long gu64;

template <int R, int CT>
struct Slab {  // one prefetched slice: image rows -1..R, distance rows 0..R-1
  float I[CT][R + 2];
  float d[R];
};

// NW waves per workgroup, R window rows per wave (register rows), P slices
// prefetched ahead, CT image channels.
// XS, the output form: 0 writes slice a in the sweep's own layout.
//   1 (XST): writes it to Dd.dxo with B contiguous (the H pass writes the
//     (W, H, D) layout of the W pass), staging each slice's window through
//     LDS so each store covers a run of consecutive B.
//   2 (XSA): writes it to Dd.dxo with A contiguous: multi-wave tiles (the W
//     pass writing the native (D, H, W) layout) stage kRunLen consecutive
//     slices per wave in LDS so that each store covers runs of kRunLen
//     consecutive A; one-row tiles on the deep ring (2D rows writing the
//     column layout and back) hold each block in registers and store a run
//     of K elements per own lane at the exchange.
// INIT: the first sweep of a transform, whose input distance is computed from
// the mask as it is loaded (SweepDesc::init_mode).
// COMM: kCommWaves more waves (indices NW, NW + 1, ...) that hold no rows and
// perform every halo exchange of the tile, each for its share of the window
// rows: the compute waves leave their front in LDS and wait at a barrier while
// they publish and poll, so the compute waves' prefetch loads stay in flight
// (a wave's wait for a halo load is also a wait for every load it issued
// earlier).
template <int NW, int R, int P, int CT, bool LAMB1, bool FAST, int XS = 0, bool INIT = false,
          bool COMM = false>
__global__ __launch_bounds__((NW + (COMM ? kCommWaves : 0)) * 64) void sweep_tiled_kernel(
    const SweepDesc *__restrict__ descs, int ndesc, SweepDesc one, int K, unsigned tag_base,
    unsigned long long *__restrict__ halo_base, int *__restrict__ status, int spin_limit,
    unsigned *__restrict__ xhist) {
  // LDS slot = ring position of the step (a compile-time index in the
  // unrolled loops, so edge addresses are immediate offsets)
  constexpr int NWL = NW > 1 ? NW : 1;
  constexpr bool XST = XS == 1, XSA = XS == 2;
  // A-contiguous output of multi-wave tiles (XSW, runs staged in LDS) and of
  // one-row tiles (XSR, each block's slices held in registers)
  constexpr bool XSR = XSA && NW == 1 && R == 1, XSW = XSA && !XSR;
  // XSW needs the LDS for its runs: two edge slots, alternating by step
  // parity (one barrier per step separates a slot's reads from its rewrite
  // two steps later)
  constexpr int ES = XSW ? 2 : P;
  __shared__ float edge_first[ES][NWL][64];
  __shared__ float edge_last[ES][NWL][64];
  __shared__ float edgeI_first[ES][NWL][CT][64];
  __shared__ float edgeI_last[ES][NWL][CT][64];
  constexpr int NWR = NW * R;  // window rows
  // one wave of one row: the plane is a single row (nB == 1), so the only
  // neighbours are lanes -1, 0, +1 of the front row itself
  constexpr bool ROW1 = NW == 1 && R == 1;
  static_assert(!XST || (NW > 1 && 64 % NWR == 0), "transposed output needs a multi-wave window of 16/32/64 rows");
  static_assert(!XSW || (NW > 1 && !COMM && !INIT && NWR > 2 * P), "A-contiguous output is a plain-sweep form");
  static_assert(!XSR || (!COMM && !INIT), "one-row deferred stores are a plain-sweep form");
  __shared__ float xtile[XST ? P : 1][XST ? NWR : 1][65];
  // XSA: slot a % kRunSlots holds slice a, one 64-lane row per own row of
  // the tile (window rows K .. K + TB - 1; the host uses this form only on
  // planes of several tiles along B, where TB <= NWR - 2K); the slots are
  // kRunPad words apart beyond their rows, so the 16 lanes that read one
  // position's run hit different banks
  constexpr int XOR = NWR - 2 * P;  // own rows at the shallowest halo (+ 1 spare row)
  __shared__ float xrun[XSW ? kRunSlots : 1][XSW ? (XOR + 1) * 64 + kRunPad : 1];
  __shared__ unsigned xhist_wg[kXHistBuckets];  // this workgroup's exchange-wait counts (measurement)
  static_assert(!COMM || (NW > 1 && NWR % kCommWaves == 0), "exchange waves serve multi-wave sweeps");
  __shared__ float xfront[COMM ? NWR : 1][64];     // COMM: the front around an exchange

  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave index: uniform
  const int tile = xcd_remap(blockIdx.x, gridDim.x);
  if (xhist) {
    if (threadIdx.x < kXHistBuckets) xhist_wg[threadIdx.x] = 0u;
    __syncthreads();
  }

  // descriptor owning this tile: the launch's only one (passed by value when
  // descs is null) or a binary search on tile0 (uniform branches and loads,
  // so the copy's fields stay in scalar registers)
  SweepDesc Dd;
  if (descs) {
    int lo = 0, hi = ndesc - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (descs[mid].tile0 <= tile) lo = mid; else hi = mid - 1;
    }
    Dd = descs[lo];
  } else {
    Dd = one;
  }
  gu64 *halo = (gu64 *)halo_base + Dd.halo0;
  const long long cs = Dd.cs, sA = Dd.sA, sB = Dd.sB;
  const int nA = Dd.nA, nB = Dd.nB, nE = Dd.nE, dir = Dd.dir;
  const float lg = Dd.lg;
  float e1[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) e1[k] = Dd.e1[k];
  const int nTB = Dd.nTB, nTE = Dd.nTE, TB = Dd.TB, TE = Dd.TE;
  const int lt = tile - Dd.tile0;
  const int tb = lt / nTE, te = lt - tb * nTE;
  const int b_own0 = tb * TB, b_own1 = min(nB, b_own0 + TB);
  const int e_own0 = te * TE, e_own1 = min(nE, e_own0 + TE);
  const int bw0 = (nTB > 1) ? b_own0 - K : 0;
  const int ew0 = (nTE > 1) ? e_own0 - K : -1;
  const int e = ew0 + lane;
  const bool lane_valid = (e >= 0) && (e < nE);
  const unsigned ec = (unsigned)min(max(e, 0), nE - 1);  // clamped: loads never leave the row
  const unsigned ecb = ec * 4u;                             // its byte offset within a row
  const bool lane_own = (e >= e_own0) && (e < e_own1);
  const long long tile_slots = (long long)TB * TE;
  // halo publication slot of this lane (own lanes) and source tile column
  // of this lane (halo lanes)
  const int lane_pub = e - e_own0;
  const int ote = lane_valid ? (e / TE) : 0;
  const int oe0 = ote * TE;
  const int ow = min(nE, oe0 + TE) - oe0;
  const int lane_src = e - oe0;

  // The block loop below advances in groups of P steps and exchanges halos
  // every K steps; it stays within the sweep only when K is a positive
  // multiple of P.  A launch that breaks this is refused before any store,
  // and so is an A-contiguous output whose own rows or staged slices would
  // not fit its LDS ring, or (one-row tiles) whose block is not one turn of
  // the slab ring.
  if (K <= 0 || K % P != 0 || (ROW1 && Dd.nB != 1) ||
      (XSW && (Dd.nTB < 2 || Dd.TB > XOR || K > kRunSlots - kRunLen + 1)) || (XSR && K != P)) {
    if (threadIdx.x == 0) __hip_atomic_store(status, FG_EINVAL, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    return;
  }
  const int nsteps = nA - 1;
  if (nsteps <= 0) return;
  const int a_first = (dir > 0) ? 1 : nA - 2;
  const int nblocks_all = (nsteps + K - 1) / K;

  // The exchange wave: the same barrier sequence as the compute waves (one
  // per step, two per exchange), publishing and polling between the two
  // barriers of each exchange for every row of the window.
  if constexpr (COMM) {
    if (w >= NW) {
      constexpr int CR = NWR / kCommWaves;  // window rows per exchange wave
      const int cr0 = (w - NW) * CR;
      bool tripped = false;
      for (int blk = 0;; ++blk) {
        for (int st = 0; st < K; ++st) __syncthreads();
        if (blk + 1 >= nblocks_all) break;
        __syncthreads();  // the compute waves' front is in xfront
        const unsigned long long t_in = xhist ? wall_clock64() : 0ull;
        const unsigned tag = tag_base + (unsigned)(blk + 1);
        const long long hslot = (long long)((blk + 1) & 1) * tile_slots;
        gu64 *pub = halo + (long long)lt * 2 * tile_slots + hslot;
        const int own_w = e_own1 - e_own0;
        const bool lane_band = (te > 0 && e - e_own0 < K) || (te < nTE - 1 && e_own1 - 1 - e < K);
        const int lane_dist = max(max(e_own0 - e, e - (e_own1 - 1)), 0);
#pragma unroll
        for (int k = 0; k < CR; ++k) {
          const int r = cr0 + k;
          const int b = bw0 + r;
          const bool row_band = (tb > 0 && b - b_own0 < K) || (tb < nTB - 1 && b_own1 - 1 - b < K);
          if (b >= b_own0 && b < b_own1 && lane_own && (row_band || lane_band))
            __hip_atomic_store(pub + (long long)(b - b_own0) * own_w + lane_pub, pack(xfront[r][lane], tag),
                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        __builtin_amdgcn_s_sleep(kHaloPollDelay);
        // whether this lane of window row r needs a neighbour's value, and
        // where that neighbour publishes it (recomputed rather than held, so
        // only the 32 loaded words stay live)
        auto row_need = [&](int r) {
          const int b = bw0 + r;
          const bool row_own = b >= b_own0 && b < b_own1;
          const int row_dist = max(max(b_own0 - b, b - (b_own1 - 1)), 0);
          return b >= 0 && b < nB && lane_valid && !(row_own && lane_own) && max(row_dist, lane_dist) <= K;
        };
        auto row_src = [&](int r) -> const gu64 * {
          const int bq = min(max(bw0 + r, 0), nB - 1);
          const int otb = bq / TB;  // uniform
          return halo + (long long)(otb * nTE + ote) * 2 * tile_slots + hslot + (long long)(bq - otb * TB) * ow +
                 lane_src;
        };
        unsigned long long x[CR];
#pragma unroll
        for (int k = 0; k < CR; ++k) {
          x[k] = (unsigned long long)tag << 32;  // lanes that need nothing count as arrived
          if (row_need(cr0 + k)) x[k] = __hip_atomic_load(row_src(cr0 + k), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        int spins = tripped ? spin_limit : 0;
        while (true) {
          bool waiting = false;
#pragma unroll
          for (int k = 0; k < CR; ++k) waiting = waiting || (unsigned)(x[k] >> 32) != tag;
          if (!waiting) break;
          if (++spins > spin_limit) {
            __hip_atomic_store(status, FG_EWATCHDOG, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            tripped = true;
            break;
          }
#pragma unroll
          for (int k = 0; k < CR; ++k)
            if ((unsigned)(x[k] >> 32) != tag)
              x[k] = __hip_atomic_load(row_src(cr0 + k), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
#pragma unroll
        for (int k = 0; k < CR; ++k)
          if (row_need(cr0 + k)) xfront[cr0 + k][lane] = __uint_as_float((unsigned)x[k]);
        if (xhist && lane == 0) {
          const unsigned long long dt = wall_clock64() - t_in;
          const int bucket = dt ? min(64 - __clzll(dt), kXHistBuckets - 1) : 0;
          atomicAdd(&xhist_wg[bucket], 1u);
        }
        __syncthreads();  // the halo is in xfront
      }
      if constexpr (XST) __syncthreads();  // the last staged slice
      if (xhist) __syncthreads();           // the workgroup's histogram flush
      return;
    }
  }

  // Row bookkeeping: rows -1..R relative to this wave.  rofs[r] is the byte
  // offset of this lane's element of row r (clamped into the plane) within
  // a slice; the host runs this kernel only on arrays of less than 4 GiB per
  // channel, so every offset in a slice fits 32 bits and all loads and
  // stores use a per-slice buffer resource plus this loop-invariant offset.
  int rb[R + 2];
  bool rvalid[R + 2];
  unsigned rofs[R + 2];
#pragma unroll
  for (int r = 0; r < R + 2; ++r) {
    rb[r] = bw0 + w * R + r - 1;
    rvalid[r] = rb[r] >= 0 && rb[r] < nB;
    rofs[r] = (unsigned)((long long)min(max(rb[r], 0), nB - 1) * sB * 4) + ecb;
  }
  // Buffer resources are raw and byte-addressed; offsets at or beyond the
  // range are out of bounds, and stores there are dropped.
  // Slice cursor of a step: its slice a and that slice's byte offset
  // a * sA * 4, advanced together by one slice per step (two scalar adds
  // instead of a 64-bit multiply per array).  A slice outside [0, nA) (the
  // prefetches and stores of steps past the end of the sweep) gets an empty
  // buffer range: its loads return 0 and its stores are dropped, without a
  // memory access.
  // FLAT (the one-row kernels, whose steps are latency-bound): one buffer
  // resource per array for the whole sweep, with the array as its range
  // (nA * sA * 4 bytes), and 32-bit offsets a * sA * 4 + lane offset, so a
  // step costs two vector adds instead of scalar resource arithmetic.  The
  // host runs these kernels only when (nA + 32) * sA * 4 <= 2^31: then the
  // offsets of slices just past either end (a in [-P, nA + P), P <= 32) wrap
  // to or stay at or beyond the range, and so do the store offsets of lanes
  // that store nothing, which are based at 2^31.
  constexpr bool FLAT = ROW1;
  struct Cursor {
    int a;
    long long off;
    unsigned foff;  // FLAT: off modulo 2^32
  };
  const long long dAb = (long long)dir * sA * 4;
  auto cursor_of = [&](int step) {
    const int a = a_first + step * dir;
    const long long off = (long long)a * sA * 4;
    return Cursor{a, off, (unsigned)off};
  };
  auto advance = [&](Cursor &c) {
    if constexpr (FLAT) {
      c.foff += (unsigned)dAb;
    } else {
      c.a += dir;
      c.off += dAb;
    }
  };
  const unsigned flat_range = FLAT ? (unsigned)((long long)nA * sA * 4) : 0u;
  auto cursor_rsrc = [&](const float *base, const Cursor &c) {
    if constexpr (FLAT) {
      return __builtin_amdgcn_make_buffer_rsrc((void *)base, 0, flat_range, 0x00020000);
    } else {
      const bool in = (unsigned)c.a < (unsigned)nA;
      return __builtin_amdgcn_make_buffer_rsrc((void *)((const char *)base + c.off), 0, in ? 0xFFFFFFF0u : 0u,
                                               0x00020000);
    }
  };
  // byte offset of a lane's element within its resource; the out-of-range
  // offset of a lane that stores nothing
  auto at = [&](unsigned lofs, const Cursor &c) { return FLAT ? lofs + c.foff : lofs; };
  auto nowhere = [&](const Cursor &c) { return FLAT ? 0x80000000u + c.foff : 0xFFFFFFFFu; };

  // Transposed output: wave w moves window columns (w * R + k) * (64 / NWR)
  // + lane / NWR, rows lane % NWR, for k < R; the byte offset of each element
  // in its output slice (out-of-range for positions this tile does not own).
  unsigned xoff[XST ? R : 1];
  if constexpr (XST) {
#pragma unroll
    for (int k = 0; k < R; ++k) {
      const int col = (w * R + k) * (64 / NWR) + lane / NWR, row = lane % NWR;
      const int eg = ew0 + col, bg = bw0 + row;
      const bool own = eg >= e_own0 && eg < e_own1 && bg >= b_own0 && bg < b_own1;
      xoff[k] = own ? (unsigned)(((long long)eg * Dd.sXe + bg) * 4) : 0xFFFFFFFFu;
    }
  }
  // stores the staged window of step sp (LDS slot ps) to the output; step -1
  // is the sweep's first slice, which the sweep itself never changes but the
  // output array must also hold
  auto xflush = [&](int ps, int sp) {
    if constexpr (XST) {
      const bool live = sp >= -1 && sp < nsteps;
      const int ap = min(max(a_first + sp * dir, 0), nA - 1);
      const __amdgpu_buffer_rsrc_t rx = __builtin_amdgcn_make_buffer_rsrc(
          (void *)(Dd.dxo + (long long)ap * Dd.sXa), 0, live ? 0xFFFFFFF0u : 0u, 0x00020000);
#pragma unroll
      for (int k = 0; k < R; ++k) {
        const int col = (w * R + k) * (64 / NWR) + lane / NWR, row = lane % NWR;
        __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(xtile[ps][row][col]), rx, (int)xoff[k], 0, 0);
      }
    }
  };

  // Prefetch loads are raw loads at clamped coordinates (always in bounds) and
  // carry no masking, so no instruction waits on them until the slab is
  // consumed.  Values at pad positions (outside the plane) are never used:
  // a pad's own result is forced to +inf after every step, and its image
  // value only enters candidates whose distance operand is that +inf.
  // With several waves, image rows -1 and R of a slab are the neighbour
  // waves' rows R-1 and 0: they arrive through LDS when the slab is consumed
  // (see step), so only rows 0..R-1 are loaded.  The initial front and the
  // one-wave kernel load all rows.
  // the initialisation from the loaded mask value: fl(v*m) (mode 1) or
  // fl(v*fl(1-m)) (mode 2), by a select rather than a branch
  const bool init_inv = Dd.init_mode == 2;
  const float init_v = Dd.init_v;
  auto init_of = [&](float m) -> float {
    if constexpr (INIT) {
      const float t = 1.0f - m;
      return init_v * (init_inv ? t : m);
    } else {
      return m;
    }
  };
  // Slab offsets per ring slot.  With K == P a block is one turn of the ring
  // and slot t always holds block step t + 1.  At block step j the window
  // rows less than j rows from a window edge that faces a neighbour tile
  // depend on positions outside the window: their values are not exact and
  // are never stored or published (the next exchange replaces them), so
  // their loads get an out-of-range offset (the load returns 0 without a
  // memory access).  The image rows of a block's last step (slot P - 1) are
  // all loaded: after the exchange they are the front image of the next
  // block's first step, whose rows are all exact again.  Computed once, so
  // the steps pay nothing for it; with K > P every row is loaded.
  constexpr int SK = (kSkipEdgeLoads && NW > 1) ? P : 1;
  unsigned slot_ofs[SK][R];
#pragma unroll
  for (int t = 0; t < SK; ++t)
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int wr = w * R + r;  // window row (uniform)
      const bool skip = SK > 1 && K == P && nTB > 1 && min(wr, NWR - 1 - wr) <= t;
      slot_ofs[t][r] = skip ? 0xFFFFFFFFu : rofs[r + 1];
    }
  // t: the ring slot the slab fills, or -1 (every row loaded)
  auto load_rows = [&](Slab<R, CT> &S, const Cursor &cur, int r0, int r1, int t) {
    const __amdgpu_buffer_rsrc_t rd = cursor_rsrc(INIT ? Dd.init_src : Dd.d, cur);
    __amdgpu_buffer_rsrc_t ri[CT];
#pragma unroll
    for (int c = 0; c < CT; ++c) ri[c] = cursor_rsrc(Dd.img + c * cs, cur);
#pragma unroll
    for (int r = 0; r < R + 2; ++r) {
      const bool slotted = SK > 1 && t >= 0 && r >= 1 && r <= R;
      const unsigned od = slotted ? slot_ofs[t][r - 1] : at(rofs[r], cur);
      const unsigned oi = (slotted && t < P - 1) ? slot_ofs[t][r - 1] : at(rofs[r], cur);
      // the one-row kernel (a plane of one row) uses only its own image row
      if (r >= r0 && r < r1 && (!ROW1 || r == 1)) {
#pragma unroll
        for (int c = 0; c < CT; ++c)
          S.I[c][r] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(ri[c], (int)oi, 0, 0));
      }
      if (r >= 1 && r <= R)
        S.d[r - 1] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rd, (int)od, 0, 0));
    }
  };
  auto load_slab = [&](Slab<R, CT> &S, const Cursor &cur, int t) {
    if constexpr (NW > 1) load_rows(S, cur, 1, R + 1, t);
    else load_rows(S, cur, 0, R + 2, t);
  };

  // per-row validity of this lane (pad positions hold +inf in the front) and
  // ownership of each wave row (uniform)
  bool pvalid[R], row_own[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    pvalid[r] = rvalid[r + 1] && lane_valid;
    row_own[r] = rb[r + 1] >= b_own0 && rb[r + 1] < b_own1;
  }

  // XSA: stage the wave's own-row values of slice a in LDS slot a % kRunSlots.
  // Runs are the kRunLen slices a0 .. a0 + kRunLen - 1 with a0 a multiple of
  // kRunLen (clipped to the sweep).  A run whose slices have all been staged
  // is stored at the next halo exchange, after the tile's publications,
  // where the wave waits for its earlier memory operations anyway (stores
  // issued inside the steps made the compiler's waits for the prefetched
  // slabs conservative on every path); the K <= kRunSlots -
  // kRunLen slices staged before then fill free slots.  Lane l moves slot
  // a0 + l % 16 of window lane 4j + l / 16, for j < 16, so every 16 lanes
  // store one 64-byte segment of the output.  Only this wave reads what it
  // staged and a wave's LDS accesses complete in order, so no barrier is
  // needed; the wave-scope fences keep the compiler from moving the accesses
  // across each other.  Slices past the end of the array (a partial first or
  // last run) and positions the tile does not own get an out-of-range
  // offset.  The host uses this form only for outputs of less than 4 GiB,
  // so 32-bit offsets address every element (the wrapped products of a
  // negative window lane are replaced before use).
  int xa_a0 = -1;  // first slice of the run awaiting its stores, or -1
  auto xa_flush = [&]() {
    if constexpr (XSW) {
      constexpr int LPR = kRunLen, PPI = 64 / kRunLen;
      static_assert(64 % kRunLen == 0, "run layout");
      if (xa_a0 >= 0) {
        const int q = lane % LPR, pl = lane / LPR;
        const int sq = (xa_a0 + q) % kRunSlots;  // LDS slot of this lane's slice
        const unsigned sXe4 = (unsigned)Dd.sXe * 4u;
        const unsigned lane_base = (unsigned)(ew0 + pl) * sXe4 + (unsigned)(xa_a0 + q) * 4u;
        const bool a_ok = xa_a0 + q < nA;
        const __amdgpu_buffer_rsrc_t rx = __builtin_amdgcn_make_buffer_rsrc((void *)Dd.dxo, 0, 0xFFFFFFF0u, 0x00020000);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
        for (int r = 0; r < R; ++r) {
          if (!row_own[r]) continue;
          const unsigned row_base = (unsigned)rb[r + 1] * (unsigned)Dd.sXb * 4u;
#pragma unroll
          for (int j = 0; j < kRunLen; ++j) {
            const int p = j * PPI + pl;  // window lane
            const int ep = ew0 + p;
            const bool own = a_ok && ep >= e_own0 && ep < e_own1;
            const float x = xrun[sq][(w * R + r - K) * 64 + p];
            const unsigned off = own ? lane_base + (unsigned)(j * PPI) * sXe4 + row_base : 0xFFFFFFFFu;
            __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(x), rx, (int)off, 0, 0);
          }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        xa_a0 = -1;
      }
    }
  };
  // Branch-free staging (a branch in the step keeps the compiler from
  // moving work across it): rows the tile does not own write a spare row,
  // and the padded steps past either end (a >= -(K - 1), K <= kXsaMaxK)
  // write slots that hold no pending slice.
  int xrow[XSW ? R : 1];
#pragma unroll
  for (int r = 0; r < (XSW ? R : 0); ++r) xrow[r] = (row_own[r] ? w * R + r - K : XOR) * 64 + lane;
  auto xa_stage = [&](const float (&v)[R], int a) {
    if constexpr (XSW) {
      const int sa = (a + kRunSlots) % kRunSlots;
#pragma unroll
      for (int r = 0; r < R; ++r) xrun[sa][xrow[r]] = v[r];
      // one run completes per block at most (K < kRunLen); a forward
      // sweep's partial last run is stored after the loop
      const int g = a & (kRunLen - 1);
      const bool done = (unsigned)a < (unsigned)nA && (dir < 0 ? g == 0 : g == kRunLen - 1);
      xa_a0 = done ? a - g : xa_a0;
    }
  };

  // XSR (one-row tiles, K == P): the block's new values stay in registers
  // (slot p = block step p) and are stored at the next exchange, after the
  // publications, element (a, e) at Dd.dxo[e * sXe + a * sXa]: the other 2D
  // layout (sXe = nA, sXa = 1), each own lane writing a run of K consecutive
  // elements.  (Deferring the stores of sweeps into their own array this way
  // measured no faster.)  Slices past either end of the
  // sweep (the padded last block) and lanes the tile does not own get an
  // out-of-range offset.  The host uses this form only for arrays of less
  // than 2 GiB (nA * nE elements).
  float xrr[XSR ? P : 1];
  int xr_s0 = 0;  // first step of the block held in xrr
  const unsigned xr_range = XSR ? (unsigned)((long long)nA * nE * 4) : 0u;
  const unsigned xr_lane = XSR ? (unsigned)e * (unsigned)Dd.sXe * 4u : 0u;  // own lanes: e >= 0
  const unsigned xr_sa4 = XSR ? (unsigned)Dd.sXa * 4u : 0u;
  auto xr_store = [&](float v, int a) {
    if constexpr (XSR) {
      const __amdgpu_buffer_rsrc_t rx = __builtin_amdgcn_make_buffer_rsrc((void *)Dd.dxo, 0, xr_range, 0x00020000);
      const unsigned off = (lane_own && (unsigned)a < (unsigned)nA) ? xr_lane + (unsigned)a * xr_sa4 : 0xFFFFFFFFu;
      __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), rx, (int)off, 0, 0);
    }
  };
  auto xr_flush = [&]() {
    if constexpr (XSR) {
#pragma unroll
      for (int p = 0; p < P; ++p) xr_store(xrr[p], a_first + (xr_s0 + p) * dir);
    }
  };

  // Slab ring: the slab of step s is S[s % P]; its image rows are also the
  // front image of step s + 1, so a slab is refilled (with the slice of step
  // s + P) only after step s + 1 has used it; the front of step 0 starts in
  // S[P - 1].  Slabs stay in fixed registers this way.
  float Fd[R];  // front distance rows 0..R-1
  Slab<R, CT> S[P];
  load_rows(S[P - 1], cursor_of(-1), 0, R + 2, -1);  // front slice: not modified by this sweep
#pragma unroll
  for (int r = 0; r < R; ++r) Fd[r] = pvalid[r] ? init_of(S[P - 1].d[r]) : kInf;
  if constexpr (INIT) {
    // the output's copy of the untouched first slice: its initialisation
    const Cursor c0 = cursor_of(-1);
    const __amdgpu_buffer_rsrc_t rs = cursor_rsrc(Dd.d, c0);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const unsigned off = (row_own[r] && lane_own) ? at(rofs[r + 1], c0) : nowhere(c0);
      __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(Fd[r]), rs, (int)off, 0, 0);
    }
  }
  if constexpr (XST) {  // stage the first slice as step -1 (ring slot P - 1)
#pragma unroll
    for (int r = 0; r < R; ++r) xtile[P - 1][w * R + r][lane] = Fd[r];
  }
  xa_stage(Fd, a_first - dir);  // XSA: the sweep's untouched first slice
  xr_store(Fd[0], a_first - dir);
#pragma unroll
  for (int p = 0; p < P - 1; ++p) load_slab(S[p], cursor_of(p), p);
  // the slices that the loop stores (step s) and prefetches (step s + P - 1)
  Cursor cur_st = cursor_of(0), cur_pf = cursor_of(P - 1);

  // Publishes the tile's own front values that neighbours read and reloads
  // the halo positions of the window from the neighbours' publications.
  //   published: own positions within K of a tile edge that has a neighbour
  //   reloaded:  non-own window positions within K of the own region (the
  //              only ones whose values reach own positions in K steps)
  // A reloaded position's source tile publishes it: it lies within K of that
  // tile's edge facing this tile.  The prefetch of the next slice is issued
  // between the halo loads and the wait for them, so that wait does not also
  // wait for the new prefetch; pending rows are polled together, so one
  // memory round trip serves all of them.
  bool wd_tripped = false;
  // COMM: hand the front to the exchange wave and take the halo back from it
  auto exchange_via_comm = [&](int blk, Slab<R, CT> &pf_slab) {
    (void)blk;
#pragma unroll
    for (int r = 0; r < R; ++r) xfront[w * R + r][lane] = Fd[r];
    load_slab(pf_slab, cur_pf, P - 2);
    advance(cur_pf);
    __syncthreads();  // front visible to the exchange wave
    __syncthreads();  // halo values written by it
    const int lane_dist = max(max(e_own0 - e, e - (e_own1 - 1)), 0);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int b = rb[r + 1];
      const bool row_own = b >= b_own0 && b < b_own1;
      const int row_dist = max(max(b_own0 - b, b - (b_own1 - 1)), 0);
      const bool need = rvalid[r + 1] && lane_valid && !(row_own && lane_own) && max(row_dist, lane_dist) <= K;
      if (need) Fd[r] = xfront[w * R + r][lane];
    }
  };
  auto exchange = [&](int blk, Slab<R, CT> &pf_slab) {
    // measurement only (xhist non-null): how long this wave spends in the
    // exchange, from its publications to the last halo value it waits for,
    // as a histogram of the device's constant-rate clock (log2 buckets)
    const unsigned long long t_in = xhist ? wall_clock64() : 0ull;
    const unsigned tag = tag_base + (unsigned)blk;
    const long long hslot = (long long)(blk & 1) * tile_slots;
    gu64 *pub = halo + (long long)lt * 2 * tile_slots + hslot;
    const int own_w = e_own1 - e_own0;
    const bool lane_band = (te > 0 && e - e_own0 < K) || (te < nTE - 1 && e_own1 - 1 - e < K);
    const int lane_dist = max(max(e_own0 - e, e - (e_own1 - 1)), 0);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int b = rb[r + 1];
      const bool row_band = (tb > 0 && b - b_own0 < K) || (tb < nTB - 1 && b_own1 - 1 - b < K);
      if (b >= b_own0 && b < b_own1 && lane_own && (row_band || lane_band))
        __hip_atomic_store(pub + (long long)(b - b_own0) * own_w + lane_pub, pack(Fd[r], tag), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    }
    // XSA: a run completed in this block (one-row tiles: the block) is
    // stored after the publications (which then leave first) and drains
    // while the wave waits to poll; the halo wait below also waits for it,
    // so every path leaves the exchange with the same loads in flight
    xa_flush();
    xr_flush();
    // Neighbours publish at about the same moment; polling right away mostly
    // finds their values not yet visible and costs a second round trip.
    // Waiting ~1300 clocks first was measured fastest for 32-row tiles
    // (16..28 x 64 clocks: 20 best by ~0.2 %, 28 0.5 % slower), ~770 for
    // 16-row tiles and 512 for one-wave row tiles.
    if constexpr (NW == 1) __builtin_amdgcn_s_sleep(kRowPollDelay);
    else if constexpr (NWR <= 16) __builtin_amdgcn_s_sleep(kHaloPollDelay16);
    else __builtin_amdgcn_s_sleep(kHaloPollDelay);
    // Only lanes that need a halo value load (measured faster than letting
    // every lane load some in-bounds line and selecting afterwards).
    const gu64 *src[R];
    bool need[R];
    unsigned long long x[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int b = rb[r + 1];
      const bool row_own = b >= b_own0 && b < b_own1;
      const int row_dist = max(max(b_own0 - b, b - (b_own1 - 1)), 0);
      need[r] = rvalid[r + 1] && lane_valid && !(row_own && lane_own) && max(row_dist, lane_dist) <= K;
      const int bq = min(max(b, 0), nB - 1);
      const int otb = bq / TB;  // uniform
      const gu64 *from = halo + (long long)(otb * nTE + ote) * 2 * tile_slots + hslot +
                         (long long)(bq - otb * TB) * ow + lane_src;
      src[r] = from;
      x[r] = 0;
      if (need[r]) x[r] = __hip_atomic_load(src[r], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    load_slab(pf_slab, cur_pf, P - 2);
    advance(cur_pf);
    // after an expiry this wave waits no more: the sweep then finishes
    // (with wrong values, reported through the status word) instead of
    // spending the full limit again at every later exchange
    int spins = wd_tripped ? spin_limit : 0;
    while (true) {
      bool waiting = false;
#pragma unroll
      for (int r = 0; r < R; ++r) waiting = waiting || (need[r] && (unsigned)(x[r] >> 32) != tag);
      if (!waiting) break;
      if (++spins > spin_limit) {
        __hip_atomic_store(status, FG_EWATCHDOG, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        wd_tripped = true;
        break;
      }
#pragma unroll
      for (int r = 0; r < R; ++r)
        if (need[r] && (unsigned)(x[r] >> 32) != tag)
          x[r] = __hip_atomic_load(src[r], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
#pragma unroll
    for (int r = 0; r < R; ++r)
      if (need[r]) Fd[r] = __uint_as_float((unsigned)x[r]);
    if (xhist && lane == 0) {
      const unsigned long long dt = wall_clock64() - t_in;
      const int bucket = dt ? min(64 - __clzll(dt), kXHistBuckets - 1) : 0;
      atomicAdd(&xhist_wg[bucket], 1u);  // LDS; added to the device histogram at the end
    }
  };

  // One slice: Sl holds the new slice (ring slot `slot`), Fr the previous
  // one (front image).  Afterwards Fr is refilled with the slice of step
  // s + P - 1.
  auto step = [&](Slab<R, CT> &Sl, Slab<R, CT> &Fr, int s, int slot, bool prefetch) {
    // steps past the end of the sweep (the last block is padded to K steps)
    // compute on slices of zeros and store nothing
    (void)s;
    // wave-boundary front rows through LDS, and the new slice's image rows
    // -1 and R (the front image of the next step) from the neighbour waves
    // (branch-free: the edge waves read their own slot and select +inf for
    // the distance; their outer image rows then only meet +inf candidates)
    float top = kInf, bot = kInf;
    if constexpr (NW > 1) {
      const int es = XSW ? (s & 1) : slot;
      edge_first[es][w][lane] = Fd[0];
      edge_last[es][w][lane] = Fd[R - 1];
#pragma unroll
      for (int c = 0; c < CT; ++c) {
        edgeI_first[es][w][c][lane] = Sl.I[c][1];
        edgeI_last[es][w][c][lane] = Sl.I[c][R];
      }
      __syncthreads();
      const float t = edge_last[es][max(w - 1, 0)][lane];
      const float u = edge_first[es][min(w + 1, NW - 1)][lane];
      top = (w > 0) ? t : kInf;
      bot = (w < NW - 1) ? u : kInf;
#pragma unroll
      for (int c = 0; c < CT; ++c) {
        Sl.I[c][0] = edgeI_last[es][max(w - 1, 0)][c][lane];
        Sl.I[c][R + 1] = edgeI_first[es][min(w + 1, NW - 1)][c][lane];
      }
      // the previous step's staged window is complete (barrier above)
      xflush((slot + P - 1) % P, s - 1);
    }
    float nd[R];
    if constexpr (ROW1) {
      float Iq[CT], In[CT];
#pragma unroll
      for (int c = 0; c < CT; ++c) {
        Iq[c] = Fr.I[c][1];
        In[c] = Sl.I[c][1];
      }
      const float v = relax_row<CT, LAMB1, FAST>(Fd[0], Iq, In, init_of(Sl.d[0]), e1, lg);
      nd[0] = pvalid[0] ? v : kInf;
    }
#pragma unroll
    for (int r = 0; r < (ROW1 ? 0 : R); ++r) {
      float dq[3];
      dq[0] = (r == 0) ? top : Fd[r - 1];
      dq[1] = Fd[r];
      dq[2] = (r == R - 1) ? bot : Fd[r + 1];
      float Iq[3][CT], In[CT];
#pragma unroll
      for (int c = 0; c < CT; ++c) {
        Iq[0][c] = Fr.I[c][r];
        Iq[1][c] = Fr.I[c][r + 1];
        Iq[2][c] = Fr.I[c][r + 2];
        In[c] = Sl.I[c][r + 1];
      }
      const float v = relax<CT, LAMB1, FAST>(dq, Iq, In, init_of(Sl.d[r]), e1, lg);
      nd[r] = pvalid[r] ? v : kInf;
    }
    // own outputs: raw buffer stores, non-own lanes and rows get an
    // out-of-range offset and the hardware drops their writes (no branch)
    if constexpr (XST) {
#pragma unroll
      for (int r = 0; r < R; ++r) xtile[slot][w * R + r][lane] = nd[r];
    } else if constexpr (XSW) {
      xa_stage(nd, cur_st.a);
    } else if constexpr (XSR) {
      xrr[slot] = nd[0];  // K == P: slot is the step's index in its block
    } else {
      // a padded step's slice is outside the array: all its stores are dropped
      const __amdgpu_buffer_rsrc_t rs = cursor_rsrc(Dd.d, cur_st);
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const unsigned off = (row_own[r] && lane_own) ? at(rofs[r + 1], cur_st) : nowhere(cur_st);
        __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(nd[r]), rs, (int)off, 0, 0);
      }
    }
#pragma unroll
    for (int r = 0; r < R; ++r) Fd[r] = nd[r];
    advance(cur_st);
    // prefetch without a branch: past the end of the sweep the loads return 0
    // and the data is never consumed (only steps that store nothing use it)
    if (prefetch) {
      load_slab(Fr, cur_pf, (slot + P - 1) % P);
      advance(cur_pf);
    }
  };

  // Blocks of exactly K steps (K a multiple of P, so a block starts at a
  // step s with s % P == 0 and slab p always sits in S[p]); the last block is
  // padded with steps that store nothing.  Every path into the group loop
  // and the block loop holds the ring in the same registers, so the waits
  // on prefetched slabs stay precise (no block boundary drains the loads).
  // The last step of a block defers its prefetch into the exchange.
  const int groups = K / P;
  const int nblocks = (nsteps + K - 1) / K;
  for (int blk = 0;; ++blk) {
    const int s0 = blk * K;
    xr_s0 = s0;
    for (int g = 0; g < groups - 1; ++g) {
#pragma unroll
      for (int p = 0; p < P; ++p) step(S[p], S[(p + P - 1) % P], s0 + g * P + p, p, true);
    }
    const int sl = s0 + K - P;
#pragma unroll
    for (int p = 0; p < P - 1; ++p) step(S[p], S[(p + P - 1) % P], sl + p, p, true);
    step(S[P - 1], S[P - 2], sl + P - 1, P - 1, false);
    if (blk + 1 >= nblocks) break;
    // deferred prefetch of step s0 + K - 1: the slice of step s0 + K + P - 2
    if constexpr (COMM) exchange_via_comm(blk + 1, S[P - 2]);
    else exchange(blk + 1, S[P - 2]);
  }
  if constexpr (XST) {
    __syncthreads();
    xflush(P - 1, nblocks * K - 1);  // the last step sits in slot P - 1
  }
  if constexpr (XSR) xr_flush();  // the last block
  if constexpr (XSW) {
    xa_flush();  // a run completed in the last block
    if (dir > 0 && ((nA - 1) & (kRunLen - 1)) != kRunLen - 1) {  // a forward sweep's partial last run
      xa_a0 = (nA - 1) & ~(kRunLen - 1);
      xa_flush();
    }
  }
  if (xhist) {
    __syncthreads();
    if (threadIdx.x < kXHistBuckets && xhist_wg[threadIdx.x]) atomicAdd(xhist + threadIdx.x, xhist_wg[threadIdx.x]);
  }
}

// ---------------------------------------------------------------------------
// Per-slice engine: one launch updates one slice from the previous one.
// Visits candidates in the reference order with the reference's min rule.
// ---------------------------------------------------------------------------

struct SliceDesc {
  const float *img;
  float *d;
  long long cs, sA, sB;
  int nB, nE, C;
  float lg;
  float e1[9];
};

__global__ void slice_kernel(SliceDesc D, int a, int aq) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long nplane = (long long)D.nB * D.nE;
  if (idx >= nplane) return;
  const int b = (int)(idx / D.nE), e = (int)(idx - (long long)b * D.nE);
  const long long p = (long long)a * D.sA + (long long)b * D.sB + e;
  float nd = D.d[p];
  for (int i = 0; i < 3; ++i) {
    const int bq = b + i - 1;
    if (bq < 0 || bq >= D.nB) continue;
    for (int j = 0; j < 3; ++j) {
      const int eq = e + j - 1;
      if (eq < 0 || eq >= D.nE) continue;
      const long long q = (long long)aq * D.sA + (long long)bq * D.sB + eq;
      float l = 0.0f;
      if (D.C == 1) {
        l = fabsf(D.img[p] - D.img[q]);
      } else {
        for (int c = 0; c < D.C; ++c) l += fabsf(D.img[c * D.cs + p] - D.img[c * D.cs + q]);
      }
      const float cur = (D.d[q] + D.e1[i * 3 + j]) + D.lg * l;
      nd = (cur < nd) ? cur : nd;
    }
  }
  D.d[p] = nd;
}

// ---------------------------------------------------------------------------
// Transposes (batched): out[c][w][h][d] = in[c][d][h][w]
// ---------------------------------------------------------------------------

constexpr int kXT = 64;

// Value of the combined sources at element k of the source layout.
__device__ __forceinline__ float xsrc(const XposeDesc &X, long long k) {
  const float a = X.src[k];
  if (X.kind == XK_SIGNED || X.kind == XK_GSF_MID) return a - X.src2[k];
  if (X.kind == XK_GSF_FINAL) return -(X.src2[k] - X.src3[k]) + (X.src4[k] - a);
  return a;
}

// Writes the (combined) value x to element k of the destination layout(s).
__device__ __forceinline__ void xdst(const XposeDesc &X, long long k, float x) {
  if (X.kind == XK_GSF_MID) {
    // x = Ds = D(M) - D(1-M); Md = Ds > theta, Me = Ds > -theta
    const float md = (x > X.theta) ? 1.0f : 0.0f;
    const float me = (x > -X.theta) ? 1.0f : 0.0f;
    const float m1 = 1.0f - md;  // mask of signed(image, 1 - Md)
    X.dst2[k] = X.v * m1;        // D(1-Md) init
    X.dst3[k] = X.v * (1.0f - m1);  // D(1-(1-Md)) init
    X.dst4[k] = X.v * me;        // D(Me) init
    X.dst[k] = X.v * (1.0f - me);   // D(1-Me) init
  } else {
    X.dst[k] = x;
  }
}

__global__ __launch_bounds__(256) void xpose_kernel(const XposeDesc *__restrict__ descs, int ndesc, XposeDesc one) {
  __shared__ float tile[kXT][kXT + 1];
  const int t = blockIdx.x;
  XposeDesc X;
  if (descs) {
    int lo = 0, hi = ndesc - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (descs[mid].tile0 <= t) lo = mid; else hi = mid - 1;
    }
    X = descs[lo];
  } else {
    X = one;
  }
  const int lt = t - X.tile0;
  const int plane = lt / X.tiles_dw, tt = lt - plane * X.tiles_dw;
  const int c = plane / X.H, h = plane - c * X.H;
  const int ntw = (X.W + kXT - 1) / kXT;
  const int td = tt / ntw, tw = tt - td * ntw;
  const int d0 = td * kXT, w0 = tw * kXT;
  const long long HW = (long long)X.H * X.W, HD = (long long)X.H * X.D;
  const long long cn = (long long)X.D * HW;
  const long long sbase = c * cn + (long long)h * X.W;  // element offsets of this (c, h) plane
  const long long dbase = c * cn + (long long)h * X.D;
  if (X.kind == XK_COPY && (X.W & 3) == 0 && (X.D & 3) == 0 && d0 + kXT <= X.D && w0 + kXT <= X.W &&
      (((uintptr_t)X.src | (uintptr_t)X.dst) & 15) == 0) {
    // full tile, 16-byte rows: each thread moves float4s (16 threads per
    // 256-byte row segment) and the LDS tile keeps its padded scalar layout
    const float *src = X.src + sbase;
    float *dst = X.dst + dbase;
    const int q = threadIdx.x & 15, r0 = threadIdx.x >> 4;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int r = r0 + 16 * k;
      const float4 v = *reinterpret_cast<const float4 *>(src + (long long)(d0 + r) * HW + w0 + 4 * q);
      tile[r][4 * q + 0] = v.x;
      tile[r][4 * q + 1] = v.y;
      tile[r][4 * q + 2] = v.z;
      tile[r][4 * q + 3] = v.w;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int r = r0 + 16 * k;  // output row (w)
      float4 v;
      v.x = tile[4 * q + 0][r];
      v.y = tile[4 * q + 1][r];
      v.z = tile[4 * q + 2][r];
      v.w = tile[4 * q + 3][r];
      *reinterpret_cast<float4 *>(dst + (long long)(w0 + r) * HD + d0 + 4 * q) = v;
    }
    return;
  }
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  for (int r = ty; r < kXT; r += 4) {
    const int d = d0 + r, w = w0 + tx;
    if (d < X.D && w < X.W) tile[r][tx] = xsrc(X, sbase + (long long)d * HW + w);
  }
  __syncthreads();
  for (int r = ty; r < kXT; r += 4) {
    const int w = w0 + r, d = d0 + tx;
    if (d < X.D && w < X.W) xdst(X, dbase + (long long)w * HD + d, tile[tx][r]);
  }
}

// ---------------------------------------------------------------------------
// Elementwise init / epilogue kernels (batched through small descriptors)
// ---------------------------------------------------------------------------

enum EwKind : int {
  EW_INIT_UNSIGNED = 0,  // out = v*m (all channels)
  EW_INIT_SIGNED = 1,    // out = v*m (all channels); a = v*(1-m) (channel 0)
  EW_FINAL_SIGNED = 2,   // out0 -= a;  out_c -= v*(1-m_c) for c >= 1
  EW_GSF_MID = 3,        // from (out0, a): thresholds -> inits of the 4 stage-2 transforms
  EW_GSF_FINAL = 4,      // out0 = -(a - b) + (c2 - out0); channels >= 1 in closed form
  EW_FILL_U64 = 5,       // zero a range of 64-bit words
  EW_CHANNELS = 6,       // mask channels >= 1 (elements [n0, n0*Cm)): the op's closed form of the mask
};

struct EwDesc {
  int kind;
  long long n0;   // voxels per channel
  int Cm;
  int op;         // EW_CHANNELS: FG_OP_*
  const float *mask;
  float *out;
  float *a, *b, *c2;
  float v, theta;
  unsigned long long *u64;
  long long nu64;
  long long blk0;   // first block of this descriptor in the launch
  long long nelem;
};

// Elements per ew_kernel block (256 threads x 4).
constexpr int kEwPerBlock = 1024;

__device__ __forceinline__ float signed_closed(float v, float m) {
  // untransformed channels of signed: v*m - v*(1-m)
  return v * m - v * (1.0f - m);
}

__device__ __forceinline__ void ew_one(const EwDesc &E, long long i);

// Each block handles kEwPerBlock consecutive elements of one descriptor (the
// block-uniform search runs on scalar loads once per block), each thread 4
// of them; the unsigned initialisation (every plain call) uses one 16-byte
// load and store when the 4 elements are aligned.
__global__ __launch_bounds__(256) void ew_kernel(const EwDesc *__restrict__ descs, int ndesc, EwDesc one) {
  const long long b = blockIdx.x;
  EwDesc E;
  if (descs) {
    int lo = 0, hi = ndesc - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (descs[mid].blk0 <= b) lo = mid; else hi = mid - 1;
    }
    E = descs[lo];
  } else {
    E = one;
  }
  const long long i0 = (b - E.blk0) * kEwPerBlock + (long long)threadIdx.x * 4;
  if (i0 >= E.nelem) return;
  if (E.kind == EW_INIT_UNSIGNED && i0 + 4 <= E.nelem &&
      (((uintptr_t)(E.out + i0) | (uintptr_t)(E.mask + i0)) & 15) == 0) {
    const float4 m = *reinterpret_cast<const float4 *>(E.mask + i0);
    float4 o;
    o.x = E.v * m.x;
    o.y = E.v * m.y;
    o.z = E.v * m.z;
    o.w = E.v * m.w;
    *reinterpret_cast<float4 *>(E.out + i0) = o;
    return;
  }
  for (int j = 0; j < 4 && i0 + j < E.nelem; ++j) ew_one(E, i0 + j);
}

__device__ __forceinline__ void ew_one(const EwDesc &E, long long i) {
  {
    const float v = E.v;
    switch (E.kind) {
      case EW_INIT_UNSIGNED:
        E.out[i] = v * E.mask[i];
        break;
      case EW_INIT_SIGNED: {
        const float m = E.mask[i];
        E.out[i] = v * m;
        if (i < E.n0) E.a[i] = v * (1.0f - m);
        break;
      }
      case EW_FINAL_SIGNED:
        if (i < E.n0) E.out[i] = E.out[i] - E.a[i];
        else E.out[i] = E.out[i] - v * (1.0f - E.mask[i]);
        break;
      case EW_GSF_MID: {
        // Ds = D(M) - D(1-M); Md = Ds > theta; Me = Ds > -theta
        const float ds = E.out[i] - E.a[i];
        const float md = (ds > E.theta) ? 1.0f : 0.0f;
        const float me = (ds > -E.theta) ? 1.0f : 0.0f;
        const float m1 = 1.0f - md;        // mask of signed(image, 1 - Md)
        E.a[i] = v * m1;                   // D(1-Md) init
        E.b[i] = v * (1.0f - m1);          // D(1-(1-Md)) init
        E.c2[i] = v * me;                  // D(Me) init
        E.out[i] = v * (1.0f - me);        // D(1-Me) init
        break;
      }
      case EW_GSF_FINAL:
        if (i < E.n0) {
          E.out[i] = -(E.a[i] - E.b[i]) + (E.c2[i] - E.out[i]);
        } else {
          const float m = E.mask[i];
          const float ds = signed_closed(v, m);
          const float md = (ds > E.theta) ? 1.0f : 0.0f;
          const float me = (ds > -E.theta) ? 1.0f : 0.0f;
          const float dd = -signed_closed(v, 1.0f - md);
          const float de = signed_closed(v, me);
          E.out[i] = dd + de;
        }
        break;
      case EW_FILL_U64:
        E.u64[i] = 0ull;
        break;
      case EW_CHANNELS: {
        // element n0 + i: an untransformed channel (reference: channels >= 1
        // keep their initialisation through every pass)
        const long long k = E.n0 + i;
        const float m = E.mask[k];
        if (E.op == FG_OP_UNSIGNED) {
          E.out[k] = v * m;
        } else if (E.op == FG_OP_SIGNED) {
          E.out[k] = signed_closed(v, m);
        } else {
          const float ds = signed_closed(v, m);
          const float md = (ds > E.theta) ? 1.0f : 0.0f;
          const float me = (ds > -E.theta) ? 1.0f : 0.0f;
          E.out[k] = -signed_closed(v, 1.0f - md) + signed_closed(v, me);
        }
        break;
      }
      default:
        break;
    }
  }
}

// Single-voxel writes (click edits of resident masks), passed by value in
// groups of up to kVoxelsPerLaunch.
constexpr int kVoxelsPerLaunch = 96;
struct VoxelBatch {
  fg_voxel_set it[kVoxelsPerLaunch];
};

__global__ __launch_bounds__(128) void set_voxels_kernel(VoxelBatch b, int n) {
  const int i = threadIdx.x;
  if (i < n) b.it[i].dst[b.it[i].index] = b.it[i].value;
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

thread_local std::string g_err;
thread_local fg_tuning g_tune = {1, 0, 0, 0};

// Requested halo depth: 0 = automatic (per plane), otherwise 1..kMaxK.
int current_halo() { return std::max(0, std::min(g_tune.halo, kMaxK)); }

// Per-launch timing of tiled sweeps (fg_profile_enable / fg_profile_collect).
struct ProfRec {
  hipEvent_t start, stop;
  double bytes;
};
thread_local bool g_prof_on = false;     // events around tiled launches
thread_local bool g_prof_xhist = false;  // exchange-wait histogram
// Checks of a pending halo value before a wait expires (0 = default); set by
// fg_set_wait_limit.
thread_local int g_wait_limit = 0;
thread_local std::vector<ProfRec> g_prof;

int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

#define FG_HIP(x)                                                                \
  do {                                                                           \
    hipError_t e_ = (x);                                                         \
    if (e_ != hipSuccess)                                                        \
      return fail(FG_EHIP, std::string(#x " failed: ") + hipGetErrorString(e_)); \
  } while (0)

// Step table of one pass: sqrt((s0 + |i-1|*s1) + |j-1|*s2) in float, in this
// summation order (reference fastgeodis_cpu.cpp:220-231).
void table3d(float s0, float s1, float s2, float ld[9]) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      float v = s0;
      v += (float)std::abs(i - 1) * s1;
      v += (float)std::abs(j - 1) * s2;
      ld[i * 3 + j] = std::sqrt(v);
    }
}

void table2d(float ld[9]) {  // {sqrt2, 1, sqrt2} (reference fastgeodis_cpu.cpp:67)
  const float r2 = std::sqrt(2.0f);
  const float row[3] = {r2, 1.0f, r2};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) ld[i * 3 + j] = row[j];
}

int ceil_div(long long a, long long b) { return (int)((a + b - 1) / b); }

// One unsigned transform to run.
struct Xform {
  const float *img;   // native (C, D, H, W) or (C, H, W)
  const float *imgT;  // transposed (C, W, H, D) or (C, W, H)
  float *d;           // native layout
  float *dT;          // transposed scratch
  int C, D, H, W;
  bool is2d;
  float lg;
  float e1[3][9];  // per pass: D, H, W (2D: [1] rows pass, [2] columns pass)
  bool lamb1, fast;
  int iters;
  int it = 0;  // the iteration being recorded
  bool xst;  // this iteration's H-backward sweep writes dT directly (no transpose before the W pass)
  bool xsa = false;  // this iteration's W-backward sweep writes d directly (no transpose after the W pass)
  // The first sweep initialises the distance from the mask (init_mode 1: v*m,
  // 2: v*(1-m)) instead of reading it; set only when that sweep is tiled.
  int init_mode = 0;
  const float *init_src = nullptr;
  float init_v = 0.f;
  // Transpose back after the last iteration: the first transform of a group
  // (signed pair, GSF stages) combines the group's distances (fin_kind) into
  // fin_dst; the other members (fin_member) contribute their dT only.
  int fin_kind = XK_COPY;
  bool fin_member = false;
  const float *fin_src[3] = {nullptr, nullptr, nullptr};
  float *fin_dst[4] = {nullptr, nullptr, nullptr, nullptr};
  float fin_v = 0.f, fin_theta = 0.f;
};

// The pass of a transform's first sweep: rows (1) in 2D, D (0) in 3D.
int first_pass(const Xform &x) { return x.is2d ? 1 : 0; }

// Tiled plan for one sweep plane.
struct Plan {
  int NW, R;
  bool comm;  // one more wave that performs the halo exchanges
  int P;  // slab ring size of the kernel that runs this plan
  int K;  // halo depth actually used: the requested depth rounded up to a multiple of P
  int nTB, nTE, TB, TE;
  bool ok;
};

struct Shape {
  int NW, R;
  int P12, P34;  // slab ring size for 1-2 / 3-4 image channels (prefetch distance P - 1)
  bool comm;     // plus kCommWaves waves that perform the halo exchanges
};
// Workgroup shapes with compiled kernels: index 0 serves 2D row sweeps
// (one row); the others tile 3D planes (selected by fg_tuning.shape).
constexpr Shape kShapes[] = {{1, 1, 6, 6, false},  {4, 8, 2, 2, false},  {8, 4, 3, 2, false},
                             {16, 2, 3, 3, false}, {8, 4, 5, 2, false},  {16, 2, 5, 3, false},
                             {8, 4, 5, 2, true},   {8, 2, 5, 3, false}};
constexpr int kNumShapes = 8;
// One-row planes at K = 30 (a lone row, or rows few enough for one tile per
// CU at that depth): the whole block in the slab ring, so the prefetches of
// a block are never waited for at a loop head.
constexpr Shape kRowDeep = {1, 1, 30, 30, false};
constexpr int kRowDeepK = 30;
constexpr int kAutoShape = 5;  // fastest on 512^2 planes: 16 x 2 with the deepest ring that keeps K = 5
// 16-row windows (8 x 2): with the automatic shape and halo, a pass whose planes (3D,
// 1-2 channels, general candidate: lamb != 1) all fit the compute units with
// these tiles at K = 5, and then occupy at least 1.5 x as many of them as
// with the 32-row windows, runs on them (see Runner::pass_shape): the
// 32-row windows leave most compute units idle on mid-size planes
// (128 x 256 x 256: 66-132 tiles) and compute 3.9 x their own tile at
// K = 10.  Measured on that volume (every pass on 16-row windows): 5.49 ->
// 4.68 ms (lamb 0.5), 5.52 -> 4.72 ms (lamb 0); lamb 1 unchanged
// (3.97 / 3.98 ms, bound by the loads rather than the arithmetic); a
// signed call with every pass on them, including the D pass whose two
// planes then exceed the device, and batches ran slower (4.18 -> 4.55 ms,
// 20.7 -> 27.2 ms).
constexpr int kSmallPlaneShape = 7;
thread_local int g_shape = kAutoShape;

// Sets this thread's shape for one pass and restores it on every way out.
struct ShapeScope {
  int saved;
  explicit ShapeScope(int shape) : saved(g_shape) { g_shape = shape; }
  ~ShapeScope() { g_shape = saved; }
};

Plan plan_fixed(int nB, int nE, int Kreq, int C, int shape = -1) {
  Plan p{};
  p.ok = false;
  const Shape s = (nB == 1) ? (Kreq > kRowDeepK - kShapes[0].P12 ? kRowDeep : kShapes[0])
                            : kShapes[shape >= 0 ? shape : g_shape];
  p.NW = s.NW;
  p.R = s.R;
  p.comm = s.comm && C <= 2;  // exchange-wave kernels exist for 1-2 channels; 3-4 run the plain 8 x 4
  const int P = C <= 2 ? s.P12 : s.P34;
  p.P = P;
  const int K = ((Kreq + P - 1) / P) * P;
  p.K = K;
  const int WR = s.NW * s.R;
  if (nE <= 62) {
    p.nTE = 1;
    p.TE = nE;
  } else {
    const int cap = 64 - 2 * K;
    if (cap < 1) return p;
    p.nTE = ceil_div(nE, cap);
    p.TE = ceil_div(nE, p.nTE);
  }
  if (nB <= WR) {
    p.nTB = 1;
    p.TB = nB;
  } else {
    const int cap = WR - 2 * K;
    if (cap < 1) return p;
    p.nTB = ceil_div(nB, cap);
    p.TB = ceil_div(nB, p.nTB);
  }
  p.ok = true;
  return p;
}

// Plan of one sweep plane.  Kreq > 0: that halo depth, rounded up to a
// multiple of the slab ring.  Kreq == 0 (automatic): the deepest halo
// (fewest exchanges) whose tiles still fit one per compute unit, as a tile
// shares its CU with no other and every tile waits for its neighbours at
// each exchange; if even the shallowest halo needs more tiles, that one.
Plan plan_plane(int nB, int nE, int Kreq, int C, int tile_budget) {
  if (Kreq > 0) return plan_fixed(nB, nE, Kreq, C);
  const Plan first = plan_fixed(nB, nE, 1, C);
  Plan best = first;
  for (int K = first.P; K <= kMaxK; K += first.P) {
    const Plan q = plan_fixed(nB, nE, K, C);
    if (!q.ok) break;
    if ((long long)q.nTB * q.nTE > tile_budget) break;
    best = q;
  }
  return best;
}

// (nB, nE) of each pass's sweep plane: 0 = D, 1 = H (2D: rows), 2 = W (2D: columns)
void pass_plane(const Xform &x, int pass, int &nA, int &nB, int &nE) {
  if (x.is2d) {
    nB = 1;
    if (pass == 1) { nA = x.H; nE = x.W; }
    else if (pass == 2) { nA = x.W; nE = x.H; }
    else { nA = 1; nE = x.W; }
    return;
  }
  if (pass == 0) { nA = x.D; nB = x.H; nE = x.W; }
  else if (pass == 1) { nA = x.H; nB = x.D; nE = x.W; }
  else { nA = x.W; nB = x.H; nE = x.D; }
}

template <int NW, int R, int P, int CT, bool L1, bool FA, int XS = 0, bool IN = false, bool CM = false>
const void *kernel_ptr() {
  return reinterpret_cast<const void *>(&sweep_tiled_kernel<NW, R, P, CT, L1, FA, XS, IN, CM>);
}

// lamb1 implies fast (l_grad == 1 and l_eucl*LD == +0 are non-negative).
// Only the channel counts a (shape, ring) pair serves are instantiated.
template <int NW, int R, int P, int CT, int XS = 0, bool IN = false, bool CM = false>
const void *pick_variant(bool lamb1, bool fast) {
  if (lamb1) return kernel_ptr<NW, R, P, CT, true, true, XS, IN, CM>();
  return fast ? kernel_ptr<NW, R, P, CT, false, true, XS, IN, CM>()
              : kernel_ptr<NW, R, P, CT, false, false, XS, IN, CM>();
}

template <int NW, int R, int P, int CT>
const void *pick_variant_init(bool lamb1, bool fast, bool init) {
  return init ? pick_variant<NW, R, P, CT, 0, true>(lamb1, fast) : pick_variant<NW, R, P, CT>(lamb1, fast);
}

// Kernels of the 8 x 4 shape with an exchange wave (ring 5, 1-2 channels).
const void *pick_comm_kernel(int C, bool lamb1, bool fast, int xs, bool init) {
  if (xs == 2) return nullptr;
  if (C == 1) {
    if (xs == 1) return pick_variant<8, 4, 5, 1, 1, false, true>(lamb1, fast);
    return init ? pick_variant<8, 4, 5, 1, 0, true, true>(lamb1, fast)
                : pick_variant<8, 4, 5, 1, 0, false, true>(lamb1, fast);
  }
  if (C == 2) {
    if (xs == 1) return pick_variant<8, 4, 5, 2, 1, false, true>(lamb1, fast);
    return init ? pick_variant<8, 4, 5, 2, 0, true, true>(lamb1, fast)
                : pick_variant<8, 4, 5, 2, 0, false, true>(lamb1, fast);
  }
  return nullptr;
}

template <int NW, int R, int P, int CLO, int CHI>
const void *pick_kernel_nr(int C, bool lamb1, bool fast, bool init) {
  if constexpr (CLO <= 1 && 1 <= CHI)
    if (C == 1) return pick_variant_init<NW, R, P, 1>(lamb1, fast, init);
  if constexpr (CLO <= 2 && 2 <= CHI)
    if (C == 2) return pick_variant_init<NW, R, P, 2>(lamb1, fast, init);
  if constexpr (CLO <= 3 && 3 <= CHI)
    if (C == 3) return pick_variant_init<NW, R, P, 3>(lamb1, fast, init);
  if constexpr (CLO <= 4 && 4 <= CHI)
    if (C == 4) return pick_variant_init<NW, R, P, 4>(lamb1, fast, init);
  return nullptr;
}

// One-row kernels with deferred stores into the other 2D layout on the ring
// P, 1-4 channels.
template <int P>
const void *pick_row_deferred(int C, bool lamb1, bool fast) {
  switch (C) {
    case 1: return pick_variant<1, 1, P, 1, 2>(lamb1, fast);
    case 2: return pick_variant<1, 1, P, 2, 2>(lamb1, fast);
    case 3: return pick_variant<1, 1, P, 3, 2>(lamb1, fast);
    case 4: return pick_variant<1, 1, P, 4, 2>(lamb1, fast);
    default: return nullptr;
  }
}

// The kernel compiled for exactly this (NW, R, P) and channel count, or null.
// xs selects the output form (0 own layout, 1 B contiguous, 2 A contiguous).
// B-contiguous kernels exist for the two 5-slice-ring 3D shapes: 8 x 4 with
// 1-2 channels, 16 x 2 with 1 channel (2 channels would exceed the LDS);
// A-contiguous ones for both shapes with 1 channel, and for one-row tiles
// (deferred stores) on the 30-slice ring with 1-4 channels.
constexpr int kXstMaxC = 2;
constexpr int kXsaMaxC = 1;
constexpr int kXsaMaxK = kRunSlots - kRunLen + 1;
const void *pick_kernel(int NW, int R, int P, int C, bool lamb1, bool fast, int xs = 0, bool init = false,
                        bool comm = false) {
  if (xs && init) return nullptr;
  if (xs == 2 && NW == 1 && R == 1) return (P == kRowDeepK && !comm) ? pick_row_deferred<kRowDeepK>(C, lamb1, fast) : nullptr;
  if (comm) return (NW == 8 && R == 4 && P == 5) ? pick_comm_kernel(C, lamb1, fast, xs, init) : nullptr;
  if (xs == 1) {
    if (NW == 8 && R == 4 && P == 5 && C == 1) return pick_variant<8, 4, 5, 1, 1>(lamb1, fast);
    if (NW == 8 && R == 4 && P == 5 && C == 2) return pick_variant<8, 4, 5, 2, 1>(lamb1, fast);
    if (NW == 16 && R == 2 && P == 5 && C == 1) return pick_variant<16, 2, 5, 1, 1>(lamb1, fast);
    return nullptr;
  }
  if (xs == 2) {
    if (NW == 8 && R == 4 && P == 5 && C == 1) return pick_variant<8, 4, 5, 1, 2>(lamb1, fast);
    if (NW == 16 && R == 2 && P == 5 && C == 1) return pick_variant<16, 2, 5, 1, 2>(lamb1, fast);
    return nullptr;
  }
#define FG_S(nw, r, p, clo, chi) \
  if (NW == nw && R == r && P == p && C >= clo && C <= chi) return pick_kernel_nr<nw, r, p, clo, chi>(C, lamb1, fast, init);
  FG_S(1, 1, 6, 1, 4)
  FG_S(1, 1, 30, 1, 4)
  FG_S(4, 8, 2, 1, 4)
  FG_S(8, 4, 3, 1, 2)
  FG_S(8, 4, 5, 1, 2)
  FG_S(16, 2, 5, 1, 2)
  FG_S(8, 4, 2, 3, 4)
  FG_S(16, 2, 3, 1, 4)
  FG_S(8, 2, 5, 1, 2)
  FG_S(8, 2, 3, 3, 4)
#undef FG_S
  return nullptr;
}

// Workspace layout helper: carves 256-byte aligned regions.
struct Carver {
  char *base;
  size_t cap, used = 0;
  bool dry;
  Carver(void *b, size_t c, bool dry_run) : base((char *)b), cap(c), dry(dry_run) {}
  void *take(size_t bytes) {
    used = (used + 255) & ~(size_t)255;
    void *p = dry ? nullptr : base + used;
    used += bytes;
    return p;
  }
  bool fits() const { return used <= cap; }
};

// Recorded launch.  Descriptor arrays live in one host blob that is copied to
// the workspace once per call, before any launch is issued.
struct Cmd {
  enum Type { TILED, SLICE, XPOSE, EW } type;
  const void *kern = nullptr;
  int grid = 0, block = 0;
  size_t off = 0;  // descriptor offset in the blob
  int ndesc = 0;
  int K = 0;
  unsigned tag_base = 0;
  SliceDesc slice{};  // SLICE: descriptor by value
  int a = 0, aq = 0;
  double alg_bytes = 0;  // TILED: algorithmic HBM bytes of the launch
  // launches of one descriptor pass it by value (no blob entry)
  bool by_value = false;
  SweepDesc sweep1{};
  XposeDesc xpose1{};
  EwDesc ew1{};
};

struct Recorder {
  std::vector<unsigned char> blob;
  std::vector<std::pair<std::string, size_t>> seen;  // identical arrays share storage
  std::vector<Cmd> cmds;

  size_t put(const void *p, size_t bytes) {
    std::string key((const char *)p, bytes);
    for (auto &s : seen)
      if (s.first == key) return s.second;
    size_t off = (blob.size() + 255) & ~(size_t)255;
    blob.resize(off + bytes);
    std::memcpy(blob.data() + off, p, bytes);
    seen.emplace_back(std::move(key), off);
    return off;
  }
};

struct Device {
  int id = -1;
  int cus = 0;
};

int query_device(Device &dev) {
  int id = 0;
  FG_HIP(hipGetDevice(&id));
  int cus = 0;
  FG_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, id));
  dev.id = id;
  dev.cus = cus;
  return FG_OK;
}

// The tiled sweep needs all workgroups of a launch resident at once.  Its
// grid never exceeds the occupancy limit (coresident_groups), so on its own
// every workgroup starts; but two such launches running together (calls on
// different streams) could each hold part of the device and wait on
// workgroups that cannot start.  Calls that launch tiled sweeps are therefore
// ordered per device: the host lock covers a whole replay, and a call on a
// different stream than the previous one first waits for that call's end.
// (Plain launches: the cooperative launch API adds ~15 us per launch.)
//
// The same per-device state holds the status word: one int in pinned,
// host-mapped, coherent host memory that every launch of the device may set
// (system-scope store) when a bounded wait expires or a launch is refused.
// Nothing clears it on the device; the next call on the device (or
// fg_status) reads and clears it from the host without synchronising, so an
// expiry is reported at the latest by the call after the one it happened in.
//
// Descriptor blobs reach the device through pinned staging slots (a ring per
// device): the copy is asynchronous, and a slot is reused only after the
// event recorded behind its last copy has completed.
struct StageSlot {
  unsigned char *host = nullptr;
  size_t cap = 0;
  hipEvent_t copied{};
  bool has_event = false;
};
constexpr int kStageSlots = 8;

// Scratch of the no-allocation entry point (fg_geodesic3d_f32_noalloc): one
// static device array per device, held from the library's load on that device
// until the process ends.  Bytes [0, 256) hold that entry point's status word
// while the device has no pinned status word; the rest is its workspace.
// Host code uses the address from hipGetSymbolAddress, never the name.
constexpr size_t kNoAllocScratchBytes = size_t(256) << 20;
constexpr size_t kNoAllocStatusBytes = 256;

struct DeviceState {
  std::mutex m;  // held for a whole replay
  bool has = false;
  hipStream_t last{};
  hipEvent_t done{};
  int *status_host = nullptr;  // pinned word written by kernels
  int *status_dev = nullptr;   // its device address
  std::atomic<int> injected{0};  // set from the host (fg_inject_status)
  StageSlot stage[kStageSlots];
  int next_slot = 0;
  unsigned char *scratch = nullptr;  // device address of the no-allocation scratch (per device)
};

std::mutex g_reg_m;
std::map<int, std::unique_ptr<DeviceState>> g_reg;

DeviceState &device_state(int id) {
  std::lock_guard<std::mutex> g(g_reg_m);
  auto &slot = g_reg[id];
  if (!slot) slot.reset(new DeviceState());
  return *slot;
}

// Allocates the device's status word on first use (the device is current).
int ensure_status_word(DeviceState &s) {
  std::lock_guard<std::mutex> g(s.m);
  if (s.status_host) return FG_OK;
  void *h = nullptr;
  FG_HIP(hipHostMalloc(&h, 256, hipHostMallocMapped | hipHostMallocCoherent));
  std::memset(h, 0, 256);
  void *d = nullptr;
  hipError_t e = hipHostGetDevicePointer(&d, h, 0);
  if (e != hipSuccess) {
    (void)hipHostFree(h);
    return fail(FG_EHIP, std::string("hipHostGetDevicePointer: ") + hipGetErrorString(e));
  }
  s.status_host = (int *)h;
  s.status_dev = (int *)d;
  return FG_OK;
}

// Reads and clears the device's pending status (no synchronisation).
int take_status(DeviceState &s) {
  int v = s.injected.exchange(0);
  if (s.status_host) {
    const int w = __atomic_exchange_n(s.status_host, 0, __ATOMIC_SEQ_CST);
    if (v == 0) v = w;
  }
  return v;
}

int status_error(int st, const char *when) {
  if (st == FG_OK) return FG_OK;
  if (st == FG_EWATCHDOG)
    return fail(FG_EWATCHDOG, std::string(when) +
                                  ": a tile's bounded halo wait expired (another persistent kernel may have held "
                                  "compute units); the results of that call are invalid");
  if (st == FG_EINVAL)
    return fail(FG_EHIP, std::string(when) +
                             ": a sweep launch was refused (halo depth not a multiple of its ring, a one-row "
                             "kernel on a plane of several rows, or a native-layout output on a plane of one "
                             "tile row)");
  return fail(FG_EHIP, std::string(when) + ": device status " + std::to_string(st));
}

// Holds a device's order for one replay; on every way out (errors included)
// marks the end of what the replay enqueued on its stream.
struct OrderedReplay {
  DeviceState *ord = nullptr;
  hipStream_t stream{};
  std::unique_lock<std::mutex> lock;
  // order: the replay contains persistent sweeps and joins the device order;
  // create: the order's event may be created here (not for callers that
  // allocate nothing, which then join only an existing order)
  int begin(DeviceState &o, hipStream_t s, bool order, bool create = true) {
    lock = std::unique_lock<std::mutex>(o.m);
    if (!order || (!o.has && !create)) return FG_OK;
    if (!o.has) {
      FG_HIP(hipEventCreateWithFlags(&o.done, hipEventDisableTiming));
      o.has = true;
    } else if (o.last != s) {
      FG_HIP(hipStreamWaitEvent(s, o.done, 0));
    }
    o.last = s;
    ord = &o;
    stream = s;
    return FG_OK;
  }
  ~OrderedReplay() {
    if (ord) (void)hipEventRecord(ord->done, stream);
  }
};

// Copies `bytes` of `src` to `dst` on `stream` through a pinned staging slot
// of the device (the device's lock is held by the caller).
int stage_copy(DeviceState &s, void *dst, const void *src, size_t bytes, hipStream_t stream) {
  StageSlot &sl = s.stage[s.next_slot];
  s.next_slot = (s.next_slot + 1) % kStageSlots;
  if (sl.has_event) FG_HIP(hipEventSynchronize(sl.copied));
  if (sl.cap < bytes) {
    if (sl.host) FG_HIP(hipHostFree(sl.host));
    sl.host = nullptr;
    sl.cap = 0;
    const size_t cap = std::max<size_t>(bytes, 64 << 10);
    void *h = nullptr;
    FG_HIP(hipHostMalloc(&h, cap, hipHostMallocDefault));
    sl.host = (unsigned char *)h;
    sl.cap = cap;
  }
  if (!sl.has_event) {
    FG_HIP(hipEventCreateWithFlags(&sl.copied, hipEventDisableTiming));
    sl.has_event = true;
  }
  std::memcpy(sl.host, src, bytes);
  FG_HIP(hipMemcpyAsync(dst, sl.host, bytes, hipMemcpyHostToDevice, stream));
  FG_HIP(hipEventRecord(sl.copied, stream));
  return FG_OK;
}

// Occupancy-limited number of co-resident workgroups per kernel (cached per
// host thread and device).
int coresident_groups(const void *kern, int threads, const Device &dev, long long &cap) {
  struct Entry { const void *k; int dev; long long cap; };
  thread_local std::vector<Entry> cache;
  for (auto &e : cache)
    if (e.k == kern && e.dev == dev.id) { cap = e.cap; return FG_OK; }
  int per_cu = 0;
  FG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, threads, 0));
  cap = (long long)per_cu * dev.cus;
  cache.push_back({kern, dev.id, cap});
  return FG_OK;
}

// Address of the exchange-wait histogram on device `dev` (the device is current).
int exchange_histogram_impl(int dev, unsigned **out);
inline int exchange_histogram(int dev, unsigned **out) { return exchange_histogram_impl(dev, out); }

// Executable graphs of repeated launch sequences.  A replay whose sequence of
// launches (kernel, grid and block of each, in order) repeats on a host
// thread and device runs as one hipGraphLaunch of a cached graph whose kernel
// nodes (a chain) get this call's arguments; the first occurrence of a
// sequence launches directly, the second builds the graph.  Graphs hold
// kernel nodes only (no memset or copy nodes).
thread_local bool g_graph_on = true;

struct GraphEntry {
  std::thread::id thread;
  uint64_t key = 0;
  int dev = -1;
  hipGraph_t graph{};
  hipGraphExec_t exec{};
  std::vector<hipGraphNode_t> nodes;
  uint64_t last_use = 0;
};
std::mutex g_graph_m;
std::vector<std::unique_ptr<GraphEntry>> g_graph_list;  // every thread's entries
thread_local std::vector<std::pair<uint64_t, int>> g_graph_seen;  // sequences seen once (key, device)
thread_local uint64_t g_graph_clock = 0;
constexpr int kGraphsPerThread = 8;

void destroy_graph(GraphEntry &e) {
  if (e.exec) (void)hipGraphExecDestroy(e.exec);
  if (e.graph) (void)hipGraphDestroy(e.graph);
  e.exec = nullptr;
  e.graph = nullptr;
}

struct SweepJob {
  Xform *x;
  int pass;
  int K;  // halo depth of this plane's plan (> 0)
};

struct Runner {
  Device dev;
  int K;
  int engine;
  int max_groups;
  size_t halo_cap;  // entries
  unsigned tag_base = 0;
  Recorder rec;

  bool no_graph = false;  // replay the launches directly (sequences that do not repeat)

  void fill_desc(const Xform &x, int pass, int dir, SweepDesc &s, int xs = 0) {
    std::memset(&s, 0, sizeof s);
    if (x.init_mode && x.it == 0 && pass == first_pass(x) && dir > 0) {
      s.init_src = x.init_src;
      s.init_v = x.init_v;
      s.init_mode = x.init_mode;
    }
    const long long N = (long long)x.D * x.H * x.W;
    s.C = x.C;
    s.cs = N;
    s.dir = dir;
    s.lg = x.lg;
    pass_plane(x, pass, s.nA, s.nB, s.nE);
    if (x.is2d) {
      if (pass == 2) { s.img = x.imgT; s.d = x.dT; s.sA = x.H; }   // transposed (W, H)
      else { s.img = x.img; s.d = x.d; s.sA = x.W; }                 // native (H, W)
      s.sB = 0;
    } else if (pass == 0) {
      s.img = x.img; s.d = x.d; s.sA = (long long)x.H * x.W; s.sB = x.W;
    } else if (pass == 1) {
      s.img = x.img; s.d = x.d; s.sA = x.W; s.sB = (long long)x.H * x.W;
    } else {
      s.img = x.imgT; s.d = x.dT; s.sA = (long long)x.H * x.D; s.sB = x.D;
    }
    for (int k = 0; k < 9; ++k) s.e1[k] = x.e1[pass][k];
    if (xs == 1) {  // H pass (A = H, B = D, E = W) written as (W, H, D)
      s.dxo = x.dT;
      s.sXe = (long long)s.nA * s.nB;
      s.sXa = s.nB;
    } else if (xs == 2 && x.is2d) {  // rows (A = H, E = W) written as (W, H), columns as (H, W)
      s.dxo = pass == 1 ? x.dT : x.d;
      s.sXe = s.nA;
      s.sXa = 1;
      s.sXb = 0;
    } else if (xs == 2) {  // W pass (A = W, B = H, E = D) written as (D, H, W)
      s.dxo = x.d;
      s.sXe = (long long)x.H * x.W;
      s.sXb = x.W;
    }
  }

  // Whether the H-backward sweep of this transform can write the transposed
  // layout itself: 3D, tiled engine with the default shape and a
  // transposed-output kernel for its channel count, and every output offset
  // within 32 bits.
  bool xstore_candidate(const Xform &x) {
    if (x.is2d) return row_xsa_candidate(x, 1);
    if (engine == 0 || x.C > kXstMaxC) return false;
    int nA, nB, nE;
    pass_plane(x, 1, nA, nB, nE);
    if (nA <= 1 || nB <= 1) return false;
    const Plan p = plan_plane(nB, nE, K, x.C, dev.cus);
    if (!p.ok || !pick_kernel(p.NW, p.R, p.P, x.C, x.lamb1, x.fast, 1, false, p.comm)) return false;
    if ((long long)x.D * x.H * x.W * 4 >= 0xFFFFFFF0ll) return false;
    return slice_offsets_fit(x, 1);
  }

  // Whether the W-backward sweep of this transform may write the native
  // layout itself (the launch then also needs a plan of several tile rows
  // and an A-contiguous kernel for its shape, checked when it is recorded):
  // 3D, tiled engine, one channel, output offsets within 32 bits, and not
  // the last iteration of a transform whose result leaves through a
  // combining transpose (signed and GSF epilogues).
  bool xsa_candidate(const Xform &x) {
    if (x.it == x.iters - 1 && (x.fin_member || x.fin_kind != XK_COPY)) return false;
    if (x.is2d) return row_xsa_candidate(x, 2);
    if (engine == 0 || x.C > kXsaMaxC) return false;
    int nA, nB, nE;
    pass_plane(x, 2, nA, nB, nE);
    if (nA <= 1 || nB <= 1) return false;
    if ((long long)x.D * x.H * x.W * 4 >= 0xFFFFFFF0ll) return false;
    return slice_offsets_fit(x, 2);
  }

  // 2D: whether the backward sweep of rows (pass 1) may write the column
  // layout, or that of columns (pass 2) the native layout, itself; the
  // launch then also needs the one-row kernel whose block is one turn of its
  // 30-slice ring (checked when it is recorded).  The output, like the input,
  // is addressed with 32-bit offsets within 2^31 bytes.
  bool row_xsa_candidate(const Xform &x, int pass) {
    if (engine == 0 || x.C > 4) return false;
    int nA, nB, nE;
    pass_plane(x, pass, nA, nB, nE);
    if (nA <= 1) return false;
    return slice_offsets_fit(x, 1) && slice_offsets_fit(x, 2);
  }

  // The tiled kernel addresses elements of a slice with 32-bit byte offsets
  // from the slice start (buffer resources with a range of 0xFFFFFFF0 bytes);
  // the one-row kernel addresses the whole array with them, within 2^31
  // bytes including 32 slices of margin past its end.
  bool slice_offsets_fit(const Xform &x, int pass) {
    SweepDesc s;
    fill_desc(x, pass, 1, s);
    const long long span = ((long long)(s.nB - 1) * s.sB + s.nE) * 4;
    if (s.nB == 1 && (long long)(s.nA + 32) * s.sA * 4 > (1ll << 31)) return false;
    return span < 0xFFFFFFF0ll;
  }

  void record_per_slice(const Xform &x, int pass, int dir) {
    SweepDesc s;
    fill_desc(x, pass, dir, s);
    Cmd c;
    c.type = Cmd::SLICE;
    SliceDesc &sd = c.slice;
    sd.img = s.img; sd.d = s.d; sd.cs = s.cs; sd.sA = s.sA; sd.sB = s.sB;
    sd.nB = s.nB; sd.nE = s.nE; sd.C = s.C; sd.lg = s.lg;
    for (int k = 0; k < 9; ++k) sd.e1[k] = s.e1[k];
    c.grid = ceil_div((long long)s.nB * s.nE, 256);
    c.block = 256;
    if (dir > 0) {
      for (int a = 1; a < s.nA; ++a) { c.a = a; c.aq = a - 1; rec.cmds.push_back(c); }
    } else {
      for (int a = s.nA - 2; a >= 0; --a) { c.a = a; c.aq = a + 1; rec.cmds.push_back(c); }
    }
  }

  int record_tiled(std::vector<SweepJob> &jobs, int dir, int NW, int R, int P, int C, bool lamb1, bool fast,
                   int xs = 0, bool in = false, bool cm = false) {
    std::vector<SweepDesc> descs;
    long long halo_used = 0;
    int tiles = 0, max_blocks = 0, kk = -1;
    double alg = 0;
    for (auto &j : jobs) {
      SweepDesc s;
      fill_desc(*j.x, j.pass, dir, s, xs);
      alg += (double)(s.nA - 1) * s.nB * s.nE * (4.0 * s.C + 8.0);
      Plan p = plan_fixed(s.nB, s.nE, j.K, s.C);
      if (!p.ok || p.NW != NW || p.R != R || p.P != P || p.comm != cm || p.K % P != 0 || (kk >= 0 && kk != p.K))
        return fail(FG_EINVAL, "internal: sweep plan does not match the selected kernel");
      s.nTB = p.nTB; s.nTE = p.nTE; s.TB = p.TB; s.TE = p.TE;
      s.tile0 = tiles;
      s.halo0 = halo_used;
      const int nt = p.nTB * p.nTE;
      tiles += nt;
      halo_used += (long long)nt * 2 * p.TB * p.TE;
      kk = p.K;
      max_blocks = std::max(max_blocks, (s.nA - 1) / p.K + 1);
      descs.push_back(s);
    }
    if (descs.empty()) return FG_OK;
    if ((size_t)halo_used > halo_cap) return fail(FG_EWORKSPACE, "workspace too small for the sweep plan");
    if ((unsigned long long)tag_base + max_blocks + 2 >= 0x7fffffffull) {
      record_zero_halo();
      tag_base = 0;
    }
    Cmd c;
    c.type = Cmd::TILED;
    c.kern = pick_kernel(NW, R, P, C, lamb1, fast, xs, in, cm);
    if (!c.kern) return fail(FG_EINVAL, "no tiled kernel for this shape");
    c.grid = tiles;
    c.block = (NW + (cm ? kCommWaves : 0)) * 64;
    c.ndesc = (int)descs.size();
    if (c.ndesc == 1) {
      c.by_value = true;
      c.sweep1 = descs[0];
    } else {
      c.off = rec.put(descs.data(), descs.size() * sizeof(SweepDesc));
    }
    c.K = kk;
    c.tag_base = tag_base;
    c.alg_bytes = alg;
    rec.cmds.push_back(c);
    tag_base += (unsigned)max_blocks + 1;
    return FG_OK;
  }

  void record_zero_halo() {
    if (halo_cap == 0) return;
    EwDesc e{};
    e.kind = EW_FILL_U64;
    e.u64 = nullptr;  // patched at replay: the halo base
    e.nelem = (long long)halo_cap;
    record_ew({e});
  }

  void record_ew(const std::vector<EwDesc> &in) {
    std::vector<EwDesc> ds = in;
    long long blocks = 0;
    for (auto &e : ds) {
      e.blk0 = blocks;
      blocks += ceil_div(e.nelem, (long long)kEwPerBlock);
    }
    if (blocks == 0) return;
    Cmd c;
    c.type = Cmd::EW;
    c.ndesc = (int)ds.size();
    if (c.ndesc == 1) {
      c.by_value = true;
      c.ew1 = ds[0];
    } else {
      c.off = rec.put(ds.data(), ds.size() * sizeof(EwDesc));
    }
    c.block = 256;
    c.grid = (int)blocks;  // < 2^31: 2^41 elements would not fit any device
    rec.cmds.push_back(c);
  }

  void record_xpose(const std::vector<XposeDesc> &in) {
    std::vector<XposeDesc> ds = in;
    int tiles = 0;
    for (auto &x : ds) {
      x.tiles_dw = ceil_div(x.D, kXT) * ceil_div(x.W, kXT);
      x.tile0 = tiles;
      tiles += x.C * x.H * x.tiles_dw;
    }
    if (tiles == 0) return;
    Cmd c;
    c.type = Cmd::XPOSE;
    c.ndesc = (int)ds.size();
    if (c.ndesc == 1) {
      c.by_value = true;
      c.xpose1 = ds[0];
    } else {
      c.off = rec.put(ds.data(), ds.size() * sizeof(XposeDesc));
    }
    c.grid = tiles;
    c.block = 256;
    rec.cmds.push_back(c);
  }

  static XposeDesc xd(const float *src, float *dst, int C, int D, int H, int W) {
    XposeDesc x{};
    x.src = src; x.dst = dst; x.C = C; x.D = D; x.H = H; x.W = W;
    return x;
  }
  static void native_dims(const Xform &x, int &D, int &H, int &W) {
    if (x.is2d) { D = x.H; H = 1; W = x.W; }  // (H, W) seen as (D'=H, H'=1, W'=W)
    else { D = x.D; H = x.H; W = x.W; }
  }

  // Whether a transform's sweeps of this pass run on the tiled engine.
  bool tiled_eligible(const Xform &x, int pass) {
    int nA, nB, nE;
    pass_plane(x, pass, nA, nB, nE);
    if (nA <= 1 || engine == 0 || x.C > 4) return false;
    return plan_plane(nB, nE, K, x.C, dev.cus).ok && slice_offsets_fit(x, pass);
  }

  // Halo depth of each transform's plane in this pass.  Requested depths
  // apply as they are.  Automatic: a lone plane takes the deepest halo whose
  // tiles fit one per CU (idle CUs cost nothing, and fewer exchanges shorten
  // the sweep).  Planes that share the pass (batches, the halves of a signed
  // or GSF call) are planned together: the deepest halo at which all their
  // tiles still fit the device, and the shallowest one when none does, as a
  // deeper halo then only adds window arithmetic (a 32 x 64 window computes
  // 1.7 x its own tile at K = 5 and 3.9 x at K = 10; a 64-lane row 1.23 x at
  // K = 6 and 16 x at K = 30).  3D tiles fill a CU each; 2D rows are one-wave
  // tiles bound by latency, so they fit one per SIMD.
  std::vector<int> pass_halos(const std::vector<Xform *> &active, int pass) {
    std::vector<int> ks(active.size(), 0);
    std::vector<int> rows, planes;  // automatic 2D rows and 3D planes
    for (size_t i = 0; i < active.size(); ++i) {
      int nA, nB, nE;
      pass_plane(*active[i], pass, nA, nB, nE);
      if (!tiled_eligible(*active[i], pass)) continue;
      ks[i] = plan_plane(nB, nE, K, active[i]->C, dev.cus).K;
      if (K == 0) (nB > 1 ? planes : rows).push_back((int)i);
    }
    if (rows.size() > 1) share_halo(active, pass, rows, 4LL * dev.cus, ks);
    if (planes.size() > 1) share_halo(active, pass, planes, dev.cus, ks);
    return ks;
  }

  // One halo depth per slab-ring length for the planes `idx` (channel
  // classes use different rings): the deepest whose tiles total at most
  // `budget`, else the shallowest.
  void share_halo(const std::vector<Xform *> &active, int pass, const std::vector<int> &idx, long long budget,
                  std::vector<int> &ks) {
    auto ring = [&](int i) {
      int nA, nB, nE;
      pass_plane(*active[i], pass, nA, nB, nE);
      return plan_fixed(nB, nE, 1, active[i]->C).P;
    };
    std::vector<int> rings;
    for (int i : idx)
      if (std::find(rings.begin(), rings.end(), ring(i)) == rings.end()) rings.push_back(ring(i));
    for (int P : rings) {
      int best = P;
      for (int k = P; k <= kMaxK; k += P) {
        long long tiles = 0;
        bool ok = true;
        for (int i : idx) {
          if (ring(i) != P) continue;
          int nA, nB, nE;
          pass_plane(*active[i], pass, nA, nB, nE);
          const Plan q = plan_fixed(nB, nE, k, active[i]->C);
          if (!q.ok) { ok = false; break; }
          tiles += (long long)q.nTB * q.nTE;
        }
        if (!ok || tiles > budget) break;
        best = k;
      }
      for (int i : idx)
        if (ring(i) == P) ks[i] = best;
    }
  }

  // The 3D workgroup shape of a pass: the tuned one, or with the automatic
  // shape the 16-row windows when the pass's planes qualify (kSmallPlaneShape).
  int pass_shape(const std::vector<Xform *> &active, int pass) {
    if (g_tune.shape != 0 || K != 0 || engine == 0) return g_shape;  // tuned shape or halo
    const std::vector<int> ks = pass_halos(active, pass);  // with the 32-row windows
    long long t32 = 0, t16 = 0;
    for (size_t i = 0; i < active.size(); ++i) {
      const Xform *x = active[i];
      int nA, nB, nE;
      pass_plane(*x, pass, nA, nB, nE);
      if (nA <= 1) continue;
      if (x->is2d || nB <= 1 || x->C > 2 || x->lamb1 || ks[i] == 0) return g_shape;
      const Plan a = plan_fixed(nB, nE, ks[i], x->C);
      const Plan b = plan_fixed(nB, nE, 5, x->C, kSmallPlaneShape);
      if (!a.ok || !b.ok) return g_shape;
      t32 += (long long)a.nTB * a.nTE;
      t16 += (long long)b.nTB * b.nTE;
    }
    return (t16 > 0 && t16 <= dev.cus && 2 * t16 >= 3 * t32) ? kSmallPlaneShape : g_shape;
  }

  // Record one pass (forward then backward) for the active transforms.
  int sweep_pass(std::vector<Xform *> &active, int pass) {
    ShapeScope shape_scope(pass_shape(active, pass));
    const std::vector<int> ks = pass_halos(active, pass);
    for (int dir : {+1, -1}) {
      struct G { int NW, R, P, K, C; bool l1, fa; int xs; bool in, cm; std::vector<SweepJob> jobs; };
      std::vector<G> groups;
      for (size_t i = 0; i < active.size(); ++i) {
        Xform *x = active[i];
        int nA, nB, nE;
        pass_plane(*x, pass, nA, nB, nE);
        if (nA <= 1) continue;
        const bool in = x->init_mode && x->it == 0 && pass == first_pass(*x) && dir > 0;
        if (ks[i] == 0) {
          if (in) return fail(FG_EINVAL, "internal: an initialising sweep was planned on the per-slice engine");
          record_per_slice(*x, pass, dir);
          if (pass == 1) x->xst = false;  // its output stays in the native layout
          if (pass == 2) x->xsa = false;  // its output stays in the W pass layout
          continue;
        }
        const Plan p = plan_fixed(nB, nE, ks[i], x->C);
        // Backward sweeps that write the next layout themselves: 3D tiles
        // stage the own rows of a tile of several along B, and the slices of
        // one block beyond a run; 2D one-row tiles defer a block's stores to
        // the exchange, on the 30-slice ring with one ring turn per block.
        // Otherwise the distance moves between layouts through a transpose.
        const bool row_other = x->is2d && p.NW == 1 && p.R == 1 && p.P == kRowDeepK && p.K == p.P &&
                               pick_kernel(p.NW, p.R, p.P, x->C, x->lamb1, x->fast, 2, in, p.comm);
        int xs = 0;
        if (pass == 1 && dir < 0 && x->xst) {
          // (the pass may run on a shape without a B-contiguous kernel)
          if (!x->is2d && pick_kernel(p.NW, p.R, p.P, x->C, x->lamb1, x->fast, 1, in, p.comm)) xs = 1;
          else if (x->is2d && row_other) xs = 2;
          else x->xst = false;
        }
        if (pass == 2 && dir < 0 && x->xsa) {
          if (x->is2d ? row_other
                      : (p.nTB > 1 && p.K <= kXsaMaxK &&
                         pick_kernel(p.NW, p.R, p.P, x->C, x->lamb1, x->fast, 2, in, p.comm)))
            xs = 2;
          else
            x->xsa = false;
        }
        bool found = false;
        for (auto &g : groups)
          if (g.NW == p.NW && g.R == p.R && g.P == p.P && g.K == p.K && g.C == x->C && g.l1 == x->lamb1 &&
              g.fa == x->fast && g.xs == xs && g.in == in && g.cm == p.comm) {
            g.jobs.push_back({x, pass, p.K});
            found = true;
            break;
          }
        if (!found) groups.push_back({p.NW, p.R, p.P, p.K, x->C, x->lamb1, x->fast, xs, in, p.comm, {{x, pass, p.K}}});
      }
      for (auto &g : groups) {
        int rc = record_group(g.jobs, dir, g.NW, g.R, g.P, g.C, g.l1, g.fa, g.xs, g.in, g.cm);
        if (rc) return rc;
      }
    }
    return FG_OK;
  }

  // A group is split into launches whose tiles can all be co-resident.
  int record_group(std::vector<SweepJob> &jobs, int dir, int NW, int R, int P, int C, bool l1, bool fa,
                   int xs = 0, bool in = false, bool cm = false) {
    const void *kern = pick_kernel(NW, R, P, C, l1, fa, xs, in, cm);
    if (!kern) return fail(FG_EINVAL, "no tiled kernel for this shape");
    long long cap = 0;
    int rc = coresident_groups(kern, (NW + (cm ? kCommWaves : 0)) * 64, dev, cap);
    if (rc) return rc;
    if (max_groups > 0) cap = std::min<long long>(cap, max_groups);
    // A launch lasts as long as its longest sweep: pack planes of similar
    // step counts together (longest first).
    std::stable_sort(jobs.begin(), jobs.end(), [](const SweepJob &u, const SweepJob &v) {
      int au, bu, eu, av, bv, ev;
      pass_plane(*u.x, u.pass, au, bu, eu);
      pass_plane(*v.x, v.pass, av, bv, ev);
      return au > av;
    });
    std::vector<SweepJob> chunk;
    long long tiles = 0;
    for (auto &j : jobs) {
      int nA, nB, nE;
      pass_plane(*j.x, j.pass, nA, nB, nE);
      Plan p = plan_fixed(nB, nE, j.K, j.x->C);
      const long long nt = (long long)p.nTB * p.nTE;
      if (nt > cap) {  // one plane alone cannot be co-resident: per-slice engine
        if (in) return fail(FG_EINVAL, "internal: an initialising sweep was planned on the per-slice engine");
        record_per_slice(*j.x, j.pass, dir);
        if (j.pass == 1) j.x->xst = false;  // its output stays in the native layout
        if (j.pass == 2) j.x->xsa = false;  // its output stays in the W pass layout
        continue;
      }
      if (tiles + nt > cap) {
        if ((rc = record_tiled(chunk, dir, NW, R, P, C, l1, fa, xs, in, cm))) return rc;
        chunk.clear();
        tiles = 0;
      }
      chunk.push_back(j);
      tiles += nt;
    }
    if (!chunk.empty()) return record_tiled(chunk, dir, NW, R, P, C, l1, fa, xs, in, cm);
    return FG_OK;
  }

  // The transposes around a W pass: into its layout (unless the H pass's
  // backward sweep wrote it) and, once the W pass is recorded, back (unless
  // its backward sweep wrote the native layout), the last iteration's
  // carrying the group's epilogue (members of a group contribute through its
  // first one).
  void w_pass_to(const std::vector<Xform *> &active, std::vector<XposeDesc> &to) {
    for (auto *x : active) {
      int D, H, W;
      native_dims(*x, D, H, W);
      if (!x->xst) to.push_back(xd(x->d, x->dT, 1, D, H, W));
    }
  }
  void w_pass_back(const std::vector<Xform *> &active, std::vector<XposeDesc> &back) {
    for (auto *x : active) {
      if (x->xsa) continue;
      int D, H, W;
      native_dims(*x, D, H, W);
      const bool last = x->it == x->iters - 1;
      if (last && x->fin_member) continue;  // part of its group's combined transpose
      XposeDesc b = xd(x->dT, x->d, 1, W, H, D);
      if (last && x->fin_kind != XK_COPY) {
        b.kind = x->fin_kind;
        b.src2 = x->fin_src[0];
        b.src3 = x->fin_src[1];
        b.src4 = x->fin_src[2];
        b.dst = x->fin_dst[0];
        b.dst2 = x->fin_dst[1];
        b.dst3 = x->fin_dst[2];
        b.dst4 = x->fin_dst[3];
        b.v = x->fin_v;
        b.theta = x->fin_theta;
      }
      back.push_back(b);
    }
  }

  int record_run(std::vector<Xform> &xs) {
    int max_it = 0;
    for (auto &x : xs) max_it = std::max(max_it, x.iters);
    for (int it = 0; it < max_it; ++it) {
      std::vector<Xform *> active;
      for (auto &x : xs)
        if (x.iters > it) {
          x.it = it;
          active.push_back(&x);
        }
      int rc;
      for (auto *x : active) {
        x->xst = xstore_candidate(*x);
        x->xsa = xsa_candidate(*x);
      }
      if ((rc = sweep_pass(active, 0))) return rc;
      if ((rc = sweep_pass(active, 1))) return rc;
      std::vector<XposeDesc> to, back;
      w_pass_to(active, to);
      record_xpose(to);
      if ((rc = sweep_pass(active, 2))) return rc;
      w_pass_back(active, back);
      record_xpose(back);
    }
    return FG_OK;
  }

  // Whether the first sweep of x runs on the tiled engine whatever halo depth
  // its pass is planned with (the plane fits the device's co-resident
  // workgroups even at the shallowest halo, which has the most tiles).
  int first_sweep_tiled(const Xform &x, bool &tiled) {
    tiled = false;
    const int pass = first_pass(x);
    if (x.iters <= 0 || !tiled_eligible(x, pass)) return FG_OK;
    int nA, nB, nE;
    pass_plane(x, pass, nA, nB, nE);
    const Plan p = plan_fixed(nB, nE, K > 0 ? K : 1, x.C);
    if (!p.ok) return FG_OK;
    const void *kern = pick_kernel(p.NW, p.R, p.P, x.C, x.lamb1, x.fast, false, true, p.comm);
    if (!kern) return FG_OK;
    long long cap = 0;
    const int rc = coresident_groups(kern, (p.NW + (p.comm ? kCommWaves : 0)) * 64, dev, cap);
    if (rc) return rc;
    if (max_groups > 0) cap = std::min<long long>(cap, max_groups);
    tiled = (long long)p.nTB * p.nTE <= cap;
    return FG_OK;
  }

  // status: the status word the launches write; no_alloc: create nothing on the
  // host side (no pinned staging, no events) and pass every descriptor by value
  int replay(hipStream_t stream, unsigned char *blob_dev, size_t blob_cap, unsigned long long *halo, int *status,
             bool no_alloc) {
    if (rec.blob.size() > blob_cap) return fail(FG_EWORKSPACE, "descriptor area too small");
    // the halo base is only known now: patch fill descriptors
    for (auto &c : rec.cmds)
      if (c.type == Cmd::EW) {
        EwDesc *ds = c.by_value ? &c.ew1 : reinterpret_cast<EwDesc *>(rec.blob.data() + c.off);
        for (int i = 0; i < c.ndesc; ++i)
          if (ds[i].kind == EW_FILL_U64) ds[i].u64 = halo;
      }
    bool tiled = false;
    for (auto &c : rec.cmds) tiled = tiled || c.type == Cmd::TILED;
    DeviceState &ds = device_state(dev.id);
    if (no_alloc && !rec.blob.empty())
      return fail(FG_EINVAL, "internal: a launch of this call needs a descriptor array");
    OrderedReplay order;
    int rc = order.begin(ds, stream, tiled, !no_alloc);
    if (rc) return rc;
    if (!rec.blob.empty() && (rc = stage_copy(ds, blob_dev, rec.blob.data(), rec.blob.size(), stream))) return rc;
    // the kernel expires a wait when a check finds a halo value pending after spin re-polls
    const int spin = g_wait_limit > 0 ? g_wait_limit - 1 : (1 << 21);

    // Argument values of every launch (the arrays point into `vals`, which
    // outlives the launches and graph updates below).
    struct Vals {
      const void *dp;
      int ndesc, k, sp, a, aq;
      unsigned tb;
    };
    const size_t n = rec.cmds.size();
    std::vector<Vals> vals(n);
    std::vector<std::array<void *, 9>> args(n);
    // exchange-wait histogram of this device while profiling
    unsigned *xhist = nullptr;
    if (g_prof_xhist && (rc = exchange_histogram(dev.id, &xhist))) return rc;
    std::vector<const void *> funcs(n);
    bool slice = false;
    for (size_t i = 0; i < n; ++i) {
      Cmd &c = rec.cmds[i];
      Vals &v = vals[i];
      v.ndesc = c.ndesc;
      auto &a = args[i];
      switch (c.type) {
        case Cmd::TILED:
          v.dp = c.by_value ? nullptr : blob_dev + c.off;
          v.k = c.K;
          v.sp = spin;
          v.tb = c.tag_base;
          a = {(void *)&v.dp, &v.ndesc, (void *)&c.sweep1, &v.k, &v.tb, (void *)&halo, (void *)&status, &v.sp,
               (void *)&xhist};
          funcs[i] = c.kern;
          break;
        case Cmd::SLICE:
          v.a = c.a;
          v.aq = c.aq;
          a = {(void *)&c.slice, &v.a, &v.aq, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
          funcs[i] = reinterpret_cast<const void *>(&slice_kernel);
          slice = true;
          break;
        case Cmd::XPOSE:
          v.dp = c.by_value ? nullptr : blob_dev + c.off;
          a = {(void *)&v.dp, &v.ndesc, (void *)&c.xpose1, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
          funcs[i] = reinterpret_cast<const void *>(&xpose_kernel);
          break;
        case Cmd::EW:
          v.dp = c.by_value ? nullptr : blob_dev + c.off;
          a = {(void *)&v.dp, &v.ndesc, (void *)&c.ew1, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
          funcs[i] = reinterpret_cast<const void *>(&ew_kernel);
          break;
      }
    }
    auto node_params = [&](size_t i) {
      hipKernelNodeParams p{};
      p.func = const_cast<void *>(funcs[i]);
      p.gridDim = dim3(rec.cmds[i].grid);
      p.blockDim = dim3(rec.cmds[i].block);
      p.sharedMemBytes = 0;
      p.kernelParams = args[i].data();
      p.extra = nullptr;
      return p;
    };

    // graph replay of a repeated sequence (not while profiling per launch, not
    // for the per-slice engine's hundreds of launches, never without allocating)
    if (g_graph_on && !no_graph && !g_prof_on && !g_prof_xhist && !slice && !no_alloc && n >= 2) {
      uint64_t key = 1469598103934665603ull;
      auto mix = [&](uint64_t x) { key = (key ^ x) * 1099511628211ull; };
      for (size_t i = 0; i < n; ++i) {
        mix((uint64_t)(uintptr_t)funcs[i]);
        mix((uint64_t)rec.cmds[i].grid);
        mix((uint64_t)rec.cmds[i].block);
      }
      GraphEntry *e = nullptr;
      {
        std::lock_guard<std::mutex> g(g_graph_m);
        for (auto &p : g_graph_list)
          if (p->thread == std::this_thread::get_id() && p->key == key && p->dev == dev.id && p->nodes.size() == n)
            e = p.get();
      }
      bool fresh = false;
      if (!e) {
        auto seen = std::find(g_graph_seen.begin(), g_graph_seen.end(), std::make_pair(key, dev.id));
        if (seen == g_graph_seen.end()) {
          g_graph_seen.push_back({key, dev.id});
          if (g_graph_seen.size() > 32) g_graph_seen.erase(g_graph_seen.begin());
        } else {
          std::unique_ptr<GraphEntry> ne(new GraphEntry());
          ne->thread = std::this_thread::get_id();
          ne->key = key;
          ne->dev = dev.id;
          bool ok = hipGraphCreate(&ne->graph, 0) == hipSuccess;
          for (size_t i = 0; ok && i < n; ++i) {
            hipGraphNode_t node{};
            const hipKernelNodeParams p = node_params(i);
            ok = hipGraphAddKernelNode(&node, ne->graph, i ? &ne->nodes.back() : nullptr, i ? 1 : 0, &p) ==
                 hipSuccess;
            if (ok) ne->nodes.push_back(node);
          }
          ok = ok && hipGraphInstantiate(&ne->exec, ne->graph, nullptr, nullptr, 0) == hipSuccess;
          if (ok) {
            std::lock_guard<std::mutex> g(g_graph_m);
            // keep at most kGraphsPerThread entries per thread: drop the least recently used
            std::vector<GraphEntry *> mine;
            for (auto &p : g_graph_list)
              if (p->thread == ne->thread) mine.push_back(p.get());
            if ((int)mine.size() >= kGraphsPerThread) {
              GraphEntry *old = *std::min_element(mine.begin(), mine.end(), [](GraphEntry *x, GraphEntry *y) {
                return x->last_use < y->last_use;
              });
              destroy_graph(*old);
              g_graph_list.erase(std::find_if(g_graph_list.begin(), g_graph_list.end(),
                                              [&](const std::unique_ptr<GraphEntry> &p) { return p.get() == old; }));
            }
            e = ne.get();
            g_graph_list.push_back(std::move(ne));
            fresh = true;
          } else {
            destroy_graph(*ne);
            (void)hipGetLastError();
          }
        }
      }
      if (e) {
        e->last_use = ++g_graph_clock;
        bool ok = true;
        for (size_t i = 0; ok && !fresh && i < n; ++i) {
          const hipKernelNodeParams p = node_params(i);
          ok = hipGraphExecKernelNodeSetParams(e->exec, e->nodes[i], &p) == hipSuccess;
        }
        if (ok) {
          FG_HIP(hipGraphLaunch(e->exec, stream));
          return FG_OK;
        }
        (void)hipGetLastError();  // an update failed: launch directly this time
      }
    }

    for (size_t i = 0; i < n; ++i) {
      Cmd &c = rec.cmds[i];
      ProfRec pr{};
      const bool prof = g_prof_on && c.type == Cmd::TILED;
      if (prof) {
        FG_HIP(hipEventCreate(&pr.start));
        FG_HIP(hipEventCreate(&pr.stop));
        FG_HIP(hipEventRecord(pr.start, stream));
      }
      FG_HIP(hipLaunchKernel(funcs[i], dim3(c.grid), dim3(c.block), args[i].data(), 0, stream));
      if (prof) {
        FG_HIP(hipEventRecord(pr.stop, stream));
        pr.bytes = c.alg_bytes;
        g_prof.push_back(pr);
      }
    }
    return FG_OK;
  }
};

// ---------------------------------------------------------------------------
// Operation planning shared by the single-call and batched entry points
// ---------------------------------------------------------------------------

struct OpSpec {
  const float *image;
  const float *imageT;  // the image already in the W pass layout (C, W, H, D) / (C, W, H), or null
  const float *mask;
  float *out;
  int C, Cm, D, H, W;
  bool is2d;
  int op;
  int iters;
  float spacing[3];
  float v, lg, le, theta;
};

bool non_negative(float x) { return !std::signbit(x); }

void make_xform(const OpSpec &o, float lg, float le, Xform &x) {
  x.C = o.C; x.D = o.D; x.H = o.H; x.W = o.W; x.is2d = o.is2d; x.lg = lg; x.iters = o.iters;
  float ld[3][9];
  if (o.is2d) {
    table2d(ld[0]);
    table2d(ld[1]);
    table2d(ld[2]);
  } else {
    float sq[3];
    for (int k = 0; k < 3; ++k) sq[k] = o.spacing[k] * o.spacing[k];
    table3d(sq[0], sq[1], sq[2], ld[0]);  // D pass: in-plane (H, W)
    table3d(sq[1], sq[0], sq[2], ld[1]);  // H pass: in-plane (D, W)
    table3d(sq[2], sq[1], sq[0], ld[2]);  // W pass: in-plane (H, D)
  }
  bool fast = non_negative(lg);
  bool e1zero = true;
  for (int p = 0; p < 3; ++p)
    for (int k = 0; k < 9; ++k) {
      x.e1[p][k] = le * ld[p][k];
      fast = fast && non_negative(x.e1[p][k]);
      uint32_t bits;
      std::memcpy(&bits, &x.e1[p][k], 4);
      e1zero = e1zero && bits == 0u;
    }
  x.fast = fast;
  x.lamb1 = (lg == 1.0f) && e1zero;
}

long long vox(const OpSpec &o) { return (long long)o.D * o.H * o.W; }

int n_dist(int op) { return op == FG_OP_UNSIGNED ? 1 : (op == FG_OP_SIGNED ? 2 : 4); }

size_t halo_entries_for(const std::vector<OpSpec> &ops, int K) {
  size_t total = 0;
  for (auto &o : ops) {
    Xform x{};
    x.D = o.D; x.H = o.H; x.W = o.W; x.is2d = o.is2d;
    size_t worst = 0;
    for (int p = 0; p < 3; ++p) {
      int nA, nB, nE;
      pass_plane(x, p, nA, nB, nE);
      for (int k = (K > 0 ? K : 1); k <= (K > 0 ? K : kMaxK); ++k) {
        for (int sh : {g_shape, (g_tune.shape == 0 && K <= 0) ? kSmallPlaneShape : g_shape}) {
          Plan pl = plan_fixed(nB, nE, k, o.C, sh);
          if (!pl.ok) continue;
          worst = std::max(worst, (size_t)pl.nTB * pl.nTE * 2 * pl.TB * pl.TE);
        }
      }
    }
    total += worst * (size_t)n_dist(o.op);
  }
  return total;
}

// Upper bound of the descriptor blob: identical descriptor arrays are stored
// once, and the active set of transforms changes at most once per distinct
// iteration count.
size_t blob_bytes_for(const std::vector<OpSpec> &ops) {
  size_t ntrans = 0;
  std::vector<int> its;
  for (auto &o : ops) {
    ntrans += n_dist(o.op);
    if (std::find(its.begin(), its.end(), o.iters) == its.end()) its.push_back(o.iters);
  }
  const size_t sets = its.size() + 1;
  const size_t nsweep = 2 * 6 * sets;  // (pass, dir) x active sets, x2 for GSF stages
  size_t bytes = 0;
  bytes += nsweep * (ntrans * sizeof(SweepDesc) + 256) * 2;
  bytes += (4 * sets + 2) * (ntrans * sizeof(XposeDesc) + 256) + (ops.size() * sizeof(XposeDesc) + 256);
  bytes += 8 * (ops.size() * sizeof(EwDesc) + 256);
  return bytes + 4096;
}

struct Layout {
  unsigned char *blob;
  size_t blob_cap;
  unsigned long long *halo;
  size_t halo_cap;
};

void layout_common(const std::vector<OpSpec> &ops, int K, Layout &L, Carver &cv) {
  L.blob_cap = blob_bytes_for(ops);
  L.blob = (unsigned char *)cv.take(L.blob_cap);
  L.halo_cap = halo_entries_for(ops, K);
  L.halo = (unsigned long long *)cv.take(L.halo_cap * sizeof(unsigned long long) + 8);
}

size_t op_buffers_bytes(const OpSpec &o) {
  const long long N = vox(o);
  const int nd = n_dist(o.op);
  // image transpose (C*N, unless given) + extra distances (nd-1) + transposed distances (nd)
  const size_t floats = (o.imageT ? 0 : (size_t)o.C * N) + (size_t)(nd - 1) * N + (size_t)nd * N;
  return floats * sizeof(float) + 16 * 256;
}

size_t workspace_for(const std::vector<OpSpec> &ops, int K) {
  Layout L{};
  Carver cv(nullptr, 0, true);
  layout_common(ops, K, L, cv);
  for (auto &o : ops) cv.take(op_buffers_bytes(o));
  return ((cv.used + 255) & ~(size_t)255) + 256;
}

int validate(const OpSpec &o) {
  if (!o.image || !o.mask || !o.out) return fail(FG_EINVAL, "null pointer");
  if (o.C < 1 || o.Cm < 1 || o.D < 1 || o.H < 1 || o.W < 1)
    return fail(FG_EINVAL, "extents and channel counts must be >= 1");
  if (o.is2d && o.D != 1) return fail(FG_EINVAL, "2D instance must have D == 1");
  if (o.op < FG_OP_UNSIGNED || o.op > FG_OP_GSF) return fail(FG_EINVAL, "unknown op");
  if (vox(o) * std::max(o.C, o.Cm) >= (1ll << 40)) return fail(FG_EINVAL, "volume too large");
  if (o.H >= (1 << 30) || o.W >= (1 << 30) || o.D >= (1 << 30)) return fail(FG_EINVAL, "extent too large");
  return FG_OK;
}

struct RunOpts {
  bool no_alloc = false;           // create nothing (no pinned status word, staging or events)
  int *fallback_status = nullptr;  // status word to use while the device has no pinned one
};

int run_ops(std::vector<OpSpec> &ops, void *ws, size_t wsb, hipStream_t stream, const RunOpts &opt = RunOpts()) {
  for (auto &o : ops) {
    int rc = validate(o);
    if (rc) return rc;
  }
  const int K = current_halo();
  const size_t need = workspace_for(ops, K);
  if (!ws || wsb < need)
    return fail(FG_EWORKSPACE, "workspace too small: need " + std::to_string(need) + " bytes");
  Layout L{};
  Carver cv(ws, wsb, false);
  layout_common(ops, K, L, cv);

  Runner run{};
  int rc = query_device(run.dev);
  if (rc) return rc;
  // an expiry left by an earlier call on this device is reported here
  DeviceState &dstate = device_state(run.dev.id);
  if (!opt.no_alloc && (rc = ensure_status_word(dstate))) return rc;
  if ((rc = status_error(take_status(dstate), "an earlier call on this device"))) return rc;
  int *status = dstate.status_dev ? dstate.status_dev : opt.fallback_status;
  if (!status) return fail(FG_EINVAL, "internal: no status word");
  run.K = K;
  run.engine = g_tune.engine;
  run.max_groups = g_tune.max_groups;
  run.halo_cap = L.halo_cap;

  struct OpBufs { float *imgT; float *dist[4]; float *dT[4]; };
  std::vector<OpBufs> bufs(ops.size());
  for (size_t i = 0; i < ops.size(); ++i) {
    const OpSpec &o = ops[i];
    const long long N = vox(o);
    const int nd = n_dist(o.op);
    const size_t start = cv.used;
    bufs[i].imgT = o.imageT ? const_cast<float *>(o.imageT) : (float *)cv.take((size_t)o.C * N * sizeof(float));
    bufs[i].dist[0] = o.out;
    for (int k = 1; k < nd; ++k) bufs[i].dist[k] = (float *)cv.take((size_t)N * sizeof(float));
    for (int k = 0; k < nd; ++k) bufs[i].dT[k] = (float *)cv.take((size_t)N * sizeof(float));
    cv.used = start;
    cv.take(op_buffers_bytes(o));
  }
  if (!cv.fits()) return fail(FG_EWORKSPACE, "workspace too small");

  // the halo tags start from zero on every call
  run.record_zero_halo();

  // image transposes (once per op)
  {
    std::vector<XposeDesc> xs;
    for (size_t i = 0; i < ops.size(); ++i) {
      const OpSpec &o = ops[i];
      if (o.imageT) continue;  // the caller keeps the transposed image
      if (o.is2d) xs.push_back(Runner::xd(o.image, bufs[i].imgT, o.C, o.H, 1, o.W));
      else xs.push_back(Runner::xd(o.image, bufs[i].imgT, o.C, o.D, o.H, o.W));
    }
    run.record_xpose(xs);
  }

  // Stage 1: initialise and transform.  With iterations, a transform's first
  // sweep computes its initialisation from the mask (when that sweep is
  // tiled), mask channels >= 1 get their closed form in one elementwise pass,
  // and each op's epilogue is applied by its last transpose back (signed:
  // D(M) - D(1-M); GSF: the thresholds that initialise stage 2).  Without
  // iterations the elementwise kernels do all of it.
  std::vector<EwDesc> init, chans;
  std::vector<Xform> xs;
  std::vector<char> fused(ops.size(), 0);
  for (size_t i = 0; i < ops.size(); ++i) {
    const OpSpec &o = ops[i];
    const long long N = vox(o);
    fused[i] = o.iters > 0;
    EwDesc e{};
    e.n0 = N;
    e.Cm = o.Cm;
    e.op = o.op;
    e.mask = o.mask;
    e.out = o.out;
    e.v = o.v;
    e.theta = o.theta;
    float lg = o.lg, le = o.le;
    if (o.op == FG_OP_GSF) le = 1 - o.lg;  // reference fastgeodis.cpp:321: 1 - lambda in float
    const int nd = (o.op == FG_OP_UNSIGNED) ? 1 : 2;
    const size_t j0 = xs.size();
    for (int k = 0; k < nd; ++k) {
      Xform x{};
      make_xform(o, lg, le, x);
      x.img = o.image;
      x.imgT = bufs[i].imgT;
      x.d = bufs[i].dist[k];
      x.dT = bufs[i].dT[k];
      xs.push_back(x);
    }
    EwDesc ie = e;
    ie.kind = (o.op == FG_OP_UNSIGNED) ? EW_INIT_UNSIGNED : EW_INIT_SIGNED;
    ie.a = bufs[i].dist[1];
    if (!fused[i]) {
      ie.nelem = N * o.Cm;  // every channel
      init.push_back(ie);
      continue;
    }
    if (o.Cm > 1) {
      EwDesc c = e;
      c.kind = EW_CHANNELS;
      c.nelem = N * (o.Cm - 1);
      chans.push_back(c);
    }
    bool tiled = false;
    if ((rc = run.first_sweep_tiled(xs[j0], tiled))) return rc;
    if (tiled) {
      for (int k = 0; k < nd; ++k) {
        xs[j0 + k].init_mode = (k == 0) ? 1 : 2;
        xs[j0 + k].init_src = o.mask;
        xs[j0 + k].init_v = o.v;
      }
    } else {
      ie.nelem = N;  // channel 0
      init.push_back(ie);
    }
    if (o.op != FG_OP_UNSIGNED) {
      Xform &x0 = xs[j0];
      x0.fin_kind = (o.op == FG_OP_SIGNED) ? XK_SIGNED : XK_GSF_MID;
      x0.fin_src[0] = xs[j0 + 1].dT;
      x0.fin_dst[0] = o.out;
      if (o.op == FG_OP_GSF)
        for (int k = 1; k < 4; ++k) x0.fin_dst[k] = bufs[i].dist[k];
      x0.fin_v = o.v;
      x0.fin_theta = o.theta;
      xs[j0 + 1].fin_member = true;
    }
  }
  run.record_ew(init);
  run.record_ew(chans);
  if ((rc = run.record_run(xs))) return rc;

  // Stage 2 (GSF): four more transforms from the thresholds.
  std::vector<EwDesc> mid, fin;
  std::vector<Xform> xs2;
  for (size_t i = 0; i < ops.size(); ++i) {
    const OpSpec &o = ops[i];
    const long long N = vox(o);
    EwDesc e{};
    e.n0 = N;
    e.Cm = o.Cm;
    e.op = o.op;
    e.mask = o.mask;
    e.out = o.out;
    e.v = o.v;
    e.theta = o.theta;
    if (o.op == FG_OP_SIGNED && !fused[i]) {
      e.kind = EW_FINAL_SIGNED;
      e.a = bufs[i].dist[1];
      e.nelem = N * o.Cm;
      fin.push_back(e);
    } else if (o.op == FG_OP_GSF) {
      if (!fused[i]) {
        EwDesc m = e;
        m.kind = EW_GSF_MID;
        m.a = bufs[i].dist[1];
        m.b = bufs[i].dist[2];
        m.c2 = bufs[i].dist[3];
        m.nelem = N;
        mid.push_back(m);
        EwDesc f = m;
        f.kind = EW_GSF_FINAL;
        f.nelem = N * o.Cm;
        fin.push_back(f);
      }
      const float le = 1 - o.lg;
      // dist[0] = out: D(1-Me); dist[1]: D(1-Md); dist[2]: D(Md); dist[3]: D(Me)
      const size_t j0 = xs2.size();
      for (int k = 0; k < 4; ++k) {
        Xform x{};
        make_xform(o, o.lg, le, x);
        x.img = o.image;
        x.imgT = bufs[i].imgT;
        x.d = bufs[i].dist[k];
        x.dT = bufs[i].dT[k];
        xs2.push_back(x);
      }
      if (fused[i]) {
        Xform &x0 = xs2[j0];
        x0.fin_kind = XK_GSF_FINAL;
        for (int k = 1; k < 4; ++k) {
          x0.fin_src[k - 1] = xs2[j0 + k].dT;
          xs2[j0 + k].fin_member = true;
        }
        x0.fin_dst[0] = o.out;
      }
    }
  }
  run.record_ew(mid);
  if (!xs2.empty() && (rc = run.record_run(xs2))) return rc;
  run.record_ew(fin);
  return run.replay(stream, L.blob, L.blob_cap, L.halo, status, opt.no_alloc);
}

OpSpec spec2d(const float *image, const float *mask, float *out, int C, int Cm, int H, int W, int op,
              int iters, float v, float lg, float le, float theta) {
  OpSpec o{};
  o.image = image; o.mask = mask; o.out = out; o.C = C; o.Cm = Cm; o.D = 1; o.H = H; o.W = W;
  o.is2d = true; o.op = op; o.iters = iters; o.v = v; o.lg = lg; o.le = le; o.theta = theta;
  o.spacing[0] = o.spacing[1] = o.spacing[2] = 1.0f;
  return o;
}

OpSpec spec3d(const float *image, const float *mask, float *out, int C, int Cm, int D, int H, int W,
              const float *spacing, int op, int iters, float v, float lg, float le, float theta) {
  OpSpec o{};
  o.image = image; o.mask = mask; o.out = out; o.C = C; o.Cm = Cm; o.D = D; o.H = H; o.W = W;
  o.is2d = false; o.op = op; o.iters = iters; o.v = v; o.lg = lg; o.le = le; o.theta = theta;
  for (int k = 0; k < 3; ++k) o.spacing[k] = spacing ? spacing[k] : 1.0f;
  return o;
}

int run_single(OpSpec o, void *ws, size_t wsb, void *stream) {
  std::vector<OpSpec> ops{o};
  return run_ops(ops, ws, wsb, (hipStream_t)stream);
}

// ---------------------------------------------------------------------------
// Stepper: transforms that advance one pass pair per step, so that new ones
// join between steps (continuous batching of a serving loop).  Each step
// records, for every active transform, its next pass pair (forward and
// backward sweep) in launches shared with every other transform at that
// pass, plus the transposes around W passes; a transform that has run its
// last W pass leaves with its epilogue applied.  Results equal the same
// transform run alone: a transform computes exactly the launches of its
// plain call, only in the company of others.
// ---------------------------------------------------------------------------

struct StepOp {
  OpSpec o;
  int ticket = 0;
  bool started = false;   // its first step (image transpose, mask channels >= 1) recorded
  int it = 0;             // iteration of its next pass pair
  int pass = 0;           // next pass: 0 D, 1 H, 2 W (2D: 1 rows, 2 columns)
  float *imgT = nullptr;  // the image in the W pass layout
  std::vector<Xform> xs;  // 1 (unsigned) or 2 (signed: D(M), D(1-M))
  size_t halo_entries = 0;
};

struct Stepper {
  int device = -1;
  unsigned long long *halo = nullptr;
  size_t halo_cap = 0;  // entries
  unsigned char *blob = nullptr;
  size_t blob_cap = 0;
  bool halo_clear = false;  // the first step clears the exchange buffer
  unsigned tag_base = 0;
  int next_ticket = 1;
  size_t halo_used = 0;  // entries reserved by the active transforms
  std::vector<StepOp> ops;
};

size_t stepper_blob_bytes(int max_transforms) {
  const size_t n = (size_t)std::max(max_transforms, 1);
  // per step: 3 passes x 2 directions of sweeps, 3 transpose launches, 2 elementwise
  return 6 * (n * sizeof(SweepDesc) + 256) * 2 + 3 * (n * sizeof(XposeDesc) + 256) +
         2 * (n * sizeof(EwDesc) + 256) + 4096;
}

int stepper_runner(const Stepper &st, Runner &run) {
  int rc = query_device(run.dev);
  if (rc) return rc;
  if (run.dev.id != st.device) return fail(FG_EINVAL, "the stepper belongs to another device");
  run.K = current_halo();
  run.engine = g_tune.engine;
  run.max_groups = g_tune.max_groups;
  run.halo_cap = st.halo_cap;
  run.tag_base = st.tag_base;
  run.no_graph = true;
  return FG_OK;
}

int stepper_add(Stepper &st, const OpSpec &o, void *ws, size_t wsb, int *ticket) {
  int rc = validate(o);
  if (rc) return rc;
  if (o.op != FG_OP_UNSIGNED && o.op != FG_OP_SIGNED) return fail(FG_EINVAL, "the stepper runs unsigned and signed transforms");
  if (o.iters < 1) return fail(FG_EINVAL, "the stepper needs iterations >= 1");
  if (!ws || wsb < op_buffers_bytes(o)) return fail(FG_EWORKSPACE, "transform workspace too small");
  Runner run{};
  if ((rc = stepper_runner(st, run))) return rc;
  const size_t hneed = halo_entries_for({o}, run.K);
  if (st.halo_used + hneed > st.halo_cap) return fail(FG_EWORKSPACE, "the stepper's exchange buffer is full");

  StepOp op;
  op.o = o;
  const long long N = vox(o);
  const int nd = n_dist(o.op);
  Carver cv(ws, wsb, false);
  op.imgT = o.imageT ? const_cast<float *>(o.imageT) : (float *)cv.take((size_t)o.C * N * sizeof(float));
  float *dist[2] = {o.out, nd > 1 ? (float *)cv.take((size_t)N * sizeof(float)) : nullptr};
  float *dT[2] = {nullptr, nullptr};
  for (int k = 0; k < nd; ++k) dT[k] = (float *)cv.take((size_t)N * sizeof(float));
  if (!cv.fits()) return fail(FG_EWORKSPACE, "transform workspace too small");
  for (int k = 0; k < nd; ++k) {
    Xform x{};
    make_xform(o, o.lg, o.le, x);
    x.img = o.image;
    x.imgT = op.imgT;
    x.d = dist[k];
    x.dT = dT[k];
    op.xs.push_back(x);
  }
  // the initialisation is computed by the first sweep, which therefore must
  // run on the tiled engine
  bool tiled = false;
  if ((rc = run.first_sweep_tiled(op.xs[0], tiled))) return rc;
  if (!tiled) return fail(FG_EINVAL, "the stepper needs a first sweep on the tiled engine (plane too large)");
  for (int k = 0; k < nd; ++k) {
    op.xs[k].init_mode = (k == 0) ? 1 : 2;
    op.xs[k].init_src = o.mask;
    op.xs[k].init_v = o.v;
  }
  if (o.op == FG_OP_SIGNED) {
    op.xs[0].fin_kind = XK_SIGNED;
    op.xs[0].fin_src[0] = op.xs[1].dT;
    op.xs[0].fin_dst[0] = o.out;
    op.xs[0].fin_v = o.v;
    op.xs[1].fin_member = true;
  }
  op.pass = first_pass(op.xs[0]);
  op.ticket = st.next_ticket++;
  if (st.next_ticket <= 0) st.next_ticket = 1;
  op.halo_entries = hneed;
  st.halo_used += hneed;
  *ticket = op.ticket;
  st.ops.push_back(std::move(op));
  return FG_OK;
}

int stepper_step(Stepper &st, hipStream_t stream, int *done, int cap, int *ndone) {
  *ndone = 0;
  if (st.ops.empty()) return FG_OK;
  Runner run{};
  int rc = stepper_runner(st, run);
  if (rc) return rc;
  DeviceState &dstate = device_state(run.dev.id);
  if ((rc = ensure_status_word(dstate))) return rc;
  if ((rc = status_error(take_status(dstate), "an earlier call on this device"))) return rc;
  if (!st.halo_clear) run.record_zero_halo();

  // first steps: image transposes and mask channels >= 1 (closed form)
  std::vector<XposeDesc> ximg;
  std::vector<EwDesc> chans;
  for (auto &op : st.ops) {
    if (op.started) continue;
    const OpSpec &o = op.o;
    if (!o.imageT) {
      if (o.is2d) ximg.push_back(Runner::xd(o.image, op.imgT, o.C, o.H, 1, o.W));
      else ximg.push_back(Runner::xd(o.image, op.imgT, o.C, o.D, o.H, o.W));
    }
    if (o.Cm > 1) {
      EwDesc c{};
      c.n0 = vox(o);
      c.Cm = o.Cm;
      c.op = o.op;
      c.mask = o.mask;
      c.out = o.out;
      c.v = o.v;
      c.theta = o.theta;
      c.kind = EW_CHANNELS;
      c.nelem = vox(o) * (o.Cm - 1);
      chans.push_back(c);
    }
  }
  run.record_xpose(ximg);
  run.record_ew(chans);

  std::vector<Xform *> act[3];
  for (auto &op : st.ops)
    for (auto &x : op.xs) {
      x.it = op.it;
      act[op.pass].push_back(&x);
    }
  for (auto *x : act[1]) x->xst = run.xstore_candidate(*x);
  for (auto *x : act[2]) x->xsa = run.xsa_candidate(*x);
  std::vector<XposeDesc> to, back;
  run.w_pass_to(act[2], to);
  run.record_xpose(to);
  for (int pass = 0; pass < 3; ++pass)
    if (!act[pass].empty() && (rc = run.sweep_pass(act[pass], pass))) return rc;
  run.w_pass_back(act[2], back);
  run.record_xpose(back);
  if ((rc = run.replay(stream, st.blob, st.blob_cap, st.halo, dstate.status_dev, false))) return rc;
  st.halo_clear = true;
  st.tag_base = run.tag_base;

  // advance; transforms past their last W pass leave
  std::vector<StepOp> keep;
  for (auto &op : st.ops) {
    op.started = true;
    if (op.pass == 2) {
      op.pass = first_pass(op.xs[0]);
      if (++op.it == op.o.iters) {
        if (*ndone < cap) done[(*ndone)++] = op.ticket;
        st.halo_used -= op.halo_entries;
        continue;
      }
    } else {
      ++op.pass;
    }
    keep.push_back(std::move(op));
  }
  st.ops.swap(keep);
  return FG_OK;
}


}  // namespace

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

// Exchange-wait histogram filled by the sweep kernels while profiling.
static __device__ unsigned xhist_dev[kXHistBuckets];

namespace {

std::mutex g_xhist_m;
std::map<int, unsigned *> g_xhist_ptr;  // device -> address of xhist_dev

int exchange_histogram_impl(int dev, unsigned **out) {
  std::lock_guard<std::mutex> g(g_xhist_m);
  auto it = g_xhist_ptr.find(dev);
  if (it == g_xhist_ptr.end()) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(xhist_dev));
    if (e != hipSuccess || !p) return fail(FG_EHIP, std::string("hipGetSymbolAddress(xhist): ") + hipGetErrorString(e));
    it = g_xhist_ptr.emplace(dev, (unsigned *)p).first;
  }
  *out = it->second;
  return FG_OK;
}

// Device address of this device's scratch array (the device is current),
// looked up once per device.
int noalloc_scratch(DeviceState &ds, unsigned char **out) {
  std::lock_guard<std::mutex> g(ds.m);
  if (!ds.scratch) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(FG_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    ds.scratch = (unsigned char *)p;
  }
  *out = ds.scratch;
  return FG_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

extern "C" {

int fg_abi_version(void) { return 4; }

const char *fg_last_error(void) { return g_err.c_str(); }

size_t fg_workspace_bytes(int op, int C, int D, int H, int W) {
  OpSpec o{};
  o.op = (op < FG_OP_UNSIGNED || op > FG_OP_GSF) ? FG_OP_GSF : op;
  o.C = std::max(C, 1); o.Cm = 1; o.D = std::max(D, 1); o.H = std::max(H, 1); o.W = std::max(W, 1);
  o.is2d = (D <= 0);
  o.iters = 1;
  std::vector<OpSpec> ops{o};
  return workspace_for(ops, current_halo());
}

int fg_geodesic2d_f32(const float *image, const float *mask, float *out, int C, int Cm, int H, int W,
                      float v, float l_grad, float l_eucl, int iterations, void *workspace,
                      size_t workspace_bytes, void *stream) {
  return run_single(spec2d(image, mask, out, C, Cm, H, W, FG_OP_UNSIGNED, iterations, v, l_grad, l_eucl, 0.f),
                    workspace, workspace_bytes, stream);
}

int fg_geodesic3d_f32(const float *image, const float *mask, float *out, int C, int Cm, int D, int H, int W,
                      const float spacing[3], float v, float l_grad, float l_eucl, int iterations,
                      void *workspace, size_t workspace_bytes, void *stream) {
  if (!spacing) return fail(FG_EINVAL, "spacing must have 3 entries");
  return run_single(
      spec3d(image, mask, out, C, Cm, D, H, W, spacing, FG_OP_UNSIGNED, iterations, v, l_grad, l_eucl, 0.f),
      workspace, workspace_bytes, stream);
}

int fg_signed_geodesic2d_f32(const float *image, const float *mask, float *out, int C, int Cm, int H, int W,
                             float v, float l_grad, float l_eucl, int iterations, void *workspace,
                             size_t workspace_bytes, void *stream) {
  return run_single(spec2d(image, mask, out, C, Cm, H, W, FG_OP_SIGNED, iterations, v, l_grad, l_eucl, 0.f),
                    workspace, workspace_bytes, stream);
}

int fg_signed_geodesic3d_f32(const float *image, const float *mask, float *out, int C, int Cm, int D, int H,
                             int W, const float spacing[3], float v, float l_grad, float l_eucl,
                             int iterations, void *workspace, size_t workspace_bytes, void *stream) {
  if (!spacing) return fail(FG_EINVAL, "spacing must have 3 entries");
  return run_single(
      spec3d(image, mask, out, C, Cm, D, H, W, spacing, FG_OP_SIGNED, iterations, v, l_grad, l_eucl, 0.f),
      workspace, workspace_bytes, stream);
}

int fg_gsf2d_f32(const float *image, const float *mask, float *out, int C, int Cm, int H, int W, float theta,
                 float v, float lambda, int iterations, void *workspace, size_t workspace_bytes,
                 void *stream) {
  return run_single(spec2d(image, mask, out, C, Cm, H, W, FG_OP_GSF, iterations, v, lambda, 0.f, theta),
                    workspace, workspace_bytes, stream);
}

int fg_gsf3d_f32(const float *image, const float *mask, float *out, int C, int Cm, int D, int H, int W,
                 float theta, const float spacing[3], float v, float lambda, int iterations, void *workspace,
                 size_t workspace_bytes, void *stream) {
  if (!spacing) return fail(FG_EINVAL, "spacing must have 3 entries");
  return run_single(
      spec3d(image, mask, out, C, Cm, D, H, W, spacing, FG_OP_GSF, iterations, v, lambda, 0.f, theta),
      workspace, workspace_bytes, stream);
}

int fg_geodesic3d_f32_noalloc(const float *image, const float *mask, float *out, int C, int Cm, int D, int H,
                              int W, const float spacing[3], float v, float l_grad, float l_eucl, int iterations,
                              void *stream) {
  if (!spacing) return fail(FG_EINVAL, "spacing must have 3 entries");
  std::vector<OpSpec> ops{
      spec3d(image, mask, out, C, Cm, D, H, W, spacing, FG_OP_UNSIGNED, iterations, v, l_grad, l_eucl, 0.f)};
  int rc = validate(ops[0]);
  if (rc) return rc;
  const size_t need = workspace_for(ops, current_halo());
  if (need > kNoAllocScratchBytes - kNoAllocStatusBytes)
    return fail(FG_EWORKSPACE, "fg_geodesic3d_f32_noalloc: this volume needs " + std::to_string(need) +
                                   " bytes of scratch; the no-allocation entry point holds " +
                                   std::to_string(kNoAllocScratchBytes - kNoAllocStatusBytes) +
                                   " (use fg_geodesic3d_f32 with a workspace)");
  int id = 0;
  hipError_t e = hipGetDevice(&id);
  if (e != hipSuccess) return fail(FG_EHIP, std::string("hipGetDevice: ") + hipGetErrorString(e));
  DeviceState &ds = device_state(id);
  unsigned char *sp = nullptr;
  if ((rc = noalloc_scratch(ds, &sp))) return rc;
  RunOpts opt;
  opt.no_alloc = true;
  opt.fallback_status = (int *)sp;
  return run_ops(ops, sp + kNoAllocStatusBytes, kNoAllocScratchBytes - kNoAllocStatusBytes, (hipStream_t)stream,
                 opt);
}

size_t fg_noalloc_scratch_bytes(void) { return kNoAllocScratchBytes - kNoAllocStatusBytes; }

static int to_specs(const fg_instance *inst, int n, std::vector<OpSpec> &ops) {
  if (n < 0 || (n > 0 && !inst)) return fail(FG_EINVAL, "bad instance array");
  for (int i = 0; i < n; ++i) {
    const fg_instance &f = inst[i];
    if (f.is2d && f.D != 1) return fail(FG_EINVAL, "2D instance must have D == 1");
    if (f.is2d)
      ops.push_back(spec2d(f.image, f.mask, f.out, f.C, f.Cm, f.H, f.W, f.op, f.iterations, f.v, f.l_grad,
                           f.l_eucl, f.theta));
    else
      ops.push_back(spec3d(f.image, f.mask, f.out, f.C, f.Cm, f.D, f.H, f.W, f.spacing, f.op, f.iterations,
                           f.v, f.l_grad, f.l_eucl, f.theta));
    ops.back().imageT = f.imageT;
  }
  return FG_OK;
}

size_t fg_batch_workspace_bytes(const fg_instance *inst, int n) {
  std::vector<OpSpec> ops;
  if (to_specs(inst, n, ops)) return 0;
  return workspace_for(ops, current_halo());
}

int fg_geodesic_batch_f32(const fg_instance *inst, int n, void *workspace, size_t workspace_bytes,
                          void *stream) {
  std::vector<OpSpec> ops;
  int rc = to_specs(inst, n, ops);
  if (rc) return rc;
  if (ops.empty()) return FG_OK;
  return run_ops(ops, workspace, workspace_bytes, (hipStream_t)stream);
}

struct fg_stepper {
  Stepper impl;
};

size_t fg_stepper_blob_bytes(int max_transforms) { return stepper_blob_bytes(max_transforms); }

size_t fg_stepper_op_bytes(const fg_instance *inst, size_t *halo_bytes) {
  std::vector<OpSpec> ops;
  if (!inst || to_specs(inst, 1, ops)) return 0;
  if (halo_bytes) *halo_bytes = halo_entries_for(ops, current_halo()) * sizeof(unsigned long long);
  return op_buffers_bytes(ops[0]);
}

int fg_stepper_create(fg_stepper **out, void *halo, size_t halo_bytes, void *blob, size_t blob_bytes) {
  if (!out || !halo || !blob) return fail(FG_EINVAL, "null argument");
  if (halo_bytes < 16 || blob_bytes < 4096) return fail(FG_EWORKSPACE, "stepper buffers too small");
  int dev = 0;
  FG_HIP(hipGetDevice(&dev));
  fg_stepper *p = new (std::nothrow) fg_stepper();
  if (!p) return fail(FG_EHIP, "out of host memory");
  p->impl.device = dev;
  p->impl.halo = (unsigned long long *)halo;
  p->impl.halo_cap = (halo_bytes - 8) / sizeof(unsigned long long);
  p->impl.blob = (unsigned char *)blob;
  p->impl.blob_cap = blob_bytes;
  *out = p;
  return FG_OK;
}

int fg_stepper_add(fg_stepper *s, const fg_instance *inst, void *workspace, size_t workspace_bytes,
                   int *ticket) {
  if (!s || !inst || !ticket) return fail(FG_EINVAL, "null argument");
  std::vector<OpSpec> ops;
  int rc = to_specs(inst, 1, ops);
  if (rc) return rc;
  return stepper_add(s->impl, ops[0], workspace, workspace_bytes, ticket);
}

int fg_stepper_step(fg_stepper *s, void *stream, int *done, int cap, int *ndone) {
  if (!s || !ndone || (cap > 0 && !done)) return fail(FG_EINVAL, "null argument");
  if (cap < (int)s->impl.ops.size()) return fail(FG_EINVAL, "done array smaller than the active transforms");
  return stepper_step(s->impl, (hipStream_t)stream, done, cap, ndone);
}

int fg_stepper_active(const fg_stepper *s) { return s ? (int)s->impl.ops.size() : 0; }

void fg_stepper_destroy(fg_stepper *s) { delete s; }

int fg_set_voxels_f32(const fg_voxel_set *items, int n, void *stream) {
  if (n < 0 || (n > 0 && !items)) return fail(FG_EINVAL, "bad voxel array");
  for (int i = 0; i < n; ++i)
    if (!items[i].dst || items[i].index < 0) return fail(FG_EINVAL, "voxel " + std::to_string(i) + ": bad target");
  for (int i0 = 0; i0 < n; i0 += kVoxelsPerLaunch) {
    VoxelBatch b;
    const int k = std::min(kVoxelsPerLaunch, n - i0);
    std::memcpy(b.it, items + i0, (size_t)k * sizeof(fg_voxel_set));
    hipLaunchKernelGGL(set_voxels_kernel, dim3(1), dim3(128), 0, (hipStream_t)stream, b, k);
    FG_HIP(hipGetLastError());
  }
  return FG_OK;
}

int fg_tuning_set(const fg_tuning *t) {
  if (!t) return fail(FG_EINVAL, "null tuning");
  if (t->engine < 0 || t->engine > 1) return fail(FG_EINVAL, "engine must be 0 or 1");
  if (t->halo < 0 || t->halo > kMaxK) return fail(FG_EINVAL, "halo must be 0 (automatic) or in [1, 30]");
  if (t->max_groups < 0) return fail(FG_EINVAL, "max_groups must be >= 0");
  if (t->shape < 0 || t->shape >= kNumShapes) return fail(FG_EINVAL, "shape must be in [0, 7]");
  g_tune = *t;
  g_shape = t->shape == 0 ? kAutoShape : t->shape;
  return FG_OK;
}

int fg_tuning_get(fg_tuning *t) {
  if (!t) return fail(FG_EINVAL, "null tuning");
  *t = g_tune;
  return FG_OK;
}

int fg_profile_enable(int on) {
  g_prof_on = (on & 1) != 0;
  g_prof_xhist = (on & 2) != 0;
  if (g_prof_xhist) {  // start a new exchange-wait histogram on the current device
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return fail(FG_EHIP, std::string("hipGetDevice: ") + hipGetErrorString(e));
    unsigned *h = nullptr;
    const int rc = exchange_histogram_impl(dev, &h);
    if (rc) return rc;
    e = hipMemset(h, 0, kXHistBuckets * sizeof(unsigned));
    if (e != hipSuccess) return fail(FG_EHIP, std::string("hipMemset: ") + hipGetErrorString(e));
  }
  return FG_OK;
}

int fg_profile_exchange_histogram(unsigned *counts, int n) {
  if (!counts || n < kXHistBuckets) return fail(FG_EINVAL, "need room for 32 buckets");
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail(FG_EHIP, std::string("hipGetDevice: ") + hipGetErrorString(e));
  unsigned *h = nullptr;
  const int rc = exchange_histogram_impl(dev, &h);
  if (rc) return rc;
  e = hipDeviceSynchronize();
  if (e == hipSuccess) e = hipMemcpy(counts, h, kXHistBuckets * sizeof(unsigned), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return fail(FG_EHIP, std::string("histogram copy: ") + hipGetErrorString(e));
  return FG_OK;
}

int fg_profile_collect(int *launches, double *total_ms, double *alg_bytes) {
  double ms = 0, bytes = 0;
  int n = 0, rc = FG_OK;
  for (auto &p : g_prof) {
    float t = 0;
    hipError_t e = hipEventSynchronize(p.stop);
    if (e == hipSuccess) e = hipEventElapsedTime(&t, p.start, p.stop);
    if (e != hipSuccess && rc == FG_OK) rc = fail(FG_EHIP, std::string("profile events: ") + hipGetErrorString(e));
    (void)hipEventDestroy(p.start);
    (void)hipEventDestroy(p.stop);
    ms += t;
    bytes += p.bytes;
    ++n;
  }
  g_prof.clear();
  if (launches) *launches = n;
  if (total_ms) *total_ms = ms;
  if (alg_bytes) *alg_bytes = bytes;
  return rc;
}

int fg_status(void *stream) {
  int id = 0;
  hipError_t e = hipGetDevice(&id);
  if (e != hipSuccess) return fail(FG_EHIP, std::string("hipGetDevice: ") + hipGetErrorString(e));
  e = hipStreamSynchronize((hipStream_t)stream);
  if (e != hipSuccess) return fail(FG_EHIP, std::string("hipStreamSynchronize: ") + hipGetErrorString(e));
  DeviceState &ds = device_state(id);
  int st = take_status(ds);
  if (ds.scratch) {  // the no-allocation entry point's word, used before a pinned one existed
    int w = 0;
    e = hipMemcpy(&w, ds.scratch, sizeof w, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return fail(FG_EHIP, std::string("hipMemcpy: ") + hipGetErrorString(e));
    if (w != 0) {
      const int zero = 0;
      e = hipMemcpy(ds.scratch, &zero, sizeof zero, hipMemcpyHostToDevice);
      if (e != hipSuccess) return fail(FG_EHIP, std::string("hipMemcpy: ") + hipGetErrorString(e));
      if (st == FG_OK) st = w;
    }
  }
  return status_error(st, "status");
}

int fg_pending_status(int device) {
  if (device < 0) return fail(FG_EINVAL, "device must be >= 0");
  return take_status(device_state(device));
}

int fg_inject_status(int device, int code) {
  if (device < 0) return fail(FG_EINVAL, "device must be >= 0");
  device_state(device).injected.store(code);
  return FG_OK;
}

int fg_set_graphs(int on) {
  g_graph_on = on != 0;
  return FG_OK;
}

int fg_set_wait_limit(int spins) {
  if (spins < 0) return fail(FG_EINVAL, "wait limit must be >= 0");
  g_wait_limit = spins;
  return FG_OK;
}

void fg_release(void) {
  int rc = fg_profile_collect(nullptr, nullptr, nullptr);
  (void)rc;
  {
    std::lock_guard<std::mutex> g(g_graph_m);
    for (auto &p : g_graph_list) destroy_graph(*p);
    g_graph_list.clear();
  }
  std::lock_guard<std::mutex> g(g_reg_m);
  for (auto &kv : g_reg) {
    DeviceState &s = *kv.second;
    std::lock_guard<std::mutex> l(s.m);
    if (s.has) (void)hipEventSynchronize(s.done);
    for (auto &sl : s.stage) {
      if (sl.has_event) {
        (void)hipEventSynchronize(sl.copied);
        (void)hipEventDestroy(sl.copied);
        sl.has_event = false;
      }
      if (sl.host) (void)hipHostFree(sl.host);
      sl.host = nullptr;
      sl.cap = 0;
    }
    if (s.has) (void)hipEventDestroy(s.done);
    s.has = false;
    if (s.status_host) (void)hipHostFree(s.status_host);
    s.status_host = s.status_dev = nullptr;
  }
}

}  // extern "C"
