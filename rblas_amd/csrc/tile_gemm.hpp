// MFMA tile machinery shared by the single-problem GEMM (gemm_mfma.hip) and
// the grouped persistent GEMM (grouped.hip). Device code only; every
// translation unit that includes it gets its own LDS stages.
//
// One workgroup (8 waves: two waves per SIMD, so one wave's barrier and LDS
// waits overlap the other's MFMAs) computes one BM x BN tile of
// C := alpha * op(A) * op(B) + beta * C (column-major), BM, BN in {64, 128},
// over a k range [kbeg, kend) (the whole k, or one split of it). Each wave
// holds (BM/32) x (BN/64) blocks of 16 x 16 MFMA accumulators, interleaved
// over the tile: wave w (on SIMD w % 4) takes the 16-row blocks w/4,
// w/4 + 2, ... and the 16-column blocks w%4, w%4 + 4, ... So the blocks a
// SIMD owns are spread over the whole tile, and on an edge tile, where the
// blocks past the matrix edge are skipped, every SIMD keeps about the same
// share of the remaining MFMAs.
//
// LDS staging. One DMA instruction moves 64 lanes x 16 bytes = 64*V elements
// (V = 16 / sizeof(T): 2 for f64, 4 for f32) into one contiguous 1 KiB LDS
// block, so a TILE x BK operand tile (BK = 8V) is TILE/8 blocks, TILE/64 per
// wave. Blocks are padded so that the four lane groups of a fragment read
// fall on different banks:
//   M/N-contiguous operand: block q holds tile k-rows q*R .. q*R+R-1
//     (R = 1024 / (TILE * sizeof(T))) as [row][TILE], lane l carrying V
//     consecutive m (or n) of one row; block stride 1024 + 128/V bytes. A
//     k-row longer than a block (f64, TILE = 256) spans BPR = TILE * 8 /
//     1024 consecutive blocks, stride 1024 + 128 bytes (the lane groups of a
//     fragment read, 2 k apart, then land 32 banks apart).
//   K-contiguous operand: block q holds the 8 m (or n) q*8 .. q*8+7, each as
//     its BK k (128 bytes = eight 16-byte chunks of V k). Lanes 8r .. 8r+7
//     fetch the eight chunks of one column, so each instruction reads whole
//     cache lines; chunk c of column mn is stored in slot c ^ (mn % 8) of
//     that column's 128 bytes, which keeps 16-lane fragment reads free of
//     bank conflicts. Block stride 1040 bytes.
// MFMA k-steps are assigned so that a lane's V consecutive k of the
// K-contiguous layout feed V consecutive MFMAs: in k-step s = V*t + p the
// lanes of group g = lane/16 use k = 4*V*t + V*g + p. Both operands follow
// the same assignment, so every k is used exactly once, and the order in
// which each C entry accumulates its k does not depend on BM or BN.
//
// A partial last k-tile is staged through registers with per-element masks
// so that every k >= kend slot holds an exact zero in both operands. Rows or
// columns past the m/n edge may hold any value: they only reach C entries
// outside the matrix, which the epilogue does not store.
//
// The MFMA is fed op(B)^T as its first operand and op(A)^T as its second, so
// the accumulator's lane index runs down a column of C: 16 consecutive lanes
// hold 16 consecutive rows of one C column and the epilogue writes runs of
// column-major C.
#pragma once
#include "fma.hpp"
#include "kernels.hpp"

namespace rblas {
namespace {

typedef __attribute__((address_space(3))) void *lds_void_ptr;

template <typename T>
struct Mfma;

template <>
struct Mfma<double> {
  typedef double acc_t __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ acc_t mma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  // Accumulator element r of lane l is row (l/16) + 4r of the 16 x 16 block
  // (measured on gfx950: the f64 form interleaves rows).
  static __device__ __forceinline__ int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};

template <>
struct Mfma<float> {
  typedef float acc_t __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ acc_t mma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  // Accumulator element r of lane l is row 4(l/16) + r of the 16 x 16 block.
  static __device__ __forceinline__ int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

constexpr int kWaves = 8;                // 2 x 4 over the C tile, two per SIMD
constexpr int kThreads = 64 * kWaves;
// A translation unit may set, before including this header, the LDS
// pipeline depth (RBLAS_TILE_STAGES, 3 or 4) and the largest tile side of
// the B operand (RBLAS_TILE_MAX_B, 128 or 256): the LDS stages below are
// sized from them. Defaults: 4 stages, 128 x 128 tiles.
#ifndef RBLAS_TILE_STAGES
#define RBLAS_TILE_STAGES 4
#endif
#ifndef RBLAS_TILE_MAX_B
#define RBLAS_TILE_MAX_B 128
#endif
constexpr int kMaxTile = 128;  // largest C tile side along m (and n unless RBLAS_TILE_MAX_B)
constexpr int kStages = RBLAS_TILE_STAGES;  // LDS pipeline depth
static_assert(kStages == 3 || kStages == 4, "LDS pipeline depth");
constexpr int kOperandBytes = (kMaxTile / 8) * (1024 + 64);  // largest operand stage
constexpr int kOperandBytesB =
    (RBLAS_TILE_MAX_B / 8) * (RBLAS_TILE_MAX_B > 128 ? 1024 + 128 : 1024 + 64);
constexpr uint32_t kMasked = 0x80000000u;  // buffer offset past num_records: reads 0

template <typename T>
struct Geo {
  static constexpr int V = 16 / (int)sizeof(T);   // elements per lane per DMA
  static constexpr int BK = 8 * V;                 // k per tile: 16 (f64) or 32 (f32)
  static constexpr int NT = BK / (4 * V);          // 2 groups of V k-steps per tile
};

// Byte stride between DMA blocks of one operand tile.
template <typename T, bool MN_MAJOR, int TILE>
constexpr int block_bytes() {
  if (!MN_MAJOR) return 1024 + 16;
  return TILE * (int)sizeof(T) > 1024 ? 1024 + 128 : 1024 + 128 / Geo<T>::V;
}

// LDS: kStages stages of (A tile, B tile), one distinct object per stage so
// the compiler can tell the stage being read from the stages being filled.
__shared__ __attribute__((aligned(16))) char g_lds_a0[kOperandBytes];
__shared__ __attribute__((aligned(16))) char g_lds_a1[kOperandBytes];
__shared__ __attribute__((aligned(16))) char g_lds_a2[kOperandBytes];
__shared__ __attribute__((aligned(16))) char g_lds_b0[kOperandBytesB];
__shared__ __attribute__((aligned(16))) char g_lds_b1[kOperandBytesB];
__shared__ __attribute__((aligned(16))) char g_lds_b2[kOperandBytesB];
#if RBLAS_TILE_STAGES > 3
__shared__ __attribute__((aligned(16))) char g_lds_a3[kOperandBytes];
__shared__ __attribute__((aligned(16))) char g_lds_b3[kOperandBytesB];
#endif

template <int S>
__device__ __forceinline__ char *lds_a() {
  static_assert(S < kStages, "stage");
  if constexpr (S == 0) return g_lds_a0;
  else if constexpr (S == 1) return g_lds_a1;
  else if constexpr (S == 2) return g_lds_a2;
#if RBLAS_TILE_STAGES > 3
  else return g_lds_a3;
#endif
}
template <int S>
__device__ __forceinline__ char *lds_b() {
  static_assert(S < kStages, "stage");
  if constexpr (S == 0) return g_lds_b0;
  else if constexpr (S == 1) return g_lds_b1;
  else if constexpr (S == 2) return g_lds_b2;
#if RBLAS_TILE_STAGES > 3
  else return g_lds_b3;
#endif
}

// Raw buffer resource over [base, base + bytes), bytes clamped to 2^31 - 1 so
// that kMasked is always out of range. The host guarantees every in-tile
// offset is below that (gemm_fits).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void *base, uint64_t bytes) {
  const int n = bytes >= 0x7fffffffull ? 0x7fffffff : (int)bytes;
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, n, 0x00020000);
}

template <typename T>
__device__ __forceinline__ T buffer_load(__amdgpu_buffer_rsrc_t r, uint32_t off);
template <>
__device__ __forceinline__ double buffer_load<double>(__amdgpu_buffer_rsrc_t r, uint32_t off) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}
template <>
__device__ __forceinline__ float buffer_load<float>(__amdgpu_buffer_rsrc_t r, uint32_t off) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}

template <typename T>
__device__ __forceinline__ void buffer_store(__amdgpu_buffer_rsrc_t r, uint32_t off, T v);
template <>
__device__ __forceinline__ void buffer_store<double>(__amdgpu_buffer_rsrc_t r, uint32_t off,
                                                     double v) {
  typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, off, 0, 0);
}
template <>
__device__ __forceinline__ void buffer_store<float>(__amdgpu_buffer_rsrc_t r, uint32_t off,
                                                    float v) {
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, off, 0, 0);
}

// Byte offset of tile element (k, mn) inside one operand's LDS tile.
template <typename T, bool MN_MAJOR, int TILE>
__device__ __forceinline__ int lds_byte(int k, int mn) {
  typedef Geo<T> G;
  constexpr int bb = block_bytes<T, MN_MAJOR, TILE>();
  if constexpr (MN_MAJOR) {
    if constexpr (TILE * (int)sizeof(T) <= 1024) {
      constexpr int R = 1024 / (TILE * (int)sizeof(T));
      return (k / R) * bb + ((k % R) * TILE + mn) * (int)sizeof(T);
    } else {
      constexpr int per = 1024 / (int)sizeof(T), bpr = TILE / per;  // elements, blocks a k-row
      return (k * bpr + mn / per) * bb + (mn % per) * (int)sizeof(T);
    }
  }
  const int r = mn % 8;
  return (mn / 8) * bb + r * 128 + (((k / G::V) ^ r) * 16) + (k % G::V) * (int)sizeof(T);
}

// One operand (op(A) as m x k or op(B)^T as n x k) of the tile loop.
template <typename T, bool MN_MAJOR, int TILE>
struct Operand {
  typedef Geo<T> G;
  static constexpr int NB = TILE / (8 * kWaves);  // DMA blocks per wave per k-tile
  const T *x;
  const T *end;  // one past the last element the stored operand may touch
  int64_t ld;
  int mn0, lim;      // tile origin along m/n, valid extent along m/n
  uint32_t off[NB];  // this lane's byte offset for each of its wave's DMA blocks
  int block0;        // first DMA block owned by this wave

  __device__ __forceinline__ void init(const T *x_, const T *end_, int64_t ld_, int mn0_,
                                       int mn_total, int wave, int lane) {
    x = x_;
    end = end_;
    ld = ld_;
    mn0 = mn0_;
    lim = mn_total - mn0_;
    block0 = wave * NB;
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      const int q = block0 + j;
      int k, mn;
      if constexpr (MN_MAJOR) {
        const int e = lane * G::V;
        if constexpr (TILE * (int)sizeof(T) <= 1024) {
          constexpr int R = 1024 / (TILE * (int)sizeof(T));  // k-rows a block
          k = q * R + e / TILE;
          mn = e % TILE;
        } else {
          constexpr int per = 1024 / (int)sizeof(T), bpr = TILE / per;  // blocks a k-row
          k = q / bpr;
          mn = (q % bpr) * per + e;
        }
      } else {
        const int r = lane / 8;            // column within the block
        k = ((lane % 8) ^ r) * G::V;       // chunk stored in slot lane % 8
        mn = q * 8 + r;
      }
      const int64_t el = MN_MAJOR ? mn + (int64_t)k * ld : k + (int64_t)mn * ld;
      off[j] = (uint32_t)(el * (int64_t)sizeof(T));
    }
  }

  __device__ __forceinline__ const T *origin(int k0) const {
    return MN_MAJOR ? x + mn0 + (int64_t)k0 * ld : x + k0 + (int64_t)mn0 * ld;
  }

  // Full k-tile: NB DMA blocks per wave straight into LDS.
  __device__ __forceinline__ void dma(char *lds, int k0) const {
    const T *o = origin(k0);
    const __amdgpu_buffer_rsrc_t rs =
        make_rsrc(o, (uint64_t)((const char *)end - (const char *)o));
#pragma unroll
    for (int j = 0; j < NB; ++j)
      __builtin_amdgcn_raw_ptr_buffer_load_lds(
          rs, (lds_void_ptr)(lds + (block0 + j) * block_bytes<T, MN_MAJOR, TILE>()), 16, off[j],
          0, 0, 0);
  }

  // Partial last k-tile: element-wise masked loads, written in the same
  // layout; k >= kend and m/n past the edge store exact zeros.
  __device__ __forceinline__ void masked(char *lds, int k0, int kend, int tid) const {
    const T *o = origin(k0);
    const __amdgpu_buffer_rsrc_t rs =
        make_rsrc(o, (uint64_t)((const char *)end - (const char *)o));
    const int krem = kend - k0;
#pragma unroll
    for (int s = 0; s < TILE * G::BK / kThreads; ++s) {
      const int e = tid + kThreads * s;
      int k, mn;
      if (MN_MAJOR) {
        mn = e % TILE;
        k = e / TILE;
      } else {
        k = e % G::BK;
        mn = e / G::BK;
      }
      const int64_t el = MN_MAJOR ? mn + (int64_t)k * ld : k + (int64_t)mn * ld;
      const bool ok = k < krem && mn < lim;
      *(T *)(lds + lds_byte<T, MN_MAJOR, TILE>(k, mn)) =
          buffer_load<T>(rs, ok ? (uint32_t)(el * (int64_t)sizeof(T)) : kMasked);
    }
  }
};

// Map tile index t of a tiles_m x tiles_n grid to (tile_m, tile_n).
// Workgroups are dispatched round-robin over the 8 XCDs, so consecutive ids
// land on different L2s; ids are first regrouped so that each XCD owns a
// contiguous run of tiles, then that run is walked in column groups of 8
// tile rows so the workgroups resident on one XCD share A and B panels in its
// L2.
__device__ __forceinline__ void tile_coords(int t, int tiles_m, int tiles_n, int &tm, int &tn) {
  constexpr int kXcds = 8;
  constexpr int kGroupM = 8;
  const int total = tiles_m * tiles_n;
  if (total % kXcds == 0) t = (t % kXcds) * (total / kXcds) + t / kXcds;
  const int per_group = kGroupM * tiles_n;
  const int group = t / per_group;
  const int first_m = group * kGroupM;
  const int gm = min(tiles_m - first_m, kGroupM);
  const int in_group = t % per_group;
  tm = first_m + in_group % gm;
  tn = in_group / gm;
}

// Triangle tiles for syrk: tile t of the tile-level triangle in row order
// (row r holds r+1 tiles): lower -> (r, c <= r), upper -> (c <= r, r).
__device__ __forceinline__ void tri_coords(int t, int tri, int &tm, int &tn) {
  int r = (int)((__builtin_sqrtf(8.f * (float)t + 1.f) - 1.f) * 0.5f);
  while ((r + 1) * (r + 2) / 2 <= t) ++r;
  while (r * (r + 1) / 2 > t) --r;
  const int c = t - r * (r + 1) / 2;
  if (tri == 2) {
    tm = r;
    tn = c;
  } else {
    tm = c;
    tn = r;
  }
}

// Workgroup barrier for the LDS pipeline: wait until at most N of this
// wave's vector-memory instructions (the DMA loads of later tiles) are
// outstanding and all its LDS stores are done, then s_barrier. Unlike
// __syncthreads(), whose release fence drains every outstanding load
// (vmcnt(0)), this keeps the later stages in flight. The empty asm statements
// stop the compiler from moving memory accesses across the barrier.
template <int N>
__device__ __forceinline__ void stage_barrier() {
  static_assert(N >= 0 && N < 64, "vmcnt range");
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_waitcnt((N & 0xf) | ((N >> 4) << 14) | 0x0070);  // vmcnt(N) lgkmcnt(0)
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
}

// One past the last element a stored rows x cols column-major operand with
// leading dimension ld can touch.
template <typename T>
__device__ __forceinline__ const T *extent_end(const T *x, int64_t rows, int64_t cols,
                                               int64_t ld) {
  return x + (cols - 1) * ld + rows;
}

// SKIP: waves skip their blocks past the matrix edge (in one k-loop copy
// per count of blocks left); false computes every block (one copy, for
// tiles whose register budget leaves no room for the copies: the 128 x 256
// tile ran 61.8 against 68.1 TFLOP/s on the 8192^3 DGEMM with them).
template <typename T, bool TA, bool TB, int BM, int BN, bool SKIP = true>
struct TileGemm {
  typedef Geo<T> G;
  typedef typename Mfma<T>::acc_t acc_t;
  static constexpr bool A_MN = !TA;   // A NoTrans is M-contiguous
  static constexpr bool B_MN = TB;    // B Trans is N-contiguous
  static constexpr int FM = BM / 32;  // 16 x 16 blocks per wave along m (row blocks 2i + wh)
  static constexpr int FN = BN / (16 * (kWaves / 2));  // ... along n (column blocks 4j + wsimd)
  static constexpr int kPerTile = BM / (8 * kWaves) + BN / (8 * kWaves);  // DMA per wave per k-tile

  Operand<T, A_MN, BM> a;
  Operand<T, B_MN, BN> b;
  acc_t acc[FM][FN];
  int kbeg, kend, ktiles, full_tiles, tid, lane, wh, wsimd;
  int nbi, nbj;  // wave-uniform: this wave's row / column blocks that reach into the matrix

  // Set up the tile at (m0, n0) over k range [kb, ke) of problem p.
  __device__ __forceinline__ void init(const GemmProblem<T> &p, int m0, int n0, int kb, int ke) {
    // Opaque copy of the thread id: in a persistent loop over tiles every
    // address derived from it is recomputed per tile instead of being hoisted
    // out of the loop for every tile variant at once (which would keep them
    // all live together).
    int t = threadIdx.x;
    asm volatile("" : "+v"(t));
    tid = t;
    lane = tid & 63;
    const int wave = tid >> 6;
    wh = wave >> 2;
    wsimd = wave & 3;
    // Blocks i < nbi (row block 2i + wh) and j < nbj (column block 4j + wsimd)
    // hold at least one row / column of C; the others are skipped.
    const int rb = ((int)p.m - m0 + 15) >> 4, cb = ((int)p.n - n0 + 15) >> 4;
    nbi = __builtin_amdgcn_readfirstlane(max(0, min(FM, (rb - wh + 1) >> 1)));
    nbj = __builtin_amdgcn_readfirstlane(max(0, min(FN, (cb - wsimd + 3) >> 2)));
    kbeg = kb;
    kend = ke;
    ktiles = (ke - kb + G::BK - 1) / G::BK;
    full_tiles = (ke - kb) / G::BK;
    // A NoTrans is stored m x k; A Trans is stored k x m.
    // B NoTrans is stored k x n; B Trans is stored n x k.
    a.init(p.a, TA ? extent_end(p.a, p.k, p.m, p.lda) : extent_end(p.a, p.m, p.k, p.lda), p.lda,
           m0, (int)p.m, wave, lane);
    b.init(p.b, TB ? extent_end(p.b, p.n, p.k, p.ldb) : extent_end(p.b, p.k, p.n, p.ldb), p.ldb,
           n0, (int)p.n, wave, lane);
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j) acc[i][j] = acc_t{0, 0, 0, 0};
  }

  // Stage k-tile kt into the LDS stage (la, lb).
  __device__ __forceinline__ void issue(int kt, char *la, char *lb) {
    if (kt >= ktiles) return;
    const int k0 = kbeg + kt * G::BK;
    if (kt < full_tiles) {
      a.dma(la, k0);
      b.dma(lb, k0);
    } else {
      a.masked(la, k0, kend, tid);
      b.masked(lb, k0, kend, tid);
    }
  }

  // Barrier after which k-tile kt is in LDS for every wave. The tiles issued
  // after kt are kt+1 .. kt+ahead, kPerTile DMA instructions per wave each;
  // vector memory instructions complete in order, so vmcnt(kPerTile * ahead)
  // means kt's have landed.
  __device__ __forceinline__ void sync_tile(int kt) {
    const int ahead = min(kStages - 2, ktiles - 1 - kt);
    if (ahead >= 2)
      stage_barrier<2 * kPerTile>();
    else if (ahead == 1)
      stage_barrier<kPerTile>();
    else
      stage_barrier<0>();
  }

  // Fragment groups [T0, T1) of the k-tile in stage S for the blocks
  // i < NI, j < NJ (the blocks of this wave that reach into the matrix;
  // NI == FM and NJ == FN on an interior tile). Compile-time counts keep the
  // code straight-line; a skipped block's accumulators stay zero and reach
  // no stored C entry.
  template <int S, int NI, int NJ, int T0 = 0, int T1 = G::NT>
  __device__ __forceinline__ void compute() {
    const char *As = lds_a<S>();
    const char *Bs = lds_b<S>();
    const int g = lane >> 4, c = lane & 15;
    typedef T vec_t __attribute__((ext_vector_type(G::V)));
#pragma unroll
    for (int t = T0; t < T1; ++t) {
      T fa[G::V][FM], fb[G::V][FN];
      const int kb = 4 * G::V * t + G::V * g;  // this lane group's first k
#pragma unroll
      for (int i = 0; i < NI; ++i) {
        const int mn = row_of(i);
        if constexpr (A_MN) {
#pragma unroll
          for (int p = 0; p < G::V; ++p)
            fa[p][i] = *(const T *)(As + lds_byte<T, true, BM>(kb + p, mn));
        } else {
          const vec_t v = *(const vec_t *)(As + lds_byte<T, false, BM>(kb, mn));
#pragma unroll
          for (int p = 0; p < G::V; ++p) fa[p][i] = v[p];
        }
      }
#pragma unroll
      for (int j = 0; j < NJ; ++j) {
        const int mn = 16 * (4 * j + wsimd) + c;
        if constexpr (B_MN) {
#pragma unroll
          for (int p = 0; p < G::V; ++p)
            fb[p][j] = *(const T *)(Bs + lds_byte<T, true, BN>(kb + p, mn));
        } else {
          const vec_t v = *(const vec_t *)(Bs + lds_byte<T, false, BN>(kb, mn));
#pragma unroll
          for (int p = 0; p < G::V; ++p) fb[p][j] = v[p];
        }
      }
#pragma unroll
      for (int p = 0; p < G::V; ++p)
#pragma unroll
        for (int i = 0; i < NI; ++i)
#pragma unroll
          for (int j = 0; j < NJ; ++j) acc[i][j] = Mfma<T>::mma(fb[p][j], fa[p][i], acc[i][j]);
    }
  }

  // Pipeline step for k-tile kt, which lives in stage S. f32 issues the DMA
  // of tile kt + kStages - 1 after its first fragment group's MFMAs, so the
  // issue's address work overlaps MFMAs already in flight (+1.4% mean over
  // the eight SGEMM 8192 layouts, measured with four stages); f64 issues it
  // right after the barrier, which measured faster there. The k order, and
  // so every result bit, is the same.
  template <int S, int NI, int NJ>
  __device__ __forceinline__ void step(int kt) {
    if (kt >= ktiles) return;
    sync_tile(kt);  // tile kt visible to all waves; the stage of kt-1 is free
    constexpr int N = (S + kStages - 1) % kStages;
    constexpr int H = sizeof(T) == 4 ? G::NT / 2 : 0;
    compute<S, NI, NJ, 0, H>();
    issue(kt + kStages - 1, lds_a<N>(), lds_b<N>());
    compute<S, NI, NJ, H, G::NT>();
  }

  template <int NI, int NJ>
  __device__ __forceinline__ void kloop() {
    for (int kt = 0; kt < ktiles; kt += kStages) {
      step<0, NI, NJ>(kt);
      step<1, NI, NJ>(kt + 1);
      step<2, NI, NJ>(kt + 2);
      if constexpr (kStages > 3) step<3, NI, NJ>(kt + 3);
    }
  }

  // The k loop for this wave's block counts nbi in [1, I], nbj == J.
  template <int I, int J>
  __device__ __forceinline__ void kloop_rows() {
    if constexpr (I > 1) {
      if (nbi < I) {
        kloop_rows<I - 1, J>();
        return;
      }
    }
    kloop<I, J>();
  }
  // ... for nbi in [1, FM], nbj in [1, J].
  template <int J>
  __device__ __forceinline__ void kloop_cols() {
    if constexpr (J > 1) {
      if (nbj < J) {
        kloop_cols<J - 1>();
        return;
      }
    }
    kloop_rows<FM, J>();
  }

  // The first kStages - 1 stages of the pipeline.
  __device__ __forceinline__ void prologue() {
    issue(0, lds_a<0>(), lds_b<0>());
    issue(1, lds_a<1>(), lds_b<1>());
    if constexpr (kStages > 3) issue(2, lds_a<2>(), lds_b<2>());
  }

  // The k loop after the prologue, unrolled by the kStages stages, in a copy for
  // each count of this wave's row and column blocks that reach into the
  // matrix (wave-uniform): all of them on an interior tile, fewer on an edge
  // tile, none for a wave wholly past the edge (it only stages and waits).
  // Every copy has the same barriers and DMA, so the waves stay in step, and
  // each is straight-line, so LDS reads overlap MFMAs as on a full tile.
  __device__ __forceinline__ void kloops() {
    if constexpr (!SKIP) {
      kloop<FM, FN>();
    } else {
      if (nbi == FM && nbj == FN)
        kloop<FM, FN>();
      else if (nbi == 0 || nbj == 0)
        kloop<0, 0>();
      else
        kloop_cols<FN>();
    }
  }

  __device__ __forceinline__ void mainloop() {
    prologue();
    kloops();
  }

  // Row of C (within the tile) held by this lane in block row i; column of
  // element r of block column j.
  __device__ __forceinline__ int row_of(int i) const {
    return 16 * (2 * i + wh) + (lane & 15);
  }
  __device__ __forceinline__ int col_of(int j, int r) const {
    return 16 * (4 * j + wsimd) + Mfma<T>::row(lane, r);
  }

  // C := alpha * acc + beta * C on the tile at (m0, n0); beta == 0 never
  // reads C; tri 1 / 2 keeps only the upper / lower triangle. C is accessed
  // through a buffer resource based at the tile's origin: each lane has one
  // byte offset, each element adds a wave-uniform offset, and an element
  // outside the matrix or the triangle uses an offset past num_records, so
  // its load returns 0 and its store is dropped -- no branches and no
  // per-element 64-bit address arithmetic (gemm_fits keeps every in-tile
  // offset of C below 2^31). The C loads of kEpiRows block rows are issued
  // together before their stores.
  static constexpr int kEpiRows = FM * FN * 4 <= 32 ? FM : FM / 2;

  __device__ __forceinline__ bool c_in(const GemmProblem<T> &p, int gi, int gj) const {
    return gi < (int)p.m && gj < (int)p.n &&
           !((p.tri == 1 && gi > gj) || (p.tri == 2 && gi < gj));
  }

  __device__ __forceinline__ void epilogue(const GemmProblem<T> &p, int m0, int n0) const {
    const T alpha = p.alpha, beta = p.beta;
    const bool read_c = beta != T(0);  // uniform
    const int64_t ldc = p.ldc;
    const T *c0 = p.c + m0 + (int64_t)n0 * ldc;
    const uint64_t span = (uint64_t)((p.n - 1 - n0) * ldc + (p.m - m0)) * sizeof(T);
    const __amdgpu_buffer_rsrc_t rs = make_rsrc(c0, span);
    // Lane part of the offset: row row_of(0), column col_of(0, 0).
    const uint32_t lane_byte = (uint32_t)((row_of(0) + (int64_t)col_of(0, 0) * ldc) * sizeof(T));
#pragma unroll
    for (int i0 = 0; i0 < FM; i0 += kEpiRows) {
      T cv[kEpiRows][FN][4];
      if (read_c) {
#pragma unroll
        for (int ii = 0; ii < kEpiRows; ++ii)
#pragma unroll
          for (int j = 0; j < FN; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r)
              cv[ii][j][r] = buffer_load<T>(rs, elem_off(p, m0, n0, i0 + ii, j, r, lane_byte));
      }
#pragma unroll
      for (int ii = 0; ii < kEpiRows; ++ii)
#pragma unroll
        for (int j = 0; j < FN; ++j)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const T v = acc[i0 + ii][j][r];
            const T out = read_c ? fma_t(alpha, v, beta * cv[ii][j][r]) : alpha * v;
            buffer_store<T>(rs, elem_off(p, m0, n0, i0 + ii, j, r, lane_byte), out);
          }
    }
  }

  // Byte offset of this lane's element (i, j, r) from the tile origin, or
  // kMasked when the element is outside C or the kept triangle.
  __device__ __forceinline__ uint32_t elem_off(const GemmProblem<T> &p, int m0, int n0, int i,
                                               int j, int r, uint32_t lane_byte) const {
    const int di = row_of(i) - row_of(0), dj = col_of(j, r) - col_of(0, 0);  // uniform
    const uint32_t u = (uint32_t)((di + (int64_t)dj * p.ldc) * sizeof(T));
    return c_in(p, m0 + row_of(i), n0 + col_of(j, r)) ? lane_byte + u : kMasked;
  }

  static constexpr int kAccPerThread = FM * FN * 4;

  // Split-k: store this split's raw accumulators in slot `slot` of `ws`
  // (kAccPerThread * kThreads values per slot, in register order, so the
  // same thread of any workgroup running this tile reads them back). Relaxed
  // agent-scope atomic stores: written through to the level coherent across
  // XCDs, so no L2 writeback is needed to hand them over.
  __device__ __forceinline__ void store_partial(T *ws, int64_t slot) const {
    T *base = ws + slot * (int64_t)(kAccPerThread * kThreads) + tid;
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r)
          __hip_atomic_store(base + ((i * FN + j) * 4 + r) * kThreads, acc[i][j][r],
                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }

  // acc := slot slot0 + slot slot0+1 + ... + slot slot0+splits-1, added in
  // that order (this split's own slot is read back too): a fixed order, so
  // the result does not depend on which workgroup finishes last. Splits are
  // the outer loop, so each pass issues all of its loads at once. Relaxed
  // agent-scope atomic loads, served from the level the stores reached.
  __device__ __forceinline__ void reduce_partials(T *ws, int64_t slot0, int splits) {
    const int64_t stride = (int64_t)(kAccPerThread * kThreads);
    for (int s = 0; s < splits; ++s) {
      T *slot = ws + (slot0 + s) * stride + tid;
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const T v = __hip_atomic_load(slot + ((i * FN + j) * 4 + r) * kThreads,
                                          __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            acc[i][j][r] = s == 0 ? v : acc[i][j][r] + v;
          }
    }
  }
};

// Split-k hand-off for one tile: every split stores its partial (agent-scope
// atomic stores), waits until those stores have completed, and once every
// wave has (barrier) counts itself in `counter` (agent-scope atomic); the
// split that counts last reduces all partials in split order (agent-scope
// atomic loads, issued after its count returned) and applies the epilogue,
// then resets the counter to 0 for the next use. Returns true in the
// workgroup that must run the epilogue. An agent-scope release / acquire
// fence pair would also order this, but writes back and then invalidates the
// XCD's whole L2 (the operand panels the other tiles are reading from it).
template <typename TG, typename T>
__device__ __forceinline__ bool split_handoff(TG &tg, T *ws, int64_t slot0, int splits, int mine,
                                              int *counter) {
  __shared__ int s_last;
  tg.store_partial(ws, slot0 + mine);
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_waitcnt(0x0070);  // vmcnt(0) lgkmcnt(0): this wave's partial is stored
  asm volatile("" ::: "memory");
  __syncthreads();
  if (threadIdx.x == 0) {
    const int prev = __hip_atomic_fetch_add(counter, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    s_last = prev == splits - 1;
  }
  __syncthreads();
  if (!s_last) return false;
  tg.reduce_partials(ws, slot0, splits);
  if (threadIdx.x == 0)
    __hip_atomic_store(counter, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return true;
}

}  // namespace
}  // namespace rblas
