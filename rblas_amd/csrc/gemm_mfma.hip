// MFMA tile GEMM for gfx950 (MI355X): C := alpha * op(A) * op(B) + beta * C,
// column-major, f32 (v_mfma_f32_16x16x4f32) and f64 (v_mfma_f64_16x16x4f64).
//
// This replaces rblas's Goto/BLIS pipeline (kernel/gemm.rs:185-474 and the
// AVX2 microkernels gemm_avx2.rs:26-155):
//   - pack_a_panel / pack_b_panel -> the buffer unit's direct global->LDS
//     loads (`buffer_load_dwordx4 ... lds`), four LDS stages in flight, one
//     barrier per k-tile, no staging registers;
//   - the MR x NR register microtile -> a 64 x 64 per-wave tile of 4 x 4
//     16 x 16 MFMA accumulators;
//   - the jc/pc/ic loops -> one workgroup per 128 x 128 C tile, the whole k
//     range accumulated in registers (no KC re-blocking of C);
//   - materialize (level3.rs:179-222) -> none: each operand is staged from
//     its stored orientation (M/N-contiguous or K-contiguous).
//
// LDS staging. One DMA instruction moves 64 lanes x 16 bytes = 64*V elements
// (V = 16 / sizeof(T): 2 for f64, 4 for f32) into one contiguous LDS block,
// so a 128 x kBK operand tile (kBK = 8V) is 16 blocks. Blocks are padded so
// that the four lane groups of a fragment read fall on different banks:
//   M/N-contiguous operand: block q holds tile k-rows q*R .. q*R+R-1 (R = V/2)
//     as [row][128], lane l carrying V consecutive m (or n) of one row;
//     block stride 1024 + 128/V bytes.
//   K-contiguous operand: block q holds the 8 m (or n) q*8 .. q*8+7, each as
//     its kBK k (128 bytes = eight 16-byte chunks of V k). Lanes 8r .. 8r+7
//     fetch the eight chunks of one column, so each instruction reads whole
//     cache lines; chunk c of column mn is stored in slot c ^ (mn % 8) of
//     that column's 128 bytes, which keeps 16-lane fragment reads free of
//     bank conflicts. Block stride 1040 bytes.
// MFMA k-steps are assigned so that a lane's V consecutive k of the
// K-contiguous layout feed V consecutive MFMAs: in k-step s = V*t + p the
// lanes of group g = lane/16 use k = 4*V*t + V*g + p. Both operands follow
// the same assignment, so every k is used exactly once.
//
// A partial last k-tile is staged through registers with per-element masks
// so that every k >= K slot holds an exact zero in both operands. Rows or
// columns past the m/n edge may hold any value: they only reach C entries
// outside the matrix, which the epilogue does not store.
//
// The MFMA is fed op(B)^T as its first operand and op(A)^T as its second, so
// the accumulator's lane index runs down a column of C: 16 consecutive lanes
// hold 16 consecutive rows of one C column and the epilogue writes runs of
// column-major C.
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

constexpr int kThreads = 256;  // 4 waves of 64, 2 x 2 over the C tile
constexpr int kTile = 128;     // C tile rows and columns
constexpr int kWave = 64;      // per-wave C tile (rows and columns)
constexpr int kFrag = kWave / 16;  // 4 x 4 MFMA blocks per wave
constexpr int kStages = 4;     // LDS pipeline depth
constexpr int kBlocks = 16;    // DMA blocks per operand tile
constexpr int kOperandBytes = kBlocks * (1024 + 64);  // room for the largest block stride
constexpr uint32_t kMasked = 0x80000000u;  // buffer offset past num_records: reads 0

template <typename T>
struct Geo {
  static constexpr int V = 16 / (int)sizeof(T);   // elements per lane per DMA
  static constexpr int BK = 8 * V;                 // k per tile: 16 (f64) or 32 (f32)
  static constexpr int R = V / 2;                  // k-rows per block (M/N-contiguous)
  static constexpr int NT = BK / (4 * V);          // 2 groups of V k-steps per tile
};

// Byte stride between DMA blocks of one operand tile.
template <typename T, bool MN_MAJOR>
constexpr int block_bytes() {
  return MN_MAJOR ? 1024 + 128 / Geo<T>::V : 1024 + 16;
}

// LDS: kStages stages of (A tile, B tile), one distinct object per stage so
// the compiler can tell the stage being read from the stages being filled.
__shared__ __attribute__((aligned(16))) char g_lds_a0[kOperandBytes];
__shared__ __attribute__((aligned(16))) char g_lds_a1[kOperandBytes];
__shared__ __attribute__((aligned(16))) char g_lds_a2[kOperandBytes];
__shared__ __attribute__((aligned(16))) char g_lds_a3[kOperandBytes];
__shared__ __attribute__((aligned(16))) char g_lds_b0[kOperandBytes];
__shared__ __attribute__((aligned(16))) char g_lds_b1[kOperandBytes];
__shared__ __attribute__((aligned(16))) char g_lds_b2[kOperandBytes];
__shared__ __attribute__((aligned(16))) char g_lds_b3[kOperandBytes];

template <int S>
__device__ __forceinline__ char *lds_a() {
  if constexpr (S == 0) return g_lds_a0;
  else if constexpr (S == 1) return g_lds_a1;
  else if constexpr (S == 2) return g_lds_a2;
  else return g_lds_a3;
}
template <int S>
__device__ __forceinline__ char *lds_b() {
  if constexpr (S == 0) return g_lds_b0;
  else if constexpr (S == 1) return g_lds_b1;
  else if constexpr (S == 2) return g_lds_b2;
  else return g_lds_b3;
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

// Byte offset of tile element (k, mn) inside one operand's LDS tile.
template <typename T, bool MN_MAJOR>
__device__ __forceinline__ int lds_byte(int k, int mn) {
  typedef Geo<T> G;
  constexpr int bb = block_bytes<T, MN_MAJOR>();
  if (MN_MAJOR) return (k / G::R) * bb + ((k % G::R) * kTile + mn) * (int)sizeof(T);
  const int r = mn % 8;
  return (mn / 8) * bb + r * 128 + (((k / G::V) ^ r) * 16) + (k % G::V) * (int)sizeof(T);
}

// One operand (op(A) as m x k or op(B)^T as n x k) of the tile loop.
template <typename T, bool MN_MAJOR>
struct Operand {
  typedef Geo<T> G;
  const T *x;
  const T *end;  // one past the last element the stored operand may touch
  int64_t ld;
  int mn0, lim;     // tile origin along m/n, valid extent along m/n
  uint32_t off[4];  // this lane's byte offset for each of its wave's 4 DMA blocks
  int block0;       // first DMA block owned by this wave

  __device__ __forceinline__ void init(const T *x_, const T *end_, int64_t ld_, int mn0_,
                                       int mn_total, int wave, int lane) {
    x = x_;
    end = end_;
    ld = ld_;
    mn0 = mn0_;
    lim = mn_total - mn0_;
    block0 = wave * 4;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int q = block0 + j;
      int k, mn;
      if (MN_MAJOR) {
        const int e = lane * G::V;
        k = q * G::R + e / kTile;
        mn = e % kTile;
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

  // Full k-tile: 4 DMA blocks per wave straight into LDS.
  __device__ __forceinline__ void dma(char *lds, int k0) const {
    const T *o = origin(k0);
    const __amdgpu_buffer_rsrc_t rs =
        make_rsrc(o, (uint64_t)((const char *)end - (const char *)o));
#pragma unroll
    for (int j = 0; j < 4; ++j)
      __builtin_amdgcn_raw_ptr_buffer_load_lds(
          rs, (lds_void_ptr)(lds + (block0 + j) * block_bytes<T, MN_MAJOR>()), 16, off[j], 0, 0,
          0);
  }

  // Partial last k-tile: element-wise masked loads, written in the same
  // layout; k >= K and m/n past the edge store exact zeros.
  __device__ __forceinline__ void masked(char *lds, int k0, int K, int tid) const {
    const T *o = origin(k0);
    const __amdgpu_buffer_rsrc_t rs =
        make_rsrc(o, (uint64_t)((const char *)end - (const char *)o));
    const int krem = K - k0;
#pragma unroll
    for (int s = 0; s < kTile * G::BK / kThreads; ++s) {
      const int e = tid + kThreads * s;
      int k, mn;
      if (MN_MAJOR) {
        mn = e % kTile;
        k = e / kTile;
      } else {
        k = e % G::BK;
        mn = e / G::BK;
      }
      const int64_t el = MN_MAJOR ? mn + (int64_t)k * ld : k + (int64_t)mn * ld;
      const bool ok = k < krem && mn < lim;
      *(T *)(lds + lds_byte<T, MN_MAJOR>(k, mn)) =
          buffer_load<T>(rs, ok ? (uint32_t)(el * (int64_t)sizeof(T)) : kMasked);
    }
  }
};

// Map the linear workgroup id to a (tile_m, tile_n) pair. Workgroups are
// dispatched round-robin over the 8 XCDs, so consecutive ids land on
// different L2s; we first regroup ids so that each XCD owns a contiguous run
// of tiles, then walk that run in column groups of kGroupM tile rows so the
// workgroups resident on one XCD share A and B panels in its L2.
__device__ __forceinline__ void tile_coords(int tiles_m, int tiles_n, int &tm, int &tn) {
  constexpr int kXcds = 8;
  constexpr int kGroupM = 8;
  const int total = tiles_m * tiles_n;
  int t = blockIdx.x;
  if (total % kXcds == 0) t = (t % kXcds) * (total / kXcds) + t / kXcds;
  const int per_group = kGroupM * tiles_n;
  const int group = t / per_group;
  const int first_m = group * kGroupM;
  const int gm = min(tiles_m - first_m, kGroupM);
  const int in_group = t % per_group;
  tm = first_m + in_group % gm;
  tn = in_group / gm;
}

// Triangle tiles for syrk: workgroup t takes the t-th tile of the tile-level
// triangle in row order (row r holds r+1 tiles): lower -> (r, c <= r),
// upper -> (c <= r, r).
__device__ __forceinline__ void tri_coords(int tri, int &tm, int &tn) {
  const int t = blockIdx.x;
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

template <typename T, bool TA, bool TB>
struct TileGemm {
  typedef Geo<T> G;
  typedef typename Mfma<T>::acc_t acc_t;
  static constexpr bool A_MN = !TA;   // A NoTrans is M-contiguous
  static constexpr bool B_MN = TB;    // B Trans is N-contiguous
  static constexpr int kPerTile = 8;  // DMA instructions per wave per k-tile

  Operand<T, A_MN> a;
  Operand<T, B_MN> b;
  acc_t acc[kFrag][kFrag];
  int K, ktiles, full_tiles, tid, lane, wm, wn;

  // Stage k-tile kt into LDS stage S (DMA when full, masked loads otherwise).
  template <int S>
  __device__ __forceinline__ void issue(int kt) {
    if (kt >= ktiles) return;
    if (kt < full_tiles) {
      a.dma(lds_a<S>(), kt * G::BK);
      b.dma(lds_b<S>(), kt * G::BK);
    } else {
      a.masked(lds_a<S>(), kt * G::BK, K, tid);
      b.masked(lds_b<S>(), kt * G::BK, K, tid);
    }
  }

  // Barrier after which k-tile kt is in LDS for every wave. The tiles issued
  // after kt are kt+1 .. kt+ahead, kPerTile DMA instructions per wave each;
  // vector memory instructions complete in order, so vmcnt(kPerTile * ahead)
  // means kt's have landed. (If one of the later tiles was the masked tail,
  // its register loads were consumed by its own LDS stores, which already
  // waited for everything issued before them.)
  __device__ __forceinline__ void sync_tile(int kt) {
    const int ahead = min(kStages - 2, ktiles - 1 - kt);
    if (ahead >= 2)
      stage_barrier<2 * kPerTile>();
    else if (ahead == 1)
      stage_barrier<kPerTile>();
    else
      stage_barrier<0>();
  }

  template <int S>
  __device__ __forceinline__ void compute() {
    const char *As = lds_a<S>();
    const char *Bs = lds_b<S>();
    const int g = lane >> 4, c = lane & 15;
    typedef T vec_t __attribute__((ext_vector_type(G::V)));
#pragma unroll
    for (int t = 0; t < G::NT; ++t) {
      T fa[G::V][kFrag], fb[G::V][kFrag];
      const int kb = 4 * G::V * t + G::V * g;  // this lane group's first k
#pragma unroll
      for (int i = 0; i < kFrag; ++i) {
        const int mn = wm + 16 * i + c;
        if constexpr (A_MN) {
#pragma unroll
          for (int p = 0; p < G::V; ++p) fa[p][i] = *(const T *)(As + lds_byte<T, true>(kb + p, mn));
        } else {
          const vec_t v = *(const vec_t *)(As + lds_byte<T, false>(kb, mn));
#pragma unroll
          for (int p = 0; p < G::V; ++p) fa[p][i] = v[p];
        }
      }
#pragma unroll
      for (int j = 0; j < kFrag; ++j) {
        const int mn = wn + 16 * j + c;
        if constexpr (B_MN) {
#pragma unroll
          for (int p = 0; p < G::V; ++p) fb[p][j] = *(const T *)(Bs + lds_byte<T, true>(kb + p, mn));
        } else {
          const vec_t v = *(const vec_t *)(Bs + lds_byte<T, false>(kb, mn));
#pragma unroll
          for (int p = 0; p < G::V; ++p) fb[p][j] = v[p];
        }
      }
#pragma unroll
      for (int p = 0; p < G::V; ++p)
#pragma unroll
        for (int i = 0; i < kFrag; ++i)
#pragma unroll
          for (int j = 0; j < kFrag; ++j)
            acc[i][j] = Mfma<T>::mma(fb[p][j], fa[p][i], acc[i][j]);
    }
  }

  // Pipeline step for k-tile kt, which lives in stage S.
  template <int S>
  __device__ __forceinline__ void step(int kt) {
    if (kt >= ktiles) return;
    sync_tile(kt);  // tile kt visible to all waves; the stage of kt-1 is free
    issue<(S + kStages - 1) % kStages>(kt + kStages - 1);
    compute<S>();
  }
};

// One past the last element a stored rows x cols column-major operand with
// leading dimension ld can touch.
template <typename T>
__device__ __forceinline__ const T *extent_end(const T *x, int64_t rows, int64_t cols,
                                               int64_t ld) {
  return x + (cols - 1) * ld + rows;
}

template <typename T, bool TA, bool TB>
__global__ __launch_bounds__(kThreads) void gemm_tile_kernel(GemmProblem<T> p) {
  typedef Geo<T> G;
  typedef typename Mfma<T>::acc_t acc_t;
  TileGemm<T, TA, TB> tg;

  const int M = (int)p.m, N = (int)p.n, K = (int)p.k;
  const int tiles_m = (M + kTile - 1) / kTile;
  const int tiles_n = (N + kTile - 1) / kTile;
  int tm, tn;
  if (p.tri == 0)
    tile_coords(tiles_m, tiles_n, tm, tn);
  else
    tri_coords(p.tri, tm, tn);
  const int m0 = tm * kTile, n0 = tn * kTile;

  tg.tid = threadIdx.x;
  tg.lane = tg.tid & 63;
  const int wave = tg.tid >> 6;
  tg.wm = (wave & 1) * kWave;
  tg.wn = (wave >> 1) * kWave;
  tg.K = K;
  tg.ktiles = (K + G::BK - 1) / G::BK;
  tg.full_tiles = K / G::BK;

  // A NoTrans is stored m x k; A Trans is stored k x m.
  // B NoTrans is stored k x n; B Trans is stored n x k.
  tg.a.init(p.a, TA ? extent_end(p.a, p.k, p.m, p.lda) : extent_end(p.a, p.m, p.k, p.lda), p.lda,
            m0, M, wave, tg.lane);
  tg.b.init(p.b, TB ? extent_end(p.b, p.n, p.k, p.ldb) : extent_end(p.b, p.k, p.n, p.ldb), p.ldb,
            n0, N, wave, tg.lane);

#pragma unroll
  for (int i = 0; i < kFrag; ++i)
#pragma unroll
    for (int j = 0; j < kFrag; ++j) tg.acc[i][j] = acc_t{0, 0, 0, 0};

  tg.template issue<0>(0);
  tg.template issue<1>(1);
  tg.template issue<2>(2);
  for (int kt = 0; kt < tg.ktiles; kt += kStages) {
    tg.template step<0>(kt);
    tg.template step<1>(kt + 1);
    tg.template step<2>(kt + 2);
    tg.template step<3>(kt + 3);
  }

  // Epilogue: element (i, j) of C gets alpha*acc + beta*C; beta == 0 never
  // reads C. Lane l of block (fi, fj) owns row i = m0+wm+16fi+l%16 and
  // columns j = n0+wn+16fj+row(l, r). Each block's C values are loaded
  // together before any of them is stored.
  const T alpha = p.alpha, beta = p.beta;
  const bool read_c = beta != T(0);
  const int c16 = tg.lane & 15;
#pragma unroll
  for (int i = 0; i < kFrag; ++i) {
    const int gi = m0 + tg.wm + 16 * i + c16;
    if (gi >= M) continue;
#pragma unroll
    for (int j = 0; j < kFrag; ++j) {
      T cv[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gj = n0 + tg.wn + 16 * j + Mfma<T>::row(tg.lane, r);
        const bool in = gj < N && !((p.tri == 1 && gi > gj) || (p.tri == 2 && gi < gj));
        cv[r] = (read_c && in) ? p.c[gi + (int64_t)gj * p.ldc] : T(0);
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gj = n0 + tg.wn + 16 * j + Mfma<T>::row(tg.lane, r);
        if (gj >= N) continue;
        if ((p.tri == 1 && gi > gj) || (p.tri == 2 && gi < gj)) continue;
        const T v = tg.acc[i][j][r];
        p.c[gi + (int64_t)gj * p.ldc] = read_c ? __builtin_fma(alpha, v, beta * cv[r]) : alpha * v;
      }
    }
  }
}

template <typename T>
__global__ void scale_tri_kernel(int64_t n, int tri, T beta, T *c, int64_t ldc) {
  const int64_t total = n * n;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = idx / n, i = idx - j * n;
    if ((tri == 1 && i > j) || (tri == 2 && i < j)) continue;
    T *cp = c + i + j * ldc;
    *cp = beta == T(0) ? T(0) : beta * *cp;
  }
}

template <typename T>
__global__ void scale_kernel(int64_t m, int64_t n, T beta, T *c, int64_t ldc) {
  const int64_t total = m * n;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = idx / m, i = idx - j * m;
    T *cp = c + i + j * ldc;
    *cp = beta == T(0) ? T(0) : beta * *cp;
  }
}

}  // namespace

template <typename T>
bool gemm_fits(const GemmProblem<T> &p) {
  // Largest in-tile byte offset for either operand orientation.
  const int64_t ld = p.lda > p.ldb ? p.lda : p.ldb;
  const int64_t span = (int64_t)(kTile - 1) * ld + kTile + Geo<T>::BK;
  return span * (int64_t)sizeof(T) < 0x7fffffffll - 16;
}

template <typename T>
hipError_t launch_gemm(const GemmProblem<T> &p, hipStream_t stream) {
  if (!gemm_fits(p)) return hipErrorInvalidValue;
  int64_t tiles = ((p.m + kTile - 1) / kTile) * ((p.n + kTile - 1) / kTile);
  if (p.tri != 0) {
    if (p.m != p.n) return hipErrorInvalidValue;
    const int64_t t = (p.m + kTile - 1) / kTile;
    tiles = t * (t + 1) / 2;
  }
  dim3 grid((unsigned)tiles), block(kThreads);
  if (!p.ta && !p.tb)
    hipLaunchKernelGGL((gemm_tile_kernel<T, false, false>), grid, block, 0, stream, p);
  else if (!p.ta && p.tb)
    hipLaunchKernelGGL((gemm_tile_kernel<T, false, true>), grid, block, 0, stream, p);
  else if (p.ta && !p.tb)
    hipLaunchKernelGGL((gemm_tile_kernel<T, true, false>), grid, block, 0, stream, p);
  else
    hipLaunchKernelGGL((gemm_tile_kernel<T, true, true>), grid, block, 0, stream, p);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_scale(int64_t m, int64_t n, T beta, T *c, int64_t ldc, hipStream_t stream) {
  const int64_t total = m * n;
  int64_t blocks = (total + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL((scale_kernel<T>), dim3((unsigned)blocks), dim3(256), 0, stream, m, n, beta,
                     c, ldc);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_scale_tri(int64_t n, int tri, T beta, T *c, int64_t ldc, hipStream_t stream) {
  int64_t blocks = (n * n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL((scale_tri_kernel<T>), dim3((unsigned)blocks), dim3(256), 0, stream, n, tri,
                     beta, c, ldc);
  return hipGetLastError();
}

template hipError_t launch_scale_tri<float>(int64_t, int, float, float *, int64_t, hipStream_t);
template hipError_t launch_scale_tri<double>(int64_t, int, double, double *, int64_t, hipStream_t);
template hipError_t launch_gemm<float>(const GemmProblem<float> &, hipStream_t);
template hipError_t launch_gemm<double>(const GemmProblem<double> &, hipStream_t);
template hipError_t launch_scale<float>(int64_t, int64_t, float, float *, int64_t, hipStream_t);
template hipError_t launch_scale<double>(int64_t, int64_t, double, double *, int64_t,
                                         hipStream_t);

}  // namespace rblas
