// Split-precision SGEMM on bf16 MFMA (gfx950): each f32 operand element x is
// written as a sum of bfloat16 terms and C = alpha * op(A) * op(B) + beta * C
// is accumulated in f32 from the cross products of those terms.
//
//   BF16x3: x = h + l            (2 planes), products h*h + h*l + l*h
//   BF16x6: x = h + m + l        (3 planes), products h*h + h*m + m*h
//                                             + h*l + m*m + l*h
// (the split and product sets of Henry, Tang and Heinecke, "Leveraging the
// bfloat16 artificial intelligence datatype for higher-precision
// computations", ARITH 2019). Each term is the round-to-nearest bf16 of the
// remainder left by the previous ones; in f32 those remainders are exact, so
// h + m + l represents x to about 2^-24 relative and h + l to about 2^-16.
//
// Two kernels:
//   split_kernel   reads op(A) (m x k) or op(B)^T (n x k) from the caller's
//                  layout, transposition and leading dimension and writes the
//                  planes k-blocked by 16, zero-padded to 256 rows and to a
//                  multiple of 48 k, so the product kernel needs no masking;
//   split_gemm_kernel  256 x 256 C tile per workgroup, 8 waves (two per SIMD)
//                  of 128 x 64 as 4 x 2 v_mfma_f32_32x32x16_bf16 blocks (128
//                  accumulator registers); every plane's 16-k slice is moved global -> LDS
//                  by buffer_load_dwordx4 ... lds (3 LDS stages, one barrier per
//                  16 k) and each fragment is reused in registers by every
//                  product it takes part in (96 MFMAs per barrier for BF16x6).
#include <algorithm>

#include "kernels.hpp"

namespace rblas {
namespace {

typedef __attribute__((address_space(3))) void *lds_void_ptr;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kTile = 256;             // C tile rows and columns; plane row padding
constexpr int kWave = 128;             // per-wave C tile rows: 4 blocks of 32

// W waves per workgroup, 2 x (W / 2) over the C tile: 8 (two per SIMD, each
// 128 x 64) or 4 (one per SIMD, each 128 x 128). Both give the same bits.
template <int W>
struct Waves {
  static constexpr int kThreads = 64 * W;
  static constexpr int kWaveN = 2 * kTile / W;  // per-wave C tile columns
  static constexpr int kFN = kWaveN / 32;       // 32-column blocks per wave
};
constexpr int kBK = 16;                // k per LDS stage (one MFMA k-step); plane k padding
constexpr int kStages = 3;
constexpr int kPlaneTileBytes = kTile * kBK * 2;  // 8 KB: 256 rows x 32 B
constexpr int kMaxPlanes = 3;
constexpr int kOperandBytes = kMaxPlanes * kPlaneTileBytes;

__shared__ __attribute__((aligned(16))) char g_sa0[kOperandBytes];
__shared__ __attribute__((aligned(16))) char g_sa1[kOperandBytes];
__shared__ __attribute__((aligned(16))) char g_sa2[kOperandBytes];
__shared__ __attribute__((aligned(16))) char g_sb0[kOperandBytes];
__shared__ __attribute__((aligned(16))) char g_sb1[kOperandBytes];
__shared__ __attribute__((aligned(16))) char g_sb2[kOperandBytes];

template <int S>
__device__ __forceinline__ char *lds_a() {
  if constexpr (S == 0) return g_sa0;
  else if constexpr (S == 1) return g_sa1;
  else return g_sa2;
}
template <int S>
__device__ __forceinline__ char *lds_b() {
  if constexpr (S == 0) return g_sb0;
  else if constexpr (S == 1) return g_sb1;
  else return g_sb2;
}

// LDS byte offset of plane s, tile row r (0..255), 16-byte k-chunk c (0..1).
// Rows are 32 B; the two chunks of a row swap places when bit 2 xor bit 3 of
// the row is set. A fragment read has lanes 0..31 on rows r..r+31 at one
// chunk: unswizzled, rows 8 apart (64 banks of 4 B) or 4 apart (32 banks)
// hit the same banks, a 2-way conflict on every read; with the swap, any 16
// consecutive rows (8 for 32 banks) cover distinct banks. The DMA writes
// lane l's 16 bytes at position l of a 1 KB block, so each lane fetches the
// chunk that belongs there.
__device__ __forceinline__ int chunk_swap(int r) { return ((r >> 2) ^ (r >> 3)) & 1; }
__device__ __forceinline__ int lds_off(int s, int r, int c) {
  return s * kPlaneTileBytes + r * 32 + ((c ^ chunk_swap(r)) << 4);
}

// ----------------------------------------------------------------- split
// Source element (r, k) of op(X) (r < rows, k < K) is x[r + k*ld] when
// R_CONTIG (A NoTrans, B Trans) and x[k + r*ld] otherwise. Plane s stores
// the padded rows_pad x Kp matrix k-blocked: element (r, k) at
//   planes + s*plane_elems + ((k/16)*rows_pad + r)*16 + k%16,
// so the 16-k slice of any run of rows is contiguous (32 B per row), and a
// product-kernel stage reads whole cache lines. Padding is zero.
template <bool R_CONTIG, int PLANES>
__global__ __launch_bounds__(256) void split_kernel(const float *__restrict__ x, int64_t ld,
                                                    int rows, int K, int rows_pad, int Kp,
                                                    __bf16 *__restrict__ planes,
                                                    int64_t plane_elems) {
  __shared__ float tile[64][65];
  const int k0 = blockIdx.x * 64, r0 = blockIdx.y * 64;
  const int t = threadIdx.x;
  // All 16 loads issued before any is used, without branches (an element
  // outside the operand reads x[0] and is replaced by 0).
  float raw[16];
#pragma unroll
  for (int s = 0; s < 16; ++s) {
    const int rl = R_CONTIG ? (t & 63) : (t >> 6) + 4 * s;
    const int kl = R_CONTIG ? (t >> 6) + 4 * s : (t & 63);
    const int r = r0 + rl, k = k0 + kl;
    const bool ok = r < rows && k < K;
    const float *q = R_CONTIG ? x + r + (int64_t)k * ld : x + k + (int64_t)r * ld;
    raw[s] = *(ok ? q : x);
  }
#pragma unroll
  for (int s = 0; s < 16; ++s) {
    const int rl = R_CONTIG ? (t & 63) : (t >> 6) + 4 * s;
    const int kl = R_CONTIG ? (t >> 6) + 4 * s : (t & 63);
    tile[rl][kl] = (r0 + rl < rows && k0 + kl < K) ? raw[s] : 0.f;
  }
  __syncthreads();
  // Thread t writes row r0 + t%64, k-block k0/16 + t/64 (16 values, 32 B per plane).
  const int rl = t & 63, kbl = t >> 6;
  const int kb = k0 / 16 + kbl;
  if (kb * 16 >= Kp) return;
  bf16x8 h[2], m[2], l[2];
#pragma unroll
  for (int e = 0; e < 16; ++e) {
    const float v = tile[rl][kbl * 16 + e];
    const __bf16 hv = (__bf16)v;
    const float r1 = v - (float)hv;
    const __bf16 mv = (__bf16)r1;
    h[e >> 3][e & 7] = hv;
    m[e >> 3][e & 7] = mv;
    if (PLANES == 3) l[e >> 3][e & 7] = (__bf16)(r1 - (float)mv);
  }
  const int64_t o = ((int64_t)kb * rows_pad + r0 + rl) * 16;
  bf16x8 *p0 = (bf16x8 *)(planes + o);
  bf16x8 *p1 = (bf16x8 *)(planes + plane_elems + o);
  p0[0] = h[0];
  p0[1] = h[1];
  p1[0] = m[0];
  p1[1] = m[1];
  if (PLANES == 3) {
    bf16x8 *p2 = (bf16x8 *)(planes + 2 * plane_elems + o);
    p2[0] = l[0];
    p2[1] = l[1];
  }
}

// ----------------------------------------------------------------- gemm
struct SplitArgs {
  const __bf16 *a;  // planes of op(A): m_pad x Kp each, k-blocked by 16
  const __bf16 *b;  // planes of op(B)^T: n_pad x Kp each, k-blocked by 16
  int64_t a_plane, b_plane;  // elements per plane
  int m, n, m_pad, n_pad, Kp;
  float alpha, beta;
  float *c;
  int64_t ldc;
  int splits;        // k splits per tile (1 = none)
  int split_ktiles;  // k-tiles per split (a multiple of kStages)
  float *partials;   // splits > 1: kTile * kTile floats per (tile, split)
  int *counters;     // splits > 1: one zero-initialised int per tile
};

template <int N>
__device__ __forceinline__ void stage_barrier() {
  static_assert(N >= 0 && N < 64, "vmcnt range");
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_waitcnt((N & 0xf) | ((N >> 4) << 14) | 0x0070);  // vmcnt(N) lgkmcnt(0)
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
}

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

// LDS address-space view of a pointer into one of the stage arrays.
__device__ __forceinline__ lds_void_ptr to_lds(char *p) { return (lds_void_ptr)p; }

// Buffer resource over one 1 KB block (wave-uniform base).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t block_rsrc(const char *base) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<char *>(base), (short)0, 1024, 0x00020000);
}

template <int PL, int W>
struct SplitGemm {
  static constexpr int kWaves = W, kFN = Waves<W>::kFN;
  static constexpr int kPerWave = 8 * PL / kWaves;  // DMA blocks (1 KB = 32 rows) per operand per wave
  static constexpr int kPerTile = 2 * kPerWave;  // DMA instructions per wave per k-tile
  const char *ga[kPerWave];  // start of each of this wave's A blocks at k-block 0
  const char *gb[kPerWave];
  int64_t sa, sb;            // bytes from one k-block of a plane to the next
  f32x16 acc[4][kFN];
  int ktiles, lane, wave, wm, wn;

  __device__ __forceinline__ void init(const SplitArgs &p, int m0, int n0) {
    // Block j is 32 rows x 32 B, one contiguous kilobyte of a plane's k-block
    // (wave-uniform address); lane l carries its bytes 16l .. 16l+15 (row
    // l/2, k-chunk l%2).
#pragma unroll
    for (int j = 0; j < kPerWave; ++j) {
      const int q = wave * kPerWave + j;  // block 0 .. 8*PL-1 (wave is uniform)
      const int s = q >> 3, blk = q & 7;
      ga[j] = (const char *)(p.a + s * p.a_plane + (int64_t)(m0 + blk * 32) * 16);
      gb[j] = (const char *)(p.b + s * p.b_plane + (int64_t)(n0 + blk * 32) * 16);
    }
    sa = (int64_t)p.m_pad * 32;
    sb = (int64_t)p.n_pad * 32;
  }

  // Block q = wave*kPerWave + j of a stage lands at byte q*1024 (plane q/8,
  // rows 32*(q%8) ..). The LDS address is formed directly from the stage's
  // own array so the compiler can see that filling stage S never touches the
  // stage being read, and inserts no wait for it.
  // (A plain member taking the stage's arrays: a device builtin inside a
  // member template fails clang's host-side pass.)
  __device__ __forceinline__ void issue(int kt, char *la, char *lb) {
    // Lane l lands at LDS position l of the block (row l / 2, physical chunk
    // l % 2) and fetches the logical chunk stored there (lds_off); blocks
    // start at multiples of 32 rows, so the row within the block decides.
    const int row = lane >> 1;
    const int src = row * 32 + (((lane & 1) ^ chunk_swap(row)) << 4);
#pragma unroll
    for (int j = 0; j < kPerWave; ++j)
      __builtin_amdgcn_raw_ptr_buffer_load_lds(block_rsrc(ga[j] + kt * sa),
                                               to_lds(la + (wave * kPerWave + j) * 1024), 16,
                                               src, 0, 0, 0);
#pragma unroll
    for (int j = 0; j < kPerWave; ++j)
      __builtin_amdgcn_raw_ptr_buffer_load_lds(block_rsrc(gb[j] + kt * sb),
                                               to_lds(lb + (wave * kPerWave + j) * 1024), 16,
                                               src, 0, 0, 0);
  }


  template <int S>
  __device__ __forceinline__ void compute() {
    const char *As = lds_a<S>();
    const char *Bs = lds_b<S>();
    const int r32 = lane & 31, c = lane >> 5;
    bf16x8 fa[PL][4], fb[PL][kFN];
#pragma unroll
    for (int s = 0; s < PL; ++s) {
#pragma unroll
      for (int i = 0; i < 4; ++i) fa[s][i] = *(const bf16x8 *)(As + lds_off(s, wm + 32 * i + r32, c));
#pragma unroll
      for (int j = 0; j < kFN; ++j) fb[s][j] = *(const bf16x8 *)(Bs + lds_off(s, wn + 32 * j + r32, c));
    }
    // Smallest terms first.
#pragma unroll
    for (int pr = 0; pr < (PL == 2 ? 3 : 6); ++pr) {
      int sa, sb;
      if constexpr (PL == 2) {
        sa = pr == 0 ? 1 : 0;
        sb = pr == 1 ? 1 : 0;
      } else {
        constexpr int kA[6] = {2, 1, 0, 1, 0, 0};
        constexpr int kB[6] = {0, 1, 2, 0, 1, 0};
        sa = kA[pr];
        sb = kB[pr];
      }
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < kFN; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fb[sb][j], fa[sa][i], acc[i][j],
                                                              0, 0, 0);
    }
  }

  // Pipeline step for k-tile kt (stage S). The k extent is padded to a
  // multiple of kStages tiles and the planes carry kStages - 1 spare k-blocks,
  // so every step issues the tile kStages - 1 ahead unconditionally and
  // exactly one tile (kPerTile instructions) is always in flight behind kt:
  // the loop body has no branches for the compiler to merge across stages.
  template <int S>
  __device__ __forceinline__ void step(int kt) {
    stage_barrier<kPerTile>();
    constexpr int N = (S + kStages - 1) % kStages;
    issue(kt + kStages - 1, lds_a<N>(), lds_b<N>());
    compute<S>();
  }
};

// Workgroup u runs k split u / tiles of tile u % tiles. With splits, every
// split stores its raw accumulators (register order: value e of thread t at
// e * kThreads + t of its slot), publishes them (device-scope release) and counts
// itself on the tile's counter; the split that counts last (acquire) sums
// the slots in split order -- a fixed order, so the bits do not depend on
// which workgroup finishes last -- runs the epilogue and resets the counter.
template <int PL, int W>
__global__ __launch_bounds__(Waves<W>::kThreads) void split_gemm_kernel(SplitArgs p) {
  constexpr int kThreads = Waves<W>::kThreads, kWaveN = Waves<W>::kWaveN, kFN = Waves<W>::kFN;
  __shared__ int s_last;
  SplitGemm<PL, W> g;
  const int tiles_m = (p.m + kTile - 1) / kTile, tiles_n = (p.n + kTile - 1) / kTile;
  const int tiles = tiles_m * tiles_n;
  int u = blockIdx.x, tm, tn;
  if (p.splits > 1) {
    // Workgroups go round-robin over the 8 XCDs. Regroup ids so each XCD
    // runs whole splits: the tiles of one split share its k-range of the A
    // and B planes, which then stream once into that XCD's L2.
    const int total = tiles * p.splits;
    if (total % 8 == 0) u = (u % 8) * (total / 8) + u / 8;
  }
  const int split = u / tiles, tile = u % tiles;
  if (p.splits > 1) {
    tm = tile % tiles_m;
    tn = tile / tiles_m;
  } else {
    tile_coords(tile, tiles_m, tiles_n, tm, tn);
  }
  const int m0 = tm * kTile, n0 = tn * kTile;
  const int tid = threadIdx.x;
  g.wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  g.lane = tid & 63;
  g.wm = (g.wave & 1) * kWave;
  g.wn = (g.wave >> 1) * kWaveN;
  g.ktiles = p.Kp / kBK;
  const int kt0 = min(g.ktiles, split * p.split_ktiles);
  const int kt1 = min(g.ktiles, kt0 + p.split_ktiles);
  g.init(p, m0, n0);
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < kFN; ++j) g.acc[i][j] = f32x16{};

  if (kt0 < kt1) {
    // Every run is a multiple of kStages tiles; the two tiles prefetched past
    // kt1 lie in the next split's range or in the planes' spare k-blocks.
    g.issue(kt0, lds_a<0>(), lds_b<0>());
    g.issue(kt0 + 1, lds_a<1>(), lds_b<1>());
    for (int kt = kt0; kt < kt1; kt += kStages) {
      g.template step<0>(kt);
      g.template step<1>(kt + 1);
      g.template step<2>(kt + 2);
    }
    stage_barrier<0>();  // the spare tiles' loads land before the LDS is reused or freed
  }

  if (p.splits > 1) {
    constexpr int kSlot = kTile * kTile;
    float *mine = p.partials + ((int64_t)tile * p.splits + split) * kSlot + threadIdx.x;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < kFN; ++j)
#pragma unroll
        for (int r = 0; r < 16; ++r)
          __hip_atomic_store(mine + ((i * kFN + j) * 16 + r) * kThreads, g.acc[i][j][r],
                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    // The partial goes through agent-scope atomic stores (written through to
    // the level coherent across XCDs) and this wave waits for them before
    // the barrier after which the split counts itself; the last split reads
    // the partials with agent-scope atomic loads. No L2 writeback or
    // invalidation (what agent-scope fences would cost every tile sharing
    // the XCD's L2).
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_waitcnt(0x0070);  // vmcnt(0) lgkmcnt(0)
    asm volatile("" ::: "memory");
    __syncthreads();
    if (threadIdx.x == 0)
      s_last = __hip_atomic_fetch_add(&p.counters[tile], 1, __ATOMIC_RELAXED,
                                      __HIP_MEMORY_SCOPE_AGENT) == p.splits - 1;
    __syncthreads();
    if (!s_last) return;
    // acc := slot 0 + slot 1 + ... in split order (this split's own slot is
    // read back too, so the sum is the same whichever split finishes last).
    // One block row (32 values) per pass, its loads issued together: a whole
    // slot at once (128 values beside the 128 accumulators) would crowd the
    // registers.
    float *base = p.partials + (int64_t)tile * p.splits * kSlot + threadIdx.x;
#pragma unroll
    for (int i = 0; i < 4; ++i)
      for (int q = 0; q < p.splits; ++q) {
        float *slot = base + (int64_t)q * kSlot;
        float v[kFN][16];
#pragma unroll
        for (int j = 0; j < kFN; ++j)
#pragma unroll
          for (int r = 0; r < 16; ++r)
            v[j][r] = __hip_atomic_load(slot + ((i * kFN + j) * 16 + r) * kThreads,
                                        __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
        for (int j = 0; j < kFN; ++j)
#pragma unroll
          for (int r = 0; r < 16; ++r)
            g.acc[i][j][r] = q == 0 ? v[j][r] : g.acc[i][j][r] + v[j][r];
      }
    if (threadIdx.x == 0)
      __hip_atomic_store(&p.counters[tile], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }

  // Element r of lane l in block (i, j): C row m0+wm+32i+l%32, column
  // n0+wn+32j+8(r/4)+4(l/32)+r%4. The C loads of a block row (2 blocks) are
  // issued together without branches (an element outside C reads C[0]
  // instead and is not stored): one memory round trip per block row.
  const bool read_c = p.beta != 0.f;  // uniform
  const int r32 = g.lane & 31, hi = g.lane >> 5;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int gi = m0 + g.wm + 32 * i + r32;
    float cv[kFN][16];
    if (read_c) {
#pragma unroll
      for (int j = 0; j < kFN; ++j)
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const int gj = n0 + g.wn + 32 * j + 8 * (r >> 2) + 4 * hi + (r & 3);
          const bool in = gi < p.m && gj < p.n;
          cv[j][r] = *(in ? p.c + gi + (int64_t)gj * p.ldc : p.c);
        }
    }
#pragma unroll
    for (int j = 0; j < kFN; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int gj = n0 + g.wn + 32 * j + 8 * (r >> 2) + 4 * hi + (r & 3);
        const float v = g.acc[i][j][r];
        const float out = read_c ? __builtin_fmaf(p.alpha, v, p.beta * cv[j][r]) : p.alpha * v;
        if (gi < p.m && gj < p.n) p.c[gi + (int64_t)gj * p.ldc] = out;
      }
  }
}

int64_t round_up(int64_t x, int64_t m) { return (x + m - 1) / m * m; }

// k extent of the planes: a multiple of kStages k-tiles, plus kStages - 1
// spare k-blocks that the pipeline prefetches past the end (never used).
int64_t padded_k(int64_t k) { return round_up(k, (int64_t)kBK * kStages); }
int64_t alloc_k(int64_t k) { return padded_k(k) + (int64_t)kBK * (kStages - 1); }

template <int PLANES>
void launch_split(bool r_contig, const float *x, int64_t ld, int rows, int K, int rows_pad, int Kp,
                  __bf16 *planes, int64_t plane_elems, hipStream_t st) {
  dim3 grid((unsigned)((Kp + 63) / 64), (unsigned)(rows_pad / 64));
  if (r_contig)
    hipLaunchKernelGGL((split_kernel<true, PLANES>), grid, dim3(256), 0, st, x, ld, rows, K,
                       rows_pad, Kp, planes, plane_elems);
  else
    hipLaunchKernelGGL((split_kernel<false, PLANES>), grid, dim3(256), 0, st, x, ld, rows, K,
                       rows_pad, Kp, planes, plane_elems);
}

__global__ void zero_counters(int *x, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = 0;
}

// k splits of a split SGEMM: when the tile grid fills under half the CUs and
// each split keeps at least 64 k-tiles, enough splits to give ~256
// workgroups (at most 16). A function of the shape only.
struct SplitK {
  int splits, per;
};
SplitK split_k(int64_t m, int64_t n, int64_t k) {
  const int64_t tiles = (round_up(m, kTile) / kTile) * (round_up(n, kTile) / kTile);
  const int64_t kt = padded_k(k) / kBK;
  int64_t s = 1;
  if (tiles < kCUs / 2) {
    s = std::min<int64_t>((kCUs + tiles - 1) / tiles, kt / 64);
    s = std::max<int64_t>(1, std::min<int64_t>(s, 16));
  }
  const int64_t per = round_up((kt + s - 1) / s, kStages);
  s = (kt + per - 1) / per;
  return SplitK{(int)s, (int)per};
}

size_t planes_bytes(int64_t m, int64_t n, int64_t k, int planes) {
  const int64_t ka = alloc_k(k), mp = round_up(m, kTile), np = round_up(n, kTile);
  return (size_t)planes * (size_t)(mp + np) * (size_t)ka * 2;
}

}  // namespace

size_t sgemm_split_workspace(int64_t m, int64_t n, int64_t k, int planes) {
  const SplitK sk = split_k(m, n, k);
  size_t bytes = round_up((int64_t)planes_bytes(m, n, k, planes), 256);
  if (sk.splits > 1) {
    const int64_t tiles = (round_up(m, kTile) / kTile) * (round_up(n, kTile) / kTile);
    bytes += (size_t)tiles * sk.splits * kTile * kTile * sizeof(float) + tiles * sizeof(int);
  }
  return bytes;
}

hipError_t launch_sgemm_split(const GemmProblem<float> &p, int planes, void *workspace,
                              hipStream_t stream) {
  if (planes != 2 && planes != 3) return hipErrorInvalidValue;
  const int64_t kp = padded_k(p.k), ka = alloc_k(p.k);
  const int64_t mp = round_up(p.m, kTile), np = round_up(p.n, kTile);
  __bf16 *wa = (__bf16 *)workspace;
  __bf16 *wb = wa + (int64_t)planes * mp * ka;
  const int64_t pa = mp * ka, pb = np * ka;
  // op(A) = A (m x k, m-contiguous) or A^T (stored k x m, k-contiguous);
  // op(B)^T = B^T (B stored k x n, k-contiguous) or B (stored n x k, n-contiguous).
  if (planes == 2) {
    launch_split<2>(!p.ta, p.a, p.lda, (int)p.m, (int)p.k, (int)mp, (int)kp, wa, pa, stream);
    launch_split<2>(p.tb, p.b, p.ldb, (int)p.n, (int)p.k, (int)np, (int)kp, wb, pb, stream);
  } else {
    launch_split<3>(!p.ta, p.a, p.lda, (int)p.m, (int)p.k, (int)mp, (int)kp, wa, pa, stream);
    launch_split<3>(p.tb, p.b, p.ldb, (int)p.n, (int)p.k, (int)np, (int)kp, wb, pb, stream);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const SplitK sk = split_k(p.m, p.n, p.k);
  const int64_t tiles = (mp / kTile) * (np / kTile);
  float *partials = nullptr;
  int *counters = nullptr;
  if (sk.splits > 1) {
    partials = (float *)((char *)workspace + round_up((int64_t)planes_bytes(p.m, p.n, p.k, planes), 256));
    counters = (int *)(partials + tiles * sk.splits * kTile * kTile);
    hipLaunchKernelGGL(zero_counters, dim3((unsigned)((tiles + 255) / 256)), dim3(256), 0, stream,
                       counters, (int)tiles);
  }
  SplitArgs a{wa, wb, pa, pb, (int)p.m, (int)p.n, (int)mp, (int)np, (int)kp, p.alpha, p.beta,
              p.c, p.ldc, sk.splits, sk.per, partials, counters};
  const dim3 grid((unsigned)(tiles * sk.splits));
  // Two waves per SIMD hide the barrier and LDS waits of the square shapes;
  // the k-split BF16x3 product (short per-split runs) measured faster with
  // one wave per SIMD.
  if (planes == 2 && sk.splits > 1)
    hipLaunchKernelGGL((split_gemm_kernel<2, 4>), grid, dim3(Waves<4>::kThreads), 0, stream, a);
  else if (planes == 2)
    hipLaunchKernelGGL((split_gemm_kernel<2, 8>), grid, dim3(Waves<8>::kThreads), 0, stream, a);
  else
    hipLaunchKernelGGL((split_gemm_kernel<3, 8>), grid, dim3(Waves<8>::kThreads), 0, stream, a);
  return hipGetLastError();
}

}  // namespace rblas
