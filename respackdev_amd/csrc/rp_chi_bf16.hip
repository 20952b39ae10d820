// Split-BF16 chi contraction for gfx950 (the mixed-precision path).
//
// chi_w(G,G') = sum_K conj(a_K(G)) d_w(K) a_K(G'), the rank-K update of
// chiqw.F:1397-1410, on v_mfma_f32_16x16x32_bf16 with every real operand split
// into two bf16 terms x = hi + lo and the products hi*hi + hi*lo + lo*hi
// (3 MFMAs per real product, FP32 accumulation).
//
// Operand format: a "packed octet" holds 8 consecutive K of one G as
// 32 bf16 = [Re hi x8][Re lo x8][Im hi x8][Im lo x8] (64 bytes); each plane is
// the eight k-slots one MFMA lane supplies.  Lane group q of a wave supplies
// octet q of a 32-K chunk, so every MFMA has real k-slots of one kind, and the
// complex product is assembled from three accumulators:
//   Re chi += ar br + ai bi,   ImP += ar bi,   ImN += ai br,   Im chi = ImP - ImN
// (conj(a) b without any sign flip or select in the loop).  The left operand a
// (shared by every frequency) is split once per q by k_presplit into octets in
// HBM; the right operand b = d_w(K) a_K(G') is formed from the same octets
// and split again per workgroup while a chunk is staged, once for all eight
// waves.
//
// Workgroup: 512 threads = 2x4 waves (two per SIMD), tile 128 G x 128 G', one
// frequency, one K range; each wave owns 4x2 blocks of 16x16.  Per chunk of 32
// complex K the left tile (32 KB of octets) and the converted right tile
// (32 KB) sit in LDS with rows padded by 16 bytes (conflict-free 16-byte
// reads), double buffered: the next chunk's global data is loaded into
// registers during the MFMAs and converted/stored after them, with one barrier
// per chunk.
//
// Schedule: one persistent workgroup per CU.  The (frequency, tile) pairs are
// dealt round-robin to the XCDs (workgroup b runs on XCD b % 8), and each XCD
// walks its pairs split by split: the 32 workgroups of an XCD work on the same
// K range at the same time, so the operand chunks they share (5 row blocks and
// 5 column blocks serve all 25 tiles at NG ~ 600) come from that XCD's L2
// instead of HBM.
// Every flush_chunks chunks (the launch's accumulation span, kFlushChunks =
// 8192 K by default) the FP32 accumulators are added into the FP64 partial
// tile of the item in HBM, which the workgroup owns exclusively (FP64
// accumulation across K blocks, FP32 within; the first flush of an item
// stores).  At the precision workload's shapes a span of 4096 K costs 4% more
// time (the flush read-modify-writes the tile) for chi errors of 4.6e-7
// instead of 6.5e-7; the host shortens the span for a q whose eps is
// ill-conditioned (the FP32 accumulation error is then what eps^-1 amplifies).
#include <hip/hip_bfloat16.h>

#include <algorithm>
#include <type_traits>

#include "rp_common.h"
#include "rp_kernels.h"

namespace rp {

typedef float f4v __attribute__((ext_vector_type(4)));
typedef __bf16 bf8v __attribute__((ext_vector_type(8)));
typedef unsigned int u4v __attribute__((ext_vector_type(4)));

namespace {
constexpr int kTM = 128, kTN = 128;   // tile (G x G')
constexpr int kKC = 32;               // complex K per chunk
constexpr int kK8 = kKC / 8;          // octets per row per chunk
constexpr int kRowBytes = kK8 * 64 + 16;  // padded LDS row (bytes)
constexpr int kFlushChunks = 256;     // default FP32 accumulation span: 8192 K
constexpr int kXcd = 8;               // XCDs of gfx950 (workgroup b -> XCD b % 8)
constexpr int kSyncChunks = 32;       // chunks between rendezvous of an XCD's workgroups
constexpr int kSyncSpin = 512;        // bound on the polls of one rendezvous
constexpr int kSyncStride = 32;       // counters 128 bytes apart (one per XCD)
}  // namespace

__device__ __forceinline__ unsigned short bf16_bits(float x) {
  // round to nearest even (the conversion __bf16(x) performs)
  __bf16 h = (__bf16)x;
  return __builtin_bit_cast(unsigned short, h);
}
__device__ __forceinline__ float bf16_to_f(unsigned short b) {
  return __builtin_bit_cast(float, (unsigned int)b << 16);
}

// Splits 8 complex values into one packed octet (4 x uint4).
__device__ __forceinline__ void make_octet(const float2* v, u4v* o) {
  unsigned short rh[8], rl[8], ih[8], il[8];
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    rh[e] = bf16_bits(v[e].x);
    rl[e] = bf16_bits(v[e].x - bf16_to_f(rh[e]));
    ih[e] = bf16_bits(v[e].y);
    il[e] = bf16_bits(v[e].y - bf16_to_f(ih[e]));
  }
#pragma unroll
  for (int w = 0; w < 4; ++w) {
    o[0][w] = (unsigned)rh[2 * w] | ((unsigned)rh[2 * w + 1] << 16);
    o[1][w] = (unsigned)rl[2 * w] | ((unsigned)rl[2 * w + 1] << 16);
    o[2][w] = (unsigned)ih[2 * w] | ((unsigned)ih[2 * w + 1] << 16);
    o[3][w] = (unsigned)il[2 * w] | ((unsigned)il[2 * w + 1] << 16);
  }
}

// A[K][lda] complex64 -> octets Ao[K/8][lda] (4 x uint4 each).  One thread per
// (octet row, G).
__global__ void __launch_bounds__(256) k_presplit(const float2* __restrict__ A, size_t lda,
                                                  size_t k8n, u4v* __restrict__ Ao) {
  const size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (idx >= k8n * lda) return;
  const size_t k8 = idx / lda, g = idx % lda;
  float2 v[8];
#pragma unroll
  for (int e = 0; e < 8; ++e) v[e] = A[(k8 * 8 + e) * lda + g];
  u4v o[4];
  make_octet(v, o);
  u4v* dst = Ao + idx * 4;
#pragma unroll
  for (int w = 0; w < 4; ++w) dst[w] = o[w];
}

__device__ __forceinline__ bf8v as_bf8(u4v x) { return __builtin_bit_cast(bf8v, x); }

// Element e (0..7) of one plane pair of an octet as hi + lo in FP32.
__device__ __forceinline__ float octet_value(u4v hi, u4v lo, int e) {
  const unsigned sh = (e & 1) ? 16u : 0u;
  const unsigned h = (hi[e >> 1] >> sh) & 0xffffu, l = (lo[e >> 1] >> sh) & 0xffffu;
  return __builtin_bit_cast(float, h << 16) + __builtin_bit_cast(float, l << 16);
}

// d (complex FP64, [nen][ldd]) -> single precision copy used by the contraction.
__global__ void k_weights_to_f32(const double2* __restrict__ dw, size_t n, float2* __restrict__ df) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    df[i] = make_float2((float)dw[i].x, (float)dw[i].y);
}

// grid: persistent, a multiple of kXcd workgroups of 512 threads (8 waves, two
// per SIMD).  Item (split z, frequency w, tile) accumulates K chunks
// [z*chunks_per_split, ...) into out, the FP64 partial [z][w][ng][ng] (element
// (G,G') at G' * ng + G); every element of it is stored by exactly one item.
// Columns of A beyond ng are zero; in edge tiles the waves whose rows or
// columns all lie beyond ng skip their MFMAs.
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_chi_bf16x3(const u4v* __restrict__ Ao, size_t lda, const float2* __restrict__ df, size_t ldd,
             size_t chunks_per_split, size_t nchunks, int ng, int ntile, int nen, int nsplit,
             int flush_chunks, double2* __restrict__ out, unsigned* __restrict__ sync) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  // two buffers of [left tile | right tile], each kTM (kTN) rows of kRowBytes
  constexpr int kBuf = (kTM + kTN) * kRowBytes;
  const int t = threadIdx.x, lane = t & 63;
  const int wave = __builtin_amdgcn_readfirstlane(t >> 6);  // wave-uniform (SGPR)
  // 2 x 4 waves of 64 G x 32 G' each.  Wave w runs on SIMD w % 4; the two
  // waves of a SIMD take different row halves and column blocks two apart, so
  // that in an edge tile (rows 64-127 or columns 64-127 past ng) each SIMD
  // keeps one wave with MFMAs.
  const int wm = (wave >> 2) & 1, wn = (wave + 2 * (wave >> 2)) & 3;
  const int li = lane & 15, q = lane >> 4;
  // this XCD's share of the (frequency, tile) pairs, walked split-major
  const int ntiles = ntile * ntile, npairs = ntiles * nen;
  const int xcd = (int)(blockIdx.x % kXcd), slot = (int)(blockIdx.x / kXcd);
  const int nslot = (int)(gridDim.x / kXcd);
  const int mypairs = (npairs - xcd + kXcd - 1) / kXcd;
  const int nitems = mypairs * nsplit;
  // Rendezvous of the XCD's workgroups every kSyncChunks chunks (split phase:
  // arrive, keep working for one chunk, then wait).  Without it the workgroups
  // drift apart by more chunks than L2 holds (measured: L2 hit rate 45% instead
  // of 84% on real pair densities) and each fetches the shared operand blocks
  // itself.  The wait is bounded (kSyncSpin polls), so a missing workgroup can
  // only cost time, never block: every wave reaches the end of the kernel.
  unsigned* const xsync = sync + xcd * kSyncStride;
  unsigned epoch = 0;
  int phase = 0;  // chunks since the last arrival; waits at phase == 1 after an arrival
  bool waiting = false;
  for (int item = slot; item < nitems; item += nslot) {
  const int z = item / mypairs, u = (item % mypairs) * kXcd + xcd;
  const int w = u / ntiles, tile = u % ntiles;
  const int g0 = (tile % ntile) * kTM, h0 = (tile / ntile) * kTN;
  const bool active = g0 + wm * 64 < ng && h0 + wn * 32 < ng;  // wave-uniform
  const float2* dww = df + (size_t)w * ldd;
  const size_t c0 = (size_t)z * chunks_per_split;
  size_t c1 = c0 + chunks_per_split;
  if (c1 > nchunks) c1 = nchunks;

  // Staging: thread t owns left octet (k8 = t>>7, row = t&127) and right
  // octet (k8 = t>>7, col = t&127) of every chunk.  Both come from the
  // presplit octets (the right one rebuilt as hi + lo, which is a to within
  // 2^-17 relative), so the left and right tiles of an XCD's workgroups share
  // one set of row blocks in L2.  sk8 is wave-uniform, so the eight weights of
  // a thread are scalar loads.
  const int srow = t & 127, sk8 = wave >> 1;
  u4v pl[4], pb[4];
  float2 pd[8];
  auto load_chunk = [&](size_t c) {
    const u4v* src = Ao + ((c * kK8 + sk8) * lda + g0 + srow) * 4;
    const u4v* srb = Ao + ((c * kK8 + sk8) * lda + h0 + srow) * 4;
#pragma unroll
    for (int x = 0; x < 4; ++x) {
      pl[x] = src[x];
      pb[x] = srb[x];
    }
    const size_t kb = c * kKC + (size_t)sk8 * 8;
#pragma unroll
    for (int e = 0; e < 8; ++e) pd[e] = dww[kb + e];
  };
  auto stage = [&](unsigned char* buf) {
    u4v* dl = (u4v*)(buf + srow * kRowBytes + sk8 * 64);
#pragma unroll
    for (int x = 0; x < 4; ++x) dl[x] = pl[x];
    float2 v[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const float ar = octet_value(pb[0], pb[1], e), ai = octet_value(pb[2], pb[3], e);
      v[e] = make_float2(pd[e].x * ar - pd[e].y * ai, pd[e].x * ai + pd[e].y * ar);
    }
    u4v ov[4];
    make_octet(v, ov);
    u4v* dr = (u4v*)(buf + (kTM + srow) * kRowBytes + sk8 * 64);
#pragma unroll
    for (int x = 0; x < 4; ++x) dr[x] = ov[x];
  };

  // Re chi = sum ar br + ai bi; Im chi = sum ar bi - sum ai br, the two terms of
  // Im in separate accumulators (ip, in) so that every MFMA takes its k-slots
  // from one plane of the octets: no selects or sign flips in the loop.
  f4v acc_re[4][2], acc_ip[4][2], acc_in[4][2];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc_re[a][b] = acc_ip[a][b] = acc_in[a][b] = f4v{0, 0, 0, 0};

  double2* o = out + ((size_t)z * nen + w) * (size_t)ng * ng;
  bool first = true;
  auto flush = [&]() {
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int gi = g0 + wm * 64 + a * 16 + 4 * q + v;
          const int hj = h0 + wn * 32 + b * 16 + li;
          if (gi < ng && hj < ng) {
            double2* p = o + (size_t)hj * ng + gi;
            double2 cur = first ? make_double2(0.0, 0.0) : *p;
            cur.x += (double)acc_re[a][b][v];
            cur.y += (double)acc_ip[a][b][v] - (double)acc_in[a][b][v];
            *p = cur;
          }
        }
        acc_re[a][b] = acc_ip[a][b] = acc_in[a][b] = f4v{0, 0, 0, 0};
      }
    first = false;
  };

  if (c0 < c1) {
    load_chunk(c0);
    stage(smem);
  }
  __syncthreads();
  int since = 0;
  // The chunk loop, in two forms: a wave whose 64 rows or 32 columns lie
  // wholly past ng (edge tiles) only stages (its accumulators would stay zero
  // and its part of the tile is never stored), the others also read their
  // operands and run the MFMAs.  Both meet the same barriers.
  auto chunks = [&](auto act) {
    constexpr bool kAct = decltype(act)::value;
    for (size_t c = c0; c < c1; ++c) {
    const int cur = (int)((c - c0) & 1);
    const unsigned char* bcur = smem + cur * kBuf;
    // next chunk's operands in flight during the MFMAs (the last chunk loads
    // itself again: no branch, so that loads, MFMAs and staging form one
    // scheduling region)
    load_chunk(c + 1 < c1 ? c + 1 : c);
    if constexpr (kAct) {
    // lane group q supplies the 8 K of octet q of the chunk (k-slots 8q..8q+7)
    bf8v Brh[2], Brl[2], Bih[2], Bil[2], Lh[4], Ll[4];
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const u4v* p = (const u4v*)(bcur + (kTM + wn * 32 + b * 16 + li) * kRowBytes + q * 64);
      Brh[b] = as_bf8(p[0]);
      Brl[b] = as_bf8(p[1]);
      Bih[b] = as_bf8(p[2]);
      Bil[b] = as_bf8(p[3]);
    }
    // Re a: acc_re += ar br, acc_ip += ar bi (hi*hi, hi*lo, lo*hi)
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const u4v* p = (const u4v*)(bcur + (wm * 64 + a * 16 + li) * kRowBytes + q * 64);
      Lh[a] = as_bf8(p[0]);
      Ll[a] = as_bf8(p[1]);
    }
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        acc_re[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Lh[a], Brh[b], acc_re[a][b], 0, 0, 0);
        acc_ip[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Lh[a], Bih[b], acc_ip[a][b], 0, 0, 0);
      }
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        acc_re[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Lh[a], Brl[b], acc_re[a][b], 0, 0, 0);
        acc_ip[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Lh[a], Bil[b], acc_ip[a][b], 0, 0, 0);
      }
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        acc_re[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ll[a], Brh[b], acc_re[a][b], 0, 0, 0);
        acc_ip[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ll[a], Bih[b], acc_ip[a][b], 0, 0, 0);
      }
    // Im a: acc_re += ai bi, acc_in += ai br
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const u4v* p = (const u4v*)(bcur + (wm * 64 + a * 16 + li) * kRowBytes + q * 64);
      Lh[a] = as_bf8(p[2]);
      Ll[a] = as_bf8(p[3]);
    }
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        acc_re[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Lh[a], Bih[b], acc_re[a][b], 0, 0, 0);
        acc_in[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Lh[a], Brh[b], acc_in[a][b], 0, 0, 0);
      }
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        acc_re[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Lh[a], Bil[b], acc_re[a][b], 0, 0, 0);
        acc_in[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Lh[a], Brl[b], acc_in[a][b], 0, 0, 0);
      }
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        acc_re[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ll[a], Bih[b], acc_re[a][b], 0, 0, 0);
        acc_in[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ll[a], Brh[b], acc_in[a][b], 0, 0, 0);
      }
    }
    stage(smem + (cur ^ 1) * kBuf);  // into the idle buffer (unused after the last chunk)
    if constexpr (kAct) {
      // Issue order: the eight global loads, the sixteen operand reads of the
      // real phase, its 48 MFMAs in pairs, the imaginary phase's reads, then
      // its 48 MFMAs each followed by up to five VALU instructions of the
      // staging (measured 2-4% faster than leaving the order to the scheduler).
      __builtin_amdgcn_sched_group_barrier(0x020, 8, 0);   // global reads
      __builtin_amdgcn_sched_group_barrier(0x100, 16, 0);  // LDS reads
#pragma unroll
      for (int s_ = 0; s_ < 24; ++s_) __builtin_amdgcn_sched_group_barrier(0x008, 2, 0);  // MFMA
      __builtin_amdgcn_sched_group_barrier(0x100, 8, 0);
#pragma unroll
      for (int s_ = 0; s_ < 48; ++s_) {
        __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
        __builtin_amdgcn_sched_group_barrier(0x002, 5, 0);  // VALU
      }
      __builtin_amdgcn_sched_group_barrier(0x200, 8, 0);  // LDS writes
      if (++since == flush_chunks) {
        since = 0;
        flush();
      }
    }
    if (t == 0) {
      if (waiting) {
        const unsigned target = epoch * (unsigned)nslot;
        for (int i = 0; i < kSyncSpin; ++i) {
          if (__hip_atomic_load(xsync, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= target) break;
          __builtin_amdgcn_s_sleep(4);
        }
      }
      if (++phase == kSyncChunks) {
        __hip_atomic_fetch_add(xsync, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        ++epoch;
        phase = 0;
        waiting = true;
      } else {
        waiting = false;
      }
    }
    __syncthreads();
    }
  };
  if (active) chunks(std::true_type{});
  else chunks(std::false_type{});
  if (active && (since || first)) flush();
  }  // item
  // A workgroup out of items releases every later rendezvous of its XCD (its
  // peers would otherwise wait out the spin bound at each one); 2^22 exceeds
  // any epoch * nslot the counter reaches (nslot <= 256, epochs < 2^14).
  if (t == 0) __hip_atomic_fetch_add(xsync, 1u << 22, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

size_t bf16x3_lda_multiple() { return kTM; }

void launch_presplit(const float2* A, size_t lda, size_t kpad, void* Ao, hipStream_t s) {
  const size_t k8n = kpad / 8;
  const size_t n = k8n * lda;
  const unsigned blocks = (unsigned)((n + 255) / 256);
  hipLaunchKernelGGL(k_presplit, dim3(blocks), dim3(256), 0, s, A, lda, k8n, (u4v*)Ao);
  RP_HIP(hipGetLastError());
}

void launch_weights_to_f32(const double2* dw, size_t n, float2* df, hipStream_t s) {
  unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 65536);
  hipLaunchKernelGGL(k_weights_to_f32, dim3(blocks), dim3(256), 0, s, dw, n, df);
  RP_HIP(hipGetLastError());
}

// Splits of K for the persistent schedule: the count (1..64) that minimises
// the rounds of items per unit of work, given nwg workgroups (kXcd groups).
int bf16x3_nsplit(int ng, int nen, size_t kpad, int nwg) {
  const int ntile = (ng + kTM - 1) / kTM;
  const long npairs = (long)ntile * ntile * nen;
  const long mypairs = (npairs + kXcd - 1) / kXcd;
  const long nslot = std::max(1, nwg / kXcd);
  const size_t nchunks = kpad / kKC;
  int best_s = 1;
  double best = 1e300;
  for (int sp = 1; sp <= 64 && (size_t)sp <= nchunks; ++sp) {
    const double rounds = (double)((mypairs * sp + nslot - 1) / nslot) / sp;
    if (rounds < best - 1e-9) { best = rounds; best_s = sp; }
  }
  return best_s;
}

int bf16x3_workgroups(int ncu) { return std::max(kXcd, ncu / kXcd * kXcd); }

size_t bf16x3_sync_words() { return (size_t)kXcd * kSyncStride; }

int bf16x3_default_span() { return kFlushChunks * kKC; }

// The FP32 accumulation error of chi, amplified in eps^-1 by about the
// condition of eps (~ |eps^-1|_1, as eps is the identity plus a modest term),
// sets the span.  Measured on the 27 q of the batched workload (NG 434-705):
// the eps^-1 error against FP64 stays below c(span) kappa with c = 1.7e-7
// (8192 K), 1.2e-7 (4096), 7.5e-8 (2048), 3.2e-8 (1024).  The spans below keep
// c kappa under 5e-6, half the mixed path's 1e-5 bar.
int bf16x3_span_for_kappa(double kappa) {
  if (kappa <= 25.0) return 8192;
  if (kappa <= 40.0) return 4096;
  if (kappa <= 60.0) return 2048;
  if (kappa <= 150.0) return 1024;
  return 512;
}

void launch_chi_bf16x3(const void* Ao, size_t lda, const float2* df, size_t ldd, size_t kpad, int ng,
                       int nen, int nsplit, int nwg, int span, double2* partial, unsigned* sync,
                       hipStream_t s) {
  RP_CHECK(span >= kKC && span % kKC == 0, "bf16x3 contraction: span must be a multiple of 32 K");
  RP_CHECK(lda % kTM == 0, "bf16x3 contraction needs lda a multiple of 128");
  RP_CHECK(kpad % kKC == 0, "bf16x3 contraction needs K padded to a multiple of 32");
  RP_CHECK(nwg % kXcd == 0 && nwg > 0, "bf16x3 contraction needs a multiple of 8 workgroups");
  RP_CHECK(sync != nullptr, "bf16x3 contraction needs its rendezvous counters");
  launch_zero(sync, sizeof(unsigned) * bf16x3_sync_words(), s);
  const int ntile = (ng + kTM - 1) / kTM;
  const size_t nchunks = kpad / kKC;
  const size_t per = (nchunks + nsplit - 1) / nsplit;
  const size_t lds = 2 * (size_t)(kTM + kTN) * kRowBytes;
  RP_HIP(hipFuncSetAttribute((const void*)k_chi_bf16x3, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)lds));
  hipLaunchKernelGGL(k_chi_bf16x3, dim3(nwg), dim3(512), lds, s, (const u4v*)Ao, lda, df, ldd, per,
                     nchunks, ng, ntile, nen, nsplit, span / kKC, partial, sync);
  RP_HIP(hipGetLastError());
}

}  // namespace rp
