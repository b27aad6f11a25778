// unet.hip — learned denoiser kernels: 3x3 convolutions as implicit GEMMs on CDNA4 matrix cores
// (v_mfma_f32_16x16x32_bf16: BF16 operands, FP32 accumulation), their data- and weight-gradient
// passes, the elementwise layers of a 3-level U-Net (2x2 average pooling, 2x nearest upsampling,
// channel concatenation), L1 loss and Adam.
//
// Activations are NHWC BF16 with the channel count a multiple of 32. Weights are kept in FP32
// (master copy for Adam) and in two BF16 layouts: forward [Cout][9][Cin] and data-gradient
// [Cin][9][Cout] with the taps mirrored (the transposed convolution). MFMA operand layout of
// 16x16x32 (wave64): lane l holds A[m = l % 16][k = 8 (l / 16) .. + 8], B[k = 8 (l / 16) .. + 8][n =
// l % 16] and receives D[m = 4 (l / 16) + i][n = l % 16], i = 0..3.
#include <algorithm>
#include "abi_util.hpp"

namespace mjpbr {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float floatx4;
typedef uint16_t bf16_t;

MJ_DEV float bf2f(bf16_t h) { return __uint_as_float((uint32_t)h << 16); }
MJ_DEV bf16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (bf16_t)((u >> 16) | 0x40);  // NaN stays NaN
  u += 0x7fffu + ((u >> 16) & 1u);                                         // round to nearest even
  return (bf16_t)(u >> 16);
}

MJ_DEV floatx4 mfma16(const uint4& a, const uint4& b, floatx4 c) {
  bf16x8 va = __builtin_bit_cast(bf16x8, a);
  bf16x8 vb = __builtin_bit_cast(bf16x8, b);
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(va, vb, c, 0, 0, 0);
}

// ------------------------------------------------------------------ forward / data-gradient conv
// y[p][n] = act(bias[n] + sum_{tap, c} x[p + off(tap)][c] * w[n][tap][c]), zero padding.
// With `mask`, the input is x * (mask > 0) (ReLU backward of the layer whose output is `mask`).
// With `x2`, the input is concat(upsample2x(x), x2) along channels (the U-Net decoder input):
// channels [0, c_split) are read from the half-resolution x at (y / 2, x / 2), the rest from x2,
// so the concatenation is never materialised.
//
// Workgroup: 4 waves, a 16x16 pixel block (a wave owns 4 rows of 16 pixels) x up to 64 output
// channels (1-4 column tiles of 16). Each MFMA multiplies a tile's weights (A: 16 output channels
// x 32 input channels of one tap) by a row of the patch (B: 32 channels x 16 pixels), so a lane
// receives 4 consecutive output channels of one pixel and the epilogue stores them straight from
// registers. The K loop runs over 32-channel chunks x 9 taps; the 18x18x32 input patch lives in
// LDS (the weights too, unless resident in registers).
constexpr int kConvBlock = 16;
constexpr int kPatch = kConvBlock + 2;
constexpr int kCinChunk = 32;
constexpr int kNTile = 64;
constexpr int kWRow = 9 * kCinChunk + 8;  // padded LDS row of one output channel's weights
constexpr int kPStride = kCinChunk + 8;   // padded LDS channel stride of a patch position
constexpr int kPatchElems = kPatch * kPatch * kPStride;
constexpr int kPatchVecs = kPatch * kPatch * (kCinChunk / 8);  // 16-byte loads of one patch
constexpr int kPrefetch = (kPatchVecs + 255) / 256;            // per thread

struct ConvArgs {
  const bf16_t* x;
  const bf16_t* x2;
  const bf16_t* mask;
  const bf16_t* w;  // [Cout][9][Cin]
  const float* bias;
  bf16_t* y;
  float* y32;  // optional FP32 output (first out_f32 channels)
  int n, h, wd, cin, cout;
  int relu;
  int out_f32;
  int c_split;
  bf16_t* pool;  // optional 2x2 average pooling of y ([n][h/2][w/2][cout])
};

MJ_DEV uint4 apply_relu_mask(uint4 v, uint4 mk) {
  uint32_t* pv = reinterpret_cast<uint32_t*>(&v);
  const uint32_t* pm = reinterpret_cast<const uint32_t*>(&mk);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    uint32_t lo = (pm[q] & 0xffffu), hi = pm[q] >> 16;
    // keep a value where the mask is positive (bf16 sign bit clear and not +0)
    uint32_t keep_lo = (lo != 0 && !(lo & 0x8000u)) ? 0xffffu : 0u;
    uint32_t keep_hi = (hi != 0 && !(hi & 0x8000u)) ? 0xffff0000u : 0u;
    pv[q] &= (keep_lo | keep_hi);
  }
  return v;
}

// Input reads go through buffer resources: an out-of-range offset reads zeros without a branch,
// so a prefetched vector is only waited for where it is stored to LDS (a branchy load - halo
// test, source select - makes the compiler wait at every join).
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
constexpr uint32_t kOutOfRange = 0x80000000u;
constexpr int kBufferWord3 = 0x00020000;  // gfx9 raw buffer: 32-bit data format, no swizzle

MJ_DEV __amdgpu_buffer_rsrc_t buffer_of(const void* base, int64_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes, kBufferWord3);
}
MJ_DEV uint4 buffer_vec(__amdgpu_buffer_rsrc_t r, uint32_t off) {
  const u32x4 v = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0);
  return make_uint4(v.x, v.y, v.z, v.w);
}

// Where one 32-channel chunk of a block's patch comes from: one image of x (or, for the decoder
// input, of the half-resolution x or of x2), `stride` channels a pixel from channel `coff`.
struct PatchSrc {
  __amdgpu_buffer_rsrc_t data, mask;
  int stride, coff, low;
};

MJ_DEV PatchSrc patch_src(const ConvArgs& a, int img, int c0) {
  PatchSrc p;
  const int64_t full = (int64_t)a.h * a.wd;
  if (a.x2 && c0 < a.c_split) {  // c_split is a multiple of the chunk: a chunk has one source
    const int64_t px = (int64_t)(a.h >> 1) * (a.wd >> 1);
    p.data = buffer_of(a.x + img * px * a.c_split, px * a.c_split * 2);
    p.stride = a.c_split;
    p.coff = c0;
    p.low = 1;
  } else if (a.x2) {
    const int c2 = a.cin - a.c_split;
    p.data = buffer_of(a.x2 + img * full * c2, full * c2 * 2);
    p.stride = c2;
    p.coff = c0 - a.c_split;
    p.low = 0;
  } else {
    p.data = buffer_of(a.x + img * full * a.cin, full * a.cin * 2);
    p.stride = a.cin;
    p.coff = c0;
    p.low = 0;
  }
  p.mask = buffer_of(a.mask ? a.mask + img * full * a.cin : a.x, a.mask ? full * a.cin * 2 : 0);
  return p;
}

// Byte offset of 16-byte vector k of the patch of block (y0, x0), or kOutOfRange for the zero
// padding and for k >= kPatchVecs.
MJ_DEV uint32_t patch_offset(const ConvArgs& a, const PatchSrc& p, int y0, int x0, int k) {
  const int pos = k >> 2, part = k & 3;
  const int yy = y0 - 1 + pos / kPatch, xx = x0 - 1 + pos % kPatch;
  const bool inside = k < kPatchVecs && yy >= 0 && yy < a.h && xx >= 0 && xx < a.wd;
  const int pix = (yy >> p.low) * (a.wd >> p.low) + (xx >> p.low);  // the half-resolution source: (y / 2, x / 2)
  const uint32_t off = (uint32_t)(pix * p.stride + p.coff + part * 8) * 2u;
  return inside ? off : kOutOfRange;
}

// 16-byte vector e of the forward weights of output channels n0.. and input channels
// c0..c0+31: output channel n0 + e / 36, tap (e % 36) / 4, 8 channels (e % 4) * 8; zeros for
// e >= `count`.
MJ_DEV uint4 load_weight_vec(const ConvArgs& a, __amdgpu_buffer_rsrc_t wr, int n0, int c0, int e, int count) {
  const int n = e / 36, rem = e % 36, tap = rem >> 2, part = rem & 3;
  const uint32_t off = (uint32_t)(((n0 + n) * 9 + tap) * a.cin + c0 + part * 8) * 2u;
  return buffer_vec(wr, e < count ? off : kOutOfRange);
}

MJ_DEV void store_patch_vec(bf16_t* patch, int k, uint4 v) {
  *reinterpret_cast<uint4*>(patch + (k >> 2) * kPStride + (k & 3) * 8) = v;
}

// NT: MFMA column tiles the accumulators are sized for (1, 2 or 4); the last column group of a
// launch may use fewer (ntile).
//
// A workgroup streams a contiguous run of `bpg` pixel blocks (so horizontally and vertically
// adjacent blocks, whose patches share halo rows and columns, are read by one workgroup through
// one XCD's L2), as (block, chunk) steps. Patches load into registers kDepth steps ahead (two
// alternating buffers for up to 2 column tiles without the ReLU mask, else one), the next step's
// weights one step ahead, all in flight while the matrix cores work on the current step.
// kRes (single-chunk layers, at most 2 column tiles): each lane keeps its B operands of all nine
// taps in registers for the whole launch, read once from global memory; otherwise the weights of
// every step go through LDS. Taps run column shift (dx) first: each of the 6 patch rows under a
// wave's 4 output rows is read once per shift and feeds the 3 row shifts (dy) it serves.
// Occupancy: 2 waves per SIMD (at 3 the compiler spills the prefetched vectors, and a spill store
// waits for its load); the registers hold the prefetch and the resident B operands.
template <int NT, bool kMask, bool kRes>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
conv3x3_kernel(ConvArgs a, int bpg) {
  extern __shared__ __attribute__((aligned(16))) bf16_t conv_lds[];
  const int n0 = blockIdx.y * NT * 16;
  const int nn = min(NT * 16, a.cout - n0);
  const int ntile = nn / 16;
  bf16_t* patch = conv_lds;
  bf16_t* wl = conv_lds + kPatchElems;  // streamed weights (not kRes)
  const int bx = a.wd / kConvBlock, by = a.h / kConvBlock;
  const int nblocks = a.n * bx * by;
  const int first = blockIdx.x * bpg;
  const int nchunk = a.cin / kCinChunk;
  const int nsteps = (min(nblocks, first + bpg) - first) * nchunk;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int lm = lane & 15, lk = lane >> 4;
  auto block_origin = [&](int blk, int* img, int* y0, int* x0) {
    *img = blk / (bx * by);
    const int r = blk - *img * (bx * by);
    *y0 = (r / bx) * kConvBlock;
    *x0 = (r % bx) * kConvBlock;
  };

  constexpr int kWVecs = NT * 16 * 9 * 4;  // 16-byte weight vectors of a chunk (all column tiles)
  constexpr int kWPrefetch = kRes ? 1 : (kWVecs + 255) / 256;
  const int wvecs = ntile * 16 * 9 * 4;
  const __amdgpu_buffer_rsrc_t wr = buffer_of(a.w, (int64_t)a.cout * 9 * a.cin * 2);
  // patches in flight: kDepth steps ahead, in as many register buffers (two where the registers
  // allow it without spilling)
  constexpr int kDepth = (NT <= 2 && !kMask) ? 2 : 1;
  constexpr int kMaskVecs = kMask ? kPrefetch : 1;
  uint4 pre0[kPrefetch], pre1[kDepth == 2 ? kPrefetch : 1];
  uint4 mpre0[kMaskVecs], mpre1[kMaskVecs];
  uint4 wpre[kWPrefetch];
  uint4 breg[kRes ? 9 : 1][NT];
  // issues the loads of the patch of step `step` (block first + step / nchunk, its chunk)
  auto prefetch_patch = [&](int step, auto& buf, auto& mbuf) {
    const int bi = step / nchunk, c0 = (step - bi * nchunk) * kCinChunk;
    int img, y0, x0;
    block_origin(first + bi, &img, &y0, &x0);
    const PatchSrc src = patch_src(a, img, c0);
#pragma unroll
    for (int i = 0; i < kPrefetch; ++i) {
      const uint32_t off = patch_offset(a, src, y0, x0, threadIdx.x + 256 * i);
      buf[i] = buffer_vec(src.data, off);
      if (kMask) mbuf[i] = buffer_vec(src.mask, off);
    }
  };
  // this lane's output channels' biases, read once (a load in the epilogue would wait for the
  // prefetch in flight: vector memory counters retire in order)
  // The column group's biases in LDS past the patch and weights (read in the epilogue with LDS
  // loads: a global load there would wait for the prefetch in flight, vector memory counters
  // retiring in order; registers would cost 4 NT a lane). Ordered by the first step's barriers.
  float* bias_lds = reinterpret_cast<float*>(conv_lds + kPatchElems + (kRes ? 0 : NT * 16 * kWRow));
  if (threadIdx.x < NT * 16) bias_lds[threadIdx.x] = (a.bias && (int)threadIdx.x < nn) ? a.bias[n0 + threadIdx.x] : 0.0f;
  if (kRes) {
#pragma unroll
    for (int t = 0; t < 9; ++t)
#pragma unroll
      for (int j = 0; j < NT; ++j) {
        const uint32_t off = (uint32_t)(((n0 + j * 16 + lm) * 9 + t) * a.cin + lk * 8) * 2u;
        breg[kRes ? t : 0][j] = buffer_vec(wr, j < ntile ? off : kOutOfRange);
      }
  } else {
#pragma unroll
    for (int i = 0; i < kWPrefetch; ++i) wpre[i] = load_weight_vec(a, wr, n0, 0, threadIdx.x + 256 * i, wvecs);
  }
  prefetch_patch(0, pre0, mpre0);
  if constexpr (kDepth == 2)
    if (nsteps > 1) prefetch_patch(1, pre1, mpre1);
  floatx4 acc[4][NT];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < NT; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
  // one (block, chunk) step whose patch is in `buf`; refills `buf` with step + kDepth's
  auto run_step = [&](int step, auto& buf, auto& mbuf) {
    const int bi = step / nchunk, chunk = step - bi * nchunk;
    __syncthreads();  // the previous step's LDS readers (taps or epilogue) are done
#pragma unroll
    for (int i = 0; i < kPrefetch; ++i) {
      const int k = threadIdx.x + 256 * i;
      if (k < kPatchVecs) store_patch_vec(patch, k, kMask ? apply_relu_mask(buf[i], mbuf[i]) : buf[i]);
    }
    if (!kRes) {
#pragma unroll
      for (int i = 0; i < kWPrefetch; ++i) {
        const int e = threadIdx.x + 256 * i;
        if (e < wvecs) {
          const int n = e / 36, rem = e % 36, tap = rem >> 2, part = rem & 3;
          *reinterpret_cast<uint4*>(wl + n * kWRow + tap * kCinChunk + part * 8) = wpre[i];
        }
      }
    }
    __syncthreads();
    // loads in flight during these MFMAs: the next step's weights, the patch kDepth steps ahead
    if (!kRes && step + 1 < nsteps) {
      const int nbi = (step + 1) / nchunk, nc0 = ((step + 1) - nbi * nchunk) * kCinChunk;
#pragma unroll
      for (int i = 0; i < kWPrefetch; ++i) wpre[i] = load_weight_vec(a, wr, n0, nc0, threadIdx.x + 256 * i, wvecs);
    }
    if (step + kDepth < nsteps) prefetch_patch(step + kDepth, buf, mbuf);
#pragma unroll
    for (int dx = 0; dx < 3; ++dx) {
      // every column tile runs (zero B operands past ntile): no branch breaks the MFMA chain
      uint4 b[3][NT];
#pragma unroll
      for (int dy = 0; dy < 3; ++dy)
#pragma unroll
        for (int j = 0; j < NT; ++j) {
          if (kRes) {
            b[dy][j] = breg[kRes ? dy * 3 + dx : 0][j];
          } else {
            b[dy][j] = *reinterpret_cast<const uint4*>(wl + (j * 16 + lm) * kWRow + (dy * 3 + dx) * kCinChunk + lk * 8);
            if (j >= ntile) b[dy][j] = make_uint4(0, 0, 0, 0);
          }
        }
      if constexpr (NT <= 2) {
        // the shift's 6 patch rows, all reads in flight before the first MFMA needs one
        uint4 aq[6];
#pragma unroll
        for (int r = 0; r < 6; ++r)
          aq[r] = *reinterpret_cast<const uint4*>(patch + ((wave * 4 + r) * kPatch + lm + dx) * kPStride + lk * 8);
#pragma unroll
        for (int r = 0; r < 6; ++r)
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const int dy = r - i;
            if (dy < 0 || dy > 2) continue;
#pragma unroll
            for (int j = 0; j < NT; ++j) acc[i][j] = mfma16(b[dy][j], aq[r], acc[i][j]);
          }
        // schedule: the shift's LDS reads (B operands unless resident, then the 6 rows) first,
        // then its 12 NT MFMAs
        __builtin_amdgcn_sched_group_barrier(0x100, kRes ? 6 : 6 + 3 * NT, 0);
        __builtin_amdgcn_sched_group_barrier(0x008, 12 * NT, 0);
      } else {  // 4 tiles: no registers for 6 rows in flight; each row read where it is used
#pragma unroll
        for (int r = 0; r < 6; ++r) {
          const uint4 aq = *reinterpret_cast<const uint4*>(patch + ((wave * 4 + r) * kPatch + lm + dx) * kPStride + lk * 8);
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const int dy = r - i;
            if (dy < 0 || dy > 2) continue;
#pragma unroll
            for (int j = 0; j < NT; ++j) acc[i][j] = mfma16(b[dy][j], aq, acc[i][j]);
          }
        }
      }
      __builtin_amdgcn_sched_barrier(0);  // nothing crosses into the next shift
    }
    if (chunk + 1 < nchunk) return;
    // epilogue of block bi, from registers: lane (lm, lk) holds output channels j * 16 + 4 lk .. + 3
    // of pixel lm of each of its rows - bias, ReLU, 4 BF16 in one 8-byte store (the 16 pixels x 16
    // channels of a tile row are one contiguous kilobyte with its neighbour tile); optional FP32
    int img, y0, x0;
    block_origin(first + bi, &img, &y0, &x0);
    float pool_sum[NT][4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int64_t pix = ((int64_t)img * a.h + y0 + wave * 4 + i) * a.wd + x0 + lm;
#pragma unroll
      for (int j = 0; j < NT; ++j) {
        const float4 bq4 = *reinterpret_cast<const float4*>(bias_lds + j * 16 + 4 * lk);
        const float bb[4] = {bq4.x, bq4.y, bq4.z, bq4.w};
        float v[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          v[q] = acc[i][j][q] + bb[q];
          if (a.relu) v[q] = fmaxf(v[q], 0.0f);
        }
        acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
        if (j >= ntile) continue;
        const int n = n0 + j * 16 + 4 * lk;
        const bf16x2 lo = {(__bf16)v[0], (__bf16)v[1]}, hi = {(__bf16)v[2], (__bf16)v[3]};
        if (a.y)
          *reinterpret_cast<uint2*>(a.y + pix * a.cout + n) = make_uint2(__builtin_bit_cast(uint32_t, lo), __builtin_bit_cast(uint32_t, hi));
        if (a.y32)
#pragma unroll
          for (int q = 0; q < 4; ++q)
            if (n + q < a.out_f32) a.y32[pix * a.out_f32 + n + q] = v[q];
        if (a.pool) {
          // the BF16 outputs of this row; with the row below (odd i) and the neighbouring column
          // (lane lm ^ 1), the 2x2 mean in avgpool2_kernel's order: rows, then columns, from 0
          const float r[4] = {(float)lo.x, (float)lo.y, (float)hi.x, (float)hi.y};
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const float nb = __shfl_xor(r[q], 1);
            const float left = (lm & 1) ? nb : r[q], right = (lm & 1) ? r[q] : nb;
            if ((i & 1) == 0) {
              pool_sum[j][q] = 0.0f;
              pool_sum[j][q] += left;
              pool_sum[j][q] += right;
            } else {
              pool_sum[j][q] += left;
              pool_sum[j][q] += right;
            }
          }
          if ((i & 1) && !(lm & 1)) {
            const int ho = a.h >> 1, wo = a.wd >> 1;
            const int64_t pp = ((int64_t)img * ho + ((y0 + wave * 4 + i) >> 1)) * wo + ((x0 + lm) >> 1);
            const bf16x2 plo = {(__bf16)(0.25f * pool_sum[j][0]), (__bf16)(0.25f * pool_sum[j][1])};
            const bf16x2 phi = {(__bf16)(0.25f * pool_sum[j][2]), (__bf16)(0.25f * pool_sum[j][3])};
            *reinterpret_cast<uint2*>(a.pool + pp * a.cout + n) =
                make_uint2(__builtin_bit_cast(uint32_t, plo), __builtin_bit_cast(uint32_t, phi));
          }
        }
      }
    }
  };
#pragma unroll 1
  for (int step = 0; step < nsteps; step += kDepth) {
    run_step(step, pre0, mpre0);
    if constexpr (kDepth == 2)
      if (step + 1 < nsteps) run_step(step + 1, pre1, mpre1);
  }
}

// ------------------------------------------------------------------ fused layer pair (inference)
// y32 = conv3x3(relu(conv3x3(x, w1) + b1), w2) + b2 for 32 -> 32 -> 16 channels, the U-Net's
// last two layers: the 32-channel intermediate never leaves the chip. A block of 16x16 outputs
// needs the intermediate on 18x18 pixels (one-pixel halo) and so a 20x20 input patch. The
// intermediate is computed as 21 MFMA tiles of 16 pixels: 18 rows of the block's 16 columns (rows
// 0..15 through the row-reuse path of conv3x3_kernel, rows 16 and 17 gathered) and 3 tiles of
// the 36 halo-column pixels (gathered), stored (bias, ReLU, BF16; zero outside the image) as an
// 18x18 patch in LDS, which the second convolution reads exactly as conv3x3_kernel reads its
// input patch. Every output accumulates its taps in conv3x3_kernel's order (column shift, then row
// shift), so the result is bit for bit the two layers run separately. Layer 1's B operands are
// resident in registers, layer 2's in LDS; workgroups stream contiguous block runs.
constexpr int kPairIn = kConvBlock + 4;  // 20: input patch side
constexpr int kPairInElems = kPairIn * kPairIn * kPStride;
constexpr int kPairInVecs = kPairIn * kPairIn * (kCinChunk / 8);
constexpr int kPairPrefetch = (kPairInVecs + 255) / 256;

struct PairArgs {
  const bf16_t* x;
  const bf16_t* w1;  // [32][9][32]
  const float* b1;
  const bf16_t* w2;  // [>= 16][9][32], the first 16 output channels used
  const float* b2;
  float* y32;        // [pixel][out_f32]
  int n, h, wd, out_f32;
};

// Byte offset of 16-byte vector k of the 20x20 input patch of block (y0, x0) in image `img`'s
// buffer, or kOutOfRange (zero padding, k past the patch).
MJ_DEV uint32_t pair_patch_offset(const PairArgs& a, int y0, int x0, int k) {
  const int pos = k >> 2, part = k & 3;
  const int yy = y0 - 2 + pos / kPairIn, xx = x0 - 2 + pos % kPairIn;
  // unsigned compares and bitwise ands keep the address computation free of branches
  const bool inside = (k < kPairInVecs) & ((unsigned)yy < (unsigned)a.h) & ((unsigned)xx < (unsigned)a.wd);
  const uint32_t off = (uint32_t)((yy * a.wd + xx) * kCinChunk + part * 8) * 2u;
  return inside ? off : kOutOfRange;
}

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
conv3x3_pair_kernel(PairArgs a, int bpg) {
  extern __shared__ __attribute__((aligned(16))) bf16_t pair_lds[];
  bf16_t* in_lds = pair_lds;                   // 20 x 20 positions x kPStride
  bf16_t* mid = pair_lds + kPairInElems;       // 18 x 18 positions x kPStride
  bf16_t* w2l = mid + kPatchElems;             // layer 2's weights: 16 rows of kWRow
  float* bias_lds = reinterpret_cast<float*>(w2l + 16 * kWRow);  // b1 [32], b2 [16]
  const int bx = a.wd / kConvBlock, by = a.h / kConvBlock;
  const int nblocks = a.n * bx * by;
  const int first = blockIdx.x * bpg;
  const int nsteps = min(nblocks, first + bpg) - first;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int lm = lane & 15, lk = lane >> 4;
  auto block_origin = [&](int blk, int* img, int* y0, int* x0) {
    *img = blk / (bx * by);
    const int r = blk - *img * (bx * by);
    *y0 = (r / bx) * kConvBlock;
    *x0 = (r % bx) * kConvBlock;
  };
  const int64_t image_bytes = (int64_t)a.h * a.wd * kCinChunk * 2;
  uint4 pre[kPairPrefetch];
  auto prefetch = [&](int step) {
    int img, y0, x0;
    block_origin(first + step, &img, &y0, &x0);
    const __amdgpu_buffer_rsrc_t src = buffer_of(a.x + (int64_t)img * (image_bytes / 2), image_bytes);
    // an opaque thread index: the patch coordinates are recomputed each step rather than kept
    // live (and spilled) across the loop
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
#pragma unroll
    for (int i = 0; i < kPairPrefetch; ++i) pre[i] = buffer_vec(src, pair_patch_offset(a, y0, x0, tid + 256 * i));
  };
  if (threadIdx.x < 48) bias_lds[threadIdx.x] = threadIdx.x < 32 ? a.b1[threadIdx.x] : a.b2[threadIdx.x - 32];
  // layer 1's B operands resident in registers (2 tiles of 16 output channels); layer 2's in LDS
  uint4 w1r[9][2];
  {
    const __amdgpu_buffer_rsrc_t r1 = buffer_of(a.w1, 32 * 9 * 32 * 2);
#pragma unroll
    for (int t = 0; t < 9; ++t)
#pragma unroll
      for (int j = 0; j < 2; ++j) w1r[t][j] = buffer_vec(r1, (uint32_t)(((j * 16 + lm) * 9 + t) * 32 + lk * 8) * 2u);
    for (int e = threadIdx.x; e < 16 * 36; e += 256) {  // 16 rows x 9 taps x 4 vectors
      const int row = e / 36, rem = e % 36;
      *reinterpret_cast<uint4*>(w2l + row * kWRow + (rem >> 2) * kCinChunk + (rem & 3) * 8) =
          *reinterpret_cast<const uint4*>(a.w2 + (row * 9 + (rem >> 2)) * 32 + (rem & 3) * 8);
    }
  }
  prefetch(0);
  // the gathered tiles of this wave: rows 16 and 17 (waves 0 and 1), halo-column tiles 0..2
  // (waves 2, 3, 3)
  const int nextra = wave == 3 ? 2 : 1;
#pragma unroll 1
  for (int step = 0; step < nsteps; ++step) {
    int img, y0, x0;
    block_origin(first + step, &img, &y0, &x0);
    __syncthreads();  // the previous step's readers of both patches are done
#pragma unroll
    for (int i = 0; i < kPairPrefetch; ++i) {
      const int k = threadIdx.x + 256 * i;
      if (k < kPairInVecs) *reinterpret_cast<uint4*>(in_lds + (k >> 2) * kPStride + (k & 3) * 8) = pre[i];
    }
    __syncthreads();
    if (step + 1 < nsteps) prefetch(step + 1);
    // ---- layer 1: rows 4 wave .. + 3 of the intermediate through the row-reuse path
    floatx4 acc[4][2], ext[2][2];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int e = 0; e < 2; ++e)
#pragma unroll
      for (int j = 0; j < 2; ++j) ext[e][j] = floatx4{0.f, 0.f, 0.f, 0.f};
    // the gathered pixels of this lane in the wave's extra tiles: intermediate (iy, ix), -1 pad
    int eiy[2], eix[2];
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      const int tile = wave < 2 ? wave : (wave == 2 ? 2 : 3 + e);  // 0, 1: rows 16, 17; 2..4: halo tiles
      if (tile < 2) {
        eiy[e] = 16 + tile;
        eix[e] = 1 + lm;
      } else {
        const int q = (tile - 2) * 16 + lm;  // halo pixel 0..47: column 0 (q < 18), 17 (q < 36), pad
        eiy[e] = q < 36 ? q % 18 : -1;
        eix[e] = q < 18 ? 0 : 17;
      }
      if (e >= nextra) eiy[e] = -1;
    }
#pragma unroll
    for (int dx = 0; dx < 3; ++dx) {
      uint4 aq[6];
#pragma unroll
      for (int r = 0; r < 6; ++r)
        aq[r] = *reinterpret_cast<const uint4*>(in_lds + ((wave * 4 + r) * kPairIn + 1 + lm + dx) * kPStride + lk * 8);
#pragma unroll
      for (int r = 0; r < 6; ++r)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int dy = r - i;
          if (dy < 0 || dy > 2) continue;
#pragma unroll
          for (int j = 0; j < 2; ++j) acc[i][j] = mfma16(w1r[dy * 3 + dx][j], aq[r], acc[i][j]);
        }
#pragma unroll
      for (int e = 0; e < 2; ++e) {
        if (e >= nextra) continue;
        const int iy = eiy[e] < 0 ? 0 : eiy[e];  // pad lanes read a valid position, stored nowhere
#pragma unroll
        for (int dy = 0; dy < 3; ++dy) {
          const uint4 bq = *reinterpret_cast<const uint4*>(in_lds + ((iy + dy) * kPairIn + eix[e] + dx) * kPStride + lk * 8);
#pragma unroll
          for (int j = 0; j < 2; ++j) ext[e][j] = mfma16(w1r[dy * 3 + dx][j], bq, ext[e][j]);
        }
      }
      __builtin_amdgcn_sched_barrier(0);
    }
    // layer 1 epilogue: bias, ReLU, BF16 into the intermediate patch; zero outside the image
    auto store_mid = [&](int iy, int ix, int j, const floatx4& v) {
      const int yy = y0 - 1 + iy, xx = x0 - 1 + ix;
      const bool inside = yy >= 0 && yy < a.h && xx >= 0 && xx < a.wd;
      const int c = j * 16 + 4 * lk;
      const float4 bb = *reinterpret_cast<const float4*>(bias_lds + c);
      const float b4[4] = {bb.x, bb.y, bb.z, bb.w};
      float o[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) o[q] = inside ? fmaxf(v[q] + b4[q], 0.0f) : 0.0f;
      const bf16x2 lo = {(__bf16)o[0], (__bf16)o[1]}, hi = {(__bf16)o[2], (__bf16)o[3]};
      *reinterpret_cast<uint2*>(mid + (iy * kPatch + ix) * kPStride + c) =
          make_uint2(__builtin_bit_cast(uint32_t, lo), __builtin_bit_cast(uint32_t, hi));
    };
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) store_mid(wave * 4 + i, 1 + lm, j, acc[i][j]);
#pragma unroll
    for (int e = 0; e < 2; ++e)
      if (e < nextra && eiy[e] >= 0)
#pragma unroll
        for (int j = 0; j < 2; ++j) store_mid(eiy[e], eix[e], j, ext[e][j]);
    __syncthreads();
    // ---- layer 2 (16 output channels) from the intermediate, as conv3x3_kernel
    floatx4 out[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) out[i] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int dx = 0; dx < 3; ++dx) {
      uint4 b[3], aq[6];
#pragma unroll
      for (int dy = 0; dy < 3; ++dy) b[dy] = *reinterpret_cast<const uint4*>(w2l + lm * kWRow + (dy * 3 + dx) * kCinChunk + lk * 8);
#pragma unroll
      for (int r = 0; r < 6; ++r) aq[r] = *reinterpret_cast<const uint4*>(mid + ((wave * 4 + r) * kPatch + lm + dx) * kPStride + lk * 8);
#pragma unroll
      for (int r = 0; r < 6; ++r)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int dy = r - i;
          if (dy < 0 || dy > 2) continue;
          out[i] = mfma16(b[dy], aq[r], out[i]);
        }
      __builtin_amdgcn_sched_group_barrier(0x100, 9, 0);
      __builtin_amdgcn_sched_group_barrier(0x008, 12, 0);
      __builtin_amdgcn_sched_barrier(0);
    }
    // layer 2 epilogue: the first out_f32 channels in FP32 (lane lk == 0 holds channels 0..3)
    if (lk == 0) {
      const float4 bb = *reinterpret_cast<const float4*>(bias_lds + 32);
      const float b4[4] = {bb.x, bb.y, bb.z, bb.w};
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int64_t pix = ((int64_t)img * a.h + y0 + wave * 4 + i) * a.wd + x0 + lm;
#pragma unroll
        for (int q = 0; q < 4; ++q)
          if (q < a.out_f32) a.y32[pix * a.out_f32 + q] = out[i][q] + b4[q];
      }
    }
  }
}

// ------------------------------------------------------------------ weight gradient
// part[chunk][n][tap][c] = sum over the chunk's pixels p of x[p + off(tap)][c] * g[p][n] * (mask[p][n] > 0)
// Workgroup: 16 input channels x 9 taps (M = 144) x up to 64 output channels, looping over 8x8
// pixel blocks. LDS holds the blocks transposed (pixel-contiguous), so MFMA operands reduce over
// pixels with 16-byte reads.
constexpr int kWgBlock = 8;
constexpr int kWgCin = 16;

struct WgradArgs {
  const bf16_t* x;
  const bf16_t* g;
  const bf16_t* mask;
  float* part;  // [nchunk][cout][9][cin]
  int n, h, wd, cin, cout;
  int blocks_per_chunk;
};

__global__ void __launch_bounds__(256) conv3x3_wgrad_kernel(WgradArgs a) {
  // xt[dx][c][r][k]: x at (block row r - 1, block col k - 1 + dx), r = 0..9, k = 0..7
  __shared__ __attribute__((aligned(16))) bf16_t xt[3 * kWgCin * 10 * 8];
  __shared__ __attribute__((aligned(16))) bf16_t gt[kNTile * 64];
  const int c0 = blockIdx.x * kWgCin;
  const int n0 = blockIdx.y * kNTile;
  const int nn = min(kNTile, a.cout - n0);
  const int ntile = nn / 16;
  const int chunk = blockIdx.z;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int lm = lane & 15, lk = lane >> 4;
  const int bx = a.wd / kWgBlock, by = a.h / kWgBlock;
  const int nblocks = a.n * bx * by;
  // tiles: t = tap * ntile + j; wave w owns t = w, w + 4, ...
  floatx4 acc[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) acc[i] = floatx4{0.f, 0.f, 0.f, 0.f};
  const int ntiles = 9 * ntile;
  int b_begin = chunk * a.blocks_per_chunk;
  int b_end = min(nblocks, b_begin + a.blocks_per_chunk);
  for (int blk = b_begin; blk < b_end; ++blk) {
    int img = blk / (bx * by), rem = blk % (bx * by);
    int y0 = (rem / bx) * kWgBlock, x0 = (rem % bx) * kWgBlock;
    __syncthreads();
    // x (transposed, three column shifts): 3 x 10 x 8 positions x 16 channels
    for (int e = threadIdx.x; e < 3 * 10 * 8 * 2; e += 256) {
      int half = e & 1, pos = e >> 1;
      int k = pos & 7, r = (pos >> 3) % 10, dx = pos / 80;
      int yy = y0 - 1 + r, xx = x0 - 1 + k + dx;
      uint4 v = make_uint4(0, 0, 0, 0);
      if (yy >= 0 && yy < a.h && xx >= 0 && xx < a.wd)
        v = *reinterpret_cast<const uint4*>(a.x + (((int64_t)img * a.h + yy) * a.wd + xx) * a.cin + c0 + half * 8);
      const bf16_t* pv = reinterpret_cast<const bf16_t*>(&v);
#pragma unroll
      for (int q = 0; q < 8; ++q) xt[((dx * kWgCin + half * 8 + q) * 10 + r) * 8 + k] = pv[q];
    }
    // g * mask (transposed): 64 pixels x nn channels
    for (int e = threadIdx.x; e < 64 * (nn / 8); e += 256) {
      int p = e % 64, part = e / 64;
      int yy = y0 + p / 8, xx = x0 + p % 8;
      int64_t off = (((int64_t)img * a.h + yy) * a.wd + xx) * a.cout + n0 + part * 8;
      uint4 v = *reinterpret_cast<const uint4*>(a.g + off);
      bf16_t* pv = reinterpret_cast<bf16_t*>(&v);
      if (a.mask) {
        uint4 mk = *reinterpret_cast<const uint4*>(a.mask + off);
        const bf16_t* pm = reinterpret_cast<const bf16_t*>(&mk);
#pragma unroll
        for (int q = 0; q < 8; ++q)
          if (pm[q] == 0 || (pm[q] & 0x8000u)) pv[q] = 0;
      }
#pragma unroll
      for (int q = 0; q < 8; ++q) gt[(part * 8 + q) * 64 + p] = pv[q];
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < 2; ++s) {  // K steps: block rows 4s .. 4s + 3, one row (8 pixels) per lane group
      int r = 4 * s + lk;
#pragma unroll
      for (int i = 0; i < 9; ++i) {
        int t = wave + 4 * i;
        if (t >= ntiles) break;
        int tap = t / ntile, j = t % ntile;
        int dy = tap / 3, dx = tap % 3;
        uint4 aq = *reinterpret_cast<const uint4*>(xt + ((dx * kWgCin + lm) * 10 + r + dy) * 8);
        uint4 bq = *reinterpret_cast<const uint4*>(gt + (j * 16 + lm) * 64 + r * 8);
        acc[i] = mfma16(aq, bq, acc[i]);
      }
    }
  }
  // D[m = cin][n = cout] of tile (tap, j): lane holds cin 4 lk + q, cout lm
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    int t = wave + 4 * i;
    if (t >= ntiles) break;
    int tap = t / ntile, j = t % ntile;
    int n = n0 + j * 16 + lm;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      int c = c0 + lk * 4 + q;
      a.part[(((int64_t)chunk * a.cout + n) * 9 + tap) * a.cin + c] = acc[i][q];
    }
  }
}

// Sums the chunk partials: grad[n][tap][c] (FP32) and the bias gradient from g * mask.
__global__ void wgrad_reduce_kernel(const float* part, int nchunk, int64_t count, float* grad) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  float s = 0.0f;
  for (int k = 0; k < nchunk; ++k) s += part[(int64_t)k * count + i];
  grad[i] = s;
}

// db[n] = sum_p g[p][n] * (mask[p][n] > 0): one workgroup per channel group of 64, a block of rows.
__global__ void bias_grad_kernel(const bf16_t* g, const bf16_t* mask, int64_t npix, int cout, float* part, int rows_per_block) {
  int n = blockIdx.y * 64 + (threadIdx.x & 63);
  int sub = threadIdx.x >> 6;  // 4 row groups
  int64_t p0 = (int64_t)blockIdx.x * rows_per_block;
  float s = 0.0f;
  if (n < cout) {
    for (int64_t p = p0 + sub; p < min(npix, p0 + rows_per_block); p += 4) {
      bf16_t gv = g[p * cout + n];
      if (mask) {
        bf16_t mv = mask[p * cout + n];
        if (mv == 0 || (mv & 0x8000u)) continue;
      }
      s += bf2f(gv);
    }
  }
  __shared__ float red[256];
  red[threadIdx.x] = s;
  __syncthreads();
  if (sub == 0 && n < cout) part[(int64_t)blockIdx.x * cout + n] = red[threadIdx.x] + red[threadIdx.x + 64] + red[threadIdx.x + 128] + red[threadIdx.x + 192];
}

// ------------------------------------------------------------------ elementwise layers
// 2x2 average pooling, NHWC.
__global__ void avgpool2_kernel(const bf16_t* x, bf16_t* y, int n, int h, int w, int c, int grad) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int ho = h / 2, wo = w / 2;
  if (!grad) {
    int64_t total = (int64_t)n * ho * wo * c;
    if (i >= total) return;
    int ch = i % c;
    int64_t p = i / c;
    int xo = p % wo, yo = (p / wo) % ho, img = p / ((int64_t)wo * ho);
    float s = 0.0f;
#pragma unroll
    for (int dy = 0; dy < 2; ++dy)
#pragma unroll
      for (int dx = 0; dx < 2; ++dx) s += bf2f(x[(((int64_t)img * h + 2 * yo + dy) * w + 2 * xo + dx) * c + ch]);
    y[i] = f2bf(0.25f * s);
  } else {  // x: gradient at the pooled resolution, y: gradient at full resolution
    int64_t total = (int64_t)n * h * w * c;
    if (i >= total) return;
    int ch = i % c;
    int64_t p = i / c;
    int xx = p % w, yy = (p / w) % h, img = p / ((int64_t)w * h);
    y[i] = f2bf(0.25f * bf2f(x[(((int64_t)img * ho + yy / 2) * wo + xx / 2) * c + ch]));
  }
}

// Decoder input: concat(upsample2x(low), skip) along channels; backward splits the gradient:
// the low-resolution part sums the 2x2 gradients, the skip part adds to `skip_grad_add`
// (which also receives the pooling path's gradient of the same tensor).
__global__ void upcat_kernel(const bf16_t* low, const bf16_t* skip, bf16_t* out, int n, int h, int w, int cl, int cs) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int c = cl + cs;
  int64_t total = (int64_t)n * h * w * c;
  if (i >= total) return;
  int ch = i % c;
  int64_t p = i / c;
  int xx = p % w, yy = (p / w) % h, img = p / ((int64_t)w * h);
  if (ch < cl)
    out[i] = low[(((int64_t)img * (h / 2) + yy / 2) * (w / 2) + xx / 2) * cl + ch];
  else
    out[i] = skip[p * cs + ch - cl];
}

__global__ void upcat_low_grad_kernel(const bf16_t* g, bf16_t* low_grad, int n, int h, int w, int cl, int cs) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int ho = h / 2, wo = w / 2;
  int64_t total = (int64_t)n * ho * wo * cl;
  if (i >= total) return;
  int ch = i % cl;
  int64_t p = i / cl;
  int xo = p % wo, yo = (p / wo) % ho, img = p / ((int64_t)wo * ho);
  float s = 0.0f;
#pragma unroll
  for (int dy = 0; dy < 2; ++dy)
#pragma unroll
    for (int dx = 0; dx < 2; ++dx) s += bf2f(g[(((int64_t)img * h + 2 * yo + dy) * w + 2 * xo + dx) * (cl + cs) + ch]);
  low_grad[i] = f2bf(s);
}

// skip_grad = pool_backward(pool_grad) + concat slice of cat_grad
__global__ void skip_grad_kernel(const bf16_t* pool_grad, const bf16_t* cat_grad, bf16_t* skip_grad, int n, int h, int w, int cs,
                                 int cl) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t total = (int64_t)n * h * w * cs;
  if (i >= total) return;
  int ch = i % cs;
  int64_t p = i / cs;
  int xx = p % w, yy = (p / w) % h, img = p / ((int64_t)w * h);
  float v = 0.25f * bf2f(pool_grad[(((int64_t)img * (h / 2) + yy / 2) * (w / 2) + xx / 2) * cs + ch]);
  v += bf2f(cat_grad[p * (cl + cs) + cl + ch]);
  skip_grad[i] = f2bf(v);
}

// L1 loss of the prediction (FP32 network output, first 3 channels, plus the residual base: the
// first 3 channels of the BF16 network input with `cb` channels) against FP32 targets: loss
// partial sums and the BF16 gradient sign(pred - target) / count in a cp-channel tensor.
__global__ void l1_kernel(const float* pred, const bf16_t* base, int cb, const float* target, int64_t npix, int cp, bf16_t* grad,
                          float* loss_part) {
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  float l = 0.0f;
  if (p < npix) {
    float inv = 1.0f / (float)(npix * 3);
    for (int c = 0; c < cp; ++c) {
      float gv = 0.0f;
      if (c < 3) {
        float d = pred[p * 3 + c] + (base ? bf2f(base[p * cb + c]) : 0.0f) - target[p * 3 + c];
        l += fabsf(d);
        gv = d > 0.0f ? inv : (d < 0.0f ? -inv : 0.0f);
      }
      grad[p * cp + c] = f2bf(gv);
    }
  }
  __shared__ float red[256];
  red[threadIdx.x] = l;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) loss_part[blockIdx.x] = red[0];
}

// Adam on FP32 master weights [cout][9][cin] (+ bias), refreshing both BF16 layouts.
__global__ void adam_kernel(float* w, float* m, float* v, const float* g, int64_t count, float lr, float b1, float b2, float eps,
                            float c1, float c2, bf16_t* wf, bf16_t* wd, int cin, int cout) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  float gi = g[i];
  float mi = b1 * m[i] + (1.0f - b1) * gi;
  float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  w[i] = wi;
  if (wf) {
    wf[i] = f2bf(wi);
    int c = i % cin;
    int tap = (i / cin) % 9;
    int n = i / ((int64_t)cin * 9);
    wd[((int64_t)c * 9 + (8 - tap)) * cout + n] = f2bf(wi);
  }
}

// BF16 copies of FP32 master weights (forward and mirrored-transposed layouts).
__global__ void weight_layouts_kernel(const float* w, int64_t count, int cin, int cout, bf16_t* wf, bf16_t* wd) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  float wi = w[i];
  wf[i] = f2bf(wi);
  int c = i % cin;
  int tap = (i / cin) % 9;
  int n = i / ((int64_t)cin * 9);
  wd[((int64_t)c * 9 + (8 - tap)) * cout + n] = f2bf(wi);
}

// Network input: [log1p(hdr * exposure), albedo, normal, depth / (1 + depth), 0...] as BF16 with
// `cp` channels, from the path tracer's per-world flat buffers of one camera.
__global__ void pack_input_kernel(const float* hdr, int hdr_stride, const uint16_t* albedo, const uint16_t* normal, const float* depth,
                                  int aov_stride, int adr, const float* exposure, int exposure_n, int n, int h, int w, int cp,
                                  bf16_t* out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t total = (int64_t)n * h * w;
  if (i >= total) return;
  int img = i / ((int64_t)h * w);
  int64_t pid = i - (int64_t)img * h * w;
  const float* hp = hdr + 3 * ((int64_t)img * hdr_stride + adr + pid);
  int64_t fa = (int64_t)img * aov_stride + adr + pid;
  float e = exposure[img % exposure_n];
  float vals[10];
  for (int c = 0; c < 3; ++c) vals[c] = log1pf(fmaxf(hp[c] * e, 0.0f));
  for (int c = 0; c < 3; ++c) vals[3 + c] = half_to_float(albedo[3 * fa + c]);
  for (int c = 0; c < 3; ++c) vals[6 + c] = half_to_float(normal[3 * fa + c]);
  float d = depth[fa];
  vals[9] = d / (1.0f + d);
  for (int c = 0; c < cp; ++c) out[i * cp + c] = f2bf(c < 10 ? vals[c] : 0.0f);
}

// Per-image statistics of the packed input and the fallback decision. One workgroup per image:
// stats[img] = {mean log radiance (channels 0-2), fraction of background pixels (zero normal),
// mean depth feature over hits, mean albedo (channels 3-5) over hits}; with `range` ({lo[4],
// hi[4]}), fallback[img] = 1 when a statistic lies outside [lo, hi] (ORed into the existing value
// when `accumulate`).
constexpr int kStats = 4;

__global__ void __launch_bounds__(256) guide_stats_kernel(const bf16_t* x, int64_t npix, int cp, const float* range, float* stats,
                                                          int32_t* fallback, int accumulate) {
  const int img = blockIdx.x;
  const bf16_t* xi = x + (int64_t)img * npix * cp;
  float rad = 0.0f, bg = 0.0f, dep = 0.0f, alb = 0.0f;
  for (int64_t p = threadIdx.x; p < npix; p += 256) {
    uint4 v = *reinterpret_cast<const uint4*>(xi + p * cp);  // channels 0-7
    uint32_t v89 = *reinterpret_cast<const uint32_t*>(xi + p * cp + 8);
    const bf16_t* c = reinterpret_cast<const bf16_t*>(&v);
    rad += bf2f(c[0]) + bf2f(c[1]) + bf2f(c[2]);
    bool hit = ((c[6] | c[7] | (bf16_t)(v89 & 0xffffu)) & 0x7fffu) != 0;
    if (hit) {
      dep += bf2f((bf16_t)(v89 >> 16));
      alb += bf2f(c[3]) + bf2f(c[4]) + bf2f(c[5]);
    } else {
      bg += 1.0f;
    }
  }
  __shared__ float red[4][256];
  red[0][threadIdx.x] = rad;
  red[1][threadIdx.x] = bg;
  red[2][threadIdx.x] = dep;
  red[3][threadIdx.x] = alb;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s)
      for (int k = 0; k < 4; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    float hits = fmaxf((float)npix - red[1][0], 1.0f);
    float st[kStats] = {red[0][0] / (3.0f * (float)npix), red[1][0] / (float)npix, red[2][0] / hits, red[3][0] / (3.0f * hits)};
    for (int k = 0; k < kStats; ++k) stats[img * kStats + k] = st[k];
    if (range && fallback) {
      int out = 0;
      for (int k = 0; k < kStats; ++k) out |= (st[k] < range[k] || st[k] > range[kStats + k]) ? 1 : 0;
      fallback[img] = accumulate ? (fallback[img] | out) : out;
    }
  }
}

// Denoised HDR from the network's log-space FP32 output (+ residual base): expm1, / exposure; and
// the display RGB. Images with skip[img] != 0 are left untouched.
__global__ void unpack_output_kernel(const float* pred, const bf16_t* base, int cb, int n, int h, int w, const float* exposure,
                                     int exposure_n, int tone_mapper, float* hdr, int hdr_stride, uint32_t* rgb, int rgb_stride,
                                     int adr, const int32_t* skip) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t total = (int64_t)n * h * w;
  if (i >= total) return;
  int img = i / ((int64_t)h * w);
  if (skip && skip[img]) return;
  int64_t pid = i - (int64_t)img * h * w;
  float e = exposure[img % exposure_n];
  float lv[3];
  for (int k = 0; k < 3; ++k) lv[k] = pred[3 * i + k] + (base ? bf2f(base[i * cb + k]) : 0.0f);
  V3 c = V3{expm1f(fmaxf(lv[0], 0.0f)), expm1f(fmaxf(lv[1], 0.0f)), expm1f(fmaxf(lv[2], 0.0f))};
  if (hdr) st3(hdr + 3 * ((int64_t)img * hdr_stride + adr + pid), c / e);
  if (rgb) rgb[(int64_t)img * rgb_stride + adr + pid] = encode_display(display_transform(c, tone_mapper));
}

// MFMA operand layout probe: D = A (16 x 32) . B (32 x 16) through the same lane mapping as the
// convolutions, for tests.
__global__ void mfma_probe_kernel(const bf16_t* A, const bf16_t* B, float* D) {
  int l = threadIdx.x;
  int lm = l & 15, lk = l >> 4;
  uint4 aq = *reinterpret_cast<const uint4*>(A + lm * 32 + lk * 8);
  bf16_t bv[8];
  for (int i = 0; i < 8; ++i) bv[i] = B[(lk * 8 + i) * 16 + lm];
  uint4 bq = *reinterpret_cast<const uint4*>(bv);
  floatx4 acc = mfma16(aq, bq, floatx4{0.f, 0.f, 0.f, 0.f});
  for (int q = 0; q < 4; ++q) D[(lk * 4 + q) * 16 + lm] = acc[q];
}

}  // namespace mjpbr

using namespace mjpbr;

namespace {
inline unsigned nblk(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }
}  // namespace

namespace {
constexpr int kConvCus = 256;  // MI355X compute units
int conv_launch(const ConvArgs& a, void* stream) {
  const int nblocks = a.n * (a.h / kConvBlock) * (a.wd / kConvBlock);
  const int tiles = a.cout / 16;
  const int nt = tiles >= 3 ? 4 : tiles;  // accumulator width of the kernel instance
  const int nn = a.cout < nt * 16 ? a.cout : nt * 16;
  const bool res = a.cin == kCinChunk && nt <= 2;  // B operands resident in registers
  // patch, weights streamed through LDS (the instance's NT tiles), biases (2 bf16 slots a float)
  const size_t elems = kPatchElems + (res ? 0 : (size_t)nt * 16 * kWRow) + 2 * nt * 16;
  // persistent workgroups: as many as fit on the compute units at the instance's occupancy per
  // column group, each streaming a contiguous run of blocks
  const int colgroups = (a.cout + nt * 16 - 1) / (nt * 16);
  const int target = std::max(1, kConvCus * 2 / colgroups);
  const int bpg = (nblocks + target - 1) / target;
  dim3 grid((unsigned)((nblocks + bpg - 1) / bpg), (unsigned)colgroups);
  hipStream_t st = (hipStream_t)stream;
  const size_t lds = elems * sizeof(bf16_t);
  using ConvFn = void (*)(ConvArgs, int);
  const bool m = a.mask != nullptr;
  ConvFn fn;
  if (nt == 1)
    fn = res ? (m ? conv3x3_kernel<1, true, true> : conv3x3_kernel<1, false, true>)
             : (m ? conv3x3_kernel<1, true, false> : conv3x3_kernel<1, false, false>);
  else if (nt == 2)
    fn = res ? (m ? conv3x3_kernel<2, true, true> : conv3x3_kernel<2, false, true>)
             : (m ? conv3x3_kernel<2, true, false> : conv3x3_kernel<2, false, false>);
  else
    fn = m ? conv3x3_kernel<4, true, false> : conv3x3_kernel<4, false, false>;
  hipLaunchKernelGGL(fn, grid, dim3(256), lds, st, a, bpg);
  return check_launch("conv3x3_kernel");
}
}  // namespace

extern "C" int mjpbr_unet_conv3x3(const uint16_t* x, const uint16_t* mask, const uint16_t* w, const float* bias, uint16_t* y,
                                  float* y32, int out_f32, int n, int h, int wd, int cin, int cout, int relu, void* stream) {
  if (h % kConvBlock || wd % kConvBlock || cin % kCinChunk || cout % 16 || n < 1)
    return set_error(MJPBR_EINVAL, "unet_conv3x3: need h, w multiples of 16, cin of 32, cout of 16 (got %dx%d, %d -> %d)", h, wd,
                     cin, cout);
  if (!y && (!y32 || out_f32 < 1)) return set_error(MJPBR_EINVAL, "unet_conv3x3: y may only be NULL when y32 receives channels");
  ConvArgs a{x, nullptr, mask, w, bias, y, y32, n, h, wd, cin, cout, relu, out_f32, 0, nullptr};
  return conv_launch(a, stream);
}

extern "C" int mjpbr_unet_conv3x3_pool(const uint16_t* x, const uint16_t* w, const float* bias, uint16_t* y, uint16_t* pool, int n, int h,
                                       int wd, int cin, int cout, int relu, void* stream) {
  if (h % kConvBlock || wd % kConvBlock || cin % kCinChunk || cout % 16 || n < 1 || !pool)
    return set_error(MJPBR_EINVAL, "unet_conv3x3_pool: need h, w multiples of 16, cin of 32, cout of 16 and a pool output (got %dx%d, %d -> %d)",
                     h, wd, cin, cout);
  ConvArgs a{x, nullptr, nullptr, w, bias, y, nullptr, n, h, wd, cin, cout, relu, 0, 0, pool};
  return conv_launch(a, stream);
}

extern "C" int mjpbr_unet_conv3x3_pair(const uint16_t* x, const uint16_t* w1, const float* b1, const uint16_t* w2, const float* b2, float* y32,
                                       int out_f32, int n, int h, int wd, void* stream) {
  if (h % kConvBlock || wd % kConvBlock || n < 1 || !x || !w1 || !b1 || !w2 || !b2 || !y32 || out_f32 < 1 || out_f32 > 4)
    return set_error(MJPBR_EINVAL, "unet_conv3x3_pair: need h, w multiples of 16, every buffer, and 1..4 FP32 channels (got %dx%d, %d)", h, wd,
                     out_f32);
  PairArgs a{reinterpret_cast<const bf16_t*>(x), reinterpret_cast<const bf16_t*>(w1), b1, reinterpret_cast<const bf16_t*>(w2), b2, y32,
             n, h, wd, out_f32};
  const int nblocks = n * (h / kConvBlock) * (wd / kConvBlock);
  const int target = kConvCus * 2;
  const int bpg = (nblocks + target - 1) / target;
  const size_t lds = (size_t)(kPairInElems + kPatchElems + 16 * kWRow) * sizeof(bf16_t) + 48 * sizeof(float);
  hipLaunchKernelGGL(conv3x3_pair_kernel, dim3((unsigned)((nblocks + bpg - 1) / bpg)), dim3(256), lds, (hipStream_t)stream, a, bpg);
  return check_launch("conv3x3_pair_kernel");
}

extern "C" int mjpbr_unet_conv3x3_upcat(const uint16_t* low, const uint16_t* skip, int c_split, const uint16_t* w, const float* bias,
                                        uint16_t* y, int n, int h, int wd, int cin, int cout, int relu, void* stream) {
  if (h % kConvBlock || wd % kConvBlock || cin % kCinChunk || cout % 16 || n < 1 || c_split <= 0 || c_split >= cin ||
      c_split % kCinChunk || h % 2 || wd % 2)
    return set_error(MJPBR_EINVAL, "unet_conv3x3_upcat: bad shape %dx%d, %d (%d low) -> %d", h, wd, cin, c_split, cout);
  ConvArgs a{low, skip, nullptr, w, bias, y, nullptr, n, h, wd, cin, cout, relu, 0, c_split, nullptr};
  return conv_launch(a, stream);
}

extern "C" int64_t mjpbr_unet_wgrad_workspace(int n, int h, int wd, int cin, int cout, int nchunk) {
  (void)n; (void)h; (void)wd;
  return (int64_t)nchunk * cout * 9 * cin;
}

extern "C" int mjpbr_unet_conv3x3_wgrad(const uint16_t* x, const uint16_t* g, const uint16_t* mask, float* part, int nchunk,
                                        float* grad, float* bias_part, float* bias_grad, int n, int h, int wd, int cin, int cout,
                                        void* stream) {
  if (h % kWgBlock || wd % kWgBlock || cin % kWgCin || cout % 16 || nchunk < 1)
    return set_error(MJPBR_EINVAL, "unet_conv3x3_wgrad: bad shape %dx%d, %d -> %d", h, wd, cin, cout);
  int nblocks = n * (h / kWgBlock) * (wd / kWgBlock);
  int per = (nblocks + nchunk - 1) / nchunk;
  WgradArgs a{x, g, mask, part, n, h, wd, cin, cout, per};
  dim3 grid((unsigned)(cin / kWgCin), (unsigned)((cout + kNTile - 1) / kNTile), (unsigned)nchunk);
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(conv3x3_wgrad_kernel, grid, dim3(256), 0, st, a);
  int rc = check_launch("conv3x3_wgrad_kernel");
  if (rc) return rc;
  int64_t count = (int64_t)cout * 9 * cin;
  hipLaunchKernelGGL(wgrad_reduce_kernel, dim3(nblk(count, 256)), dim3(256), 0, st, part, nchunk, count, grad);
  rc = check_launch("wgrad_reduce_kernel");
  if (rc || !bias_grad) return rc;
  int64_t npix = (int64_t)n * h * wd;
  int rows = 4096;
  int nb = (int)((npix + rows - 1) / rows);
  hipLaunchKernelGGL(bias_grad_kernel, dim3((unsigned)nb, (unsigned)((cout + 63) / 64)), dim3(256), 0, st, g, mask, npix, cout,
                     bias_part, rows);
  rc = check_launch("bias_grad_kernel");
  if (rc) return rc;
  // reduce nb partials of cout values each
  hipLaunchKernelGGL(wgrad_reduce_kernel, dim3(nblk(cout, 256)), dim3(256), 0, st, bias_part, nb, (int64_t)cout, bias_grad);
  return check_launch("wgrad_reduce_kernel");
}

extern "C" int64_t mjpbr_unet_bias_workspace(int n, int h, int wd, int cout) {
  int64_t npix = (int64_t)n * h * wd;
  return ((npix + 4095) / 4096) * cout;
}

extern "C" int mjpbr_unet_avgpool2(const uint16_t* x, uint16_t* y, int n, int h, int wd, int c, int grad, void* stream) {
  int64_t total = grad ? (int64_t)n * h * wd * c : (int64_t)n * (h / 2) * (wd / 2) * c;
  hipLaunchKernelGGL(avgpool2_kernel, dim3(nblk(total, 256)), dim3(256), 0, (hipStream_t)stream, x, y, n, h, wd, c, grad);
  return check_launch("avgpool2_kernel");
}

extern "C" int mjpbr_unet_upcat(const uint16_t* low, const uint16_t* skip, uint16_t* out, int n, int h, int wd, int cl, int cs,
                                void* stream) {
  int64_t total = (int64_t)n * h * wd * (cl + cs);
  hipLaunchKernelGGL(upcat_kernel, dim3(nblk(total, 256)), dim3(256), 0, (hipStream_t)stream, low, skip, out, n, h, wd, cl, cs);
  return check_launch("upcat_kernel");
}

extern "C" int mjpbr_unet_upcat_low_grad(const uint16_t* g, uint16_t* low_grad, int n, int h, int wd, int cl, int cs,
                                         void* stream) {
  int64_t total = (int64_t)n * (h / 2) * (wd / 2) * cl;
  hipLaunchKernelGGL(upcat_low_grad_kernel, dim3(nblk(total, 256)), dim3(256), 0, (hipStream_t)stream, g, low_grad, n, h, wd, cl, cs);
  return check_launch("upcat_low_grad_kernel");
}

extern "C" int mjpbr_unet_skip_grad(const uint16_t* pool_grad, const uint16_t* cat_grad, uint16_t* skip_grad, int n, int h, int wd,
                                    int cs, int cl, void* stream) {
  int64_t total = (int64_t)n * h * wd * cs;
  hipLaunchKernelGGL(skip_grad_kernel, dim3(nblk(total, 256)), dim3(256), 0, (hipStream_t)stream, pool_grad, cat_grad, skip_grad, n,
                     h, wd, cs, cl);
  return check_launch("skip_grad_kernel");
}

extern "C" int mjpbr_unet_l1(const float* pred, const uint16_t* base, int cb, const float* target, int64_t npix, int cp,
                             uint16_t* grad, float* loss_part, void* stream) {
  hipLaunchKernelGGL(l1_kernel, dim3(nblk(npix, 256)), dim3(256), 0, (hipStream_t)stream, pred, base, cb, target, npix, cp, grad,
                     loss_part);
  return check_launch("l1_kernel");
}

extern "C" int mjpbr_unet_adam(float* w, float* m, float* v, const float* g, int64_t count, float lr, float b1, float b2, float eps,
                               int step, uint16_t* wf, uint16_t* wd, int cin, int cout, void* stream) {
  float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
  hipLaunchKernelGGL(adam_kernel, dim3(nblk(count, 256)), dim3(256), 0, (hipStream_t)stream, w, m, v, g, count, lr, b1, b2, eps, c1,
                     c2, wf, wd, cin, cout);
  return check_launch("adam_kernel");
}

extern "C" int mjpbr_unet_weight_layouts(const float* w, int cin, int cout, uint16_t* wf, uint16_t* wd, void* stream) {
  int64_t count = (int64_t)cout * 9 * cin;
  hipLaunchKernelGGL(weight_layouts_kernel, dim3(nblk(count, 256)), dim3(256), 0, (hipStream_t)stream, w, count, cin, cout, wf, wd);
  return check_launch("weight_layouts_kernel");
}

extern "C" int mjpbr_unet_pack_input(const float* hdr, int hdr_stride, const uint16_t* albedo, const uint16_t* normal,
                                     const float* depth, int aov_stride, int adr, const float* exposure, int exposure_n, int n, int h,
                                     int wd, int cp, uint16_t* out, void* stream) {
  int64_t total = (int64_t)n * h * wd;
  hipLaunchKernelGGL(pack_input_kernel, dim3(nblk(total, 256)), dim3(256), 0, (hipStream_t)stream, hdr, hdr_stride, albedo, normal,
                     depth, aov_stride, adr, exposure, exposure_n, n, h, wd, cp, out);
  return check_launch("pack_input_kernel");
}

extern "C" int mjpbr_unet_unpack_output(const float* pred, const uint16_t* base, int cb, int n, int h, int wd,
                                        const float* exposure, int exposure_n, int tone_mapper, float* hdr, int hdr_stride,
                                        uint32_t* rgb, int rgb_stride, int adr, const int32_t* skip, void* stream) {
  int64_t total = (int64_t)n * h * wd;
  hipLaunchKernelGGL(unpack_output_kernel, dim3(nblk(total, 256)), dim3(256), 0, (hipStream_t)stream, pred, base, cb, n, h, wd, exposure,
                     exposure_n, tone_mapper, hdr, hdr_stride, rgb, rgb_stride, adr, skip);
  return check_launch("unpack_output_kernel");
}

extern "C" int mjpbr_unet_guide_stats(const uint16_t* x, int n, int h, int wd, int cp, const float* range, float* stats,
                                      int32_t* fallback, int accumulate, void* stream) {
  if (cp < 10 || cp % 8 || n < 1) return set_error(MJPBR_EINVAL, "unet_guide_stats: need cp >= 10 and a multiple of 8 (got %d)", cp);
  hipLaunchKernelGGL(guide_stats_kernel, dim3((unsigned)n), dim3(256), 0, (hipStream_t)stream, x, (int64_t)h * wd, cp, range, stats,
                     fallback, accumulate);
  return check_launch("guide_stats_kernel");
}

extern "C" int mjpbr_unet_mfma_probe(const uint16_t* A, const uint16_t* B, float* D, void* stream) {
  hipLaunchKernelGGL(mfma_probe_kernel, dim3(1), dim3(64), 0, (hipStream_t)stream, A, B, D);
  return check_launch("mfma_probe_kernel");
}
