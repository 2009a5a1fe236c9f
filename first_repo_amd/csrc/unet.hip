// Learned pressure preconditioner: a 3-level 3-D U-Net (16/32/64 channels,
// 3x3x3 convolutions) whose convolutions run on the gfx950 matrix cores in
// BF16 with FP32 accumulation (v_mfma_f32_16x16x32_bf16), trained on the GPU
// by the kernels below (forward, backward-data, backward-weights, Adam).
//
// No reference counterpart: the reference solves the variable-coefficient
// pressure equation with Hypre only (poisson.cpp:115-388,
// poisson_IO.cpp:247-452).  The network maps (RHS, 1/rho, previous P) to the
// pressure increment; mg_pcg.hip uses it for the initial guess and inside a
// flexible PCG, with the plain multigrid preconditioner as the fallback.
//
// Layout: activations are channels-last BF16, pixel p = (z*ny + y)*nx + x of
// one level's interior grid, element p*stride + offset + c (the two
// concatenations of the U-Net are channel ranges of one buffer).  Gradients
// of activations are FP32 in the same layout.  Master weights are FP32,
// conv weights W[tap][ci][co] with tap = (dz+1)*9 + (dy+1)*3 + (dx+1), zero
// padding at the domain edge.
//
// Matrix-core tiling of a 3x3x3 convolution as an implicit GEMM
// D[co][pixel] = sum_k A[co][k] * B[k][pixel], k = (tap, ci):
//   * one wave = 64 consecutive x pixels of one (y, z) row (4 N-blocks of 16)
//     x all output channels (COUT/16 M-blocks), a workgroup = 4 rows;
//   * K runs in chunks of 32: two taps of 16 channels (CIN = 16), one tap of
//     32 (CIN = 32) or half a tap of 64 (CIN = 64); the 27th tap of CIN = 16
//     is padded with zero weights;
//   * operand lane layout of the 16x16x32 instruction: lane l holds
//     A[l%16][8(l/16) .. +7], B[8(l/16) .. +7][l%16] and returns
//     D[4(l/16) + i][l%16], i = 0..3.  B is therefore 8 consecutive channels
//     of one input pixel (one 16-byte load), A is pre-packed in lane order
//     (one 16-byte load per lane per chunk), and D is 4 consecutive output
//     channels of one pixel (one 8-byte store).
#include <cmath>
#include <cstring>
#include <vector>

#include "fr_common.hpp"
#include "fr_internal.hpp"

namespace fr {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef unsigned short u16;

struct U3 {
  int nx, ny, nz;
  long n;
};

static U3 make_u3(int nx, int ny, int nz) { return U3{nx, ny, nz, (long)nx * ny * nz}; }

// (x, y, z) of pixel t = (z*ny + y)*nx + x in 32-bit arithmetic (grids here
// have fewer than 2^31 pixels).
__device__ __forceinline__ void xyz(long t, int nx, int ny, int& x, int& y, int& z) {
  const unsigned u = (unsigned)t, q = u / (unsigned)nx;
  x = (int)(u - q * (unsigned)nx);
  y = (int)(q % (unsigned)ny);
  z = (int)(q / (unsigned)ny);
}

// FP32 -> BF16, round to nearest even (NaN kept quiet).
__host__ __device__ __forceinline__ u16 f2bf(float f) {
  uint32_t u = __builtin_bit_cast(uint32_t, f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (u16)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
  u += 0x7fffu + ((u >> 16) & 1u);
  return (u16)(u >> 16);
}
__host__ __device__ __forceinline__ float bf2f(u16 h) {
  return __builtin_bit_cast(float, (uint32_t)h << 16);
}
__host__ __device__ __forceinline__ float bfround(float f) { return bf2f(f2bf(f)); }

template <int CIN>
struct Chunks {
  static constexpr int n = CIN == 16 ? 14 : 27 * CIN / 32;
};

// Tap and first channel of chunk c for lane group g (g = lane / 16).
__host__ __device__ __forceinline__ int chunk_tap(int CIN, int c, int g, int& ci0) {
  if (CIN == 16) {
    ci0 = 8 * (g & 1);
    return 2 * c + (g >> 1);
  }
  if (CIN == 32) {
    ci0 = 8 * g;
    return c;
  }
  ci0 = 32 * (c & 1) + 8 * g;
  return c >> 1;
}

__host__ __device__ inline int nchunks(int CIN) { return CIN == 16 ? 14 : 27 * CIN / 32; }

// Packs W[tap][ci][co] (FP32, ci < cin_w, co < cout_w) into the lane order of
// the A operand for a convolution with CIN input and COUT output channels.
// With `flip` the packed operator is the backward-data one: input channels
// are the forward outputs and tap t reads W[26 - t][co][ci].
__global__ void k_pack_conv(const float* __restrict__ W, int cin_w, int cout_w, int CIN, int COUT,
                            int flip, u16* __restrict__ out) {
  const int NB = COUT / 16, NCH = nchunks(CIN);
  const long total = (long)NCH * NB * 64 * 8;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  const int e = (int)(t & 7), lane = (int)((t >> 3) & 63);
  const long cb = t >> 9;
  const int nb = (int)(cb % NB), c = (int)(cb / NB);
  const int g = lane >> 4, m = lane & 15;
  int ci0;
  const int tap = chunk_tap(CIN, c, g, ci0);
  const int ci = ci0 + e, co = nb * 16 + m;
  float v = 0.f;
  if (tap < 27) {
    if (!flip)
      v = W[((long)tap * cin_w + ci) * cout_w + co];
    else  // forward weight of tap 26-tap, forward input channel co, forward output ci
      v = W[((long)(26 - tap) * cin_w + co) * cout_w + ci];
  }
  out[t] = f2bf(v);
}

// MODE 0: BF16 out = act(acc + bias); MODE 1: FP32 out = acc; MODE 2: FP32 out += acc.
template <int CIN, int COUT, int MODE, bool RELU>
__global__ void __launch_bounds__(256) k_conv3(U3 d, const u16* __restrict__ in, int is, int io,
                                               const uint4* __restrict__ wp,
                                               const float* __restrict__ bias, void* out, int os,
                                               int oo) {
  constexpr int NB = COUT / 16, NCH = Chunks<CIN>::n;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int y = blockIdx.y * 4 + w, z = blockIdx.z;
  if (y >= d.ny) return;  // whole wave
  const int x0 = blockIdx.x * 64;
  const int g = lane >> 4, m = lane & 15;
  floatx4 acc[4][NB];
#pragma unroll
  for (int pb = 0; pb < 4; pb++)
#pragma unroll
    for (int nb = 0; nb < NB; nb++) acc[pb][nb] = floatx4{0.f, 0.f, 0.f, 0.f};
  for (int c = 0; c < NCH; c++) {
    int ci0;
    const int tap = chunk_tap(CIN, c, g, ci0);
    const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
    const int yy = y + dy, zz = z + dz;
    const bool rowok = tap < 27 && yy >= 0 && yy < d.ny && zz >= 0 && zz < d.nz;
    const long rowbase = rowok ? ((long)zz * d.ny + yy) * d.nx : 0;
    bf16x8 b[4];
#pragma unroll
    for (int pb = 0; pb < 4; pb++) {
      const int xx = x0 + 16 * pb + m + dx;
      uint4 v = make_uint4(0u, 0u, 0u, 0u);
      if (rowok && xx >= 0 && xx < d.nx)
        v = *reinterpret_cast<const uint4*>(in + (rowbase + xx) * is + io + ci0);
      b[pb] = __builtin_bit_cast(bf16x8, v);
    }
#pragma unroll
    for (int nb = 0; nb < NB; nb++) {
      const bf16x8 a = __builtin_bit_cast(bf16x8, wp[((long)c * NB + nb) * 64 + lane]);
#pragma unroll
      for (int pb = 0; pb < 4; pb++)
        acc[pb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[pb], acc[pb][nb], 0, 0, 0);
    }
  }
  const long rowp = ((long)z * d.ny + y) * d.nx;
#pragma unroll
  for (int pb = 0; pb < 4; pb++) {
    const int xx = x0 + 16 * pb + m;
    if (xx >= d.nx) continue;
    const long p = rowp + xx;
#pragma unroll
    for (int nb = 0; nb < NB; nb++) {
      const int co0 = nb * 16 + 4 * g;
      if (MODE == 0) {
        u16 h[4];
#pragma unroll
        for (int i = 0; i < 4; i++) {
          float v = acc[pb][nb][i] + bias[co0 + i];
          if (RELU) v = v > 0.f ? v : 0.f;
          h[i] = f2bf(v);
        }
        *reinterpret_cast<uint2*>(static_cast<u16*>(out) + p * os + oo + co0) =
            make_uint2(h[0] | ((uint32_t)h[1] << 16), h[2] | ((uint32_t)h[3] << 16));
      } else {
        float4* o = reinterpret_cast<float4*>(static_cast<float*>(out) + p * os + oo + co0);
        float4 v = make_float4(acc[pb][nb][0], acc[pb][nb][1], acc[pb][nb][2], acc[pb][nb][3]);
        if (MODE == 2) {
          const float4 q = *o;
          v.x += q.x;
          v.y += q.y;
          v.z += q.z;
          v.w += q.w;
        }
        *o = v;
      }
    }
  }
}

// Partial weight gradients of a 3x3x3 convolution over a chunk of output
// rows: part[chunk][tap][ci][co] = sum over the chunk's pixels p of
// x[p + off(tap)][ci] * dp[p][co], on the matrix cores with the 32 pixels of
// one x tile as K.  The input halo (9 rows x 34 pixels x CIN) and the dp tile
// (32 x COUT) are staged in LDS; each wave owns NPW (tap, ci-block,
// co-block) products.
template <int CIN, int COUT, int NPW>
__global__ void __launch_bounds__(256) k_conv3_dw(U3 d, const u16* __restrict__ x, int xs, int xo,
                                                  const u16* __restrict__ dp, int rows_per_chunk,
                                                  float* __restrict__ part) {
  constexpr int CB = CIN / 16, OB = COUT / 16, BP = 27 * CB * OB;
  __shared__ __align__(16) u16 lx[9][34][CIN];
  __shared__ __align__(16) u16 ld[32][COUT];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, g = lane >> 4, m = lane & 15;
  const int chunk = blockIdx.x;
  const int bp0 = (blockIdx.y * 4 + w) * NPW;
  floatx4 acc[NPW];
#pragma unroll
  for (int q = 0; q < NPW; q++) acc[q] = floatx4{0.f, 0.f, 0.f, 0.f};
  const int rows = d.ny * d.nz;
  const int r0 = chunk * rows_per_chunk;
  const int r1 = min(rows, r0 + rows_per_chunk);
  for (int r = r0; r < r1; r++) {
    const int y = r % d.ny, z = r / d.ny;
    for (int x0 = 0; x0 < d.nx; x0 += 32) {
      __syncthreads();
      for (int t = threadIdx.x; t < 9 * 34 * (CIN / 8); t += 256) {
        const int v8 = t % (CIN / 8), rest = t / (CIN / 8);
        const int px = rest % 34, hr = rest / 34;
        const int xx = x0 - 1 + px, yy = y + hr % 3 - 1, zz = z + hr / 3 - 1;
        uint4 v = make_uint4(0u, 0u, 0u, 0u);
        if (xx >= 0 && xx < d.nx && yy >= 0 && yy < d.ny && zz >= 0 && zz < d.nz)
          v = *reinterpret_cast<const uint4*>(x + (((long)zz * d.ny + yy) * d.nx + xx) * xs + xo +
                                              8 * v8);
        *reinterpret_cast<uint4*>(&lx[hr][px][8 * v8]) = v;
      }
      for (int t = threadIdx.x; t < 32 * (COUT / 8); t += 256) {
        const int v8 = t % (COUT / 8), px = t / (COUT / 8);
        const int xx = x0 + px;
        uint4 v = make_uint4(0u, 0u, 0u, 0u);
        if (xx < d.nx)
          v = *reinterpret_cast<const uint4*>(dp + (((long)z * d.ny + y) * d.nx + xx) * COUT +
                                              8 * v8);
        *reinterpret_cast<uint4*>(&ld[px][8 * v8]) = v;
      }
      __syncthreads();
#pragma unroll
      for (int q = 0; q < NPW; q++) {
        const int bp = bp0 + q;
        if (bp < BP) {
          const int tap = bp / (CB * OB), rem = bp % (CB * OB);
          const int cib = rem / OB, cob = rem % OB;
          const int hr = tap / 3, dx = tap % 3 - 1;
          uint4 av, bv;
          u16* ah = reinterpret_cast<u16*>(&av);
          u16* bh = reinterpret_cast<u16*>(&bv);
#pragma unroll
          for (int e = 0; e < 8; e++) {
            ah[e] = ld[8 * g + e][cob * 16 + m];
            bh[e] = lx[hr][8 * g + e + dx + 1][cib * 16 + m];
          }
          acc[q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, av),
                                                           __builtin_bit_cast(bf16x8, bv), acc[q],
                                                           0, 0, 0);
        }
      }
    }
  }
#pragma unroll
  for (int q = 0; q < NPW; q++) {
    const int bp = bp0 + q;
    if (bp >= BP) continue;
    const int tap = bp / (CB * OB), rem = bp % (CB * OB);
    const int cib = rem / OB, cob = rem % OB;
    const int ci = cib * 16 + m;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int co = cob * 16 + 4 * g + i;
      part[(((long)chunk * 27 + tap) * CIN + ci) * COUT + co] = acc[q][i];
    }
  }
}

// out[j] (+)= sum over k < nparts of part[k*m + j], in k order.
__global__ void k_sum_parts(int nparts, long m, const float* __restrict__ part,
                            float* __restrict__ out, int accumulate) {
  const long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= m) return;
  float s = 0.f;
  for (int k = 0; k < nparts; k++) s += part[(long)k * m + j];
  out[j] = accumulate ? out[j] + s : s;
}

// dp = bf16(g * (y > 0)) (compact [n][C]); y == nullptr: no mask.
__global__ void k_relu_bwd(long n, int C, const float* __restrict__ gr, int gs, int go,
                           const u16* __restrict__ y, int ys, int yo, u16* __restrict__ dp) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * C) return;
  const long p = t / C;
  const int c = (int)(t - p * C);
  float v = gr[p * gs + go + c];
  if (y && !(bf2f(y[p * ys + yo + c]) > 0.f)) v = 0.f;
  dp[t] = f2bf(v);
}

// Per-block channel sums of g * (y > 0) (bias gradients): part[block][c].
// Each block covers `per` consecutive pixels.
__global__ void __launch_bounds__(256) k_chan_sum(long n, int C, long per,
                                                  const float* __restrict__ gr, int gs, int go,
                                                  const u16* __restrict__ y, int ys, int yo,
                                                  float* __restrict__ part) {
  __shared__ float red[256];
  const int c = threadIdx.x % C, lanes = 256 / C, pl = threadIdx.x / C;
  const long p0 = (long)blockIdx.x * per, p1 = min(n, p0 + per);
  float s = 0.f;
  if (pl < lanes)
    for (long p = p0 + pl; p < p1; p += lanes) {
      float v = gr[p * gs + go + c];
      if (y && !(bf2f(y[p * ys + yo + c]) > 0.f)) v = 0.f;
      s += v;
    }
  red[threadIdx.x] = pl < lanes ? s : 0.f;
  __syncthreads();
  if (threadIdx.x < C) {
    float t = 0.f;
    for (int k = 0; k < lanes; k++) t += red[k * C + threadIdx.x];
    part[(long)blockIdx.x * C + threadIdx.x] = t;
  }
}

// Per-block sums of a[p][c] * g[p] (the 1x1 output conv's weight gradient).
__global__ void __launch_bounds__(256) k_chan_dot(long n, int C, long per,
                                                  const u16* __restrict__ a,
                                                  const float* __restrict__ gr,
                                                  float* __restrict__ part) {
  __shared__ float red[256];
  const int c = threadIdx.x % C, lanes = 256 / C, pl = threadIdx.x / C;
  const long p0 = (long)blockIdx.x * per, p1 = min(n, p0 + per);
  float s = 0.f;
  if (pl < lanes)
    for (long p = p0 + pl; p < p1; p += lanes) s += bf2f(a[p * C + c]) * gr[p];
  red[threadIdx.x] = pl < lanes ? s : 0.f;
  __syncthreads();
  if (threadIdx.x < C) {
    float t = 0.f;
    for (int k = 0; k < lanes; k++) t += red[k * C + threadIdx.x];
    part[(long)blockIdx.x * C + threadIdx.x] = t;
  }
}

// 2x2x2 average pooling: coarse dims dc, fine dims 2*dc.
__global__ void k_pool(U3 dc, int C, const u16* __restrict__ in, int is, int io,
                       u16* __restrict__ out) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= dc.n * C) return;
  const long q = t / C;
  const int c = (int)(t - q * C);
  int x, y, z;
  xyz(q, dc.nx, dc.ny, x, y, z);
  const int fx = 2 * dc.nx, fy = 2 * dc.ny;
  float s = 0.f;
#pragma unroll
  for (int o = 0; o < 8; o++) {
    const long p = ((long)(2 * z + (o >> 2)) * fy + (2 * y + ((o >> 1) & 1))) * fx + 2 * x + (o & 1);
    s += bf2f(in[p * is + io + c]);
  }
  out[t] = f2bf(s * 0.125f);
}

// Adjoint of k_pool: gin[child] += gout[parent] / 8 (fine dims 2*dc).
__global__ void k_pool_bwd(U3 dc, int C, const float* __restrict__ gout, float* __restrict__ gin,
                           int gs, int go) {
  const long fn = 8 * dc.n;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= fn * C) return;
  const long p = t / C;
  const int c = (int)(t - p * C);
  const int fx = 2 * dc.nx, fy = 2 * dc.ny;
  int x, y, z;
  xyz(p, fx, fy, x, y, z);
  const long q = ((long)(z >> 1) * dc.ny + (y >> 1)) * dc.nx + (x >> 1);
  gin[p * gs + go + c] += gout[q * C + c] * 0.125f;
}

// 2x2x2 stride-2 transposed convolution (coarse dc -> fine 2*dc):
// out[2q + o][co] = b[co] + sum_ci in[q][ci] * W[o][ci][co], BF16 weights.
__global__ void k_upconv(U3 dc, int CIN, int COUT, const u16* __restrict__ in,
                         const float* __restrict__ W, const float* __restrict__ b,
                         u16* __restrict__ out, int os, int oo) {
  const long fn = 8 * dc.n;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= fn * COUT) return;
  const long p = t / COUT;
  const int co = (int)(t - p * COUT);
  const int fx = 2 * dc.nx, fy = 2 * dc.ny;
  int x, y, z;
  xyz(p, fx, fy, x, y, z);
  const long q = ((long)(z >> 1) * dc.ny + (y >> 1)) * dc.nx + (x >> 1);
  const int o = ((z & 1) << 2) | ((y & 1) << 1) | (x & 1);
  float s = 0.f;
  for (int ci = 0; ci < CIN; ci++)
    s += bf2f(in[q * CIN + ci]) * bfround(W[((long)o * CIN + ci) * COUT + co]);
  out[p * os + oo + co] = f2bf(s + b[co]);
}

// Backward-data of k_upconv: gin[q][ci] = sum_o sum_co W[o][ci][co] gout[2q+o][co].
__global__ void k_upconv_bwd(U3 dc, int CIN, int COUT, const float* __restrict__ gout, int gs,
                             int go, const float* __restrict__ W, float* __restrict__ gin) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= dc.n * CIN) return;
  const long q = t / CIN;
  const int ci = (int)(t - q * CIN);
  int x, y, z;
  xyz(q, dc.nx, dc.ny, x, y, z);
  const int fx = 2 * dc.nx, fy = 2 * dc.ny;
  float s = 0.f;
  for (int o = 0; o < 8; o++) {
    const long p = ((long)(2 * z + (o >> 2)) * fy + (2 * y + ((o >> 1) & 1))) * fx + 2 * x + (o & 1);
    for (int co = 0; co < COUT; co++)
      s += bfround(W[((long)o * CIN + ci) * COUT + co]) * gout[p * gs + go + co];
  }
  gin[t] = s;
}

// Partial weight gradients of k_upconv over chunks of coarse pixels:
// part[chunk][o][ci][co] = sum_q in[q][ci] * gout[2q+o][co].
__global__ void k_upconv_dw(U3 dc, int CIN, int COUT, const u16* __restrict__ in,
                            const float* __restrict__ gout, int gs, int go, long per,
                            float* __restrict__ part) {
  const long m = 8L * CIN * COUT;
  const long j = (long)blockIdx.y * blockDim.x + threadIdx.x;
  if (j >= m) return;
  const int co = (int)(j % COUT), ci = (int)((j / COUT) % CIN), o = (int)(j / ((long)CIN * COUT));
  const long q0 = (long)blockIdx.x * per, q1 = min(dc.n, q0 + per);
  const int fx = 2 * dc.nx, fy = 2 * dc.ny;
  float s = 0.f;
  for (long q = q0; q < q1; q++) {
    int x, y, z;
  xyz(q, dc.nx, dc.ny, x, y, z);
    const long p =
        ((long)(2 * z + (o >> 2)) * fy + (2 * y + ((o >> 1) & 1))) * fx + 2 * x + (o & 1);
    s += bf2f(in[q * CIN + ci]) * gout[p * gs + go + co];
  }
  part[(long)blockIdx.x * m + j] = s;
}

// 1x1 output convolution 16 -> 1 (FP32 output) and its input gradient.
__global__ void k_out_fwd(long n, const u16* __restrict__ a, const float* __restrict__ w,
                          const float* __restrict__ b, float* __restrict__ y) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  float s = 0.f;
#pragma unroll
  for (int c = 0; c < 16; c++) s += bf2f(a[p * 16 + c]) * bfround(w[c]);
  y[p] = s + b[0];
}
__global__ void k_out_bwd(long n, const float* __restrict__ gy, const float* __restrict__ w,
                          float* __restrict__ ga) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * 16) return;
  ga[t] = bfround(w[t & 15]) * gy[t >> 4];
}

// Mean squared error: gy = 2 (y - t) / n, part[block] = sum (y - t)^2.
__global__ void __launch_bounds__(256) k_mse(long n, const float* __restrict__ y,
                                             const float* __restrict__ tg, float* __restrict__ gy,
                                             double* __restrict__ part) {
  __shared__ double red[256];
  const long p = (long)blockIdx.x * 256 + threadIdx.x;
  double s = 0.0;
  if (p < n) {
    const float e = y[p] - tg[p];
    gy[p] = 2.f * e / (float)n;
    s = (double)e * e;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void k_scale_f(long n, float a, float* __restrict__ x) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) x[t] *= a;
}

__global__ void __launch_bounds__(256) k_sumsq_bf16(long n, const u16* __restrict__ a,
                                                    double* __restrict__ part) {
  __shared__ double red[256];
  const long per = (n + gridDim.x - 1) / gridDim.x;
  const long p0 = (long)blockIdx.x * per, p1 = min(n, p0 + per);
  double s = 0.0;
  for (long p = p0 + threadIdx.x; p < p1; p += 256) {
    const double v = bf2f(a[p]);
    s += v * v;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void k_sum_d(int n, const double* __restrict__ part, double* __restrict__ out) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    double s = 0.0;
    for (int k = 0; k < n; k++) s += part[k];
    out[0] = s;
  }
}

// Adam on the FP32 master weights; lr_t carries the bias corrections.
__global__ void k_adam(long n, float* __restrict__ w, const float* __restrict__ gr,
                       float* __restrict__ m1, float* __restrict__ m2, float lr_t, float b1,
                       float b2, float eps) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const float gg = gr[t];
  const float a = b1 * m1[t] + (1.f - b1) * gg;
  const float v = b2 * m2[t] + (1.f - b2) * gg * gg;
  m1[t] = a;
  m2[t] = v;
  w[t] -= lr_t * a / (sqrtf(v) + eps);
}

// Network input from the solver's ghosted FP64 arrays: channel 0 = RHS / sr,
// 1 = (1/rho at the cell, mean of the six face values' 1/Rho_face) / si,
// 2 = P / sp (sp = 0: channel 2 is zero), 3..15 = 0.
__global__ void k_unet_input(Dev d, const double* __restrict__ rhs, const double* __restrict__ rfx,
                             const double* __restrict__ rfy, const double* __restrict__ rfz,
                             const double* __restrict__ P, double isr, double isi, double isp,
                             u16* __restrict__ x0) {
  const int nx = d.Nx - 2, ny = d.Ny - 2, nz = d.Nz - 2;
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= (long)nx * ny * nz) return;
  int i, j, k;
  xyz(p, nx, ny, i, j, k);
  i++;
  j++;
  k++;
  const long c = at(d, i, j, k);
  const double ir = (1. / rfx[c] + 1. / rfx[c + 1] + 1. / rfy[c] + 1. / rfy[c + d.Nx] +
                     1. / rfz[c] + 1. / rfz[c + d.NxNy]) / 6.;
  uint4 v0 = make_uint4(0u, 0u, 0u, 0u), v1 = make_uint4(0u, 0u, 0u, 0u);
  v0.x = f2bf((float)(rhs[c] * isr)) | ((uint32_t)f2bf((float)(ir * isi)) << 16);
  v0.y = P && isp != 0.0 ? f2bf((float)(P[c] * isp)) : 0u;
  uint4* o = reinterpret_cast<uint4*>(x0 + p * 16);
  o[0] = v0;
  o[1] = v1;
}

// P[c] += s * y[p] on the interior (ghosts are refreshed by the caller).
__global__ void k_unet_apply(Dev d, const float* __restrict__ y, double s, double* __restrict__ P,
                             int accumulate) {
  const int nx = d.Nx - 2, ny = d.Ny - 2, nz = d.Nz - 2;
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= (long)nx * ny * nz) return;
  int i, j, k;
  xyz(p, nx, ny, i, j, k);
  i++;
  j++;
  k++;
  const long c = at(d, i, j, k);
  const double v = s * (double)y[p];
  P[c] = accumulate ? P[c] + v : v;
}

// Training sample crop: 3 input channels and the target of a stored pair.
__global__ void k_crop(U3 full, const u16* __restrict__ xs, const u16* __restrict__ ts, int ox,
                       int oy, int oz, U3 cr, int drop_p, float itau, u16* __restrict__ x0,
                       float* __restrict__ tg) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= cr.n) return;
  int x, y, z;
  xyz(p, cr.nx, cr.ny, x, y, z);
  const long q = ((long)(z + oz) * full.ny + (y + oy)) * full.nx + (x + ox);
  uint4 v0 = make_uint4(0u, 0u, 0u, 0u), v1 = make_uint4(0u, 0u, 0u, 0u);
  v0.x = xs[q * 3] | ((uint32_t)xs[q * 3 + 1] << 16);
  v0.y = drop_p ? 0u : xs[q * 3 + 2];
  uint4* o = reinterpret_cast<uint4*>(x0 + p * 16);
  o[0] = v0;
  o[1] = v1;
  tg[p] = bf2f(ts[q]) * itau;
}

// A stored solver pair: the 3 input channels (copied from x0) ...
__global__ void k_store_input(long n, const u16* __restrict__ x0, u16* __restrict__ xs) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  xs[p * 3] = x0[p * 16];
  xs[p * 3 + 1] = x0[p * 16 + 1];
  xs[p * 3 + 2] = x0[p * 16 + 2];
}
// ... and the target (Pnew - Pprev) * is on the interior, as BF16.
__global__ void k_store_target(Dev d, const double* __restrict__ Pprev,
                               const double* __restrict__ Pnew, double is, u16* __restrict__ ts) {
  const int nx = d.Nx - 2, ny = d.Ny - 2, nz = d.Nz - 2;
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= (long)nx * ny * nz) return;
  int i, j, k;
  xyz(p, nx, ny, i, j, k);
  i++;
  j++;
  k++;
  const long c = at(d, i, j, k);
  ts[p] = f2bf((float)((Pnew[c] - Pprev[c]) * is));
}

__global__ void k_bf16_to_f32(long n, const u16* __restrict__ a, int as, int ao, int C,
                              float* __restrict__ out) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * C) return;
  const long p = t / C;
  out[t] = bf2f(a[p * as + ao + (int)(t - p * C)]);
}
__global__ void k_f32_to_bf16(long n, const float* __restrict__ a, int C, u16* __restrict__ out,
                              int os, int oo) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * C) return;
  const long p = t / C;
  out[p * os + oo + (int)(t - p * C)] = f2bf(a[t]);
}

}  // namespace fr

// ---- host side ------------------------------------------------------------------
namespace fr {

// Layers in parameter order: 3x3x3 convolutions c1a c1b c2a c2b c3a c3b
// (encoder), transposed 2x2x2 convolutions up2 (64->32) and up1 (32->16),
// decoder convolutions d2a d2b d1a d1b, and the 1x1 output convolution.
enum { C1A, C1B, C2A, C2B, C3A, C3B, D2A, D2B, D1A, D1B, NCONV };
static const int kConvIO[NCONV][2] = {{16, 16}, {16, 16}, {16, 32}, {32, 32}, {32, 64},
                                      {64, 64}, {64, 32}, {32, 32}, {32, 16}, {16, 16}};

struct UnetWS {
  U3 l1{}, l2{}, l3{};
  bool train = false;
  u16 *x0 = nullptr, *e1a = nullptr, *cat1 = nullptr, *p1 = nullptr, *e2a = nullptr,
      *cat2 = nullptr, *p2 = nullptr, *b1 = nullptr, *b2 = nullptr, *d2a = nullptr,
      *d2 = nullptr, *d1a = nullptr, *d1 = nullptr, *dp = nullptr;
  float *y = nullptr, *tg = nullptr, *gy = nullptr;
  float *g_d1 = nullptr, *g_d1a = nullptr, *g_cat1 = nullptr, *g_d2 = nullptr, *g_d2a = nullptr,
        *g_cat2 = nullptr, *g_b2 = nullptr, *g_b1 = nullptr, *g_p2 = nullptr, *g_e2a = nullptr,
        *g_p1 = nullptr, *g_e1a = nullptr;
  double *dpart = nullptr, *dsum = nullptr;
  std::vector<void*> all;
};

}  // namespace fr

struct fr_unet {
  long nparam = 0;
  float *w = nullptr, *g = nullptr, *m1 = nullptr, *m2 = nullptr;
  long step = 0;
  long cw[fr::NCONV], cb[fr::NCONV];  // offsets of conv weights / biases
  long uw[2], ub[2];                  // up2, up1
  long ow = 0, ob = 0;                // output 1x1
  fr::u16* pf[fr::NCONV] = {};        // packed forward operands
  fr::u16* pb[fr::NCONV] = {};        // packed backward-data operands
  bool packed = false;
  float* part = nullptr;
  long part_cap = 0;
  fr::UnetWS ws[2];  // [0] inference size, [1] training crop size
  double tau = 0.0;  // output scale: RMS of the stored targets (set by training)
};

struct fr_unet_data {
  fr::U3 full{};
  int cap = 0, count = 0;
  fr::u16 *xs = nullptr, *ts = nullptr;  // [cap][n][3], [cap][n]
  fr::u16* x0 = nullptr;                 // staging, [n][16]
  double* red = nullptr;
  std::vector<double> msq;               // mean square of each stored target
};

namespace fr {

static unsigned nblk(long n, int b = 256) { return (unsigned)((n + b - 1) / b); }

template <typename T>
static int dalloc(UnetWS& s, T** p, long count) {
  if (count <= 0) {
    *p = nullptr;
    return FR_OK;
  }
  FR_CHECK(hipMalloc((void**)p, sizeof(T) * (size_t)count));
  s.all.push_back((void*)*p);
  return FR_OK;
}

static void ws_free(UnetWS& s) {
  for (void* p : s.all) (void)hipFree(p);
  s = UnetWS();
}

static int ws_init(UnetWS& s, int nx, int ny, int nz, bool train) {
  if (nx % 4 || ny % 4 || nz % 4 || nx < 4 || ny < 4 || nz < 4) return FR_EINVAL;
  if (s.l1.nx == nx && s.l1.ny == ny && s.l1.nz == nz && (s.train || !train)) return FR_OK;
  ws_free(s);
  s.l1 = make_u3(nx, ny, nz);
  s.l2 = make_u3(nx / 2, ny / 2, nz / 2);
  s.l3 = make_u3(nx / 4, ny / 4, nz / 4);
  s.train = train;
  const long n1 = s.l1.n, n2 = s.l2.n, n3 = s.l3.n;
  FR_RET(dalloc(s, &s.x0, n1 * 16));
  FR_RET(dalloc(s, &s.e1a, n1 * 16));
  FR_RET(dalloc(s, &s.cat1, n1 * 32));
  FR_RET(dalloc(s, &s.p1, n2 * 16));
  FR_RET(dalloc(s, &s.e2a, n2 * 32));
  FR_RET(dalloc(s, &s.cat2, n2 * 64));
  FR_RET(dalloc(s, &s.p2, n3 * 32));
  FR_RET(dalloc(s, &s.b1, n3 * 64));
  FR_RET(dalloc(s, &s.b2, n3 * 64));
  FR_RET(dalloc(s, &s.d2a, n2 * 32));
  FR_RET(dalloc(s, &s.d2, n2 * 32));
  FR_RET(dalloc(s, &s.d1a, n1 * 16));
  FR_RET(dalloc(s, &s.d1, n1 * 16));
  FR_RET(dalloc(s, &s.y, n1));
  FR_RET(dalloc(s, &s.dpart, (long)nblk(n1) + 1));
  FR_RET(dalloc(s, &s.dsum, 1));
  FR_CHECK(hipMemset(s.x0, 0, sizeof(u16) * n1 * 16));
  if (train) {
    FR_RET(dalloc(s, &s.dp, n1 * 16));
    FR_RET(dalloc(s, &s.tg, n1));
    FR_RET(dalloc(s, &s.gy, n1));
    FR_RET(dalloc(s, &s.g_d1, n1 * 16));
    FR_RET(dalloc(s, &s.g_d1a, n1 * 16));
    FR_RET(dalloc(s, &s.g_cat1, n1 * 32));
    FR_RET(dalloc(s, &s.g_d2, n2 * 32));
    FR_RET(dalloc(s, &s.g_d2a, n2 * 32));
    FR_RET(dalloc(s, &s.g_cat2, n2 * 64));
    FR_RET(dalloc(s, &s.g_b2, n3 * 64));
    FR_RET(dalloc(s, &s.g_b1, n3 * 64));
    FR_RET(dalloc(s, &s.g_p2, n3 * 32));
    FR_RET(dalloc(s, &s.g_e2a, n2 * 32));
    FR_RET(dalloc(s, &s.g_p1, n2 * 16));
    FR_RET(dalloc(s, &s.g_e1a, n1 * 16));
  }
  return FR_OK;
}

// ---- launches, dispatched on the channel pair ------------------------------------
template <int CI, int CO>
static void conv_launch(int mode, U3 d, const u16* in, int is, int io, const u16* wp,
                        const float* b, void* out, int os, int oo, hipStream_t st) {
  const dim3 grid(nblk(d.nx, 64), nblk(d.ny, 4), d.nz);
  if (mode == 0)
    hipLaunchKernelGGL((k_conv3<CI, CO, 0, true>), grid, dim3(256), 0, st, d, in, is, io,
                       (const uint4*)wp, b, out, os, oo);
  else
    hipLaunchKernelGGL((k_conv3<CI, CO, 1, false>), grid, dim3(256), 0, st, d, in, is, io,
                       (const uint4*)wp, b, out, os, oo);
}

template <int CI, int CO>
static int dw_launch(U3 d, const u16* x, int xs, int xo, const u16* dp, float* part, long cap,
                     float* g, hipStream_t st) {
  constexpr int BP = 27 * (CI / 16) * (CO / 16);
  const int rows = d.ny * d.nz;
  const int chunks = rows < 256 ? rows : 256;
  const int per = (rows + chunks - 1) / chunks;
  const long m = 27L * CI * CO;
  if ((long)chunks * m > cap) return FR_EINVAL;
  hipLaunchKernelGGL((k_conv3_dw<CI, CO, 8>), dim3(chunks, nblk(BP, 32)), dim3(256), 0, st, d, x,
                     xs, xo, dp, per, part);
  hipLaunchKernelGGL(k_sum_parts, dim3(nblk(m)), dim3(256), 0, st, chunks, m, part, g, 1);
  return FR_OK;
}

#define FR_UNET_PAIRS(X) X(16, 16) X(16, 32) X(32, 32) X(32, 64) X(64, 64) X(64, 32) X(32, 16)

static int conv(int ci, int co, int mode, U3 d, const u16* in, int is, int io, const u16* wp,
                const float* b, void* out, int os, int oo, hipStream_t st) {
#define FR_X(A, B)                                                      \
  if (ci == A && co == B) {                                             \
    conv_launch<A, B>(mode, d, in, is, io, wp, b, out, os, oo, st);     \
    return FR_OK;                                                       \
  }
  FR_UNET_PAIRS(FR_X)
#undef FR_X
  return FR_EINVAL;
}

static int conv_dw(int ci, int co, U3 d, const u16* x, int xs, int xo, const u16* dp, float* part,
                   long cap, float* g, hipStream_t st) {
#define FR_X(A, B) \
  if (ci == A && co == B) return dw_launch<A, B>(d, x, xs, xo, dp, part, cap, g, st);
  FR_UNET_PAIRS(FR_X)
#undef FR_X
  return FR_EINVAL;
}

static long conv_packed_size(int ci, int co) { return (long)nchunks(ci) * (co / 16) * 64 * 8; }

static int unet_pack(fr_unet* u, hipStream_t st) {
  for (int l = 0; l < NCONV; l++) {
    const int ci = kConvIO[l][0], co = kConvIO[l][1];
    const long nf = conv_packed_size(ci, co), nb = conv_packed_size(co, ci);
    hipLaunchKernelGGL(k_pack_conv, dim3(nblk(nf)), dim3(256), 0, st, u->w + u->cw[l], ci, co, ci,
                       co, 0, u->pf[l]);
    hipLaunchKernelGGL(k_pack_conv, dim3(nblk(nb)), dim3(256), 0, st, u->w + u->cw[l], ci, co, co,
                       ci, 1, u->pb[l]);
  }
  FR_LAUNCH_CHECK();
  u->packed = true;
  return FR_OK;
}

// Bias gradient of a layer: g[boff + c] += sum_p gr[p] * (y[p] > 0).
static int bias_grad(fr_unet* u, long n, int C, const float* gr, int gs, int go, const u16* y,
                     int ys, int yo, float* gout, hipStream_t st) {
  long nb = (n + 4095) / 4096;
  if (nb > 1024) nb = 1024;
  const long per = (n + nb - 1) / nb;
  if (nb * C > u->part_cap) return FR_EINVAL;
  hipLaunchKernelGGL(k_chan_sum, dim3((unsigned)nb), dim3(256), 0, st, n, C, per, gr, gs, go, y, ys,
                     yo, u->part);
  hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(256), 0, st, (int)nb, (long)C, u->part, gout, 1);
  return FR_OK;
}

static int upconv_grads(fr_unet* u, int k, U3 dc, int ci, int co, const u16* in, const float* gout,
                        int gs, int go, hipStream_t st) {
  const long m = 8L * ci * co;
  long chunks = dc.n < 256 ? dc.n : 256;
  const long per = (dc.n + chunks - 1) / chunks;
  if (chunks * m > u->part_cap) return FR_EINVAL;
  hipLaunchKernelGGL(k_upconv_dw, dim3((unsigned)chunks, nblk(m)), dim3(256), 0, st, dc, ci, co, in,
                     gout, gs, go, per, u->part);
  hipLaunchKernelGGL(k_sum_parts, dim3(nblk(m)), dim3(256), 0, st, (int)chunks, m, u->part,
                     u->g + u->uw[k], 1);
  const U3 df = make_u3(2 * dc.nx, 2 * dc.ny, 2 * dc.nz);
  return bias_grad(u, df.n, co, gout, gs, go, nullptr, 0, 0, u->g + u->ub[k], st);
}

int unet_forward(fr_unet* u, UnetWS& s, hipStream_t st) {
  if (!u->packed) FR_RET(unet_pack(u, st));
  const U3 L1 = s.l1, L2 = s.l2, L3 = s.l3;
  auto C = [&](int l, U3 d, const u16* in, int is, int io, u16* out, int os, int oo) {
    return conv(kConvIO[l][0], kConvIO[l][1], 0, d, in, is, io, u->pf[l], u->w + u->cb[l], out, os,
                oo, st);
  };
  FR_RET(C(C1A, L1, s.x0, 16, 0, s.e1a, 16, 0));
  FR_RET(C(C1B, L1, s.e1a, 16, 0, s.cat1, 32, 16));
  hipLaunchKernelGGL(k_pool, dim3(nblk(L2.n * 16)), dim3(256), 0, st, L2, 16, s.cat1, 32, 16, s.p1);
  FR_RET(C(C2A, L2, s.p1, 16, 0, s.e2a, 32, 0));
  FR_RET(C(C2B, L2, s.e2a, 32, 0, s.cat2, 64, 32));
  hipLaunchKernelGGL(k_pool, dim3(nblk(L3.n * 32)), dim3(256), 0, st, L3, 32, s.cat2, 64, 32, s.p2);
  FR_RET(C(C3A, L3, s.p2, 32, 0, s.b1, 64, 0));
  FR_RET(C(C3B, L3, s.b1, 64, 0, s.b2, 64, 0));
  hipLaunchKernelGGL(k_upconv, dim3(nblk(L2.n * 32)), dim3(256), 0, st, L3, 64, 32, s.b2,
                     u->w + u->uw[0], u->w + u->ub[0], s.cat2, 64, 0);
  FR_RET(C(D2A, L2, s.cat2, 64, 0, s.d2a, 32, 0));
  FR_RET(C(D2B, L2, s.d2a, 32, 0, s.d2, 32, 0));
  hipLaunchKernelGGL(k_upconv, dim3(nblk(L1.n * 16)), dim3(256), 0, st, L2, 32, 16, s.d2,
                     u->w + u->uw[1], u->w + u->ub[1], s.cat1, 32, 0);
  FR_RET(C(D1A, L1, s.cat1, 32, 0, s.d1a, 16, 0));
  FR_RET(C(D1B, L1, s.d1a, 16, 0, s.d1, 16, 0));
  hipLaunchKernelGGL(k_out_fwd, dim3(nblk(L1.n)), dim3(256), 0, st, L1.n, s.d1, u->w + u->ow,
                     u->w + u->ob, s.y);
  FR_LAUNCH_CHECK();
  return FR_OK;
}

// Backward pass after unet_forward on the same workspace; s.gy holds dL/dy.
// Parameter gradients are added to u->g.
int unet_backward(fr_unet* u, UnetWS& s, hipStream_t st) {
  if (!s.train) return FR_EINVAL;
  const U3 L1 = s.l1, L2 = s.l2, L3 = s.l3;
  float* G = u->g;
  // one conv layer: output (gr, y) -> dp -> weight/bias grads, input grad
  auto layer = [&](int l, U3 d, const float* gr, int gs, int go, const u16* y, int ys, int yo,
                   const u16* x, int xs, int xo, float* gin) -> int {
    const int ci = kConvIO[l][0], co = kConvIO[l][1];
    hipLaunchKernelGGL(k_relu_bwd, dim3(nblk(d.n * co)), dim3(256), 0, st, d.n, co, gr, gs, go, y,
                       ys, yo, s.dp);
    FR_RET(conv_dw(ci, co, d, x, xs, xo, s.dp, u->part, u->part_cap, G + u->cw[l], st));
    FR_RET(bias_grad(u, d.n, co, gr, gs, go, y, ys, yo, G + u->cb[l], st));
    if (gin) FR_RET(conv(co, ci, 1, d, s.dp, co, 0, u->pb[l], nullptr, gin, ci, 0, st));
    return FR_OK;
  };
  // output 1x1
  hipLaunchKernelGGL(k_out_bwd, dim3(nblk(L1.n * 16)), dim3(256), 0, st, L1.n, s.gy, u->w + u->ow,
                     s.g_d1);
  {
    long nb = (L1.n + 4095) / 4096;
    if (nb > 1024) nb = 1024;
    const long per = (L1.n + nb - 1) / nb;
    hipLaunchKernelGGL(k_chan_dot, dim3((unsigned)nb), dim3(256), 0, st, L1.n, 16, per, s.d1, s.gy,
                       u->part);
    hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(256), 0, st, (int)nb, 16L, u->part, G + u->ow, 1);
    FR_RET(bias_grad(u, L1.n, 1, s.gy, 1, 0, nullptr, 0, 0, G + u->ob, st));
  }
  FR_RET(layer(D1B, L1, s.g_d1, 16, 0, s.d1, 16, 0, s.d1a, 16, 0, s.g_d1a));
  FR_RET(layer(D1A, L1, s.g_d1a, 16, 0, s.d1a, 16, 0, s.cat1, 32, 0, s.g_cat1));
  // up1: cat1[:, 0:16] = upconv(d2)
  hipLaunchKernelGGL(k_upconv_bwd, dim3(nblk(L2.n * 32)), dim3(256), 0, st, L2, 32, 16, s.g_cat1,
                     32, 0, u->w + u->uw[1], s.g_d2);
  FR_RET(upconv_grads(u, 1, L2, 32, 16, s.d2, s.g_cat1, 32, 0, st));
  FR_RET(layer(D2B, L2, s.g_d2, 32, 0, s.d2, 32, 0, s.d2a, 32, 0, s.g_d2a));
  FR_RET(layer(D2A, L2, s.g_d2a, 32, 0, s.d2a, 32, 0, s.cat2, 64, 0, s.g_cat2));
  // up2: cat2[:, 0:32] = upconv(b2)
  hipLaunchKernelGGL(k_upconv_bwd, dim3(nblk(L3.n * 64)), dim3(256), 0, st, L3, 64, 32, s.g_cat2,
                     64, 0, u->w + u->uw[0], s.g_b2);
  FR_RET(upconv_grads(u, 0, L3, 64, 32, s.b2, s.g_cat2, 64, 0, st));
  FR_RET(layer(C3B, L3, s.g_b2, 64, 0, s.b2, 64, 0, s.b1, 64, 0, s.g_b1));
  FR_RET(layer(C3A, L3, s.g_b1, 64, 0, s.b1, 64, 0, s.p2, 32, 0, s.g_p2));
  // p2 = pool(e2), e2 = cat2[:, 32:64]
  hipLaunchKernelGGL(k_pool_bwd, dim3(nblk(L2.n * 32)), dim3(256), 0, st, L3, 32, s.g_p2,
                     s.g_cat2, 64, 32);
  FR_RET(layer(C2B, L2, s.g_cat2, 64, 32, s.cat2, 64, 32, s.e2a, 32, 0, s.g_e2a));
  FR_RET(layer(C2A, L2, s.g_e2a, 32, 0, s.e2a, 32, 0, s.p1, 16, 0, s.g_p1));
  // p1 = pool(e1), e1 = cat1[:, 16:32]
  hipLaunchKernelGGL(k_pool_bwd, dim3(nblk(L1.n * 16)), dim3(256), 0, st, L2, 16, s.g_p1,
                     s.g_cat1, 32, 16);
  FR_RET(layer(C1B, L1, s.g_cat1, 32, 16, s.cat1, 32, 16, s.e1a, 16, 0, s.g_e1a));
  FR_RET(layer(C1A, L1, s.g_e1a, 16, 0, s.e1a, 16, 0, s.x0, 16, 0, nullptr));
  FR_LAUNCH_CHECK();
  return FR_OK;
}

int unet_adam(fr_unet* u, float lr, hipStream_t st) {
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  u->step++;
  const double c1 = 1.0 - std::pow((double)b1, (double)u->step);
  const double c2 = 1.0 - std::pow((double)b2, (double)u->step);
  const float lr_t = (float)(lr * std::sqrt(c2) / c1);
  hipLaunchKernelGGL(k_adam, dim3(nblk(u->nparam)), dim3(256), 0, st, u->nparam, u->w, u->g,
                     u->m1, u->m2, lr_t, b1, b2, eps);
  FR_LAUNCH_CHECK();
  u->packed = false;
  return FR_OK;
}

// Deterministic He-normal initialisation (splitmix64 + Box-Muller); the
// padded input channels 3..15 of c1a and all biases start at zero, the
// output layer small so that an untrained network adds almost nothing.
static void unet_init_host(const fr_unet* u, unsigned seed, std::vector<float>& h) {
  h.assign(u->nparam, 0.f);
  uint64_t st = 0x9E3779B97F4A7C15ull ^ ((uint64_t)seed * 0xBF58476D1CE4E5B9ull);
  auto next = [&]() {
    uint64_t z = (st += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  };
  auto normal = [&]() {
    const double u1 = ((next() >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    const double u2 = ((next() >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  };
  for (int l = 0; l < NCONV; l++) {
    const int ci = kConvIO[l][0], co = kConvIO[l][1];
    const int real_ci = l == C1A ? 3 : ci;
    const double sd = std::sqrt(2.0 / (27.0 * real_ci));
    for (int t = 0; t < 27; t++)
      for (int a = 0; a < ci; a++)
        for (int b = 0; b < co; b++)
          h[u->cw[l] + ((long)t * ci + a) * co + b] = a < real_ci ? (float)(sd * normal()) : 0.f;
  }
  const int uio[2][2] = {{64, 32}, {32, 16}};
  for (int k = 0; k < 2; k++) {
    const double sd = std::sqrt(1.0 / uio[k][0]);
    for (long j = 0; j < 8L * uio[k][0] * uio[k][1]; j++) h[u->uw[k] + j] = (float)(sd * normal());
  }
  for (int c = 0; c < 16; c++) h[u->ow + c] = (float)(0.01 * normal());
}

// Network input of a pressure correction: r0 = sign * r (the right-hand side
// of S dP = r0 with S = div(C grad)), 1/rho from Rho_face, Pprev (or none);
// the scales make each channel O(1) and give the output's physical scale.
static int unet_input(UnetWS& s, const Dev& d, const double* r, double sign, double rrms,
                      const double* rfx, const double* rfy, const double* rfz, const double* Pprev,
                      double* out_scale, u16* x0, hipStream_t st) {
  double rfmin = 0, pmax = 0;
  FR_RET(reduce_sync(d, rfx, 3, &rfmin, st));
  if (Pprev) FR_RET(reduce_sync(d, Pprev, 4, &pmax, st));
  const double sr = rrms > 0 ? rrms : 1.0;
  const double si = rfmin > 0 ? 1.0 / rfmin : 1.0;
  const double lap = d.idx2 + d.idy2 + d.idz2;
  *out_scale = sr / (si * lap);
  hipLaunchKernelGGL(k_unet_input, dim3(nblk(s.l1.n)), dim3(256), 0, st, d, r, rfx, rfy, rfz,
                     Pprev, sign / sr, 1.0 / si, pmax > 0 ? 1.0 / pmax : 0.0, x0);
  FR_LAUNCH_CHECK();
  return FR_OK;
}

int unet_infer(fr_unet* u, const Dev& d, const double* r, double sign, double rrms,
               const double* rfx, const double* rfy, const double* rfz, const double* Pprev,
               const float** y, double* scale, hipStream_t st) {
  UnetWS& s = u->ws[0];
  FR_RET(ws_init(s, d.Nx - 2, d.Ny - 2, d.Nz - 2, false));
  FR_RET(unet_input(s, d, r, sign, rrms, rfx, rfy, rfz, Pprev, scale, s.x0, st));
  FR_RET(unet_forward(u, s, st));
  *y = s.y;
  if (u->tau > 0) *scale *= u->tau;
  return FR_OK;
}

bool unet_data_room(const fr_unet_data* D) { return D && D->count < D->cap; }

// Pair collection around one solve: the input at entry, the target
// (Pfinal - Pstart) / scale at exit.
int unet_data_begin(fr_unet_data* D, const Dev& d, const double* r, double sign, double rrms,
                    const double* rfx, const double* rfy, const double* rfz, const double* Pprev,
                    double* scale, hipStream_t st) {
  if (D->count >= D->cap) return FR_EOVERFLOW;
  if (D->full.nx != d.Nx - 2 || D->full.ny != d.Ny - 2 || D->full.nz != d.Nz - 2) return FR_EINVAL;
  UnetWS tmp;
  tmp.l1 = D->full;
  FR_RET(unet_input(tmp, d, r, sign, rrms, rfx, rfy, rfz, Pprev, scale, D->x0, st));
  const long n = D->full.n;
  hipLaunchKernelGGL(k_store_input, dim3(nblk(n)), dim3(256), 0, st, n, D->x0,
                     D->xs + (long)D->count * 3 * n);
  FR_LAUNCH_CHECK();
  return FR_OK;
}

int unet_data_end(fr_unet_data* D, const Dev& d, const double* Pstart, const double* Pfinal,
                  double scale, hipStream_t st) {
  const long n = D->full.n;
  u16* t = D->ts + (long)D->count * n;
  hipLaunchKernelGGL(k_store_target, dim3(nblk(n)), dim3(256), 0, st, d, Pstart, Pfinal,
                     1.0 / scale, t);
  hipLaunchKernelGGL(k_sumsq_bf16, dim3(256), dim3(256), 0, st, n, t, D->red);
  hipLaunchKernelGGL(k_sum_d, dim3(1), dim3(1), 0, st, 256, D->red, D->red + 256);
  FR_LAUNCH_CHECK();
  double h = 0;
  FR_CHECK(hipMemcpyAsync(&h, D->red + 256, sizeof(double), hipMemcpyDeviceToHost, st));
  FR_CHECK(hipStreamSynchronize(st));
  D->msq.push_back(h / (double)n);
  D->count++;
  return FR_OK;
}

}  // namespace fr

using namespace fr;

extern "C" {

int fr_unet_create(int seed, fr_unet** out) {
  if (!out) return FR_EINVAL;
  fr_unet* u = new fr_unet();
  long off = 0;
  for (int l = 0; l < NCONV; l++) {
    u->cw[l] = off;
    off += 27L * kConvIO[l][0] * kConvIO[l][1];
    u->cb[l] = off;
    off += kConvIO[l][1];
  }
  const int uio[2][2] = {{64, 32}, {32, 16}};
  for (int k = 0; k < 2; k++) {
    u->uw[k] = off;
    off += 8L * uio[k][0] * uio[k][1];
    u->ub[k] = off;
    off += uio[k][1];
  }
  u->ow = off;
  off += 16;
  u->ob = off;
  off += 1;
  u->nparam = off;
  int rc = FR_OK;
  auto A = [&](void** p, size_t bytes) {
    if (!rc && hipMalloc(p, bytes) != hipSuccess) rc = FR_ENOMEM;
  };
  A((void**)&u->w, sizeof(float) * off);
  A((void**)&u->g, sizeof(float) * off);
  A((void**)&u->m1, sizeof(float) * off);
  A((void**)&u->m2, sizeof(float) * off);
  for (int l = 0; l < NCONV && !rc; l++) {
    A((void**)&u->pf[l], sizeof(u16) * conv_packed_size(kConvIO[l][0], kConvIO[l][1]));
    A((void**)&u->pb[l], sizeof(u16) * conv_packed_size(kConvIO[l][1], kConvIO[l][0]));
  }
  // largest partial-sum buffer: 256 row chunks of the 64x64 weight gradient
  u->part_cap = 256L * 27 * 64 * 64;
  A((void**)&u->part, sizeof(float) * u->part_cap);
  if (rc) {
    fr_unet_destroy(u);
    return rc;
  }
  std::vector<float> h;
  unet_init_host(u, (unsigned)seed, h);
  if (hipMemcpy(u->w, h.data(), sizeof(float) * off, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(u->g, 0, sizeof(float) * off) != hipSuccess ||
      hipMemset(u->m1, 0, sizeof(float) * off) != hipSuccess ||
      hipMemset(u->m2, 0, sizeof(float) * off) != hipSuccess) {
    fr_unet_destroy(u);
    return FR_ENOMEM;
  }
  *out = u;
  return FR_OK;
}

int fr_unet_destroy(fr_unet* u) {
  if (!u) return FR_OK;
  for (auto& s : u->ws) ws_free(s);
  for (int l = 0; l < NCONV; l++) {
    (void)hipFree(u->pf[l]);
    (void)hipFree(u->pb[l]);
  }
  for (void* p : {(void*)u->w, (void*)u->g, (void*)u->m1, (void*)u->m2, (void*)u->part})
    (void)hipFree(p);
  delete u;
  return FR_OK;
}

int fr_unet_num_params(fr_unet* u, long long* n) {
  if (!u || !n) return FR_EINVAL;
  *n = u->nparam;
  return FR_OK;
}

int fr_unet_get_params(fr_unet* u, float* host) {
  if (!u || !host) return FR_EINVAL;
  FR_CHECK(hipMemcpy(host, u->w, sizeof(float) * u->nparam, hipMemcpyDeviceToHost));
  return FR_OK;
}

int fr_unet_set_params(fr_unet* u, const float* host) {
  if (!u || !host) return FR_EINVAL;
  FR_CHECK(hipMemcpy(u->w, host, sizeof(float) * u->nparam, hipMemcpyHostToDevice));
  u->packed = false;
  return FR_OK;
}

int fr_unet_get_grads(fr_unet* u, float* host) {
  if (!u || !host) return FR_EINVAL;
  FR_CHECK(hipMemcpy(host, u->g, sizeof(float) * u->nparam, hipMemcpyDeviceToHost));
  return FR_OK;
}

int fr_unet_zero_grads(fr_unet* u, void* stream) {
  if (!u) return FR_EINVAL;
  FR_CHECK(hipMemsetAsync(u->g, 0, sizeof(float) * u->nparam, (hipStream_t)stream));
  return FR_OK;
}

int fr_unet_forward(fr_unet* u, int nx, int ny, int nz, const float* x, float* y, void* stream) {
  if (!u || !x || !y) return FR_EINVAL;
  hipStream_t st = (hipStream_t)stream;
  UnetWS& s = u->ws[1];
  FR_RET(ws_init(s, nx, ny, nz, true));
  hipLaunchKernelGGL(k_f32_to_bf16, dim3(nblk(s.l1.n * 3)), dim3(256), 0, st, s.l1.n, x, 3, s.x0,
                     16, 0);
  FR_RET(unet_forward(u, s, st));
  FR_CHECK(hipMemcpyAsync(y, s.y, sizeof(float) * s.l1.n, hipMemcpyDeviceToDevice, st));
  return FR_OK;
}

int fr_unet_backward(fr_unet* u, int nx, int ny, int nz, const float* gy, void* stream) {
  if (!u || !gy) return FR_EINVAL;
  hipStream_t st = (hipStream_t)stream;
  UnetWS& s = u->ws[1];
  if (!s.train || s.l1.nx != nx || s.l1.ny != ny || s.l1.nz != nz) return FR_EINVAL;
  FR_CHECK(hipMemcpyAsync(s.gy, gy, sizeof(float) * s.l1.n, hipMemcpyDeviceToDevice, st));
  return unet_backward(u, s, st);
}

int fr_unet_adam_step(fr_unet* u, float lr, void* stream) {
  if (!u) return FR_EINVAL;
  return unet_adam(u, lr, (hipStream_t)stream);
}

int fr_unet_data_create(int nx, int ny, int nz, int capacity, fr_unet_data** out) {
  if (!out || capacity <= 0 || nx % 4 || ny % 4 || nz % 4) return FR_EINVAL;
  fr_unet_data* d = new fr_unet_data();
  d->full = make_u3(nx, ny, nz);
  d->cap = capacity;
  const long n = d->full.n;
  if (hipMalloc((void**)&d->xs, sizeof(u16) * 3 * n * capacity) != hipSuccess ||
      hipMalloc((void**)&d->ts, sizeof(u16) * n * capacity) != hipSuccess ||
      hipMalloc((void**)&d->x0, sizeof(u16) * 16 * n) != hipSuccess ||
      hipMalloc((void**)&d->red, sizeof(double) * 260) != hipSuccess) {
    fr_unet_data_destroy(d);
    return FR_ENOMEM;
  }
  *out = d;
  return FR_OK;
}

int fr_unet_data_destroy(fr_unet_data* d) {
  if (!d) return FR_OK;
  for (void* p : {(void*)d->xs, (void*)d->ts, (void*)d->x0, (void*)d->red}) (void)hipFree(p);
  delete d;
  return FR_OK;
}

int fr_unet_data_count(fr_unet_data* d, int* n) {
  if (!d || !n) return FR_EINVAL;
  *n = d->count;
  return FR_OK;
}

int fr_unet_train(fr_unet* u, fr_unet_data* D, int iters, int batch, int cx, int cy, int cz,
                  float lr, unsigned seed, double* loss_first, double* loss_last, void* stream) {
  if (!u || !D || iters < 0 || batch <= 0 || D->count == 0) return FR_EINVAL;
  if (cx > D->full.nx || cy > D->full.ny || cz > D->full.nz) return FR_EINVAL;
  hipStream_t st = (hipStream_t)stream;
  UnetWS& s = u->ws[1];
  FR_RET(ws_init(s, cx, cy, cz, true));
  uint64_t rs = 0x2545F4914F6CDD1Dull ^ ((uint64_t)seed << 17);
  auto rnd = [&](uint64_t m) {  // xorshift64*
    rs ^= rs >> 12;
    rs ^= rs << 25;
    rs ^= rs >> 27;
    return (uint64_t)((rs * 0x2545F4914F6CDD1Dull) >> 11) % m;
  };
  const long n = s.l1.n;
  const unsigned nbm = nblk(n);
  if (u->tau <= 0) {  // output scale from the stored targets, fixed at the first training
    double m = 0;
    for (double v : D->msq) m += v;
    u->tau = D->msq.empty() || m <= 0 ? 1.0 : std::sqrt(m / (double)D->msq.size());
  }
  const float itau = (float)(1.0 / u->tau);
  const int win = iters < 20 ? (iters > 0 ? iters : 1) : 20;
  double lf = 0, ll = 0;
  for (int it = 0; it < iters; it++) {
    FR_CHECK(hipMemsetAsync(u->g, 0, sizeof(float) * u->nparam, st));
    double loss = 0;
    for (int b = 0; b < batch; b++) {
      const int k = (int)rnd((uint64_t)D->count);
      const int ox = (int)rnd((uint64_t)(D->full.nx - cx + 1));
      const int oy = (int)rnd((uint64_t)(D->full.ny - cy + 1));
      const int oz = (int)rnd((uint64_t)(D->full.nz - cz + 1));
      const int drop = (int)rnd(2);
      hipLaunchKernelGGL(k_crop, dim3(nblk(n)), dim3(256), 0, st, D->full,
                         D->xs + (long)k * 3 * D->full.n, D->ts + (long)k * D->full.n, ox, oy, oz,
                         s.l1, drop, itau, s.x0, s.tg);
      FR_RET(unet_forward(u, s, st));
      hipLaunchKernelGGL(k_mse, dim3(nbm), dim3(256), 0, st, n, s.y, s.tg, s.gy, s.dpart);
      hipLaunchKernelGGL(k_sum_d, dim3(1), dim3(1), 0, st, (int)nbm, s.dpart, s.dsum);
      // batch mean: gy carries 2 e / n, scaled by 1 / batch
      if (batch > 1)
        hipLaunchKernelGGL(k_scale_f, dim3(nbm), dim3(256), 0, st, n, 1.f / batch, s.gy);
      FR_RET(unet_backward(u, s, st));
      double h = 0;
      FR_CHECK(hipMemcpyAsync(&h, s.dsum, sizeof(double), hipMemcpyDeviceToHost, st));
      FR_CHECK(hipStreamSynchronize(st));
      loss += h / (double)n;
    }
    loss /= batch;
    if (it < win) lf += loss / win;
    if (it >= iters - win) ll += loss / win;
    // cosine decay of the learning rate over the run
    const float lr_it = (float)(0.5 * lr * (1.0 + std::cos(3.141592653589793 * it / iters)));
    FR_RET(unet_adam(u, lr_it, st));
  }
  FR_CHECK(hipStreamSynchronize(st));
  if (loss_first) *loss_first = lf;
  if (loss_last) *loss_last = ll;
  return FR_OK;
}

}  // extern "C"
