// Band-point convolution kernel of the nonlocal response (one definition:
// included by engine.hip only; the geometry and arguments are in
// nonlocal.hpp).
#pragma once
#include "nonlocal.hpp"

namespace nlse {

// Direct convolution at the points of one tile (blockIdx.x) of one slab
// (blockIdx.y).  The intensities the tile needs are staged in LDS with a
// row pitch of PITCH floats (0: the tile's own width), per chunk of kc
// kernel rows.  The BAND_WAVES waves split the kernel columns; each lane
// slides a register window of RB_BAND + 3 rows down its column and reads
// four kernel rows at once, accumulating FP32 partial sums of at most
// 4 x 8 products that are added in FP64.
template <int PITCH>
__global__ __launch_bounds__(64 * BAND_WAVES) void k_nl_band(BandArgs b) {
  extern __shared__ float smem[];
  const BandTile tl = b.tiles[blockIdx.x];
  const int64_t slab = blockIdx.y;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wv = tid >> 6;
  const int kw = 2 * b.r + 1;
  const int groups = 64 / tl.ncols;
  const int wcols = tl.ncols + 2 * b.r;
  const int pitch = PITCH > 0 ? PITCH : wcols;
  const int srows = groups * RB_BAND + b.kc - 1;
  float* img = smem;
  float* kt = smem + (((size_t)srows * pitch + 3) & ~(size_t)3);
  double* red = reinterpret_cast<double*>(kt + (((size_t)kw * b.kc + 1) & ~(size_t)1));
  const int c = lane % tl.ncols;
  const int g = lane / tl.ncols;
  const bool lane_ok = g < groups;
  const float2* a = b.A + slab * (int64_t)b.ny * b.nx;
  const int vper = (kw + BAND_WAVES - 1) / BAND_WAVES;
  const int v0 = wv * vper;
  const int v1 = v0 + vper < kw ? v0 + vper : kw;

  double acc[RB_BAND];
#pragma unroll
  for (int j = 0; j < RB_BAND; ++j) acc[j] = 0.0;

  const int total = srows * wcols;
  for (int u0 = 0; u0 < b.kr; u0 += b.kc) {
    const int kc = (b.kr - u0) < b.kc ? (b.kr - u0) : b.kc;
    __syncthreads();
    // kernel chunk, transposed: kt[v][u - u0]
    for (int e = tid; e < kw * kc; e += 64 * BAND_WAVES) {
      const int v = e / kc, u = e - v * kc;
      kt[e] = b.KT[(int64_t)v * b.kr + u0 + u];
    }
    // staged row i holds field row y0 - r + u0 + i; column j holds x0 - r + j
    for (int base = 0; base < total; base += 8 * 64 * BAND_WAVES) {
      float vals[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const int e = base + k * 64 * BAND_WAVES + tid;
        const int i = e / wcols, j = e - i * wcols;
        const int y = tl.y0 - b.r + u0 + i, x = tl.x0 - b.r + j;
        vals[k] = (e < total && y >= 0 && y < b.ny && x >= 0 && x < b.nx) ? sqmag(a[(int64_t)y * b.nx + x]) : 0.f;
      }
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const int e = base + k * 64 * BAND_WAVES + tid;
        if (e < total) {
          const int i = e / wcols;
          img[i * pitch + (e - i * wcols)] = vals[k];
        }
      }
    }
    __syncthreads();
    if (lane_ok) {
      const float* col = img + g * RB_BAND * pitch + c;
      for (int q = 0; q < kc; q += KU_BAND) {
        for (int vc = v0; vc < v1; vc += 8) {
          const int vend = vc + 8 < v1 ? vc + 8 : v1;
          float s[RB_BAND];
#pragma unroll
          for (int j = 0; j < RB_BAND; ++j) s[j] = 0.f;
          for (int vv = vc; vv < vend; ++vv) {
            const float4 k4 = *reinterpret_cast<const float4*>(kt + vv * kc + q);
            const float* p = col + q * pitch + vv;
            float w[RB_BAND + KU_BAND - 1];
#pragma unroll
            for (int i = 0; i < RB_BAND + KU_BAND - 1; ++i) w[i] = p[i * pitch];
#pragma unroll
            for (int j = 0; j < RB_BAND; ++j) {
              s[j] = fmaf(k4.x, w[j], s[j]);
              s[j] = fmaf(k4.y, w[j + 1], s[j]);
              s[j] = fmaf(k4.z, w[j + 2], s[j]);
              s[j] = fmaf(k4.w, w[j + 3], s[j]);
            }
          }
#pragma unroll
          for (int j = 0; j < RB_BAND; ++j) acc[j] += (double)s[j];
        }
      }
    }
  }
  // sum the partial results of the waves
#pragma unroll
  for (int j = 0; j < RB_BAND; ++j) red[(wv * 64 + lane) * RB_BAND + j] = acc[j];
  __syncthreads();
  if (wv != 0 || !lane_ok) return;
  const int x = tl.x0 + c;
#pragma unroll
  for (int j = 0; j < RB_BAND; ++j) {
    double sum = 0.0;
#pragma unroll
    for (int w2 = 0; w2 < BAND_WAVES; ++w2) sum += red[(w2 * 64 + lane) * RB_BAND + j];
    const int yo = g * RB_BAND + j;
    if (yo >= tl.nrows) break;
    const int y = tl.y0 + yo;
    const int bi = b.rowmap[y];
    float* out;
    if (bi >= 0)
      out = b.rowband + (slab * b.nb + bi) * (int64_t)b.nx + x;
    else
      out = b.colband + (slab * b.ny + y) * (int64_t)2 * b.r + (x < b.r ? x : x - b.nx + 2 * b.r);
    *out = (float)sum;
  }
}

}  // namespace nlse
