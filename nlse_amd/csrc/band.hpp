// Band-point convolution kernel of the nonlocal response (one definition:
// included by engine.hip only; the geometry and arguments are in
// nonlocal.hpp).
#pragma once
#include "nonlocal.hpp"

namespace nlse {

// Direct FP64 convolution at the points of one tile (blockIdx.x) of one
// slab (blockIdx.y).  The intensities the tile needs are staged in LDS per
// chunk of kc kernel rows (batched loads, 8 per thread in flight); the
// BAND_WAVES waves split the kernel columns and their partial sums are
// added at the end.  Each lane slides a register window of RB_BAND + KU_BAND - 1 rows
// down its column, so one LDS read feeds KU_BAND FMAs.
__global__ __launch_bounds__(64 * BAND_WAVES) void k_nl_band(BandArgs b) {
  extern __shared__ double smem_d[];
  float* smem = reinterpret_cast<float*>(smem_d);
  const BandTile tl = b.tiles[blockIdx.x];
  const int64_t slab = blockIdx.y;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wv = tid >> 6;
  const int kw = 2 * b.r + 1;
  const int groups = 64 / tl.ncols;
  const int wcols = tl.ncols + 2 * b.r;
  const int srows = groups * RB_BAND + b.kc - 1;
  const size_t img_n = band_img_floats(tl.ncols, b.r, b.kc);
  float* img = smem;
  double* kch = reinterpret_cast<double*>(smem + img_n);
  double* red = kch + (size_t)b.kc * kw;
  const int c = lane % tl.ncols;
  const int g = lane / tl.ncols;
  const bool lane_ok = g < groups;
  const float2* a = b.A + slab * (int64_t)b.ny * b.nx;
  // kernel columns of this wave
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
    // staged row i holds field row y0 - r + u0 + i; column j holds x0 - r + j
    for (int base = 0; base < total; base += 8 * 64 * BAND_WAVES) {
      float vals[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const int e = base + k * 64 * BAND_WAVES + tid;
        const int i = e / wcols, j = e - (e / wcols) * wcols;
        const int y = tl.y0 - b.r + u0 + i, x = tl.x0 - b.r + j;
        vals[k] = (e < total && y >= 0 && y < b.ny && x >= 0 && x < b.nx) ? sqmag(a[(int64_t)y * b.nx + x]) : 0.f;
      }
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const int e = base + k * 64 * BAND_WAVES + tid;
        if (e < total) img[e] = vals[k];
      }
    }
    for (int e = tid; e < kc * kw; e += 64 * BAND_WAVES) kch[e] = b.K[(int64_t)u0 * kw + e];
    __syncthreads();
    if (lane_ok) {
      for (int q = 0; q < kc; q += KU_BAND) {
        for (int vv = v0; vv < v1; ++vv) {
          float w[RB_BAND + KU_BAND - 1];
#pragma unroll
          for (int i = 0; i < RB_BAND + KU_BAND - 1; ++i) w[i] = img[(g * RB_BAND + q + i) * wcols + c + vv];
#pragma unroll
          for (int u = 0; u < KU_BAND; ++u) {
            const double k = kch[(q + u) * kw + vv];
#pragma unroll
            for (int j = 0; j < RB_BAND; ++j) acc[j] = fma(k, (double)w[j + u], acc[j]);
          }
        }
      }
    }
  }
  // sum the partial results of the four waves
#pragma unroll
  for (int j = 0; j < RB_BAND; ++j) red[(wv * 64 + lane) * RB_BAND + j] = acc[j];
  __syncthreads();
  if (wv != 0 || !lane_ok) return;
#pragma unroll
  for (int j = 0; j < RB_BAND; ++j) {
    double s = 0.0;
#pragma unroll
    for (int w2 = 0; w2 < BAND_WAVES; ++w2) s += red[(w2 * 64 + lane) * RB_BAND + j];
    acc[j] = s;
  }
  const int x = tl.x0 + c;
#pragma unroll
  for (int j = 0; j < RB_BAND; ++j) {
    const int yo = g * RB_BAND + j;
    if (yo >= tl.nrows) break;
    const int y = tl.y0 + yo;
    const int bi = b.rowmap[y];
    float* out;
    if (bi >= 0)
      out = b.rowband + (slab * b.nb + bi) * (int64_t)b.nx + x;
    else
      out = b.colband + (slab * b.ny + y) * (int64_t)2 * b.r + (x < b.r ? x : x - b.nx + 2 * b.r);
    *out = (float)acc[j];
  }
}

}  // namespace nlse
