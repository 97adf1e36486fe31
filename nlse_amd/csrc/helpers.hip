// Initial-state and drive helpers on the device: the vortex phase imprint of
// kernels_gpu.py:159-176, the noise of DDGPE.add_noise (ddgpe.py:96-125)
// from a counter-based generator (Philox4x64-10) and the dense 2-D
// propagator table of nlse.py:188-197.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "../../include/nlse_hip.h"
#include "common.hpp"

using namespace nlse;

namespace {

unsigned grid_for(int64_t n) {
  int64_t b = (n + 255) / 256;
  if (b > 65536) b = 65536;
  if (b < 1) b = 1;
  return (unsigned)b;
}

// ------------------------------------------------------------ vortex
// Coordinates and phase images come as float32, float64, int32 or int64.
__device__ __forceinline__ double load_num(const void* p, int kind, int64_t i) {
  switch (kind) {
    case NLSE_NUM_F32: return (double)static_cast<const float*>(p)[i];
    case NLSE_NUM_F64: return static_cast<const double*>(p)[i];
    case NLSE_NUM_I32: return (double)static_cast<const int32_t*>(p)[i];
    default: return (double)static_cast<const int64_t*>(p)[i];
  }
}

// arg(z^l) for z = x + iy: the principal value of l * atan2(y, x), in
// (-pi, pi].  At z = 0 the result is 0 (the phase of a vortex core).
__device__ __forceinline__ double vortex_phase(double x, double y, int l) {
  const double two_pi = 6.283185307179586476925286766559;
  const double t = (double)l * atan2(y, x);
  double w = t - two_pi * rint(t / two_pi);
  if (w <= -3.1415926535897932384626433832795) w += two_pi;
  return w;
}

__global__ void k_vortex(void* im, int im_kind, int64_t n, double i0, double j0, const void* ii, const void* jj,
                         int c_kind, int l) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
    const double ph = vortex_phase(load_num(ii, c_kind, k) - i0, load_num(jj, c_kind, k) - j0, l);
    if (im_kind == NLSE_NUM_F32) {
      float* f = static_cast<float*>(im);
      f[k] = (float)((double)f[k] + ph);
    } else {
      double* d = static_cast<double*>(im);
      d[k] += ph;
    }
  }
}

// ------------------------------------------------------------ Philox4x64-10
// Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3"
// (SC11): ten rounds of two 64x64->128 multiplies with a Weyl key schedule.
struct U4 {
  uint64_t v[4];
};

__device__ __forceinline__ U4 philox4x64_10(U4 c, uint64_t k0, uint64_t k1) {
  const uint64_t M0 = 0xD2E7470EE14C6C93ull, M1 = 0xCA5A826395121157ull;
  const uint64_t W0 = 0x9E3779B97F4A7C15ull, W1 = 0xBB67AE8584CAA73Bull;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r) {
      k0 += W0;
      k1 += W1;
    }
    const uint64_t lo0 = M0 * c.v[0], hi0 = __umul64hi(M0, c.v[0]);
    const uint64_t lo1 = M1 * c.v[2], hi1 = __umul64hi(M1, c.v[2]);
    U4 o;
    o.v[0] = hi1 ^ c.v[1] ^ k0;
    o.v[1] = lo1;
    o.v[2] = hi0 ^ c.v[3] ^ k1;
    o.v[3] = lo0;
    c = o;
  }
  return c;
}

// block b of a stream: counter (c0 + b + 1, c1, c2, c3) with carry, the
// counter increment-then-encrypt order of numpy.random.Philox
__device__ __forceinline__ U4 philox_block(const uint64_t* ctr, uint64_t k0, uint64_t k1, uint64_t b) {
  U4 c;
  c.v[0] = ctr[0] + b + 1;
  uint64_t carry = c.v[0] < ctr[0] ? 1 : 0;
  c.v[1] = ctr[1] + carry;
  carry = carry && c.v[1] == 0 ? 1 : 0;
  c.v[2] = ctr[2] + carry;
  carry = carry && c.v[2] == 0 ? 1 : 0;
  c.v[3] = ctr[3] + carry;
  return philox4x64_10(c, k0, k1);
}

struct Ctr {
  uint64_t v[4];
};

__global__ void k_philox(uint64_t* out, int64_t nblocks, Ctr ctr, uint64_t k0, uint64_t k1) {
  for (int64_t b = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; b < nblocks;
       b += (int64_t)gridDim.x * blockDim.x) {
    const U4 r = philox_block(ctr.v, k0, k1, (uint64_t)b);
    // four consecutive words of one block, two 16-byte stores
    reinterpret_cast<ulonglong2*>(out)[2 * b] = make_ulonglong2(r.v[0], r.v[1]);
    reinterpret_cast<ulonglong2*>(out)[2 * b + 1] = make_ulonglong2(r.v[2], r.v[3]);
  }
}

// uniform double in [0, 1) from the top 53 bits (numpy's random_raw -> random)
__device__ __forceinline__ double u53(uint64_t w) { return (double)(w >> 11) * (1.0 / 9007199254740992.0); }

// two standard normals from two uniforms (Box and Muller, 1958), with 1 - u1
// in (0, 1] so the logarithm is finite
__device__ __forceinline__ void box_muller(uint64_t w1, uint64_t w2, double& z0, double& z1) {
  const double r = sqrt(-2.0 * log(1.0 - u53(w1)));
  double s, c;
  sincospi(2.0 * u53(w2), &s, &c);
  z0 = r * c;
  z1 = r * s;
}

// DDGPE.add_noise: A[..., 0, :, :] += amp0 * rand1, A[..., 1, :, :] += amp1 * rand2
// with rand = dz * (z_re + i z_im).  Point p of the transverse grid draws
// block p of the stream: words (0, 1) give rand1, words (2, 3) rand2; the
// same draw is added to every item, as the reference's (NY, NX) draw
// broadcasts over the leading axes.  Complex128 arithmetic, rounded to
// complex64 on the store (numpy's in-place add).
__global__ void k_add_noise(float2* A, int64_t items, int64_t isz, Ctr ctr, uint64_t k0, uint64_t k1, double dz,
                            double amp0, double amp1) {
#pragma clang fp contract(off)  // numpy rounds amp * (dz * z) before adding it to A
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < isz; p += (int64_t)gridDim.x * blockDim.x) {
    const U4 r = philox_block(ctr.v, k0, k1, (uint64_t)p);
    double a, b, c, d;
    box_muller(r.v[0], r.v[1], a, b);
    box_muller(r.v[2], r.v[3], c, d);
    const double t0r = amp0 * (dz * a), t0i = amp0 * (dz * b);
    const double t1r = amp1 * (dz * c), t1i = amp1 * (dz * d);
    for (int64_t it = 0; it < items; ++it) {
      float2* e = A + it * 2 * isz + p;
      float2* h = e + isz;
      const float2 x = *e, y = *h;
      *e = make_float2((float)((double)x.x + t0r), (float)((double)x.y + t0i));
      *h = make_float2((float)((double)y.x + t1r), (float)((double)y.y + t1i));
    }
  }
}

// ------------------------------------------------------------ propagator
// Dense table of nlse.py:188-197, exp(-1j * 0.5 * (Kxx**2 + Kyy**2) / k * dz)
// cast to complex64, with numpy's FP64 operation order: the product with
// -0.5j leaves a zero real part and an imaginary part -0.5 S exactly, the
// complex division by the real k multiplies by 1/k (numpy's division loop
// scales by the reciprocal of the divisor), the product with dz is one
// rounding, and exp of (0 + i theta) is (cos theta, sin theta).
__global__ void k_prop_dense(float2* out, const double* kx, const double* ky, int64_t nx, int64_t n, double inv_k,
                             double dz) {
#pragma clang fp contract(off)  // every product and sum rounds as numpy's do
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x) {
    const double x = kx[p % nx], y = ky[p / nx];
    const double s = x * x + y * y;
    const double theta = ((-0.5 * s) * inv_k) * dz;
    double sn, cs;
    sincos(theta, &sn, &cs);
    out[p] = make_float2((float)cs, (float)sn);
  }
}

bool num_kind_ok(int k) { return k == NLSE_NUM_F32 || k == NLSE_NUM_F64 || k == NLSE_NUM_I32 || k == NLSE_NUM_I64; }

}  // namespace

extern "C" {

int nlse_vortex(void* im, int im_kind, int64_t n, double i0, double j0, const void* ii, const void* jj,
                int coord_kind, int charge, void* stream) {
  if (!im || !ii || !jj || n < 0 || (im_kind != NLSE_NUM_F32 && im_kind != NLSE_NUM_F64) ||
      !num_kind_ok(coord_kind))
    return fail(NLSE_ERR_ARG, "bad vortex arguments");
  if (n == 0) return NLSE_OK;
  hipLaunchKernelGGL(k_vortex, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, im, im_kind, n, i0, j0, ii, jj,
                     coord_kind, charge);
  return launch_status("k_vortex");
}

int nlse_propagator_dense(void* out, const double* kx, const double* ky, int64_t nx, int64_t ny, double inv_k,
                          double dz, void* stream) {
  if (!out || !kx || !ky || nx < 0 || ny < 0) return fail(NLSE_ERR_ARG, "bad propagator arguments");
  if (nx == 0 || ny == 0) return NLSE_OK;
  const int64_t n = nx * ny;
  hipLaunchKernelGGL(k_prop_dense, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, (float2*)out, kx, ky, nx, n,
                     inv_k, dz);
  return launch_status("k_prop_dense");
}

int nlse_philox4x64(uint64_t* out, int64_t nblocks, const uint64_t* counter, const uint64_t* key, void* stream) {
  if (!out || !counter || !key || nblocks < 0) return fail(NLSE_ERR_ARG, "bad philox arguments");
  if (nblocks == 0) return NLSE_OK;
  Ctr c;
  for (int i = 0; i < 4; ++i) c.v[i] = counter[i];
  hipLaunchKernelGGL(k_philox, dim3(grid_for(nblocks)), dim3(256), 0, (hipStream_t)stream, out, nblocks, c, key[0],
                     key[1]);
  return launch_status("k_philox");
}

int nlse_add_noise(void* A, int64_t items, int64_t isz, const uint64_t* counter, const uint64_t* key, double dz,
                   double amp0, double amp1, void* stream) {
  if (!A || !counter || !key || items < 0 || isz < 0) return fail(NLSE_ERR_ARG, "bad add_noise arguments");
  if (items == 0 || isz == 0) return NLSE_OK;
  Ctr c;
  for (int i = 0; i < 4; ++i) c.v[i] = counter[i];
  hipLaunchKernelGGL(k_add_noise, dim3(grid_for(isz)), dim3(256), 0, (hipStream_t)stream, (float2*)A, items, isz, c,
                     key[0], key[1], dz, amp0, amp1);
  return launch_status("k_add_noise");
}

}  // extern "C"
