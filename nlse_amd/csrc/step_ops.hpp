// Pointwise operators of the split step (device side), shared by the fused
// line passes and the standalone operator kernels.
//
// Reference semantics: NLSE/kernels_cpu.py (numba, FP64 intermediates,
// one rounding at the store) and NLSE/kernels_gpu.py (saturated loss).
// The factor exp(dz*(-a + i theta)) is applied as A + A*(f - 1) with
//   f - 1 = expm1(-a)*cos(theta) - 2 sin^2(theta/2)
//         + i * exp(-a) * 2 sin(theta/2) cos(theta/2)
// so that f - 1 (small) carries full relative precision and the update
// rounds once at the magnitude of A.  A plain FP32 factor exp(-a) would
// carry a fixed rounding error (up to 3e-8) that a 1000-step run turns into
// a 3e-5 amplitude bias; the numba reference has no such bias.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nlse {

// One nonlinear application, single component (nl_prop / nl_prop_without_V).
struct NLParams1 {
  float adz;       // dz * alpha'        (alpha' = alpha/2)
  float gdz;       // dz * g             (g = k/2 n2 c eps0)
  float inv_isat;  // 1 / Isat'          (Isat' = 2 I_sat/(eps0 c)); 0 for inf
  float vdz;       // dz, multiplier of V' (k/2 V)
};

// One nonlinear application for one component of a two-component field
// (nl_prop_c / nl_prop_without_V_c, kernels_cpu.py:66-141).
struct NLParams2 {
  float adz;         // dz * alpha'
  float gs_dz;       // dz * g_self   (multiplies I_self * sat)
  float gc_dz;       // dz * g_cross  (multiplies I_other * sat)
  float inv_is_s;    // 1/Isat applied to I_self
  float inv_is_o;    // 1/Isat applied to I_other
  float vdz;         // dz * (multiplier of V for this component)
};

// Rabi rotation A1' = c A1 - i s A2, A2' = c A2 - i s A1 with c - 1 and s
// precomputed in FP64 (kernels_cpu.py:144-161).
struct RabiParams {
  float cm1;  // cos(omega dz) - 1
  float sn;   // sin(omega dz)
};

__device__ __forceinline__ float2 apply_phase_loss(float2 A, float a, float theta) {
  const float em = expm1f(-a);
  float s2, c2;
  sincosf(0.5f * theta, &s2, &c2);
  const float cosm1 = -2.0f * s2 * s2;
  const float dr = fmaf(em, 1.0f + cosm1, cosm1);
  const float di = (1.0f + em) * (2.0f * s2 * c2);
  const float cx = fmaf(A.x, dr, -A.y * di);
  const float cy = fmaf(A.y, dr, A.x * di);
  return make_float2(A.x + cx, A.y + cy);
}

// V kinds
enum : int { V_NONE = 0, V_REAL = 1, V_CPLX = 2 };

__device__ __forceinline__ void load_v(const void* V, int vkind, int64_t idx, float& vr, float& vi) {
  vr = 0.f;
  vi = 0.f;
  if (vkind == V_REAL) {
    vr = reinterpret_cast<const float*>(V)[idx];
  } else if (vkind == V_CPLX) {
    float2 w = reinterpret_cast<const float2*>(V)[idx];
    vr = w.x;
    vi = w.y;
  }
}

// nl_prop: A *= exp(dz*(-alpha' [*sat] + i g I sat + i V'))
// A complex V' = vr + i vi contributes i*V' = -vi + i vr.
__device__ __forceinline__ float2 nl1(float2 A, const NLParams1& p, bool loss_sat, float vr, float vi) {
  const float I = A.x * A.x + A.y * A.y;
  const float sat = 1.0f / (1.0f + I * p.inv_isat);
  const float theta = fmaf(p.gdz * I, sat, p.vdz * vr);
  const float a = fmaf(p.vdz, vi, loss_sat ? p.adz * sat : p.adz);
  return apply_phase_loss(A, a, theta);
}

// nl_prop_c for one component from the shared intensity snapshot.
__device__ __forceinline__ float2 nl2(float2 A, float Is, float Io, const NLParams2& p, float vr, float vi) {
  const float sat = 1.0f / (1.0f + Is * p.inv_is_s + Io * p.inv_is_o);
  const float theta = fmaf(p.gs_dz * Is, sat, fmaf(p.gc_dz * Io, sat, p.vdz * vr));
  const float a = fmaf(p.vdz, vi, p.adz * sat);
  return apply_phase_loss(A, a, theta);
}

__device__ __forceinline__ void rabi(float2& A1, float2& A2, const RabiParams& r) {
  // A1' = A1 + (c-1) A1 - i s A2 ;  -i s (x + i y) = s y - i s x
  const float2 a1 = A1, a2 = A2;
  A1 = make_float2(a1.x + fmaf(r.cm1, a1.x, r.sn * a2.y), a1.y + fmaf(r.cm1, a1.y, -r.sn * a2.x));
  A2 = make_float2(a2.x + fmaf(r.cm1, a2.x, r.sn * a1.y), a2.y + fmaf(r.cm1, a2.y, -r.sn * a1.x));
}

// A_c64 * P where P is the complex64 rounding of the FP64 product px*py
// (the reference builds exp(-i(kx^2+ky^2) dz/2k) in FP64 and casts it to
// complex64, nlse.py:194-197).  `scale` is an exact power of two (1/N).
__device__ __forceinline__ float2 prop_sep_c64(float2 A, double2 px, double2 py, float scale) {
  const double pr = fma(px.x, py.x, -px.y * py.y);
  const double pi = fma(px.x, py.y, px.y * py.x);
  const float2 P = make_float2((float)pr * scale, (float)pi * scale);
  return make_float2(fmaf(A.x, P.x, -A.y * P.y), fmaf(A.x, P.y, A.y * P.x));
}

__device__ __forceinline__ float2 prop_c64(float2 A, float2 P, float scale) {
  P.x *= scale;
  P.y *= scale;
  return make_float2(fmaf(A.x, P.x, -A.y * P.y), fmaf(A.x, P.y, A.y * P.x));
}

// complex64 field times a complex128 propagator, computed in FP64 with one
// rounding (numpy multiplies complex64 by complex128 in complex128 and casts
// back, nlse_1d.py:104, nlse_3d.py:112).
__device__ __forceinline__ float2 prop_c128(float2 A, double2 P, double scale) {
  const double ax = A.x, ay = A.y;
  const double px = P.x * scale, py = P.y * scale;
  return make_float2((float)fma(ax, px, -ay * py), (float)fma(ax, py, ay * px));
}

}  // namespace nlse
