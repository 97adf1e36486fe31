// Per-line-length kernel dispatch: twiddle tables, launch configurations
// and launchers of every kernel templated on the line length N.  Each
// lines_*.hip translation unit instantiates make_ops<N> for its lengths, so
// the kernels of different lengths compile in parallel.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <map>
#include <mutex>
#include <vector>

#include "common.hpp"
#include "passes_2d.hpp"

#ifndef NLSE_COL_LDS_KB
#define NLSE_COL_LDS_KB 150  // LDS budget of one column-sweep block
#endif

namespace nlse {

// Runtime-N dispatch table.
struct LineOps {
  int (*twiddles)(const float4**);
  int (*row)(int, const RowArgs&, hipStream_t);
  int (*line1d)(int, const Line1DArgs&, hipStream_t);
  int (*nlc)(int, const NlcArgs&, hipStream_t);
  int (*col)(int, const ColArgs&, hipStream_t);
  int (*colw)(int64_t, int);
  int lpb1, lpb2, t;
};

// ------------------------------------------------------ twiddle tables
// One table per (device, line length), layout of Plan<N,P>::tw_offset:
// stage s (s >= 1), entry [(k-1)*Ns + jm] = exp(-2 pi i jm k / (Ns R)),
// evaluated in FP64 and stored as {c_hi, s_hi, c_lo, s_lo} with
// c = c_hi + c_lo to ~2^-48.
template <int N, int P>
std::vector<float4> build_twiddles() {
  using PL = Plan<N, P>;
  std::vector<float4> tw(PL::TW_SIZE > 0 ? PL::TW_SIZE : 1);
  for (int s = 1; s < PL::NSTAGES; ++s) {
    const int R = PL::radix(s), NS = PL::ns(s), off = PL::tw_offset(s);
    for (int k = 1; k < R; ++k)
      for (int jm = 0; jm < NS; ++jm) {
        // angle = -2 pi * jm*k/(NS*R); reduce the rational first
        const long num = (long)jm * k, den = (long)NS * R;
        const double ang = -2.0 * M_PI * (double)num / (double)den;
        const double c = std::cos(ang), sn = std::sin(ang);
        const float ch = (float)c, sh = (float)sn;
        tw[off + (k - 1) * NS + jm] = make_float4(ch, sh, (float)(c - (double)ch), (float)(sn - (double)sh));
      }
  }
  return tw;
}

static std::mutex g_tw_mu;
static std::map<std::pair<int, int>, float4*> g_tw;  // (device, N) -> table

template <int N, int P>
static int get_twiddles(const float4** out) {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_tw_mu);
  auto key = std::make_pair(dev, N);
  auto it = g_tw.find(key);
  if (it != g_tw.end()) {
    *out = it->second;
    return NLSE_OK;
  }
  std::vector<float4> h = build_twiddles<N, P>();
  float4* d = nullptr;
  HIP_TRY(hipMalloc(&d, h.size() * sizeof(float4)));
  HIP_TRY(hipMemcpy(d, h.data(), h.size() * sizeof(float4), hipMemcpyHostToDevice));
  g_tw[key] = d;
  *out = d;
  return NLSE_OK;
}

// -------------------------------------------------- kernel dispatchers
template <int N>
struct LineCfg {
  static constexpr int P = N >= 16 ? 16 : N;
  static constexpr int T = N / P;
  static constexpr int LINE = LdsLine<N, 3, 3>::SIZE;  // column-pass lines
  // rows per block: 256 threads for one component, 128 for pairs
  static constexpr int LPB1 = (256 / T) >= 1 ? (256 / T) : 1;
  static constexpr int LPB2 = (128 / T) >= 1 ? (128 / T) : 1;
  // default column tile: 8 columns (64-byte row segments; wider tiles
  // spill registers at 512 threads)
  static constexpr int WDEF = 8;
};

// Column tile of W columns: WT threads across (<= 512 threads per block,
// i.e. up to 256 VGPRs per lane), G = W/WT columns per thread, LDS rounds
// of GR columns per thread column (<= 150 KiB of LDS).
template <int N, int W>
struct ColCfg {
  using C = LineCfg<N>;
  static constexpr int WTMAX = (512 / C::T) >= 1 ? (512 / C::T) : 1;
  static constexpr int WT = W < WTMAX ? W : WTMAX;
  static constexpr int G = W / WT;
  static constexpr int GRMAX = (NLSE_COL_LDS_KB * 1024) / (WT * C::LINE * 8);
  static constexpr int gr(int g) { return (g <= GRMAX || g == 1) ? g : gr(g / 2); }
  static constexpr int GR = gr(G);
  static constexpr bool OK = G >= 1 && G <= 4 && GR >= 1 && WT * GR * C::LINE * 8 <= 160 * 1024;
};

template <int N>
static int row_launch(int G, const RowArgs& a, hipStream_t s) {
  using C = LineCfg<N>;
  if (G == 1) {
    constexpr int L = C::LPB1;
    const int64_t blocks = (a.ngroups + L - 1) / L;
    hipLaunchKernelGGL((k_row<N, C::P, L, 1>), dim3((unsigned)blocks), dim3(L * C::T), 0, s, a);
  } else {
    constexpr int L = C::LPB2;
    const int64_t blocks = (a.ngroups + L - 1) / L;
    hipLaunchKernelGGL((k_row<N, C::P, L, 2>), dim3((unsigned)blocks), dim3(L * C::T), 0, s, a);
  }
  return launch_status("k_row");
}

template <int N>
static int line1d_launch(int G, const Line1DArgs& a, hipStream_t s) {
  using C = LineCfg<N>;
  if (G == 1) {
    constexpr int L = C::LPB1;
    const int64_t blocks = (a.ngroups + L - 1) / L;
    hipLaunchKernelGGL((k_line1d<N, C::P, L, 1>), dim3((unsigned)blocks), dim3(L * C::T), 0, s, a);
  } else {
    constexpr int L = C::LPB2;
    const int64_t blocks = (a.ngroups + L - 1) / L;
    hipLaunchKernelGGL((k_line1d<N, C::P, L, 2>), dim3((unsigned)blocks), dim3(L * C::T), 0, s, a);
  }
  return launch_status("k_line1d");
}

// Nonlocal line kernels: stage 1 (k_nlc_z) or stage 2 (k_nlc_nl), one
// line pair per T threads.
template <int N>
static int nlc_launch(int stage, const NlcArgs& a, hipStream_t s) {
  using C = LineCfg<N>;
  constexpr int L = C::LPB2;
  const int64_t blocks = (a.ngroups + L - 1) / L;
  const bool full = a.full && a.ncomp == 1;
  if (stage == 1 && full)
    hipLaunchKernelGGL((k_nlc_z<N, C::P, L, 1>), dim3((unsigned)blocks), dim3(L * C::T), 0, s, a);
  else if (stage == 1)
    hipLaunchKernelGGL((k_nlc_z<N, C::P, L>), dim3((unsigned)blocks), dim3(L * C::T), 0, s, a);
  else if (full)
    hipLaunchKernelGGL((k_nlc_nl<N, C::P, L, 1, 1>), dim3((unsigned)blocks), dim3(L * C::T), 0, s, a);
  else if (a.ncomp == 1)
    hipLaunchKernelGGL((k_nlc_nl<N, C::P, L, 1>), dim3((unsigned)blocks), dim3(L * C::T), 0, s, a);
  else
    hipLaunchKernelGGL((k_nlc_nl<N, C::P, L, 2>), dim3((unsigned)blocks), dim3(L * C::T), 0, s, a);
  return launch_status(stage == 1 ? "k_nlc_z" : "k_nlc_nl");
}

template <int N, int W>
static bool col_launch_w(const ColArgs& a, hipStream_t s) {
  using C = LineCfg<N>;
  using K = ColCfg<N, W>;
  if constexpr (K::OK) {
    const int64_t blocks = a.nslabs * (a.NX / W);
    // the separable 2-D propagator and the real kernel spectrum of the
    // nonlocal response get their own instances (fewer live registers
    // than the run-time dispatch over every kind)
    const bool prop = a.do_prop != 0;
    if (prop && a.pkind == PROP_SEP64)
      hipLaunchKernelGGL((k_col<N, C::P, K::WT, K::G, K::GR, PROP_SEP64>), dim3((unsigned)blocks),
                         dim3(K::WT * C::T), 0, s, a);
    else if (prop && a.pkind == PROP_KREAL)
      hipLaunchKernelGGL((k_col<N, C::P, K::WT, K::G, K::GR, PROP_KREAL>), dim3((unsigned)blocks),
                         dim3(K::WT * C::T), 0, s, a);
    else
      hipLaunchKernelGGL((k_col<N, C::P, K::WT, K::G, K::GR>), dim3((unsigned)blocks), dim3(K::WT * C::T), 0, s,
                         a);
    return true;
  } else {
    return false;
  }
}

template <int N>
static bool col_ok(int w) {
  switch (w) {
    case 1: return ColCfg<N, 1>::OK;
    case 2: return ColCfg<N, 2>::OK;
    case 4: return ColCfg<N, 4>::OK;
    case 8: return ColCfg<N, 8>::OK;
    case 16: return ColCfg<N, 16>::OK;
    case 32: return ColCfg<N, 32>::OK;
    case 64: return ColCfg<N, 64>::OK;
    default: return false;
  }
}

template <int N>
static int col_width_for(int64_t ncols, int want) {
  int w = want > 0 ? want : LineCfg<N>::WDEF;
  if (w > 64) w = 64;
  while (w > 1 && (w > ncols || (ncols % w) != 0 || !col_ok<N>(w))) w >>= 1;
  return w;
}

template <int N>
static int col_launch(int W, const ColArgs& a, hipStream_t s) {
  bool ok = false;
  switch (W) {
    case 1: ok = col_launch_w<N, 1>(a, s); break;
    case 2: ok = col_launch_w<N, 2>(a, s); break;
    case 4: ok = col_launch_w<N, 4>(a, s); break;
    case 8: ok = col_launch_w<N, 8>(a, s); break;
    case 16: ok = col_launch_w<N, 16>(a, s); break;
    case 32: ok = col_launch_w<N, 32>(a, s); break;
    case 64: ok = col_launch_w<N, 64>(a, s); break;
    default: break;
  }
  if (!ok) return fail(NLSE_ERR_ARG, "unsupported column width %d for length %d", W, N);
  return launch_status("k_col");
}

template <int N>
LineOps make_ops() {
  using C = LineCfg<N>;
  return LineOps{&get_twiddles<N, C::P>, &row_launch<N>,     &line1d_launch<N>, &nlc_launch<N>,
                 &col_launch<N>,         &col_width_for<N>, C::LPB1,           C::LPB2,
                 C::T};
}


// Dispatch table entry of a line length (engine.hip), or null.
const LineOps* line_ops_public(int64_t n);

// Defined in the lines_*.hip units.
LineOps make_ops_8();
LineOps make_ops_16();
LineOps make_ops_32();
LineOps make_ops_64();
LineOps make_ops_128();
LineOps make_ops_256();
LineOps make_ops_512();
LineOps make_ops_1024();
LineOps make_ops_2048();
LineOps make_ops_4096();

}  // namespace nlse
