// mp_adjoint.hip -- launch of the adjoint kernel (mp_adjoint.h) and the
// reduction of its per-wavefront parameter-gradient partial sums.
#include <mutex>

#include "mp_adjoint.h"

namespace shield {

// the EXT instances are compiled in mp_adjoint_ext.hip
extern template hipError_t launch_adjoint_kernel<1, true>(unsigned, const AdjArgs&, hipStream_t);
extern template hipError_t launch_adjoint_kernel<2, true>(unsigned, const AdjArgs&, hipStream_t);
extern template hipError_t launch_adjoint_kernel<3, true>(unsigned, const AdjArgs&, hipStream_t);

// Device scratch of shield_mp_adjoint_noalloc: the tapes and parameter
// partial sums of as many wavefronts as fit (SHIELD_MP_ADJOINT_SCRATCH_BYTES,
// each GPU holds its own copy for the life of the process).  Host code never
// uses the array's name as an address: the per-device address comes from
// hipGetSymbolAddress (scratch_address).
static __device__ __align__(256) double scratch[SHIELD_MP_ADJOINT_SCRATCH_BYTES / sizeof(double)];

struct AdjLayout {
  int S, P;  // slots, columns per wavefront
};

__global__ __launch_bounds__(64) void adj_param_reduce(const double* part, int64_t nw, double* out) {
  const int p = threadIdx.x;
  if (p >= AP_N) return;
  double acc = 0.0;
  for (int64_t w = 0; w < nw; ++w) acc += part[w * AP_N + p];
  out[p] = acc;
}

// P columns in S slots: the most busy lanes over S = 1..3 (the fewest slots
// on a tie); 2 columns in 3 slots at nk = 79
AdjLayout adjoint_layout(int64_t nk) {
  AdjLayout best{0, 0};
  double util = -1.0;
  for (int S = 1; S <= 3; ++S) {
    const int64_t p0 = (64 * S) / nk;
    if (p0 < 1) continue;
    const int P = (int)(p0 > pmax_for(S) ? pmax_for(S) : p0);
    const double u = double(P * nk) / (64.0 * S);
    if (u > util + 1e-9) {
      util = u;
      best = AdjLayout{S, P};
    }
  }
  return best;
}

int adjoint_cols_per_wave(int64_t nk) { return adjoint_layout(nk).P; }

// doubles of tape per wavefront (with do_sedi_w, SW_FIELDS more per sub-step)
int64_t adjoint_tape_doubles(int64_t nk, int ntimes, bool sedi_w) {
  return (int64_t)ntimes * (REC_FIELDS + (sedi_w ? SW_FIELDS : 0)) * 64 * adjoint_layout(nk).S;
}

hipError_t launch_adjoint(const shield_mp_shape& sh, const shield_mp_params& p, const shield_mp_fields& f,
                          const shield_mp_adjoint_fields& a, void* ws, size_t ws_bytes, hipStream_t stream) {
  const int64_t ncol = sh.ni * sh.nj;
  const AdjLayout lay = adjoint_layout(sh.nk);
  const int64_t groups = (ncol + lay.P - 1) / lay.P;
  const int64_t tape = adjoint_tape_doubles(sh.nk, p.ntimes, p.do_sedi_w != 0);
  const int64_t per_wave = (tape + AP_N) * (int64_t)sizeof(double);
  int64_t nw = (int64_t)(ws_bytes / (size_t)per_wave);
  if (nw > groups) nw = groups;
  if (nw < 1) return hipErrorInvalidValue;
  double* base = static_cast<double*>(ws);
  double* part = base + nw * tape;
  const AdjArgs args{sh, p, make_coef<double>(p), f, a, base, tape, part, lay.P};
  const unsigned grid = (unsigned)nw;
  // fix_negative, the ice slope limiter and the w transport have an instance
  // of their own
  const bool ext = p.do_sedi_w || p.nml.z_slope_ice || p.nml.fix_negative;
  hipError_t e;
  if (lay.S == 1)
    e = ext ? launch_adjoint_kernel<1, true>(grid, args, stream) : launch_adjoint_kernel<1, false>(grid, args, stream);
  else if (lay.S == 2)
    e = ext ? launch_adjoint_kernel<2, true>(grid, args, stream) : launch_adjoint_kernel<2, false>(grid, args, stream);
  else
    e = ext ? launch_adjoint_kernel<3, true>(grid, args, stream) : launch_adjoint_kernel<3, false>(grid, args, stream);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(adj_param_reduce, dim3(1), dim3(64), 0, stream, part, nw, a.params);
  return hipGetLastError();
}

// The address of this device's copy of `scratch`, looked up once per device
// with that device current; nullptr on failure.
static void* scratch_address() {
  constexpr int kMaxDevices = 64;
  static void* addr[kMaxDevices] = {};
  static std::mutex mu;
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess) return nullptr;
    addr[dev] = p;
  }
  return addr[dev];
}

hipError_t launch_adjoint_noalloc(const shield_mp_shape& sh, const shield_mp_params& p, const shield_mp_fields& f,
                                  const shield_mp_adjoint_fields& a, hipStream_t stream) {
  void* ws = scratch_address();
  if (!ws) return hipErrorInvalidSymbol;
  return launch_adjoint(sh, p, f, a, ws, SHIELD_MP_ADJOINT_SCRATCH_BYTES, stream);
}

}  // namespace shield
