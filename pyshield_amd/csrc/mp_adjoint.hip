// mp_adjoint.hip -- launch of the adjoint kernel (mp_adjoint.h) and the
// reduction of its per-wavefront parameter-gradient partial sums.
#include "mp_adjoint.h"

namespace shield {

__global__ __launch_bounds__(64) void adj_param_reduce(const double* part, int64_t nw, double* out) {
  const int p = threadIdx.x;
  if (p >= AP_N) return;
  double acc = 0.0;
  for (int64_t w = 0; w < nw; ++w) acc += part[w * AP_N + p];
  out[p] = acc;
}

int adjoint_slots(int64_t nk) { return (int)((nk + 63) / 64); }

// doubles of tape per wavefront
int64_t adjoint_tape_doubles(int64_t nk, int ntimes) {
  return (int64_t)ntimes * REC_FIELDS * 64 * adjoint_slots(nk);
}

hipError_t launch_adjoint(const shield_mp_shape& sh, const shield_mp_params& p, const shield_mp_fields& f,
                          const shield_mp_adjoint_fields& a, void* ws, size_t ws_bytes, hipStream_t stream) {
  const int64_t ncol = sh.ni * sh.nj;
  const int64_t tape = adjoint_tape_doubles(sh.nk, p.ntimes);
  const int64_t per_wave = (tape + AP_N) * (int64_t)sizeof(double);
  int64_t nw = (int64_t)(ws_bytes / (size_t)per_wave);
  if (nw > ncol) nw = ncol;
  if (nw < 1) return hipErrorInvalidValue;
  double* base = static_cast<double*>(ws);
  double* part = base + nw * tape;
  const Coef<double> c = make_coef<double>(p);
  const dim3 grid((unsigned)nw), block(64);
  if (adjoint_slots(sh.nk) == 1)
    hipLaunchKernelGGL(mp_adjoint_kernel<1>, grid, block, 0, stream, sh, p, c, f, a, base, tape, part);
  else
    hipLaunchKernelGGL(mp_adjoint_kernel<2>, grid, block, 0, stream, sh, p, c, f, a, base, tape, part);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(adj_param_reduce, dim3(1), dim3(64), 0, stream, part, nw, a.params);
  return hipGetLastError();
}

}  // namespace shield
