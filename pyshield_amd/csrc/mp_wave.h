// mp_wave.h -- column operations for the "levels across lanes" layout.
//
// One 64-lane wavefront owns `ncols` consecutive columns.  Their cells are
// flattened column-major (cell = col_local * nz + k, k = 0 at the model top)
// and dealt to the wavefront in S slots: slot s, lane l holds cell 64 s + l.
// With the reference's [X, Y, Z] storage (k contiguous) one slot of one field
// is a contiguous 512-byte run, so every global load and store is coalesced.
//
// Vertical operations of the reference's FORWARD/BACKWARD blocks become:
//   * column any / first-hit  -> wave ballots + scalar bit arithmetic,
//   * k-1 / k+1 / surface / top values -> one LDS round trip,
//   * linear recurrences x[k] = a[k] x[k-1] + b[k] -> segmented scans over
//     the 64 lanes of a slot (DPP row shifts and row broadcasts) with a carry
//     between slots; a segment starts at k = 0, so columns never mix.
// The workgroup is exactly one wavefront, so __syncthreads() is the only
// synchronisation the LDS round trips need.
#pragma once

#include <utility>

#include "mp_common.h"

namespace shield {

template <int S>
struct Lay {
  int k[S];   // level of the cell (0 for cells past the last column)
  int kb[S];  // nz-1-k (0 for cells past the last column)
  int cl[S];  // local column index (0 for cells past the last column)
  bool v[S];  // cell belongs to a column of this wavefront
  int nz, ncols;
};

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }

// bits [lo, hi) of a 64-bit mask, clipped to [0, 64)
__device__ __forceinline__ uint64_t range_mask(int lo, int hi) {
  lo = lo < 0 ? 0 : lo;
  hi = hi > 64 ? 64 : hi;
  if (hi <= lo) return 0ull;
  const int n = hi - lo;
  const uint64_t m = (n == 64) ? ~0ull : ((1ull << n) - 1ull);
  return m << lo;
}

// Calls f(std::integral_constant<int, s>) for s = 0 .. S-1: a loop over the
// slots that is unrolled by construction, for bodies too large for the
// unroller's size limit (a rolled slot loop would index the per-slot state
// arrays dynamically and move them to scratch memory).
template <class F, int... I>
__device__ __forceinline__ void for_slots_impl(F&& f, std::integer_sequence<int, I...>) {
  (f(std::integral_constant<int, I>{}), ...);
}
template <int S, class F>
__device__ __forceinline__ void for_slots(F&& f) {
  for_slots_impl(f, std::make_integer_sequence<int, S>{});
}

// number of set bits of m below this lane
__device__ __forceinline__ int bits_below(uint64_t m) {
  return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// A stable counting sort of the wavefront's 64 S cells by a class 0, 1 or 2:
// pos[s] is the sorted position of this lane's cell in slot s, src[s] the
// cell (64 s' + lane') that sorted position 64 s + lane holds.  `ib` is LDS
// scratch of 64 S ints; the workgroup is synchronised on return.
template <int S>
__device__ __forceinline__ void sort_cells3(const int (&cls)[S], int (&pos)[S], int (&src)[S], int* ib) {
  const int lane = lane_id();
  uint64_t b0[S], b1[S];
  int p0[S], p1[S], n0 = 0, n1 = 0;
#pragma unroll
  for (int s = 0; s < S; ++s) {
    b0[s] = __ballot(cls[s] == 0);
    b1[s] = __ballot(cls[s] == 1);
    p0[s] = n0;
    p1[s] = n1;
    n0 += __popcll(b0[s]);
    n1 += __popcll(b1[s]);
  }
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const int idx = 64 * s + lane;
    const int r0 = p0[s] + bits_below(b0[s]), r1 = p1[s] + bits_below(b1[s]);
    pos[s] = cls[s] == 0 ? r0 : cls[s] == 1 ? n0 + r1 : n0 + n1 + (idx - r0 - r1);
    ib[pos[s]] = idx;
  }
  __syncthreads();
#pragma unroll
  for (int s = 0; s < S; ++s) src[s] = ib[64 * s + lane];
  __syncthreads();
}

// The value of v at lane `from` of the wavefront (two ds_bpermute_b32, which
// move data through the LDS crossbar without using LDS memory).  Every lane
// must be active.
__device__ __forceinline__ double pull(double v, int from) {
  const int a = from << 2;
  const int lo = __builtin_amdgcn_ds_bpermute(a, __double2loint(v));
  const int hi = __builtin_amdgcn_ds_bpermute(a, __double2hiint(v));
  return __hiloint2double(hi, lo);
}

// The cells of the wavefront where pred holds, packed into the lanes 0 .. n-1
// of one slot (when n <= 64), in cell order.  A branch that only those cells
// take then runs once instead of once per slot that holds one of them.
template <int S>
struct Packed {
  int n;       // cells packed (the same in every lane)
  int src;     // packed lane j < n: its cell, 64 s + l
  int pos[S];  // this lane's cell of slot s: its packed lane (where pred holds)
};
template <int S>
__device__ __forceinline__ Packed<S> pack_cells(const bool (&pred)[S]) {
  Packed<S> P;
  const int lane = lane_id();
  uint64_t m[S];
  int pre[S], cnt[S], n = 0;
#pragma unroll
  for (int s = 0; s < S; ++s) {
    m[s] = __ballot(pred[s]);
    pre[s] = n;
    cnt[s] = __popcll(m[s]);
    n += cnt[s];
  }
  P.n = n;
  P.src = 0;
  if (n > 64) return P;
#pragma unroll
  for (int s = 0; s < S; ++s) {
    // slot s sends its cells to lanes pre .. pre+cnt-1 and, to make the
    // forward permute a bijection, the other lanes to the remaining ones
    const int below = bits_below(m[s]);
    const int p = pred[s] ? pre[s] + below : (pre[s] + cnt[s] + (lane - below)) & 63;
    P.pos[s] = p;
    const int t = __builtin_amdgcn_ds_permute(p << 2, 64 * s + lane);
    if (lane >= pre[s] && lane < pre[s] + cnt[s]) P.src = t;
  }
  return P;
}
// the value of the per-slot v at cell `src` (every lane active)
template <int S>
__device__ __forceinline__ double gather_packed(const double (&v)[S], int src) {
  const int sl = src >> 6, l = src & 63;
  double r = pull(v[0], l);
#pragma unroll
  for (int s = 1; s < S; ++s) {
    const double t = pull(v[s], l);
    r = sl == s ? t : r;
  }
  return r;
}
// v[s] = the packed value of this lane's cell, where pred holds
template <int S>
__device__ __forceinline__ void scatter_packed(double (&v)[S], double pv, const Packed<S>& P,
                                               const bool (&pred)[S]) {
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const double t = pull(pv, P.pos[s]);
    if (pred[s]) v[s] = t;
  }
}

// Per-slot values through LDS: put_cells writes them in cell order, get_cells
// reads cell from[s] into slot s (after a workgroup synchronisation).
template <int S, class T>
__device__ __forceinline__ void put_cells(const T (&v)[S], T* buf) {
  const int lane = lane_id();
#pragma unroll
  for (int s = 0; s < S; ++s) buf[64 * s + lane] = v[s];
}
template <int S, class T>
__device__ __forceinline__ void get_cells(T (&v)[S], const T* buf, const int (&from)[S]) {
#pragma unroll
  for (int s = 0; s < S; ++s) v[s] = buf[from[s]];
}

__device__ __forceinline__ double readlane(double v, int l) {
  int lo = __double2loint(v), hi = __double2hiint(v);
  lo = __builtin_amdgcn_readlane(lo, l);
  hi = __builtin_amdgcn_readlane(hi, l);
  return __hiloint2double(hi, lo);
}

template <int S>
__device__ __forceinline__ bool wave_any(const bool (&p)[S]) {
  bool a = false;
#pragma unroll
  for (int s = 0; s < S; ++s) a = a || (__ballot(p[s]) != 0ull);
  return a;
}

// out = "some cell of my column satisfies pred" (pred must be false on
// cells past the last column).  FORWARD+BACKWARD no_fall pair of
// microphysics.py:349-368 and its three analogues.
template <int S>
__device__ __forceinline__ void col_any(const Lay<S>& L, const bool (&pred)[S], bool (&out)[S]) {
  uint64_t m[S], res[S];
#pragma unroll
  for (int s = 0; s < S; ++s) {
    m[s] = __ballot(pred[s]);
    res[s] = 0ull;
  }
  for (int c = 0; c < L.ncols; ++c) {
    const int lo = c * L.nz, hi = lo + L.nz;
    uint64_t hit = 0ull;
#pragma unroll
    for (int s = 0; s < S; ++s) hit |= m[s] & range_mask(lo - 64 * s, hi - 64 * s);
    if (hit) {
#pragma unroll
      for (int s = 0; s < S; ++s) res[s] |= range_mask(lo - 64 * s, hi - 64 * s);
    }
  }
  const int lane = lane_id();
#pragma unroll
  for (int s = 0; s < S; ++s) out[s] = ((res[s] >> lane) & 1ull) != 0ull;
}

// out = "k is at or below the first level (from the top) where pred holds,
// or k is the surface layer": the stop_k sweep of microphysics.py:706-724.
template <int S>
__device__ __forceinline__ void col_from_first(const Lay<S>& L, const bool (&pred)[S], bool (&out)[S]) {
  uint64_t m[S], res[S];
#pragma unroll
  for (int s = 0; s < S; ++s) {
    m[s] = __ballot(pred[s]);
    res[s] = 0ull;
  }
  for (int c = 0; c < L.ncols; ++c) {
    const int lo = c * L.nz, hi = lo + L.nz;
    int first = hi - 1;
    bool found = false;
#pragma unroll
    for (int s = 0; s < S; ++s) {
      const uint64_t h = m[s] & range_mask(lo - 64 * s, hi - 64 * s);
      if (!found && h) {
        first = 64 * s + __builtin_ctzll(h);
        found = true;
      }
    }
#pragma unroll
    for (int s = 0; s < S; ++s) res[s] |= range_mask(first - 64 * s, hi - 64 * s);
  }
  const int lane = lane_id();
#pragma unroll
  for (int s = 0; s < S; ++s) out[s] = ((res[s] >> lane) & 1ull) != 0ull;
}

// ---------------------------------------------------------- LDS round trips
// value of the cell above (k-1); undefined at k = 0
template <int S, class T>
__device__ __forceinline__ void col_prev(T* buf, const T (&x)[S], T (&out)[S]) {
  const int lane = lane_id();
#pragma unroll
  for (int s = 0; s < S; ++s) buf[64 * s + lane] = x[s];
  __syncthreads();
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const int j = 64 * s + lane - 1;
    out[s] = buf[j < 0 ? 0 : j];
  }
  __syncthreads();
}
// value of the cell below (k+1); undefined at k = nz-1
template <int S, class T>
__device__ __forceinline__ void col_next(T* buf, const T (&x)[S], T (&out)[S]) {
  const int lane = lane_id();
#pragma unroll
  for (int s = 0; s < S; ++s) buf[64 * s + lane] = x[s];
  __syncthreads();
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const int j = 64 * s + lane + 1;
    out[s] = buf[j > 64 * S - 1 ? 64 * S - 1 : j];
  }
  __syncthreads();
}
// value of the surface layer (k = nz-1) of the cell's column
template <int S, class T>
__device__ __forceinline__ void col_bottom(const Lay<S>& L, T* buf, const T (&x)[S], T (&out)[S]) {
  const int lane = lane_id();
#pragma unroll
  for (int s = 0; s < S; ++s) buf[64 * s + lane] = x[s];
  __syncthreads();
#pragma unroll
  for (int s = 0; s < S; ++s) out[s] = buf[L.cl[s] * L.nz + L.nz - 1];
  __syncthreads();
}

// ---------------------------------------------------------------- scans
// One DPP move of a double (two 32-bit halves).  Lanes whose source lane is
// outside its row read 0 and lanes of rows outside ROW_MASK are not written;
// the scan never uses either.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_move(double src) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(src), CTRL, ROW_MASK, 0xF, true);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(src), CTRL, ROW_MASK, 0xF, true);
  return __hiloint2double(hi, lo);
}
// DPP controls (gfx9): row_shr:d shifts lanes up by d inside each 16-lane
// row; row_bcast:15 / :31 hand lane 15 of each row (lane 31) to later rows.
constexpr int DPP_ROW_SHR = 0x110, DPP_ROW_BCAST15 = 0x142, DPP_ROW_BCAST31 = 0x143;

// One step of the segmented affine scan: compose (A, B) with the map of
// source lane j, fetched by DPP, where `combine` holds: the source lane is
// valid for the move and the lane's column reaches back to it.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ void scan_step(double& A, double& B, bool combine) {
  const double Ap = dpp_move<CTRL, ROW_MASK>(A);
  const double Bp = dpp_move<CTRL, ROW_MASK>(B);
  if (combine) {
    B = A * Bp + B;
    A = A * Ap;
  }
}

// x[k] = a[k] x[k-1] + b[k] for k > 0, x[0] = b[0]   (top -> surface).
// Within a slot: an inclusive scan of the affine maps x -> a x + b with DPP
// (row shifts 1, 2, 4, 8, then the row broadcasts), segmented by column: the
// lane at level k combines with source lane j only if k >= lane - j.  Across
// slots: the last lane's value of slot s-1 carries into slot s.
template <int S>
__device__ __forceinline__ void scan_fwd(const Lay<S>& L, const double (&a)[S], const double (&b)[S],
                                         double (&x)[S]) {
  const int lane = lane_id();
  const int in_row = lane & 15;
  double carry = 0.0;
#pragma unroll
  for (int s = 0; s < S; ++s) {
    double A = a[s], B = b[s];
    const int k = L.k[s];
    scan_step<DPP_ROW_SHR + 1, 0xF>(A, B, in_row >= 1 && k >= 1);
    scan_step<DPP_ROW_SHR + 2, 0xF>(A, B, in_row >= 2 && k >= 2);
    scan_step<DPP_ROW_SHR + 4, 0xF>(A, B, in_row >= 4 && k >= 4);
    scan_step<DPP_ROW_SHR + 8, 0xF>(A, B, in_row >= 8 && k >= 8);
    // rows 1 and 3 take lane 15 / 47, then rows 2 and 3 take lane 31
    scan_step<DPP_ROW_BCAST15, 0xA>(A, B, (lane & 16) && k >= in_row + 1);
    scan_step<DPP_ROW_BCAST31, 0xC>(A, B, lane >= 32 && k >= lane - 31);
    x[s] = (k <= lane) ? B : A * carry + B;
    carry = readlane(x[s], 63);
  }
}

// Two scans with the same maps' multipliers a (the momentum transport's u and
// v): each output gets exactly the operations of scan_fwd; the multipliers'
// composition and its DPP moves are shared.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ void scan_step2(double& A, double& B, double& C, bool combine) {
  const double Ap = dpp_move<CTRL, ROW_MASK>(A);
  const double Bp = dpp_move<CTRL, ROW_MASK>(B);
  const double Cp = dpp_move<CTRL, ROW_MASK>(C);
  if (combine) {
    B = A * Bp + B;
    C = A * Cp + C;
    A = A * Ap;
  }
}
template <int S>
__device__ __forceinline__ void scan_fwd2(const Lay<S>& L, const double (&a)[S], const double (&b)[S],
                                          const double (&c)[S], double (&x)[S], double (&y)[S]) {
  const int lane = lane_id();
  const int in_row = lane & 15;
  double carry_x = 0.0, carry_y = 0.0;
#pragma unroll
  for (int s = 0; s < S; ++s) {
    double A = a[s], B = b[s], C = c[s];
    const int k = L.k[s];
    scan_step2<DPP_ROW_SHR + 1, 0xF>(A, B, C, in_row >= 1 && k >= 1);
    scan_step2<DPP_ROW_SHR + 2, 0xF>(A, B, C, in_row >= 2 && k >= 2);
    scan_step2<DPP_ROW_SHR + 4, 0xF>(A, B, C, in_row >= 4 && k >= 4);
    scan_step2<DPP_ROW_SHR + 8, 0xF>(A, B, C, in_row >= 8 && k >= 8);
    scan_step2<DPP_ROW_BCAST15, 0xA>(A, B, C, (lane & 16) && k >= in_row + 1);
    scan_step2<DPP_ROW_BCAST31, 0xC>(A, B, C, lane >= 32 && k >= lane - 31);
    x[s] = (k <= lane) ? B : A * carry_x + B;
    y[s] = (k <= lane) ? C : A * carry_y + C;
    carry_x = readlane(x[s], 63);
    carry_y = readlane(y[s], 63);
  }
}

}  // namespace shield
