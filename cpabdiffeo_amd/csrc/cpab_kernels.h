// CDNA4 kernels of the CPA integration path.
//
//  * k_prep        sign * As, dt scaling and the per-cell matrix exponential
//                  in one launch (cpab/cpaNd/_CpaCalcs.py:243-313 +
//                  cpab/gpu/expm/expm_affine_by_series_{2D,3D}.py).
//  * k_integrate   calc_T / calc_trajectory (cpab/gpu/calc_2Dto2D.cu:348-408,
//                  499-556).  One lane owns one point at a time; the per-cell
//                  tables live in LDS.  The reference runs every step as
//                  "Trel proposal, and on a cell change nStepsODE RK2
//                  substeps", which in SIMT form idles 63 lanes of a wave
//                  while one lane does its ten substeps.  Here every loop trip
//                  of a lane is ONE unit of that schedule (a proposal or a
//                  substep), each with exactly one cell lookup, so lanes that
//                  cross a facet drift out of step with their neighbours
//                  instead of stalling them, and a lane that finishes its
//                  point takes the next point of its wave's range.  Per
//                  point, the sequence of IEEE operations is the reference's.
//  * k_simple      calc_T_simple (calc_2Dto2D.cu:412-445).
//  * k_velocity    calc_v (calc_2Dto2D.cu:561-602).
//  * k_cell_idx    calc_cell_idx (calc_2Dto2D.cu:323-344).
//  * k_remap2d / k_remap3d / k_ssd_partial / k_reduce_ll: remap_fwd and the
//                  Gaussian SSD likelihood in FP32 with an FP64 final sum.
#pragma once
#include "cpab_device.h"

#pragma clang fp contract(off)

namespace cpab {

// ------------------------------------------------------------------ expm
// expm_affine_by_series_2D.py:28-74
__device__ inline void expm_2x2(double a, double b, double c, double d, double* E) {
  const double delta_tmp = (a - d) * (a - d) + 4 * b * c;
  const double e = exp((a + d) / 2);
  if (delta_tmp == 0) {
    E[0] = (1 + (a - d) / 2) * e;
    E[1] = b * e;
    E[2] = c * e;
    E[3] = (1 - (a - d) / 2) * e;
  } else if (delta_tmp > 0) {
    const double delta = sqrt(delta_tmp) / 2;
    const double ch = cosh(delta), s = sinh(delta) / delta;
    E[0] = (ch + (a - d) / 2 * s) * e;
    E[1] = b * s * e;
    E[2] = c * s * e;
    E[3] = (ch - (a - d) / 2 * s) * e;
  } else {
    const double delta = sqrt(-delta_tmp) / 2;
    const double co = cos(delta), s = sin(delta) / delta;
    E[0] = (co + (a - d) / 2 * s) * e;
    E[1] = b * s * e;
    E[2] = c * s * e;
    E[3] = (co - (a - d) / 2 * s) * e;
  }
}

// expm_affine_by_series_2D.py:85-237 (closed form when the linear part is
// invertible, else the p-term Taylor series of the 2x3 block)
__device__ inline void expm_affine2(const double* A, double* T, int p) {
  const double A00 = A[0], A01 = A[1], A02 = A[2], A10 = A[3], A11 = A[4], A12 = A[5];
  const double det = A00 * A11 - A01 * A10;
  if (det != 0) {
    double E[4];
    expm_2x2(A00, A01, A10, A11, E);
    T[0] = E[0];
    T[1] = E[1];
    T[3] = E[2];
    T[4] = E[3];
    const double i00 = A11 / det, i11 = A00 / det, i01 = -A01 / det, i10 = -A10 / det;
    const double a = i00 * A02 + i01 * A12;
    const double b = i10 * A02 + i11 * A12;
    T[2] = (T[0] - 1) * a + T[1] * b;
    T[5] = T[3] * a + (T[4] - 1) * b;
    return;
  }
  double S[6] = {1, 0, 0, 0, 1, 0};  // identity, then + A^1 + A^2/2! + ...
  double An[6] = {A00, A01, A02, A10, A11, A12};
  for (int k = 0; k < 6; k++) S[k] += An[k];
  int fact = 1;
  for (int i = 2; i <= p; i++) {
    const double B0 = An[0], B1 = An[1], B2 = An[2], B3 = An[3], B4 = An[4], B5 = An[5];
    fact *= i;
    An[0] = A00 * B0 + A01 * B3;
    An[1] = A00 * B1 + A01 * B4;
    An[2] = A00 * B2 + A01 * B5;
    An[3] = A10 * B0 + A11 * B3;
    An[4] = A10 * B1 + A11 * B4;
    An[5] = A10 * B2 + A11 * B5;
    for (int k = 0; k < 6; k++) S[k] += An[k] / fact;
  }
  for (int k = 0; k < 6; k++) T[k] = S[k];
}

// expm_affine_by_series_3D.py:27-219 (series on the 3x4 block)
__device__ inline void expm_affine3(const double* A, double* T, int p) {
  double S[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
  double An[12];
  for (int k = 0; k < 12; k++) An[k] = A[k];
  for (int k = 0; k < 12; k++) S[k] += An[k];
  int fact = 1;
  for (int i = 2; i <= p; i++) {
    double B[12];
    for (int k = 0; k < 12; k++) B[k] = An[k];
    fact *= i;
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
      for (int col = 0; col < 4; col++)
        An[r * 4 + col] = A[r * 4 + 0] * B[col] + A[r * 4 + 1] * B[4 + col] + A[r * 4 + 2] * B[8 + col];
    for (int k = 0; k < 12; k++) S[k] += An[k] / fact;
  }
  for (int k = 0; k < 12; k++) T[k] = S[k];
}

// Fused prep of cpab/cpaNd/_CpaCalcs.py:243-313: signedAs = sign*As (the
// reference multiplies by -1 only for the inverse), As_dt = signedAs*dt,
// Trels = expm(As_dt); in 1D the closed form of :285-292.  One lane per
// cell; `inst` strides over independent instances (batched callers).
template <int DIM>
__device__ __forceinline__ void prep_cell(const double* __restrict__ A, int sign, double dt, double* sA,
                                          double* T, int p) {
  constexpr int NE = DIM * (DIM + 1);
  double sAdt[NE];
#pragma unroll
  for (int k = 0; k < NE; k++) {
    double v = A[k];
    if (sign == -1) v *= -1;
    sA[k] = v;
    sAdt[k] = v * dt;
  }
  if (DIM == 1) {
    T[0] = exp(sAdt[0]);
    T[1] = sAdt[1];
    if (sA[0] != 0) T[1] = T[1] * (T[0] - 1) / sAdt[0];
  } else if (DIM == 2) {
    expm_affine2(sAdt, T, p);
  } else {
    expm_affine3(sAdt, T, p);
  }
}

template <int DIM>
__global__ void k_prep(const double* __restrict__ As, int nC, int sign, double dt,
                       double* __restrict__ sA_out, double* __restrict__ T_out, int p,
                       int64_t inst_stride) {
  constexpr int NE = DIM * (DIM + 1);
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nC) return;
  const int64_t off = (int64_t)blockIdx.y * inst_stride + (int64_t)c * NE;
  double sA[NE], T[NE];
  prep_cell<DIM>(As + off, sign, dt, sA, T, p);
#pragma unroll
  for (int k = 0; k < NE; k++) {
    if (sA_out) sA_out[off + k] = sA[k];
    T_out[off + k] = T[k];
  }
}

// Avees = B theta for m thetas (CpaSpace.theta2Avees, cpab/cpaNd/_CpaSpace.py:366-376):
// one lane per (theta, row), the row's dot product summed in index order, so
// a theta gives the same Avees whatever else is in the batch.  The products'
// operands are loaded eight at a time ahead of the dependent sum (one load
// at a time left the call latency bound: ~30 us for a 3840-row table).
__global__ void k_theta2avees(const double* __restrict__ B, const double* __restrict__ theta,
                              double* __restrict__ Av, int D, int d, int m) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)m * D) return;
  const int64_t i = idx / D;
  const int r = (int)(idx - i * D);
  const double* row = B + (int64_t)r * d;
  const double* th = theta + i * d;
  double acc = 0.0;
  int j = 0;
  for (; j + 8 <= d; j += 8) {
    double a[8], t[8];
#pragma unroll
    for (int u = 0; u < 8; u++) {
      a[u] = row[j + u];
      t[u] = th[j + u];
    }
#pragma unroll
    for (int u = 0; u < 8; u++) acc += a[u] * t[u];
  }
  for (; j < d; j++) acc += row[j] * th[j];
  Av[idx] = acc;
}

// The reference's gpu_expm(As, Ts, p) on an already scaled table.
template <int DIM>
__global__ void k_expm(const double* __restrict__ As, double* __restrict__ Ts, int nC, int p) {
  constexpr int NE = DIM * (DIM + 1);
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nC) return;
  double A[NE], T[NE];
#pragma unroll
  for (int k = 0; k < NE; k++) A[k] = As[(int64_t)c * NE + k];
  if (DIM == 2) expm_affine2(A, T, p);
  else expm_affine3(A, T, p);
#pragma unroll
  for (int k = 0; k < NE; k++) Ts[(int64_t)c * NE + k] = T[k];
}

// ------------------------------------------------------------- helpers
__device__ __forceinline__ int lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0)); }

__device__ __forceinline__ uint64_t ballot(bool b) { return __ballot(b); }

__device__ __forceinline__ int rank_in(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0));
}

template <int DIM>
__device__ __forceinline__ void load_pt(const double* __restrict__ src, int64_t i, double* p) {
  if (DIM == 2) {
    const double2 v = *reinterpret_cast<const double2*>(src + i * 2);
    p[0] = v.x;
    p[1] = v.y;
  } else {
#pragma unroll
    for (int d = 0; d < DIM; d++) p[d] = src[i * DIM + d];
  }
}

template <int DIM>
__device__ __forceinline__ void store_pt(double* __restrict__ dst, int64_t i, const double* p) {
  if (DIM == 2) {
    *reinterpret_cast<double2*>(dst + i * 2) = make_double2(p[0], p[1]);
  } else {
#pragma unroll
    for (int d = 0; d < DIM; d++) dst[i * DIM + d] = p[d];
  }
}

// Copy the Trel and A tables into LDS (both when they fit).
template <int NE>
__device__ __forceinline__ void stage_tables(double* lds, const double* __restrict__ Trels,
                                             const double* __restrict__ As, int nC) {
  const int n = nC * NE;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    lds[i] = Trels ? Trels[i] : 0.0;
    lds[n + i] = As[i];
  }
  __syncthreads();
}

// The per-cell prep fused into a launch (cpab_calc_T_prep): the raw Avees
// table with k_prep's sign and dt.  Every block forms the tables itself
// (prep_cell, the same operations as k_prep) into LDS; block 0 also writes
// them out as cpab_prep_trels would.  raw == nullptr: the tables come ready.
struct PrepSrc {
  const double* raw;
  double* sA_out;
  double* T_out;
  double dt;
  int sign;
};

template <int DIM>
__device__ __forceinline__ void stage_tables_prep(double* lds, const PrepSrc& ps, int nC) {
  constexpr int NE = DIM * (DIM + 1);
  const int n = nC * NE;
  for (int c = threadIdx.x; c < nC; c += blockDim.x) {
    double sA[NE], T[NE];
    prep_cell<DIM>(ps.raw + (int64_t)c * NE, ps.sign, ps.dt, sA, T, 12);
#pragma unroll
    for (int e = 0; e < NE; e++) {
      lds[c * NE + e] = T[e];
      lds[n + c * NE + e] = sA[e];
    }
    if (blockIdx.x == 0) {
#pragma unroll
      for (int e = 0; e < NE; e++) {
        ps.T_out[(int64_t)c * NE + e] = T[e];
        ps.sA_out[(int64_t)c * NE + e] = sA[e];
      }
    }
  }
  __syncthreads();
}

// ----------------------------------------------------------------- trip
// One unit of the reference's step schedule for one point (see k_integrate):
//   k == 0: y = Trel_c [p;1]; if cell(y) == c accept, else start the RK2 redo
//           (k = nStepsODE; p and c unchanged)   (calc_2Dto2D.cu:389-401)
//   k >  0: one RK2 midpoint substep of A_c from p (solveODE, :179-194);
//           p = y, c = cell(y), k--.
// Both kinds of trip share the table read, the first affine product and the
// cell lookup, so lanes in different modes diverge only over the midpoint
// arithmetic.  (Measured on gfx950: evaluating the midpoint in every lane
// and selecting, or chaining several Trel proposals speculatively so their
// lookups overlap, are both slower than this.)
// Returns true when the step is complete (the caller advances t).
template <int DIM, int TESS, bool POW2, bool FAST = false, bool EARLY = true>
__device__ __forceinline__ bool trip(double* p, int& c, int& k, const double*& row, int& nfb, const double* Ttab,
                                     const double* Atab, double h, int nS, const Geom& g) {
  constexpr int NE = DIM * (DIM + 1);
  const bool fast = (k == 0);
  double y[DIM];
  affine<DIM>(y, row, p);
  if (!fast) {  // RK2 midpoint substep (solveODE, calc_2Dto2D.cu:182-194)
    // The reference's h * v / 2 is fl(h v) halved; (h / 2) * v rounds the
    // same real number h v / 2, and halving commutes with rounding, so the
    // two agree bit for bit unless |h v / 2| is subnormal (below 2^-1022).
    // A nonzero v = A [p; 1] is at least ~2^-53 of its largest term, so that
    // takes table entries near 1e-290.  One multiply per coordinate fewer.
    const double hh = h * 0.5;
    double m[DIM];
#pragma unroll
    for (int d = 0; d < DIM; d++) m[d] = p[d] + hh * y[d];
    affine<DIM>(y, row, m);
#pragma unroll
    for (int d = 0; d < DIM; d++) y[d] = p[d] + y[d] * h;
  }
  const int cy = cell_of<DIM, TESS, POW2, FAST, EARLY>(y, g);
  // (Measured on gfx950: the state update below written as selects instead
  // of branches ran 3-4 % slower here and 11 % slower for 256 registration
  // chains.)
  if (!fast) {
#pragma unroll
    for (int d = 0; d < DIM; d++) p[d] = y[d];
    const bool done = --k == 0;
    row = (done ? Ttab : Atab) + mul24(cy, NE);  // (the same row when the cell did not change)
    c = cy;
    return done;
  }
  if (cy == c) {
#pragma unroll
    for (int d = 0; d < DIM; d++) p[d] = y[d];
    return true;
  }
  nfb++;
  if (nS > 0) {
    k = nS;
    row = Atab + mul24(c, NE);
    return false;
  }
  return true;  // zero substeps: the step leaves p where it was
}

// ------------------------------------------------------------ lone lane
// A wave left with a single unfinished lane runs one dependent chain of
// instructions, so its trips cost their latency (a dependent FP64 operation
// takes 16-22 cycles on MI355X), not their issue.  The point that gets there
// is the domain corner (0, 0): its velocity is zero up to rounding, it
// drifts just outside the domain, and the Trel proposal leaves the cell at
// most steps, so the lane runs up to 1 100 trips while its neighbours need
// ~104.  (In the registration step, where only its FP32 sample matters, a
// bound on its path usually fixes that sample and the pixel is left out:
// corner_sample_is_src00.)  Here the lane loops on its own: its branches are
// uniform (one lane), each trip does only its own kind of work, and the lookup is branch-free
// (cell2_flat, no strict-interior test).  Measured on the one-chain
// registration: ~770 instead of ~940 cycles a trip, +5 % at 40 slots, +11 %
// at 24.  (Holding the table row in registers took the step kernel past 64
// VGPRs; reading the row with scalar loads from a global copy of the tables,
// the cell being wave-uniform here, measured 1-2 % slower than LDS.)  Same
// schedule and IEEE operations as `trip`.  Runs until t >= tb and leaves row
// consistent with (c, k).
template <int DIM, int TESS, bool POW2>
__device__ __forceinline__ int lone_cell(const double* y, const Geom& g) {
  if (DIM == 2 && TESS == TESS_I) return cell2_flat<POW2>(y, g);
  return cell_of<DIM, TESS, POW2>(y, g);
}

template <int DIM, int TESS, bool POW2>
__device__ __forceinline__ void lone_trips(double* p, int& c, int& k, int& t, int tb, const double*& row, int& nfb,
                                           const double* Ttab, const double* Atab, double h, int nS,
                                           const Geom& g) {
  constexpr int NE = DIM * (DIM + 1);
  const double hh = h * 0.5;  // as in `trip`
  while (t < tb) {
    const double* r = row;
    double y[DIM];
    affine<DIM>(y, r, p);
    if (k == 0) {  // Trel proposal
      const int cy = lone_cell<DIM, TESS, POW2>(y, g);
      if (cy == c) {
#pragma unroll
        for (int d = 0; d < DIM; d++) p[d] = y[d];
        t++;
      } else {
        nfb++;
        if (nS > 0) {
          k = nS;
          row = Atab + mul24(c, NE);
        } else {
          t++;
        }
      }
    } else {  // one RK2 midpoint substep
      double m[DIM];
#pragma unroll
      for (int d = 0; d < DIM; d++) m[d] = p[d] + hh * y[d];
      affine<DIM>(y, r, m);
#pragma unroll
      for (int d = 0; d < DIM; d++) y[d] = p[d] + y[d] * h;
      const int cy = lone_cell<DIM, TESS, POW2>(y, g);
#pragma unroll
      for (int d = 0; d < DIM; d++) p[d] = y[d];
      const bool done = --k == 0;
      if (done | (cy != c)) row = (done ? Ttab : Atab) + mul24(cy, NE);
      c = cy;
      if (done) t++;
    }
  }
}

// --------------------------------------------------------- integration
// MODE 0: calc_T (write the end point).  MODE 1: calc_trajectory (write the
// point after every step into traj[(t*nPts + i)*DIM]).
//
// Lane state: p (current point), c (cell of p), t (steps done),
// k (RK2 substeps still to run for the current step; 0 = next trip is a
// Trel proposal); one loop trip is one `trip`.  solveODE recomputes cell(p)
// at the top of every substep; here it is the lookup at the end of the
// previous trip, of the same p: same value.
//
// Trips run in an inner loop until some lane finishes its point; only then
// do lanes retire and refill from the wave's range (the bookkeeping costs
// ballots and branches, so it stays out of the per-trip path).
// Processing index i -> point index when the points form a row-major grid
// of row length w (a multiple of 8): the first nTiled = 8 w floor(nPts / 8 w)
// indices walk 8 x 8 tiles (64 consecutive indices one tile, the tiles of
// each band of 8 rows in turn), the rest map to themselves.  A bijection of
// [0, nPts) whatever w is.
__device__ __forceinline__ int64_t tile_order(int64_t i, int w, int64_t nTiled) {
  if (i >= nTiled) return i;
  const int64_t t = i >> 6;
  const int u = (int)(i & 63), tpr = w >> 3;
  const int64_t ty = t / tpr;
  const int tx = (int)(t - ty * tpr);
  return (ty * 8 + (u >> 3)) * (int64_t)w + tx * 8 + (u & 7);
}

// One wave integrates the points [begin, end) of the processing order
// (indices into pos0, out and nfb_out; with gw > 0 through tile_order, with
// lord through lbase + lord[i - lbase], a tile's order held in LDS).
// Wave-uniform arguments; every lane of the wave calls it.
template <int DIM, int TESS, bool POW2, int MODE, bool FAST = false>
__device__ __forceinline__ void integrate_range(int64_t begin, int64_t end, const double* __restrict__ pos0,
                                                double* __restrict__ out, const double* Ttab, const double* Atab,
                                                double h, int nT, int nS, int64_t nPts, const Geom& g,
                                                int32_t* __restrict__ nfb_out, int gw = 0,
                                                const uint16_t* lord = nullptr, int64_t lbase = 0) {
  const int64_t nTiled = gw > 0 ? nPts / (8 * (int64_t)gw) * (8 * (int64_t)gw) : 0;
  constexpr int NE = DIM * (DIM + 1);
  int64_t next = begin;  // wave-uniform cursor into [begin, end)
  double p[DIM];
  int c = 0, t = 0, k = 0, nfb = 0;
  const double* row = Ttab;  // the table row the next trip reads
  int64_t my = 0;
  bool active = false;
  for (;;) {
    if (active && t >= nT) {  // retire a finished point
      if (MODE == 0) store_pt<DIM>(out, my, p);
      if (nfb_out) nfb_out[my] = nfb;
      active = false;
    }
    const uint64_t idle = ballot(!active);
    if (idle != 0 && next < end) {  // refill idle lanes from the wave's range
      if (!active) {
        const int64_t cand = next + rank_in(idle);
        if (cand < end) {
          my = gw > 0 ? tile_order(cand, gw, nTiled) : (lord ? lbase + lord[cand - lbase] : cand);
          active = true;
          load_pt<DIM>(pos0, my, p);
          c = cell_of<DIM, TESS, POW2>(p, g);
          row = Ttab + mul24(c, NE);
          t = 0;
          k = 0;
          nfb = 0;
        }
      }
      next += __popcll(idle);
    }
    if (ballot(active) == 0) break;
    if (!active || t >= nT) continue;
    if (MODE == 0 && DIM == 2 && __popcll(__builtin_amdgcn_read_exec()) == 1) {  // one lane left: its own loop
      lone_trips<DIM, TESS, POW2>(p, c, k, t, nT, row, nfb, Ttab, Atab, h, nS, g);
      continue;
    }
    // Trips until some lane of the region finishes its point, two per loop
    // iteration with the exit test between them (written out: the compiler
    // does not unroll a loop around a ballot), so the point's registers
    // alternate instead of being copied back at the end of every trip.
    for (;;) {
      if (trip<DIM, TESS, POW2, FAST>(p, c, k, row, nfb, Ttab, Atab, h, nS, g)) {
        if (MODE == 1) store_pt<DIM>(out, (int64_t)t * nPts + my, p);
        t++;
      }
      if (ballot(t >= nT) != 0) break;
      if (trip<DIM, TESS, POW2, FAST>(p, c, k, row, nfb, Ttab, Atab, h, nS, g)) {
        if (MODE == 1) store_pt<DIM>(out, (int64_t)t * nPts + my, p);
        t++;
      }
      if (ballot(t >= nT) != 0) break;
    }
  }
}

template <int DIM, int TESS, bool POW2, bool LDS, int MODE, bool FAST = false>
__global__ void __launch_bounds__(1024)
    k_integrate(const double* __restrict__ pos0, double* __restrict__ out,
                const double* __restrict__ Trels, const double* __restrict__ As, double h,
                int nT, int nS, int64_t nPts, Geom g, int nC, int ptsPerWave,
                int32_t* __restrict__ nfb_out, PrepSrc ps) {
  constexpr int NE = DIM * (DIM + 1);
  extern __shared__ double lds[];
  // calc_T of 2D points: when pos0 looks like a row-major grid, its row
  // length W (the first index, a multiple of 8 up to 8 x blockDim.x, where y
  // changes from point 0's), so that waves take 8 x 8 tiles: neighbouring
  // points flow alike and cross facets at about the same steps, so a wave's
  // lanes need similar numbers of trips.  Only the processing order depends
  // on it (every point's result is its own), so a wrong guess costs time
  // only.  The loads are issued before the tables are staged.
  __shared__ int grid_w;
  constexpr bool kDetect = MODE == 0 && DIM == 2;
  double yc = 0.0, yp = 0.0, y0 = 0.0;
  const int64_t ic = 8 * ((int64_t)threadIdx.x + 1);
  if (kDetect) {
    if (threadIdx.x == 0) grid_w = 0x7fffffff;
    if (ic < nPts) {
      y0 = pos0[1];
      yc = pos0[ic * 2 + 1];
      yp = pos0[(ic - 1) * 2 + 1];
    }
    __syncthreads();
  }
  const double* Ttab = Trels;
  const double* Atab = As;
  if (LDS) {
    if (ps.raw) stage_tables_prep<DIM>(lds, ps, nC);  // (a kernel argument: uniform)
    else stage_tables<NE>(lds, Trels, As, nC);
    Ttab = lds;
    Atab = lds + nC * NE;
  }
  int gw = 0;
  if (kDetect) {
    if (ic < nPts && yc != y0 && yp == y0) atomicMin(&grid_w, (int)ic);
    __syncthreads();
    gw = grid_w != 0x7fffffff && 8 * (int64_t)grid_w <= nPts ? grid_w : 0;
  }
  const int64_t wave = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int64_t begin = wave * (int64_t)ptsPerWave;
  if (begin >= nPts) return;
  const int64_t end = begin + ptsPerWave < nPts ? begin + ptsPerWave : nPts;
  integrate_range<DIM, TESS, POW2, MODE, FAST>(begin, end, pos0, out, Ttab, Atab, h, nT, nS, nPts, g, nfb_out, gw);
}

// ---------------------------------------------------------- sensitivity
// calc_grad_theta (cpab/gpu/calc_2Dto2D.cu:202-319, 448-491): calc_T_simple
// together with the forward sensitivity q_j = dp/dtheta_j of every basis
// column j, integrated by the same RK2 (q' = B_j [p;1] + A_lin q).  Grid
// (point blocks, column chunks): a thread owns one point and JC columns,
// whose basis tables (layout (d, nC, DIM, DIM+1)) sit in LDS with As; the
// reference keeps q in global memory and re-reads it every substep, here it
// stays in registers (same values).  Output grad (nPts, DIM, d) and, from
// the first chunk, the end points.  JC = columns per thread (the widest
// whose tables fit in LDS); LDS = false reads the tables from global memory.
template <int DIM, int TESS, bool POW2, int JC, bool LDS>
__global__ void __launch_bounds__(256)
    k_grad_theta(const double* __restrict__ pos0, double* __restrict__ pos, const double* __restrict__ As,
                 const double* __restrict__ Bs, double* __restrict__ grad, int d, double h, int nSteps,
                 int64_t nPts, Geom g, int nC) {
  constexpr int NE = DIM * (DIM + 1);
  extern __shared__ double lds[];
  const int j0 = blockIdx.y * JC;
  const int nj = d - j0 < JC ? d - j0 : JC;
  const double* Al = As;
  const double* Bl = Bs + (int64_t)j0 * nC * NE;
  if (LDS) {
    double* a = lds;
    double* b = lds + nC * NE;
    for (int i = threadIdx.x; i < nC * NE; i += blockDim.x) a[i] = As[i];
    for (int i = threadIdx.x; i < nj * nC * NE; i += blockDim.x) b[i] = Bs[(int64_t)j0 * nC * NE + i];
    __syncthreads();
    Al = a;
    Bl = b;
  }
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nPts) return;
  double p[DIM], v[DIM], pMid[DIM], vMid[DIM];
  double q[JC][DIM];
#pragma unroll
  for (int jj = 0; jj < JC; jj++)
#pragma unroll
    for (int i = 0; i < DIM; i++) q[jj][i] = 0.0;
  load_pt<DIM>(pos0, idx, p);
  for (int s = 0; s < nSteps; s++) {
    const int c = cell_of<DIM, TESS, POW2>(p, g);
    const double* A = Al + c * NE;
    affine<DIM>(v, A, p);
#pragma unroll
    for (int i = 0; i < DIM; i++) pMid[i] = p[i] + h * v[i] / 2.;
    affine<DIM>(vMid, A, pMid);
#pragma unroll
    for (int jj = 0; jj < JC; jj++) {
      if (jj < nj) {
        const double* Bj = Bl + (jj * nC + c) * NE;
        double BtT[DIM], AdT[DIM], u[DIM], qMid[DIM];
        affine<DIM>(BtT, Bj, p);
        linear<DIM>(AdT, A, q[jj]);
#pragma unroll
        for (int i = 0; i < DIM; i++) u[i] = BtT[i] + AdT[i];
#pragma unroll
        for (int i = 0; i < DIM; i++) qMid[i] = q[jj][i] + h * u[i] / 2.;
        affine<DIM>(BtT, Bj, pMid);
        linear<DIM>(AdT, A, qMid);
#pragma unroll
        for (int i = 0; i < DIM; i++) u[i] = BtT[i] + AdT[i];
#pragma unroll
        for (int i = 0; i < DIM; i++) q[jj][i] += u[i] * h;
      }
    }
#pragma unroll
    for (int i = 0; i < DIM; i++) p[i] += vMid[i] * h;
  }
  if (blockIdx.y == 0) store_pt<DIM>(pos, idx, p);
#pragma unroll
  for (int jj = 0; jj < JC; jj++)
    if (jj < nj)
#pragma unroll
      for (int i = 0; i < DIM; i++) grad[(idx * DIM + i) * d + j0 + jj] = q[jj][i];
}

// calc_T_simple: nStepsODE * nTimeSteps RK2 substeps (calc_2Dto2D.cu:412-445).
template <int DIM, int TESS, bool POW2, bool LDS>
__global__ void __launch_bounds__(1024)
    k_simple(const double* __restrict__ pos0, double* __restrict__ out, const double* __restrict__ As,
             double h, int nSteps, int64_t nPts, Geom g, int nC) {
  constexpr int NE = DIM * (DIM + 1);
  extern __shared__ double lds[];
  const double* Atab = As;
  if (LDS) {
    for (int i = threadIdx.x; i < nC * NE; i += blockDim.x) lds[i] = As[i];
    __syncthreads();
    Atab = lds;
  }
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nPts) return;
  double p[DIM], y[DIM], m[DIM];
  load_pt<DIM>(pos0, i, p);
  for (int s = 0; s < nSteps; s++) {
    const double* A = Atab + cell_of<DIM, TESS, POW2>(p, g) * NE;
    affine<DIM>(y, A, p);
#pragma unroll
    for (int d = 0; d < DIM; d++) m[d] = p[d] + h * y[d] / 2.;
    affine<DIM>(y, A, m);
#pragma unroll
    for (int d = 0; d < DIM; d++) p[d] += y[d] * h;
  }
  store_pt<DIM>(out, i, p);
}

template <int DIM, int TESS, bool POW2, bool LDS>
__global__ void __launch_bounds__(1024)
    k_velocity(const double* __restrict__ pts, double* __restrict__ vel, const double* __restrict__ As,
               int64_t nPts, Geom g, int nC) {
  constexpr int NE = DIM * (DIM + 1);
  extern __shared__ double lds[];
  const double* Atab = As;
  if (LDS) {
    for (int i = threadIdx.x; i < nC * NE; i += blockDim.x) lds[i] = As[i];
    __syncthreads();
    Atab = lds;
  }
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nPts) return;
  double p[DIM], v[DIM];
  load_pt<DIM>(pts, i, p);
  affine<DIM>(v, Atab + cell_of<DIM, TESS, POW2>(p, g) * NE, p);
  store_pt<DIM>(vel, i, v);
}

template <int DIM, int TESS, bool POW2>
__global__ void k_cell_idx(const double* __restrict__ pts, int32_t* __restrict__ cell, int64_t nPts, Geom g) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nPts) return;
  double p[DIM];
  load_pt<DIM>(pts, i, p);
  cell[i] = cell_of<DIM, TESS, POW2>(p, g);
}

// ---------------------------------------------------- remap + likelihood
// Bilinear sample, zero outside (OpenCV INTER_LINEAR, BORDER_CONSTANT 0),
// weights from the double coordinates' fractional parts, float arithmetic.
__device__ __forceinline__ float sample2(const float* __restrict__ img, int H, int W, double x, double y) {
  const double xf = floor(x), yf = floor(y);
  // a full pixel or more outside, or not a number: zero
  if (!(xf >= -1.0 && xf < (double)W && yf >= -1.0 && yf < (double)H)) return 0.f;
  const float fx = (float)(x - xf), fy = (float)(y - yf);
  float v00 = 0.f, v01 = 0.f, v10 = 0.f, v11 = 0.f;
  const int x0 = (int)xf, y0 = (int)yf, x1 = x0 + 1, y1 = y0 + 1;
  const bool okx0 = x0 >= 0, okx1 = x1 < W, oky0 = y0 >= 0, oky1 = y1 < H;
  if (oky0 && okx0) v00 = img[(int64_t)y0 * W + x0];
  if (oky0 && okx1) v01 = img[(int64_t)y0 * W + x1];
  if (oky1 && okx0) v10 = img[(int64_t)y1 * W + x0];
  if (oky1 && okx1) v11 = img[(int64_t)y1 * W + x1];
  const float gx = 1.f - fx, gy = 1.f - fy;
  const float a = gx * v00 + fx * v01;
  const float b = gx * v10 + fx * v11;
  return gy * a + fy * b;
}

__device__ __forceinline__ float sample3(const float* __restrict__ vol, int D, int H, int W, double x,
                                         double y, double z) {
  const double xf = floor(x), yf = floor(y), zf = floor(z);
  if (!(xf >= -1.0 && xf < (double)W && yf >= -1.0 && yf < (double)H && zf >= -1.0 && zf < (double)D))
    return 0.f;
  const float fx = (float)(x - xf), fy = (float)(y - yf), fz = (float)(z - zf);
  float v[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  const int x0 = (int)xf, y0 = (int)yf, z0 = (int)zf;
#pragma unroll
  for (int dz = 0; dz < 2; dz++)
#pragma unroll
    for (int dy = 0; dy < 2; dy++)
#pragma unroll
      for (int dx = 0; dx < 2; dx++) {
        const int xx = x0 + dx, yy = y0 + dy, zz = z0 + dz;
        if (xx >= 0 && xx < W && yy >= 0 && yy < H && zz >= 0 && zz < D)
          v[dz * 4 + dy * 2 + dx] = vol[((int64_t)zz * H + yy) * W + xx];
      }
  const float gx = 1.f - fx, gy = 1.f - fy, gz = 1.f - fz;
  const float a0 = gx * v[0] + fx * v[1];
  const float a1 = gx * v[2] + fx * v[3];
  const float a2 = gx * v[4] + fx * v[5];
  const float a3 = gx * v[6] + fx * v[7];
  const float b0 = gy * a0 + fy * a1;
  const float b1 = gy * a2 + fy * a3;
  return gz * b0 + fz * b1;
}

__global__ void k_remap2d(const double* __restrict__ pts, const float* __restrict__ img,
                          float* __restrict__ out, int H, int W) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)H * W) return;
  const double2 q = *reinterpret_cast<const double2*>(pts + 2 * i);
  out[i] = sample2(img, H, W, q.x, q.y);
}

__global__ void k_remap3d(const double* __restrict__ pts, const float* __restrict__ vol,
                          float* __restrict__ out, int D, int H, int W) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)D * H * W) return;
  out[i] = sample3(vol, D, H, W, pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
}

// Block sum of a per-lane double through LDS; result valid in thread 0.
__device__ __forceinline__ double block_sum(double v, double* red) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < nw; i++) s += red[i];
  return s;
}

// Per-block partial sums of (x - y)^2 over float images; the difference and
// its square are taken in double (two FP64 ops per pixel, and the sum of
// 10^5 terms keeps ~1e-15 relative accuracy instead of ~1e-7).
__global__ void k_ssd_partial(const float* __restrict__ x, const float* __restrict__ y, int64_t n,
                              double* __restrict__ partial) {
  __shared__ double red[16];
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double e = (double)x[i] - (double)y[i];
    acc += e * e;
  }
  const double s = block_sum(acc, red);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// ll = -0.5 / sigma^2 * sum(partials), one wave (gaussian.py:30-34 + sum).
__global__ void k_reduce_ll(const double* __restrict__ partial, int n, double sigma, double* __restrict__ ll) {
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += 64) acc += partial[i];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
  if (threadIdx.x == 0) *ll = acc * (-0.5 / (sigma * sigma));
}

}  // namespace cpab
