// Device-side geometry, cell lookup and affine products for the CPA kernels.
//
// Rounding contract: every floating-point operation here is a single IEEE
// double operation in the order the reference writes it (no FMA
// contraction; the translation unit is built with -ffp-contract=off and the
// pragma below), so a lookup on the GPU returns the same cell as the
// reference's compute_cell_idx for every input (cpab/gpu/calc_2Dto2D.cu:38-153,
// calc_1Dto1D.cu:62-65, calc_3Dto3D.cu:43-137).
//
// Two forms of each lookup:
//   * LIT  : the reference expression restated literally (fmod, IEEE division).
//   * POW2 : used when every increment is a power of two.  Then x / inc and
//            x * (1/inc) are the same real number, hence round identically,
//            and fmod(x, inc) / inc == s - trunc(s) with s = x * (1/inc)
//            exactly, while (x - fmod(x, inc)) / inc == trunc(s) exactly
//            (|s| < 2^53; the clamps keep it so or make the difference
//            irrelevant).  The POW2 form therefore returns bit-identical cell
//            ids at a fraction of the cost: no divisions, no fmod loop.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#pragma clang fp contract(off)

namespace cpab {

constexpr int TESS_I = 1;   // triangles / tetrahedra (reference TESS_TYPE 1)
constexpr int TESS_II = 2;  // intervals / rectangles / boxes (TESS_TYPE 2)

// Everything the lookups need, precomputed on the host the way the
// reference kernels compute it inline.
struct Geom {
  int nC0, nC1, nC2, pad;
  double inc_x, inc_y, inc_z;
  double rinc_x, rinc_y, rinc_z;  // exact 1/inc (POW2 form only)
  double top_x, top_y, top_z;     // clamps nC0*inc_x-1e-10, nC1*inc_y-1e-10, nC2*inc_x-1e-10
  double wall_x, wall_y;          // nC0*inc_x, nC1*inc_y (2D type I out-of-domain tests)
};

__device__ __forceinline__ double dmin(double a, double b) { return __builtin_fmin(a, b); }

// Product of two non-negative ints below 2^24 (cell and box indices are far
// below that): the masks let the compiler use v_mul_u32_u24 (full rate)
// instead of v_mul_lo_u32.
__device__ __forceinline__ int mul24(int a, int b) {
  return (int)(((unsigned)a & 0xffffffu) * ((unsigned)b & 0xffffffu));
}
__device__ __forceinline__ double dmax(double a, double b) { return __builtin_fmax(a, b); }

// fmod(x, y) for y > 0, exact (the IEEE remainder with the sign of x).
// q = trunc(fl(x/y)) is the true quotient or one more (fl rounded up across
// an integer); x - q*y is then r or r - y, both representable, so the fma is
// exact, and the fix-up adds y back exactly.  |x/y| >= 2^53 only happens for
// unclamped type II / 1D coordinates whose cell is clamped anyway.
__device__ __forceinline__ double fmod_exact(double x, double y) {
  double q = __builtin_trunc(x / y);
  double r = __builtin_fma(-q, y, x);
  if (x >= 0.0) {
    if (r < 0.0) r = r + y;
  } else {
    if (r > 0.0) r = r - y;
  }
  return r;
}

// calc_2Dto2D.cu:20-22
__device__ __forceinline__ int mymin(int a, double b) { return !(b < a) ? a : (int)__builtin_round(b); }

// One axis of the type II / 1D lookup:
//   round(min(nC-1, max(0, (p - fmod(p, inc)) / inc)))
template <bool POW2>
__device__ __forceinline__ int axis_box(double p, int nC, double inc, double rinc) {
  double q;
  if (POW2) {
    // (p - fmod(p,inc))/inc == trunc(p/inc); fmod of +-inf or NaN is NaN,
    // which max(0, NaN) turns into 0.
    q = __builtin_isfinite(p) ? __builtin_trunc(p * rinc) : 0.0;
  } else {
    q = (p - fmod_exact(p, inc)) / inc;
  }
  return (int)__builtin_round(dmin((double)(nC - 1), dmax(0.0, q)));
}

// One clamped axis of the type I lookups: returns the box index (mymin of
// (p0 - xmod)/inc) and the in-box coordinate xmod/inc.
template <bool POW2>
__device__ __forceinline__ int axis_simplex(double p0, int nC, double inc, double rinc, double& frac) {
  if (POW2) {
    // p0 >= 0 (clamped), so s >= 0: s - trunc(s) is exact and equals
    // fract(s) = s - floor(s) (one v_fract_f64), and (int)s truncates.
    const double s = p0 * rinc;
    frac = __builtin_amdgcn_fract(s);
    const int i = (int)s;
    return i < nC - 1 ? i : nC - 1;
  } else {
    double m = fmod_exact(p0, inc);
    frac = m / inc;
    return mymin(nC - 1, (p0 - m) / inc);
  }
}

template <bool POW2>
__device__ __forceinline__ double divinc(double v, double inc, double rinc) {
  return POW2 ? v * rinc : v / inc;
}

// ------------------------------------------------------------------ 1D
// calc_1Dto1D.cu:62-65
template <bool POW2>
__device__ __forceinline__ int cell1(const double* p, const Geom& g) {
  return axis_box<POW2>(p[0], g.nC0, g.inc_x, g.rinc_x);
}

// ------------------------------------------------------------------ 2D
// calc_2Dto2D.cu:38-153
// FAST (POW2 type I only): when every lane of the wave is strictly inside
// (0, top) on both axes, min(top, max(0, p)) is p, trunc(p / inc) <= nC - 1
// and no outside rule applies, so the lookup is done without the clamps,
// the index mins and the min/max operand canonicalisations IEEE mode needs.
// The test is wave-uniform: a wave with a lane on or past the boundary (or
// NaN) runs the general lookup only.  Measured on MI355X the gain depends
// on the launch: 256 registration chains +14-17 %, one chain with deep
// speculation +1-2 %, the 512x512 calc_T +3 %, the 1 024-instance batch
// +1.5 %; but a 65536-point calc_T -28 % and the serving mix's work items
// -5 %.  So each launch chooses it.
template <int TESS, bool POW2, bool FAST = false, bool EARLY = true>
__device__ __forceinline__ int cell2(const double* p, const Geom& g) {
  if (TESS == TESS_II) {
    return axis_box<POW2>(p[0], g.nC0, g.inc_x, g.rinc_x) +
           mul24(axis_box<POW2>(p[1], g.nC1, g.inc_y, g.rinc_y), g.nC0);
  }
  const double px = p[0], py = p[1];
  if (POW2 && FAST) {
    // The interior form's cell.  EARLY: formed before the test that allows
    // it, so that the test's scalar chain and branch are off the cell's
    // critical path (a lane outside gets a value that is not used; measured
    // on MI355X: the 512 x 512 calc_T -1.5 %, one registration chain +1.2 %,
    // the batch -5.5 %, but 256 registration chains one step per launch 2 %
    // slower, so those launches form it after the test).
    const uint64_t ex = __builtin_amdgcn_read_exec();
    if (EARLY) {
      const double sx = px * g.rinc_x, sy = py * g.rinc_y;  // > 0 inside: fract as in axis_simplex
      const double x = __builtin_amdgcn_fract(sx), y = __builtin_amdgcn_fract(sy);
      const int c = ((int)sx + mul24((int)sy, g.nC0)) * 4;
      const double omx = 1.0 - x;
      const int cf = c + ((x < y) ? ((omx < y) ? 2 : 3) : ((omx < y) ? 1 : 0));
      // the strict-interior test as lane masks (fcmp predicates: 2 ordered
      // greater, 4 ordered less; inactive lanes read 0), combined in scalar code
      const uint64_t ok = __builtin_amdgcn_fcmp(px, 0.0, 2) & __builtin_amdgcn_fcmp(px, g.top_x, 4) &
                          __builtin_amdgcn_fcmp(py, 0.0, 2) & __builtin_amdgcn_fcmp(py, g.top_y, 4);
      if ((ok & ex) == ex) return cf;
    } else {
      const uint64_t ok = __builtin_amdgcn_fcmp(px, 0.0, 2) & __builtin_amdgcn_fcmp(px, g.top_x, 4) &
                          __builtin_amdgcn_fcmp(py, 0.0, 2) & __builtin_amdgcn_fcmp(py, g.top_y, 4);
      if ((ok & ex) == ex) {
        const double sx = px * g.rinc_x, sy = py * g.rinc_y;
        const double x = __builtin_amdgcn_fract(sx), y = __builtin_amdgcn_fract(sy);
        const int c = ((int)sx + mul24((int)sy, g.nC0)) * 4;
        const double omx = 1.0 - x;
        if (x < y) return c + ((omx < y) ? 2 : 3);
        return c + ((omx < y) ? 1 : 0);
      }
    }
  }
  const double p0 = dmin(g.top_x, dmax(0.0, px));
  const double p1 = dmin(g.top_y, dmax(0.0, py));
  double x, y;
  const int i = axis_simplex<POW2>(p0, g.nC0, g.inc_x, g.rinc_x, x);
  const int j = axis_simplex<POW2>(p1, g.nC1, g.inc_y, g.rinc_y, y);
  int c = (i + mul24(j, g.nC0)) * 4;  // triangles: 0 top, 1 right, 2 bottom, 3 left
  const bool left = px <= 0.0, right = px >= g.wall_x, up = py <= 0.0, down = py >= g.wall_y;
  if (left | right | up | down) {  // rare: outside the domain
    // the reference's rules, as selects on the four sides' tests
    const double nC0 = (double)g.nC0, nC1 = (double)g.nC1;
    const double sx = divinc<POW2>(px, g.inc_x, g.rinc_x), sy = divinc<POW2>(py, g.inc_y, g.rinc_y);
    const double nx = divinc<POW2>(-px, g.inc_x, g.rinc_x), ny = divinc<POW2>(-py, g.inc_y, g.rinc_y);
    const bool upL = up & (sy < sx), dnL = down & (sy - nC1 > nx);
    const bool upR = up & (ny > sx - nC0), dnR = down & (sy - nC1 > sx - nC0);
    const int tl = upL ? 0 : (dnL ? 2 : 3);
    const int tr = upR ? 0 : (dnR ? 2 : 1);
    const int tm = up ? 0 : 2;
    return c + (left ? tl : (right ? tr : tm));
  }
  const double omx = 1.0 - x;
  if (x < y) return c + ((omx < y) ? 2 : 3);
  return c + ((omx < y) ? 1 : 0);
}

// cell2's general type I path without branches: the inside and outside
// classifications are both computed and one is selected, so a lane that is
// alone in its wave (a single dependent chain, see lone_trips) pays no
// branch and no strict-interior test.  Same result as cell2 for every point.
template <bool POW2>
__device__ __forceinline__ int cell2_flat(const double* p, const Geom& g) {
  const double px = p[0], py = p[1];
  const double p0 = dmin(g.top_x, dmax(0.0, px));
  const double p1 = dmin(g.top_y, dmax(0.0, py));
  double x, y;
  const int i = axis_simplex<POW2>(p0, g.nC0, g.inc_x, g.rinc_x, x);
  const int j = axis_simplex<POW2>(p1, g.nC1, g.inc_y, g.rinc_y, y);
  const int c = (i + mul24(j, g.nC0)) * 4;
  const bool left = px <= 0.0, right = px >= g.wall_x, up = py <= 0.0, down = py >= g.wall_y;
  const double nC0 = (double)g.nC0, nC1 = (double)g.nC1;
  const double sx = divinc<POW2>(px, g.inc_x, g.rinc_x), sy = divinc<POW2>(py, g.inc_y, g.rinc_y);
  const double nx = divinc<POW2>(-px, g.inc_x, g.rinc_x), ny = divinc<POW2>(-py, g.inc_y, g.rinc_y);
  const bool upL = up & (sy < sx), dnL = down & (sy - nC1 > nx);
  const bool upR = up & (ny > sx - nC0), dnR = down & (sy - nC1 > sx - nC0);
  const int tl = upL ? 0 : (dnL ? 2 : 3);
  const int tr = upR ? 0 : (dnR ? 2 : 1);
  const int tm = up ? 0 : 2;
  const int tout = left ? tl : (right ? tr : tm);
  const double omx = 1.0 - x;
  const int tin = (x < y) ? ((omx < y) ? 2 : 3) : ((omx < y) ? 1 : 0);
  return c + ((left | right | up | down) ? tout : tin);
}

// ------------------------------------------------------------------ 3D
// calc_3Dto3D.cu:43-137
template <int TESS, bool POW2>
__device__ __forceinline__ int cell3(const double* p, const Geom& g) {
  if (TESS == TESS_II) {
    const int a = axis_box<POW2>(p[0], g.nC0, g.inc_x, g.rinc_x);
    const int b = axis_box<POW2>(p[1], g.nC1, g.inc_y, g.rinc_y);
    const int cc = axis_box<POW2>(p[2], g.nC2, g.inc_z, g.rinc_z);
    return a + mul24(b, g.nC0) + mul24(cc, mul24(g.nC1, g.nC0));
  }
  const double p0 = dmin(g.top_x, dmax(0.0, p[0]));
  const double p1 = dmin(g.top_y, dmax(0.0, p[1]));
  const double p2 = dmin(g.top_z, dmax(0.0, p[2]));  // top_z uses inc_x, as the reference
  double x, y, z;
  const int i = axis_simplex<POW2>(p0, g.nC0, g.inc_x, g.rinc_x, x);
  const int j = axis_simplex<POW2>(p1, g.nC1, g.inc_y, g.rinc_y, y);
  const int k = axis_simplex<POW2>(p2, g.nC2, g.inc_z, g.rinc_z, z);
  int c = mul24(i + mul24(j, g.nC0) + mul24(k, mul24(g.nC1, g.nC0)), 5);
  // alternating orientation: flip when the parity of i + j + k is odd
  const bool tf = ((i + j + k) & 1) != 0;
  if (tf) {
    const double tmp = x;
    x = y;
    y = 1.0 - tmp;
  }
  if (-x - y + z >= 0.0) return c + 1;
  if (x + y + z - 2.0 >= 0.0) return c + 2;
  if (-x + y - z >= 0.0) return c + 3;
  if (x - y - z >= 0.0) return c + 4;
  return c;
}

template <int DIM, int TESS, bool POW2, bool FAST = false, bool EARLY = true>
__device__ __forceinline__ int cell_of(const double* p, const Geom& g) {
  if (DIM == 1) return cell1<POW2>(p, g);
  if (DIM == 2) return cell2<TESS, POW2, FAST, EARLY>(p, g);
  return cell3<TESS, POW2>(p, g);
}

// ------------------------------------------------------ affine products
// calc_2Dto2D.cu:24-36, calc_1Dto1D.cu:18-52, calc_3Dto3D.cu:25-39
template <int DIM>
__device__ __forceinline__ void affine(double* x, const double* A, const double* b) {
  if (DIM == 1) {
    x[0] = A[0] * b[0] + A[1];
  } else if (DIM == 2) {
    x[0] = A[0] * b[0] + A[1] * b[1] + A[2];
    x[1] = A[3] * b[0] + A[4] * b[1] + A[5];
  } else {
    x[0] = A[0] * b[0] + A[1] * b[1] + A[2] * b[2] + A[3];
    x[1] = A[4] * b[0] + A[5] * b[1] + A[6] * b[2] + A[7];
    x[2] = A[8] * b[0] + A[9] * b[1] + A[10] * b[2] + A[11];
  }
}

template <int DIM>
__device__ __forceinline__ void linear(double* x, const double* A, const double* b) {
  if (DIM == 1) {
    x[0] = A[0] * b[0];
  } else if (DIM == 2) {
    x[0] = A[0] * b[0] + A[1] * b[1];
    x[1] = A[3] * b[0] + A[4] * b[1];
  } else {
    x[0] = A[0] * b[0] + A[1] * b[1] + A[2] * b[2];
    x[1] = A[4] * b[0] + A[5] * b[1] + A[6] * b[2];
    x[2] = A[8] * b[0] + A[9] * b[1] + A[10] * b[2];
  }
}

}  // namespace cpab
