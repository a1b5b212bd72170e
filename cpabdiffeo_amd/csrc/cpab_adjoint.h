// Reverse-mode derivative (adjoint) of the 2D calc_T_simple map and of the
// SSD loss of an image warped by it, with respect to the per-cell affine
// tables.
//
// The map is calc_T_simple (cpab/gpu/calc_2Dto2D.cu:412-445): nT*nS RK2
// midpoint substeps, each with the cell of its start point
// (solveODE, calc_2Dto2D.cu:167-197):
//     v1 = A_c [p;1],  m = p + h v1 / 2,  v2 = A_c [m;1],  p' = p + v2 h.
// It is the map the reference differentiates (its forward sensitivity
// solveODE2, calc_2Dto2D.cu:202-319), so the adjoint here and that
// sensitivity are two evaluations of the same derivative.  For a cotangent
// lambda' of p' (A_c = [L | b]):
//     g2 = h lambda',  gm = L^T g2,  g1 = (h/2) gm,
//     lambda = lambda' + gm + L^T g1,
//     dJ/dA_c += g2 [m;1]^T + g1 [p;1]^T.
// (Cell changes are measure-zero events and contribute nothing, as in the
// sensitivity.)
//
// Schedule, one lane per point: the forward sweep writes the point at the
// start of every time step to a checkpoint buffer (nT points per point, not
// the nT*nS substeps); the backward sweep walks the time steps in reverse,
// recomputes the step's nS substep start points from its checkpoint into
// registers (NSMAX of them, unrolled so they stay in registers), and runs
// the substep adjoints in reverse.  dJ/dA accumulates per lane in registers
// for the lane's current cell and is added to the block's LDS copy (one
// nC x 6 table) when the cell changes, so LDS atomics are rare.  Per-block
// tables and loss sums go to a partial buffer; k_adj_reduce sums them in a
// fixed order.
//
// MODE 0 (dT/dA^T lambda): points and cotangents given.
// MODE 1 (SSD): points are the H x W pixel grid (x = column, y = row); the
// loss is 0.5 * sum (src(T(x)) - dst(x))^2 with the bilinear sample taken in
// FP64 (the convention of k_remap2d: OpenCV INTER_LINEAR, zero outside),
// and lambda = (w - dst) * grad w at T(x).
#pragma once
#include "cpab_kernels.h"

#pragma clang fp contract(off)

namespace cpab {

constexpr int ADJ_NSMAX = 16;  // substeps per time step kept in registers

struct AdjDesc {
  int nInst, nC;
  int64_t nPts;          // points per instance (H*W in MODE 1)
  int H, W;
  const double* As;      // (nInst, nC, 6): the tables the map integrates (signed As)
  const double* pos0;    // MODE 0: (nInst, nPts, 2)
  const double* lam;     // MODE 0: (nInst, nPts, 2) cotangents of the end points
  double* pos_out;       // optional (nInst, nPts, 2): end points
  const float* src;      // MODE 1: (nInst, H, W)
  const float* dst;      // MODE 1: (nInst, H, W)
  double* ckpt;          // (nInst, nT, nPts, 2) workspace
  double* partial;       // (nInst, blocks per instance, nC*6 + 1)
  double h;
  int nT, nS;
  Geom g;
};

// The FP64 bilinear sample and its gradient (the k_remap2d convention).
__device__ __forceinline__ double sample2_grad(const float* __restrict__ img, int H, int W, double x, double y,
                                               double* gx_out, double* gy_out) {
  const double xf = floor(x), yf = floor(y);
  *gx_out = 0.0;
  *gy_out = 0.0;
  if (!(xf >= -1.0 && xf < (double)W && yf >= -1.0 && yf < (double)H)) return 0.0;
  const double fx = x - xf, fy = y - yf;
  double v00 = 0.0, v01 = 0.0, v10 = 0.0, v11 = 0.0;
  const int x0 = (int)xf, y0 = (int)yf, x1 = x0 + 1, y1 = y0 + 1;
  const bool okx0 = x0 >= 0, okx1 = x1 < W, oky0 = y0 >= 0, oky1 = y1 < H;
  if (oky0 && okx0) v00 = img[(int64_t)y0 * W + x0];
  if (oky0 && okx1) v01 = img[(int64_t)y0 * W + x1];
  if (oky1 && okx0) v10 = img[(int64_t)y1 * W + x0];
  if (oky1 && okx1) v11 = img[(int64_t)y1 * W + x1];
  const double gx = 1.0 - fx, gy = 1.0 - fy;
  const double a = gx * v00 + fx * v01;
  const double b = gx * v10 + fx * v11;
  *gx_out = gy * (v01 - v00) + fy * (v11 - v10);
  *gy_out = b - a;
  return gy * a + fy * b;
}

template <int TESS, bool POW2>
__device__ __forceinline__ void rk2_substep(double* p, const double* At, double h, const Geom& g) {
  const int c = cell2<TESS, POW2>(p, g);
  const double* A = At + c * 6;
  double v[2], m[2];
  affine<2>(v, A, p);
  m[0] = p[0] + h * v[0] / 2.;
  m[1] = p[1] + h * v[1] / 2.;
  affine<2>(v, A, m);
  p[0] = p[0] + v[0] * h;
  p[1] = p[1] + v[1] * h;
}

template <int TESS, bool POW2, int MODE>
__global__ void __launch_bounds__(256) k_adjoint2d(AdjDesc D) {
  extern __shared__ double lds[];
  __shared__ double red[4];
  const int inst = blockIdx.y;
  const int nC = D.nC, NT = nC * 6;
  double* At = lds;
  double* Gs = lds + NT;
  const double* Ain = D.As + (int64_t)inst * NT;
  for (int i = threadIdx.x; i < NT; i += blockDim.x) {
    At[i] = Ain[i];
    Gs[i] = 0.0;
  }
  __syncthreads();
  const Geom& g = D.g;
  const double h = D.h;
  const int nT = D.nT, nS = D.nS;
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nPts = D.nPts;
  double loss = 0.0;
  if (idx < nPts) {
    const int64_t ib = (int64_t)inst * nPts;
    double p[2];
    if (MODE == 0) {
      p[0] = D.pos0[(ib + idx) * 2];
      p[1] = D.pos0[(ib + idx) * 2 + 1];
    } else {
      p[0] = (double)(idx % D.W);
      p[1] = (double)(idx / D.W);
    }
    // forward sweep, one checkpoint per time step
    double* ck = D.ckpt + (int64_t)inst * nT * nPts * 2;
    for (int k = 0; k < nT; k++) {
      ck[((int64_t)k * nPts + idx) * 2] = p[0];
      ck[((int64_t)k * nPts + idx) * 2 + 1] = p[1];
      for (int j = 0; j < nS; j++) rk2_substep<TESS, POW2>(p, At, h, g);
    }
    if (D.pos_out) {
      D.pos_out[(ib + idx) * 2] = p[0];
      D.pos_out[(ib + idx) * 2 + 1] = p[1];
    }
    double lam[2];
    if (MODE == 0) {
      lam[0] = D.lam[(ib + idx) * 2];
      lam[1] = D.lam[(ib + idx) * 2 + 1];
    } else {
      const int64_t io = (int64_t)inst * D.H * D.W;
      double gx, gy;
      const double w = sample2_grad(D.src + io, D.H, D.W, p[0], p[1], &gx, &gy);
      const double r = w - (double)D.dst[io + idx];
      loss = 0.5 * r * r;
      lam[0] = r * gx;
      lam[1] = r * gy;
    }
    // backward sweep
    int cl = -1;
    double G[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int k = nT - 1; k >= 0; k--) {
      double ps[ADJ_NSMAX][2];
      double q[2] = {ck[((int64_t)k * nPts + idx) * 2], ck[((int64_t)k * nPts + idx) * 2 + 1]};
#pragma unroll
      for (int j = 0; j < ADJ_NSMAX; j++) {
        if (j < nS) {
          ps[j][0] = q[0];
          ps[j][1] = q[1];
          rk2_substep<TESS, POW2>(q, At, h, g);
        }
      }
#pragma unroll
      for (int j = ADJ_NSMAX - 1; j >= 0; j--) {
        if (j < nS) {
          const double* pj = ps[j];
          const int c = cell2<TESS, POW2>(pj, g);
          const double* A = At + c * 6;
          double v1[2], m[2];
          affine<2>(v1, A, pj);
          m[0] = pj[0] + h * v1[0] / 2.;
          m[1] = pj[1] + h * v1[1] / 2.;
          const double g2x = h * lam[0], g2y = h * lam[1];
          const double gmx = A[0] * g2x + A[3] * g2y, gmy = A[1] * g2x + A[4] * g2y;
          const double g1x = 0.5 * h * gmx, g1y = 0.5 * h * gmy;
          lam[0] = lam[0] + gmx + (A[0] * g1x + A[3] * g1y);
          lam[1] = lam[1] + gmy + (A[1] * g1x + A[4] * g1y);
          if (c != cl) {
            if (cl >= 0)
#pragma unroll
              for (int e = 0; e < 6; e++) atomicAdd(Gs + cl * 6 + e, G[e]);
#pragma unroll
            for (int e = 0; e < 6; e++) G[e] = 0.0;
            cl = c;
          }
          G[0] += g2x * m[0] + g1x * pj[0];
          G[1] += g2x * m[1] + g1x * pj[1];
          G[2] += g2x + g1x;
          G[3] += g2y * m[0] + g1y * pj[0];
          G[4] += g2y * m[1] + g1y * pj[1];
          G[5] += g2y + g1y;
        }
      }
    }
    if (cl >= 0)
#pragma unroll
      for (int e = 0; e < 6; e++) atomicAdd(Gs + cl * 6 + e, G[e]);
  }
  const double ls = block_sum(loss, red);
  __syncthreads();
  double* out = D.partial + ((int64_t)inst * gridDim.x + blockIdx.x) * (NT + 1);
  for (int i = threadIdx.x; i < NT; i += blockDim.x) out[i] = Gs[i];
  if (threadIdx.x == 0) out[NT] = ls;
}

// Sum the per-block partials of every instance in a fixed order:
// gradA (nInst, nC*6) and loss (nInst,).
__global__ void k_adj_reduce(const double* __restrict__ partial, int nBlocks, int NT, double* __restrict__ gradA,
                             double* __restrict__ loss) {
  const int inst = blockIdx.x;
  const double* pi = partial + (int64_t)inst * nBlocks * (NT + 1);
  for (int e = threadIdx.x; e <= NT; e += blockDim.x) {
    double s = 0.0;
    for (int b = 0; b < nBlocks; b++) s += pi[(int64_t)b * (NT + 1) + e];
    if (e < NT) {
      gradA[(int64_t)inst * NT + e] = s;
    } else if (loss) {
      loss[inst] = s;
    }
  }
}

}  // namespace cpab
