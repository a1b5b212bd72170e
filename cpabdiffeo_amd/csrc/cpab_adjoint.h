// Reverse-mode derivative (adjoint) of the 2D CPA maps, and of the SSD loss
// of an image warped by them, with respect to the per-cell affine tables.
//
// Two maps, both evaluated as the reference evaluates them:
//  * calc_T_simple (cpab/gpu/calc_2Dto2D.cu:412-445): nT*nS RK2 midpoint
//    substeps, each with the cell of its start point (solveODE,
//    calc_2Dto2D.cu:167-197):
//        v1 = A_c [p;1],  m = p + h v1 / 2,  v2 = A_c [m;1],  p' = p + v2 h.
//    For a cotangent lambda' of p' (A_c = [L | b]):
//        g2 = h lambda',  gm = L^T g2,  g1 = (h/2) gm,
//        lambda = lambda' + gm + L^T g1,
//        dJ/dA_c += g2 [m;1]^T + g1 [p;1]^T.
//  * calc_T (calc_2Dto2D.cu:348-408), the map of calc_T_fwd / calc_T_inv:
//    per time step the proposal pNew = Trel_c [p;1] with Trel_c =
//    expm(dt A~_c) (the top two rows), kept when cell(pNew) == c, else the
//    step is redone by nS RK2 substeps (as above).  An accepted step gives
//        dJ/dTrel_c += lambda' [p;1]^T,  lambda = E_c^T lambda'
//    (E_c the linear part of Trel_c); the gradient in the tables then
//    follows through the matrix exponential: dJ/dA~_c += dt L(dt A~_c^T, G~)
//    with G~ the 3x3 whose top rows are dJ/dTrel_c and L(X, E) the Frechet
//    derivative of expm at X in the direction E (the adjoint of L(X, .) is
//    L(X^T, .)), evaluated by k_adj_expm.
// The accept/redo decisions and the cells of the substeps are replayed
// from the forward sweep; they change only on a set of measure zero and
// contribute nothing to the derivative (as in the reference's forward
// sensitivity solveODE2, calc_2Dto2D.cu:202-319).
//
// Schedule, one lane per point: the forward sweep writes the point at the
// start of every time step to a checkpoint buffer (nT points per point) and,
// for calc_T, one bit per step recording whether the step was redone; the
// backward sweep walks the time steps in reverse (each lane its own, for
// calc_T's map: accepted steps advance together, lanes at a redone step wait
// for others and run theirs together), recomputes a redone step's nS
// substep start points from its checkpoint into registers (NSM of them,
// unrolled so they stay in registers), then runs its substeps' adjoints
// j = nS-1 .. 0; an accepted step is one Trel adjoint unit.
//
// Accumulation, bit-reproducible: each lane keeps two register
// accumulators, dJ/dA of its current RK2 cell and dJ/dTrel of its current
// Trel cell, and flushes one when that cell changes.  Flushes go into
// tables private to the wave (LDS), one flushing lane at a time in lane
// order; at the end a block sums its
// waves' tables in wave order and k_adj_reduce sums the blocks in block
// order.  The order of every addition is fixed by the data, so the gradient
// is the same bit for bit on every run.
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

constexpr int ADJ_NSMAX = 16;  // substeps per time step kept in registers (the kernel's NSM is 10 or 16)

struct AdjDesc {
  int nInst, nC;
  int64_t nPts;          // points per instance (H*W in MODE 1)
  int H, W;
  const double* As;      // (nInst, nC, 6): the signed tables the RK2 integrates
  const double* Trels;   // (nInst, nC, 6): expm(dt As) for the calc_T map, else null
  const double* pos0;    // MODE 0: (nInst, nPts, 2)
  const double* lam;     // MODE 0: (nInst, nPts, 2) cotangents of the end points
  double* pos_out;       // optional (nInst, nPts, 2): end points
  const float* src;      // MODE 1: (nInst, H, W)
  const float* dst;      // MODE 1: (nInst, H, W)
  uint32_t* sig;         // optional (nInst, nPts): hash of every cell the forward sweep looked up
  double* ckpt;          // (nInst, nT, nPts, 2) workspace
  uint64_t* redo_bits;   // calc_T map: (nInst, ceil(nT / 64), nPts) workspace, bit k % 64 of word k / 64:
                         // step k was redone by RK2
  double* partial;       // (nInst, blocks per instance, NG*nC*6 + 2): tables, loss, redone steps
  double h;
  int nT, nS;
  int tile8;             // MODE 1: a wave's 64 pixels are an 8 x 8 tile (W and H multiples of 8)
  int inst_major;        // grid (nInst, blocks): every instance's first block (the domain corner) starts first
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

__device__ __forceinline__ uint32_t sig_mix(uint32_t s, int c) { return s * 0x9E3779B1u + (uint32_t)(c + 1); }

// One RK2 substep from p, its cell looked up at p (the same IEEE operations
// as solveODE); returns the cell.
template <int TESS, bool POW2>
__device__ __forceinline__ int rk2_substep(double* p, const double* At, double h, const Geom& g) {
  const int c = cell2<TESS, POW2>(p, g);
  const double* A = At + c * 6;
  double v[2], m[2];
  affine<2>(v, A, p);
  m[0] = p[0] + h * v[0] / 2.;
  m[1] = p[1] + h * v[1] / 2.;
  affine<2>(v, A, m);
  p[0] = p[0] + v[0] * h;
  p[1] = p[1] + v[1] * h;
  return c;
}

// The same substep from p in its known cell c; returns the cell of the new
// point (the cell the next substep or step starts in).
template <int TESS, bool POW2>
__device__ __forceinline__ int rk2_substep_in(double* p, int c, const double* At, double h, const Geom& g) {
  const double* A = At + c * 6;
  double v[2], m[2];
  affine<2>(v, A, p);
  m[0] = p[0] + h * v[0] / 2.;
  m[1] = p[1] + h * v[1] / 2.;
  affine<2>(v, A, m);
  p[0] = p[0] + v[0] * h;
  p[1] = p[1] + v[1] * h;
  return cell2<TESS, POW2>(p, g);
}

// Active lanes with `need` add their accumulator G into row `cell` of the
// wave's table `tab`, one lane at a time in lane order.  Which lanes are
// active where the call sits is fixed by the data and the code, so the order
// of the additions is too.
__device__ __forceinline__ void adj_flush(bool need, int cell, const double* G, double* tab) {
  uint64_t m = __ballot(need);
  const int lane = threadIdx.x & 63;
  while (m) {
    const int l = __builtin_ctzll(m);
    if (lane == l) {
      double* t = tab + cell * 6;
#pragma unroll
      for (int e = 0; e < 6; e++) t[e] = t[e] + G[e];
    }
    __builtin_amdgcn_wave_barrier();
    m &= m - 1;
  }
}

// LDS of k_adjoint2d (doubles): the tables (As, and Trels for calc_T), then
// per wave NG gradient tables.
__host__ __device__ inline size_t adj_lds_doubles(int nC, bool fwd, int waves) {
  const size_t NT = (size_t)nC * 6, NG = fwd ? 2 : 1;
  return NG * NT + (size_t)waves * NG * NT;
}

template <int TESS, bool POW2, int MODE, bool FWD, int NSM>
__global__ void __launch_bounds__(256, 4) k_adjoint2d(AdjDesc D) {
  extern __shared__ double lds[];
  __shared__ double red[4];
  constexpr int NG = FWD ? 2 : 1;
  // grid (blocks, nInst), or (nInst, blocks) with inst_major: the hardware
  // dispatches x fastest, so the instances' first blocks, which hold the
  // domain corner (a pixel that can redo most of its steps: the longest
  // chain of trips), then start before any instance's other blocks
  const int inst = D.inst_major ? (int)blockIdx.x : (int)blockIdx.y;
  const int64_t blk = D.inst_major ? (int64_t)blockIdx.y : (int64_t)blockIdx.x;
  const int64_t nblk = D.inst_major ? (int64_t)gridDim.y : (int64_t)gridDim.x;
  const int nC = D.nC, NT = nC * 6;
  const int nW = blockDim.x >> 6, wave = threadIdx.x >> 6;
  double* At = lds;
  double* Tr = lds + NT;                        // FWD only
  double* Gall = lds + NG * NT;                 // nW * NG tables
  double* GA = Gall + (size_t)wave * NG * NT;   // this wave's dJ/dA
  double* GT = GA + NT;                         // this wave's dJ/dTrel (FWD)
  const double* Ain = D.As + (int64_t)inst * NT;
  for (int i = threadIdx.x; i < NT; i += blockDim.x) {
    At[i] = Ain[i];
    if (FWD) Tr[i] = D.Trels[(int64_t)inst * NT + i];
  }
  for (int i = threadIdx.x; i < nW * NG * NT; i += blockDim.x) Gall[i] = 0.0;
  __syncthreads();
  const Geom& g = D.g;
  const double h = D.h;
  const int nT = D.nT, nS = D.nS;
  const int64_t idx = blk * blockDim.x + threadIdx.x;
  const int64_t nPts = D.nPts;
  double loss = 0.0;
  int nredo = 0;  // time steps redone by RK2 (every step of calc_T_simple)
  if (idx < nPts) {
    const int64_t ib = (int64_t)inst * nPts;
    // the point's index in the caller's arrays: idx, or in MODE 1 with tiles
    // the pixel of lane idx & 63 in 8 x 8 tile idx >> 6 (neighbouring pixels
    // cross facets at about the same steps, so fewer of a wave's steps have
    // a redone lane; the checkpoints stay indexed by idx)
    int64_t pix = idx;
    if (MODE == 1 && D.tile8) {
      const int64_t t = idx >> 6;
      const int u = (int)(idx & 63), tw = D.W >> 3;
      pix = ((t / tw) * 8 + (u >> 3)) * (int64_t)D.W + (t % tw) * 8 + (u & 7);
    }
    double p[2];
    if (MODE == 0) {
      p[0] = D.pos0[(ib + idx) * 2];
      p[1] = D.pos0[(ib + idx) * 2 + 1];
    } else {
      p[0] = (double)(pix % D.W);
      p[1] = (double)(pix / D.W);
    }
    // forward sweep, one checkpoint per time step (and, for calc_T, the
    // step's decision, 64 steps to a word)
    double* ck = D.ckpt + (int64_t)inst * nT * nPts * 2;
    const int nWords = (nT + 63) >> 6;
    uint64_t* rb = FWD ? D.redo_bits + (int64_t)inst * nWords * nPts : nullptr;
    uint32_t sg = 0;
    uint64_t bits = 0;
    // the cell of p, carried: an accepted step ends in the cell it started
    // in, and each substep looks up the cell of its end point
    int c = cell2<TESS, POW2>(p, g);
    for (int k = 0; k < nT; k++) {
      ck[((int64_t)k * nPts + idx) * 2] = p[0];
      ck[((int64_t)k * nPts + idx) * 2 + 1] = p[1];
      bool redo = true;
      if (FWD) {
        double pn[2];
        affine<2>(pn, Tr + c * 6, p);
        const int cn = cell2<TESS, POW2>(pn, g);
        sg = sig_mix(sig_mix(sg, c), cn);
        redo = cn != c;
        if (!redo) {
          p[0] = pn[0];
          p[1] = pn[1];
        }
        bits |= (uint64_t)redo << (k & 63);
        if ((k & 63) == 63 || k == nT - 1) {
          rb[(int64_t)(k >> 6) * nPts + idx] = bits;
          bits = 0;
        }
      }
      if (redo) {
        nredo++;
        for (int j = 0; j < nS; j++) {
          sg = sig_mix(sg, c);
          c = rk2_substep_in<TESS, POW2>(p, c, At, h, g);
        }
      }
    }
    if (D.sig) D.sig[ib + pix] = sg;
    if (D.pos_out) {
      D.pos_out[(ib + pix) * 2] = p[0];
      D.pos_out[(ib + pix) * 2 + 1] = p[1];
    }
    double lam[2];
    if (MODE == 0) {
      lam[0] = D.lam[(ib + idx) * 2];
      lam[1] = D.lam[(ib + idx) * 2 + 1];
    } else {
      const int64_t io = (int64_t)inst * D.H * D.W;
      double gx, gy;
      const double w = sample2_grad(D.src + io, D.H, D.W, p[0], p[1], &gx, &gy);
      const double r = w - (double)D.dst[io + pix];
      loss = 0.5 * r * r;
      lam[0] = r * gx;
      lam[1] = r * gy;
    }
    // backward sweep
    int ca = -1, ct = -1;  // cells of the RK2 and Trel accumulators (-1: empty)
    // the cell of the end point of the step being processed, carried back:
    // an accepted step starts in the cell it ends in, and a redone step's
    // start cell is the one its first substep looks up
    int cq = c;
    double Ga[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    double Gt[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    // the substeps' start cells as the recompute looks them up, two 16-bit
    // cells a register (the host keeps nC below 2^16), for the adjoint
    // substeps; with 16 substep registers the lookup is repeated there
    // instead (the kernel would spill more than it saves)
    constexpr bool kKeepCells = NSM <= 10;
    // One redone step from its start point q, in reverse: the recompute of
    // its substep start points (and cells), then the adjoint of each
    // substep.  Sets cq to the cell of q.
    auto redone_step = [&](const double* q) {
      double ps[NSM][2];
      uint32_t cs[(NSM + 1) / 2];
#pragma unroll
      for (int i = 0; i < (NSM + 1) / 2; i++) cs[i] = 0u;
      double r[2] = {q[0], q[1]};
#pragma unroll
      for (int j = 0; j < NSM; j++) {
        if (j < nS) {
          ps[j][0] = r[0];
          ps[j][1] = r[1];
          const int cj = rk2_substep<TESS, POW2>(r, At, h, g);
          if (kKeepCells) cs[j >> 1] |= (uint32_t)cj << ((j & 1) * 16);
        }
      }
#pragma unroll
      for (int j = NSM - 1; j >= 0; j--) {
        if (j < nS) {
          const double* pj = ps[j];
          const int c = kKeepCells ? (int)((cs[j >> 1] >> ((j & 1) * 16)) & 0xffffu) : cell2<TESS, POW2>(pj, g);
          if (j == 0) cq = c;
          adj_flush(ca >= 0 && c != ca, ca, Ga, GA);
          if (c != ca) {
#pragma unroll
            for (int e = 0; e < 6; e++) Ga[e] = 0.0;
            ca = c;
          }
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
          Ga[0] += g2x * m[0] + g1x * pj[0];
          Ga[1] += g2x * m[1] + g1x * pj[1];
          Ga[2] += g2x + g1x;
          Ga[3] += g2y * m[0] + g1y * pj[0];
          Ga[4] += g2y * m[1] + g1y * pj[1];
          Ga[5] += g2y + g1y;
        }
      }
    };
    if (FWD) {
      // Each lane walks its own steps back.  A trip either takes one
      // accepted step on every lane whose current step is accepted, or
      // (a redo trip) runs the redone step of every lane parked at one;
      // redo trips run once kRedoBatch lanes are parked or nothing else can
      // move, so a redone step costs its wave one pass of its nS substeps
      // shared by several lanes rather than one pass per lane and step.
      // Every choice is a ballot of the data: the sums keep a fixed order.
      // (12 lanes measured best on MI355X: 6, 8 and 16 within 4 %.)
      constexpr int kRedoBatch = 12;
      int k = nT - 1;  // this lane's step (-1: done)
      double q[2] = {ck[((int64_t)k * nPts + idx) * 2], ck[((int64_t)k * nPts + idx) * 2 + 1]};
      bits = rb[(int64_t)(k >> 6) * nPts + idx];
      bool redo = (bits >> (k & 63)) & 1;
      auto next_step = [&]() {  // the step before: its start point (checkpoint) and decision bit
        if (--k >= 0) {
          q[0] = ck[((int64_t)k * nPts + idx) * 2];
          q[1] = ck[((int64_t)k * nPts + idx) * 2 + 1];
          if ((k & 63) == 63) bits = rb[(int64_t)(k >> 6) * nPts + idx];
          redo = (bits >> (k & 63)) & 1;
        }
      };
      for (;;) {
        const uint64_t parked = __ballot(k >= 0 && redo);
        const uint64_t ready = __ballot(k >= 0 && !redo);
        if ((parked | ready) == 0) break;
        if (ready == 0 || __popcll(parked) >= kRedoBatch) {
          if (k >= 0 && redo) {
            redone_step(q);
            next_step();
          }
        } else if (k >= 0 && !redo) {  // the accepted Trel step: it starts in the cell it ends in
          const int c0 = cq;
          adj_flush(ct >= 0 && c0 != ct, ct, Gt, GT);
          if (c0 != ct) {
#pragma unroll
            for (int e = 0; e < 6; e++) Gt[e] = 0.0;
            ct = c0;
          }
          const double* T = Tr + c0 * 6;
          Gt[0] += lam[0] * q[0];
          Gt[1] += lam[0] * q[1];
          Gt[2] += lam[0];
          Gt[3] += lam[1] * q[0];
          Gt[4] += lam[1] * q[1];
          Gt[5] += lam[1];
          const double l0 = T[0] * lam[0] + T[3] * lam[1];
          const double l1 = T[1] * lam[0] + T[4] * lam[1];
          lam[0] = l0;
          lam[1] = l1;
          next_step();
        }
      }
    } else {  // calc_T_simple's map: every step is nS substeps, the same on every lane
      // (the checkpoint of the step before is loaded while a step is processed)
      double qn[2] = {ck[((int64_t)(nT - 1) * nPts + idx) * 2], ck[((int64_t)(nT - 1) * nPts + idx) * 2 + 1]};
      for (int k = nT - 1; k >= 0; k--) {
        const double q[2] = {qn[0], qn[1]};
        if (k > 0) {
          qn[0] = ck[((int64_t)(k - 1) * nPts + idx) * 2];
          qn[1] = ck[((int64_t)(k - 1) * nPts + idx) * 2 + 1];
        }
        redone_step(q);
      }
    }
    adj_flush(ca >= 0, ca, Ga, GA);
    if (FWD) adj_flush(ct >= 0, ct, Gt, GT);
  }
  const double ls = block_sum(loss, red);
  __syncthreads();
  const double nr = block_sum((double)nredo, red);
  __syncthreads();
  // the block's tables: its waves' tables summed in wave order
  double* out = D.partial + ((int64_t)inst * nblk + blk) * (NG * NT + 2);
  for (int i = threadIdx.x; i < NG * NT; i += blockDim.x) {
    double s = 0.0;
    for (int w = 0; w < nW; w++) s += Gall[(size_t)w * NG * NT + i];
    out[i] = s;
  }
  if (threadIdx.x == 0) {
    out[NG * NT] = ls;
    out[NG * NT + 1] = nr;
  }
}

// Sum the per-block partials of every instance in a fixed order:
// gradA (nInst, nC*6), gradT (nInst, nC*6; calc_T map only), loss (nInst,)
// and redone steps (nInst,), the last two optional.
// Grid (ceil((NG*NT + 2) / blockDim.x), nInst): one thread per entry, its
// blocks' partials loaded eight at a time ahead of the dependent sum, which
// adds them in block order.
__global__ void k_adj_reduce(const double* __restrict__ partial, int nBlocks, int NT, int NG,
                             double* __restrict__ gradA, double* __restrict__ gradT, double* __restrict__ loss,
                             double* __restrict__ nredo) {
  const int inst = blockIdx.y;
  const int stride = NG * NT + 2;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= stride) return;
  const double* pi = partial + (int64_t)inst * nBlocks * stride + e;
  double s = 0.0;
  int b = 0;
  for (; b + 8 <= nBlocks; b += 8) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; u++) v[u] = pi[(int64_t)(b + u) * stride];
#pragma unroll
    for (int u = 0; u < 8; u++) s += v[u];
  }
  for (; b < nBlocks; b++) s += pi[(int64_t)b * stride];
  if (e < NT) {
    gradA[(int64_t)inst * NT + e] = s;
  } else if (e < NG * NT) {
    gradT[(int64_t)inst * NT + e - NT] = s;
  } else if (e == NG * NT) {
    if (loss) loss[inst] = s;
  } else if (nredo) {
    nredo[inst] = s;
  }
}

// 3x3 row-major products.
__device__ inline void mat3_mul(const double* a, const double* b, double* c) {
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) c[i * 3 + j] = a[i * 3] * b[j] + a[i * 3 + 1] * b[3 + j] + a[i * 3 + 2] * b[6 + j];
}

// The Frechet derivative L(X, E) of the matrix exponential (3x3): the upper
// right block of exp([[X, E], [0, X]]) (Van Loan's block form), by scaling
// and squaring with a degree-20 Taylor polynomial in Horner form on the
// block pairs (D, U) ~ [[D, U], [0, D]].  E is normalised first (L is
// linear in E), so the scaling follows the size of X alone.
__device__ inline void expm_frechet3(const double* X, const double* E, double* L) {
  double en = 0.0, xn = 0.0;
  for (int j = 0; j < 3; j++) {
    double se = 0.0, sx = 0.0;
    for (int i = 0; i < 3; i++) {
      se += fabs(E[i * 3 + j]);
      sx += fabs(X[i * 3 + j]);
    }
    en = se > en ? se : en;
    xn = sx > xn ? sx : xn;
  }
  for (int i = 0; i < 9; i++) L[i] = 0.0;
  if (en == 0.0) return;
  int s = 0;
  while (xn > 0.5 && s < 64) {
    xn *= 0.5;
    s++;
  }
  const double sc = ldexp(1.0, -s);
  double x[9], e[9], Dm[9], U[9], t1[9], t2[9];
  for (int i = 0; i < 9; i++) {
    x[i] = X[i] * sc;
    e[i] = E[i] / en * sc;
    Dm[i] = (i % 4 == 0) ? 1.0 : 0.0;
    U[i] = 0.0;
  }
  for (int k = 20; k >= 1; k--) {
    // (D, U) <- (I + x D / k, (x U + e D) / k)
    mat3_mul(x, U, t1);
    mat3_mul(e, Dm, t2);
    for (int i = 0; i < 9; i++) U[i] = (t1[i] + t2[i]) / (double)k;
    mat3_mul(x, Dm, t1);
    for (int i = 0; i < 9; i++) Dm[i] = ((i % 4 == 0) ? 1.0 : 0.0) + t1[i] / (double)k;
  }
  for (int q = 0; q < s; q++) {
    // (D, U) <- (D D, D U + U D)
    mat3_mul(Dm, U, t1);
    mat3_mul(U, Dm, t2);
    for (int i = 0; i < 9; i++) U[i] = t1[i] + t2[i];
    mat3_mul(Dm, Dm, t1);
    for (int i = 0; i < 9; i++) Dm[i] = t1[i];
  }
  for (int i = 0; i < 9; i++) L[i] = U[i] * en;
}

// gradA_c += dt * (top rows of L(dt A~_c^T, G~_c)), G~_c the 3x3 with the
// rows dJ/dTrel_c: the chain rule through Trel_c = expm(dt A~_c).  One lane
// per (instance, cell).
__global__ void k_adj_expm(const double* __restrict__ As, const double* __restrict__ gradT, double* __restrict__ gradA,
                           int nC, int nInst, double dt) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)nC * nInst) return;
  const double* a = As + i * 6;
  const double* gt = gradT + i * 6;
  const double X[9] = {dt * a[0], dt * a[3], 0.0, dt * a[1], dt * a[4], 0.0, dt * a[2], dt * a[5], 0.0};
  const double E[9] = {gt[0], gt[1], gt[2], gt[3], gt[4], gt[5], 0.0, 0.0, 0.0};
  double L[9];
  expm_frechet3(X, E, L);
  double* ga = gradA + i * 6;
  for (int e = 0; e < 6; e++) ga[e] = ga[e] + dt * L[e];
}

}  // namespace cpab
