// Device-resident Metropolis sampling of 2D image registration.
//
// One step of Metropolis.step (cpab/cpaNd/inference/Metropolis.py:41-85) for
// many independent chains, with the Gaussian SSD likelihood of
// cpab/cpa2d/model/transformations/register/_ScaleDependentLogLikelihoodGaussian.py:34-109:
//
//   k_mcmc_propose   theta' = theta + L z (z ~ N(0, I), L lower triangular:
//                    Proposal.__call__, cpab/cpaNd/inference/Proposal.py:58-64),
//                    Avees = B theta' (CpaSpace.theta2Avees, _CpaSpace.py:366-376),
//                    sign -1, dt scaling and expm per cell (calc_T_inv prep).
//   k_integrate_ssd  calc_T_inv of the pixel grid (x = column, y = row), then
//                    at each finished point the bilinear sample of the source
//                    image and (warped - target)^2, reduced per block.
//   k_mcmc_accept    ll' = -0.5/sigma^2 sum (divided by the pixel count when
//                    use_ave_ll, Metropolis.py:58-59), log_ratio = ll' - ll,
//                    accept when log_ratio > 0 or log(u) < log_ratio (with
//                    `inverted` the reference's test log(u) > log_ratio,
//                    Metropolis.py:77-82).
//
// Random numbers: Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11), keyed by
// the seed, counter (step, chain, slot); Box-Muller for normals.  Everything
// is indexed by per-chain device counters, so a captured step replays.
#pragma once
#include "cpab_kernels.h"

#pragma clang fp contract(off)

namespace cpab {

struct Philox {
  uint32_t x0, x1, x2, x3;
};

__host__ __device__ inline Philox philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                                                uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
  for (int r = 0; r < 10; r++) {
    const uint64_t p0 = (uint64_t)M0 * c0;
    const uint64_t p1 = (uint64_t)M1 * c2;
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    const uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
    c0 = n0;
    c1 = lo1;
    c2 = n2;
    c3 = lo0;
    k0 += W0;
    k1 += W1;
  }
  return Philox{c0, c1, c2, c3};
}

// 53-bit uniform in [0, 1) from two words
__host__ __device__ inline double u53(uint32_t a, uint32_t b) {
  return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}

constexpr uint32_t SLOT_ACCEPT = 0x80000000u;

struct McmcDesc {
  int nChains, d, nC, nC0, nC1, tess;
  double inc_x, inc_y;
  int H, W;
  double sigma, dt, h;
  int nT, nS;
  int use_ave_ll, inverted, trace_len;
  uint32_t seed_lo, seed_hi;
  const double* B;      // (nC*6, d) row-major
  const double* L;      // (d, d) row-major lower triangular
  const float* src;     // (H, W)
  const float* dst;     // (H, W)
  double* theta;        // (nChains, d)
  double* theta_prop;   // (nChains, d)
  double* ll;           // (nChains,)
  double* ll_prop;      // (nChains,)
  int64_t* n_acc;       // (nChains,)
  int64_t* step;        // (nChains,)
  double* tables;       // (nChains, 2, nC, 6): Trels then signed As
  double* partial;      // (nChains, nBlocks)
  double* trace;        // (nChains, trace_len) or null
  unsigned long long* nfb;  // (nChains,) rejected Trel proposals, accumulated; or null
  int nBlocks;          // integration blocks per chain
  int ptsPerWave;
  Geom g;
};

// One block per chain.  with_noise == 0 (initialisation) takes theta' = theta.
__global__ void __launch_bounds__(256) k_mcmc_propose(McmcDesc D, int with_noise) {
  extern __shared__ double sh[];  // z[d] then theta'[d] then Avees[nC*6]
  const int ch = blockIdx.x;
  const int d = D.d;
  double* z = sh;
  double* th = sh + d;
  double* av = sh + 2 * d;
  const int64_t step = D.step[ch];
  const double* theta = D.theta + (int64_t)ch * d;
  if (with_noise) {
    for (int t = threadIdx.x; 2 * t < d; t += blockDim.x) {
      const Philox r = philox4x32_10((uint32_t)step, (uint32_t)(step >> 32), (uint32_t)ch, (uint32_t)t, D.seed_lo,
                                     D.seed_hi);
      const double u1 = 1.0 - u53(r.x0, r.x1), u2 = u53(r.x2, r.x3);
      const double rad = sqrt(-2.0 * log(u1));
      const double ang = 6.283185307179586 * u2;
      z[2 * t] = rad * cos(ang);
      if (2 * t + 1 < d) z[2 * t + 1] = rad * sin(ang);
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < d; i += blockDim.x) {
    double v = theta[i];
    if (with_noise) {
      double acc = 0.0;
      for (int j = 0; j <= i; j++) acc += D.L[(int64_t)i * d + j] * z[j];
      v = v + acc;
    }
    th[i] = v;
    D.theta_prop[(int64_t)ch * d + i] = v;
  }
  __syncthreads();
  const int ne = 6;
  for (int r = threadIdx.x; r < D.nC * ne; r += blockDim.x) {
    double acc = 0.0;
    const double* row = D.B + (int64_t)r * d;
    for (int j = 0; j < d; j++) acc += row[j] * th[j];
    av[r] = acc;
  }
  __syncthreads();
  double* T = D.tables + (int64_t)ch * 2 * D.nC * ne;
  double* sA = T + D.nC * ne;
  for (int c = threadIdx.x; c < D.nC; c += blockDim.x) {
    double a[6], adt[6], t6[6];
#pragma unroll
    for (int k = 0; k < 6; k++) {
      double v = av[c * 6 + k];
      v *= -1;  // calc_T_inv
      a[k] = v;
      adt[k] = v * D.dt;
    }
    expm_affine2(adt, t6, 12);
#pragma unroll
    for (int k = 0; k < 6; k++) {
      T[c * 6 + k] = t6[k];
      sA[c * 6 + k] = a[k];
    }
  }
}

// calc_T_inv of the H x W pixel grid for chain blockIdx.y, fused with the
// bilinear remap of the source and the squared error against the target.
template <int TESS, bool POW2>
__global__ void __launch_bounds__(256) k_integrate_ssd(McmcDesc D) {
  constexpr int NE = 6;
  extern __shared__ double lds[];
  __shared__ double red[4];
  const int ch = blockIdx.y;
  const int nC = D.nC;
  const double* tab = D.tables + (int64_t)ch * 2 * nC * NE;
  for (int i = threadIdx.x; i < 2 * nC * NE; i += blockDim.x) lds[i] = tab[i];
  __syncthreads();
  const double* Ttab = lds;
  const double* Atab = lds + nC * NE;
  const Geom& g = D.g;
  const int64_t nPts = (int64_t)D.H * D.W;
  const int64_t wave = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int64_t begin = wave * (int64_t)D.ptsPerWave;
  const int64_t end = begin + D.ptsPerWave < nPts ? begin + D.ptsPerWave : nPts;
  const int nT = D.nT, nS = D.nS;
  const double h = D.h;
  double acc = 0.0;
  unsigned int nfb = 0;
  if (begin < nPts) {
    int64_t next = begin;
    double p[2];
    int c = 0, t = 0, k = 0;
    int64_t my = 0;
    bool active = false;
    for (;;) {
      if (active && t >= nT) {
        const float w = sample2(D.src, D.H, D.W, p[0], p[1]);
        const double e = (double)w - (double)D.dst[my];
        acc += e * e;
        active = false;
      }
      const uint64_t idle = ballot(!active);
      if (idle != 0 && next < end) {
        if (!active) {
          const int64_t cand = next + rank_in(idle);
          if (cand < end) {
            my = cand;
            active = true;
            p[0] = (double)(my % D.W);
            p[1] = (double)(my / D.W);
            c = cell2<TESS, POW2>(p, g);
            t = 0;
            k = 0;
          }
        }
        next += __popcll(idle);
      }
      if (ballot(active) == 0) break;
      if (!active || t >= nT) continue;
      const double* M = (k == 0 ? Ttab : Atab) + c * NE;
      double y[2];
      affine<2>(y, M, p);
      if (k != 0) {
        double m[2];
        m[0] = p[0] + h * y[0] / 2.;
        m[1] = p[1] + h * y[1] / 2.;
        affine<2>(y, M, m);
        y[0] = p[0] + y[0] * h;
        y[1] = p[1] + y[1] * h;
      }
      const int cy = cell2<TESS, POW2>(y, g);
      bool stepped = false;
      if (k == 0) {
        if (cy == c) {
          p[0] = y[0];
          p[1] = y[1];
          stepped = true;
        } else {
          nfb++;
          if (nS > 0) k = nS;
          else stepped = true;
        }
      } else {
        p[0] = y[0];
        p[1] = y[1];
        c = cy;
        if (--k == 0) stepped = true;
      }
      if (stepped) t++;
    }
  }
  if (D.nfb) {  // one atomic per wave
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) nfb += __shfl_xor(nfb, o, 64);
    if ((threadIdx.x & 63) == 0 && nfb) atomicAdd(D.nfb + ch, (unsigned long long)nfb);
  }
  // block reduction (fixed order) -> one partial per block
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int i = 0; i < (int)(blockDim.x >> 6); i++) s += red[i];
    D.partial[(int64_t)ch * D.nBlocks + blockIdx.x] = s;
  }
}

// One wave per chain.  init == 1 sets the current log-likelihood.
__global__ void __launch_bounds__(64) k_mcmc_accept(McmcDesc D, int init) {
  const int ch = blockIdx.x;
  double s = 0.0;
  for (int i = threadIdx.x; i < D.nBlocks; i += 64) s += D.partial[(int64_t)ch * D.nBlocks + i];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  double llp = s * (-0.5 / (D.sigma * D.sigma));
  if (D.use_ave_ll) llp = llp / (double)((int64_t)D.H * D.W);
  const int64_t step = D.step[ch];
  bool acc;
  if (init) {
    acc = true;
  } else {
    const double lr = llp - D.ll[ch];
    if (lr > 0) {
      acc = true;
    } else {
      const Philox r = philox4x32_10((uint32_t)step, (uint32_t)(step >> 32), (uint32_t)ch, SLOT_ACCEPT, D.seed_lo,
                                     D.seed_hi);
      const double lu = log(u53(r.x0, r.x1));
      acc = D.inverted ? (lu > lr) : (lu < lr);
    }
  }
  if (acc) {
    for (int i = threadIdx.x; i < D.d; i += 64) D.theta[(int64_t)ch * D.d + i] = D.theta_prop[(int64_t)ch * D.d + i];
  }
  if (threadIdx.x == 0) {
    D.ll_prop[ch] = llp;
    if (acc) D.ll[ch] = llp;
    if (!init) {
      if (acc) D.n_acc[ch] += 1;
      if (D.trace && step < D.trace_len) D.trace[(int64_t)ch * D.trace_len + step] = acc ? llp : D.ll[ch];
      D.step[ch] = step + 1;
    }
  }
}

}  // namespace cpab
