// Device-resident Metropolis sampling of 2D image registration.
//
// One step of Metropolis.step (cpab/cpaNd/inference/Metropolis.py:41-85) for
// many independent chains, with the Gaussian SSD likelihood of
// cpab/cpa2d/model/transformations/register/_ScaleDependentLogLikelihoodGaussian.py:34-109,
// is ONE kernel launch, k_mcmc_step, with grid (blocks per chain, chains):
//
//   prologue  every block loads the chain's armed proposal tables into LDS.
//   body      calc_T_inv of the block's pixels (x = column, y = row); at each
//             finished point the bilinear sample of the source image and
//             (warped - target)^2.
//   epilogue  block sums are published (release), the last block of the chain
//             (acquire) sums them in a fixed order: ll' = -0.5/sigma^2 sum
//             (divided by the pixel count when use_ave_ll, Metropolis.py:58-59),
//             log_ratio = ll' - ll, accept when log_ratio > 0 or log(u) <
//             log_ratio (with `inverted` the reference's test log(u) >
//             log_ratio, Metropolis.py:77-82), advances the chain's step
//             counter, and arms the next proposal once for the chain:
//             theta' = theta + L z (z ~ N(0, I), L lower triangular:
//             Proposal.__call__, cpab/cpaNd/inference/Proposal.py:58-64),
//             Avees = B theta' (CpaSpace.theta2Avees, _CpaSpace.py:366-376),
//             sign -1, dt scaling and expm per cell (the calc_T_inv prep).
// k_mcmc_prep arms the first (noise-free) proposal before the init launch.
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
  double* tables;       // (nChains, 2, nC, 6): the armed proposal's Trels then signed As
  double* partial;      // (nChains, nBlocks) block sums of squared errors
  double* partial_nfb;  // (nChains, nBlocks) block counts of rejected proposals
  unsigned int* done;   // (nChains,) blocks finished in the current launch (0 between launches)
  double* trace;        // (nChains, trace_len) or null
  unsigned long long* nfb;  // (nChains,) rejected Trel proposals, accumulated; or null
  int nBlocks;          // integration blocks per chain
  int ptsPerWave;
  Geom g;
};

// LDS workspace of the proposal (doubles): z, theta', Avees.
__host__ __device__ inline int mcmc_prop_doubles(int nC, int d) { return 2 * d + nC * 6; }

// The proposal of chain `ch` for the step with counter `step`, from the
// state theta (LDS or global), by the whole block: theta' = theta + L z,
// Avees = B theta', and the calc_T_inv tables (Trels then signed As) into
// D.tables[ch]; theta' into D.theta_prop[ch].  `w` is LDS workspace.
__device__ inline void mcmc_propose(const McmcDesc& D, int ch, int64_t step, int with_noise, const double* theta,
                                    double* w) {
  const int d = D.d, nC = D.nC;
  double* z = w;
  double* th = z + d;
  double* av = th + d;
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
  for (int r = threadIdx.x; r < nC * 6; r += blockDim.x) {
    double acc = 0.0;
    const double* row = D.B + (int64_t)r * d;
    for (int j = 0; j < d; j++) acc += row[j] * th[j];
    av[r] = acc;
  }
  __syncthreads();
  double* T = D.tables + (int64_t)ch * 2 * nC * 6;
  double* sA = T + nC * 6;
  for (int c = threadIdx.x; c < nC; c += blockDim.x) {
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

// One block per chain: arm the proposal of the chain's current step.
__global__ void __launch_bounds__(256) k_mcmc_prep(McmcDesc D, int with_noise) {
  extern __shared__ double w[];
  const int ch = blockIdx.x;
  mcmc_propose(D, ch, D.step[ch], with_noise, D.theta + (int64_t)ch * D.d, w);
}

// The last block of the chain: reduce, decide, advance, then arm the next
// proposal.  `w` is LDS workspace of d + mcmc_prop_doubles doubles.
__device__ inline void mcmc_finish(const McmcDesc& D, int ch, int64_t step, int init, double* w) {
  double* theta_new = w;
  if (threadIdx.x < 64) {
    double s = 0.0, nf = 0.0;
    for (int i = threadIdx.x; i < D.nBlocks; i += 64) {
      s += __hip_atomic_load(D.partial + (int64_t)ch * D.nBlocks + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      nf += __hip_atomic_load(D.partial_nfb + (int64_t)ch * D.nBlocks + i, __ATOMIC_RELAXED,
                              __HIP_MEMORY_SCOPE_AGENT);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      s += __shfl_xor(s, o, 64);
      nf += __shfl_xor(nf, o, 64);
    }
    double llp = s * (-0.5 / (D.sigma * D.sigma));
    if (D.use_ave_ll) llp = llp / (double)((int64_t)D.H * D.W);
    bool acc;
    if (init) {
      acc = true;
    } else {
      const double lr = llp - D.ll[ch];
      if (lr > 0) {
        acc = true;
      } else {
        const Philox r = philox4x32_10((uint32_t)step, (uint32_t)(step >> 32), (uint32_t)ch, SLOT_ACCEPT,
                                       D.seed_lo, D.seed_hi);
        const double lu = log(u53(r.x0, r.x1));
        acc = D.inverted ? (lu > lr) : (lu < lr);
      }
    }
    for (int i = threadIdx.x; i < D.d; i += 64) {
      const double v = acc ? D.theta_prop[(int64_t)ch * D.d + i] : D.theta[(int64_t)ch * D.d + i];
      theta_new[i] = v;
      D.theta[(int64_t)ch * D.d + i] = v;
    }
    if (threadIdx.x == 0) {
      D.ll_prop[ch] = llp;
      if (acc) D.ll[ch] = llp;
      if (!init) {
        if (acc) D.n_acc[ch] += 1;
        if (D.trace && step < D.trace_len) D.trace[(int64_t)ch * D.trace_len + step] = D.ll[ch];
        D.step[ch] = step + 1;
      }
      if (D.nfb) D.nfb[ch] += (unsigned long long)nf;
      D.done[ch] = 0;  // re-arm for the next launch (kernel boundaries order it)
    }
  }
  __syncthreads();
  // the next step's proposal, from the state just decided
  mcmc_propose(D, ch, init ? step : step + 1, 1, theta_new, w + D.d);
}

// init == 1: evaluate the armed (noise-free) proposal as the current state.
template <int TESS, bool POW2>
__global__ void __launch_bounds__(256) k_mcmc_step(McmcDesc D, int init) {
  extern __shared__ double lds[];
  __shared__ double red[4], rednf[4];
  __shared__ int last;
  const int ch = blockIdx.y;
  const int nC = D.nC;
  const int64_t step = D.step[ch];
  const double* tab = D.tables + (int64_t)ch * 2 * nC * 6;
  for (int i = threadIdx.x; i < 2 * nC * 6; i += blockDim.x) lds[i] = tab[i];
  __syncthreads();
  const double* Ttab = lds;
  const double* Atab = lds + nC * 6;
  const Geom& g = D.g;
  const int64_t nPts = (int64_t)D.H * D.W;
  const int64_t wave = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int64_t begin = wave * (int64_t)D.ptsPerWave;
  const int64_t end = begin + D.ptsPerWave < nPts ? begin + D.ptsPerWave : nPts;
  const int nT = D.nT, nS = D.nS;
  const double h = D.h;
  double acc = 0.0;
  int nfb = 0;
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
      do {
        if (trip<2, TESS, POW2>(p, c, k, nfb, Ttab, Atab, h, nS, g)) t++;
      } while (ballot(t >= nT) == 0);
    }
  }
  // block sums (fixed order), published with release; the last block
  // acquires.  No same-address atomics besides the one completion count per
  // block: on a multi-XCD part those serialise at the memory side.
  double nf = (double)nfb;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    acc += __shfl_xor(acc, o, 64);
    nf += __shfl_xor(nf, o, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    red[threadIdx.x >> 6] = acc;
    rednf[threadIdx.x >> 6] = nf;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double sum = 0.0, sumnf = 0.0;
    for (int i = 0; i < (int)(blockDim.x >> 6); i++) {
      sum += red[i];
      sumnf += rednf[i];
    }
    __hip_atomic_store(D.partial + (int64_t)ch * D.nBlocks + blockIdx.x, sum, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(D.partial_nfb + (int64_t)ch * D.nBlocks + blockIdx.x, sumnf, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
    const unsigned int prev = __hip_atomic_fetch_add(D.done + ch, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    last = (prev == (unsigned int)(D.nBlocks - 1));
  }
  __syncthreads();
  if (last) mcmc_finish(D, ch, step, init, lds);  // block-uniform branch
}

}  // namespace cpab
