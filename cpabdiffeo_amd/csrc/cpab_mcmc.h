// Device-resident Metropolis sampling of 2D image registration.
//
// Steps of Metropolis.step (cpab/cpaNd/inference/Metropolis.py:41-85) for
// many independent chains, with the Gaussian SSD likelihood of
// cpab/cpa2d/model/transformations/register/_ScaleDependentLogLikelihoodGaussian.py:34-109.
// One launch of k_mcmc_step, grid (K x blocks per chain, chains), evaluates K
// armed proposals per chain and decides as many steps as stay valid:
//
//   prologue  every block loads its slot's proposal tables into LDS.
//   body      calc_T_inv of the block's pixels (x = column, y = row); at each
//             finished point the bilinear sample of the source image and
//             (warped - target)^2.
//   epilogue  block sums are published (release); the last block of each
//             slot (acquire) sums them in a fixed order: ll' = -0.5/sigma^2
//             sum (divided by the pixel count when use_ave_ll,
//             Metropolis.py:58-59); the last block of the chain (acquire)
//             forms log_ratio = ll' - ll (+ wlp (lp' - lp) with a prior,
//             Metropolis.py:64-70), accepts when log_ratio > 0 or
//             log(u) < log_ratio (with `inverted` the reference's test
//             log(u) > log_ratio, Metropolis.py:77-82), slot after slot until
//             the first rejection, advances the step counter by the slots
//             decided, and arms the next K proposals (theta') for the chain.
//   then      k_mcmc_tables, a launch of (2K (+1) x chains) blocks, forms
//             the armed proposals' per-cell tables (B theta', expm) for the
//             next step launch, the L z vectors its last block will arm
//             from, and (several pixels per lane) the next dispatch order.
//
// Speculation (K > 1, prefetching in the sense of Brockwell, J. Comput.
// Graph. Stat. 15(1), 2006): slot m holds the proposal of step s + m on the
// path where steps s .. s+m-1 are all accepted.  Slots are decided in order;
// the first rejected slot ends the launch's steps and the later slots are
// discarded -- and, once the rejection is known during the launch, the
// blocks of later slots that have not started skip their pixels (early
// abort).  Every decided step uses exactly the numbers the K = 1 chain uses
// (same proposal sums, same random counters), so the chain is the same bit
// for bit.
// k_mcmc_prep (+ k_mcmc_tables) arms the first (noise-free) proposal before
// the init launch.
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
// Spacing of the per-slot completion counters (unsigned ints): one 128-byte
// line per counter.
constexpr int DONE_STRIDE = 32;
constexpr int kNoAbort = 0x7fffffff;

// Threads per block of the registration step: 256 (4 waves) or 64.  A block
// keeps its CU resources until its slowest wave ends, so small blocks free
// them sooner.  Measured on MI355X: one chain with speculation is fastest at
// 256 (64: -2 %, 1 024: -20 %), 256 chains one step per launch at 64 (+12 %).
constexpr int MCMC_BLOCK = 256;

struct McmcDesc {
  int nChains, d, nC, nC0, nC1, tess;
  double inc_x, inc_y;
  int H, W;
  double sigma, dt, h;
  int nT, nS;
  int use_ave_ll, inverted, trace_len;
  uint32_t seed_lo, seed_hi;
  int K;                // proposals evaluated per chain and launch (speculation depth)
  const double* Bt;     // (d, nC*6) row-major: the basis B (Avees = B theta) transposed
  const double* L;      // (d, d) row-major lower triangular
  const float* src;     // (H, W)
  const float* dst;     // (H, W)
  double* theta;        // (nChains, d)
  double* theta_prop;   // (nChains, K, d): armed proposals, slot m = step + m
  double* ll;           // (nChains,)
  double* ll_prop;      // (nChains,)
  int64_t* n_acc;       // (nChains,)
  int64_t* step;        // (nChains,)
  int64_t* stop;        // (nChains,) step counter at which the current run ends
  double* tables;       // (nChains, K, 2, nC, 6): per slot the Trels then the signed As
  double* partial;      // (nChains, K, nBlocks) block sums of squared errors
  double* partial_px;   // (nChains, K, nBlocks) pixels each block integrated (0: not started)
  double* slot_res;     // (nChains, K, 3): ll', rejected Trel proposals and pixels of each slot
  unsigned long long* ready;  // (nChains,) bit m: slot m's ll' is published (current launch; 0 between)
  int* abort_after;     // (nChains,) a slot above it is not started: some slot at or below it
                        // was found rejected on the all-accepted path (kNoAbort between launches)
  double* lz;           // (nChains, 2K, d): L z_{step+j}, j < 2K, for the current step counter
  double* partial_nfb;  // (nChains, K, nBlocks) block counts of rejected Trel proposals
  unsigned int* done;   // (nChains,) slots whose blocks have all finished in the current launch
  unsigned int* done_slot;  // (nChains, K, DONE_STRIDE) blocks finished per slot in the current
                            // launch, one 128-byte line each (all 0 between launches)
  double* trace;        // (nChains, trace_len) or null
  unsigned long long* nfb;  // (nChains,) rejected Trel proposals of decided steps, accumulated; or null
  int64_t* stats;       // (nChains, 5) or null: proposals evaluated in full, rejected Trel
                        // proposals of every pixel integrated, launches, steps decided,
                        // pixels integrated
  int nBlocks;          // integration blocks per chain and slot
  int ptsPerWave;
  const uint32_t* pixmap;  // (H*W) the pixel of each work index: the domain's top row and left
                           // column first, then the interior's partial strips, then its 8 x 8 tiles
  int poll;             // abort checkpoints per pixel run (one pixel per lane only; 0: none)
  int* order;           // (nChains, nBlocks): the pixel block each dispatch position integrates,
                        // or null: dispatch position = pixel block
  double wlp;           // weight of the log-prior ratio
  const double* J;      // (d, d) row-major prior precision, or null: no prior term
  double* lp;           // (nChains,) log-prior of the current theta (with J)
  double* lp_prop;      // (nChains, K) log-prior of each armed proposal (with J)
  Geom g;
  int integrate_corner;  // 1: the domain corner is integrated even when its sample is fixed
};

// LDS of the last block's epilogue (doubles): ll', fallback count,
// acceptance uniform, pixel count and log-prior per slot.
__host__ __device__ inline int mcmc_finish_doubles(int K) { return 5 * K; }
// LDS of k_mcmc_prep (doubles): normals and proposals of every slot.
__host__ __device__ inline int mcmc_arm_doubles(int K, int d) { return 2 * K * d; }
// LDS of k_mcmc_tables (doubles): one proposal, its Avees and its log-prior
// terms, two normals, or the block counts of the dispatch ordering.
__host__ __device__ inline int mcmc_tables_doubles(int nC, int d, int nBlocks) {
  int n = 2 * d + 6 * nC;
  const int o = nBlocks < 1024 ? nBlocks : 1024;
  return n > o ? n : o;
}

// Normals 2t and 2t+1 of z_s ~ N(0, I_d) into z (Philox counter (s, chain,
// t), Box-Muller).
__device__ inline void mcmc_normal_pair(const McmcDesc& D, int ch, int64_t s, int t, double* z) {
  const int d = D.d;
  const Philox r = philox4x32_10((uint32_t)s, (uint32_t)(s >> 32), (uint32_t)ch, (uint32_t)t, D.seed_lo, D.seed_hi);
  const double u1 = 1.0 - u53(r.x0, r.x1), u2 = u53(r.x2, r.x3);
  const double rad = sqrt(-2.0 * log(u1));
  const double ang = 6.283185307179586 * u2;
  z[2 * t] = rad * cos(ang);
  if (2 * t + 1 < d) z[2 * t + 1] = rad * sin(ang);
}

// (L z)_i, summed in index order, loads of L issued eight at a time ahead of
// the dependent sum.
__device__ inline double mcmc_lz_row(const double* __restrict__ Li, const double* zm, int i) {
  double acc = 0.0;
  int j = 0;
  for (; j + 8 <= i + 1; j += 8) {
    double l[8];
#pragma unroll
    for (int u = 0; u < 8; u++) l[u] = Li[j + u];
#pragma unroll
    for (int u = 0; u < 8; u++) acc += l[u] * zm[j + u];
  }
  for (; j <= i; j++) acc += Li[j] * zm[j];
  return acc;
}

// Arm the K proposals that follow `step` from the state theta, by the whole
// block.  Slot m is the proposal of step + m on the path where every earlier
// slot is accepted: theta'_0 = theta + L z_step, theta'_m = theta'_{m-1} +
// L z_{step+m} -- the same sums, in the same order, as the one-step-at-a-time
// chain forms them (Proposal.__call__, cpab/cpaNd/inference/Proposal.py:58-64).
// Used where a chain set is (re)armed outside the step launches; the step's
// last block forms the same sums from L z vectors k_mcmc_tables prepared.
// with_noise == 0 arms slot 0 with theta itself.  z, th: K*d doubles of LDS each.
__device__ inline void mcmc_arm_theta(const McmcDesc& D, int ch, int64_t step, int with_noise, const double* theta,
                                      double* z, double* th) {
  const int d = D.d, K = D.K;
  double* tp = D.theta_prop + (int64_t)ch * K * d;
  if (!with_noise) {
    for (int i = threadIdx.x; i < d; i += blockDim.x) tp[i] = theta[i];
    return;
  }
  const int hd = (d + 1) / 2;
  for (int idx = threadIdx.x; idx < K * hd; idx += blockDim.x) {
    const int m = idx / hd, t = idx - m * hd;
    mcmc_normal_pair(D, ch, step + m, t, z + m * d);
  }
  __syncthreads();
  // the slot index runs fastest over the lanes, so a wave's loads of L touch
  // a few rows
  for (int idx = threadIdx.x; idx < K * d; idx += blockDim.x) {
    const int i = idx / K, m = idx - i * K;
    th[m * d + i] = mcmc_lz_row(D.L + (int64_t)i * d, z + m * d, i);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < d; i += blockDim.x) {
    double v = theta[i];
    for (int m = 0; m < K; m++) {
      v = v + th[m * d + i];
      tp[m * d + i] = v;
    }
  }
}

// Dispatch order of a chain's pixel blocks for the next step launch, when
// lanes take several pixels each (refill keeps them busy, so a block's
// duration follows its total work): by descending count of rejected Trel
// proposals in slot 0 of the last launch (ties by index).  The longest
// blocks start first and the shortest fill the end of the launch
// (longest-processing-time first; 256 chains one step per launch: +7 %,
// measured on MI355X).  With one pixel per lane a block's duration follows
// its slowest pixel, and ordering by that measured no faster than the
// identity (the corner pixel's block, the longest chain of trips when the
// corner is integrated, is dispatched first anyway), so such launches keep
// the identity and skip the
// lookup.  Only the timing depends on the order: every block stores its
// sums at its pixel block's index.  Above kOrderMax blocks the quadratic
// ranking is skipped and the order is the identity.
constexpr int kOrderMax = 1024;
__device__ inline void mcmc_order(const McmcDesc& D, int ch, double* w) {
  const int n = D.nBlocks;
  int* ord = D.order + (int64_t)ch * n;
  if (n > kOrderMax) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) ord[i] = i;
    return;
  }
  const double* nf = D.partial_nfb + (int64_t)ch * D.K * n;  // slot 0
  for (int i = threadIdx.x; i < n; i += blockDim.x) w[i] = nf[i];
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const double v = w[i];
    int rank = 0;
    for (int j = 0; j < n; j++) rank += (w[j] > v) | ((w[j] == v) & (j < i));
    ord[rank] = i;
  }
}

// Per-cell tables of the first nk armed proposals of every chain, the L z
// vectors the step's last block will need, and the next dispatch order:
// grid (nk + K (+ 1 with an order), chains).
// Avees = B theta' (CpaSpace.theta2Avees, cpab/cpaNd/_CpaSpace.py:366-376),
// each row summed in index order, with B
// read transposed (Bt, d x 6nC, so the lanes of a wave load consecutive
// doubles); sign -1, dt scaling and the per-cell expm of the calc_T_inv prep
// (cpab/cpaNd/_CpaCalcs.py:243-313) into D.tables: Trels, then signed As.
// Blocks nk .. nk+K-1 form L z_{step+j} for j < 2K (two vectors each) from
// the step counter the decision left: the last block of the next step arms
// slot m from vector j0 + m (j0 = steps it decided, 0 after the init launch)
// -- the sums mcmc_arm_theta forms, in the same order.
// The kernel is on the critical path between two step launches and runs on
// few CUs, so it is written for latency: one row per thread (blocks of up to
// 1024 threads) and sixteen loads of B in flight per thread.
__global__ void __launch_bounds__(1024) k_mcmc_tables(McmcDesc D, int nk) {
  extern __shared__ double w[];
  const int m = blockIdx.x, ch = blockIdx.y;
  const int d = D.d, nC = D.nC, R = 6 * nC;
  if (m == nk + D.K) {  // dispatch order of the next step's pixel blocks (when D.order)
    mcmc_order(D, ch, w);
    return;
  }
  if (m >= nk) {
    const int q = m - nk;
    const int64_t step = D.step[ch];
    const int hd = (d + 1) / 2;
    for (int idx = threadIdx.x; idx < 2 * hd; idx += blockDim.x) {
      const int jj = idx / hd, t = idx - jj * hd;
      mcmc_normal_pair(D, ch, step + 2 * q + jj, t, w + jj * d);
    }
    __syncthreads();
    double* lz = D.lz + ((int64_t)ch * 2 * D.K + 2 * q) * d;
    for (int idx = threadIdx.x; idx < 2 * d; idx += blockDim.x) {
      const int i = idx >> 1, jj = idx & 1;
      lz[jj * d + i] = mcmc_lz_row(D.L + (int64_t)i * d, w + jj * d, i);
    }
    return;
  }
  double* th = w;
  double* a = w + d;
  const double* tp = D.theta_prop + ((int64_t)ch * D.K + m) * d;
  for (int i = threadIdx.x; i < d; i += blockDim.x) th[i] = tp[i];
  __syncthreads();
  double* T = D.tables + ((int64_t)ch * D.K + m) * 2 * R;
  for (int r = threadIdx.x; r < R; r += blockDim.x) {
    const double* b = D.Bt + r;
    double acc = 0.0;
    int j = 0;
    for (; j + 16 <= d; j += 16) {
      double v[16];
#pragma unroll
      for (int u = 0; u < 16; u++) v[u] = b[(int64_t)(j + u) * R];
#pragma unroll
      for (int u = 0; u < 16; u++) acc += v[u] * th[j + u];
    }
    for (; j < d; j++) acc += b[(int64_t)j * R] * th[j];
    double v = acc;
    v *= -1;  // calc_T_inv
    a[r] = v;
    T[R + r] = v;
  }
  if (D.J) {
    // lp' = -0.5 (theta'^T J) theta' (ScaleDependentLogPrior.calc_lp, mu = 0):
    // (theta'^T J)_j summed over i in index order, then the products summed
    // in index order, so a proposal's lp' does not depend on its slot
    double* pr = a + R;
    for (int j = threadIdx.x; j < d; j += blockDim.x) {
      double v = 0.0;
      for (int i = 0; i < d; i++) v += th[i] * D.J[(int64_t)i * d + j];
      pr[j] = v * th[j];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      double q = 0.0;
      for (int j = 0; j < d; j++) q += pr[j];
      D.lp_prop[(int64_t)ch * D.K + m] = -0.5 * q;
    }
  }
  __syncthreads();
  for (int c = threadIdx.x; c < nC; c += blockDim.x) {
    double adt[6], t6[6];
#pragma unroll
    for (int k = 0; k < 6; k++) adt[k] = a[c * 6 + k] * D.dt;
    expm_affine2(adt, t6, 12);
#pragma unroll
    for (int k = 0; k < 6; k++) T[c * 6 + k] = t6[k];
  }
}

// out (cols x rows) = in (rows x cols) transposed, both row-major.
__global__ void k_transpose(const double* __restrict__ in, double* __restrict__ out, int rows, int cols) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)rows * cols) return;
  const int64_t r = i / cols, c = i - r * cols;
  out[c * rows + r] = in[i];
}

// One block per chain: arm the proposals of the chain's current step.
__global__ void __launch_bounds__(256) k_mcmc_prep(McmcDesc D, int with_noise) {
  extern __shared__ double w[];
  const int ch = blockIdx.x;
  mcmc_arm_theta(D, ch, D.step[ch], with_noise, D.theta + (int64_t)ch * D.d, w, w + D.K * D.d);
}

// stop = step + n for every chain (the end of a run of n steps).
__global__ void k_mcmc_set_stop(McmcDesc D, int64_t n) {
  const int ch = blockIdx.x * blockDim.x + threadIdx.x;
  if (ch < D.nChains) D.stop[ch] = D.step[ch] + n;
}

// The last block of the chain: decide the evaluated slots (each reduced by
// its own last block) in step order while they stay valid (slot m is valid
// when slots 0..m-1 were accepted), advance, then arm the next K proposals:
// theta'_m = theta + L z_{s} + ... + L z_{s+m}, s the new step counter, from
// the L z vectors k_mcmc_tables prepared for this launch.
__device__ inline void mcmc_finish(const McmcDesc& D, int ch, int64_t step, int kact, int init, double* w,
                                   int* sh) {
  const int K = D.K, d = D.d;
  double* llp = w;
  double* nfs = w + K;
  double* lus = nfs + K;
  double* pxs = lus + K;
  double* lpp = pxs + K;
  for (int m = threadIdx.x; m < kact; m += blockDim.x) {
    const double* r = D.slot_res + ((int64_t)ch * K + m) * 3;
    llp[m] = __hip_atomic_load(r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    nfs[m] = __hip_atomic_load(r + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    pxs[m] = __hip_atomic_load(r + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (D.J) lpp[m] = D.lp_prop[(int64_t)ch * K + m];  // written by the tables launch before this one
    // log of the slot's acceptance uniform (read only if its ll does not improve)
    if (!init) {
      const int64_t s = step + m;
      const Philox u = philox4x32_10((uint32_t)s, (uint32_t)(s >> 32), (uint32_t)ch, SLOT_ACCEPT, D.seed_lo,
                                     D.seed_hi);
      lus[m] = log(u53(u.x0, u.x1));
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double ll = D.ll[ch];
    double lp = D.J ? D.lp[ch] : 0.0;
    int nd = 0, la = -1;
    int64_t nacc = 0;
    double nf_dec = 0.0, nf_all = 0.0, px_all = 0.0;
    for (int m = 0; m < kact; m++) {
      nf_all += nfs[m];
      px_all += pxs[m];
    }
    for (int m = 0; m < kact; m++) {
      bool acc;
      if (init) {
        acc = true;
      } else {
        double lr = llp[m] - ll;
        if (D.J) lr += D.wlp * (lpp[m] - lp);  // Metropolis.py:64-70
        if (lr > 0) {
          acc = true;
        } else {
          const double lu = lus[m];
          acc = D.inverted ? (lu > lr) : (lu < lr);
        }
      }
      nd++;
      nf_dec += nfs[m];
      if (acc) {
        ll = llp[m];
        if (D.J) lp = lpp[m];
        la = m;
        nacc++;
      }
      if (!init && D.trace && step + m < D.trace_len) D.trace[(int64_t)ch * D.trace_len + step + m] = ll;
      if (!acc) break;
    }
    D.ll_prop[ch] = llp[nd - 1];
    D.ll[ch] = ll;
    if (D.J) D.lp[ch] = lp;
    if (!init) {
      D.n_acc[ch] += nacc;
      D.step[ch] = step + nd;
    }
    if (D.nfb) D.nfb[ch] += (unsigned long long)nf_dec;
    const int ab = __hip_atomic_load(D.abort_after + ch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (D.stats) {
      int64_t* st = D.stats + (int64_t)ch * 5;
      st[0] += ab < kact ? ab + 1 : kact;  // slots at or below abort_after all ran in full
      st[1] += (int64_t)nf_all;
      st[2] += 1;
      st[3] += init ? 0 : nd;
      st[4] += llround(px_all);
    }
    // re-arm for the next launch (kernel boundaries order it)
    D.done[ch] = 0;
    D.ready[ch] = 0ull;
    D.abort_after[ch] = kNoAbort;
    sh[0] = la;
    sh[1] = init ? 0 : nd;
  }
  __syncthreads();
  const int la = sh[0], j0 = sh[1];
  double* tp = D.theta_prop + (int64_t)ch * K * d;
  const double* lz = D.lz + ((int64_t)ch * 2 * K + j0) * d;
  for (int i = threadIdx.x; i < d; i += blockDim.x) {
    double v = la >= 0 ? tp[(int64_t)la * d + i] : D.theta[(int64_t)ch * d + i];
    D.theta[(int64_t)ch * d + i] = v;
    int m = 0;
    for (; m + 8 <= K; m += 8) {
      double t[8];
#pragma unroll
      for (int u = 0; u < 8; u++) t[u] = lz[(int64_t)(m + u) * d + i];
#pragma unroll
      for (int u = 0; u < 8; u++) {
        v = v + t[u];
        tp[(int64_t)(m + u) * d + i] = v;
      }
    }
    for (; m < K; m++) {
      v = v + lz[(int64_t)m * d + i];
      tp[(int64_t)m * d + i] = v;
    }
  }
}

// The test of slot j on the all-accepted path (slots 0..j-1 accepted, so the
// state is ll'_{j-1}, or the chain's ll for j = 0): Metropolis.py:64-82.
__device__ inline bool mcmc_accepts(const McmcDesc& D, int ch, int64_t step, int j, double ll_prev, double ll_j) {
  const double lr = ll_j - ll_prev;
  if (lr > 0) return true;
  const int64_t s = step + j;
  const Philox u = philox4x32_10((uint32_t)s, (uint32_t)(s >> 32), (uint32_t)ch, SLOT_ACCEPT, D.seed_lo, D.seed_hi);
  const double lu = log(u53(u.x0, u.x1));
  return D.inverted ? (lu > lr) : (lu < lr);
}

// x^n for x >= 1, rounded up (each product inflated past its rounding error).
__device__ inline double pow_up(double x, int n) {
  constexpr double up = 1.0 + 0x1p-40;
  double r = 1.0;
  for (; n > 0; n >>= 1) {
    if (n & 1) r = r * x * up;
    x = x * x * up;
  }
  return r;
}

// Whether the pixel at the domain corner (0, 0) can be left out of the
// integration because its sample is src(0, 0) bit for bit, whatever path
// the trips take.  With zero velocity on the boundary the corner's velocity
// is zero up to the rounding of B theta, so its point stays within ~1e-12 px
// of (0, 0) -- yet it falls back to the RK2 substeps at nearly every step
// (~1 100 dependent trips where other pixels need ~104), which would bound
// the launch.
//  * The path: every point it visits lies within rho of (0, 0), by
//    induction over the trips.  A point within rho (below half the
//    increments) looks up a cell of box (0, 0) that touches the corner: the
//    box itself (type II); of type I's triangles 0 (top) or 3 (left) -- inside
//    the box 1 - x > y rules out 1 and 2, outside the left rule gives 0 or 3
//    and the top rule 0 (cell2, calc_2Dto2D.cu:38-153).  With
//    gT, tT, a, b the largest |T_r0| + |T_r1|, |T_r2|, |A_r0| + |A_r1|,
//    |A_r2| over those cells' rows: a Trel step maps rho to gT rho + tT; an
//    RK2 substep (v = A[p;1], m = p + h/2 v, p' = p + h A[m;1]) to
//    rho (1 + h a (1 + h/2 a)) + h b (1 + h/2 a); each factor inflated by
//    2^-40 for the rounding of a trip's ten IEEE operations, 2^-1000 added
//    per step for underflow.  Over nT steps: rho_T <= tau nT g^nT, g >= 1
//    the larger growth of a step (one Trel step, or nS substeps), tau its
//    larger offset.
//  * The sample (sample2, FP32 weights): for |x|, |y| <= rho <= 2^-27 the
//    weights are exactly 0 or 1, or below 2^-25 (x in [0, rho): fx = x,
//    1 - fx rounds to 1; x in [-rho, 0): x - floor(x) rounds to 1, and the
//    pixel at -1 is outside, zero), so each of the sample's two additions is
//    v00 (or 0) plus a term of at most rho M (1 + 2^-20), M the largest
//    |v| of the four pixels; below a quarter of v00's FP32 ulp that term
//    rounds away.  So w = v00 when v00 is a normal number and
//    4 rho M (1 + 2^-20) <= ulp(v00).
// The path depends on the slot's tables, the sample on src: every lane of
// the wave computes the same answer.  NaN or infinite tables give false.
template <int TESS>
__device__ inline bool corner_sample_is_src00(const double* Ttab, const double* Atab, const McmcDesc& D) {
  constexpr double up = 1.0 + 0x1p-40;
  const int ncell = TESS == TESS_I ? 2 : 1;
  double gT = 0.0, tT = 0.0, a = 0.0, b = 0.0, probe = 0.0;
  for (int i = 0; i < ncell; i++) {
    const int c = 3 * i;  // cells 0 and 3 (type I), 0 (type II)
#pragma unroll
    for (int r = 0; r < 2; r++) {
      const double* T = Ttab + c * 6 + r * 3;
      const double* A = Atab + c * 6 + r * 3;
      const double g1 = fabs(T[0]) + fabs(T[1]), t1 = fabs(T[2]);
      const double a1 = fabs(A[0]) + fabs(A[1]), b1 = fabs(A[2]);
      gT = fmax(gT, g1);
      tT = fmax(tT, t1);
      a = fmax(a, a1);
      b = fmax(b, b1);
      probe += g1 + t1 + a1 + b1;  // (fmax drops a NaN; the sum keeps it)
    }
  }
  if (!(probe <= 1e300) || !(D.h > 0.0) || D.nT < 0 || D.nS < 0) return false;
  gT *= up;
  tT *= up;
  a *= up;
  b *= up;
  const double h = D.h, ga = (1.0 + 0.5 * h * a) * up;
  const double gs = (1.0 + h * a * ga) * up * up, ts = h * b * ga * up * up;  // one substep
  const double gsn = pow_up(gs, D.nS);                                         // nS substeps
  const double tsn = ts * (double)D.nS * gsn * up;
  const double g = fmax(fmax(gT * up, gsn), 1.0);
  const double tau = fmax(tT * up, tsn) + 0x1p-1000;
  const double rho = tau * (double)D.nT * pow_up(g, D.nT) * up;
  if (!(rho <= 0x1p-27) || !(rho < 0.5 * fmin(D.g.inc_x, D.g.inc_y))) return false;
  const int W = D.W, H = D.H;
  const float v00 = D.src[0];
  const float v01 = W > 1 ? D.src[1] : 0.f;
  const float v10 = H > 1 ? D.src[W] : 0.f;
  const float v11 = (W > 1 && H > 1) ? D.src[W + 1] : 0.f;
  // (a NaN or infinite neighbour would reach the sample through its tiny weight)
  if (!(isfinite(v00) && isfinite(v01) && isfinite(v10) && isfinite(v11))) return false;
  const double m = fmax(fmax(fabs((double)v00), fabs((double)v01)), fmax(fabs((double)v10), fabs((double)v11)));
  if (!(fabs((double)v00) >= 0x1p-126)) return false;
  int e;
  (void)frexp((double)v00, &e);  // |v00| in [2^(e-1), 2^e): its FP32 ulp is 2^(e-24)
  return 4.0 * rho * m * (1.0 + 0x1p-20) <= ldexp(1.0, e - 24);
}

// One launch evaluates up to K proposals of every chain: grid (K * nBlocks,
// nChains), or (nChains, K * nBlocks) for several chains.  Dispatch order:
// the first pixel block of every slot (it holds the domain corner, whose
// pixel, when it is integrated, falls back at most steps: the longest chain
// of trips), then the
// slots one after another; with several chains, chain-fastest at each
// place, so every chain's corner blocks start first.  Slot m is accepted on
// the all-accepted path or not as soon as slots m-1 and m are summed, so a
// rejection is usually known while later slots have not started: their
// blocks then skip their pixels, and waves already running stop at the next
// of a few step checkpoints (early abort; every slot at or below the first
// rejection always runs in full, so the chain does not depend on the timing
// -- only the counts of work done do).  A wave's 64 pixels are an
// 8 x 8 tile when the image allows: neighbouring pixels flow alike and cross
// facets at about the same steps, so the lanes of a wave need similar
// numbers of trips (one chain, 20 slots: +11 % over 64 pixels of a row,
// measured on MI355X).  init == 1: evaluate the armed (noise-free) slot 0 as
// the current state.  A chain whose run has ended (step >= stop) is left as
// it is.  Every block of a live chain, evaluating or not, counts its
// arrival, so the last arrival is the only block that touches the chain's
// state.
template <int TESS, bool POW2, int BLOCK, bool FAST, bool SWAP>
__global__ void __launch_bounds__(BLOCK) k_mcmc_step(McmcDesc D, int init) {
  extern __shared__ double lds[];
  __shared__ double red[BLOCK / 64], rednf[BLOCK / 64], redpx[BLOCK / 64];
  __shared__ int last, slot_last;
  __shared__ int fin[2];
  // SWAP: grid (chains, K x nBlocks), so the hardware, which dispatches x
  // fastest, starts the first pixel block of every chain's slots first
  const int ch = SWAP ? blockIdx.x : blockIdx.y;
  const int bi = SWAP ? (int)blockIdx.y : (int)blockIdx.x;
  const int K = D.K;
  const int nB = D.nBlocks;
  int slot, pos;  // pos: the block's place in its slot's sequence of pixel blocks
  if (bi < K) {
    slot = bi;
    pos = 0;
  } else {
    const int r = bi - K;
    slot = r / (nB - 1);
    pos = 1 + r - slot * (nB - 1);
  }
  const int bx = D.order ? D.order[(int64_t)ch * nB + pos] : pos;
  const int nC = D.nC;
  const int64_t step = D.step[ch];
  int kact = 1;
  if (!init) {
    const int64_t left = D.stop[ch] - step;
    if (left <= 0) return;  // grid-uniform per chain: no block of this chain counts
    kact = left < K ? (int)left : K;
  }
  // whether this block integrates: read once by thread 0 (abort_after may
  // change under the block) so the decision, and the barriers after it, are
  // block-uniform
  if (threadIdx.x == 0)
    fin[0] = slot < kact &&
             (init || slot <= __hip_atomic_load(D.abort_after + ch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
  __syncthreads();
  const bool run = fin[0] != 0;
  double acc = 0.0, pxw = 0.0;  // SSD, pixels' worth integrated
  int nfb = 0;
  if (run) {
    const double* tab = D.tables + ((int64_t)ch * K + slot) * 2 * nC * 6;
    for (int i = threadIdx.x; i < 2 * nC * 6; i += blockDim.x) lds[i] = tab[i];
    __syncthreads();
    const double* Ttab = lds;
    const double* Atab = lds + nC * 6;
    const Geom& g = D.g;
    const int64_t nPts = (int64_t)D.H * D.W;
    const int64_t wave = (int64_t)bx * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int64_t begin = wave * (int64_t)D.ptsPerWave;
    const int64_t end = begin + D.ptsPerWave < nPts ? begin + D.ptsPerWave : nPts;
    const int nT = D.nT, nS = D.nS;
    const double h = D.h;
    if (begin < nPts) {
      // The domain corner's sample, fixed by a bound on its path: its lane
      // takes no trips (its rejected Trel proposals are then not counted).
      // Its lane takes no further pixel either way, so the other lanes take
      // the same pixels, and every lane's sum the same terms in the same
      // order, whether or not the corner is integrated.
      const bool corner_fixed = begin == 0 && !D.integrate_corner && corner_sample_is_src00<TESS>(Ttab, Atab, D);
      bool corner_lane = false;
      // Abort checkpoints: with one pixel per lane every lane runs steps 0..nT
      // once, so the wave can stop at fixed steps when its slot lies above a
      // rejection found meanwhile (no cost per trip: the trip loop's bound
      // is the next checkpoint instead of nT).
      const int nCk = (!init && kact > 1 && D.ptsPerWave == 64 && nT > D.poll) ? D.poll : 0;
      const int tstep = nCk > 0 ? nT / (nCk + 1) : nT;  // steps between checkpoints
      int cp = 0, tb = tstep;                             // checkpoint index, trip loop bound
      int64_t next = begin;
      double p[2];
      int c = 0, t = 0, k = 0;
      const double* row = Ttab;
      int64_t my = 0, pix = 0;  // work index, pixel index
      bool active = false;
      for (;;) {
        if (active && t >= nT) {
          const float w = sample2(D.src, D.H, D.W, p[0], p[1]);
          const double e = (double)w - (double)D.dst[pix];
          acc += e * e;
          pxw += 1.0;
          active = false;
        }
        const uint64_t idle = ballot(!active && !corner_lane);
        if (idle != 0 && next < end) {
          if (!active && !corner_lane) {
            const int64_t cand = next + rank_in(idle);
            if (cand < end) {
              my = cand;
              active = true;
              pix = (int64_t)D.pixmap[my];
              p[0] = (double)(pix % D.W);
              p[1] = (double)(pix / D.W);
              c = cell2<TESS, POW2, FAST, !SWAP>(p, g);
              row = Ttab + mul24(c, 6);
              t = 0;
              k = 0;
              if (pix == 0) {
                corner_lane = true;
                if (corner_fixed) {  // its sample: src(0, 0) (corner_sample_is_src00)
                  const double e = (double)D.src[0] - (double)D.dst[0];
                  acc += e * e;
                  pxw += 1.0;
                  active = false;
                }
              }
            }
          }
          next += __popcll(idle);
        }
        if (ballot(active) == 0) break;
        if (cp < nCk && ballot(active && t >= tb) != 0) {  // a lane reached checkpoint cp
          const int ab = __builtin_amdgcn_readfirstlane(
              __hip_atomic_load(D.abort_after + ch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
          if (slot > ab) {
            if (active) pxw += (double)t / (double)nT;  // the wave stops: its slot is not needed
            break;
          }
          do {
            cp++;
            tb = cp < nCk ? tb + tstep : nT;
          } while (cp < nCk && ballot(active && t >= tb) != 0);
        }
        if (!active || t >= nT) continue;
        if (__popcll(__builtin_amdgcn_read_exec()) == 1) {  // one lane left: it loops on its own
          lone_trips<2, TESS, POW2>(p, c, k, t, tb, row, nfb, Ttab, Atab, h, nS, g);
          continue;
        }
        for (;;) {  // two trips per iteration, as in integrate_range
          if (trip<2, TESS, POW2, FAST, !SWAP>(p, c, k, row, nfb, Ttab, Atab, h, nS, g)) t++;
          if (ballot(t >= tb) != 0) break;
          if (trip<2, TESS, POW2, FAST, !SWAP>(p, c, k, row, nfb, Ttab, Atab, h, nS, g)) t++;
          if (ballot(t >= tb) != 0) break;
        }
      }
    }
  }
  // Block sums (fixed order), published with release.  Completion is
  // counted in two levels: every block on its slot's counter (release); the
  // slot's last block acquires (a fence), sums the slot's block sums and
  // counts on the chain's counter (release); the chain's last block
  // acquires and decides.  Only those K + 1 blocks pay an acquire: at agent
  // scope it invalidates the XCD's L2 (buffer_inv sc1), which per block cost
  // ~0.1 ms per launch at 20 slots x 256 blocks on MI355X.  The release
  // sequences of the counters and the fences order every sum before its
  // reader (C++ fence rules).
  double nf = (double)nfb;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    acc += __shfl_xor(acc, o, 64);
    nf += __shfl_xor(nf, o, 64);
    pxw += __shfl_xor(pxw, o, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    red[threadIdx.x >> 6] = acc;
    rednf[threadIdx.x >> 6] = nf;
    redpx[threadIdx.x >> 6] = pxw;
  }
  __syncthreads();
  unsigned int* ds = D.done_slot + ((int64_t)ch * K + slot) * DONE_STRIDE;
  if (threadIdx.x == 0) {
    if (slot < kact) {  // a block that did not start stores zeros
      double sum = 0.0, sumnf = 0.0, px = 0.0;
      for (int i = 0; i < (int)(blockDim.x >> 6); i++) {
        sum += red[i];
        sumnf += rednf[i];
        px += redpx[i];
      }
      const int64_t at = ((int64_t)ch * K + slot) * nB + bx;
      __hip_atomic_store(D.partial + at, sum, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(D.partial_nfb + at, sumnf, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(D.partial_px + at, px, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    const unsigned int prev = __hip_atomic_fetch_add(ds, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    slot_last = prev == (unsigned int)(D.nBlocks - 1);
    if (slot_last) {  // the slot's last block
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      __hip_atomic_store(ds, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);  // no other arrival this launch
    }
  }
  __syncthreads();
  if (!slot_last) return;  // block-uniform
  if (slot < kact && threadIdx.x < 64) {  // the first wave sums the slot's block sums, lane-strided
    const int lane = threadIdx.x;
    const double* ps = D.partial + ((int64_t)ch * K + slot) * nB;
    const double* pn = D.partial_nfb + ((int64_t)ch * K + slot) * nB;
    const double* pp = D.partial_px + ((int64_t)ch * K + slot) * nB;
    double s = 0.0, f = 0.0, x = 0.0;
    for (int i = lane; i < nB; i += 64) {
      s += __hip_atomic_load(ps + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      f += __hip_atomic_load(pn + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      x += __hip_atomic_load(pp + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      s += __shfl_xor(s, o, 64);
      f += __shfl_xor(f, o, 64);
      x += __shfl_xor(x, o, 64);
    }
    if (lane == 0) {  // thread 0: its releases below publish these
      double l = s * (-0.5 / (D.sigma * D.sigma));
      if (D.use_ave_ll) l = l / (double)((int64_t)D.H * D.W);
      double* r = D.slot_res + ((int64_t)ch * K + slot) * 3;
      __hip_atomic_store(r, l, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(r + 1, f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(r + 2, x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (!init && kact > 1) {
        // Publish the slot, then test every slot j in {slot, slot + 1} whose
        // own sum and predecessor's are both published: of two neighbours
        // the later to publish sees both.  A rejection stops the slots above
        // it from starting.  A slot that did not run in full has a partial
        // sum, but it lies above the first rejection, which is always found
        // (its slot and the one before it run in full), so its test can only
        // lower the bound to a value at or above the first rejection.
        const unsigned long long mask =
            __hip_atomic_fetch_or(D.ready + ch, 1ull << slot, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) |
            (1ull << slot);
        for (int j = slot; j <= slot + 1 && j < kact; j++) {
          if (!((mask >> j) & 1ull) || (j > 0 && !((mask >> (j - 1)) & 1ull))) continue;
          const double lj = __hip_atomic_load(D.slot_res + ((int64_t)ch * K + j) * 3, __ATOMIC_RELAXED,
                                              __HIP_MEMORY_SCOPE_AGENT);
          const double lp = j == 0 ? D.ll[ch]
                                   : __hip_atomic_load(D.slot_res + ((int64_t)ch * K + j - 1) * 3, __ATOMIC_RELAXED,
                                                       __HIP_MEMORY_SCOPE_AGENT);
          if (!mcmc_accepts(D, ch, step, j, lp, lj))
            __hip_atomic_fetch_min(D.abort_after + ch, j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
      }
    }
  }
  if (threadIdx.x == 0) {
    const unsigned int p2 = __hip_atomic_fetch_add(D.done + ch, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    last = p2 == (unsigned int)(K - 1);
    if (last) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");  // the chain's last block
  }
  __syncthreads();
  if (last) mcmc_finish(D, ch, step, kact, init, lds, fin);  // block-uniform branch
}

}  // namespace cpab
