// Many independent calc_T problems in one persistent launch.
//
// Each instance has its own tessellation (cell counts, increments), its own
// tables (Trels then signed As, staged per instance in one table buffer) and
// its own point range in one concatenated point buffer.  The host cuts every
// instance into tiles of at most `tile` points (in instance order) and the
// kernel's blocks take `chunk` consecutive tiles at a time from a queue (one
// atomic per chunk; a block leaves when the queue is exhausted, so every
// block ends).  A block keeps the tables of the instance it is working on in
// LDS and reloads them only when a tile of another instance comes; the waves
// of a block split the tile's points and integrate them with the lane-skewed
// loop of k_integrate.  Per point the arithmetic is exactly calc_T's, so a
// point gives the same bits batched or alone.
#pragma once
#include "cpab_kernels.h"

#pragma clang fp contract(off)

namespace cpab {

struct BatchInst {
  Geom g;
  int nC;
  int pad;
  int64_t tab;   // offset (doubles) of the instance's tables: Trels (nC x NE), then signed As
  int64_t pts;   // offset (points) of the instance's first point
  int64_t nPts;
  int64_t cell;  // offset of the instance's first cell in the concatenated As input
};

struct BatchTile {
  int inst;
  int pad;
  int64_t begin, end;  // point range, absolute indices into the point buffer
};

// Signed As and Trels of every cell of every instance, into the table buffer.
// One lane per cell of the concatenated input As (cells in instance order).
template <int DIM>
__global__ void k_batch_prep(const BatchInst* __restrict__ insts, int nInst, const double* __restrict__ As,
                             int64_t nCells, int sign, double dt, double* __restrict__ tables, int p) {
  constexpr int NE = DIM * (DIM + 1);
  const int64_t cg = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (cg >= nCells) return;
  int lo = 0, hi = nInst - 1;  // the instance holding cell cg
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (insts[mid].cell <= cg) lo = mid;
    else hi = mid - 1;
  }
  const BatchInst& I = insts[lo];
  const int64_t c = cg - I.cell;
  double sA[NE], T[NE];
  prep_cell<DIM>(As + cg * NE, sign, dt, sA, T, p);
  double* Tt = tables + I.tab;
#pragma unroll
  for (int k = 0; k < NE; k++) {
    Tt[c * NE + k] = T[k];
    Tt[(int64_t)I.nC * NE + c * NE + k] = sA[k];
  }
}

__global__ void k_batch_reset(unsigned int* queue) {
  if (threadIdx.x == 0) *queue = 0u;
}

// Processing order of a tile's points by position: a counting sort of the
// tile's point indices over a grid of buckets on the domain (nb per axis, a
// power of two up to 1 024 in 1D, 32 in 2D, 8 in 3D, at least 16 points a
// bucket on average), buckets in row order, built in LDS by the whole block.
// Neighbouring points flow alike and cross facets at about the same steps,
// so a wave's lanes then need closer numbers of trips than for points in
// input order.  Within a bucket the order follows the LDS atomics, which may
// differ from run to run: that changes which lane integrates a point and
// when, never the point's result.
constexpr int kBucketTileMax = 16384;  // the most points of a sorted tile (16-bit indices)
constexpr int kBucketMax = 1024;       // buckets

template <int DIM>
__device__ __forceinline__ int bucket_of(const double* p, const double* sc, int nb) {
  // (NaN coordinates go to a border bucket: fmin/fmax return the other operand)
  int b = 0;
#pragma unroll
  for (int d = DIM - 1; d >= 0; d--) b = b * nb + (int)fmax(0.0, fmin(p[d] * sc[d], (double)(nb - 1)));
  return b;
}

template <int DIM>
__device__ void tile_bucket_order(const double* __restrict__ pos0, int64_t begin, int n, const Geom& g,
                                  unsigned int* hist, uint16_t* idx) {
  constexpr int NBMAX = DIM == 1 ? 1024 : (DIM == 2 ? 32 : 8);
  int nb = 1, nbk = 1;
  while (nb < NBMAX && 16 * (nbk << DIM) <= n) {
    nb <<= 1;
    nbk <<= DIM;
  }
  double sc[DIM];
  sc[0] = (double)nb / ((double)g.nC0 * g.inc_x);
  if (DIM > 1) sc[DIM > 1 ? 1 : 0] = (double)nb / ((double)g.nC1 * g.inc_y);
  if (DIM > 2) sc[DIM > 2 ? 2 : 0] = (double)nb / ((double)g.nC2 * g.inc_z);
  __syncthreads();  // every wave is done with the previous tile's order
  for (int i = threadIdx.x; i < nbk; i += blockDim.x) hist[i] = 0u;
  __syncthreads();
  for (int j = threadIdx.x; j < n; j += blockDim.x) {
    double p[DIM];
    load_pt<DIM>(pos0, begin + j, p);
    atomicAdd(&hist[bucket_of<DIM>(p, sc, nb)], 1u);
  }
  __syncthreads();
  if (threadIdx.x < 64) {  // exclusive scan by the first wave, each lane a run of buckets
    const int per = (nbk + 63) >> 6;
    const int lo = (int)threadIdx.x * per;
    const int hi = lo + per < nbk ? lo + per : nbk;
    unsigned int sum = 0u;
    for (int i = lo; i < hi; i++) sum += hist[i];
    unsigned int x = sum;
    for (int d = 1; d < 64; d <<= 1) {
      const unsigned int y = __shfl_up(x, d, 64);
      if ((int)threadIdx.x >= d) x += y;
    }
    unsigned int run = x - sum;
    for (int i = lo; i < hi; i++) {
      const unsigned int c = hist[i];
      hist[i] = run;
      run += c;
    }
  }
  __syncthreads();
  for (int j = threadIdx.x; j < n; j += blockDim.x) {
    double p[DIM];
    load_pt<DIM>(pos0, begin + j, p);
    idx[atomicAdd(&hist[bucket_of<DIM>(p, sc, nb)], 1u)] = (uint16_t)j;
  }
  __syncthreads();
}

template <int DIM, int TESS, bool POW2>
__global__ void __launch_bounds__(1024)
    k_integrate_batch(const BatchInst* __restrict__ insts, const BatchTile* __restrict__ tiles, int nTiles, int chunk,
                      unsigned int* __restrict__ queue, const double* __restrict__ tables,
                      const double* __restrict__ pos0, double* __restrict__ out, double h, int nT, int nS,
                      int32_t* __restrict__ nfb_out, int sort_off) {
  constexpr int NE = DIM * (DIM + 1);
  extern __shared__ double lds[];
  __shared__ int s_t0;
  // sort_off > 0: the tiles' processing order by position, in LDS after
  // the tables (sort_off doubles in): bucket counters, then indices
  unsigned int* hist = reinterpret_cast<unsigned int*>(lds + sort_off);
  uint16_t* idx = reinterpret_cast<uint16_t*>(hist + kBucketMax);
  const int nw = blockDim.x >> 6, w = threadIdx.x >> 6;
  int cur = -1;  // instance whose tables are in LDS (block-uniform)
  for (;;) {
    __syncthreads();  // every wave is done with the previous chunk's s_t0
    if (threadIdx.x == 0) s_t0 = (int)atomicAdd(queue, (unsigned int)chunk);
    __syncthreads();
    const int t0 = s_t0;
    if (t0 >= nTiles) break;
    const int t1 = t0 + chunk < nTiles ? t0 + chunk : nTiles;
    for (int t = t0; t < t1; t++) {
      const BatchTile tl = tiles[t];
      const int nC = insts[tl.inst].nC;
      if (tl.inst != cur) {
        const double* src = tables + insts[tl.inst].tab;
        __syncthreads();  // the previous tables are no longer read
        for (int i = threadIdx.x; i < 2 * nC * NE; i += blockDim.x) lds[i] = src[i];
        __syncthreads();
        cur = tl.inst;
      }
      const Geom g = insts[tl.inst].g;
      const int64_t n = tl.end - tl.begin;
      const uint16_t* lord = nullptr;
      if (sort_off > 0 && n <= kBucketTileMax) {  // (block-uniform)
        tile_bucket_order<DIM>(pos0, tl.begin, (int)n, g, hist, idx);
        lord = idx;
      }
      const int64_t per = (n + nw - 1) / nw;
      const int64_t b = tl.begin + (int64_t)w * per;
      const int64_t e = b + per < tl.end ? b + per : tl.end;
      if (b < e)
        integrate_range<DIM, TESS, POW2, 0, true>(b, e, pos0, out, lds, lds + nC * NE, h, nT, nS, 0, g, nfb_out, 0,
                                                  lord, tl.begin);
    }
  }
}

}  // namespace cpab

namespace cpab {

// A tile of one request of a served mix: its own tables, points and
// geometry, so a launch can take tiles of any set of requests.
struct WorkItem {
  Geom g;
  int nC;
  int pad;
  const double* trels;   // (nC, NE)
  const double* as;      // (nC, NE) signed As
  const double* pos0;
  double* out;
  int64_t begin, end;    // point range within the request's arrays
};

template <int DIM, int TESS, bool POW2>
__global__ void __launch_bounds__(1024)
    k_integrate_items(const WorkItem* __restrict__ items, int nItems, int chunk, unsigned int* __restrict__ queue,
                      double h, int nT, int nS, int sort_off) {
  constexpr int NE = DIM * (DIM + 1);
  extern __shared__ double lds[];
  __shared__ int s_t0;
  // sort_off > 0: each item's points in an order by position (tile_bucket_order)
  unsigned int* hist = reinterpret_cast<unsigned int*>(lds + sort_off);
  uint16_t* idx = reinterpret_cast<uint16_t*>(hist + kBucketMax);
  const int nw = blockDim.x >> 6, w = threadIdx.x >> 6;
  const double* cur = nullptr;  // Trels in LDS (block-uniform)
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0) s_t0 = (int)atomicAdd(queue, (unsigned int)chunk);
    __syncthreads();
    const int t0 = s_t0;
    if (t0 >= nItems) break;
    const int t1 = t0 + chunk < nItems ? t0 + chunk : nItems;
    for (int t = t0; t < t1; t++) {
      const WorkItem& it = items[t];
      const int nC = it.nC;
      const double* tr = it.trels;
      if (tr != cur) {
        const double* sa = it.as;
        __syncthreads();
        for (int i = threadIdx.x; i < nC * NE; i += blockDim.x) {
          lds[i] = tr[i];
          lds[nC * NE + i] = sa[i];
        }
        __syncthreads();
        cur = tr;
      }
      const Geom g = it.g;
      const int64_t n = it.end - it.begin;
      const uint16_t* lord = nullptr;
      if (sort_off > 0 && n <= kBucketTileMax) {  // (block-uniform)
        tile_bucket_order<DIM>(it.pos0, it.begin, (int)n, g, hist, idx);
        lord = idx;
      }
      const int64_t per = (n + nw - 1) / nw;
      const int64_t b = it.begin + (int64_t)w * per;
      const int64_t e = b + per < it.end ? b + per : it.end;
      if (b < e)
        integrate_range<DIM, TESS, POW2, 0, false>(b, e, it.pos0, it.out, lds, lds + nC * NE, h, nT, nS, 0, g, nullptr,
                                                   0, lord, it.begin);
    }
  }
}

}  // namespace cpab
