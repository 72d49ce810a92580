// Kernels of the particle path. One lane per particle for advection and
// location, one workgroup per run of cells for redistribution, one lane per
// (cell, vertex) for the projection sums and one lane per DoF for the node
// quotient. All particle kernels are grid-stride loops that read the current
// particle count from device memory, so a step never waits on the host.
#pragma once

#include "pfem2gpu_device.h"

namespace pfem2gpu {

// ---------------------------------------------------------------- utilities

__global__ void k_fill_int(int* __restrict__ a, long n, int value) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) a[i] = value;
}

// dV = V - V_old per DoF (src/pfem2particlehandler.cpp:120-124)
template <int D>
__global__ void k_sub(int ndofs, DField V, DField Vold, double* __restrict__ d0, double* __restrict__ d1,
                      double* __restrict__ d2) {
  double* out[3] = {d0, d1, d2};
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < ndofs; j += gridDim.x * blockDim.x) {
#pragma unroll
    for (int k = 0; k < D; ++k) out[k][j] = V.c[k][j] - Vold.c[k][j];
  }
}

// Advecting field in cell-major order: Vc[c][i][k] = V_k[dof(c, i)], so the
// move kernel reads a cell's nodal values with one contiguous load.
template <int D>
__global__ void k_cellfield(int nc, const int* __restrict__ cd, DField V, double* __restrict__ Vc) {
  constexpr int NV = 1 << D;
  const long total = (long)nc * NV;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
    const int dof = cd[t];
#pragma unroll
    for (int k = 0; k < D; ++k) Vc[t * D + k] = V.c[k][dof];
  }
}

// ------------------------------------------------------------- prefix scans
// Exclusive scan of in[0, n) into out[0, n], out[n] = total. Three kernels:
// per-chunk sums, a one-workgroup scan of the chunk sums, per-chunk scans.
constexpr int kScanItems = 8;
constexpr int kScanChunk = kBlock * kScanItems;

__device__ __forceinline__ int block_exclusive_scan(int v, int* lds /* >= 4 ints */, int* total) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  int incl = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int t = __shfl_up(incl, off, 64);
    if (lane >= off) incl += t;
  }
  if (lane == 63) lds[wave] = incl;
  __syncthreads();
  int wbase = 0, tot = 0;
  const int nw = blockDim.x >> 6;
  for (int w = 0; w < nw; ++w) {
    const int s = lds[w];
    if (w < wave) wbase += s;
    tot += s;
  }
  __syncthreads();
  *total = tot;
  return wbase + incl - v;
}

__global__ __launch_bounds__(kBlock) void k_scan_reduce(const int* __restrict__ in, int n, int* __restrict__ blk) {
  __shared__ int lds[kBlock / 64];
  const int b = blockIdx.x;
  const int base = b * kScanChunk + threadIdx.x * kScanItems;
  int s = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i)
    if (base + i < n) s += in[base + i];
  int tot;
  block_exclusive_scan(s, lds, &tot);
  if (threadIdx.x == 0) blk[b] = tot;
}

__global__ __launch_bounds__(kBlock) void k_scan_blocks(int* __restrict__ blk, int nb, int* __restrict__ out_total) {
  __shared__ int lds[kBlock / 64];
  int carry = 0;
  for (int c0 = 0; c0 < nb; c0 += kBlock) {
    const int i = c0 + threadIdx.x;
    const int v = (i < nb) ? blk[i] : 0;
    int tot;
    const int ex = block_exclusive_scan(v, lds, &tot);
    if (i < nb) blk[i] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0) *out_total = carry;
}

__global__ __launch_bounds__(kBlock) void k_scan_down(const int* __restrict__ in, int n, const int* __restrict__ blk,
                                                     int* __restrict__ out) {
  __shared__ int lds[kBlock / 64];
  const int b = blockIdx.x;
  const int base = b * kScanChunk + threadIdx.x * kScanItems;
  int vals[kScanItems];
  int s = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    vals[i] = (base + i < n) ? in[base + i] : 0;
    s += vals[i];
  }
  int tot;
  int run = blk[b] + block_exclusive_scan(s, lds, &tot);
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    if (base + i < n) out[base + i] = run;
    run += vals[i];
  }
}

// -------------------------------------------------------------- move passes
// Adds, for every distinct key >= 0 held by the lanes of the wavefront, the
// number of lanes holding it to arr[key] (one atomic per key) and returns each
// lane's old value plus its rank among the lanes with the same key. Must be
// called by all 64 lanes.
__device__ __forceinline__ int wave_group_add(int* __restrict__ arr, int key) {
  const int lane = threadIdx.x & 63;
  const unsigned long long below = (1ULL << lane) - 1ULL;
  int result = 0;
  unsigned long long pending = __ballot(key >= 0);
  while (pending) {
    const int leader = __ffsll((unsigned long long)pending) - 1;
    const int lk = __shfl(key, leader, 64);
    const unsigned long long grp = __ballot(key == lk) & pending;
    int base = 0;
    if (lane == leader) base = atomicAdd(&arr[lk], __popcll(grp));
    base = __shfl(base, leader, 64);
    if ((grp >> lane) & 1ULL) result = base + __popcll(grp & below);
    pending &= ~grp;
  }
  return result;
}

// Spill binning: particles that did not fit their chunk's bucket are counted
// per chunk and copied, grouped by chunk, into the second spill region.
__global__ void k_spill_count(const int* __restrict__ spill_n, long spill_base, const int* __restrict__ tcell,
                              int cshift, int* __restrict__ scount) {
  const int n = *spill_n;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    atomicAdd(&scount[tcell[spill_base + i] >> cshift], 1);
}

template <int D>
__global__ void k_spill_scatter(const int* __restrict__ spill_n, long spill_base, long spill2_base, int cshift,
                                const int* __restrict__ sstart, int* __restrict__ sfill, DTmp T) {
  const int n = *spill_n;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const long src = spill_base + i;
    const int c = T.cell[src];
    const int ch = c >> cshift;
    const long dst = spill2_base + sstart[ch] + atomicAdd(&sfill[ch], 1);
#pragma unroll
    for (int k = 0; k < D; ++k) {
      T.x[k][dst] = T.x[k][src];
      T.v[k][dst] = T.v[k][src];
    }
    T.cell[dst] = c;
    if (T.src) T.src[dst] = T.src[src];
  }
}

// Advances every particle through the n_sub Euler substeps with location
// (move_particles, src/pfem2particlehandler.cpp:141-171) and appends it to the
// bucket of its destination chunk of 2^cshift cells in T. `advect` = 0 turns
// the pass into a pure re-binning (used by insertion). A pending velocity
// correction is applied from the pre-move cell and reference position, as
// correct_particle_velocity() precedes move_particles().
//
// Each wavefront owns a run of `run` consecutive particles (sorted by
// cell, so lanes share cells and their cached data comes from L1/L2) and
// processes it as a work queue: every lane runs its particle's Euler
// substeps until it leaves its cell or finishes; the lanes that left their
// cell then run the neighbour walk together; lanes whose particle is done
// take the next particle of the run. The walk thus executes with many lanes
// active instead of once per substep for a few.
template <int D>
__global__ __launch_bounds__(kBlock) void k_move(int advect, DMesh m, const double* __restrict__ Vc,
                                                const int* __restrict__ n_ptr, int n_add, DParts P, DField dV,
                                                int use_dv, int cshift, int capc, int* __restrict__ bcount,
                                                long spill_base, int* __restrict__ spill_n, DTmp T, int run) {
  const int n = *n_ptr + n_add;
  const int lane = threadIdx.x & 63;
  const unsigned long long below = (1ULL << lane) - 1ULL;
  const long gw = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const long nw = ((long)gridDim.x * blockDim.x) >> 6;
  for (long wb = gw * run; wb < n; wb += nw * run) {
    const int wend = (int)min((long)n, wb + run);
    int next = (int)wb + 64;  // first particle of the run not yet taken
    int p = (int)wb + lane;
    bool fresh = true;        // lane must load particle p
    CellCache<D> cc;
    double x[D], xi[D], v[D];
    double dt_sub = 0.0;
    int n_sub = 0, s = 0;
    bool walking = false, dead = false;
    for (;;) {
      if (fresh) {
        fresh = false;
        dead = false;
        walking = false;
        s = 0;
        if (p < wend) {
          const int c = P.cell[p];
          if (c < 0) {
            dead = true;
          } else {
            load_cell<D>(m, Vc, c, cc, advect != 0);
            const Prm& pr = m.prm[inst_of(m, c)];
            dt_sub = pr.dt_sub;
            n_sub = advect ? pr.n_sub : 0;
#pragma unroll
            for (int k = 0; k < D; ++k) x[k] = P.x[k][p];
            inv_cached<D>(m, cc, x, xi);
            snap_cached<D>(m, pr, cc, x, xi);
#pragma unroll
            for (int k = 0; k < D; ++k) v[k] = P.v[k][p];
            if (use_dv) {
              double corr[D];
              int dof[1 << D];
              load_ids<D>(m.cd, c, dof);
              gather<D>(dV, dof, xi, corr);
#pragma unroll
              for (int k = 0; k < D; ++k) v[k] = v[k] + corr[k];
            }
          }
        } else {
          p = -1;
        }
      }
      if (__ballot(p >= 0) == 0ULL) break;
      // Euler substeps in the current cell until the particle leaves it
      if (p >= 0 && !dead) {
        while (s < n_sub) {
          double u[D];
          gather_cached<D>(Vc, cc, xi, u);
#pragma unroll
          for (int k = 0; k < D; ++k) {
            const double t = u[k] * dt_sub;
            x[k] = x[k] + t;
          }
          double xn[D];
          if (inv_cached<D>(m, cc, x, xn) && inside<D>(xn)) {
#pragma unroll
            for (int k = 0; k < D; ++k) xi[k] = xn[k];
            ++s;
          } else {
            walking = true;
            break;
          }
        }
      }
      // neighbour walk of the lanes that left their cell, all together
      if (walking) {
        double u[D], xn[D];
        gather_cached<D>(Vc, cc, xi, u);  // the substep's advecting velocity (xi is pre-substep)
        const int c2 = walk_cached<D>(m, cc, xi, x, u, xn);
        walking = false;
        if (c2 < 0) {
          dead = true;
        } else {
          load_cell<D>(m, Vc, c2, cc, true);
#pragma unroll
          for (int k = 0; k < D; ++k) xi[k] = xn[k];
          ++s;
        }
      }
      const bool done = (p >= 0) && !walking && (dead || s >= n_sub);
      // append the moved particle to its destination chunk's bucket (one
      // atomic per distinct chunk among the finishing lanes), or to the
      // spill area when the bucket is full
      const int key = (done && !dead) ? (cc.c >> cshift) : -1;
      const int slot = wave_group_add(bcount, key);
      if (key >= 0) {
        long dst = (long)key * capc + slot;
        if (slot >= capc) dst = spill_base + atomicAdd(spill_n, 1);
#pragma unroll
        for (int k = 0; k < D; ++k) {
          T.x[k][dst] = x[k];
          T.v[k][dst] = v[k];
        }
        T.cell[dst] = cc.c;
        if (T.src) T.src[dst] = p;
      }
      const unsigned long long dm = __ballot(done);
      if (done) {
        p = next + __popcll(dm & below);
        fresh = true;
      }
      next += __popcll(dm);
    }
  }
}

// ---------------------------------------------------------------- settle
// check_particle_distribution (src/pfem2particlehandler.cpp:355-410) on the
// chunk buckets written by k_move, fused with the per-(cell, vertex)
// projection sums of project_particle_fields (:191-199). A workgroup owns
// one chunk of 2^cshift consecutive cells; its particles are the chunk's
// bucket plus its share of the spill region.
// Chunks are taken in ticket order. The workgroup stages the chunk's
// particles in LDS with their reference position, part and canonical key and
// counts them per (cell, part); the final count of each cell is survivors
// min(count, max+1) per part plus one refill per empty part of a non-empty
// cell. The chunk's output offset comes from a decoupled look-back over the
// chunk totals (no separate count pass and scan). Each particle is ranked
// inside its (cell, part) group by key; survivors and refills get their final
// slot (cell base + part base + rank, i.e. (part, key) order); the chunk is
// written out coalesced in that order, with the new cell_start, and the
// (cell, vertex) sums are accumulated sequentially in the same order.
// A chunk with more than `nmax` particles takes the same steps from global
// memory, recomputing parts and keys, and leaves its projection sums to
// k_contrib_dirty.
struct SettleLayout {
  int off_x, off_v, off_xi, off_key, off_pl, off_perm, off_ord, off_cnt, off_base, off_roff, off_rfill, off_cell,
      bytes;
};

__host__ __device__ inline SettleLayout settle_layout(int D, int cpw, int pmax, int nmax, int nslots) {
  SettleLayout L;
  int o = 0;
  L.off_x = o;
  o += 8 * D * nslots;
  L.off_v = o;
  o += 8 * D * nslots;
  L.off_xi = o;
  o += 8 * D * nslots;
  L.off_key = o;
  o += 8 * nmax;
  L.off_pl = o;
  o += 4 * nslots;
  L.off_perm = o;
  o += 4 * nslots;
  L.off_ord = o;
  o += 4 * nmax;
  L.off_cnt = o;
  o += 4 * cpw * pmax;
  L.off_base = o;
  o += 4 * cpw * pmax;
  L.off_roff = o;
  o += 4 * cpw * pmax;
  L.off_rfill = o;
  o += 4 * cpw * pmax;
  L.off_cell = o;
  o += 4 * (4 * cpw + 8);  // per-cell count, raw offset, final offset, total; misc
  L.bytes = o;
  return L;
}

// global index of chunk-local particle j: bucket first, then spill share
__device__ __forceinline__ long chunk_index(int j, int nb, long bucket0, long spill0) {
  return j < nb ? bucket0 + j : spill0 + (j - nb);
}

// strict order of two particles of one (cell, part) with equal keys:
// position bits, then velocity bits, then chunk-local index (which only
// separates bitwise identical particles, so it never changes the output)
template <int D>
__device__ __forceinline__ bool tie_less(const double* xa, const double* va, const double* xb, const double* vb, int ja,
                                         int jb) {
#pragma unroll
  for (int k = 0; k < D; ++k) {
    const unsigned long long a = (unsigned long long)__double_as_longlong(xa[k]);
    const unsigned long long b = (unsigned long long)__double_as_longlong(xb[k]);
    if (a != b) return a < b;
  }
#pragma unroll
  for (int k = 0; k < D; ++k) {
    const unsigned long long a = (unsigned long long)__double_as_longlong(va[k]);
    const unsigned long long b = (unsigned long long)__double_as_longlong(vb[k]);
    if (a != b) return a < b;
  }
  return ja < jb;
}

// Decoupled look-back status word of a chunk: (status << 32) | value, status
// 1 = chunk total available, 2 = inclusive prefix available.
__device__ __forceinline__ void lb_publish(unsigned long long* lb, int ch, unsigned status, unsigned value) {
  __hip_atomic_store(&lb[ch], ((unsigned long long)status << 32) | value, __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
}

// Exclusive prefix of the chunk totals before chunk `ch`. Chunks are taken in
// ticket order, so every chunk waited on belongs to a workgroup that is
// already running and will publish; the wait is still bounded, and a timeout
// is reported through *err instead of hanging.
__device__ __forceinline__ unsigned lb_prefix(unsigned long long* lb, int ch, int* err) {
  unsigned prefix = 0;
  int k = ch - 1;
  long spins = 0;
  while (k >= 0) {
    const unsigned long long w = __hip_atomic_load(&lb[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned st = (unsigned)(w >> 32);
    if (st == 0) {
      if (++spins > (1L << 26)) {
        atomicOr(err, 1);
        break;
      }
      __builtin_amdgcn_s_sleep(1);
      continue;
    }
    prefix += (unsigned)(w & 0xffffffffULL);
    if (st == 2) break;
    --k;
  }
  return prefix;
}

template <int D>
__global__ __launch_bounds__(kBlock) void k_settle(int apply_cap, int do_refill, DMesh m, int nc, int cshift, int pshift,
                                                  int nmax, int nslots, int capc, const int* __restrict__ bcount,
                                                  const int* __restrict__ sstart, long spill2_base, DTmp T,
                                                  int* __restrict__ ticket, unsigned long long* __restrict__ lb,
                                                  int* __restrict__ err, int* __restrict__ cs_out, DParts P, DField Vf,
                                                  double* __restrict__ contrib, unsigned char* __restrict__ dirty) {
  constexpr int NV = 1 << D;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int cpw = 1 << cshift;
  const int pmax = 1 << pshift;  // part buckets per cell, padded to a power of two
  const SettleLayout L = settle_layout(D, cpw, pmax, nmax, nslots);
  double* s_x = reinterpret_cast<double*>(smem_raw + L.off_x);
  double* s_v = reinterpret_cast<double*>(smem_raw + L.off_v);
  double* s_xi = reinterpret_cast<double*>(smem_raw + L.off_xi);
  unsigned long long* s_key = reinterpret_cast<unsigned long long*>(smem_raw + L.off_key);
  int* s_pl = reinterpret_cast<int*>(smem_raw + L.off_pl);
  int* s_perm = reinterpret_cast<int*>(smem_raw + L.off_perm);
  int* s_ord = reinterpret_cast<int*>(smem_raw + L.off_ord);
  int* s_cnt = reinterpret_cast<int*>(smem_raw + L.off_cnt);
  int* s_base = reinterpret_cast<int*>(smem_raw + L.off_base);
  int* s_roff = reinterpret_cast<int*>(smem_raw + L.off_roff);
  int* s_rfill = reinterpret_cast<int*>(smem_raw + L.off_rfill);
  int* s_ccnt = reinterpret_cast<int*>(smem_raw + L.off_cell);
  int* s_coff = s_ccnt + cpw;  // raw (staged) offset of each cell
  int* s_cfin = s_coff + cpw;  // final offset of each cell inside the chunk
  int* s_misc = s_cfin + cpw;  // [0] refill slots, [1] ticket, [2] chunk total, [3] chunk prefix

  const int nchunks = (nc + cpw - 1) >> cshift;
  for (;;) {
    if (threadIdx.x == 0) s_misc[1] = atomicAdd(ticket, 1);
    __syncthreads();
    const int ch = s_misc[1];
    if (ch >= nchunks) break;
    const int c0 = ch << cshift;
    const int c1 = min(nc, c0 + cpw);
    const int ncell = c1 - c0;
    const int nb = min(bcount[ch], capc);
    const int ns = sstart[ch + 1] - sstart[ch];
    const int n = nb + ns;
    const long bucket0 = (long)ch * capc;
    const long spill0 = spill2_base + sstart[ch];
    const bool staged = n <= nmax;
    for (int i = threadIdx.x; i < ncell * pmax; i += blockDim.x) {
      s_cnt[i] = 0;
      s_rfill[i] = 0;
    }
    for (int i = threadIdx.x; i < cpw; i += blockDim.x) s_ccnt[i] = 0;
    if (threadIdx.x == 0) s_misc[0] = 0;
    __syncthreads();
    // 1. reference position, part and key of every particle of the chunk
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
      const long i = chunk_index(j, nb, bucket0, spill0);
      const int c = T.cell[i];
      const int lc = c - c0;
      const Prm& pr = m.prm[inst_of(m, c)];
      double x[D], xi[D];
#pragma unroll
      for (int k = 0; k < D; ++k) x[k] = T.x[k][i];
      recover_xi_c<D>(m, pr, c, x, xi);
      const int part = part_lin<D>(pr, xi, nullptr);
      atomicAdd(&s_cnt[(lc << pshift) + part], 1);
      atomicAdd(&s_ccnt[lc], 1);
      if (staged) {
#pragma unroll
        for (int k = 0; k < D; ++k) {
          s_x[k * nslots + j] = x[k];
          s_v[k * nslots + j] = T.v[k][i];
          s_xi[k * nslots + j] = xi[k];
        }
        s_pl[j] = (lc << 16) | part;
        s_key[j] = particle_key<D>(x);
      }
    }
    __syncthreads();
    // 2. per (cell, part): survivors + refill; per cell the exclusive prefixes
    //    over parts of the final (part bases) and raw (groups) counts
    for (int t = threadIdx.x; t < ncell * pmax; t += blockDim.x) {
      const int lc = t >> pshift, pp = t & (pmax - 1);
      const int c = c0 + lc;
      const int inst = inst_of(m, c);
      const int k = s_cnt[t];
      const int maxpp = m.prm[inst].maxpp;
      const int kept = apply_cap ? min(k, maxpp + 1) : k;
      const bool valid = m.ptab[(inst << pshift) + pp] < 0;
      const int refill = (do_refill && s_ccnt[lc] > 0 && k == 0 && valid) ? 1 : 0;
      s_base[t] = kept + refill;  // padding parts have k == 0 and are not valid
    }
    __syncthreads();
    for (int lc = threadIdx.x; lc < ncell; lc += blockDim.x) {
      int run = 0, raw = 0;
      for (int pp = 0; pp < pmax; ++pp) {
        const int t = (lc << pshift) + pp;
        const int a = s_base[t];
        s_base[t] = run;
        run += a;
        s_roff[t] = raw;
        raw += s_cnt[t];
      }
      s_cfin[lc] = run;                               // cell total for now
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      int run = 0, raw = 0;
      for (int lc = 0; lc < ncell; ++lc) {
        const int tot = s_cfin[lc];
        s_cfin[lc] = run;
        run += tot;
        s_coff[lc] = raw;
        raw += s_ccnt[lc];
      }
      s_misc[2] = run;
      lb_publish(lb, ch, ch == 0 ? 2u : 1u, (unsigned)run);
      const unsigned prefix = ch == 0 ? 0u : lb_prefix(lb, ch, err);
      if (ch > 0) lb_publish(lb, ch, 2u, prefix + (unsigned)run);
      s_misc[3] = (int)prefix;
    }
    __syncthreads();
    const int o0 = s_misc[3];
    for (int lc = threadIdx.x; lc < ncell; lc += blockDim.x) cs_out[c0 + lc] = o0 + s_cfin[lc];
    if (c1 == nc && threadIdx.x == 0) cs_out[nc] = o0 + s_misc[2];
    if (staged) {
      // group the staged particles by (cell, part)
      for (int j = threadIdx.x; j < n; j += blockDim.x) {
        const int lc = s_pl[j] >> 16;
        const int t = (lc << pshift) + (s_pl[j] & 0xffff);
        s_ord[s_coff[lc] + s_roff[t] + atomicAdd(&s_rfill[t], 1)] = j;
      }
      __syncthreads();
    }
    // 3. rank inside (cell, part) and final slot of every particle
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
      int lc, part, r = 0;
      if (staged) {
        lc = s_pl[j] >> 16;
        part = s_pl[j] & 0xffff;
        const unsigned long long key = s_key[j];
        const int t = (lc << pshift) + part;
        const int a = s_coff[lc] + s_roff[t], b = a + s_cnt[t];
        for (int u = a; u < b; ++u) {
          const int q = s_ord[u];
          if (q == j) continue;
          const unsigned long long kq = s_key[q];
          if (kq < key) {
            ++r;
          } else if (kq == key) {
            double xq[D], vq[D], xj[D], vj[D];
#pragma unroll
            for (int k = 0; k < D; ++k) {
              xq[k] = s_x[k * nslots + q];
              vq[k] = s_v[k * nslots + q];
              xj[k] = s_x[k * nslots + j];
              vj[k] = s_v[k * nslots + j];
            }
            if (tie_less<D>(xq, vq, xj, vj, q, j)) ++r;
          }
        }
      } else {
        // more particles than the LDS stage holds: recompute from global
        const long i = chunk_index(j, nb, bucket0, spill0);
        const int c = T.cell[i];
        lc = c - c0;
        const Prm& pr = m.prm[inst_of(m, c)];
        double xj[D], vj[D], xi[D];
#pragma unroll
        for (int k = 0; k < D; ++k) {
          xj[k] = T.x[k][i];
          vj[k] = T.v[k][i];
        }
        recover_xi_c<D>(m, pr, c, xj, xi);
        part = part_lin<D>(pr, xi, nullptr);
        const unsigned long long key = particle_key<D>(xj);
        for (int q = 0; q < n; ++q) {
          if (q == j) continue;
          const long iq = chunk_index(q, nb, bucket0, spill0);
          if (T.cell[iq] != c) continue;
          double xq[D], vq[D], xiq[D];
#pragma unroll
          for (int k = 0; k < D; ++k) {
            xq[k] = T.x[k][iq];
            vq[k] = T.v[k][iq];
          }
          recover_xi_c<D>(m, pr, c, xq, xiq);
          if (part_lin<D>(pr, xiq, nullptr) != part) continue;
          const unsigned long long kq = particle_key<D>(xq);
          if (kq < key || (kq == key && tie_less<D>(xq, vq, xj, vj, q, j))) ++r;
        }
      }
      const int c = c0 + lc;
      if (apply_cap && r > m.prm[inst_of(m, c)].maxpp) continue;
      const int fpos = s_cfin[lc] + s_base[(lc << pshift) + part] + r;
      if (staged) {
        s_perm[fpos] = j;
      } else {
        const long i = chunk_index(j, nb, bucket0, spill0);
#pragma unroll
        for (int k = 0; k < D; ++k) {
          P.x[k][o0 + fpos] = T.x[k][i];
          P.v[k][o0 + fpos] = T.v[k][i];
        }
        P.cell[o0 + fpos] = c;
        if (P.src) P.src[o0 + fpos] = T.src[i];
      }
    }
    // 4. refills: one particle at the centre of each empty part of a
    //    non-empty cell, velocity interpolated from the advecting field
    if (do_refill) {
      for (int t = threadIdx.x; t < ncell * pmax; t += blockDim.x) {
        const int lc = t >> pshift, pp = t & (pmax - 1);
        const int c = c0 + lc;
        if (s_cnt[t] != 0 || s_ccnt[lc] == 0) continue;
        const int inst = inst_of(m, c);
        const int code = m.ptab[(inst << pshift) + pp];
        if (code >= 0) continue;  // not a refillable part (or padding)
        const Prm& pr = m.prm[inst];
        int pi[D];
#pragma unroll
        for (int d = 0; d < D; ++d) pi[d] = (code >> (8 * d)) & 0xff;
        double xi[D], x[D], v[D];
#pragma unroll
        for (int k = 0; k < D; ++k) xi[k] = ((double)pi[k] + 1.0 / 2) * pr.h[k];
        int vid[NV], dof[NV];
        load_ids<D>(m.cv, c, vid);
        load_ids<D>(m.cd, c, dof);
        fwd<D>(m, vid, xi, x);
        gather<D>(Vf, dof, xi, v);
        const int fpos = s_cfin[lc] + s_base[t];
        if (staged) {
          const int slot = n + atomicAdd(&s_misc[0], 1);
#pragma unroll
          for (int k = 0; k < D; ++k) {
            s_x[k * nslots + slot] = x[k];
            s_v[k * nslots + slot] = v[k];
            s_xi[k * nslots + slot] = xi[k];
          }
          s_pl[slot] = (lc << 16) | pp;
          s_perm[fpos] = slot;
        } else {
#pragma unroll
          for (int k = 0; k < D; ++k) {
            P.x[k][o0 + fpos] = x[k];
            P.v[k][o0 + fpos] = v[k];
          }
          P.cell[o0 + fpos] = c;
          if (P.src) P.src[o0 + fpos] = -1 - pp;
        }
      }
    }
    __syncthreads();
    if (staged) {
      // 5. coalesced write-out in final order
      const int nout = s_misc[2];
      for (int t = threadIdx.x; t < nout; t += blockDim.x) {
        const int slot = s_perm[t];
#pragma unroll
        for (int k = 0; k < D; ++k) {
          P.x[k][o0 + t] = s_x[k * nslots + slot];
          P.v[k][o0 + t] = s_v[k * nslots + slot];
        }
        P.cell[o0 + t] = c0 + (s_pl[slot] >> 16);
        if (P.src) P.src[o0 + t] = slot < n ? T.src[chunk_index(slot, nb, bucket0, spill0)] : -1 - (s_pl[slot] & 0xffff);
      }
      // 6. per (cell, vertex) sums, sequential in final order
      for (int t = threadIdx.x; t < ncell * NV; t += blockDim.x) {
        const int lc = t / NV, iv = t - lc * NV;
        const int c = c0 + lc;
        const int a = s_cfin[lc], b = (lc + 1 < ncell) ? s_cfin[lc + 1] : nout;
        double S[D], W = 0.0;
#pragma unroll
        for (int k = 0; k < D; ++k) S[k] = 0.0;
        for (int f = a; f < b; ++f) {
          const int slot = s_perm[f];
          double N = 1.0;
#pragma unroll
          for (int k = 0; k < D; ++k) {
            const double xk = s_xi[k * nslots + slot];
            const double l = ((iv >> k) & 1) ? xk : (1.0 - xk);
            N = (k == 0) ? l : N * l;
          }
#pragma unroll
          for (int k = 0; k < D; ++k) S[k] = S[k] + N * s_v[k * nslots + slot];
          W = W + N;
        }
        double* e = contrib + ((size_t)c * NV + iv) * (D + 1);
#pragma unroll
        for (int k = 0; k < D; ++k) e[k] = S[k];
        e[D] = W;
      }
    }
    if (threadIdx.x == 0) dirty[ch] = staged ? 0 : 1;
    __syncthreads();
  }
}

// ------------------------------------------------------------- projection
// Per (cell, vertex) sums S_k = sum_p N_i(xi_p) v_p,k and W = sum_p N_i(xi_p),
// sequential over the cell's particles in stored order.
// src/pfem2particlehandler.cpp:191-199
template <int D>
__device__ __forceinline__ void contrib_one(const DMesh& m, int c, int i, const int* __restrict__ cs, const DParts& P,
                                            double* __restrict__ contrib) {
  constexpr int NV = 1 << D;
  const Prm pr = m.prm[inst_of(m, c)];
  int vid[NV];
  load_ids<D>(m.cv, c, vid);
  double S[D], W = 0.0;
#pragma unroll
  for (int k = 0; k < D; ++k) S[k] = 0.0;
  const int s = cs[c], e = cs[c + 1];
  for (int p = s; p < e; ++p) {
    double x[D], xi[D], N[NV];
#pragma unroll
    for (int k = 0; k < D; ++k) x[k] = P.x[k][p];
    recover_xi_c<D>(m, pr, c, x, xi);
    shape<D>(xi, N);
    double Ni = N[0];
#pragma unroll
    for (int q = 1; q < NV; ++q)
      if (q == i) Ni = N[q];
#pragma unroll
    for (int k = 0; k < D; ++k) S[k] = S[k] + Ni * P.v[k][p];
    W = W + Ni;
  }
  double* e_out = contrib + ((size_t)c * NV + i) * (D + 1);
#pragma unroll
  for (int k = 0; k < D; ++k) e_out[k] = S[k];
  e_out[D] = W;
}

template <int D>
__global__ __launch_bounds__(kBlock) void k_contrib(DMesh m, int nc, const int* __restrict__ cs, DParts P,
                                                   double* __restrict__ contrib) {
  constexpr int NV = 1 << D;
  const long total = (long)nc * NV;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x)
    contrib_one<D>(m, (int)(t / NV), (int)(t % NV), cs, P, contrib);
}

// The same sums for the cells of chunks the settle kernel could not stage.
template <int D>
__global__ __launch_bounds__(kBlock) void k_contrib_dirty(DMesh m, int nc, int cpw, const unsigned char* __restrict__ dirty,
                                                         const int* __restrict__ cs, DParts P,
                                                         double* __restrict__ contrib) {
  constexpr int NV = 1 << D;
  const int nchunks = (nc + cpw - 1) / cpw;
  for (int ch = blockIdx.x; ch < nchunks; ch += gridDim.x) {
    if (!dirty[ch]) continue;
    const int c0 = ch * cpw, ncell = min(nc, c0 + cpw) - c0;
    for (int t = threadIdx.x; t < ncell * NV; t += blockDim.x) contrib_one<D>(m, c0 + t / NV, t % NV, cs, P, contrib);
  }
}

// V_k[j] = (sum over incident cells, ascending id, of S_k) / (same for W);
// 0/0 gives NaN as in src/pfem2particlehandler.cpp:207-209.
template <int D>
__global__ __launch_bounds__(kBlock) void k_node(int ndofs, const int* __restrict__ d2c_off,
                                                const int* __restrict__ d2c, const double* __restrict__ contrib,
                                                double* o0, double* o1, double* o2) {
  double* out[3] = {o0, o1, o2};
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < ndofs; j += gridDim.x * blockDim.x) {
    double S[D], W = 0.0;
#pragma unroll
    for (int k = 0; k < D; ++k) S[k] = 0.0;
    for (int e = d2c_off[j]; e < d2c_off[j + 1]; ++e) {
      const double* c = contrib + (size_t)d2c[e] * (D + 1);
#pragma unroll
      for (int k = 0; k < D; ++k) S[k] = S[k] + c[k];
      W = W + c[D];
    }
#pragma unroll
    for (int k = 0; k < D; ++k) out[k][j] = S[k] / W;
  }
}

// ------------------------------------------------------ per-particle kernels

// v += interp(dV) (flush of a pending correction outside move)
template <int D>
__global__ __launch_bounds__(kBlock) void k_correct(DMesh m, const int* __restrict__ n_ptr, DParts P, DField dV) {
  constexpr int NV = 1 << D;
  const int n = *n_ptr;
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
    const int c = P.cell[p];
    const Prm pr = m.prm[inst_of(m, c)];
    double x[D], xi[D], corr[D];
#pragma unroll
    for (int k = 0; k < D; ++k) x[k] = P.x[k][p];
    int vid[NV], dof[NV];
    load_ids<D>(m.cv, c, vid);
    load_ids<D>(m.cd, c, dof);
    recover_xi_c<D>(m, pr, c, x, xi);
    gather<D>(dV, dof, xi, corr);
#pragma unroll
    for (int k = 0; k < D; ++k) P.v[k][p] = P.v[k][p] + corr[k];
  }
}

// v = interp(V)
template <int D>
__global__ __launch_bounds__(kBlock) void k_setvel(DMesh m, const int* __restrict__ n_ptr, DParts P, DField V) {
  constexpr int NV = 1 << D;
  const int n = *n_ptr;
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
    const int c = P.cell[p];
    const Prm pr = m.prm[inst_of(m, c)];
    double x[D], xi[D], u[D];
#pragma unroll
    for (int k = 0; k < D; ++k) x[k] = P.x[k][p];
    int vid[NV], dof[NV];
    load_ids<D>(m.cv, c, vid);
    load_ids<D>(m.cd, c, dof);
    recover_xi_c<D>(m, pr, c, x, xi);
    gather<D>(V, dof, xi, u);
#pragma unroll
    for (int k = 0; k < D; ++k) P.v[k][p] = u[k];
  }
}

// reference positions for download
template <int D>
__global__ __launch_bounds__(kBlock) void k_xi(DMesh m, const int* __restrict__ n_ptr, DParts P, double* xi0,
                                              double* xi1, double* xi2) {
  constexpr int NV = 1 << D;
  double* out[3] = {xi0, xi1, xi2};
  const int n = *n_ptr;
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
    const int c = P.cell[p];
    const Prm pr = m.prm[inst_of(m, c)];
    double x[D], xi[D];
#pragma unroll
    for (int k = 0; k < D; ++k) x[k] = P.x[k][p];
    int vid[NV];
    load_ids<D>(m.cv, c, vid);
    recover_xi_c<D>(m, pr, c, x, xi);
#pragma unroll
    for (int k = 0; k < D; ++k) out[k][p] = xi[k];
  }
}

// Appends n_new staged particles after the current ones, rejecting those whose
// cell id is out of range or whose position is not inside that cell.
template <int D>
__global__ __launch_bounds__(kBlock) void k_append(DMesh m, int nc, const int* __restrict__ n_ptr, int n_new,
                                                  const double* sx0, const double* sx1, const double* sx2,
                                                  const double* sv0, const double* sv1, const double* sv2,
                                                  const int* __restrict__ scell, DParts P,
                                                  unsigned long long* __restrict__ rejected) {
  constexpr int NV = 1 << D;
  const double* sx[3] = {sx0, sx1, sx2};
  const double* sv[3] = {sv0, sv1, sv2};
  const int n = *n_ptr;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n_new; i += gridDim.x * blockDim.x) {
    int c = scell[i];
    double x[D], xi[D];
#pragma unroll
    for (int k = 0; k < D; ++k) x[k] = sx[k][i];
    bool ok = (c >= 0) && (c < nc);
    if (ok) {
      int vid[NV];
      load_ids<D>(m.cv, c, vid);
      ok = inv_map<D>(m, c, vid, x, xi) && inside<D>(xi);
    }
    if (!ok) {
      atomicAdd(rejected, 1ULL);
      c = -1;
    }
#pragma unroll
    for (int k = 0; k < D; ++k) {
      P.x[k][n + i] = x[k];
      P.v[k][n + i] = sv[0] ? sv[k][i] : 0.0;
    }
    P.cell[n + i] = c;
  }
}

// seed_particles: particle at the centre of every valid part of every cell,
// velocity interpolated from V, staged for k_append.
template <int D>
__global__ __launch_bounds__(kBlock) void k_seed(DMesh m, int nc, int pvalid, DField V, double* sx0, double* sx1,
                                                double* sx2, double* sv0, double* sv1, double* sv2,
                                                int* __restrict__ scell) {
  constexpr int NV = 1 << D;
  double* sx[3] = {sx0, sx1, sx2};
  double* sv[3] = {sv0, sv1, sv2};
  const long total = (long)nc * pvalid;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
    const int c = (int)(t / pvalid);
    int rem = (int)(t % pvalid);
    const Prm pr = m.prm[inst_of(m, c)];
    int pi[D];
#pragma unroll
    for (int d = D - 1; d >= 0; --d) {
      pi[d] = rem % pr.q[d];
      rem /= pr.q[d];
    }
    double xi[D], x[D], v[D];
#pragma unroll
    for (int k = 0; k < D; ++k) xi[k] = ((double)pi[k] + 1.0 / 2) * pr.h[k];
    int vid[NV], dof[NV];
    load_ids<D>(m.cv, c, vid);
    load_ids<D>(m.cd, c, dof);
    fwd<D>(m, vid, xi, x);
    gather<D>(V, dof, xi, v);
#pragma unroll
    for (int k = 0; k < D; ++k) {
      sx[k][t] = x[k];
      sv[k][t] = v[k];
    }
    scell[t] = c;
  }
}

}  // namespace pfem2gpu
