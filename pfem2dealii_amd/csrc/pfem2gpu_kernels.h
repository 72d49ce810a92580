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

// Zeroes three int arrays in one launch (the per-move counters).
__global__ void k_zero3(int* __restrict__ a, long na, int* __restrict__ b, long nb, int* __restrict__ c, long nc) {
  const long n = na + nb + nc;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    if (i < na)
      a[i] = 0;
    else if (i < na + nb)
      b[i - na] = 0;
    else
      c[i - na - nb] = 0;
  }
}

// dV = V - V_old (src/pfem2particlehandler.cpp:120-124), stored cell-major
// like the advecting field: dVc[c][i][k] = V_k[dof(c, i)] - V_old_k[dof(c, i)]
// (the same subtraction per DoF, repeated for each incident cell), so the
// correction reads a cell's values with one contiguous load.
template <int D>
__global__ void k_subc(int nc, const int* __restrict__ cd, DField V, DField Vold, double* __restrict__ dVc) {
  constexpr int NV = 1 << D;
  const long total = (long)nc * NV;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
    const int dof = cd[t];
#pragma unroll
    for (int k = 0; k < D; ++k) dVc[t * D + k] = V.c[k][dof] - Vold.c[k][dof];
  }
}

// u_k = sum_i N_i(xi) F[c][i][k] from a cell-major table, vertex-outer,
// component-inner, from 0.0.
template <int D>
__device__ __forceinline__ void gather_table(const double* __restrict__ F, int c, const double* xi, double* u) {
  constexpr int NV = 1 << D;
  double N[NV];
  shape<D>(xi, N);
  const double* f = F + (size_t)c * NV * D;
#pragma unroll
  for (int k = 0; k < D; ++k) u[k] = 0.0;
#pragma unroll
  for (int i = 0; i < NV; ++i)
#pragma unroll
    for (int k = 0; k < D; ++k) u[k] = u[k] + N[i] * f[i * D + k];
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
// Groups the lanes of the wavefront by key (ballots only, no memory access):
// for a lane with key >= 0, leader_of is the lowest lane holding the same key
// and rank its position among those lanes; a leader's cnt is the size of its
// group. The caller's group leaders then add cnt to a per-key counter, all in
// one instruction, so the wavefront waits for one round trip however many
// keys it holds. Must be called by all 64 lanes.
__device__ __forceinline__ void wave_groups(int key, int& leader_of, int& rank, int& cnt) {
  const int lane = threadIdx.x & 63;
  const unsigned long long below = (1ULL << lane) - 1ULL;
  leader_of = lane;
  rank = 0;
  cnt = 0;
  unsigned long long pending = __ballot(key >= 0);
  while (pending) {
    const int leader = __ffsll((unsigned long long)pending) - 1;
    const int lk = __builtin_amdgcn_readlane(key, leader);  // (a wave-uniform lane: a scalar read)
    const unsigned long long grp = __ballot(key == lk) & pending;
    if ((grp >> lane) & 1ULL) {
      leader_of = leader;
      rank = __popcll(grp & below);
    }
    if (lane == leader) cnt = __popcll(grp);
    pending &= ~grp;
  }
}

// Spill binning: particles that did not fit their chunk's bucket are counted
// per chunk and copied, grouped by chunk, into the second spill region.
__global__ void k_spill_count(const int* __restrict__ spill_n, long spill_base, const int* __restrict__ tcell,
                              int cshift, int* __restrict__ scount) {
  const int n = *spill_n;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    atomicAdd(&scount[cell_of(tcell[spill_base + i]) >> cshift], 1);
}

template <int D>
__global__ void k_spill_scatter(const int* __restrict__ spill_n, long spill_base, long spill2_base, int cshift,
                                const int* __restrict__ sstart, int* __restrict__ sfill, DTmp T) {
  const int n = *spill_n;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const long src = spill_base + i;
    const int c = T.cell[src];  // cell word (placed flag kept)
    const int ch = cell_of(c) >> cshift;
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

// -------------------------------------------------------------------- move
// move_particles (src/pfem2particlehandler.cpp:141-171): per particle the
// correction of the carried velocity (fused, :111-139), then n_sub explicit
// Euler substeps x <- x + dt_sub u(x) (:150-163), each followed by the
// current-cell test and, when the particle left its cell, the neighbour walk
// (:418-492). Scheduled in lockstep: k_move_cell runs every particle's
// substeps while it stays in its cell and queues the particles that leave
// it; each k_move_walk round takes the queue, walks every queued particle to
// its new cell together, continues its substeps and queues it again if it
// leaves once more. A particle walks at most once per substep, so n_sub
// rounds drain the queue. Every particle ends in the bucket of its
// destination chunk (2^cshift consecutive cells), or in the spill region when
// the bucket is full; the order inside a bucket is not deterministic and the
// settle kernel does not depend on it.

// The walk queue is split into kQStripes stripes of `qreg` entries with one
// counter each; a workgroup writes to stripe blockIdx.x % kQStripes, so the
// counters see little contention and a stripe never overflows (it receives
// at most the particles of the workgroups mapped to it).
constexpr int kQStripes = 1024;

// Appends one particle per lane with key >= 0 to its chunk bucket (or the
// spill region) and queues the lanes with walk set. Wave-uniform. With
// capc < 0 (replay of a recorded move) the particle is stored at its source
// index in T.x / T.v instead of a bucket. The source index is queued only
// when the move needs it (keep_src: source maps, paths or a replay).
template <int D>
__device__ __forceinline__ void move_emit(int key, bool walk, int cell, int tflag, int s, int src, bool keep_src,
                                          const double* x, int qcell, const double* v, int capc,
                                          int* __restrict__ bcount, long spill_base, int* __restrict__ spill_n,
                                          DTmp& T, DWalk& Q, int* __restrict__ qn, long qreg) {
  const int lane = threadIdx.x & 63;
  const unsigned long long below = (1ULL << lane) - 1ULL;
  const unsigned long long wm = __ballot(walk);
  const int stripe = blockIdx.x % kQStripes;
  const int qleader = wm ? __ffsll((unsigned long long)wm) - 1 : 0;
  // lanes grouped by bucket first; then the queue reservation and the bucket
  // reservations are issued together and waited for once
  int leader_of = lane, rank = 0, cnt = 0;
  if (capc >= 0) wave_groups(key, leader_of, rank, cnt);
  int qb = 0, base = 0;
  // (the stripe index passes through a vector register: with a wave-uniform
  // address the compiler would rebuild the atomic as a wave reduction whose
  // result is read back at once, a second round trip)
  int qi = stripe;
  asm volatile("" : "+v"(qi));
  if (wm && lane == qleader) qb = atomicAdd(&qn[qi], __popcll(wm));
  if (capc >= 0 && key >= 0 && leader_of == lane) base = atomicAdd(&bcount[key], cnt);
  const int slot = __shfl(base, leader_of, 64) + rank;
  if (wm) {
    qb = __builtin_amdgcn_readlane(qb, qleader);
    if (qb + __popcll(wm) > qreg) {  // the stripe is full: report, never write past it
      if (lane == qleader) atomicOr(Q.err, kErrQueue);
    } else if (walk) {
      const long q = (long)stripe * qreg + qb + __popcll(wm & below);
#pragma unroll
      for (int k = 0; k < D; ++k) {
        Q.x[k][q] = x[k];
        Q.v[k][q] = v[k];
      }
      Q.cell[q] = qcell;
      Q.s[q] = s;
      if (keep_src) Q.src[q] = src;
    }
  }
  if (capc < 0) {  // replay: position and velocity stored at the source index
    if (key >= 0) {
#pragma unroll
      for (int k = 0; k < D; ++k) {
        T.x[k][src] = x[k];
        T.v[k][src] = v[k];
      }
    }
    return;
  }
  if (key >= 0) {
    long dst = (long)key * capc + slot;
    if (slot >= capc) dst = spill_base + atomicAdd(spill_n, 1);
#pragma unroll
    for (int k = 0; k < D; ++k) {
      T.x[k][dst] = x[k];
      T.v[k][dst] = v[k];
    }
    T.cell[dst] = cell | tflag;
    if (T.src) T.src[dst] = src;
  }
}

// Substeps s, s+1, ... of a particle in cell cc until n_sub or until it
// leaves the cell (returns true; x and xi are then the position and the
// reference position before the failing substep, u its velocity).
template <int D>
__device__ __forceinline__ bool substeps_in_cell(const DMesh& m, const double* __restrict__ Vc, const CellCache<D>& cc,
                                                 int n_sub, double dt_sub, int& s, double* x, double* xi, double* u,
                                                 const DPath& path, int pidx) {
  for (; s < n_sub; ++s) {
    if (path.c) {
      const long o = s * path.stride + pidx;
      path.c[o] = cc.c;
#pragma unroll
      for (int k = 0; k < D; ++k) path.xi[k][o] = xi[k];
    }
    gather_cached<D>(Vc, cc, xi, u);
    double xe[D];
#pragma unroll
    for (int k = 0; k < D; ++k) {
      const double t = u[k] * dt_sub;
      xe[k] = x[k] + t;
    }
    double xn[D];
    if (inv_cached<D>(m, cc, xe, xn) && inside<D>(xn)) {
#pragma unroll
      for (int k = 0; k < D; ++k) {
        x[k] = xe[k];
        xi[k] = xn[k];
      }
    } else {
      return true;
    }
  }
  return false;
}

// In-cell pass, one particle per thread. `advect` = 0 turns it into a pure
// re-binning (used by insertion). A pending velocity correction is applied
// from the pre-move cell and reference position first, as
// correct_particle_velocity would have done before the move.
template <int D, int B>
__global__ __launch_bounds__(B) void k_move_cell(int advect, DMesh m, const double* __restrict__ Vc,
                                                     const int* __restrict__ n_ptr, int n_add, DParts P,
                                                     const double* __restrict__ dVc, int use_dv, int cshift, int capc,
                                                     int* __restrict__ bcount,
                                                     long spill_base, int* __restrict__ spill_n, DTmp T, DWalk Q,
                                                     int* __restrict__ qn, long qreg, DPath path) {
  constexpr int NV = 1 << D;
  const int n = *n_ptr + n_add;
  const bool keep_src = T.src || capc < 0 || path.c;
  // a handle with one instance holds its move parameters in scalar registers
  const bool single = m.cell_inst == nullptr;
  MovePrm<D> mp1{};
  if (single) mp1 = uniform_move_prm<D>(m.prm);
  for (long base = (long)blockIdx.x * blockDim.x; base < n; base += (long)gridDim.x * blockDim.x) {
    const int p = (int)(base + threadIdx.x);
    int key = -1, s = 0, tflag = 0, qflag = 0;
    bool walk = false;
    CellCache<D> cc;
    cc.c = -1;
    double x[D], xi[D], v[D], u[D];
#pragma unroll
    for (int k = 0; k < D; ++k) {
      x[k] = 0.0;
      xi[k] = 0.0;
      v[k] = 0.0;
      u[k] = 0.0;
    }
    // first round of loads: the cell word, position and velocity together
    int w = -1;
    if (p < n) {
      w = P.cell[p];
#pragma unroll
      for (int k = 0; k < D; ++k) {
        x[k] = P.x[k][p];
        v[k] = P.v[k][p];
      }
    }
    if (w >= 0) {
      const int c = cell_of(w);
      // second round: the cell's geometry, field, correction table and
      // (batch) instance parameters, all before the first use
      load_cell<D>(m, Vc, c, cc, advect != 0);
      // (2D: the correction table too; 3D reads it after the inverse map,
      // where its 24 values need not be held across it)
      double dvc[D == 2 ? NV * D : 1];
      if constexpr (D == 2) {
        if (use_dv) {
          const double2* f = reinterpret_cast<const double2*>(dVc + (size_t)c * NV * D);
#pragma unroll
          for (int i = 0; i < NV; ++i) {
            const double2 t = f[i];
            dvc[2 * i] = t.x;
            dvc[2 * i + 1] = t.y;
          }
        }
      }
      const MovePrm<D> mp = single ? mp1 : load_move_prm<D>(m.prm + m.cell_inst[c]);
      inv_cached<D>(m, cc, x, xi);
      if (placed_of(w)) snap_part<D>(mp, xi);
      // a move locates every particle at least once (n_sub >= 1); a
      // re-binning keeps the flag
      if (!advect) tflag = w & kPlaced;
      if (use_dv) {
        double corr[D];
        if constexpr (D == 2) {
          // gather_table's sum from the preloaded table
          double N[NV];
          shape<D>(xi, N);
#pragma unroll
          for (int k = 0; k < D; ++k) corr[k] = 0.0;
#pragma unroll
          for (int i = 0; i < NV; ++i)
#pragma unroll
            for (int k = 0; k < D; ++k) corr[k] = corr[k] + N[i] * dvc[i * D + k];
        } else {
          gather_table<D>(dVc, c, xi, corr);
        }
#pragma unroll
        for (int k = 0; k < D; ++k) v[k] = v[k] + corr[k];
      }
      walk = substeps_in_cell<D>(m, Vc, cc, advect ? mp.n_sub : 0, mp.dt_sub, s, x, xi, u, path, p);
      if (!walk) key = c >> cshift;
      // a leaver's queued reference position is recomputed from x by the
      // walk: the snap of the first substep travels as the cell word's flag
      if (walk && s == 0 && placed_of(w)) qflag = kPlaced;
    }
    move_emit<D>(key, walk, cc.c, tflag, s, p, keep_src, x, cc.c | qflag, v, capc, bcount, spill_base, spill_n, T, Q,
                 qn, qreg);
  }
}

// Walk round: workgroup b takes stripe b % kQStripes of the input queue
// (the grid is a multiple of kQStripes) and writes to the same stripe of the
// output queue.
template <int D, int B>
__global__ __launch_bounds__(B) __attribute__((amdgpu_waves_per_eu(D == 2 ? 3 : 1))) void k_move_walk(DMesh m, const double* __restrict__ Vc, DWalk Qin,
                                                     const int* __restrict__ qn_in, long qreg, int cshift, int capc,
                                                     int* __restrict__ bcount, long spill_base,
                                                     int* __restrict__ spill_n, DTmp T, DWalk Qout,
                                                     int* __restrict__ qn_out, DPath path) {
  const int stripe = blockIdx.x % kQStripes;
  const int per = gridDim.x / kQStripes;
  const int nq = qn_in[stripe];
  const bool keep_src = T.src || capc < 0 || path.c;
  const bool single = m.cell_inst == nullptr;
  MovePrm<D> mp1{};
  if (single) mp1 = uniform_move_prm<D>(m.prm);
  for (long base = (long)(blockIdx.x / kQStripes) * blockDim.x; base < nq; base += (long)per * blockDim.x) {
    const int i = (int)(stripe * qreg + base + threadIdx.x);
    const bool in = base + threadIdx.x < nq;
    int key = -1, s = 0, src = 0;
    bool walk = false;
    CellCache<D> cc;
    cc.c = -1;
    double x[D], xi[D], v[D], u[D];
#pragma unroll
    for (int k = 0; k < D; ++k) {
      x[k] = 0.0;
      xi[k] = 0.0;
      v[k] = 0.0;
      u[k] = 0.0;
    }
    if (in) {
      const int cw = Qin.cell[i];
      const int c = cell_of(cw);
      s = Qin.s[i];
      if (keep_src) src = Qin.src[i];
#pragma unroll
      for (int k = 0; k < D; ++k) {
        x[k] = Qin.x[k][i];
        v[k] = Qin.v[k][i];
      }
      int4 r0 = make_int4(-2, -1, -1, -1), r1 = r0, r2 = r0, r3 = r0;
      if constexpr (D == 2) {
        if (m.cvc) {
          r0 = m.cvc[(size_t)c * 4 + 0];
          r1 = m.cvc[(size_t)c * 4 + 1];
          r2 = m.cvc[(size_t)c * 4 + 2];
          r3 = m.cvc[(size_t)c * 4 + 3];
        }
      }
      load_cell<D>(m, Vc, c, cc, true);
      const MovePrm<D> pr = single ? mp1 : load_move_prm<D>(m.prm + m.cell_inst[c]);
      // the failing substep again, bit for bit: its reference position (the
      // left cell's inverse map of the queued position, snapped to the part
      // centre for a placed particle's first substep), its velocity and its
      // end position
      double xo[D];
      inv_cached<D>(m, cc, x, xo);
      if (placed_of(cw)) snap_part<D>(pr, xo);
      gather_cached<D>(Vc, cc, xo, u);
#pragma unroll
      for (int k = 0; k < D; ++k) {
        const double t = u[k] * pr.dt_sub;
        x[k] = x[k] + t;
      }
      double xn[D];
      int c2 = -2;
      bool have_g = false;
      double gw[6];
      if constexpr (D == 2) {
        if (m.cvc && cc.box) {
          c2 = walk_box2(m, cc, xo, x, u, r0, r1, r2, r3, xn, gw);
          have_g = c2 >= 0 && gw[0] == gw[0];
        }
      }
      if (c2 == -2) c2 = walk_cached<D>(m, cc, xo, x, u, xn);
      if (c2 >= 0) {
        if (have_g) {  // winner's corners already read: only its field
          cc.c = c2;
          cc.box = true;
#pragma unroll
          for (int k = 0; k < D; ++k) {
            cc.X0[k] = gw[k];
            cc.X1[k] = gw[D + k];
            cc.RH[k] = gw[2 * D + k];
          }
          if constexpr (D == 2) {
            const double2* f = reinterpret_cast<const double2*>(Vc + (size_t)c2 * 8);
#pragma unroll
            for (int i2 = 0; i2 < 4; ++i2) {
              const double2 t = f[i2];
              cc.Vn[i2][0] = t.x;
              cc.Vn[i2][1] = t.y;
            }
          }
        } else {
          load_cell<D>(m, Vc, c2, cc, true);
        }
#pragma unroll
        for (int k = 0; k < D; ++k) xi[k] = xn[k];
        ++s;
        walk = substeps_in_cell<D>(m, Vc, cc, pr.n_sub, pr.dt_sub, s, x, xi, u, path, src);
        if (!walk) key = c2 >> cshift;
      }
    }
    move_emit<D>(key, walk, cc.c, 0, s, src, keep_src, x, cc.c, v, capc, bcount, spill_base, spill_n, T, Qout, qn_out,
                 qreg);
  }
}

// ---------------------------------------------------------------- settle
// check_particle_distribution (src/pfem2particlehandler.cpp:355-410) on the
// chunk buckets written by the move, fused with the per-(cell, vertex)
// projection sums of project_particle_fields (:191-199). A workgroup owns
// one chunk of 2^cshift consecutive cells at a time; its particles are the
// chunk's bucket plus its share of the spill region. Chunks are taken in
// ticket order (the next ticket and the next chunk's bucket and spill extents
// are read while the current chunk is processed).
//  A. the chunk's cell data (box geometry, advecting field, instance, part
//     codes) and its particles (position, velocity, cell) are read in one
//     round of loads; each particle's reference position, part and
//     canonical key are computed and counted per (cell, part) in LDS;
//  B. per (cell, part): survivors min(count, max+1) plus one refill per
//     empty valid part of a non-empty cell; one workgroup-wide scan gives
//     the final and the raw (grouping) offsets of every (cell, part); the
//     chunk total is published for the decoupled look-back;
//  D. particles are grouped by (cell, part) and ranked inside their group by
//     key, which gives every survivor its final slot (cell base + part base
//     + rank, i.e. (part, key) order); refills are built in their slots;
//  E1. the (cell, vertex) sums are accumulated in final order, each in
//      kSumParts contiguous parts (one thread per part) added in order;
//  C. only then is the chunk's output offset resolved from the totals of
//     the chunks before it (decoupled look-back, its first loads issued
//     before E1), so that the wait overlaps the chunk's own work;
//  E2. the chunk is written out coalesced in final order, with the new
//     cell_start, by every thread after its part of E1.
// A chunk with more than `nmax` particles resolves its offset in B and takes
// the remaining steps with its particles in global memory (grouping through
// the `gord` scratch array, ranking over the group only); it leaves its
// projection sums to k_contrib_dirty.
struct SettleLayout {
  int off_fin;                  // [nslots] final order: x[D], v[D], xi[D] (doubles), src (int), chunk-local
                                // cell (byte); before
                                // the final placement the same bytes hold key (u64), ord (int), xi[D]
                                // and v[D] (doubles) in arrival order and the keys in group order (u64)
  int off_cnt, off_base, off_roff;  // [cpw * pmax] per (cell, part)
  int off_cbuf[2];              // cell data of the current / next chunk
  int off_geo, off_vc, off_prm, off_bpos;  // offsets inside a cell data buffer
  int off_cell;                 // per-cell count and final offset; misc; scan partials
  int bytes;
};

__host__ __device__ inline SettleLayout settle_layout(int D, int cpw, int pmax, int nslots) {
  SettleLayout L;
  int o = 0;
  L.off_fin = o;
  o += (8 * 3 * D + 4 + 1) * nslots;
  o = (o + 15) & ~15;
  L.off_cnt = o;
  o += 4 * cpw * pmax;
  L.off_base = o;
  o += 4 * cpw * pmax;
  L.off_roff = o;
  o += 4 * cpw * pmax;
  int c = 0;
  L.off_geo = c;
  c += 8 * 3 * D * cpw;
  L.off_vc = c;
  c += 8 * (1 << D) * D * cpw;
  L.off_prm = c;
  c += (int)sizeof(Prm) * cpw;
  L.off_bpos = c;
  c += 4 * cpw * (1 << D);
  c = (c + 15) & ~15;
  o = (o + 15) & ~15;
  L.off_cbuf[0] = o;
  o += c;
  L.off_cbuf[1] = o;
  o += c;
  L.off_cell = o;
  o += 4 * (2 * cpw + 32);  // per-cell count, final offset; misc; scan partials
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
        atomicOr(err, kErrLookback);
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

// The value of part j of the lane's group of QP consecutive lanes (QP = 2 or
// 4, groups aligned inside quads of lanes), by a DPP quad permute of each
// 32-bit half.
template <int QP>
__device__ __forceinline__ double quad_from(double v, int j) {
  static_assert(QP == 2 || QP == 4, "parts of a sum within a quad");
  const long long b = __double_as_longlong(v);
  int lo = (int)(b & 0xffffffffLL), hi = (int)(b >> 32);
  // quad_perm controls: lane j of each quad to all four (QP = 4), or lane j
  // of each pair to both (QP = 2)
  switch (QP == 4 ? j : 4 + j) {
    case 0: lo = __builtin_amdgcn_mov_dpp(lo, 0x00, 0xf, 0xf, false); hi = __builtin_amdgcn_mov_dpp(hi, 0x00, 0xf, 0xf, false); break;
    case 1: lo = __builtin_amdgcn_mov_dpp(lo, 0x55, 0xf, 0xf, false); hi = __builtin_amdgcn_mov_dpp(hi, 0x55, 0xf, 0xf, false); break;
    case 2: lo = __builtin_amdgcn_mov_dpp(lo, 0xaa, 0xf, 0xf, false); hi = __builtin_amdgcn_mov_dpp(hi, 0xaa, 0xf, 0xf, false); break;
    case 3: lo = __builtin_amdgcn_mov_dpp(lo, 0xff, 0xf, 0xf, false); hi = __builtin_amdgcn_mov_dpp(hi, 0xff, 0xf, 0xf, false); break;
    case 4: lo = __builtin_amdgcn_mov_dpp(lo, 0xa0, 0xf, 0xf, false); hi = __builtin_amdgcn_mov_dpp(hi, 0xa0, 0xf, 0xf, false); break;
    default: lo = __builtin_amdgcn_mov_dpp(lo, 0xf5, 0xf, 0xf, false); hi = __builtin_amdgcn_mov_dpp(hi, 0xf5, 0xf, 0xf, false); break;
  }
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Inclusive scan of one 32-bit value per lane over the wavefront with DPP
// moves (row shifts inside each row of 16 lanes, then the row broadcasts of
// lanes 15 and 31), no LDS traffic.
__device__ __forceinline__ unsigned wave_scan_u32(unsigned v) {
  v += (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, false);  // row_shr:1
  v += (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, false);  // row_shr:2
  v += (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, false);  // row_shr:4
  v += (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, false);  // row_shr:8
  v += (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false);  // row_bcast:15
  v += (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false);  // row_bcast:31
  return v;
}

// Inclusive-to-exclusive scan of one 64-bit value per thread over the
// workgroup, as two 32-bit halves (the callers' halves are counts that stay
// far below 2^32, so no carry crosses them); `part` holds one value per
// wavefront. Returns the exclusive prefix and sets *total.
__device__ __forceinline__ unsigned long long block_scan_u64(unsigned long long v, unsigned long long* part,
                                                             unsigned long long* total) {
  int tid = (int)threadIdx.x;
  asm volatile("" : "+v"(tid));  // (rebuilt here rather than held: see cells_load)
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const unsigned lo = wave_scan_u32((unsigned)(v & 0xffffffffULL));
  const unsigned hi = wave_scan_u32((unsigned)(v >> 32));
  const unsigned long long incl = ((unsigned long long)hi << 32) | lo;
  if (lane == 63) part[wave] = incl;
  __syncthreads();
  unsigned long long wbase = 0, tot = 0;
  const int nw = blockDim.x >> 6;
  for (int w = 0; w < nw; ++w) {
    const unsigned long long s = part[w];
    if (w < wave) wbase += s;
    tot += s;
  }
  // (`part` is written again only after later barriers of the caller)
  *total = tot;
  return wbase + incl - v;
}

// Status word of chunk k (chunks before the first count as an inclusive
// prefix of 0).
__device__ __forceinline__ unsigned long long lb_word(unsigned long long* lb, int k) {
  return k >= 0 ? __hip_atomic_load(&lb[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : (2ULL << 32);
}

// The same exclusive prefix looked up by one whole wavefront: lane i reads
// the status word of chunk ch-1-i, so one round of loads covers 64
// predecessors; the prefix is the sum over the lanes up to the nearest chunk
// with an inclusive prefix (status 2), once none of them is still unpublished.
// Every lane of the calling wavefront must call it; all get the prefix.
// `first` is this lane's status word of chunk ch-1-lane, loaded by the caller
// ahead of time (lb_word), so that the first round's latency overlaps the
// caller's work in between.
__device__ __forceinline__ unsigned lb_prefix_wave(unsigned long long* lb, int ch, unsigned long long first,
                                                   int* err) {
  const int lane = threadIdx.x & 63;
  unsigned prefix = 0;
  int base = ch - 1;
  long spins = 0;
  bool fresh = true;
  while (base >= 0) {
    const unsigned long long w = fresh ? first : lb_word(lb, base - lane);
    fresh = false;
    const unsigned st = (unsigned)(w >> 32);
    const unsigned long long m2 = __ballot(st == 2);
    const unsigned long long m0 = __ballot(st == 0);
    const int first2 = m2 ? __ffsll((unsigned long long)m2) - 1 : 64;
    const unsigned long long need = first2 >= 63 ? ~0ULL : ((1ULL << (first2 + 1)) - 1ULL);
    if (m0 & need) {
      if (++spins > (1L << 26)) {
        if (lane == 0) atomicOr(err, kErrLookback);
        break;
      }
      __builtin_amdgcn_s_sleep(1);
      continue;
    }
    unsigned part = lane <= first2 ? (unsigned)(w & 0xffffffffULL) : 0u;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) part += __shfl_xor(part, off, 64);
    prefix += part;
    if (first2 < 64) break;
    base -= 64;
  }
  return prefix;
}

// Stored reference position of x in chunk cell c from the LDS geometry (box)
// or the generic path, the part centre for a placed particle (recover_xi_c).
template <int D>
__device__ __forceinline__ void settle_xi(const DMesh& m, const Prm& pr, int c, bool placed,
                                          const double* __restrict__ geo, const double* x, double* xi) {
  CellCache<D> cc;
  cc.c = c;
#pragma unroll
  for (int k = 0; k < D; ++k) {
    cc.X0[k] = geo[k];
    cc.X1[k] = geo[D + k];
    cc.RH[k] = geo[2 * D + k];
  }
  cc.box = (cc.X0[0] == cc.X0[0]);
  inv_cached<D>(m, cc, x, xi);
  if (placed) snap_part<D>(pr, xi);
}

// Particles per thread a staged chunk holds in registers between the
// classification and the final placement.
template <int D>
struct SettlePre {
  static constexpr int n = D == 2 ? 3 : 2;
};

// Settle workgroup: four wavefronts per chunk of 2^kSettleShift cells.
constexpr int kSettleBlock = 256;
constexpr int kSettleShift = 4;
// The look-back (publishing the chunk's total and prefix, and waiting for
// its predecessors) runs on the last wavefront, the one with the fewest of a
// chunk's particles (particle j sits on thread j % kSettleBlock, so the last
// wavefront holds the high j of the second round, usually past the chunk's
// count); thread 0 draws the tickets and reads the chunk headers.
constexpr int kLbLead = kSettleBlock - 64;
// the final-order cell byte holds the chunk-local cell in bits 0-6 and the
// placed flag in bit 7
static_assert(kSettleShift <= 7, "chunk-local cell ids must fit seven bits");
// the arrival-order arrays (key, order, xi, v, group-order key) share the
// bytes of the final-order ones (x, v, xi, src, cell byte)
static_assert(8 + 4 + 16 * 2 + 8 <= 24 * 2 + 5 && 8 + 4 + 16 * 3 + 8 <= 24 * 3 + 5, "arrival arrays exceed the slot");
// The projection sum of a (cell, vertex) is formed in kSumParts<D> parts:
// part j holds the cell's particles at positions [n j / Q, n (j + 1) / Q)
// (integer division) of its n particles in final order, summed sequentially
// from zero, and the parts are added in order j = 0, 1, ... The settle kernel
// takes one thread per (cell, vertex, part) of a chunk, the parts of a sum on
// neighbouring lanes of one wavefront.
template <int D>
constexpr int kSumParts = kSettleBlock / ((1 << kSettleShift) << D);
static_assert(kSumParts<2> == 4 && kSumParts<3> == 2, "parts per projection sum");
static_assert(64 % kSumParts<2> == 0 && 64 % kSumParts<3> == 0, "the parts of a sum share a wavefront");

// Cell data of a chunk (box geometry, advecting field, instance and its
// parameters), carried in registers from its loads to its stores into an LDS
// cell data buffer, one value per thread and array.
constexpr int kPrmWords = (int)(sizeof(Prm) / 8);
static_assert(sizeof(Prm) % 8 == 0 && (1 << kSettleShift) * kPrmWords <= kSettleBlock,
              "Prm words per chunk exceed the workgroup");
static_assert((1 << kSettleShift) * 3 * 3 <= kSettleBlock && (1 << kSettleShift) * 8 * 3 <= 2 * kSettleBlock,
              "cell data per chunk exceeds the workgroup's loads");

template <int D>
struct CellRegs {
  double geo;
  double vc[2];
  double prm;
  int bpos;  // place of a (cell, vertex) projection sum in the d2c order
};

template <int D>
__device__ __forceinline__ void cells_load(const DMesh& m, const double* __restrict__ Vc, int do_refill, int c0,
                                           int ncell, CellRegs<D>& r) {
  // (a handle with one instance keeps its parameter words in both buffers
  // from the start: no dependent load here)
  const bool single = m.cell_inst == nullptr;
  constexpr int NV = 1 << D;
  int t = threadIdx.x;
  // the per-lane addresses below are rebuilt at every call rather than kept
  // across the settle loop (where they would be spilled)
  asm volatile("" : "+v"(t));
  // the parameter words first: their address depends on a load (the cell's
  // instance) in a batch, and the wait for it then covers no other load
  r.prm = (!single && t < ncell * kPrmWords)
              ? reinterpret_cast<const double*>(&m.prm[m.cell_inst[c0 + t / kPrmWords]])[t % kPrmWords]
              : 0.0;
  r.bpos = t < ncell * NV ? m.bpos[(size_t)c0 * NV + t] : 0;
  r.geo = t < ncell * 3 * D ? m.geom[(size_t)c0 * 3 * D + t] : 0.0;
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const int i = t + q * kSettleBlock;
    r.vc[q] = (do_refill && i < ncell * NV * D) ? Vc[(size_t)c0 * NV * D + i] : 0.0;
  }
}

template <int D>
__device__ __forceinline__ void cells_store(const CellRegs<D>& r, bool single, int ncell, double* geo, double* vc,
                                            Prm* prm, int* bpos) {
  constexpr int NV = 1 << D;
  int t = threadIdx.x;
  asm volatile("" : "+v"(t));  // (rebuilt here rather than held: see cells_load)
  if (t < ncell * 3 * D) geo[t] = r.geo;
#pragma unroll
  for (int q = 0; q < 2; ++q)
    if (t + q * kSettleBlock < ncell * NV * D) vc[t + q * kSettleBlock] = r.vc[q];
  if (!single && t < ncell * kPrmWords) reinterpret_cast<double*>(prm)[t] = r.prm;
  if (t < ncell * NV) bpos[t] = r.bpos;
}

// Part index lin of a part code: pi_k = (lin / stride_k) % Q_k with x
// slowest; valid when every pi_k < q_k.
template <int D>
__device__ __forceinline__ bool part_code(const Prm& pr, int lin, int* pi) {
  // the divisions by Q_k as multiplications by Qm_k: exact for lin < 2^16
  // (part buckets per cell stay below 2^8, bounded by the settle's LDS)
  bool ok = true;
  unsigned l = (unsigned)lin;
#pragma unroll
  for (int k = D - 1; k >= 0; --k) {
    const unsigned qd = __umulhi(l, pr.Qm[k]);
    pi[k] = (int)(l - qd * (unsigned)pr.Q[k]);
    l = qd;
    ok = ok && pi[k] < pr.q[k];
  }
  return ok;
}

template <int D>
__global__ __launch_bounds__(kSettleBlock, D == 2 ? 3 : 2) void k_settle(int apply_cap, int do_refill, DMesh m, int nc, int cshift, int pshift,
                                                  int nmax, int nslots, int capc, const int* __restrict__ bcount,
                                                  const int* __restrict__ sstart, long spill2_base, DTmp T,
                                                  int* __restrict__ ticket, unsigned long long* __restrict__ lb,
                                                  int* __restrict__ err, int* __restrict__ cs_out, DParts P,
                                                  const double* __restrict__ Vc, int* __restrict__ gord,
                                                  double* __restrict__ contrib, unsigned char* __restrict__ dirty) {
  constexpr int NV = 1 << D;
  constexpr int PRE = SettlePre<D>::n;
  constexpr int PRV = D == 2 ? 2 : 0;  // register particles whose velocities are prefetched
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int cpw = 1 << cshift;
  const int pmax = 1 << pshift;  // part buckets per cell, padded to a power of two
  const SettleLayout L = settle_layout(D, cpw, pmax, nslots);
  const int ns = nslots;  // row stride of the per-slot arrays
  double* s_fx = reinterpret_cast<double*>(smem_raw + L.off_fin);  // [D][ns]
  double* s_fv = s_fx + D * ns;                                    // [D][ns]
  double* s_fxi = s_fv + D * ns;                                   // [D][ns]
  int* s_fsrc = reinterpret_cast<int*>(s_fxi + D * ns);            // [ns]
  unsigned char* s_fcell = reinterpret_cast<unsigned char*>(s_fsrc + ns);  // [ns]
  // before the final placement: key, group order, reference position and
  // velocity of the staged particles, in arrival order (same bytes)
  unsigned long long* s_key = reinterpret_cast<unsigned long long*>(smem_raw + L.off_fin);  // [nmax]
  int* s_ord = reinterpret_cast<int*>(s_key + ns);                                        // [nmax]
  double* s_sxi = reinterpret_cast<double*>(s_ord + ns);                                  // [D][ns]
  double* s_sv = s_sxi + D * ns;                                                          // [D][ns]
  unsigned long long* s_gkey = reinterpret_cast<unsigned long long*>(s_sv + D * ns);      // [ns] keys in group order
  int* s_cnt = reinterpret_cast<int*>(smem_raw + L.off_cnt);
  int* s_base = reinterpret_cast<int*>(smem_raw + L.off_base);
  int* s_roff = reinterpret_cast<int*>(smem_raw + L.off_roff);
  int* s_ccnt = reinterpret_cast<int*>(smem_raw + L.off_cell);  // particles per cell
  int* s_cfin = s_ccnt + cpw;                                   // final offset per cell (chunk-relative)
  // [0] chunk, [1..3] bucket count, spill start, spill end of the chunk;
  // [4] next chunk, [5..7] its header; [8] chunk prefix; [9] staged
  int* s_misc = s_cfin + cpw;
  unsigned long long* s_part = reinterpret_cast<unsigned long long*>(s_misc + 16);  // [4] scan partials

  const int nchunks = (nc + cpw - 1) >> cshift;
  // with one instance every cell's parameter words are the same: read those
  // of cell 0, one LDS address for the whole wavefront (a broadcast)
  const int psel = m.cell_inst == nullptr ? 0 : ~0;
  const int per_thread = ((pmax << cshift) + blockDim.x - 1) / blockDim.x;

  // registers: the particles of this thread in the current chunk
  int pc[PRE];
  double px[PRE][D];
  double pvr[PRV > 0 ? PRV : 1][D];  // velocities of the first PRV register particles, loaded with their positions
  auto load_parts = [&](int n, int nb, long bucket0, long spill0) {
#pragma unroll
    for (int u = 0; u < PRE; ++u) {
      const int j = threadIdx.x + u * blockDim.x;
      pc[u] = -1;
#pragma unroll
      for (int k = 0; k < D; ++k) px[u][k] = 0.0;
      if (u < PRV)
#pragma unroll
        for (int k = 0; k < D; ++k) pvr[u < PRV ? u : 0][k] = 0.0;
      if (j < n) {
        const long i = chunk_index(j, nb, bucket0, spill0);
        pc[u] = T.cell[i];
#pragma unroll
        for (int k = 0; k < D; ++k) px[u][k] = T.x[k][i];
        if (u < PRV)
#pragma unroll
          for (int k = 0; k < D; ++k) pvr[u < PRV ? u : 0][k] = T.v[k][i];
      }
    }
  };
  const int cb0 = L.off_cbuf[0], cb1 = L.off_cbuf[1];
  auto cbuf_geo = [&](int b) { return reinterpret_cast<double*>(smem_raw + (b ? cb1 : cb0) + L.off_geo); };
  auto cbuf_vc = [&](int b) { return reinterpret_cast<double*>(smem_raw + (b ? cb1 : cb0) + L.off_vc); };
  auto cbuf_prm = [&](int b) { return reinterpret_cast<Prm*>(smem_raw + (b ? cb1 : cb0) + L.off_prm); };
  auto cbuf_bpos = [&](int b) { return reinterpret_cast<int*>(smem_raw + (b ? cb1 : cb0) + L.off_bpos); };

  // prologue: the first chunk's header, cell data and particles
  if (threadIdx.x == 0) {
    const int ch = atomicAdd(ticket, 1);
    s_misc[0] = ch;
    if (ch < nchunks) {
      s_misc[1] = bcount[ch];
      s_misc[2] = sstart[ch];
      s_misc[3] = sstart[ch + 1];
    }
  }
  __syncthreads();
  int buf = 0;
  {
    const int ch = s_misc[0];
    if (ch < nchunks) {
      const int c0 = ch << cshift;
      CellRegs<D> cr;
      cells_load<D>(m, Vc, do_refill, c0, min(nc, c0 + cpw) - c0, cr);
      cells_store<D>(cr, m.cell_inst == nullptr, min(nc, c0 + cpw) - c0, cbuf_geo(0), cbuf_vc(0), cbuf_prm(0),
                     cbuf_bpos(0));
      const int nb = min(s_misc[1], capc);
      load_parts(nb + s_misc[3] - s_misc[2], nb, (long)ch * capc, spill2_base + s_misc[2]);
    }
#pragma unroll
    for (int u = 0; u < PRE; ++u) {  // complete before the loop (as at the end of every iteration)
      asm volatile("" ::"v"(pc[u]));
#pragma unroll
      for (int k = 0; k < D; ++k) asm volatile("" ::"v"(px[u][k]));
      if (u < PRV)
#pragma unroll
        for (int k = 0; k < D; ++k) asm volatile("" ::"v"(pvr[u < PRV ? u : 0][k]));
    }
    for (int i = threadIdx.x; i < cpw * pmax; i += blockDim.x) s_cnt[i] = 0;
    for (int i = threadIdx.x; i < cpw; i += blockDim.x) s_ccnt[i] = 0;
    if (m.cell_inst == nullptr)  // one instance: its parameters for every cell of both buffers, once
      for (int i = threadIdx.x; i < cpw * kPrmWords; i += blockDim.x) {
        const double w = reinterpret_cast<const double*>(m.prm)[i % kPrmWords];
        reinterpret_cast<double*>(cbuf_prm(0))[i] = w;
        reinterpret_cast<double*>(cbuf_prm(1))[i] = w;
      }
  }
  __syncthreads();
  for (;;) {
    const int ch = s_misc[0];
    if (ch >= nchunks) break;
    int next = 0;
    if (threadIdx.x == 0) {  // the next chunk, drawn early and waited for only where its header is read
      // (the counter's index passes through a vector register: with a
      // wave-uniform address the compiler would rebuild the atomic as a wave
      // reduction whose result is read back, and waited for, at once)
      int z = 0;
      asm volatile("" : "+v"(z));
      next = atomicAdd(ticket + z, 1);
    }
    const int c0 = ch << cshift;
    const int c1 = min(nc, c0 + cpw);
    const int ncell = c1 - c0;
    const int nb = min(s_misc[1], capc);
    const int n = nb + s_misc[3] - s_misc[2];
    const long bucket0 = (long)ch * capc;
    const long spill0 = spill2_base + s_misc[2];
    const bool staged_in = n <= nmax;  // particles fit the registers and the key/group arrays
    const double* s_geo = cbuf_geo(buf);
    const double* s_vc = cbuf_vc(buf);
    const Prm* s_prm = cbuf_prm(buf);
    // (the per-(cell, part) and per-cell counts were cleared by the previous
    // chunk's last phase)
    // A. reference position, part and key of every particle; counts per
    //    (cell, part)
    int ppl[PRE];
    double pv[PRE][D];  // velocities: loaded here, stored with the classification
    // (loaded by every lane: a lane past the chunk's particles reads slot 0
    // of the chunk's bucket, in bounds and unused, so that the number of
    // loads in flight is the same on every path and the waits for them stay
    // at their uses)
#pragma unroll
    for (int u = 0; u < PRE; ++u) {
      if (u < PRV) {
#pragma unroll
        for (int k = 0; k < D; ++k) pv[u][k] = pvr[u < PRV ? u : 0][k];
      } else {
        const int j = threadIdx.x + u * blockDim.x;
        const long i = (j < n && staged_in) ? chunk_index(j, nb, bucket0, spill0) : bucket0;
#pragma unroll
        for (int k = 0; k < D; ++k) pv[u][k] = T.v[k][i];
      }
    }
#pragma unroll
    for (int u = 0; u < PRE; ++u) {
      const int j = threadIdx.x + u * blockDim.x;
      ppl[u] = 0;
      if (j < n) {
        const int lc = cell_of(pc[u]) - c0;
        const Prm pr = prm_regs(&s_prm[lc & psel]);
        double xi[D];
        settle_xi<D>(m, pr, cell_of(pc[u]), placed_of(pc[u]), s_geo + lc * 3 * D, px[u], xi);
        const int part = part_lin<D>(pr, xi, nullptr);
        ppl[u] = (lc << 16) | (placed_of(pc[u]) ? 0x8000 : 0) | part;  // part < 2^15 (LDS budget)
        atomicAdd(&s_cnt[(lc << pshift) + part], 1);
        atomicAdd(&s_ccnt[lc], 1);
        if (staged_in) {
          s_key[j] = particle_key<D>(px[u]);
#pragma unroll
          for (int k = 0; k < D; ++k) s_sxi[k * ns + j] = xi[k];
        }
      }
    }
    if (staged_in) {  // the velocity loads overlap the classification above
#pragma unroll
      for (int u = 0; u < PRE; ++u) {
        const int j = threadIdx.x + u * blockDim.x;
        if (j < n)
#pragma unroll
          for (int k = 0; k < D; ++k) s_sv[k * ns + j] = pv[u][k];
      }
    }
    for (int j = threadIdx.x + PRE * blockDim.x; j < n; j += blockDim.x) {  // beyond the registers: not staged
      const long i = chunk_index(j, nb, bucket0, spill0);
      const int cw = T.cell[i];
      const int c = cell_of(cw);
      const int lc = c - c0;
      const Prm pr = prm_regs(&s_prm[lc & psel]);
      double x[D], xi[D];
#pragma unroll
      for (int k = 0; k < D; ++k) x[k] = T.x[k][i];
      settle_xi<D>(m, pr, c, placed_of(cw), s_geo + lc * 3 * D, x, xi);
      const int part = part_lin<D>(pr, xi, nullptr);
      atomicAdd(&s_cnt[(lc << pshift) + part], 1);
      atomicAdd(&s_ccnt[lc], 1);
    }
    int nh0 = 0, nh1 = 0, nh2 = 0;
    if (threadIdx.x == 0 && next < nchunks) {  // header of the next chunk, read while this one runs
      nh0 = bcount[next];
      nh1 = sstart[next];
      nh2 = sstart[next + 1];
    }
    __syncthreads();
    // B. final (survivors + refill) and raw counts per (cell, part), scanned
    //    over the chunk in (cell, part) order; the chunk total is published
    // (a thread's per_thread entries lie in one cell: per_thread divides
    // pmax)
    unsigned long long mine = 0;
    {
      const int t0 = threadIdx.x * per_thread;
      const int lc = t0 >> pshift;
      const Prm pr = prm_regs(&s_prm[(lc < cpw ? lc : 0) & psel]);
      const int ccnt = s_ccnt[lc < cpw ? lc : 0];
      for (int e = 0; e < per_thread; ++e) {
        const int t = t0 + e;
        if (t >= ncell * pmax) break;
        const int k = s_cnt[t];
        int pi[D];
        const bool valid = ((t & (pmax - 1)) < pr.P) && part_code<D>(pr, t & (pmax - 1), pi);
        const int kept = apply_cap ? min(k, pr.maxpp + 1) : k;
        const int refill = (do_refill && ccnt > 0 && k == 0 && valid) ? 1 : 0;
        s_base[t] = kept + refill;
        mine += ((unsigned long long)(kept + refill) << 32) | (unsigned)k;
      }
    }
    unsigned long long tot = 0;
    unsigned long long run = block_scan_u64(mine, s_part, &tot);
    for (int e = 0; e < per_thread; ++e) {
      const int t = threadIdx.x * per_thread + e;
      if (t >= ncell * pmax) break;
      const int fin = s_base[t];
      const int k = s_cnt[t];
      s_base[t] = (int)(run >> 32);
      s_roff[t] = (int)(run & 0xffffffffULL);
      if ((t & (pmax - 1)) == 0) s_cfin[t >> pshift] = (int)(run >> 32);
      run += ((unsigned long long)fin << 32) | (unsigned)k;
    }
    const int nout = (int)(tot >> 32);
    const bool staged = staged_in && nout <= nslots;
    if (threadIdx.x == kLbLead) {
      lb_publish(lb, ch, ch == 0 ? 2u : 1u, (unsigned)nout);
      if (ch == 0) s_misc[8] = 0;
      if (!staged && ch > 0) {  // the global-memory path writes as it goes
        const unsigned prefix = lb_prefix(lb, ch, err);
        lb_publish(lb, ch, 2u, prefix + (unsigned)nout);
        s_misc[8] = (int)prefix;
      }
    }
    if (threadIdx.x == 0) {
      s_misc[4] = next;
      s_misc[5] = nh0;
      s_misc[6] = nh1;
      s_misc[7] = nh2;
    }
    __syncthreads();
    // D. group by (cell, part) (s_roff becomes the end of each group), rank
    //    by key inside the group
    if (staged) {
#pragma unroll
      for (int u = 0; u < PRE; ++u) {
        const int j = threadIdx.x + u * blockDim.x;
        if (j < n) {
          const int t = ((ppl[u] >> 16) << pshift) + (ppl[u] & 0x7fff);
          const int g = atomicAdd(&s_roff[t], 1);
          s_ord[g] = j;
          s_gkey[g] = s_key[j];  // the ranking reads the group's keys without the indirection
        }
      }
    } else {
      for (int j = threadIdx.x; j < n; j += blockDim.x) {
        const long i = chunk_index(j, nb, bucket0, spill0);
        const int cw = T.cell[i];
        const int c = cell_of(cw);
        const int lc = c - c0;
        double x[D], xi[D];
#pragma unroll
        for (int k = 0; k < D; ++k) x[k] = T.x[k][i];
        settle_xi<D>(m, s_prm[lc & psel], c, placed_of(cw), s_geo + lc * 3 * D, x, xi);
        const int t = (lc << pshift) + part_lin<D>(s_prm[lc & psel], xi, nullptr);
        gord[chunk_index(atomicAdd(&s_roff[t], 1), nb, bucket0, spill0)] = j;
      }
    }
    __syncthreads();
    const int o0n = staged ? 0 : s_misc[8];
    int pf[PRE];  // final slot of each register particle (-1: capped away)
    if (staged) {
#pragma unroll
      for (int u = 0; u < PRE; ++u) {
        const int j = threadIdx.x + u * blockDim.x;
        pf[u] = -1;
        if (j < n) {
          const int lc = ppl[u] >> 16;
          const int t = (lc << pshift) + (ppl[u] & 0x7fff);
          const unsigned long long key = s_key[j];
          const int b = s_roff[t], a = b - s_cnt[t];
          const int maxpp = s_prm[lc & psel].maxpp;
          // rank = keys below this one in the group, four keys read per
          // round; equal keys (the particle's own and any other) counted too
          int r = 0, eq = 0;
          for (int q = a; q < b; q += 4) {
            unsigned long long kq[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) kq[e] = s_gkey[min(q + e, b - 1)];
#pragma unroll
            for (int e = 0; e < 4; ++e)
              if (q + e < b) {
                r += kq[e] < key ? 1 : 0;
                eq += kq[e] == key ? 1 : 0;
              }
          }
          if (eq > 1) {  // another particle with the same key: position, velocity, index decide (rare)
            r = 0;
            for (int q = a; q < b; ++q) {
              const unsigned long long kq = s_gkey[q];
              if (kq < key) {
                ++r;
              } else if (kq == key) {
                const int o = s_ord[q];
                if (o == j) continue;
                double xq[D], vq[D];
                const long iq = chunk_index(o, nb, bucket0, spill0);
#pragma unroll
                for (int k = 0; k < D; ++k) {
                  xq[k] = T.x[k][iq];
                  vq[k] = T.v[k][iq];
                }
                double vj[D];
#pragma unroll
                for (int k = 0; k < D; ++k) vj[k] = s_sv[k * ns + j];
                if (tie_less<D>(xq, vq, px[u], vj, o, j)) ++r;
              }
            }
          }
          if (!(apply_cap && r > maxpp)) pf[u] = s_base[t] + r;
        }
      }
    } else {
      for (int j = threadIdx.x; j < n; j += blockDim.x) {
        const long i = chunk_index(j, nb, bucket0, spill0);
        const int cw = T.cell[i];
        const int c = cell_of(cw);
        const int lc = c - c0;
        double xj[D], vj[D], xi[D];
#pragma unroll
        for (int k = 0; k < D; ++k) {
          xj[k] = T.x[k][i];
          vj[k] = T.v[k][i];
        }
        settle_xi<D>(m, s_prm[lc & psel], c, placed_of(cw), s_geo + lc * 3 * D, xj, xi);
        const int t = (lc << pshift) + part_lin<D>(s_prm[lc & psel], xi, nullptr);
        const unsigned long long key = particle_key<D>(xj);
        const int b = s_roff[t], a = b - s_cnt[t];
        int r = 0;
        for (int q = a; q < b; ++q) {
          const int o = gord[chunk_index(q, nb, bucket0, spill0)];
          if (o == j) continue;
          const long iq = chunk_index(o, nb, bucket0, spill0);
          double xq[D], vq[D];
#pragma unroll
          for (int k = 0; k < D; ++k) {
            xq[k] = T.x[k][iq];
            vq[k] = T.v[k][iq];
          }
          const unsigned long long kq = particle_key<D>(xq);
          if (kq < key || (kq == key && tie_less<D>(xq, vq, xj, vj, o, j))) ++r;
        }
        if (apply_cap && r > s_prm[lc & psel].maxpp) continue;
        const int fpos = o0n + s_base[t] + r;
#pragma unroll
        for (int k = 0; k < D; ++k) {
          P.x[k][fpos] = xj[k];
          P.v[k][fpos] = vj[k];
        }
        P.cell[fpos] = cw;
        if (P.src) P.src[fpos] = T.src[i];
      }
    }
    double qxi[PRE][D], qv[PRE][D];
    if (staged) {
#pragma unroll
      for (int u = 0; u < PRE; ++u) {
        const int j = threadIdx.x + u * blockDim.x;
#pragma unroll
        for (int k = 0; k < D; ++k) {
          qxi[u][k] = pf[u] >= 0 ? s_sxi[k * ns + j] : 0.0;
          qv[u][k] = pf[u] >= 0 ? s_sv[k * ns + j] : 0.0;
        }
      }
    }
    // C. output offset of a staged chunk by decoupled look-back on the last
    //    wavefront: its first round of status loads is issued here and
    //    waited for after the wavefront's placement and refills
    // (the source indices of a move that records them are read before the
    // look-back's loads, so that the placement below waits for neither)
    int psrc[PRE];
#pragma unroll
    for (int u = 0; u < PRE; ++u) {
      psrc[u] = 0;
      if (P.src && staged && pf[u] >= 0) psrc[u] = T.src[chunk_index(threadIdx.x + u * blockDim.x, nb, bucket0, spill0)];
    }
    const bool lb_on = staged && (int)threadIdx.x >= kLbLead && (int)threadIdx.x < kLbLead + 64 && ch > 0;
    unsigned long long lbw = 0;
    if (lb_on) lbw = lb_word(lb, ch - 1 - ((int)threadIdx.x & 63));
    __syncthreads();  // the arrival-order arrays are dead from here on
    // final placement of the staged survivors
    if (staged) {
#pragma unroll
      for (int u = 0; u < PRE; ++u) {
        const int f = pf[u];
        if (f >= 0) {
#pragma unroll
          for (int k = 0; k < D; ++k) {
            s_fx[k * ns + f] = px[u][k];
            s_fv[k * ns + f] = qv[u][k];
            s_fxi[k * ns + f] = qxi[u][k];
          }
          s_fcell[f] = (unsigned char)((ppl[u] >> 16) | ((ppl[u] & 0x8000) ? 0x80 : 0));
          if (P.src) s_fsrc[f] = psrc[u];
        }
      }
    }
    // refills: one particle at the centre of each empty part of a non-empty
    // cell, velocity interpolated from the advecting field
    if (do_refill) {
      for (int t = threadIdx.x; t < ncell * pmax; t += blockDim.x) {
        const int lc = t >> pshift, pp = t & (pmax - 1);
        if (s_cnt[t] != 0 || s_ccnt[lc] == 0) continue;
        const Prm pr = prm_regs(&s_prm[lc & psel]);
        int pi[D];
        if (pp >= pr.P || !part_code<D>(pr, pp, pi)) continue;  // padding or no refillable part
        const int c = c0 + lc;
        double xi[D], x[D], v[D];
#pragma unroll
        for (int k = 0; k < D; ++k) xi[k] = ((double)pi[k] + 1.0 / 2) * pr.h[k];
        const double* g = s_geo + lc * 3 * D;
        if (g[0] == g[0]) {
          CellCache<D> cc;
#pragma unroll
          for (int k = 0; k < D; ++k) {
            cc.X0[k] = g[k];
            cc.X1[k] = g[D + k];
          }
          fwd_box<D>(cc, xi, x);
        } else {
          int vid[NV];
          load_ids<D>(m.cv, c, vid);
          fwd<D>(m, vid, xi, x);
        }
        double N[NV];
        shape<D>(xi, N);
        const double* fv = s_vc + lc * NV * D;
#pragma unroll
        for (int k = 0; k < D; ++k) v[k] = 0.0;
#pragma unroll
        for (int i = 0; i < NV; ++i)
#pragma unroll
          for (int k = 0; k < D; ++k) v[k] = v[k] + N[i] * fv[i * D + k];
        const int f = s_base[t];
        if (staged) {
#pragma unroll
          for (int k = 0; k < D; ++k) {
            s_fx[k * ns + f] = x[k];
            s_fv[k * ns + f] = v[k];
            s_fxi[k * ns + f] = xi[k];
          }
          s_fsrc[f] = -1 - pp;
          s_fcell[f] = (unsigned char)(lc | 0x80);  // placed at the part centre
        } else {
#pragma unroll
          for (int k = 0; k < D; ++k) {
            P.x[k][o0n + f] = x[k];
            P.v[k][o0n + f] = v[k];
          }
          P.cell[o0n + f] = c | kPlaced;
          if (P.src) P.src[o0n + f] = -1 - pp;
        }
      }
    }
    if (lb_on) {  // (its loads complete before the next chunk's, issued below)
      const unsigned prefix = lb_prefix_wave(lb, ch, lbw, err);
      if (threadIdx.x == kLbLead) {
        lb_publish(lb, ch, 2u, prefix + (unsigned)nout);
        s_misc[8] = (int)prefix;
      }
    }
    __syncthreads();  // final-order arrays complete, the output offset known
    const int o0 = s_misc[8];
    // the next chunk's loads first, so that they arrive during the sums and
    // the write-out: particles into the registers (free since the
    // placement), cell data into registers and from there into the other
    // buffer at the end
    const int nch = s_misc[4];
    const int nc0 = nch << cshift;
    const int nncell = nch < nchunks ? min(nc, nc0 + cpw) - nc0 : 0;
    CellRegs<D> cr;
    if (nch < nchunks) {
      cells_load<D>(m, Vc, do_refill, nc0, nncell, cr);
      const int nnb = min(s_misc[5], capc);
      load_parts(nnb + s_misc[7] - s_misc[6], nnb, (long)nch * capc, spill2_base + s_misc[6]);
    }
    // The next chunk's cell data into the other buffer, before any store of
    // this chunk (a wait for these loads also waits for every store issued
    // before it); its particles, loaded with them, are complete too: waiting
    // for them here (an empty asm that reads them) keeps the next iteration's
    // first wait from covering its velocity loads.
    auto next_chunk_in = [&]() {
      if (nch < nchunks)
        cells_store<D>(cr, m.cell_inst == nullptr, nncell, cbuf_geo(buf ^ 1), cbuf_vc(buf ^ 1), cbuf_prm(buf ^ 1),
                       cbuf_bpos(buf ^ 1));
#pragma unroll
      for (int u = 0; u < PRE; ++u) {
        asm volatile("" ::"v"(pc[u]));
#pragma unroll
        for (int k = 0; k < D; ++k) asm volatile("" ::"v"(px[u][k]));
      if (u < PRV)
#pragma unroll
        for (int k = 0; k < D; ++k) asm volatile("" ::"v"(pvr[u < PRV ? u : 0][k]));
      }
    };
    if (staged) {
      // E1. per (cell, vertex) sums S_k = sum N_i v_k and W = sum N_i over
      //     the cell's particles in final order, each formed in kSumParts
      //     contiguous parts (sequential inside a part, the parts added in
      //     order): one thread per (cell, vertex, part), the parts of one
      //     sum on neighbouring lanes
      constexpr int QP = kSumParts<D>;
      constexpr int G = 4;  // the loads of a group of particles issued together
      // (the lane's chain and part are rebuilt from the thread index here:
      // values held across the chunk loop are spilled, and a reload at this
      // point would wait for the next chunk's loads issued above)
      int tix = (int)threadIdx.x;
      asm volatile("" : "+v"(tix));
      const int chain = tix / QP, part = tix % QP;
      const int lc = chain >> D, iv = chain & (NV - 1);
      int a = 0, b = 0;
      if (lc < ncell) {
        const int ca = s_cfin[lc], cb = (lc + 1 < ncell) ? s_cfin[lc + 1] : nout;
        a = ca + (part * (cb - ca)) / QP;
        b = ca + ((part + 1) * (cb - ca)) / QP;
      }
      double S[D], W = 0.0;
#pragma unroll
      for (int k = 0; k < D; ++k) S[k] = 0.0;
      int f = a;
      for (; f + G <= b; f += G) {
        double N[G], w[G][D];
#pragma unroll
        for (int e = 0; e < G; ++e) {
#pragma unroll
          for (int k = 0; k < D; ++k) {
            const double xk = s_fxi[k * ns + f + e];
            const double l = ((iv >> k) & 1) ? xk : (1.0 - xk);
            N[e] = (k == 0) ? l : N[e] * l;
            w[e][k] = s_fv[k * ns + f + e];
          }
        }
#pragma unroll
        for (int e = 0; e < G; ++e) {
#pragma unroll
          for (int k = 0; k < D; ++k) S[k] = S[k] + N[e] * w[e][k];
          W = W + N[e];
        }
      }
      {  // the last G - 1 or fewer, their loads issued together
        double N[G - 1], w[G - 1][D];
#pragma unroll
        for (int e = 0; e < G - 1; ++e) {
          const int g = f + e < b ? f + e : f;
#pragma unroll
          for (int k = 0; k < D; ++k) {
            const double xk = s_fxi[k * ns + g];
            const double l = ((iv >> k) & 1) ? xk : (1.0 - xk);
            N[e] = (k == 0) ? l : N[e] * l;
            w[e][k] = s_fv[k * ns + g];
          }
        }
#pragma unroll
        for (int e = 0; e < G - 1; ++e) {
          if (f + e < b) {
#pragma unroll
            for (int k = 0; k < D; ++k) S[k] = S[k] + N[e] * w[e][k];
            W = W + N[e];
          }
        }
      }
      // the parts of this sum, read from their lanes (the QP lanes of a sum
      // are aligned inside a quad of lanes) with DPP quad permutes, added in
      // order
      double PS[QP][D + 1];
#pragma unroll
      for (int j = 0; j < QP; ++j) {
#pragma unroll
        for (int k = 0; k < D; ++k) PS[j][k] = quad_from<QP>(S[k], j);
        PS[j][D] = quad_from<QP>(W, j);
      }
      double T[D], TW = PS[0][D];
#pragma unroll
      for (int k = 0; k < D; ++k) T[k] = PS[0][k];
#pragma unroll
      for (int j = 1; j < QP; ++j) {
#pragma unroll
        for (int k = 0; k < D; ++k) T[k] = T[k] + PS[j][k];
        TW = TW + PS[j][D];
      }
      // (the next chunk's cell data goes to the other buffer here, between
      // the sums and their store, see below)
      next_chunk_in();
      if (part == 0 && lc < ncell) {
        double* co = contrib + (size_t)cbuf_bpos(buf)[lc * NV + iv] * (D + 1);
#pragma unroll
        for (int k = 0; k < D; ++k) co[k] = T[k];
        co[D] = TW;
      }
    } else {
      next_chunk_in();
    }
    if (staged) {
      // E2. coalesced write-out in final order, two particles per thread
      //     and round (their LDS reads issued together)
      for (int t0 = threadIdx.x; t0 < nout; t0 += 2 * kSettleBlock) {
        double ox[2][D], ov[2][D];
        int fc[2], fs[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int t = min(t0 + h * kSettleBlock, nout - 1);
#pragma unroll
          for (int k = 0; k < D; ++k) {
            ox[h][k] = s_fx[k * ns + t];
            ov[h][k] = s_fv[k * ns + t];
          }
          fc[h] = s_fcell[t];
          fs[h] = s_fsrc[t];
        }
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int t = t0 + h * kSettleBlock;
          if (t < nout) {
#pragma unroll
            for (int k = 0; k < D; ++k) {
              P.x[k][o0 + t] = ox[h][k];
              P.v[k][o0 + t] = ov[h][k];
            }
            P.cell[o0 + t] = (c0 + (fc[h] & 0x7f)) | ((fc[h] & 0x80) ? kPlaced : 0);
            if (P.src) P.src[o0 + t] = fs[h];
          }
        }
      }
    }
    for (int lc = threadIdx.x; lc < ncell; lc += blockDim.x) cs_out[c0 + lc] = o0 + s_cfin[lc];
    if (c1 == nc && threadIdx.x == 0) cs_out[nc] = o0 + nout;
    // the counts of the next chunk start at zero (read last by the refills)
    for (int i = threadIdx.x; i < cpw * pmax; i += blockDim.x) s_cnt[i] = 0;
    for (int i = threadIdx.x; i < cpw; i += blockDim.x) s_ccnt[i] = 0;
    if (threadIdx.x == 0) {
      dirty[ch] = staged ? 0 : 1;
      s_misc[0] = s_misc[4];
      s_misc[1] = s_misc[5];
      s_misc[2] = s_misc[6];
      s_misc[3] = s_misc[7];
    }
    buf ^= 1;
    __syncthreads();
  }
}

// ------------------------------------------------------------------ maps
// initialize_maps (src/pfem2particlehandler.cpp:319-324): for every vertex
// the cells that contain it, ascending and each once (a degenerate cell can
// list a vertex twice), and the unit direction from the vertex to each such
// cell's centre (vertex average), built on the device: count, scatter,
// per-vertex sort and de-duplication, then the compacted lists with their
// directions; plus the per (cell, vertex) views of the lists used by the walk.

__global__ void k_v2c_count(const int* __restrict__ cv, long ne, int* __restrict__ cnt) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < ne; e += (long)gridDim.x * blockDim.x)
    atomicAdd(&cnt[cv[e]], 1);
}

__global__ void k_v2c_fill(const int* __restrict__ cv, long ne, int nvc, const int* __restrict__ off,
                           int* __restrict__ fill, int* __restrict__ lst) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < ne; e += (long)gridDim.x * blockDim.x) {
    const int v = cv[e];
    lst[off[v] + atomicAdd(&fill[v], 1)] = (int)(e / nvc);
  }
}

// each vertex's list sorted ascending (insertion sort: lists are short),
// repeats dropped in place; cnt2[v] = cells kept
__global__ void k_v2c_sort(int nv, const int* __restrict__ off, int* __restrict__ lst, int* __restrict__ cnt2) {
  for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < nv; v += gridDim.x * blockDim.x) {
    const int a = off[v], b = off[v + 1];
    for (int i = a + 1; i < b; ++i) {
      const int key = lst[i];
      int j = i - 1;
      while (j >= a && lst[j] > key) {
        lst[j + 1] = lst[j];
        --j;
      }
      lst[j + 1] = key;
    }
    int w = a;
    for (int i = a; i < b; ++i)
      if (i == a || lst[i] != lst[w - 1]) lst[w++] = lst[i];
    cnt2[v] = w - a;
  }
}

// compacted lists and directions d = c - X_v, c = (sum_i X_i) * (1 / 2^D)
// (sequential sum over the cell's vertices), normalised as d * (1 / |d|)
template <int D>
__global__ void k_v2c_emit(DMesh m, int nv, const int* __restrict__ off, const int* __restrict__ lst,
                           const int* __restrict__ off2, int* __restrict__ lst2, double* __restrict__ dirs,
                           double inv_n) {
  constexpr int NV = 1 << D;
  for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < nv; v += gridDim.x * blockDim.x) {
    const int a = off[v], o = off2[v], cnt = off2[v + 1] - o;
    for (int j = 0; j < cnt; ++j) {
      const int c = lst[a + j];
      lst2[o + j] = c;
      double ctr[D], d[D];
#pragma unroll
      for (int k = 0; k < D; ++k) ctr[k] = 0.0;
#pragma unroll
      for (int i = 0; i < NV; ++i) {
        const int vi = m.cv[(size_t)c * NV + i];
#pragma unroll
        for (int k = 0; k < D; ++k) ctr[k] = ctr[k] + m.X[k][vi];
      }
#pragma unroll
      for (int k = 0; k < D; ++k) d[k] = ctr[k] * inv_n - m.X[k][v];
      double m2 = 0.0;
#pragma unroll
      for (int k = 0; k < D; ++k) m2 = m2 + d[k] * d[k];
      const double inv_m = 1.0 / sqrt(m2);
#pragma unroll
      for (int k = 0; k < D; ++k) dirs[(size_t)(o + j) * D + k] = d[k] * inv_m;
    }
  }
}

// per (cell, vertex): (offset, count) of the vertex's list, and the list
// itself when it has at most 2^dim cells (first id -2 otherwise), -1 padded:
// one int4 per (cell, vertex) in 2D, two in 3D
__global__ void k_cell_vertex_maps(long ne, int nvc, const int* __restrict__ cv, const int* __restrict__ off2,
                                   const int* __restrict__ lst2, int2* __restrict__ cvo, int4* __restrict__ cvc) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < ne; e += (long)gridDim.x * blockDim.x) {
    const int v = cv[e];
    const int o = off2[v], n = off2[v + 1] - o;
    cvo[e] = make_int2(o, n);
    int ids[8] = {-1, -1, -1, -1, -1, -1, -1, -1};
    if (n > nvc)
      ids[0] = -2;
    else
      for (int j = 0; j < n; ++j) ids[j] = lst2[o + j];
    cvc[e * (nvc / 4)] = make_int4(ids[0], ids[1], ids[2], ids[3]);
    if (nvc == 8) cvc[e * 2 + 1] = make_int4(ids[4], ids[5], ids[6], ids[7]);
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
  constexpr int QP = kSumParts<D>;
  double T[D], TW = 0.0;
#pragma unroll
  for (int k = 0; k < D; ++k) T[k] = 0.0;
  const int s = cs[c], n = cs[c + 1] - s;
  for (int j = 0; j < QP; ++j) {  // the parts of the settle kernel's sums
    double S[D], W = 0.0;
#pragma unroll
    for (int k = 0; k < D; ++k) S[k] = 0.0;
    for (int p = s + (j * n) / QP; p < s + ((j + 1) * n) / QP; ++p) {
      double x[D], xi[D], N[NV];
#pragma unroll
      for (int k = 0; k < D; ++k) x[k] = P.x[k][p];
      recover_xi_c<D>(m, pr, c, placed_of(P.cell[p]), x, xi);
      shape<D>(xi, N);
      double Ni = N[0];
#pragma unroll
      for (int q = 1; q < NV; ++q)
        if (q == i) Ni = N[q];
#pragma unroll
      for (int k = 0; k < D; ++k) S[k] = S[k] + Ni * P.v[k][p];
      W = W + Ni;
    }
    if (j == 0) {
#pragma unroll
      for (int k = 0; k < D; ++k) T[k] = S[k];
      TW = W;
    } else {
#pragma unroll
      for (int k = 0; k < D; ++k) T[k] = T[k] + S[k];
      TW = TW + W;
    }
  }
  double* e_out = contrib + (size_t)m.bpos[(size_t)c * NV + i] * (D + 1);
#pragma unroll
  for (int k = 0; k < D; ++k) e_out[k] = T[k];
  e_out[D] = TW;
}

template <int D>
__global__ __launch_bounds__(kBlock) void k_contrib(DMesh m, int nc, const int* __restrict__ cs, DParts P,
                                                   double* __restrict__ contrib) {
  constexpr int NV = 1 << D;
  const long total = (long)nc * NV;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x)
    contrib_one<D>(m, (int)(t / NV), (int)(t % NV), cs, P, contrib);
}

// The same sums for the cells of chunks the settle kernel could not stage
// (rare): a workgroup reads the flags of blockDim.x chunks in one coalesced
// pass, lists its flagged chunks in LDS and sums each with all its threads.
template <int D>
__global__ __launch_bounds__(kBlock) void k_contrib_dirty(DMesh m, int nc, int cpw, const unsigned char* __restrict__ dirty,
                                                         const int* __restrict__ cs, DParts P,
                                                         double* __restrict__ contrib) {
  constexpr int NV = 1 << D;
  __shared__ int s_list[kBlock];
  __shared__ int s_n;
  const int nchunks = (nc + cpw - 1) / cpw;
  for (int base = blockIdx.x * blockDim.x; base < nchunks; base += gridDim.x * blockDim.x) {
    if (threadIdx.x == 0) s_n = 0;
    __syncthreads();
    const int ch = base + threadIdx.x;
    if (ch < nchunks && dirty[ch]) s_list[atomicAdd(&s_n, 1)] = ch;
    __syncthreads();
    const int nd = s_n;
    for (int l = 0; l < nd; ++l) {
      const int c0 = s_list[l] * cpw, ncell = min(nc, c0 + cpw) - c0;
      for (int t = threadIdx.x; t < ncell * NV; t += blockDim.x) contrib_one<D>(m, c0 + t / NV, t % NV, cs, P, contrib);
    }
    __syncthreads();
  }
}

// V_k[j] = (sum over incident cells, ascending id, of S_k) / (same for W);
// 0/0 gives NaN as in src/pfem2particlehandler.cpp:207-209. The sums of
// DoF j's (cell, vertex) pairs are the contiguous run d2c_off[j] ..
// d2c_off[j+1] of contrib (stored in d2c order).
template <int D>
__global__ __launch_bounds__(kBlock) void k_node(int ndofs, const int* __restrict__ d2c_off,
                                                const double* __restrict__ contrib, double* o0, double* o1,
                                                double* o2) {
  double* out[3] = {o0, o1, o2};
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < ndofs; j += gridDim.x * blockDim.x) {
    double S[D], W = 0.0;
#pragma unroll
    for (int k = 0; k < D; ++k) S[k] = 0.0;
    for (int e = d2c_off[j]; e < d2c_off[j + 1]; ++e) {
      const double* c = contrib + (size_t)e * (D + 1);
#pragma unroll
      for (int k = 0; k < D; ++k) S[k] = S[k] + c[k];
      W = W + c[D];
    }
#pragma unroll
    for (int k = 0; k < D; ++k) out[k][j] = S[k] / W;
  }
}

// ------------------------------------------------------ per-particle kernels

// Removal of listed particles (remove_particle, an erase that keeps the
// order of the others, src/pfem2particlehandler.cpp:332-336): keep flags,
// then each kept particle moves to the exclusive prefix of the flags; the
// cell segments follow because the particles are sorted by cell.
__global__ void k_keep_flags(const int* __restrict__ n_ptr, int* __restrict__ keep) {
  const int n = *n_ptr;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) keep[i] = 1;
}

__global__ void k_drop_listed(long n_rm, const long long* __restrict__ rm, int* __restrict__ keep) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n_rm; i += (long)gridDim.x * blockDim.x)
    keep[rm[i]] = 0;
}

template <int D>
__global__ __launch_bounds__(kBlock) void k_compact_kept(const int* __restrict__ n_ptr, int nc,
                                                        const int* __restrict__ keep, const int* __restrict__ pos,
                                                        const int* __restrict__ cs, int* __restrict__ cs_new,
                                                        DParts P, DTmp T) {
  const int n = *n_ptr;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i <= (long)nc; i += (long)gridDim.x * blockDim.x)
    cs_new[i] = pos[cs[i]];
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
    if (!keep[p]) continue;
    const int q = pos[p];
#pragma unroll
    for (int k = 0; k < D; ++k) {
      T.x[k][q] = P.x[k][p];
      T.v[k][q] = P.v[k][p];
    }
    T.cell[q] = P.cell[p];
  }
}

// v += interp(dV) (flush of a pending correction outside move)
template <int D>
__global__ __launch_bounds__(kBlock) void k_correct(DMesh m, const int* __restrict__ n_ptr, DParts P,
                                                   const double* __restrict__ dVc) {
  const int n = *n_ptr;
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
    const int w = P.cell[p];
    const int c = cell_of(w);
    const Prm pr = m.prm[inst_of(m, c)];
    double x[D], xi[D], corr[D];
#pragma unroll
    for (int k = 0; k < D; ++k) x[k] = P.x[k][p];
    recover_xi_c<D>(m, pr, c, placed_of(w), x, xi);
    gather_table<D>(dVc, c, xi, corr);
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
    const int w = P.cell[p];
    const int c = cell_of(w);
    const Prm pr = m.prm[inst_of(m, c)];
    double x[D], xi[D], u[D];
#pragma unroll
    for (int k = 0; k < D; ++k) x[k] = P.x[k][p];
    int dof[NV];
    load_ids<D>(m.cd, c, dof);
    recover_xi_c<D>(m, pr, c, placed_of(w), x, xi);
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
    const int w = P.cell[p];
    const int c = cell_of(w);
    const Prm pr = m.prm[inst_of(m, c)];
    double x[D], xi[D];
#pragma unroll
    for (int k = 0; k < D; ++k) x[k] = P.x[k][p];
    recover_xi_c<D>(m, pr, c, placed_of(w), x, xi);
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
      if (ok && at_part_centre<D>(m, m.prm[inst_of(m, c)], vid, x, xi)) c |= kPlaced;
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
