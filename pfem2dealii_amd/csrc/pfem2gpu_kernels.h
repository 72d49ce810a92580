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
// Pass 0 counts particles per destination cell; pass 1 recomputes the same
// trajectory bit for bit and scatters the particle into its destination cell
// segment with its part bucket and canonical key. `advect` = 0 turns both
// passes into a pure re-binning (used by insertion). The velocity correction
// (if pending) is applied in pass 1 from the pre-move cell and reference
// position, as correct_particle_velocity() precedes move_particles().
template <int D>
__global__ __launch_bounds__(kBlock) void k_move(int pass, int advect, DMesh m, const int* __restrict__ n_ptr,
                                                int n_add, DParts P, DField Vf, DField dV, int use_dv,
                                                int* __restrict__ count, const int* __restrict__ tstart,
                                                int* __restrict__ fill, DTmp T) {
  constexpr int NV = 1 << D;
  const int n = *n_ptr + n_add;
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
    int c = P.cell[p];
    if (c < 0) continue;
    const Prm pr = m.prm[inst_of(m, c)];
    double x[D], xi[D];
#pragma unroll
    for (int k = 0; k < D; ++k) x[k] = P.x[k][p];
    int vid[NV];
    load_ids<D>(m.cv, c, vid);
    recover_xi<D>(m, pr, c, vid, x, xi);
    double v[D];
    if (pass == 1) {
#pragma unroll
      for (int k = 0; k < D; ++k) v[k] = P.v[k][p];
      if (use_dv) {
        int dof[NV];
        load_ids<D>(m.cd, c, dof);
        double corr[D];
        gather<D>(dV, dof, xi, corr);
#pragma unroll
        for (int k = 0; k < D; ++k) v[k] = v[k] + corr[k];
      }
    }
    bool alive = true;
    if (advect) {
      for (int s = 0; s < pr.n_sub; ++s) {
        int dof[NV];
        load_ids<D>(m.cd, c, dof);
        double u[D];
        gather<D>(Vf, dof, xi, u);
#pragma unroll
        for (int k = 0; k < D; ++k) {
          const double t = u[k] * pr.dt_sub;
          x[k] = x[k] + t;
        }
        if (!locate<D>(m, c, vid, xi, x, u)) {
          alive = false;
          break;
        }
      }
    }
    if (!alive) continue;
    if (pass == 0) {
      atomicAdd(&count[c], 1);
      continue;
    }
    const Prm pr2 = m.prm[inst_of(m, c)];
    const int part = part_lin<D>(pr2, xi, nullptr);
    const int slot = tstart[c] + atomicAdd(&fill[c], 1);
#pragma unroll
    for (int k = 0; k < D; ++k) {
      T.x[k][slot] = x[k];
      T.v[k][slot] = v[k];
    }
    T.cell[slot] = c;
    T.part[slot] = (unsigned short)part;
    T.key[slot] = particle_key<D>(x);
  }
}

// Strict order of two particles of the same cell and part with equal keys:
// position bits, then velocity bits, then slot (slot order only separates
// particles that are bitwise identical, so it never changes the output).
template <int D>
__device__ __forceinline__ bool tie_less(const DTmp& T, int j, int i) {
#pragma unroll
  for (int k = 0; k < D; ++k) {
    const unsigned long long a = (unsigned long long)__double_as_longlong(T.x[k][j]);
    const unsigned long long b = (unsigned long long)__double_as_longlong(T.x[k][i]);
    if (a != b) return a < b;
  }
#pragma unroll
  for (int k = 0; k < D; ++k) {
    const unsigned long long a = (unsigned long long)__double_as_longlong(T.v[k][j]);
    const unsigned long long b = (unsigned long long)__double_as_longlong(T.v[k][i]);
    if (a != b) return a < b;
  }
  return j < i;
}

// ---------------------------------------------------------------- settle
// check_particle_distribution (src/pfem2particlehandler.cpp:355-410) on the
// cell-binned particles in T. A workgroup owns `cpw` consecutive cells and
// counts particles per (cell, part) in LDS. Phase 0 writes each cell's final
// count. Phase 1 writes survivors (the max+1 smallest keys of each part) and
// one refill per empty part of a non-empty cell, in (part, key) order, at
// fstart[c] + the part's base offset + the rank within the part.
template <int D>
__global__ __launch_bounds__(kBlock) void k_settle(int phase, int apply_cap, int do_refill, DMesh m, int nc, int cpw,
                                                  int pmax, const int* __restrict__ tstart, DTmp T,
                                                  int* __restrict__ fcount, const int* __restrict__ fstart, DParts P,
                                                  DField Vf) {
  constexpr int NV = 1 << D;
  extern __shared__ int smem[];
  int* cnt = smem;               // [cpw][pmax]
  int* base = smem + cpw * pmax;  // [cpw][pmax]
  const int nchunks = (nc + cpw - 1) / cpw;
  for (int ch = blockIdx.x; ch < nchunks; ch += gridDim.x) {
    const int c0 = ch * cpw;
    const int c1 = min(nc, c0 + cpw);
    for (int i = threadIdx.x; i < cpw * pmax; i += blockDim.x) cnt[i] = 0;
    __syncthreads();
    const int s0 = tstart[c0], s1 = tstart[c1];
    for (int i = s0 + threadIdx.x; i < s1; i += blockDim.x) atomicAdd(&cnt[(T.cell[i] - c0) * pmax + T.part[i]], 1);
    __syncthreads();
    for (int lc = threadIdx.x; lc < c1 - c0; lc += blockDim.x) {
      const int c = c0 + lc;
      const Prm pr = m.prm[inst_of(m, c)];
      const int nci = tstart[c + 1] - tstart[c];
      int run = 0;
      for (int pp = 0; pp < pr.P; ++pp) {
        const int k = cnt[lc * pmax + pp];
        const int kept = apply_cap ? min(k, pr.maxpp + 1) : k;
        bool valid = true;
        int rem = pp;
#pragma unroll
        for (int d = D - 1; d >= 0; --d) {
          const int pk = rem % pr.Q[d];
          rem /= pr.Q[d];
          valid = valid && (pk < pr.q[d]);
        }
        const int refill = (do_refill && nci > 0 && k == 0 && valid) ? 1 : 0;
        base[lc * pmax + pp] = run;
        run += kept + refill;
      }
      if (phase == 0) fcount[c] = run;
    }
    __syncthreads();
    if (phase == 0) continue;
    for (int i = s0 + threadIdx.x; i < s1; i += blockDim.x) {
      const int c = T.cell[i];
      const int part = T.part[i];
      const unsigned long long key = T.key[i];
      const int cs = tstart[c], ce = tstart[c + 1];
      int r = 0;
      for (int j = cs; j < ce; ++j) {
        if (j == i || T.part[j] != part) continue;
        const unsigned long long kj = T.key[j];
        if (kj < key || (kj == key && tie_less<D>(T, j, i))) ++r;
      }
      const Prm pr = m.prm[inst_of(m, c)];
      if (apply_cap && r > pr.maxpp) continue;
      const int out = fstart[c] + base[(c - c0) * pmax + part] + r;
#pragma unroll
      for (int k = 0; k < D; ++k) {
        P.x[k][out] = T.x[k][i];
        P.v[k][out] = T.v[k][i];
      }
      P.cell[out] = c;
    }
    if (do_refill) {
      for (int t = threadIdx.x; t < (c1 - c0) * pmax; t += blockDim.x) {
        const int lc = t / pmax, pp = t - lc * pmax;
        const int c = c0 + lc;
        const Prm pr = m.prm[inst_of(m, c)];
        if (pp >= pr.P || cnt[t] != 0 || tstart[c + 1] == tstart[c]) continue;
        int pi[D];
        bool valid = true;
        int rem = pp;
#pragma unroll
        for (int d = D - 1; d >= 0; --d) {
          pi[d] = rem % pr.Q[d];
          rem /= pr.Q[d];
          valid = valid && (pi[d] < pr.q[d]);
        }
        if (!valid) continue;
        double xi[D], x[D], v[D];
#pragma unroll
        for (int k = 0; k < D; ++k) xi[k] = ((double)pi[k] + 1.0 / 2) * pr.h[k];
        int vid[NV], dof[NV];
        load_ids<D>(m.cv, c, vid);
        load_ids<D>(m.cd, c, dof);
        fwd<D>(m, vid, xi, x);
        gather<D>(Vf, dof, xi, v);
        const int out = fstart[c] + base[t];
#pragma unroll
        for (int k = 0; k < D; ++k) {
          P.x[k][out] = x[k];
          P.v[k][out] = v[k];
        }
        P.cell[out] = c;
      }
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------- projection
// Per (cell, vertex) sums S_k = sum_p N_i(xi_p) v_p,k and W = sum_p N_i(xi_p),
// sequential over the cell's particles in stored order.
// src/pfem2particlehandler.cpp:191-199
template <int D>
__global__ __launch_bounds__(kBlock) void k_contrib(DMesh m, int nc, const int* __restrict__ cs, DParts P,
                                                   double* __restrict__ contrib) {
  constexpr int NV = 1 << D;
  const long total = (long)nc * NV;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
    const int c = (int)(t / NV), i = (int)(t % NV);
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
      recover_xi<D>(m, pr, c, vid, x, xi);
      shape<D>(xi, N);
      double Ni = N[0];
#pragma unroll
      for (int q = 1; q < NV; ++q)
        if (q == i) Ni = N[q];
#pragma unroll
      for (int k = 0; k < D; ++k) S[k] = S[k] + Ni * P.v[k][p];
      W = W + Ni;
    }
    double* e_out = contrib + t * (D + 1);
#pragma unroll
    for (int k = 0; k < D; ++k) e_out[k] = S[k];
    e_out[D] = W;
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
    recover_xi<D>(m, pr, c, vid, x, xi);
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
    recover_xi<D>(m, pr, c, vid, x, xi);
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
    recover_xi<D>(m, pr, c, vid, x, xi);
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
