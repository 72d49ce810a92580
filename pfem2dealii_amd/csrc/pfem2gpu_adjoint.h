// Adjoint kernels of advect + project (no counterpart in the reference).
//
// Forward model differentiated (per step, with the cell of every particle at
// every substep, the survivor set and the refills frozen at the values the
// forward pass chose): n_sub explicit-Euler substeps x <- x + dt_sub * u(x),
// u(x) = sum_i N_i(xi(x)) V[dof_i], xi(x) = (x - X_0) / h on the particle's
// (box) cell; the redistribution keeps survivors unchanged and creates
// refills at sub-cell centres with v = sum_i N_i(xi_c) V[dof_i]; after the
// last step the nodal projection P_j = S_j / W_j, S_j = sum_p N_j v_p,
// W_j = sum_p N_j, and the loss L = 1/2 |P - target|^2.
// Reverse: the projection's scatter becomes a gather (k_adj_init), the
// interpolation's gather becomes a scatter into a per-(cell, vertex)
// accumulator (k_adj_move, k_adj_unmap; integer atomics on 128-bit fixed-point
// values after aggregation over each wavefront), gathered into dL/dV per DoF
// at the end (k_adj_gather). Integer sums do not depend on the order of the
// additions, so the gradient and the loss are the same bits on every run.
#pragma once

#include "pfem2gpu_kernels.h"

namespace pfem2gpu {

constexpr int kAdjMaxSub = 8;

// 128-bit two's-complement fixed-point accumulator of a double sum, as two
// 64-bit words (low, high) of value (hi * 2^64 + lo) * 2^-kFixBits. A term is
// rounded once (toward zero, at 2^-80) when it is converted; the additions
// are exact integer additions, so the result does not depend on their order.
// Range: |sum| < 2^(127 - kFixBits) = 2^47.
constexpr int kFixBits = 80;

// Returns false (and converts 0) for a non-finite term or one of magnitude
// 2^46 or more, which the accumulator cannot hold.
__device__ __forceinline__ bool fix_from(double x, unsigned long long& lo, unsigned long long& hi) {
  double a = fabs(x) * 0x1p80;                     // exact (a power-of-two scale)
  const bool ok = a < 0x1p126;                     // false for NaN too
  if (!ok) a = 0.0;
  const double hd = floor(a * 0x1p-64);            // exact
  const double r = a - hd * 0x1p64;                // exact, in [0, 2^64)
  lo = (unsigned long long)r;
  hi = (unsigned long long)hd;
  if (x < 0.0) {  // two's complement negation
    lo = ~lo + 1ULL;
    hi = ~hi + (lo == 0ULL ? 1ULL : 0ULL);
  }
  return ok;
}

__device__ __forceinline__ double fix_to(unsigned long long lo, unsigned long long hi) {
  return ((double)(long long)hi * 0x1p64 + (double)lo) * 0x1p-80;
}

// acc[0..1] += (lo, hi): a wrap of the low word carries into the high word;
// the number of wraps is that of the total sum, whatever the order.
__device__ __forceinline__ void fix_add_words(unsigned long long* acc, unsigned long long lo, unsigned long long hi) {
  const unsigned long long old = atomicAdd(&acc[0], lo);
  const unsigned long long carry = (old + lo < old) ? 1ULL : 0ULL;
  if (hi + carry != 0ULL) atomicAdd(&acc[1], hi + carry);
}

__device__ __forceinline__ bool fix_add(unsigned long long* acc, double x) {
  unsigned long long lo, hi;
  const bool ok = fix_from(x, lo, hi);
  fix_add_words(acc, lo, hi);
  return ok;
}

// dst[i] = the double value of the fixed-point accumulator src[2i..2i+1]
__global__ void k_fix_to_double(long n, const unsigned long long* __restrict__ src, double* __restrict__ dst) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    dst[i] = fix_to(src[2 * i], src[2 * i + 1]);
}

__global__ void k_fill_u64(unsigned long long* __restrict__ a, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) a[i] = 0ULL;
}

__global__ void k_fill_double(double* __restrict__ a, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) a[i] = 0.0;
}

// nodal projection P = S / W and the weights W (same sums as k_node)
template <int D>
__global__ __launch_bounds__(kBlock) void k_node_w(int ndofs, const int* __restrict__ d2c_off,
                                                  const double* __restrict__ contrib,
                                                  double* p0, double* p1, double* p2, double* __restrict__ W) {
  double* out[3] = {p0, p1, p2};
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < ndofs; j += gridDim.x * blockDim.x) {
    double S[D], w = 0.0;
#pragma unroll
    for (int k = 0; k < D; ++k) S[k] = 0.0;
    for (int e = d2c_off[j]; e < d2c_off[j + 1]; ++e) {
      const double* c = contrib + (size_t)e * (D + 1);
#pragma unroll
      for (int k = 0; k < D; ++k) S[k] = S[k] + c[k];
      w = w + c[D];
    }
#pragma unroll
    for (int k = 0; k < D; ++k) out[k][j] = S[k] / w;
    W[j] = w;
  }
}

// L = 1/2 sum_{k,j} (P_k[j] - T_k[j])^2: block partial sums in a fixed tree,
// added to the fixed-point accumulator out[0..1]
template <int D>
__global__ __launch_bounds__(kBlock) void k_loss(int ndofs, DField P, DField T, unsigned long long* __restrict__ out,
                                                int* __restrict__ err) {
  __shared__ double part[kBlock];
  double s = 0.0;
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < ndofs; j += gridDim.x * blockDim.x)
#pragma unroll
    for (int k = 0; k < D; ++k) {
      const double r = P.c[k][j] - T.c[k][j];
      s = s + r * r;
    }
  part[threadIdx.x] = s;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0 && !fix_add(out, 0.5 * part[0])) atomicOr(err, kErrRange);
}

// Post-step state of a replayed move from the maps the forward recorded:
// slot q takes the moved position and the velocity of its source particle
// (k_move_* in replay mode stored them by source index in M), or, for a
// refill (src = -1 - part), the particle the redistribution built: the part's
// sub-cell centre and the advecting field interpolated there (the same
// expressions as the refill branch of k_settle). Writes the count.
template <int D>
__global__ __launch_bounds__(kBlock) void k_adj_replay(DMesh m, const double* __restrict__ Vc, int pshift, int n_end,
                                                      const int* __restrict__ src, const int* __restrict__ end_cell,
                                                      DTmp M, DParts P, int* __restrict__ count) {
  constexpr int NV = 1 << D;
  if (blockIdx.x == 0 && threadIdx.x == 0) *count = n_end;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n_end; q += gridDim.x * blockDim.x) {
    const int s = src[q], cw = end_cell[q], c = cell_of(cw);
    double x[D], v[D];
    if (s >= 0) {
#pragma unroll
      for (int k = 0; k < D; ++k) {
        x[k] = M.x[k][s];
        v[k] = M.v[k][s];
      }
    } else {
      const int inst = inst_of(m, c);
      const Prm& pr = m.prm[inst];
      const int code = m.ptab[(inst << pshift) + (-1 - s)];
      double xi[D], N[NV];
#pragma unroll
      for (int k = 0; k < D; ++k) xi[k] = ((double)((code >> (8 * k)) & 0xff) + 1.0 / 2) * pr.h[k];
      const double* g = m.geom + (size_t)c * 3 * D;
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
      shape<D>(xi, N);
      const double* fv = Vc + (size_t)c * NV * D;
#pragma unroll
      for (int k = 0; k < D; ++k) v[k] = 0.0;
#pragma unroll
      for (int i = 0; i < NV; ++i)
#pragma unroll
        for (int k = 0; k < D; ++k) v[k] = v[k] + N[i] * fv[i * D + k];
    }
#pragma unroll
    for (int k = 0; k < D; ++k) {
      P.x[k][q] = x[k];
      P.v[k][q] = v[k];
    }
    P.cell[q] = cw;  // refills carry the placed flag
  }
}

// Box-cell geometry of cell c from the table (the adjoint supports box cells).
template <int D>
__device__ __forceinline__ void box_geom(const DMesh& m, int c, double* X0, double* RH) {
  const double* g = m.geom + (size_t)c * 3 * D;
#pragma unroll
  for (int k = 0; k < D; ++k) {
    X0[k] = g[k];
    RH[k] = g[2 * D + k];
  }
}

// Advecting field at vertex i, component k of a cached cell (registers in
// 2D, the cell-major table otherwise).
template <int D>
__device__ __forceinline__ double cell_field(const double* __restrict__ Vc, const CellCache<D>& cc, int i, int k) {
  if constexpr (D == 2)
    return cc.Vn[i][k];
  else
    return Vc[(size_t)cc.c * (1 << D) * D + i * D + k];
}

// Shape values and their reference-coordinate gradients.
template <int D>
__device__ __forceinline__ void shape_and_grad(const double* xi, double* N, double (*dN)[3]) {
  shape<D>(xi, N);
  shape_grad<D>(xi, dN);
}

// Adjoint of the final projection: for each particle q of the final state,
// dL/dv_q = sum_j N_j r_j / W_j and dL/dx_q = (1/h) sum_j grad N_j
// sum_k (r_jk / W_j)(v_qk - P_jk), with r = P - target.
template <int D>
__global__ __launch_bounds__(kBlock) void k_adj_init(DMesh m, const int* __restrict__ n_ptr, DParts Pt, DField Pn,
                                                    const double* __restrict__ W, DField Tg, double* xt0,
                                                    double* xt1, double* xt2, double* vt0, double* vt1,
                                                    double* vt2) {
  constexpr int NV = 1 << D;
  double* xt[3] = {xt0, xt1, xt2};
  double* vt[3] = {vt0, vt1, vt2};
  const int n = *n_ptr;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
    const int w = Pt.cell[q];
    const int c = cell_of(w);
    const Prm& pr = m.prm[inst_of(m, c)];
    double x[D], v[D], xi[D], N[NV], dN[NV][3], X0[D], RH[D];
#pragma unroll
    for (int k = 0; k < D; ++k) {
      x[k] = Pt.x[k][q];
      v[k] = Pt.v[k][q];
    }
    recover_xi_c<D>(m, pr, c, placed_of(w), x, xi);
    shape_and_grad<D>(xi, N, dN);
    box_geom<D>(m, c, X0, RH);
    int dof[NV];
    load_ids<D>(m.cd, c, dof);
    double av[D], ax[D];
#pragma unroll
    for (int k = 0; k < D; ++k) {
      av[k] = 0.0;
      ax[k] = 0.0;
    }
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      const int j = dof[i];
      const double w = W[j];
      double coef = 0.0;
#pragma unroll
      for (int k = 0; k < D; ++k) {
        const double pj = Pn.c[k][j];
        const double a = (pj - Tg.c[k][j]) / w;
        av[k] += N[i] * a;
        coef += a * (v[k] - pj);
      }
#pragma unroll
      for (int d = 0; d < D; ++d) ax[d] += dN[i][d] * coef;
    }
#pragma unroll
    for (int k = 0; k < D; ++k) {
      vt[k][q] = av[k];
      xt[k][q] = ax[k] * RH[k];
    }
  }
}

// dL/dV[k][j] = the sum over the (cell, vertex) entries of DoF j of the
// per-(cell, vertex) fixed-point accumulators (an exact integer sum,
// converted once).
template <int D>
__global__ __launch_bounds__(kBlock) void k_adj_gather(int ndofs, const int* __restrict__ d2c_off,
                                                      const int* __restrict__ d2c,
                                                      const unsigned long long* __restrict__ gcv, double* g0,
                                                      double* g1, double* g2) {
  double* g[3] = {g0, g1, g2};
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < ndofs; j += gridDim.x * blockDim.x) {
    unsigned long long lo[D], hi[D];
#pragma unroll
    for (int k = 0; k < D; ++k) lo[k] = hi[k] = 0ULL;
    for (int e = d2c_off[j]; e < d2c_off[j + 1]; ++e)
#pragma unroll
      for (int k = 0; k < D; ++k) {
        const unsigned long long* a = gcv + 2 * ((size_t)d2c[e] * D + k);
        const unsigned long long l = lo[k] + a[0];
        hi[k] = hi[k] + a[1] + (l < lo[k] ? 1ULL : 0ULL);
        lo[k] = l;
      }
#pragma unroll
    for (int k = 0; k < D; ++k) g[k][j] = fix_to(lo[k], hi[k]);
  }
}

// Scatter of per-lane contributions add[i][k] into the per-(cell, vertex)
// fixed-point accumulator gcv (two words per value) of the lane's cell
// (key < 0: nothing), aggregated over the lanes of the wavefront that share
// the cell: up to kAdjGroups distinct cells are summed in LDS (fixed point
// too) and flushed with one global addition per (cell, vertex, component);
// lanes beyond that add directly. Wave-uniform.
constexpr int kAdjGroups = 16;

// Orders this wavefront's LDS accesses around the point (waits for its LDS
// operations only).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0); vmcnt and expcnt left at their maxima
  __builtin_amdgcn_wave_barrier();
}

template <int D>
__device__ __forceinline__ void adj_scatter(int key, const double (*add)[3], unsigned long long* __restrict__ gcv,
                                            int* s_gcell, unsigned long long* s_acc, int* __restrict__ err) {
  bool ok = true;
  constexpr int NV = 1 << D;
  constexpr int W = NV * D;
  const int lane = threadIdx.x & 63;
  int gid = -1, ng = 0;
  unsigned long long pending = __ballot(key >= 0);
  while (pending && ng < kAdjGroups) {
    const int leader = __ffsll((unsigned long long)pending) - 1;
    const int lk = __shfl(key, leader, 64);
    const unsigned long long grp = __ballot(key == lk) & pending;
    if ((grp >> lane) & 1ULL) gid = ng;
    if (lane == leader) s_gcell[ng] = lk;
    pending &= ~grp;
    ++ng;
  }
  for (int l = lane; l < 2 * ng * W; l += 64) s_acc[l] = 0ULL;
  wave_lds_sync();
  if (gid >= 0) {
#pragma unroll
    for (int i = 0; i < NV; ++i)
#pragma unroll
      for (int k = 0; k < D; ++k) ok = fix_add(&s_acc[2 * (gid * W + i * D + k)], add[i][k]) && ok;
  } else if (key >= 0) {
#pragma unroll
    for (int i = 0; i < NV; ++i)
#pragma unroll
      for (int k = 0; k < D; ++k) ok = fix_add(&gcv[2 * (((size_t)key * NV + i) * D + k)], add[i][k]) && ok;
  }
  if (!ok) atomicOr(err, kErrRange);
  wave_lds_sync();
  for (int l = lane; l < ng * W; l += 64) {
    const int gi = l / W;
    fix_add_words(&gcv[2 * ((size_t)s_gcell[gi] * W + (l - gi * W))], s_acc[2 * l], s_acc[2 * l + 1]);
  }
  wave_lds_sync();
}

// Reverse of the redistribution: survivors hand their adjoints back to the
// particle they came from; refills (constant position, velocity interpolated
// from V) add N_i(xi_c) * dL/dv to the per-(cell, vertex) accumulator gcv
// through the aggregated scatter (the refills of a cell are neighbours in the
// end state, so the lanes of one cell add once per entry). Every lane of a
// wavefront reaches the scatter.
template <int D>
__global__ __launch_bounds__(kBlock) void k_adj_unmap(DMesh m, int pshift, int n_end, const int* __restrict__ src,
                                                     const int* __restrict__ end_cell, const double* xe0,
                                                     const double* xe1, const double* xe2, const double* ve0,
                                                     const double* ve1, const double* ve2, double* xs0, double* xs1,
                                                     double* xs2, double* vs0, double* vs1, double* vs2,
                                                     unsigned long long* __restrict__ gcv, int* __restrict__ err) {
  constexpr int NV = 1 << D;
  __shared__ int s_gcell[kBlock / 64][kAdjGroups];
  __shared__ unsigned long long s_acc[kBlock / 64][2 * kAdjGroups * NV * D];
  const int wave = threadIdx.x >> 6;
  const double* xe[3] = {xe0, xe1, xe2};
  const double* ve[3] = {ve0, ve1, ve2};
  double* xs[3] = {xs0, xs1, xs2};
  double* vs[3] = {vs0, vs1, vs2};
  for (long base = (long)blockIdx.x * blockDim.x; base < n_end; base += (long)gridDim.x * blockDim.x) {
    const int q = (int)(base + threadIdx.x);
    const int s = q < n_end ? src[q] : 0;
    int key = -1;
    double add[NV][3];
#pragma unroll
    for (int i = 0; i < NV; ++i)
#pragma unroll
      for (int k = 0; k < 3; ++k) add[i][k] = 0.0;
    if (q < n_end && s >= 0) {
#pragma unroll
      for (int k = 0; k < D; ++k) {
        xs[k][s] = xe[k][q];
        vs[k][s] = ve[k][q];
      }
    } else if (q < n_end) {
      const int c = cell_of(end_cell[q]);
      const int inst = inst_of(m, c);
      const Prm& pr = m.prm[inst];
      const int code = m.ptab[(inst << pshift) + (-1 - s)];
      double xi[D], N[NV];
#pragma unroll
      for (int d = 0; d < D; ++d) xi[d] = ((double)((code >> (8 * d)) & 0xff) + 1.0 / 2) * pr.h[d];
      shape<D>(xi, N);
#pragma unroll
      for (int i = 0; i < NV; ++i)
#pragma unroll
        for (int k = 0; k < D; ++k) add[i][k] = N[i] * ve[k][q];
      key = c;
    }
    adj_scatter<D>(key, add, gcv, s_gcell[wave], s_acc[wave], err);
  }
}

// Reverse of the substeps of one move for every particle of the step's start
// state, from the path the recomputed move recorded (cell and reference
// position at the start of each substep), walked backwards:
//   dL/dV[dof_i]  += dt N_i(xi_s) a            (interpolation reversed: scatter
//                                                into the per-(cell, vertex)
//                                                accumulator gcv)
//   a             += dt (du/dx)^T a, du_c/dx_d = sum_i dN_i/dxi_d V_c[i] / h_d
// `xt` holds dL/dx after the move on entry and before it on exit. A particle
// the move deleted has no adjoint (it is no survivor), so lanes with a zero
// adjoint skip their substeps (a zero adjoint contributes nothing). The
// aggregated scatters (at a cell change and at the end) are reached by every
// lane of the wavefront. MS bounds the substep count (max_sub <= MS): the
// path is held in registers for MS substeps.
template <int D, int MS>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3))) void k_adj_move(DMesh m, const double* __restrict__ Vc, int n_start, int max_sub,
                                                    DPath path, double* xt0, double* xt1, double* xt2,
                                                    unsigned long long* __restrict__ gcv, int* __restrict__ err) {
  constexpr int NV = 1 << D;
  __shared__ int s_gcell[kBlock / 64][kAdjGroups];
  __shared__ unsigned long long s_acc[kBlock / 64][2 * kAdjGroups * NV * D];
  const int wave = threadIdx.x >> 6;
  double* xt[3] = {xt0, xt1, xt2};
  for (long base = (long)blockIdx.x * blockDim.x; base < n_start; base += (long)gridDim.x * blockDim.x) {
    const int p = (int)(base + threadIdx.x);
    double a[D];
    bool live = false;
#pragma unroll
    for (int k = 0; k < D; ++k) {
      a[k] = p < n_start ? xt[k][p] : 0.0;
      live = live || a[k] != 0.0;
    }
    // the whole path in one round of loads, issued together with the
    // adjoint's (a lane that turns out dead ignores what it loaded; a lane
    // past the end or a substep past max_sub reads entry (0, 0), in bounds, so
    // that every load is issued unconditionally)
    int pc[MS];
    double pxi[MS][D];
#pragma unroll
    for (int s = 0; s < MS; ++s) {
      const bool in = p < n_start && s < max_sub;
      const long o = in ? s * path.stride + p : 0;
      pc[s] = path.c[o];
#pragma unroll
      for (int k = 0; k < D; ++k) pxi[s][k] = path.xi[k][o];
      if (!in) pc[s] = -1;
    }
    // substep count and length: one instance's (a uniform load) or the
    // instance of the particle's first cell
    int n_sub = m.prm[0].n_sub;
    double dt = m.prm[0].dt_sub;
    if (m.cell_inst && live) {
      const Prm& pr = m.prm[m.cell_inst[pc[0]]];
      n_sub = pr.n_sub;
      dt = pr.dt_sub;
    }
    CellCache<D> cc;
    cc.c = -1;
    // contributions summed per lane over consecutive substeps in one cell
    // (acell), handed to the aggregated scatter when the cell changes and
    // once more at the end
    double acc[NV][3];
    int acell = -1;
#pragma unroll
    for (int i = 0; i < NV; ++i)
#pragma unroll
      for (int k = 0; k < 3; ++k) acc[i][k] = 0.0;
#pragma unroll
    for (int s = MS - 1; s >= 0; --s) {
      if (s >= max_sub) continue;
      const bool act = live && s < n_sub;
      // a lane whose cell changes here hands its sums for the previous cell
      // to the aggregated scatter (wave-uniform: every lane takes part), so
      // lanes of one cell flushing together add once per (cell, vertex,
      // component) instead of contending on the same addresses
      const bool fl = act && pc[s] != acell && acell >= 0;
      if (__ballot(fl)) {
        adj_scatter<D>(fl ? acell : -1, acc, gcv, s_gcell[wave], s_acc[wave], err);
        if (fl) {
#pragma unroll
          for (int i = 0; i < NV; ++i)
#pragma unroll
            for (int k = 0; k < D; ++k) acc[i][k] = 0.0;
        }
      }
      if (act) {
        const int c = pc[s];
        if (c != cc.c) load_cell<D>(m, Vc, c, cc, true);  // particles mostly keep their cell
        acell = c;
        double N[NV], dN[NV][3];
        shape_and_grad<D>(pxi[s], N, dN);
        double gx[D];
#pragma unroll
        for (int d = 0; d < D; ++d) gx[d] = 0.0;
#pragma unroll
        for (int i = 0; i < NV; ++i) {
          double av = 0.0;
#pragma unroll
          for (int k = 0; k < D; ++k) {
            acc[i][k] += dt * N[i] * a[k];
            av += a[k] * cell_field<D>(Vc, cc, i, k);
          }
#pragma unroll
          for (int d = 0; d < D; ++d) gx[d] += dN[i][d] * av;
        }
#pragma unroll
        for (int d = 0; d < D; ++d) a[d] += dt * gx[d] * cc.RH[d];
      }
    }
    adj_scatter<D>(acell, acc, gcv, s_gcell[wave], s_acc[wave], err);
    if (p < n_start) {
#pragma unroll
      for (int k = 0; k < D; ++k) xt[k][p] = a[k];
    }
  }
}

// Copies the first *n_ptr entries of up to six double arrays and one int
// array (the count lives on the device, so no host round trip is needed).
__global__ void k_copy_state(const int* __restrict__ n_ptr, const double* s0, const double* s1, const double* s2,
                             const double* s3, const double* s4, const double* s5, const int* sc, double* d0,
                             double* d1, double* d2, double* d3, double* d4, double* d5, int* dc) {
  const long n = *n_ptr;
  const double* src[6] = {s0, s1, s2, s3, s4, s5};
  double* dst[6] = {d0, d1, d2, d3, d4, d5};
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
#pragma unroll
    for (int a = 0; a < 6; ++a)
      if (src[a]) dst[a][i] = src[a][i];
    if (sc) dc[i] = sc[i];
  }
}

// Zeroes the velocity and position adjoints of a step's start state.
__global__ void k_fill_adj(long n, double* a0, double* a1, double* a2, double* b0, double* b1, double* b2) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    a0[i] = 0.0;
    b0[i] = 0.0;
    if (a1) {
      a1[i] = 0.0;
      b1[i] = 0.0;
    }
    if (a2) {
      a2[i] = 0.0;
      b2[i] = 0.0;
    }
  }
}

}  // namespace pfem2gpu
