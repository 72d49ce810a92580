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
// interpolation's gather becomes a scatter into dL/dV (k_adj_move,
// k_adj_unmap). dL/dV is accumulated with FP64 atomics.
#pragma once

#include "pfem2gpu_kernels.h"

namespace pfem2gpu {

constexpr int kAdjMaxSub = 8;

__global__ void k_fill_double(double* __restrict__ a, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) a[i] = 0.0;
}

// nodal projection P = S / W and the weights W (same sums as k_node)
template <int D>
__global__ __launch_bounds__(kBlock) void k_node_w(int ndofs, const int* __restrict__ d2c_off,
                                                  const int* __restrict__ d2c, const double* __restrict__ contrib,
                                                  double* p0, double* p1, double* p2, double* __restrict__ W) {
  double* out[3] = {p0, p1, p2};
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < ndofs; j += gridDim.x * blockDim.x) {
    double S[D], w = 0.0;
#pragma unroll
    for (int k = 0; k < D; ++k) S[k] = 0.0;
    for (int e = d2c_off[j]; e < d2c_off[j + 1]; ++e) {
      const double* c = contrib + (size_t)d2c[e] * (D + 1);
#pragma unroll
      for (int k = 0; k < D; ++k) S[k] = S[k] + c[k];
      w = w + c[D];
    }
#pragma unroll
    for (int k = 0; k < D; ++k) out[k][j] = S[k] / w;
    W[j] = w;
  }
}

// L = 1/2 sum_{k,j} (P_k[j] - T_k[j])^2 (block partial sums, one atomic each)
template <int D>
__global__ __launch_bounds__(kBlock) void k_loss(int ndofs, DField P, DField T, double* __restrict__ out) {
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
  if (threadIdx.x == 0) atomicAdd(out, 0.5 * part[0]);
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
    const int c = Pt.cell[q];
    const Prm& pr = m.prm[inst_of(m, c)];
    double x[D], v[D], xi[D], N[NV], dN[NV][3], X0[D], RH[D];
#pragma unroll
    for (int k = 0; k < D; ++k) {
      x[k] = Pt.x[k][q];
      v[k] = Pt.v[k][q];
    }
    recover_xi_c<D>(m, pr, c, x, xi);
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

// Reverse of the redistribution: survivors hand their adjoints back to the
// particle they came from; refills (constant position, velocity interpolated
// from V) scatter N_i(xi_c) * dL/dv into dL/dV.
template <int D>
__global__ __launch_bounds__(kBlock) void k_adj_unmap(DMesh m, int pshift, int n_end, const int* __restrict__ src,
                                                     const int* __restrict__ end_cell, const double* xe0,
                                                     const double* xe1, const double* xe2, const double* ve0,
                                                     const double* ve1, const double* ve2, double* xs0, double* xs1,
                                                     double* xs2, double* vs0, double* vs1, double* vs2, double* g0,
                                                     double* g1, double* g2) {
  constexpr int NV = 1 << D;
  const double* xe[3] = {xe0, xe1, xe2};
  const double* ve[3] = {ve0, ve1, ve2};
  double* xs[3] = {xs0, xs1, xs2};
  double* vs[3] = {vs0, vs1, vs2};
  double* g[3] = {g0, g1, g2};
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n_end; q += gridDim.x * blockDim.x) {
    const int s = src[q];
    if (s >= 0) {
#pragma unroll
      for (int k = 0; k < D; ++k) {
        xs[k][s] = xe[k][q];
        vs[k][s] = ve[k][q];
      }
      continue;
    }
    const int c = end_cell[q];
    const int inst = inst_of(m, c);
    const Prm& pr = m.prm[inst];
    const int code = m.ptab[(inst << pshift) + (-1 - s)];
    double xi[D], N[NV];
#pragma unroll
    for (int d = 0; d < D; ++d) xi[d] = ((double)((code >> (8 * d)) & 0xff) + 1.0 / 2) * pr.h[d];
    shape<D>(xi, N);
    int dof[NV];
    load_ids<D>(m.cd, c, dof);
#pragma unroll
    for (int i = 0; i < NV; ++i)
#pragma unroll
      for (int k = 0; k < D; ++k) atomicAdd(&g[k][dof[i]], N[i] * ve[k][q]);
  }
}

// Reverse of the substeps of one move for every particle of the step's start
// state: the forward substeps are recomputed from the checkpointed start
// position (same arithmetic, so the same cells), then walked backwards:
//   dL/dV[dof_i]  += dt N_i(xi_k) a            (interpolation reversed: scatter)
//   a             += dt (du/dx)^T a, du_c/dx_d = sum_i dN_i/dxi_d V_c[i] / h_d
// `xt` holds dL/dx after the move on entry and before it on exit.
template <int D>
__global__ __launch_bounds__(kBlock) void k_adj_move(DMesh m, const double* __restrict__ Vc, int n_start,
                                                    const double* xs0, const double* xs1, const double* xs2,
                                                    const int* __restrict__ cs, double* xt0, double* xt1,
                                                    double* xt2, double* g0, double* g1, double* g2) {
  constexpr int NV = 1 << D;
  const double* xs[3] = {xs0, xs1, xs2};
  double* xt[3] = {xt0, xt1, xt2};
  double* g[3] = {g0, g1, g2};
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n_start; p += gridDim.x * blockDim.x) {
    const int c0 = cs[p];
    CellCache<D> cc;
    load_cell<D>(m, Vc, c0, cc, true);
    const Prm& pr = m.prm[inst_of(m, c0)];
    const int n_sub = pr.n_sub;
    const double dt = pr.dt_sub;
    double x[D], xi[D];
#pragma unroll
    for (int k = 0; k < D; ++k) x[k] = xs[k][p];
    inv_cached<D>(m, cc, x, xi);
    snap_cached<D>(m, pr, cc, x, xi);
    double path_xi[kAdjMaxSub][D];
    int path_c[kAdjMaxSub];
    bool alive = true;
    for (int s = 0; s < n_sub; ++s) {
#pragma unroll
      for (int k = 0; k < D; ++k) path_xi[s][k] = xi[k];
      path_c[s] = cc.c;
      double u[D];
      gather_cached<D>(Vc, cc, xi, u);
#pragma unroll
      for (int k = 0; k < D; ++k) {
        const double t = u[k] * dt;
        x[k] = x[k] + t;
      }
      double xn[D];
      if (inv_cached<D>(m, cc, x, xn) && inside<D>(xn)) {
#pragma unroll
        for (int k = 0; k < D; ++k) xi[k] = xn[k];
      } else {
        const int c2 = walk_cached<D>(m, cc, xi, x, u, xn);
        if (c2 < 0) {
          alive = false;
          break;
        }
        load_cell<D>(m, Vc, c2, cc, true);
#pragma unroll
        for (int k = 0; k < D; ++k) xi[k] = xn[k];
      }
    }
    double a[D];
#pragma unroll
    for (int k = 0; k < D; ++k) a[k] = alive ? xt[k][p] : 0.0;
    if (alive) {
      for (int s = n_sub - 1; s >= 0; --s) {
        const int c = path_c[s];
        double N[NV], dN[NV][3], X0[D], RH[D];
        shape_and_grad<D>(path_xi[s], N, dN);
        box_geom<D>(m, c, X0, RH);
        int dof[NV];
        load_ids<D>(m.cd, c, dof);
        const double* f = Vc + (size_t)c * NV * D;
        double gx[D];
#pragma unroll
        for (int d = 0; d < D; ++d) gx[d] = 0.0;
#pragma unroll
        for (int i = 0; i < NV; ++i) {
          double av = 0.0;
#pragma unroll
          for (int k = 0; k < D; ++k) {
            atomicAdd(&g[k][dof[i]], dt * N[i] * a[k]);
            av += a[k] * f[i * D + k];
          }
#pragma unroll
          for (int d = 0; d < D; ++d) gx[d] += dN[i][d] * av;
        }
#pragma unroll
        for (int d = 0; d < D; ++d) a[d] += dt * gx[d] * RH[d];
      }
    }
#pragma unroll
    for (int k = 0; k < D; ++k) xt[k][p] = a[k];
  }
}

}  // namespace pfem2gpu
