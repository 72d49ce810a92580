// Finite-element assembly of the pressure equation and of the velocity
// correction (src/pfem2fem.cpp:392-458 and :460-517) on the GPU: Q1 elements,
// the 2-point Gauss rule per direction (QUAD_POINTS_PER_DIRECTION,
// src/pfem2fem.h:24), the Q1 mapping, and the reference's arithmetic inside a
// cell (quadrature point, then test function i, then trial function j).
//
// k_fem_cells: a workgroup takes kFemBlock / 2^dim cells, one thread per
// (cell, r). Each thread first stages vertex r's coordinates and DoF r's field
// values in the cell's LDS record, then computes the geometry at quadrature
// point q = r (Jacobian, JxW and the real gradients of all 2^dim shape
// functions) into the record, then row i = r of the cell matrix and entry i of
// the cell right-hand side from the record. Each cell-matrix entry (c, i, j)
// is stored at its place apos[(c 2^d + i) 2^d + j] in contribution order (by
// global entry, then ascending cell), each right-hand side entry (c, i) at
// bpos[c 2^d + i] in the order of the DoF -> (cell, vertex) lists (by DoF,
// then ascending cell; one plane of nc 2^d entries per component).
//
// k_fem_gather_matrix: one thread per entry of the global CSR matrix sums its
// contiguous run of contributions, and k_fem_gather_rhs one thread per DoF
// its contiguous run of cell entries: the sums of distribute_local_to_global
// over the cells in ascending order, without atomics, so results are the
// same bits on every run.
// Constraints are not applied: the result is the unconstrained global system.
#pragma once

#include "pfem2gpu_kernels.h"

namespace pfem2gpu {

constexpr int kFemBlock = 128;

// Points of QGauss<1>(2) on [0, 1]: (1 -+ 1/sqrt(3)) / 2, weight 1/2 each.
__device__ __forceinline__ double gauss_pt(int b) {
  return b ? 0.78867513459481288225 : 0.21132486540518711775;
}

// Layout of a cell's LDS record, in doubles. The size is odd so that the
// records of the cells of one wavefront start in different LDS banks.
template <int D>
struct FemRec {
  static constexpr int NV = 1 << D;
  static constexpr int NQ = 1 << D;
  static constexpr int G = 0;                 // [NQ][NV][D] real shape gradients at the points
  static constexpr int J = G + NQ * NV * D;   // [NQ] JxW
  static constexpr int X = J + NQ;            // [NV][D] vertex coordinates
  static constexpr int V = X + NV * D;        // [NV][D] first field at the cell's DoFs
  static constexpr int P = V + NV * D;        // [NV] scalar field at the DoFs
  static constexpr int U = P + NV;            // [NV][D] second vector field at the DoFs
  static constexpr int size = (U + NV * D) | 1;
};

// Jacobian J[k][l] = sum_i X_i[k] dN_i/dxi_l of the Q1 map (sequential in i),
// its determinant and its inverse (adjugate entries / det).
template <int D>
struct Q1Jac {
  double det;
  double inv[D][D];  // inv[l][k] = (J^-1)[l][k]
};

template <int D>
__device__ __forceinline__ Q1Jac<D> q1_jacobian(const double* Xr, const double (*dN)[3]) {
  constexpr int NV = 1 << D;
  double J[D][D];
#pragma unroll
  for (int k = 0; k < D; ++k)
#pragma unroll
    for (int l = 0; l < D; ++l) J[k][l] = 0.0;
#pragma unroll
  for (int i = 0; i < NV; ++i)
#pragma unroll
    for (int k = 0; k < D; ++k)
#pragma unroll
      for (int l = 0; l < D; ++l) J[k][l] = J[k][l] + Xr[i * D + k] * dN[i][l];
  Q1Jac<D> r;
  if constexpr (D == 2) {
    r.det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
    r.inv[0][0] = J[1][1] / r.det;
    r.inv[0][1] = -J[0][1] / r.det;
    r.inv[1][0] = -J[1][0] / r.det;
    r.inv[1][1] = J[0][0] / r.det;
  } else {
    const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
    const double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
    const double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
    r.det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
    r.inv[0][0] = c00 / r.det;
    r.inv[1][0] = c01 / r.det;
    r.inv[2][0] = c02 / r.det;
    r.inv[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / r.det;
    r.inv[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / r.det;
    r.inv[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / r.det;
    r.inv[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / r.det;
    r.inv[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / r.det;
    r.inv[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / r.det;
  }
  return r;
}

// Real-space gradient of a shape function: g_k = sum_l dN/dxi_l (J^-1)[l][k].
template <int D>
__device__ __forceinline__ void q1_grad(const Q1Jac<D>& jac, const double* dNi, double* g) {
#pragma unroll
  for (int k = 0; k < D; ++k) {
    double s = dNi[0] * jac.inv[0][k];
#pragma unroll
    for (int l = 1; l < D; ++l) s = s + dNi[l] * jac.inv[l][k];
    g[k] = s;
  }
}

template <int D>
__device__ __forceinline__ double dot_d(const double* a, const double* b) {
  double s = a[0] * b[0];
#pragma unroll
  for (int k = 1; k < D; ++k) s = s + a[k] * b[k];
  return s;
}

// Reference point of point q of face f = 2 k + side (xi_k = side; the other
// coordinates Gauss points, the lower direction in the lower bit of q).
template <int D>
__device__ __forceinline__ void face_point(int f, int q, double* xi) {
  const int kf = f >> 1;
  int b = 0;
#pragma unroll
  for (int k = 0; k < D; ++k) {
    if (k == kf) {
      xi[k] = (double)(f & 1);
    } else {
      xi[k] = gauss_pt((q >> b) & 1);
      ++b;
    }
  }
}

// Geometry at quadrature point q of the cell whose vertex coordinates the
// record holds: JxW into rec[J + q], the real gradients of all shape
// functions into rec[G + (q 2^d + j) d + k].
template <int D>
__device__ __forceinline__ void fem_geometry(double* rec, int q) {
  using R = FemRec<D>;
  constexpr int NV = 1 << D;
  double xi[D], w = 1.0;
#pragma unroll
  for (int k = 0; k < D; ++k) {
    xi[k] = gauss_pt((q >> k) & 1);
    w = w * 0.5;
  }
  double dN[NV][3];
  shape_grad<D>(xi, dN);
  const Q1Jac<D> jac = q1_jacobian<D>(rec + R::X, dN);
  rec[R::J + q] = jac.det * w;
#pragma unroll
  for (int j = 0; j < NV; ++j) {
    double g[D];
    q1_grad<D>(jac, dN[j], g);
#pragma unroll
    for (int k = 0; k < D; ++k) rec[R::G + (q * NV + j) * D + k] = g[k];
  }
}

// Face point f, q of a cell: shape values N, real gradients g (rows j),
// unit outward normal n and JxW (Nanson), from the vertex coordinates Xr.
template <int D>
__device__ __forceinline__ double fem_face_geometry(const double* Xr, int f, int q, double* N, double (*g)[3],
                                                    double* nv) {
  constexpr int NV = 1 << D;
  double xi[D], w = 1.0;
  face_point<D>(f, q, xi);
#pragma unroll
  for (int k = 0; k < D - 1; ++k) w = w * 0.5;
  double dN[NV][3];
  shape<D>(xi, N);
  shape_grad<D>(xi, dN);
  const Q1Jac<D> jac = q1_jacobian<D>(Xr, dN);
  const int kf = f >> 1;
  const double sgn = (f & 1) ? 1.0 : -1.0;
  double n2 = 0.0;
#pragma unroll
  for (int k = 0; k < D; ++k) {
    double a = jac.inv[0][k];
#pragma unroll
    for (int l = 1; l < D; ++l)
      if (l == kf) a = jac.inv[l][k];
    nv[k] = sgn * a;
    n2 = n2 + nv[k] * nv[k];
  }
  const double nn = sqrt(n2);
#pragma unroll
  for (int k = 0; k < D; ++k) nv[k] = nv[k] / nn;
#pragma unroll
  for (int j = 0; j < NV; ++j) {
    double gj[D];
    q1_grad<D>(jac, dN[j], gj);
#pragma unroll
    for (int k = 0; k < D; ++k) g[j][k] = gj[k];
  }
  return jac.det * nn * w;
}

// Boundary term of assemble_pressure_equation (src/pfem2fem.cpp:436-449),
// after the cell term of the same right-hand side entries: one thread per
// cell with faces flagged in faces[c]; face by face in increasing number and
// point by point, the geometry once, then for every row i
//   b_i -= coeff N_i (V_q . n) JxW,  V_q = sum_l N_l V(l),
// n = J^-T n_ref / |J^-T n_ref| and JxW = det J |J^-T n_ref| w (Nanson).
template <int D>
__global__ __launch_bounds__(kBlock) void k_fem_faces(DMesh m, int nc, DField Vp, double coeff,
                                                     const unsigned char* __restrict__ faces,
                                                     const int* __restrict__ bpos, double* __restrict__ bcell) {
  constexpr int NV = 1 << D;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < nc; c += gridDim.x * blockDim.x) {
    const int fm = faces[c];
    if (!fm) continue;
    int vid[NV], dof[NV];
    load_ids<D>(m.cv, c, vid);
    load_ids<D>(m.cd, c, dof);
    double Xr[NV * D], Vr[NV * D], b[NV];
#pragma unroll
    for (int l = 0; l < NV; ++l) {
#pragma unroll
      for (int k = 0; k < D; ++k) {
        Xr[l * D + k] = m.X[k][vid[l]];
        Vr[l * D + k] = Vp.c[k][dof[l]];
      }
      b[l] = bcell[bpos[(size_t)c * NV + l]];
    }
    for (int f = 0; f < 2 * D; ++f) {
      if (!((fm >> f) & 1)) continue;
      for (int q = 0; q < (1 << (D - 1)); ++q) {
        double N[NV], g[NV][3], nv[D];
        const double JxW = fem_face_geometry<D>(Xr, f, q, N, g, nv);
        double Vq[D];
#pragma unroll
        for (int k = 0; k < D; ++k) Vq[k] = 0.0;
#pragma unroll
        for (int l = 0; l < NV; ++l)
#pragma unroll
          for (int k = 0; k < D; ++k) Vq[k] = Vq[k] + N[l] * Vr[l * D + k];
#pragma unroll
        for (int i = 0; i < NV; ++i) {
          double tv[D];
#pragma unroll
          for (int k = 0; k < D; ++k) tv[k] = coeff * N[i] * Vq[k];
          b[i] = b[i] - dot_d<D>(tv, nv) * JxW;
        }
      }
    }
#pragma unroll
    for (int i = 0; i < NV; ++i) bcell[bpos[(size_t)c * NV + i]] = b[i];
  }
}

// PRESSURE: assemble_pressure_equation (src/pfem2fem.cpp:392-458, SCHEMEB
// off), c0 = coeff = rho / dt:
//   A_ij = sum_q grad N_i . grad N_j JxW
//   b_i  = sum_q sum_j coeff (sum_k V*_k(j) d_k N_i) N_j JxW
// (the boundary term follows in k_fem_faces)
// otherwise: assemble_velocity_correction (src/pfem2fem.cpp:460-517, SCHEMEB
// off), c0 = rho, c1 = dt, the matrix common to the components:
//   A_ij  = sum_q rho N_i N_j JxW
//   b_ik  = sum_q sum_j [rho N_i N_j JxW V*_k(j) - dt N_i JxW d_k N_j P(j)]
template <int D, bool PRESSURE>
__global__ __launch_bounds__(kFemBlock) void k_fem_cells(DMesh m, int nc, DField Vp, const double* __restrict__ P,
                                                        double c0, double c1, const int* __restrict__ apos,
                                                        const int* __restrict__ bpos, double* __restrict__ Acell,
                                                        double* __restrict__ bcell) {
  using R = FemRec<D>;
  constexpr int NV = 1 << D;
  constexpr int NQ = 1 << D;
  constexpr int CPB = kFemBlock / NV;
  __shared__ double s[CPB * R::size];
  const int lc = threadIdx.x / NV, r = threadIdx.x % NV;
  double* rec = s + lc * R::size;
  for (long base = (long)blockIdx.x * CPB; base < nc; base += (long)gridDim.x * CPB) {
    const long c = base + lc;
    const bool on = c < nc;
    if (on) {
      const int vid = m.cv[c * NV + r], dof = m.cd[c * NV + r];
#pragma unroll
      for (int k = 0; k < D; ++k) {
        rec[R::X + r * D + k] = m.X[k][vid];
        rec[R::V + r * D + k] = Vp.c[k][dof];
      }
      if constexpr (!PRESSURE) rec[R::P + r] = P[dof];
    }
    __syncthreads();
    if (on) fem_geometry<D>(rec, r);  // geometry at quadrature point q = r
    __syncthreads();
    if (on) {  // row i = r
      const int i = r;
      double A[NV];
#pragma unroll
      for (int j = 0; j < NV; ++j) A[j] = 0.0;
      if constexpr (PRESSURE) {
        const double coeff = c0;
        double b = 0.0;
#pragma unroll 1
        for (int q = 0; q < NQ; ++q) {
          double xi[D];
#pragma unroll
          for (int k = 0; k < D; ++k) xi[k] = gauss_pt((q >> k) & 1);
          double N[NV];
          shape<D>(xi, N);
          const double JxW = rec[R::J + q];
          double gi[D];
#pragma unroll
          for (int k = 0; k < D; ++k) gi[k] = rec[R::G + (q * NV + i) * D + k];
#pragma unroll
          for (int j = 0; j < NV; ++j) {
            double gj[D];
#pragma unroll
            for (int k = 0; k < D; ++k) gj[k] = rec[R::G + (q * NV + j) * D + k];
            A[j] = A[j] + dot_d<D>(gi, gj) * JxW;
            double aux = 0.0;
#pragma unroll
            for (int k = 0; k < D; ++k) aux = aux + rec[R::V + j * D + k] * gi[k];
            b = b + coeff * aux * N[j] * JxW;
          }
        }
        bcell[bpos[c * NV + i]] = b;
      } else {
        const double rho = c0, dt = c1;
        double b[D];
#pragma unroll
        for (int k = 0; k < D; ++k) b[k] = 0.0;
#pragma unroll 1
        for (int q = 0; q < NQ; ++q) {
          double xi[D];
#pragma unroll
          for (int k = 0; k < D; ++k) xi[k] = gauss_pt((q >> k) & 1);
          double N[NV];
          shape<D>(xi, N);
          const double JxW = rec[R::J + q];
          const double aux2 = dt * N[i] * JxW;
#pragma unroll
          for (int j = 0; j < NV; ++j) {
            const double aux = rho * N[i] * N[j] * JxW;
            A[j] = A[j] + aux;
            const double pj = rec[R::P + j];
#pragma unroll
            for (int k = 0; k < D; ++k) {
              b[k] = b[k] + aux * rec[R::V + j * D + k];
              b[k] = b[k] - aux2 * rec[R::G + (q * NV + j) * D + k] * pj;
            }
          }
        }
#pragma unroll
        for (int k = 0; k < D; ++k) bcell[(size_t)k * nc * NV + bpos[c * NV + i]] = b[k];
      }
#pragma unroll
      for (int j = 0; j < NV; ++j) Acell[apos[(c * NV + i) * NV + j]] = A[j];
    }
    __syncthreads();  // the records are rewritten by the next cells
  }
}

// assemble_velocity_prediction (src/pfem2fem.cpp:246-390, SCHEMEB off), row
// i of cell c for every component k, Vo = the old velocity, U = the current
// one, a = mu dt JxW:
//   A^k_ij = sum_q rho N_i N_j JxW + a (viscous stencil of component k)
//   b^k_i  = sum_q sum_j rho N_i N_j JxW Vo_k(j) - a (explicit cross terms) U(j)
// with the stencils and cross terms of the reference, term by term in its
// order (the boundary terms follow in k_fem_predict_faces). Component k's
// cell matrices are plane k of Acell (nc 4^d entries each, at apos), its
// right-hand sides plane k of bcell (nc 2^d entries each, at bpos).
template <int D>
__global__ __launch_bounds__(kFemBlock) void k_fem_predict(DMesh m, int nc, DField Vo, DField U, double rho,
                                                          double mu, double dt, const int* __restrict__ apos,
                                                          const int* __restrict__ bpos, double* __restrict__ Acell,
                                                          double* __restrict__ bcell) {
  using R = FemRec<D>;
  constexpr int NV = 1 << D;
  constexpr int NQ = 1 << D;
  constexpr int CPB = kFemBlock / NV;
  __shared__ double s[CPB * R::size];
  const int lc = threadIdx.x / NV, r = threadIdx.x % NV;
  double* rec = s + lc * R::size;
  const long aplane = (long)nc * NV * NV, bplane = (long)nc * NV;
  for (long base = (long)blockIdx.x * CPB; base < nc; base += (long)gridDim.x * CPB) {
    const long c = base + lc;
    const bool on = c < nc;
    if (on) {
      const int vid = m.cv[c * NV + r], dof = m.cd[c * NV + r];
#pragma unroll
      for (int k = 0; k < D; ++k) {
        rec[R::X + r * D + k] = m.X[k][vid];
        rec[R::V + r * D + k] = Vo.c[k][dof];
        rec[R::U + r * D + k] = U.c[k][dof];
      }
    }
    __syncthreads();
    if (on) fem_geometry<D>(rec, r);
    __syncthreads();
    if (on) {
      const int i = r;
      double A[D][NV], b[D];
#pragma unroll
      for (int k = 0; k < D; ++k) {
        b[k] = 0.0;
#pragma unroll
        for (int j = 0; j < NV; ++j) A[k][j] = 0.0;
      }
#pragma unroll 1
      for (int q = 0; q < NQ; ++q) {
        double xi[D];
#pragma unroll
        for (int k = 0; k < D; ++k) xi[k] = gauss_pt((q >> k) & 1);
        double N[NV];
        shape<D>(xi, N);
        const double weight = rec[R::J + q];
        double gi[D];
#pragma unroll
        for (int k = 0; k < D; ++k) gi[k] = rec[R::G + (q * NV + i) * D + k];
#pragma unroll
        for (int j = 0; j < NV; ++j) {
          double gj[D], u[D];
#pragma unroll
          for (int k = 0; k < D; ++k) {
            gj[k] = rec[R::G + (q * NV + j) * D + k];
            u[k] = rec[R::U + j * D + k];
          }
          double aux = rho * N[i] * N[j] * weight;
#pragma unroll
          for (int k = 0; k < D; ++k) {
            A[k][j] = A[k][j] + aux;
            b[k] = b[k] + aux * rec[R::V + j * D + k];
          }
          aux = mu * dt * weight;
          if constexpr (D == 2) {
            A[0][j] = A[0][j] + aux * (gi[1] * gj[1] + 4.0 / 3.0 * gi[0] * gj[0]);
            A[1][j] = A[1][j] + aux * (gi[0] * gj[0] + 4.0 / 3.0 * gi[1] * gj[1]);
            b[0] = b[0] - aux * (gi[1] * gj[0] - 2.0 / 3.0 * gi[0] * gj[1]) * u[1];
            b[1] = b[1] - aux * (gi[0] * gj[1] - 2.0 / 3.0 * gi[1] * gj[0]) * u[0];
          } else {
            A[0][j] = A[0][j] + aux * (4.0 / 3.0 * gi[0] * gj[0] + gi[1] * gj[1] + gi[2] * gj[2]);
            A[1][j] = A[1][j] + aux * (gj[0] * gi[0] + 4.0 / 3.0 * gi[1] * gj[1] + gi[2] * gj[2]);
            A[2][j] = A[2][j] + aux * (gj[0] * gi[0] + gj[1] * gi[1] + 4.0 / 3.0 * gi[2] * gj[2]);
            b[0] = b[0] - aux * ((gi[1] * gj[0] - 2.0 / 3.0 * gi[0] * gj[1]) * u[1] +
                                 (gi[2] * gj[0] - 2.0 / 3.0 * gi[0] * gj[2]) * u[2]);
            b[1] = b[1] - aux * ((gi[0] * gj[1] - 2.0 / 3.0 * gi[1] * gj[0]) * u[0] +
                                 (gi[2] * gj[1] - 2.0 / 3.0 * gi[1] * gj[2]) * u[2]);
            b[2] = b[2] - aux * ((gi[0] * gj[2] - 2.0 / 3.0 * gi[2] * gj[0]) * u[0] +
                                 (gi[1] * gj[2] - 2.0 / 3.0 * gi[2] * gj[1]) * u[1]);
          }
        }
      }
#pragma unroll
      for (int k = 0; k < D; ++k) {
#pragma unroll
        for (int j = 0; j < NV; ++j) Acell[k * aplane + apos[(c * NV + i) * NV + j]] = A[k][j];
        bcell[k * bplane + bpos[c * NV + i]] = b[k];
      }
    }
    __syncthreads();  // the records are rewritten by the next cells
  }
}

// Boundary terms of assemble_velocity_prediction (src/pfem2fem.cpp:317-381)
// on the faces flagged in faces[c] (boundaries without a velocity Dirichlet
// condition), one thread per flagged cell, face by face and point by point,
// after the cell terms. In 2D, for every (i, j), with c = mu dt:
//   A^0_ij -= c N_i 4/3 d_0 N_j n_0 JxW,   b^0_i += c N_i (-2/3) d_1 N_j U_1(j) n_0 JxW
//   A^1_ij -= c N_i d_0 N_j n_0 JxW,       b^1_i += c N_i d_1 N_j U_0(j) n_0 JxW
// In 3D, right-hand sides only: the traction rows tX, tY, tZ summed over the
// cell's DoFs from the velocity with its first component taken as zero (the
// reference fills components 1..dim-1 only), then b^k_i += (c N_i JxW t_k) . n.
template <int D>
__global__ __launch_bounds__(kBlock) void k_fem_predict_faces(DMesh m, int nc, DField U, double mu, double dt,
                                                             const unsigned char* __restrict__ faces,
                                                             const int* __restrict__ apos,
                                                             const int* __restrict__ bpos, double* __restrict__ Acell,
                                                             double* __restrict__ bcell) {
  constexpr int NV = 1 << D;
  const long aplane = (long)nc * NV * NV, bplane = (long)nc * NV;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < nc; c += gridDim.x * blockDim.x) {
    const int fm = faces[c];
    if (!fm) continue;
    int vid[NV], dof[NV];
    load_ids<D>(m.cv, c, vid);
    load_ids<D>(m.cd, c, dof);
    double Xr[NV * D], Ur[NV][D], b[D][NV];
#pragma unroll
    for (int l = 0; l < NV; ++l)
#pragma unroll
      for (int k = 0; k < D; ++k) {
        Xr[l * D + k] = m.X[k][vid[l]];
        Ur[l][k] = U.c[k][dof[l]];
        b[k][l] = bcell[k * bplane + bpos[(size_t)c * NV + l]];
      }
    const double cm = mu * dt;
    for (int f = 0; f < 2 * D; ++f) {
      if (!((fm >> f) & 1)) continue;
      for (int q = 0; q < (1 << (D - 1)); ++q) {
        double N[NV], g[NV][3], nv[D];
        const double JxW = fem_face_geometry<D>(Xr, f, q, N, g, nv);
        if constexpr (D == 2) {
          for (int i = 0; i < NV; ++i)
            for (int j = 0; j < NV; ++j) {
              const long a0 = apos[((size_t)c * NV + i) * NV + j];
              Acell[a0] = Acell[a0] - cm * N[i] * 4.0 / 3.0 * g[j][0] * nv[0] * JxW;
              b[0][i] = b[0][i] + cm * N[i] * (-2.0 / 3.0) * g[j][1] * Ur[j][1] * nv[0] * JxW;
              Acell[aplane + a0] = Acell[aplane + a0] - cm * N[i] * g[j][0] * nv[0] * JxW;
              b[1][i] = b[1][i] + cm * N[i] * g[j][1] * Ur[j][0] * nv[0] * JxW;
            }
        } else {
          double tX[3] = {0.0, 0.0, 0.0}, tY[3] = {0.0, 0.0, 0.0}, tZ[3] = {0.0, 0.0, 0.0};
          for (int i = 0; i < NV; ++i) {
            const double* gi = g[i];
            const double v0 = 0.0, v1 = Ur[i][1], v2 = Ur[i][2];
            tX[0] = tX[0] + ((4.0 / 3.0) * gi[0] * v0 - (2.0 / 3.0) * gi[1] * v1 - (2.0 / 3.0) * gi[2] * v2);
            tX[1] = tX[1] + (gi[1] * v0 + gi[0] * v1);
            tX[2] = tX[2] + (gi[2] * v0 + gi[0] * v2);
            tY[0] = tY[0] + (gi[1] * v0 + gi[0] * v1);
            tY[1] = tY[1] + ((-2.0 / 3.0) * gi[0] * v0 + (4.0 / 3.0) * gi[1] * v1 - (2.0 / 3.0) * gi[2] * v2);
            tY[2] = tY[2] + (gi[2] * v1 + gi[1] * v2);
            tZ[0] = tZ[0] + (gi[2] * v0 + gi[0] * v2);
            tZ[1] = tZ[1] + (gi[2] * v1 + gi[1] * v2);
            tZ[2] = tZ[2] + ((-2.0 / 3.0) * gi[0] * v0 - (2.0 / 3.0) * gi[1] * v1 + (4.0 / 3.0) * gi[2] * v2);
          }
          for (int i = 0; i < NV; ++i) {
            const double aux = cm * N[i] * JxW;
            const double* t[3] = {tX, tY, tZ};
            for (int k = 0; k < 3; ++k) {
              double dotn = (aux * t[k][0]) * nv[0];
              dotn = dotn + (aux * t[k][1]) * nv[1];
              dotn = dotn + (aux * t[k][2]) * nv[2];
              b[k][i] = b[k][i] + dotn;
            }
          }
        }
      }
    }
#pragma unroll
    for (int l = 0; l < NV; ++l)
#pragma unroll
      for (int k = 0; k < D; ++k) bcell[k * bplane + bpos[(size_t)c * NV + l]] = b[k][l];
  }
}

// Segmented sums for the global matrix and right-hand sides: segment p of
// component k is src[k * plane + off[p] .. k * plane + off[p+1]), summed in
// order (ascending cell) into out_k[p]. The segments of a workgroup's kBlock
// consecutive p form one contiguous run of src, staged through LDS in tiles of
// kFemTile doubles with coalesced loads; each thread then sums its own
// segment's part of every tile in order.
constexpr int kFemTile = 2048;

template <typename OffT>
__global__ __launch_bounds__(kBlock) void k_fem_segsum(long nseg, const OffT* __restrict__ off, int ncomp,
                                                      long plane, const double* __restrict__ src, double* o0,
                                                      double* o1, double* o2) {
  __shared__ double tile[kFemTile];
  double* out[3] = {o0, o1, o2};
  for (long base = (long)blockIdx.x * kBlock; base < nseg; base += (long)gridDim.x * kBlock) {
    const long p = base + threadIdx.x;
    const long lo = off[base], hi = off[base + kBlock < nseg ? base + kBlock : nseg];
    const long mlo = p < nseg ? (long)off[p] : hi, mhi = p < nseg ? (long)off[p + 1] : hi;
    for (int k = 0; k < ncomp; ++k) {
      const double* sk = src + k * plane;
      double s = 0.0;
      for (long t0 = lo; t0 < hi; t0 += kFemTile) {
        const long t1 = t0 + kFemTile < hi ? t0 + kFemTile : hi;
        __syncthreads();  // the previous tile is consumed
        for (long e = t0 + threadIdx.x; e < t1; e += kBlock) tile[e - t0] = sk[e];
        __syncthreads();
        const long a = mlo > t0 ? mlo : t0, b = mhi < t1 ? mhi : t1;
        for (long e = a; e < b; ++e) s = s + tile[e - t0];
      }
      if (p < nseg) out[k][p] = s;
    }
  }
}

}  // namespace pfem2gpu
