// TEST INFRASTRUCTURE ONLY — the CPU oracle of the FEM assembly.
//
// A from-scratch CPU restatement of pfem2Fem's pressure-equation and
// velocity-correction assembly, the parity checker of the HIP assembly
// kernels. Nothing in the product links, loads or calls this file.
//
// Reference files followed (paths relative to the reference repository):
//   src/pfem2fem.h:24              QUAD_POINTS_PER_DIRECTION = 2 (QGauss, cells and faces)
//   src/pfem2fem.cpp:18-22         FEValues flags (values, gradients, JxW; faces also normals)
//   src/pfem2fem.cpp:78-95         DoFTools::make_sparsity_pattern (DoFs sharing a cell)
//   src/pfem2fem.cpp:246-390       assemble_velocity_prediction (SCHEMEB off)
//   src/pfem2fem.cpp:392-458       assemble_pressure_equation (SCHEMEB off)
//   src/pfem2fem.cpp:460-517       assemble_velocity_correction (SCHEMEB off)
// deal.II functions restated from their mathematics (not in the reference tree):
//   QGauss<1>(2) on [0,1]: points (1 -+ 1/sqrt(3))/2, weights 1/2; tensor
//     points with direction 0 fastest.
//   FEValues with MappingQ1: J = sum_i X_i (x) grad_xi N_i, JxW = det J * w,
//     shape_grad = J^-T grad_xi N (J^-1 from the adjugate).
//   FEFaceValues: face f = 2k + side at xi_k = side; normal J^-T n_ref
//     normalised, JxW = det J |J^-T n_ref| w (Nanson's formula).
//   AffineConstraints::distribute_local_to_global without constraints: global
//     (i, j) += local (i, j) and b_i += local b_i, cell by cell in ascending
//     order, into matrices and vectors that start at zero.
// Parity with deal.II itself is unpinned (it is not available to run); the
// restatement is pinned by closed forms in tests/test_fem_oracle.py.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

namespace {

const double kGp[2] = {0.21132486540518711775, 0.78867513459481288225};

struct Geo {  // FEValues at one point of one cell
  double N[8];
  double grad[8][3];
  double JxW;
  double normal[3];
};

struct Cell {
  int dim, nv;
  double X[8][3];
};

// Q1 shape values and reference gradients at xi.
void q1(int dim, const double* xi, double* N, double (*dN)[3]) {
  const int nv = 1 << dim;
  for (int i = 0; i < nv; ++i) {
    double v = ((i & 1) ? xi[0] : 1.0 - xi[0]);
    for (int k = 1; k < dim; ++k) v = v * (((i >> k) & 1) ? xi[k] : 1.0 - xi[k]);
    N[i] = v;
    for (int l = 0; l < dim; ++l) {
      double g = 1.0;
      for (int k = 0; k < dim; ++k) {
        const double f = (k == l) ? (((i >> k) & 1) ? 1.0 : -1.0) : (((i >> k) & 1) ? xi[k] : 1.0 - xi[k]);
        g = g * f;
      }
      dN[i][l] = g;
    }
  }
}

// Jacobian, determinant and inverse; inv[l][k] = (J^-1)[l][k].
double jac_inv(const Cell& c, const double (*dN)[3], double (*inv)[3]) {
  const int dim = c.dim;
  double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  for (int i = 0; i < c.nv; ++i)
    for (int k = 0; k < dim; ++k)
      for (int l = 0; l < dim; ++l) J[k][l] = J[k][l] + c.X[i][k] * dN[i][l];
  double det;
  if (dim == 2) {
    det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
    inv[0][0] = J[1][1] / det;
    inv[0][1] = -J[0][1] / det;
    inv[1][0] = -J[1][0] / det;
    inv[1][1] = J[0][0] / det;
  } else {
    const double a = J[1][1] * J[2][2] - J[1][2] * J[2][1];
    const double b = J[1][2] * J[2][0] - J[1][0] * J[2][2];
    const double d = J[1][0] * J[2][1] - J[1][1] * J[2][0];
    det = J[0][0] * a + J[0][1] * b + J[0][2] * d;
    inv[0][0] = a / det;
    inv[1][0] = b / det;
    inv[2][0] = d / det;
    inv[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / det;
    inv[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / det;
    inv[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / det;
    inv[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / det;
    inv[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / det;
    inv[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / det;
  }
  return det;
}

void real_grads(const Cell& c, const double (*dN)[3], const double (*inv)[3], double (*grad)[3]) {
  for (int i = 0; i < c.nv; ++i)
    for (int k = 0; k < c.dim; ++k) {
      double s = dN[i][0] * inv[0][k];
      for (int l = 1; l < c.dim; ++l) s = s + dN[i][l] * inv[l][k];
      grad[i][k] = s;
    }
}

// fe_values.reinit(cell): every cell quadrature point.
void reinit(const Cell& c, std::vector<Geo>& g) {
  const int nq = 1 << c.dim;
  g.resize(nq);
  for (int q = 0; q < nq; ++q) {
    double xi[3], w = 1.0, dN[8][3], inv[3][3];
    for (int k = 0; k < c.dim; ++k) {
      xi[k] = kGp[(q >> k) & 1];
      w = w * 0.5;
    }
    q1(c.dim, xi, g[q].N, dN);
    const double det = jac_inv(c, dN, inv);
    g[q].JxW = det * w;
    real_grads(c, dN, inv, g[q].grad);
  }
}

// fe_face_values.reinit(cell, f): the points of face f.
void reinit_face(const Cell& c, int f, std::vector<Geo>& g) {
  const int nq = 1 << (c.dim - 1), kf = f >> 1;
  g.resize(nq);
  for (int q = 0; q < nq; ++q) {
    double xi[3], w = 1.0, dN[8][3], inv[3][3];
    int b = 0;
    for (int k = 0; k < c.dim; ++k) {
      if (k == kf) {
        xi[k] = (double)(f & 1);
      } else {
        xi[k] = kGp[(q >> b) & 1];
        ++b;
      }
    }
    for (int k = 0; k < c.dim - 1; ++k) w = w * 0.5;
    q1(c.dim, xi, g[q].N, dN);
    const double det = jac_inv(c, dN, inv);
    const double sgn = (f & 1) ? 1.0 : -1.0;
    double n2 = 0.0;
    for (int k = 0; k < c.dim; ++k) {
      g[q].normal[k] = sgn * inv[kf][k];
      n2 = n2 + g[q].normal[k] * g[q].normal[k];
    }
    const double nn = std::sqrt(n2);
    for (int k = 0; k < c.dim; ++k) g[q].normal[k] = g[q].normal[k] / nn;
    g[q].JxW = det * nn * w;
    real_grads(c, dN, inv, g[q].grad);
  }
}

struct Sys {
  const int64_t* rowptr;
  const int32_t* cols;
  double* vals;  // may be null
  long find(int r, int c) const {
    const int32_t* b = cols + rowptr[r];
    const int32_t* e = cols + rowptr[r + 1];
    const int32_t* it = std::lower_bound(b, e, c);
    return (it != e && *it == c) ? (long)(rowptr[r] + (it - b)) : -1;
  }
};

Cell load_cell(int dim, int64_t nverts, const double* X, const int* cv, long c) {
  Cell cell;
  cell.dim = dim;
  cell.nv = 1 << dim;
  for (int i = 0; i < cell.nv; ++i)
    for (int k = 0; k < dim; ++k) cell.X[i][k] = X[k * nverts + cv[c * cell.nv + i]];
  return cell;
}

}  // namespace

extern "C" {

// Sparsity pattern of the DoFs that share a cell: row_ptr [ndofs+1]; cols
// written when not null (call once with null for the size).
int64_t orc_fem_pattern(int dim, int64_t nc, int64_t ndofs, const int* cd, int64_t* rowptr, int32_t* cols) {
  const int nv = 1 << dim;
  std::vector<std::vector<int>> rows(ndofs);
  for (long c = 0; c < nc; ++c)
    for (int i = 0; i < nv; ++i)
      for (int j = 0; j < nv; ++j) rows[cd[c * nv + i]].push_back(cd[c * nv + j]);
  int64_t n = 0;
  rowptr[0] = 0;
  for (long r = 0; r < ndofs; ++r) {
    std::vector<int>& row = rows[r];
    std::sort(row.begin(), row.end());
    row.erase(std::unique(row.begin(), row.end()), row.end());
    for (int col : row) {
      if (cols) cols[n] = col;
      ++n;
    }
    rowptr[r + 1] = n;
    std::vector<int>().swap(row);
  }
  return n;
}

// assemble_pressure_equation: vals [nnz] (may be null), rhs [ndofs]. V is
// [dim][ndofs]; faces [nc] bit f = face f flagged (may be null). Returns 0,
// or -1 when a cell entry is missing from the pattern.
int orc_fem_pressure(int dim, int64_t nverts, const double* X, int64_t nc, const int* cv, const int* cd,
                     int64_t ndofs, const double* V, const unsigned char* faces, double rho, double dt,
                     const int64_t* rowptr, const int32_t* cols, double* vals, double* rhs) {
  const int nv = 1 << dim, nq = 1 << dim;
  Sys S{rowptr, cols, vals};
  if (vals) std::fill(vals, vals + rowptr[ndofs], 0.0);
  std::fill(rhs, rhs + ndofs, 0.0);
  const double coeff = rho / dt;
  std::vector<Geo> g, gf;
  for (long c = 0; c < nc; ++c) {
    const Cell cell = load_cell(dim, nverts, X, cv, c);
    reinit(cell, g);
    double A[8][8] = {}, b[8] = {};
    const int* dof = cd + c * nv;
    for (int q = 0; q < nq; ++q) {
      const double weight = g[q].JxW;
      for (int i = 0; i < nv; ++i) {
        const double* Ni = g[q].grad[i];
        for (int j = 0; j < nv; ++j) {
          const double* Nj = g[q].grad[j];
          double d = Ni[0] * Nj[0];
          for (int k = 1; k < dim; ++k) d = d + Ni[k] * Nj[k];
          const double aux0 = d * weight;
          A[i][j] = A[i][j] + aux0;
          double aux = 0.0;
          for (int k = 0; k < dim; ++k) aux = aux + V[k * ndofs + dof[j]] * Ni[k];
          b[i] = b[i] + coeff * aux * g[q].N[j] * weight;
        }
      }
    }
    const int fm = faces ? faces[c] : 0;
    for (int f = 0; f < 2 * dim; ++f) {
      if (!((fm >> f) & 1)) continue;
      reinit_face(cell, f, gf);
      for (size_t q = 0; q < gf.size(); ++q) {
        double Vq[3] = {0.0, 0.0, 0.0};
        for (int i = 0; i < nv; ++i)
          for (int k = 0; k < dim; ++k) Vq[k] = Vq[k] + gf[q].N[i] * V[k * ndofs + dof[i]];
        for (int i = 0; i < nv; ++i) {
          const double s = coeff * gf[q].N[i];
          double d = (s * Vq[0]) * gf[q].normal[0];
          for (int k = 1; k < dim; ++k) d = d + (s * Vq[k]) * gf[q].normal[k];
          b[i] = b[i] - d * gf[q].JxW;
        }
      }
    }
    for (int i = 0; i < nv; ++i) {
      if (vals)
        for (int j = 0; j < nv; ++j) {
          const long p = S.find(dof[i], dof[j]);
          if (p < 0) return -1;
          vals[p] = vals[p] + A[i][j];
        }
      rhs[dof[i]] = rhs[dof[i]] + b[i];
    }
  }
  return 0;
}

// assemble_velocity_prediction (src/pfem2fem.cpp:246-390): vals [dim][nnz]
// (may be null), rhs [dim][ndofs]; Vo (old velocity) and U (current) are
// [dim][ndofs]; faces [nc] bit f = face f without a velocity Dirichlet
// condition (may be null).
int orc_fem_prediction(int dim, int64_t nverts, const double* X, int64_t nc, const int* cv, const int* cd,
                       int64_t ndofs, const double* Vo, const double* U, const unsigned char* faces, double rho,
                       double mu, double time_step, const int64_t* rowptr, const int32_t* cols, double* vals,
                       double* rhs) {
  const int nv = 1 << dim, nq = 1 << dim;
  const int64_t nnz = rowptr[ndofs];
  Sys S{rowptr, cols, vals};
  if (vals) std::fill(vals, vals + dim * nnz, 0.0);
  std::fill(rhs, rhs + dim * ndofs, 0.0);
  std::vector<Geo> g, gf;
  for (long c = 0; c < nc; ++c) {
    const Cell cell = load_cell(dim, nverts, X, cv, c);
    reinit(cell, g);
    double A[3][8][8] = {}, b[3][8] = {};
    const int* dof = cd + c * nv;
    for (int q = 0; q < nq; ++q) {
      const double weight = g[q].JxW;
      for (int i = 0; i < nv; ++i) {
        const double Ni = g[q].N[i];
        const double* gi = g[q].grad[i];
        for (int j = 0; j < nv; ++j) {
          const int jd = dof[j];
          const double Nj = g[q].N[j];
          const double* gj = g[q].grad[j];
          double aux = rho * Ni * Nj * weight;
          for (int k = 0; k < dim; ++k) {
            A[k][i][j] += aux;
            b[k][i] += aux * Vo[k * ndofs + jd];
          }
          aux = mu * time_step * weight;
          if (dim == 2) {
            A[0][i][j] += aux * (gi[1] * gj[1] + 4.0 / 3.0 * gi[0] * gj[0]);
            A[1][i][j] += aux * (gi[0] * gj[0] + 4.0 / 3.0 * gi[1] * gj[1]);
            b[0][i] -= aux * (gi[1] * gj[0] - 2.0 / 3.0 * gi[0] * gj[1]) * U[1 * ndofs + jd];
            b[1][i] -= aux * (gi[0] * gj[1] - 2.0 / 3.0 * gi[1] * gj[0]) * U[0 * ndofs + jd];
          } else {
            A[0][i][j] += aux * (4.0 / 3.0 * gi[0] * gj[0] + gi[1] * gj[1] + gi[2] * gj[2]);
            A[1][i][j] += aux * (gj[0] * gi[0] + 4.0 / 3.0 * gi[1] * gj[1] + gi[2] * gj[2]);
            A[2][i][j] += aux * (gj[0] * gi[0] + gj[1] * gi[1] + 4.0 / 3.0 * gi[2] * gj[2]);
            b[0][i] -= aux * ((gi[1] * gj[0] - 2.0 / 3.0 * gi[0] * gj[1]) * U[1 * ndofs + jd] +
                              (gi[2] * gj[0] - 2.0 / 3.0 * gi[0] * gj[2]) * U[2 * ndofs + jd]);
            b[1][i] -= aux * ((gi[0] * gj[1] - 2.0 / 3.0 * gi[1] * gj[0]) * U[0 * ndofs + jd] +
                              (gi[2] * gj[1] - 2.0 / 3.0 * gi[1] * gj[2]) * U[2 * ndofs + jd]);
            b[2][i] -= aux * ((gi[0] * gj[2] - 2.0 / 3.0 * gi[2] * gj[0]) * U[0 * ndofs + jd] +
                              (gi[1] * gj[2] - 2.0 / 3.0 * gi[2] * gj[1]) * U[1 * ndofs + jd]);
          }
        }
      }
    }
    const int fm = faces ? faces[c] : 0;
    for (int f = 0; f < 2 * dim; ++f) {
      if (!((fm >> f) & 1)) continue;
      reinit_face(cell, f, gf);
      for (size_t q = 0; q < gf.size(); ++q) {
        const Geo& G = gf[q];
        if (dim == 2) {
          for (int i = 0; i < nv; ++i)
            for (int j = 0; j < nv; ++j) {
              A[0][i][j] -= mu * time_step * G.N[i] * 4.0 / 3.0 * G.grad[j][0] * G.normal[0] * G.JxW;
              b[0][i] += mu * time_step * G.N[i] * (-2.0 / 3.0) * G.grad[j][1] * U[1 * ndofs + dof[j]] *
                         G.normal[0] * G.JxW;
              A[1][i][j] -= mu * time_step * G.N[i] * G.grad[j][0] * G.normal[0] * G.JxW;
              b[1][i] += mu * time_step * G.N[i] * G.grad[j][1] * U[0 * ndofs + dof[j]] * G.normal[0] * G.JxW;
            }
        } else {
          double tX[3] = {0, 0, 0}, tY[3] = {0, 0, 0}, tZ[3] = {0, 0, 0};
          for (int i = 0; i < nv; ++i) {
            const double* gi = G.grad[i];
            double vel[3] = {0.0, 0.0, 0.0};  // the reference fills components 1..dim-1 only
            for (int j = 1; j < dim; ++j) vel[j] = U[j * ndofs + dof[i]];
            tX[0] += (4.0 / 3.0) * gi[0] * vel[0] - (2.0 / 3.0) * gi[1] * vel[1] - (2.0 / 3.0) * gi[2] * vel[2];
            tX[1] += gi[1] * vel[0] + gi[0] * vel[1];
            tX[2] += gi[2] * vel[0] + gi[0] * vel[2];
            tY[0] += gi[1] * vel[0] + gi[0] * vel[1];
            tY[1] += (-2.0 / 3.0) * gi[0] * vel[0] + (4.0 / 3.0) * gi[1] * vel[1] - (2.0 / 3.0) * gi[2] * vel[2];
            tY[2] += gi[2] * vel[1] + gi[1] * vel[2];
            tZ[0] += gi[2] * vel[0] + gi[0] * vel[2];
            tZ[1] += gi[2] * vel[1] + gi[1] * vel[2];
            tZ[2] += (-2.0 / 3.0) * gi[0] * vel[0] - (2.0 / 3.0) * gi[1] * vel[1] + (4.0 / 3.0) * gi[2] * vel[2];
          }
          const double* t[3] = {tX, tY, tZ};
          for (int i = 0; i < nv; ++i) {
            const double aux = mu * time_step * G.N[i] * G.JxW;
            for (int k = 0; k < 3; ++k) {
              double d = (aux * t[k][0]) * G.normal[0];
              for (int l = 1; l < 3; ++l) d += (aux * t[k][l]) * G.normal[l];
              b[k][i] += d;
            }
          }
        }
      }
    }
    for (int i = 0; i < nv; ++i) {
      if (vals)
        for (int j = 0; j < nv; ++j) {
          const long p = S.find(dof[i], dof[j]);
          if (p < 0) return -1;
          for (int k = 0; k < dim; ++k) vals[k * nnz + p] += A[k][i][j];
        }
      for (int k = 0; k < dim; ++k) rhs[k * ndofs + dof[i]] += b[k][i];
    }
  }
  return 0;
}

// assemble_velocity_correction: vals [nnz] (may be null), rhs [dim][ndofs].
int orc_fem_vcorr(int dim, int64_t nverts, const double* X, int64_t nc, const int* cv, const int* cd, int64_t ndofs,
                  const double* V, const double* P, double rho, double dt, const int64_t* rowptr,
                  const int32_t* cols, double* vals, double* rhs) {
  const int nv = 1 << dim, nq = 1 << dim;
  Sys S{rowptr, cols, vals};
  if (vals) std::fill(vals, vals + rowptr[ndofs], 0.0);
  std::fill(rhs, rhs + dim * ndofs, 0.0);
  std::vector<Geo> g;
  for (long c = 0; c < nc; ++c) {
    const Cell cell = load_cell(dim, nverts, X, cv, c);
    reinit(cell, g);
    double A[8][8] = {}, b[3][8] = {};
    const int* dof = cd + c * nv;
    for (int q = 0; q < nq; ++q) {
      const double weight = g[q].JxW;
      for (int i = 0; i < nv; ++i) {
        const double Ni = g[q].N[i];
        const double aux2 = dt * Ni * weight;
        for (int j = 0; j < nv; ++j) {
          const double aux = rho * Ni * g[q].N[j] * weight;
          A[i][j] = A[i][j] + aux;
          for (int k = 0; k < dim; ++k) {
            b[k][i] = b[k][i] + aux * V[k * ndofs + dof[j]];
            b[k][i] = b[k][i] - aux2 * g[q].grad[j][k] * P[dof[j]];
          }
        }
      }
    }
    for (int i = 0; i < nv; ++i) {
      if (vals)
        for (int j = 0; j < nv; ++j) {
          const long p = S.find(dof[i], dof[j]);
          if (p < 0) return -1;
          vals[p] = vals[p] + A[i][j];
        }
      for (int k = 0; k < dim; ++k) rhs[k * ndofs + dof[i]] = rhs[k * ndofs + dof[i]] + b[k][i];
    }
  }
  return 0;
}

}  // extern "C"
