// Device-side building blocks of the particle path (gfx950, FP64).
//
// Every floating-point expression here is evaluated with contraction off
// (-ffp-contract=off and the pragma below): the reference rounds each product
// and each sum on its own, and bitwise parity of cell assignments depends on
// reproducing that rounding.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#pragma clang fp contract(off)

namespace pfem2gpu {

constexpr int kBlock = 256;
constexpr int kNewtonMaxIt = 30;
constexpr double kNewtonTol = 1e-12;

// Per-instance particle parameters (one entry for a plain handle, one per
// instance for a batch). src/pfem2particlehandler.cpp:34-39,47,362-364
struct Prm {
  double dt_sub;   // dt / (double)n_sub
  double h[3];     // 1.0 / q_k (sub-cell part width in reference coordinates)
  int n_sub;
  int maxpp;       // max_particles_per_cell_part
  int q[3];        // parts per direction
  int Q[3];        // q_k + 1: part buckets per direction (xi_k == 1 gives index q_k)
  int P;           // prod Q_k: part buckets per cell
  int Pvalid;      // prod q_k: refillable parts per cell
};

struct DMesh {
  const double* X[3];            // vertex coordinates, SoA
  const int* cv;                 // [nc][2^D] vertex ids
  const int* cd;                 // [nc][2^D] dof ids
  const unsigned char* box;      // [nc] axis-aligned box flag
  const int* v2c_off;            // [nv+1]
  const int* v2c;                // vertex -> cells, ascending
  const double* v2c_dir;         // [nnz][D] unit directions vertex -> cell centre
  const int* cell_inst;          // [nc] instance of each cell, or null (single)
  const Prm* prm;                // per-instance parameters
};

struct DField {
  const double* c[3];
};

struct DParts {
  double* x[3];
  double* v[3];
  int* cell;
};

struct DTmp {
  double* x[3];
  double* v[3];
  int* cell;
  unsigned short* part;
  unsigned long long* key;
};

__device__ __forceinline__ int inst_of(const DMesh& m, int c) { return m.cell_inst ? m.cell_inst[c] : 0; }

template <int D>
__device__ __forceinline__ void load_ids(const int* __restrict__ base, int c, int* out) {
  if constexpr (D == 2) {
    const int4 a = *reinterpret_cast<const int4*>(base + 4 * (size_t)c);
    out[0] = a.x; out[1] = a.y; out[2] = a.z; out[3] = a.w;
  } else {
    const int4* p = reinterpret_cast<const int4*>(base + 8 * (size_t)c);
    const int4 a = p[0], b = p[1];
    out[0] = a.x; out[1] = a.y; out[2] = a.z; out[3] = a.w;
    out[4] = b.x; out[5] = b.y; out[6] = b.z; out[7] = b.w;
  }
}

// Q1 shape values, Lagrange product form in k order (FE_Q<dim>(1)::shape_value).
template <int D>
__device__ __forceinline__ void shape(const double* xi, double* N) {
  double l0[D], l1[D];
#pragma unroll
  for (int k = 0; k < D; ++k) {
    l0[k] = 1.0 - xi[k];
    l1[k] = xi[k];
  }
#pragma unroll
  for (int i = 0; i < (1 << D); ++i) {
    double val = (i & 1) ? l1[0] : l0[0];
#pragma unroll
    for (int k = 1; k < D; ++k) val = val * (((i >> k) & 1) ? l1[k] : l0[k]);
    N[i] = val;
  }
}

template <int D>
__device__ __forceinline__ void shape_grad(const double* xi, double (*dN)[3]) {
#pragma unroll
  for (int i = 0; i < (1 << D); ++i)
#pragma unroll
    for (int l = 0; l < D; ++l) {
      double val = 1.0;
#pragma unroll
      for (int k = 0; k < D; ++k) {
        double f;
        if (k == l)
          f = ((i >> k) & 1) ? 1.0 : -1.0;
        else
          f = ((i >> k) & 1) ? xi[k] : (1.0 - xi[k]);
        val = val * f;
      }
      dN[i][l] = val;
    }
}

// Forward Q1 map x = sum_i N_i(xi) X_i, sequential in i.
template <int D>
__device__ __forceinline__ void fwd(const DMesh& m, const int* vid, const double* xi, double* x) {
  double N[1 << D];
  shape<D>(xi, N);
#pragma unroll
  for (int k = 0; k < D; ++k) x[k] = 0.0;
#pragma unroll
  for (int i = 0; i < (1 << D); ++i)
#pragma unroll
    for (int k = 0; k < D; ++k) x[k] = x[k] + N[i] * m.X[k][vid[i]];
}

// Cramer solve of A d = r; returns det.
template <int D>
__device__ __forceinline__ double solve(const double (*A)[3], const double* r, double* d) {
  if constexpr (D == 2) {
    const double det = A[0][0] * A[1][1] - A[0][1] * A[1][0];
    d[0] = (r[0] * A[1][1] - A[0][1] * r[1]) / det;
    d[1] = (A[0][0] * r[1] - r[0] * A[1][0]) / det;
    return det;
  } else {
    const double c00 = A[1][1] * A[2][2] - A[1][2] * A[2][1];
    const double c01 = A[1][2] * A[2][0] - A[1][0] * A[2][2];
    const double c02 = A[1][0] * A[2][1] - A[1][1] * A[2][0];
    const double det = A[0][0] * c00 + A[0][1] * c01 + A[0][2] * c02;
    const double d0n = r[0] * c00 + A[0][1] * (A[1][2] * r[2] - r[1] * A[2][2]) +
                       A[0][2] * (r[1] * A[2][1] - A[1][1] * r[2]);
    const double d1n = A[0][0] * (r[1] * A[2][2] - A[1][2] * r[2]) + r[0] * c01 +
                       A[0][2] * (A[1][0] * r[2] - r[1] * A[2][0]);
    const double d2n = A[0][0] * (A[1][1] * r[2] - r[1] * A[2][1]) +
                       A[0][1] * (r[1] * A[2][0] - A[1][0] * r[2]) + r[0] * c02;
    d[0] = d0n / det;
    d[1] = d1n / det;
    d[2] = d2n / det;
    return det;
  }
}

// Inverse Q1 map (MappingQ1::transform_real_to_unit_cell restated): closed
// form on axis-aligned boxes, Newton from the affine guess otherwise. Returns
// false when the transformation fails.
template <int D>
__device__ bool inv_map(const DMesh& m, int c, const int* vid, const double* x, double* xi) {
  if (m.box[c]) {
#pragma unroll
    for (int k = 0; k < D; ++k) {
      const double x0 = m.X[k][vid[0]];
      const double x1 = m.X[k][vid[1 << k]];
      xi[k] = (x[k] - x0) / (x1 - x0);
    }
    return true;
  }
  double A[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  double r[3] = {0, 0, 0};
  double X0[D];
#pragma unroll
  for (int k = 0; k < D; ++k) X0[k] = m.X[k][vid[0]];
#pragma unroll
  for (int l = 0; l < D; ++l)
#pragma unroll
    for (int k = 0; k < D; ++k) A[k][l] = m.X[k][vid[1 << l]] - X0[k];
#pragma unroll
  for (int k = 0; k < D; ++k) r[k] = x[k] - X0[k];
  double det = solve<D>(A, r, xi);
  if (!(det != 0.0) || !isfinite(det)) return false;
  for (int it = 0; it < kNewtonMaxIt; ++it) {
    double N[1 << D], dN[1 << D][3];
    shape<D>(xi, N);
    shape_grad<D>(xi, dN);
    double Fx[3] = {0, 0, 0};
    double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
#pragma unroll
    for (int i = 0; i < (1 << D); ++i)
#pragma unroll
      for (int k = 0; k < D; ++k) {
        const double Xik = m.X[k][vid[i]];
        Fx[k] = Fx[k] + N[i] * Xik;
#pragma unroll
        for (int l = 0; l < D; ++l) J[k][l] = J[k][l] + dN[i][l] * Xik;
      }
    double res[3] = {0, 0, 0}, d[3] = {0, 0, 0};
#pragma unroll
    for (int k = 0; k < D; ++k) res[k] = x[k] - Fx[k];
    det = solve<D>(J, res, d);
    if (!(det != 0.0) || !isfinite(det)) return false;
    double dmax = 0.0;
#pragma unroll
    for (int k = 0; k < D; ++k) {
      xi[k] = xi[k] + d[k];
      const double a = fabs(d[k]);
      dmax = (a > dmax) ? a : dmax;
    }
    if (!(dmax == dmax)) return false;
    if (dmax <= kNewtonTol) return true;
  }
  return false;
}

// GeometryInfo<dim>::is_inside_unit_cell: exact, NaN is outside.
template <int D>
__device__ __forceinline__ bool inside(const double* xi) {
  bool in = true;
#pragma unroll
  for (int k = 0; k < D; ++k) in = in && (xi[k] >= 0.0) && (xi[k] <= 1.0);
  return in;
}

// u_k = sum_i N_i(xi) F_k[dof_i], vertex-outer, component-inner, from 0.0.
template <int D>
__device__ __forceinline__ void gather(const DField& F, const int* dof, const double* xi, double* u) {
  double N[1 << D];
  shape<D>(xi, N);
#pragma unroll
  for (int k = 0; k < D; ++k) u[k] = 0.0;
#pragma unroll
  for (int i = 0; i < (1 << D); ++i)
#pragma unroll
    for (int k = 0; k < D; ++k) u[k] = u[k] + N[i] * F.c[k][dof[i]];
}

// Part index per direction, (int)(xi_k / h_k) as in
// src/pfem2particlehandler.cpp:369-370; linearised with x slowest.
template <int D>
__device__ __forceinline__ int part_lin(const Prm& p, const double* xi, bool* valid) {
  int lin = 0;
  bool ok = true;
#pragma unroll
  for (int k = 0; k < D; ++k) {
    const int pk = (int)(xi[k] / p.h[k]);
    ok = ok && (pk >= 0) && (pk < p.q[k]);
    lin = lin * p.Q[k] + pk;
  }
  if (valid) *valid = ok;
  return lin;
}

// Reference position of a located particle from (cell, x): the inverse map,
// snapped to the exact sub-cell centre when x is bitwise the forward map of
// that centre (particles placed by seeding or refill keep the exact centre,
// src/pfem2particlehandler.cpp:88-91,390-394).
template <int D>
__device__ __forceinline__ void recover_xi(const DMesh& m, const Prm& p, int c, const int* vid, const double* x,
                                           double* xi) {
  inv_map<D>(m, c, vid, x, xi);
  int pi[D];
  bool valid = true;
#pragma unroll
  for (int k = 0; k < D; ++k) {
    pi[k] = (int)(xi[k] / p.h[k]);
    valid = valid && (pi[k] >= 0) && (pi[k] < p.q[k]);
  }
  if (!valid) return;
  double xic[D], xc[D];
#pragma unroll
  for (int k = 0; k < D; ++k) xic[k] = ((double)pi[k] + 1.0 / 2) * p.h[k];
  fwd<D>(m, vid, xic, xc);
  bool eq = true;
#pragma unroll
  for (int k = 0; k < D; ++k) eq = eq && (xc[k] == x[k]);
  if (eq) {
#pragma unroll
    for (int k = 0; k < D; ++k) xi[k] = xic[k];
  }
}

__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
  z ^= z >> 30;
  z *= 0xbf58476d1ce4e5b9ULL;
  z ^= z >> 27;
  z *= 0x94d049bb133111ebULL;
  z ^= z >> 31;
  return z;
}

// Canonical key: hash of the position bits (survivor choice and in-cell order).
template <int D>
__device__ __forceinline__ unsigned long long particle_key(const double* x) {
  unsigned long long h = 0x9E3779B97F4A7C15ULL;
#pragma unroll
  for (int k = 0; k < D; ++k) h = mix64(h ^ (unsigned long long)__double_as_longlong(x[k]));
  return h;
}

template <int D>
__device__ __forceinline__ double dot_dir(const double* __restrict__ dir, const double* a) {
  double s = dir[0] * a[0];
#pragma unroll
  for (int k = 1; k < D; ++k) s = s + dir[k] * a[k];
  return s;
}

// Locate after an Euler substep (src/pfem2particlehandler.cpp:420-492 with
// src/pfem2particle.cpp:186-211). On entry (c, vid, xi) describe the cell and
// reference position before the substep and u is the advecting velocity of
// the substep. Returns false if the particle is to be deleted.
template <int D>
__device__ bool locate(const DMesh& m, int& c, int* vid, double* xi, const double* x, const double* u) {
  constexpr int NV = 1 << D;
  double xn[D];
  if (inv_map<D>(m, c, vid, x, xn) && inside<D>(xn)) {
#pragma unroll
    for (int k = 0; k < D; ++k) xi[k] = xn[k];
    return true;
  }
  // closest vertex of the current cell along the velocity, from the old position
  double xold[D];
  fwd<D>(m, vid, xi, xold);
  double n2 = 0.0;
#pragma unroll
  for (int k = 0; k < D; ++k) n2 = n2 + u[k] * u[k];
  const double inv_un = 1.0 / sqrt(n2);
  double vn[D];
#pragma unroll
  for (int k = 0; k < D; ++k) vn[k] = u[k] * inv_un;
  double best = 0.0;
  int bestv = 0;
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    double d[D];
#pragma unroll
    for (int k = 0; k < D; ++k) d[k] = m.X[k][vid[v]] - xold[k];
    double m2 = 0.0;
#pragma unroll
    for (int k = 0; k < D; ++k) m2 = m2 + d[k] * d[k];
    const double inv_m = 1.0 / sqrt(m2);
#pragma unroll
    for (int k = 0; k < D; ++k) d[k] = d[k] * inv_m;
    double a = vn[0] * d[0];
#pragma unroll
    for (int k = 1; k < D; ++k) a = a + vn[k] * d[k];
    if (v == 0) {
      best = a;
    } else if (a > best) {
      best = a;
      bestv = v;
    }
  }
  int cvg = vid[0];
#pragma unroll
  for (int v = 1; v < NV; ++v)
    if (bestv == v) cvg = vid[v];
  double vtp[D];
#pragma unroll
  for (int k = 0; k < D; ++k) vtp[k] = x[k] - m.X[k][cvg];
  double t2 = 0.0;
#pragma unroll
  for (int k = 0; k < D; ++k) t2 = t2 + vtp[k] * vtp[k];
  const double inv_t = 1.0 / sqrt(t2);
#pragma unroll
  for (int k = 0; k < D; ++k) vtp[k] = vtp[k] * inv_t;

  const int o0 = m.v2c_off[cvg];
  const int nn = m.v2c_off[cvg + 1] - o0;
  int vid2[NV];
  constexpr int MAXC = 8;
  if (nn <= MAXC) {
    double dots[MAXC];
#pragma unroll
    for (int j = 0; j < MAXC; ++j) {
      dots[j] = 0.0;
      if (j < nn) dots[j] = dot_dir<D>(m.v2c_dir + (size_t)(o0 + j) * D, vtp);
    }
    unsigned used = 0;
    for (int r = 0; r < nn; ++r) {
      int bj = -1;
      double bd = 0.0;
#pragma unroll
      for (int j = 0; j < MAXC; ++j) {
        if (j < nn && !((used >> j) & 1u)) {
          if (bj < 0 || dots[j] > bd) {
            bj = j;
            bd = dots[j];
          }
        }
      }
      used |= 1u << bj;
      const int cc = m.v2c[o0 + bj];
      if (cc == c) continue;  // the current cell already failed the same test
      load_ids<D>(m.cv, cc, vid2);
      if (inv_map<D>(m, cc, vid2, x, xn) && inside<D>(xn)) {
        c = cc;
#pragma unroll
        for (int i = 0; i < NV; ++i) vid[i] = vid2[i];
#pragma unroll
        for (int k = 0; k < D; ++k) xi[k] = xn[k];
        return true;
      }
    }
    return false;
  }
  // more than MAXC cells at this vertex: same order, selected without storage
  const bool nan_mode = !(vtp[0] == vtp[0]);
  int pj = -1;
  double pd = 0.0;
  for (int r = 0; r < nn; ++r) {
    int bj = -1;
    double bd = 0.0;
    for (int j = 0; j < nn; ++j) {
      const double dj = dot_dir<D>(m.v2c_dir + (size_t)(o0 + j) * D, vtp);
      bool after;
      if (r == 0)
        after = true;
      else if (nan_mode)
        after = j > pj;
      else
        after = (dj < pd) || (dj == pd && j > pj);
      if (!after) continue;
      if (bj < 0 || dj > bd) {
        bj = j;
        bd = dj;
      }
    }
    pj = bj;
    pd = bd;
    const int cc = m.v2c[o0 + bj];
    if (cc == c) continue;
    load_ids<D>(m.cv, cc, vid2);
    if (inv_map<D>(m, cc, vid2, x, xn) && inside<D>(xn)) {
      c = cc;
#pragma unroll
      for (int i = 0; i < NV; ++i) vid[i] = vid2[i];
#pragma unroll
      for (int k = 0; k < D; ++k) xi[k] = xn[k];
      return true;
    }
  }
  return false;
}

}  // namespace pfem2gpu
