// Device-side building blocks of the particle path (gfx950, FP64).
//
// Every floating-point expression here is evaluated with contraction off
// (-ffp-contract=off and the pragma below): the reference rounds each product
// and each sum on its own, and bitwise parity of cell assignments depends on
// reproducing that rounding.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#pragma clang fp contract(off)

namespace pfem2gpu {

constexpr int kBlock = 256;
constexpr int kNewtonMaxIt = 30;
constexpr double kNewtonTol = 1e-12;

// Per-instance particle parameters (one entry for a plain handle, one per
// instance for a batch). src/pfem2particlehandler.cpp:34-39,47,362-364
// The fields a move reads per particle come first, in the first 64 bytes, so
// that four 16-byte loads issued together fetch them (load_move_prm).
struct alignas(16) Prm {
  double dt_sub;   // dt / (double)n_sub
  int n_sub;
  int qpow2;       // every q_k is a power of two: xi / h_k == xi * q_k exactly
  double h[3];     // 1.0 / q_k (sub-cell part width in reference coordinates)
  double qf[3];    // q_k as double
  int maxpp;       // max_particles_per_cell_part
  int q[3];        // parts per direction
  int Q[3];        // q_k + 1: part buckets per direction (xi_k == 1 gives index q_k)
  int P;           // prod Q_k: part buckets per cell
  int Pvalid;      // prod q_k: refillable parts per cell
  unsigned Qm[3];  // ceil(2^32 / Q_k): l / Q_k == umulhi(l, Qm_k) for 0 <= l < 2^16 (0 for k >= dim)
};
static_assert(sizeof(Prm) == 112 && offsetof(Prm, n_sub) == 8 && offsetof(Prm, qpow2) == 12 &&
                  offsetof(Prm, h) == 16 && offsetof(Prm, qf) == 40,
              "load_move_prm reads the first 64 bytes of a Prm row in this layout");

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
  const double* geom;            // [nc][3D]: X_0, X_{2^D-1}, 1/(X_{2^D-1} - X_0) of box cells; X_0[0] = NaN otherwise
  const int2* cvo;               // [nc][2^D]: (v2c offset, count) of each cell vertex
  const int4* cvc;               // [nc][2^D][2^D / 4]: the (up to 2^D) cells around each cell vertex, in
                                 // v2c order, -1 padded; first id -2 when the vertex has more
  const int* ptab;               // [ninst][pmax]: refillable part -> INT_MIN | pi_0 | pi_1<<8 | pi_2<<16, else 0
  const int* bpos;               // [nc][2^D]: place of (cell, vertex) in the DoF -> (cell, vertex) lists
};

struct DField {
  const double* c[3];
};

struct DParts {
  double* x[3];
  double* v[3];
  int* cell;
  int* src;  // optional: source particle of each slot after a move (-1 - part for refills)
};

struct DTmp {
  double* x[3];
  double* v[3];
  int* cell;
  int* src;  // optional: particle index before the move
};

// Walk queue: particles that left their cell in a substep, with the state
// the neighbour walk needs (the position before the failing substep, the
// left cell, the substep index) and the carried velocity and source index.
// The walk recomputes the failing substep bit for bit from the queued
// position: the left cell's inverse map gives its reference position (the
// cell word's kPlaced flag marks a first substep that started from a snapped
// part centre), the field gives its velocity, and the Euler update its end.
struct DWalk {
  double* x[3];
  double* v[3];
  int* cell;
  int* s;
  int* src;
  int* err;  // device error flags: kErrQueue is set when a stripe would overflow
};

// Device-side error flags (one int per handle, read back by the host).
constexpr int kErrLookback = 1;  // a redistribution look-back wait timed out
constexpr int kErrQueue = 2;     // a walk-queue stripe was too small (particles dropped)
constexpr int kErrRange = 4;     // an adjoint term outside the fixed-point accumulator's range

// Optional record of every particle's path through a move: the cell and the
// reference position at the start of each substep, [s * stride + p] for the
// particle at index p of the move's input (c == nullptr: off).
struct DPath {
  int* c;
  double* xi[3];
  long stride;
};

__device__ __forceinline__ int inst_of(const DMesh& m, int c) { return m.cell_inst ? m.cell_inst[c] : 0; }

// Stored particle cell word: the cell id, plus kPlaced when the particle's
// reference position is the exact centre of its sub-cell part because
// seeding, a refill or an insertion at F(centre) put it there and no locate
// has run on it since (src/pfem2particlehandler.cpp:88-91,390-394: those
// store xi = centre, while every locate stores xi = F^-1(x),
// :420-436,458-478). Cell ids stay below 2^28 (int32 indexing of
// cell * 2^dim), so bit 30 is free; -1 (no cell) keeps every bit set.
constexpr int kPlaced = 1 << 30;
constexpr int kCellMask = kPlaced - 1;

__device__ __forceinline__ int cell_of(int word) { return word & kCellMask; }
__device__ __forceinline__ bool placed_of(int word) { return (word & kPlaced) != 0; }

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

// Newton inverse of a general (non-box) Q1 cell from the affine guess of
// vertex 0 and its edges. Kept out of line: the hot loop only runs the box
// closed form, and inlining this would raise register pressure everywhere.
template <int D>
__device__ bool inv_map_newton(const DMesh& m, const int* vid, const double* x, double* xi);

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
  return inv_map_newton<D>(m, vid, x, xi);
}

template <int D>
__device__ bool inv_map_newton(const DMesh& m, const int* vid, const double* x, double* xi) {
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

// Correctly rounded d / h from the correctly rounded reciprocal r = 1/h:
// q0 = d*r is refined by one FMA residual step (Markstein), which returns
// RN(d/h) for normal operands. Zero, tiny, non-finite or non-normal inputs
// take the IEEE division, so the result always equals d / h bitwise.
__device__ __forceinline__ double div_by(double d, double h, double r) {
  const double q0 = d * r;
  if (!(fabs(d) >= 0x1p-960) || !(fabs(q0) <= 0x1p+960)) return d / h;
  const double rem = __builtin_fma(-q0, h, d);
  return __builtin_fma(rem, r, q0);
}

// part index (int)(xi_k / h_k) with h_k = 1.0/q_k (src/pfem2particlehandler.cpp:370);
// PP is Prm or MovePrm<D>
template <class PP>
__device__ __forceinline__ int part_index(const PP& p, int k, double xi) {
  return p.qpow2 ? (int)(xi * p.qf[k]) : (int)(xi / p.h[k]);
}

// A whole Prm row copied into registers with its 16-byte reads issued
// together (from LDS or global memory), so that code branching on some of
// its fields does not wait for each read in turn.
__device__ __forceinline__ Prm prm_regs(const Prm* __restrict__ p) {
  union Rows {
    Prm prm;
    double2 v[sizeof(Prm) / 16];
  } u;
  const double2* q = reinterpret_cast<const double2*>(p);
#pragma unroll
  for (int i = 0; i < (int)(sizeof(Prm) / 16); ++i) u.v[i] = q[i];
  return u.prm;
}

// The parameters of a move (substep count and length, part widths), held in
// registers: loaded in one round of four 16-byte loads from a Prm row, or
// made wave-uniform (scalar registers) for a handle with one instance.
template <int D>
struct MovePrm {
  double dt_sub;
  int n_sub, qpow2;
  double h[D], qf[D];
};

template <int D>
__device__ __forceinline__ MovePrm<D> load_move_prm(const Prm* __restrict__ p) {
  const double2* q = reinterpret_cast<const double2*>(p);
  const double2 a = q[0], b = q[1], c = q[2], d = q[3];  // bytes 0-63
  MovePrm<D> r;
  r.dt_sub = a.x;
  const long long nq = __double_as_longlong(a.y);
  r.n_sub = (int)(nq & 0xffffffffLL);
  r.qpow2 = (int)(nq >> 32);
  const double hq[6] = {b.x, b.y, c.x, c.y, d.x, d.y};  // h[0..2], qf[0..2]
#pragma unroll
  for (int k = 0; k < D; ++k) {
    r.h[k] = hq[k];
    r.qf[k] = hq[3 + k];
  }
  return r;
}

__device__ __forceinline__ double uniform_double(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readfirstlane((int)(b & 0xffffffffLL));
  const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// The same values in scalar registers (every lane loads the same row).
template <int D>
__device__ __forceinline__ MovePrm<D> uniform_move_prm(const Prm* __restrict__ p) {
  MovePrm<D> r = load_move_prm<D>(p);
  r.dt_sub = uniform_double(r.dt_sub);
  r.n_sub = __builtin_amdgcn_readfirstlane(r.n_sub);
  r.qpow2 = __builtin_amdgcn_readfirstlane(r.qpow2);
#pragma unroll
  for (int k = 0; k < D; ++k) {
    r.h[k] = uniform_double(r.h[k]);
    r.qf[k] = uniform_double(r.qf[k]);
  }
  return r;
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
    const int pk = part_index(p, k, xi[k]);
    ok = ok && (pk >= 0) && (pk < p.q[k]);
    lin = lin * p.Q[k] + pk;
  }
  if (valid) *valid = ok;
  return lin;
}

// Exact centre of the sub-cell part containing xi: the stored reference
// position of a placed particle (kPlaced), whose inverse-mapped position lies
// within a few ulps of that centre, well inside the part. The centre is
// (idx + 1/2) * h_k as seeding and refill compute it
// (src/pfem2particlehandler.cpp:88-91,390-394).
template <int D, class PP>
__device__ __forceinline__ void snap_part(const PP& p, double* xi) {
#pragma unroll
  for (int k = 0; k < D; ++k) xi[k] = ((double)part_index(p, k, xi[k]) + 1.0 / 2) * p.h[k];
}

// Whether x is bitwise the forward map of the centre of the part that holds
// xi = F^-1(x): an insertion there is a placed particle, like a seeded one.
// The forward map is only evaluated when xi is within 1e-9 of the centre
// (x == F(centre) bitwise puts F^-1(x) within a few ulps of it).
template <int D>
__device__ __forceinline__ bool at_part_centre(const DMesh& m, const Prm& p, const int* vid, const double* x,
                                               const double* xi) {
  bool valid = true, near = true;
  double xic[D];
#pragma unroll
  for (int k = 0; k < D; ++k) {
    const int pk = part_index(p, k, xi[k]);
    valid = valid && (pk >= 0) && (pk < p.q[k]);
    xic[k] = ((double)pk + 1.0 / 2) * p.h[k];
    near = near && (fabs(xi[k] - xic[k]) <= 1e-9);
  }
  if (!valid || !near) return false;
  double xc[D];
  fwd<D>(m, vid, xic, xc);
  bool eq = true;
#pragma unroll
  for (int k = 0; k < D; ++k) eq = eq && (xc[k] == x[k]);
  return eq;
}

// Per-lane copy of what the substep loop reads about the particle's current
// cell: the box corners (vertex 0 and vertex 2^D-1; every vertex coordinate
// of a box is one of them) and, in 2D, the advecting field at the cell's four
// DoFs. It is reloaded only when the particle changes cell, so the 2D
// Euler/locate loop runs from registers. In 3D the 24 field values would cost
// more occupancy than the reloads they save, so they are read from the
// cell-major field table (L1/L2) each substep.
template <int D>
struct CellCache {
  int c;
  bool box;
  double X0[D];
  double X1[D];
  double RH[D];  // correctly rounded 1 / (X1 - X0)
  double Vn[D == 2 ? 4 : 1][D];
};

// Vc: advecting field in cell-major order, [nc][2^D][D] (rebuilt per move).
template <int D>
__device__ __forceinline__ void load_cell(const DMesh& m, const double* __restrict__ Vc, int c, CellCache<D>& cc,
                                          bool with_field) {
  cc.c = c;
  const double* g = m.geom + (size_t)c * 3 * D;
#pragma unroll
  for (int k = 0; k < D; ++k) {
    cc.X0[k] = g[k];
    cc.X1[k] = g[D + k];
    cc.RH[k] = g[2 * D + k];
  }
  cc.box = (cc.X0[0] == cc.X0[0]);
  if constexpr (D == 2) {
    if (!with_field) return;
    const double2* f = reinterpret_cast<const double2*>(Vc + (size_t)c * 8);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const double2 t = f[i];
      cc.Vn[i][0] = t.x;
      cc.Vn[i][1] = t.y;
    }
  }
}

template <int D>
__device__ __forceinline__ bool inv_cached(const DMesh& m, const CellCache<D>& cc, const double* x, double* xi) {
  if (cc.box) {
#pragma unroll
    for (int k = 0; k < D; ++k) xi[k] = div_by(x[k] - cc.X0[k], cc.X1[k] - cc.X0[k], cc.RH[k]);
    return true;
  }
  int vid[1 << D];
  load_ids<D>(m.cv, cc.c, vid);
  return inv_map_newton<D>(m, vid, x, xi);
}

// u_k = sum_i N_i(xi) V_k[dof_i] from the cached / cell-major field
template <int D>
__device__ __forceinline__ void gather_cached(const double* __restrict__ Vc, const CellCache<D>& cc, const double* xi,
                                              double* u) {
  double N[1 << D];
  shape<D>(xi, N);
#pragma unroll
  for (int k = 0; k < D; ++k) u[k] = 0.0;
  if constexpr (D == 2) {
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int k = 0; k < 2; ++k) u[k] = u[k] + N[i] * cc.Vn[i][k];
  } else {
    const double* f = Vc + (size_t)cc.c * (8 * D);
#pragma unroll
    for (int i = 0; i < (1 << D); ++i)
#pragma unroll
      for (int k = 0; k < D; ++k) u[k] = u[k] + N[i] * f[i * D + k];
  }
}

// Forward map of a box cell from its corners (bitwise the generic forward map:
// the same vertex coordinates in the same order).
template <int D>
__device__ __forceinline__ void fwd_box(const CellCache<D>& cc, const double* xi, double* x) {
  double N[1 << D];
  shape<D>(xi, N);
#pragma unroll
  for (int k = 0; k < D; ++k) x[k] = 0.0;
#pragma unroll
  for (int i = 0; i < (1 << D); ++i)
#pragma unroll
    for (int k = 0; k < D; ++k) x[k] = x[k] + N[i] * (((i >> k) & 1) ? cc.X1[k] : cc.X0[k]);
}

// Stored reference position of a particle in cell c from (c, x) and its
// placed flag: the inverse map from the geometry table (box cells: no vertex
// loads at all), or the exact part centre for a placed particle.
template <int D>
__device__ __forceinline__ void recover_xi_c(const DMesh& m, const Prm& p, int c, bool placed, const double* x,
                                             double* xi) {
  CellCache<D> cc;
  cc.c = c;
  const double* g = m.geom + (size_t)c * 3 * D;
#pragma unroll
  for (int k = 0; k < D; ++k) {
    cc.X0[k] = g[k];
    cc.X1[k] = g[D + k];
    cc.RH[k] = g[2 * D + k];
  }
  cc.box = (cc.X0[0] == cc.X0[0]);
  inv_cached<D>(m, cc, x, xi);
  if (placed) snap_part<D>(p, xi);
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

// Neighbour walk after a particle left its cell (src/pfem2particlehandler.cpp:
// 438-492 with src/pfem2particle.cpp:186-211): closest vertex of the old cell
// along the substep velocity u, seen from the old position F(xi_old); the
// cells around that vertex in descending alignment of their centre direction
// with the vertex->particle direction (ties in ascending cell id); the first
// one whose inverse map contains x wins. Returns the new cell (xi_new set) or
// -1 when the particle is to be deleted. Out of line: it runs for the ~10% of
// particles that cross a face in a substep.
template <int D>
__device__ int walk(const DMesh& m, int c, const int* vid, const double* xi, const double* x,
                                 const double* u, double* xi_new) {
  constexpr int NV = 1 << D;
  double xn[D];
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
#pragma unroll
        for (int k = 0; k < D; ++k) xi_new[k] = xn[k];
        return cc;
      }
    }
    return -1;
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
#pragma unroll
      for (int k = 0; k < D; ++k) xi_new[k] = xn[k];
      return cc;
    }
  }
  return -1;
}

// Closest vertex of a box cell along u, seen from xold: the reference's
// argmax_v of a_v = sum_k vn_k (d_vk inv_m_v), vn = u / |u|, d_v = X_v - xold,
// inv_m_v = 1 / |d_v|, strict > scan from v = 0
// (src/pfem2particle.cpp:186-211).
// The argmax is first decided from estimates e_v = (u . d_v) rsq(|d_v|^2)
// rsq(|u|^2) with single-precision reciprocal square roots (relative error
// below 1e-6, so |e_v - cos_v| < 3e-6): when the best estimate leads every
// other one by more than kVertexGap the exact cosines, which lie within 1e-14
// of cos_v, have the same unique maximum. Otherwise (near ties, zero or
// non-finite velocity, a vertex at xold, magnitudes outside single
// precision's range) the exact cosines are computed as in the reference.
constexpr double kVertexGap = 1e-4;

template <int D>
__device__ __forceinline__ int closest_vertex_box(const CellCache<D>& cc, const double* xold, const double* u) {
  constexpr int NV = 1 << D;
  double n2 = 0.0;
#pragma unroll
  for (int k = 0; k < D; ++k) n2 = n2 + u[k] * u[k];
  bool sure = (n2 >= 1e-30) && (n2 <= 1e30);
  const double ru = (double)__builtin_amdgcn_rsqf((float)n2);
  double e1 = -4.0, e2 = -4.0;
  int b1 = 0;
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    double dot = 0.0, m2 = 0.0;
#pragma unroll
    for (int k = 0; k < D; ++k) {
      const double d = (((v >> k) & 1) ? cc.X1[k] : cc.X0[k]) - xold[k];
      dot = dot + u[k] * d;
      m2 = m2 + d * d;
    }
    sure = sure && (m2 >= 1e-30) && (m2 <= 1e30);
    const double e = dot * (double)__builtin_amdgcn_rsqf((float)m2) * ru;
    if (e > e1) {
      e2 = e1;
      e1 = e;
      b1 = v;
    } else if (e > e2) {
      e2 = e;
    }
  }
  if (sure && e1 - e2 > kVertexGap) return b1;
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
    for (int k = 0; k < D; ++k) d[k] = (((v >> k) & 1) ? cc.X1[k] : cc.X0[k]) - xold[k];
    double m2 = 0.0;
#pragma unroll
    for (int k = 0; k < D; ++k) m2 = m2 + d[k] * d[k];
    const double inv_m = 1.0 / sqrt(m2);
    double a = vn[0] * (d[0] * inv_m);
#pragma unroll
    for (int k = 1; k < D; ++k) a = a + vn[k] * (d[k] * inv_m);
    if (v == 0) {
      best = a;
    } else if (a > best) {
      best = a;
      bestv = v;
    }
  }
  return bestv;
}

// Reference position of x in candidate cell c2 (box: from the geometry
// table; otherwise Newton); false when the inverse map fails.
template <int D>
__device__ __forceinline__ bool inv_candidate(const DMesh& m, int c2, const double* x, double* xn) {
  const double* g = m.geom + (size_t)c2 * 3 * D;
  if (g[0] == g[0]) {
#pragma unroll
    for (int k = 0; k < D; ++k) xn[k] = div_by(x[k] - g[k], g[D + k] - g[k], g[2 * D + k]);
    return true;
  }
  int vid2[1 << D];
  load_ids<D>(m.cv, c2, vid2);
  return inv_map_newton<D>(m, vid2, x, xn);
}

// Neighbour walk from a box cell: the same result as walk(), with the current
// cell's vertex coordinates taken from its corners and the vertex's candidate
// list from the per-cell table. The candidates (the cells around the closest
// vertex except the current one) are tested in list order first; the
// reference's order (descending alignment of the centre direction with the
// vertex->particle direction, ties by cell id) only decides between two or
// more containing candidates, so it is computed only then.
template <int D>
__device__ int walk_box(const DMesh& m, const CellCache<D>& cc, const double* xi, const double* x, const double* u,
                        double* xi_new) {
  double xold[D];
  fwd_box<D>(cc, xi, xold);
  const int bestv = closest_vertex_box<D>(cc, xold, u);
  const int2 oc = m.cvo[(size_t)cc.c * (1 << D) + bestv];
  // 3D: the list itself in the same round of loads as its extent (the
  // per-(cell, vertex) copy; read before the extent is tested, so that the
  // two loads are waited for together)
  int ids[8];
  if constexpr (D == 3) {
    const int4* cl = m.cvc + ((size_t)cc.c * 8 + bestv) * 2;
    const int4 l0 = cl[0], l1 = cl[1];
    ids[0] = l0.x; ids[1] = l0.y; ids[2] = l0.z; ids[3] = l0.w;
    ids[4] = l1.x; ids[5] = l1.y; ids[6] = l1.z; ids[7] = l1.w;
  }
  const int o0 = oc.x, nn = oc.y;
  if (nn > 8 || (D == 3 && ids[0] == -2)) return -2;  // caller falls back to the generic walk
  unsigned hits = 0;
  int first = -1, cfirst = -1;
  double xf[D];
  auto test = [&](int j, int c2) {
    if (c2 == cc.c) return;  // the current cell already failed the same test
    double xn[D];
    if (inv_candidate<D>(m, c2, x, xn) && inside<D>(xn)) {
      hits |= 1u << j;
      if (first < 0) {
        first = j;
        cfirst = c2;
#pragma unroll
        for (int k = 0; k < D; ++k) xf[k] = xn[k];
      }
    }
  };
  if constexpr (D == 3) {
    // the candidates in list order, the corners of kCandRound of them read
    // in one round (an absent candidate, or the current cell, reads the
    // current cell's row and is skipped)
    constexpr int kCandRound = 2;
#pragma unroll
    for (int j0 = 0; j0 < 8; j0 += kCandRound) {
      double g[kCandRound][3 * D];
#pragma unroll
      for (int e = 0; e < kCandRound; ++e) {
        const int j = j0 + e;
        const int c2 = (j < nn) ? ids[j] : cc.c;
        const double* gp = m.geom + (size_t)c2 * 3 * D;
#pragma unroll
        for (int q = 0; q < 3 * D; ++q) g[e][q] = gp[q];
      }
#pragma unroll
      for (int e = 0; e < kCandRound; ++e) {
        const int j = j0 + e;
        if (j >= nn || ids[j] == cc.c) continue;  // the current cell already failed the same test
        double xn[D];
        bool ok;
        if (g[e][0] == g[e][0]) {
#pragma unroll
          for (int k = 0; k < D; ++k) xn[k] = div_by(x[k] - g[e][k], g[e][D + k] - g[e][k], g[e][2 * D + k]);
          ok = true;
        } else {
          ok = inv_candidate<D>(m, ids[j], x, xn);
        }
        if (ok && inside<D>(xn)) {
          hits |= 1u << j;
          if (first < 0) {
            first = j;
            cfirst = ids[j];
#pragma unroll
            for (int k = 0; k < D; ++k) xf[k] = xn[k];
          }
        }
      }
    }
  } else {
    for (int j = 0; j < nn; ++j) test(j, m.v2c[o0 + j]);
  }
  if (hits == 0) return -1;
  if ((hits & (hits - 1)) == 0) {
#pragma unroll
    for (int k = 0; k < D; ++k) xi_new[k] = xf[k];
    return cfirst;
  }
  // two or more containing candidates: the first in the reference's order
  double vtp[D];
#pragma unroll
  for (int k = 0; k < D; ++k) vtp[k] = x[k] - (((bestv >> k) & 1) ? cc.X1[k] : cc.X0[k]);
  double t2 = 0.0;
#pragma unroll
  for (int k = 0; k < D; ++k) t2 = t2 + vtp[k] * vtp[k];
  const double inv_t = 1.0 / sqrt(t2);
#pragma unroll
  for (int k = 0; k < D; ++k) vtp[k] = vtp[k] * inv_t;
  double dots[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    dots[j] = 0.0;
    if (j < nn) dots[j] = dot_dir<D>(m.v2c_dir + (size_t)(o0 + j) * D, vtp);
  }
  unsigned used = 0;
  for (int r = 0; r < nn; ++r) {
    int bj = -1;
    double bd = 0.0;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      if (j < nn && !((used >> j) & 1u)) {
        if (bj < 0 || dots[j] > bd) {
          bj = j;
          bd = dots[j];
        }
      }
    }
    used |= 1u << bj;
    if ((hits >> bj) & 1u) {
      const int c2 = m.v2c[o0 + bj];
      inv_candidate<D>(m, c2, x, xi_new);
      return c2;
    }
  }
  return -1;
}

// 2D neighbour walk from a box cell with the candidate lists of the cell's
// four vertices already loaded (r0..r3, from the per-cell table): the same
// result as walk_box. The candidates' geometry is read in one round of
// loads; the winner's corners are returned in gw (X0, X1, RH) so that the
// caller does not read them again. Returns -2 when the closest vertex has
// more than four cells (the caller then walks through walk_cached).
__device__ __forceinline__ int sel4(int b, int a0, int a1, int a2, int a3) {
  return b == 3 ? a3 : (b == 2 ? a2 : (b == 1 ? a1 : a0));
}

__device__ __forceinline__ int walk_box2(const DMesh& m, const CellCache<2>& cc, const double* xi, const double* x,
                                         const double* u, int4 r0, int4 r1, int4 r2, int4 r3, double* xi_new,
                                         double* gw) {
  double xold[2];
  fwd_box<2>(cc, xi, xold);
  const int bestv = closest_vertex_box<2>(cc, xold, u);
  const int ids[4] = {sel4(bestv, r0.x, r1.x, r2.x, r3.x), sel4(bestv, r0.y, r1.y, r2.y, r3.y),
                      sel4(bestv, r0.z, r1.z, r2.z, r3.z), sel4(bestv, r0.w, r1.w, r2.w, r3.w)};
  if (ids[0] == -2) return -2;
  unsigned hits = 0;
  int first = -1;
  double xf[2] = {0.0, 0.0};
  double gf[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};  // corners of the first containing candidate
  // The current cell is one of the four (it holds the vertex) and already
  // failed the same test: the other three candidates' corners are read in one
  // round of loads (an absent candidate reads the current cell's row and is
  // skipped), slot k of the three being list position k + (k >= self).
  int self = -1;
#pragma unroll
  for (int j = 0; j < 4; ++j)
    if (ids[j] == cc.c) self = j;
  if (self < 0) return -2;  // (not a list of this cell's vertex: the generic walk decides)
  int cid[3];
  double gc[3][6];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    cid[k] = k >= self ? ids[k + 1] : ids[k];
    const int id = cid[k] < 0 ? cc.c : cid[k];
    const double2* gp = reinterpret_cast<const double2*>(m.geom + (size_t)id * 6);
    const double2 a = gp[0], b = gp[1], c = gp[2];
    gc[k][0] = a.x;
    gc[k][1] = a.y;
    gc[k][2] = b.x;
    gc[k][3] = b.y;
    gc[k][4] = c.x;
    gc[k][5] = c.y;
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    if (cid[k] < 0) continue;
    const int j = k >= self ? k + 1 : k;
    const double* g = gc[k];
    double xn[2];
    bool ok;
    if (g[0] == g[0]) {
#pragma unroll
      for (int q = 0; q < 2; ++q) xn[q] = div_by(x[q] - g[q], g[2 + q] - g[q], g[4 + q]);
      ok = true;
    } else {
      int vid2[4];
      load_ids<2>(m.cv, cid[k], vid2);
      ok = inv_map_newton<2>(m, vid2, x, xn);
    }
    if (ok && inside<2>(xn)) {
      hits |= 1u << j;
      if (first < 0) {
        first = j;
        xf[0] = xn[0];
        xf[1] = xn[1];
#pragma unroll
        for (int q = 0; q < 6; ++q) gf[q] = g[q];
      }
    }
  }
  if (hits == 0) return -1;
  int win = first;  // a single containing candidate wins in any order
  if ((hits & (hits - 1)) != 0) {
    // two or more containing candidates: the first in the reference's order
    const int o0 = m.cvo[(size_t)cc.c * 4 + bestv].x;
    const int nn = m.cvo[(size_t)cc.c * 4 + bestv].y;
    double vtp[2];
#pragma unroll
    for (int k = 0; k < 2; ++k) vtp[k] = x[k] - (((bestv >> k) & 1) ? cc.X1[k] : cc.X0[k]);
    const double t2 = vtp[0] * vtp[0] + vtp[1] * vtp[1];
    const double inv_t = 1.0 / sqrt(t2);
#pragma unroll
    for (int k = 0; k < 2; ++k) vtp[k] = vtp[k] * inv_t;
    double dots[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) dots[j] = j < nn ? dot_dir<2>(m.v2c_dir + (size_t)(o0 + j) * 2, vtp) : 0.0;
    unsigned used = 0;
    win = -1;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      if (r >= nn || win >= 0) break;
      int bj = -1;
      double bd = 0.0;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        if (j < nn && !((used >> j) & 1u)) {
          if (bj < 0 || dots[j] > bd) {
            bj = j;
            bd = dots[j];
          }
        }
      }
      used |= 1u << bj;
      if ((hits >> bj) & 1u) win = bj;
    }
  }
  int wid = ids[0];
#pragma unroll
  for (int j = 1; j < 4; ++j)
    if (win == j) wid = ids[j];
  if (win != first) {  // the reference's order picked another containing candidate
    inv_candidate<2>(m, wid, x, xf);
    const double* gp = m.geom + (size_t)wid * 6;
#pragma unroll
    for (int q = 0; q < 6; ++q) gf[q] = gp[q];
  }
#pragma unroll
  for (int k = 0; k < 2; ++k) xi_new[k] = xf[k];
#pragma unroll
  for (int q = 0; q < 6; ++q) gw[q] = gf[q];
  return wid;
}

// Neighbour walk for the particle in cache cc (box fast path or generic).
template <int D>
__device__ __forceinline__ int walk_cached(const DMesh& m, const CellCache<D>& cc, const double* xi, const double* x,
                                           const double* u, double* xi_new) {
  if (cc.box) {
    const int r = walk_box<D>(m, cc, xi, x, u, xi_new);
    if (r != -2) return r;
  }
  int vid[1 << D];
  load_ids<D>(m.cv, cc.c, vid);
  return walk<D>(m, cc.c, vid, xi, x, u, xi_new);
}

}  // namespace pfem2gpu
