// TEST INFRASTRUCTURE ONLY — runs the oracle's entry points under the
// address and undefined-behaviour sanitizers (built by `make -C oracle
// sanitize`, g++ with -fsanitize=address,undefined; no GPU code).
//
// Both modes of the particle oracle on 2D and 3D box meshes and a perturbed
// 2D mesh: maps, lattice seeding, insertion with rejects, correction, moves
// through an inflow/outflow field, projection with weights, removal and
// download; then the three FEM assemblies with face flags. Exit status 0
// when every call returned and the checked invariants hold.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

extern "C" {
void* orc_create(int dim, int64_t nv, const double* X, int64_t nc, const int* cell_vertices, const int* cell_dofs,
                 int64_t ndofs, unsigned n_sub, double dt, unsigned max_per_part, const unsigned* q, int mode);
void orc_destroy(void* h);
void orc_initialize_maps(void* h);
int64_t orc_insert(void* h, int64_t n, const double* x, const int* cell, const double* v);
void orc_seed_lattice(void* h, const double* V);
void orc_correct(void* h, const double* V, const double* Vold);
int orc_move(void* h, const double* V);
void orc_project_weights(void* h, double* Vout, double* W);
void orc_remove(void* h, int64_t n, const int64_t* index);
int64_t orc_count(void* h);
void orc_download(void* h, double* x, double* xi, double* v, int* cell);
int64_t orc_fem_pattern(int dim, int64_t nc, int64_t ndofs, const int* cd, int64_t* rowptr, int32_t* cols);
int orc_fem_pressure(int dim, int64_t nverts, const double* X, int64_t nc, const int* cv, const int* cd, int64_t ndofs,
                     const double* V, const unsigned char* faces, double rho, double dt, const int64_t* rowptr,
                     const int32_t* cols, double* vals, double* rhs);
int orc_fem_vcorr(int dim, int64_t nverts, const double* X, int64_t nc, const int* cv, const int* cd, int64_t ndofs,
                  const double* V, const double* P, double rho, double dt, const int64_t* rowptr,
                  const int32_t* cols, double* vals, double* rhs);
int orc_fem_prediction(int dim, int64_t nverts, const double* X, int64_t nc, const int* cv, const int* cd,
                       int64_t ndofs, const double* Vo, const double* U, const unsigned char* faces, double rho,
                       double mu, double time_step, const int64_t* rowptr, const int32_t* cols, double* vals,
                       double* rhs);
}

namespace {

int g_fail = 0;

void check(bool ok, const char* what) {
  if (!ok) {
    std::fprintf(stderr, "FAILED: %s\n", what);
    g_fail = 1;
  }
}

struct Mesh {
  int dim;
  int n[3];
  std::vector<double> X;  // [dim][nv]
  std::vector<int> cv;    // [nc][2^dim]
  int64_t nv = 0, nc = 0;
  std::vector<unsigned char> faces;  // every boundary face flagged
};

Mesh box(int dim, const int* n, double perturb) {
  Mesh m;
  m.dim = dim;
  int nvx[3] = {1, 1, 1}, ncx[3] = {1, 1, 1};
  for (int k = 0; k < dim; ++k) {
    m.n[k] = n[k];
    nvx[k] = n[k] + 1;
    ncx[k] = n[k];
  }
  m.nv = (int64_t)nvx[0] * nvx[1] * nvx[2];
  m.nc = (int64_t)ncx[0] * ncx[1] * ncx[2];
  m.X.assign((size_t)dim * m.nv, 0.0);
  uint64_t s = 12345;
  for (int64_t v = 0; v < m.nv; ++v) {
    int64_t r = v;
    for (int k = 0; k < dim; ++k) {
      const int i = (int)(r % nvx[k]);
      r /= nvx[k];
      double x = (double)i / n[k];
      if (perturb > 0 && i > 0 && i < n[k]) {
        s = s * 6364136223846793005ULL + 1442695040888963407ULL;
        x += perturb * ((double)(s >> 11) / 9007199254740992.0 - 0.5) / n[k];
      }
      m.X[(size_t)k * m.nv + v] = x;
    }
  }
  const int nvc = 1 << dim;
  m.cv.resize((size_t)m.nc * nvc);
  m.faces.assign(m.nc, 0);
  for (int64_t c = 0; c < m.nc; ++c) {
    int64_t r = c;
    int ci[3] = {0, 0, 0};
    for (int k = 0; k < dim; ++k) {
      ci[k] = (int)(r % ncx[k]);
      r /= ncx[k];
      if (ci[k] == 0) m.faces[c] |= (unsigned char)(1 << (2 * k));
      if (ci[k] == n[k] - 1) m.faces[c] |= (unsigned char)(1 << (2 * k + 1));
    }
    for (int b = 0; b < nvc; ++b) {
      int64_t vid = 0, stride = 1;
      for (int k = 0; k < dim; ++k) {
        vid += (ci[k] + ((b >> k) & 1)) * stride;
        stride *= nvx[k];
      }
      m.cv[(size_t)c * nvc + b] = (int)vid;
    }
  }
  return m;
}

// a field with inflow and outflow: particles enter and leave the box
std::vector<double> field(const Mesh& m, double a) {
  std::vector<double> V((size_t)m.dim * m.nv);
  for (int64_t v = 0; v < m.nv; ++v)
    for (int k = 0; k < m.dim; ++k) {
      const double x = m.X[(size_t)v];
      const double y = m.X[(size_t)m.nv + v];
      V[(size_t)k * m.nv + v] = k == 0 ? a * (0.5 + y) : (k == 1 ? 0.3 * a * std::sin(3.0 * x) : 0.1 * a);
    }
  return V;
}

void particles(const Mesh& m, int mode) {
  const unsigned q[3] = {2, 2, 2};
  void* h = orc_create(m.dim, m.nv, m.X.data(), m.nc, m.cv.data(), m.cv.data(), m.nv, 3, 0.05, 2, q, mode);
  check(h != nullptr, "create");
  orc_initialize_maps(h);
  const std::vector<double> V = field(m, 1.0), Vold = field(m, 0.9);
  orc_seed_lattice(h, V.data());
  const int64_t n0 = orc_count(h);
  check(n0 == m.nc * (1 << m.dim), "lattice seeds one particle per part");
  // insertion: every third particle outside its cell (rejected)
  const int nvc = 1 << m.dim;
  std::vector<double> x((size_t)m.dim * 30), v((size_t)m.dim * 30, 0.25);
  std::vector<int> cell(30);
  for (int i = 0; i < 30; ++i) {
    cell[i] = (int)((i * 7) % m.nc);
    for (int k = 0; k < m.dim; ++k) {
      double lo = 1e9, hi = -1e9;
      for (int b = 0; b < nvc; ++b) {
        const double xv = m.X[(size_t)k * m.nv + m.cv[(size_t)cell[i] * nvc + b]];
        lo = std::fmin(lo, xv);
        hi = std::fmax(hi, xv);
      }
      x[(size_t)k * 30 + i] = (i % 3 == 0) ? hi + 1.0 : lo + 0.5 * (hi - lo);
    }
  }
  const int64_t rej = orc_insert(h, 30, x.data(), cell.data(), v.data());
  check(rej >= 10 && rej <= 30, "insertion rejects the outside particles");
  for (int s = 0; s < 4; ++s) {
    orc_correct(h, V.data(), Vold.data());
    check(orc_move(h, V.data()) >= 0, "move");
    std::vector<double> P((size_t)m.dim * m.nv), W(m.nv);
    orc_project_weights(h, P.data(), W.data());
    for (int64_t j = 0; j < m.nv; ++j)
      if (W[j] > 0)
        for (int k = 0; k < m.dim; ++k) check(std::isfinite(P[(size_t)k * m.nv + j]), "finite projection");
  }
  const int64_t n = orc_count(h);
  check(n > 0, "particles remain");
  std::vector<int64_t> rm = {0, n / 2, n - 1, n / 2};
  orc_remove(h, (int64_t)rm.size(), rm.data());
  check(orc_count(h) == n - 3, "removal of three listed particles");
  const int64_t n2 = orc_count(h);
  std::vector<double> dx((size_t)m.dim * n2), dxi((size_t)m.dim * n2), dv((size_t)m.dim * n2);
  std::vector<int> dc(n2);
  orc_download(h, dx.data(), dxi.data(), dv.data(), dc.data());
  for (int64_t i = 0; i < n2; ++i) check(dc[i] >= 0 && dc[i] < m.nc, "downloaded cells in range");
  orc_destroy(h);
}

void fem(const Mesh& m) {
  std::vector<int64_t> rp(m.nv + 1);
  const int64_t nnz = orc_fem_pattern(m.dim, m.nc, m.nv, m.cv.data(), rp.data(), nullptr);
  std::vector<int32_t> cols(nnz);
  orc_fem_pattern(m.dim, m.nc, m.nv, m.cv.data(), rp.data(), cols.data());
  const std::vector<double> V = field(m, 1.0), U = field(m, 0.7);
  std::vector<double> P(m.nv, 0.0);
  for (int64_t v = 0; v < m.nv; ++v) P[v] = m.X[(size_t)v];
  std::vector<double> vals((size_t)m.dim * nnz), rhs((size_t)m.dim * m.nv);
  check(orc_fem_pressure(m.dim, m.nv, m.X.data(), m.nc, m.cv.data(), m.cv.data(), m.nv, V.data(), m.faces.data(), 1.2,
                         0.01, rp.data(), cols.data(), vals.data(), rhs.data()) == 0,
        "pressure assembly");
  check(orc_fem_vcorr(m.dim, m.nv, m.X.data(), m.nc, m.cv.data(), m.cv.data(), m.nv, V.data(), P.data(), 1.2, 0.01,
                      rp.data(), cols.data(), vals.data(), rhs.data()) == 0,
        "velocity correction assembly");
  check(orc_fem_prediction(m.dim, m.nv, m.X.data(), m.nc, m.cv.data(), m.cv.data(), m.nv, U.data(), V.data(),
                           m.faces.data(), 1.2, 0.3, 0.01, rp.data(), cols.data(), vals.data(), rhs.data()) == 0,
        "velocity prediction assembly");
  for (double a : vals) check(std::isfinite(a), "finite matrix entries");
  for (double a : rhs) check(std::isfinite(a), "finite right-hand sides");
}

}  // namespace

int main() {
  const int n2[3] = {12, 9, 1}, n3[3] = {5, 4, 3};
  const Mesh b2 = box(2, n2, 0.0), p2 = box(2, n2, 0.4), b3 = box(3, n3, 0.0);
  for (int mode = 0; mode < 2; ++mode) {
    particles(b2, mode);
    particles(p2, mode);
    particles(b3, mode);
  }
  fem(b2);
  fem(p2);
  fem(b3);
  std::printf(g_fail ? "sanitized oracle run: FAILED\n" : "sanitized oracle run: ok\n");
  return g_fail;
}
