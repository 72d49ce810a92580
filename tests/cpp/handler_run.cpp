// Runs the particle path through the C++ handler class (include/
// pfem2gpu_handler.hpp) on the case in an input file and writes the final
// particles and the last projected field to an output file; the GPU test
// compares them with the same sequence run through the Python binding.
//
//   handler_run <input> <output>
//
// Input (little-endian): int32 dim, int32 steps, int32 order; int64
// n_vertices, n_cells, n_dofs, n_particles; uint32 n_sub; double dt; uint32
// max_per_part; uint32 q[3]; double X[dim][n_vertices]; int32
// cell_vertices[n_cells][2^dim]; int32 cell[n_particles]; double
// x[dim][n_particles]; double V[dim][n_dofs].
// order 0: initialize_maps, insert_particle for each particle (zero
// velocity), then per step move_particles(V), P = project_particle_fields(),
// correct_particle_velocity(V, P).
// order 1, the reference solver's order (src/pfem2solver.cpp:151-166) with
// the FEM step replaced by a prescribed field: seed_particles(V) before
// initialize_maps(), V_old = 0 (the FEM's initial old solution,
// src/pfem2fem.cpp:586-594), then per step correct_particle_velocity(V,
// V_old), move_particles(V), P = project_particle_fields(), V_old = P; the
// reference's four timer sections are printed at the end.
// Output: int64 n, rejected; double x[dim][n], xi[dim][n], v[dim][n];
// int32 cell[n]; double P[dim][n_dofs].
#include <cstdio>
#include <cstring>
#include <exception>
#include <vector>

#include "pfem2gpu_handler.hpp"

namespace {

struct Reader {
  FILE* f;
  template <typename T>
  T get() {
    T v;
    if (std::fread(&v, sizeof(T), 1, f) != 1) throw std::runtime_error("input file too short");
    return v;
  }
  template <typename T>
  std::vector<T> vec(std::size_t n) {
    std::vector<T> v(n);
    if (n && std::fread(v.data(), sizeof(T), n, f) != n) throw std::runtime_error("input file too short");
    return v;
  }
};

template <typename T>
void put(FILE* f, const T* p, std::size_t n) {
  if (n && std::fwrite(p, sizeof(T), n, f) != n) throw std::runtime_error("cannot write output");
}

template <int dim>
int run(Reader& in, int steps, int order, FILE* out) {
  const int64_t nv = in.get<int64_t>(), nc = in.get<int64_t>(), nd = in.get<int64_t>(), np = in.get<int64_t>();
  pfem2gpu::Parameters prm;
  prm.particle_integration_steps = in.get<uint32_t>();
  prm.time_step = in.get<double>();
  prm.max_particles_per_cell_part = in.get<uint32_t>();
  for (int k = 0; k < 3; ++k) prm.particles_per_cell[k] = in.get<uint32_t>();
  pfem2gpu::MeshData mesh;
  mesh.dim = dim;
  mesh.vertices = in.vec<double>((std::size_t)dim * nv);
  mesh.cell_vertices = in.vec<int32_t>((std::size_t)nc << dim);
  mesh.n_dofs = nd;
  const std::vector<int32_t> cell = in.vec<int32_t>((std::size_t)np);
  const std::vector<double> x = in.vec<double>((std::size_t)dim * np);
  pfem2gpu::NodalField<dim> V;
  for (int k = 0; k < dim; ++k) V[k] = in.vec<double>((std::size_t)nd);

  pfem2gpu::pfem2ParticleHandler<dim> h(mesh, prm);
  pfem2gpu::NodalField<dim> P;
  if (order == 0) {
    h.initialize_maps();
    for (int64_t p = 0; p < np; ++p) {
      std::array<double, dim> loc;
      for (int k = 0; k < dim; ++k) loc[k] = x[k * np + p];
      h.insert_particle(loc, cell[p]);
    }
    for (int s = 0; s < steps; ++s) {
      h.move_particles(V);
      P = h.project_particle_fields();
      h.correct_particle_velocity(V, P);
    }
  } else {
    h.enable_timer(true);
    h.seed_particles(V);
    h.initialize_maps();
    pfem2gpu::NodalField<dim> V_old;
    for (int k = 0; k < dim; ++k) V_old[k].assign((std::size_t)nd, 0.0);
    for (int s = 0; s < steps; ++s) {
      h.correct_particle_velocity(V, V_old);
      h.move_particles(V);
      P = h.project_particle_fields();
      V_old = P;
    }
    h.print_timer_summary(stdout);
  }
  const std::vector<pfem2gpu::Particle<dim>> parts = h.particles();
  const int64_t n = (int64_t)parts.size(), rej = h.rejected_insertions();
  put(out, &n, 1);
  put(out, &rej, 1);
  std::vector<double> buf((std::size_t)n);
  for (int what = 0; what < 3; ++what)
    for (int k = 0; k < dim; ++k) {
      for (int64_t p = 0; p < n; ++p)
        buf[p] = what == 0 ? parts[p].location[k] : what == 1 ? parts[p].reference_location[k] : parts[p].velocity[k];
      put(out, buf.data(), (std::size_t)n);
    }
  std::vector<int32_t> c((std::size_t)n);
  for (int64_t p = 0; p < n; ++p) c[p] = parts[p].cell;
  put(out, c.data(), (std::size_t)n);
  for (int k = 0; k < dim; ++k) put(out, P[k].data(), P[k].size());
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 3) {
    std::fprintf(stderr, "usage: %s <input> <output>\n", argv[0]);
    return 2;
  }
  FILE* fi = std::fopen(argv[1], "rb");
  if (!fi) {
    std::fprintf(stderr, "cannot open %s\n", argv[1]);
    return 2;
  }
  FILE* fo = std::fopen(argv[2], "wb");
  if (!fo) {
    std::fclose(fi);
    std::fprintf(stderr, "cannot open %s\n", argv[2]);
    return 2;
  }
  int rc = 1;
  try {
    Reader in{fi};
    const int dim = in.get<int32_t>(), steps = in.get<int32_t>(), order = in.get<int32_t>();
    if (dim == 2) rc = run<2>(in, steps, order, fo);
    else if (dim == 3) rc = run<3>(in, steps, order, fo);
    else throw std::runtime_error("dim must be 2 or 3");
  } catch (const std::exception& e) {
    std::fprintf(stderr, "handler_run: %s\n", e.what());
    rc = 1;
  }
  std::fclose(fi);
  if (std::fclose(fo) != 0) rc = 1;
  return rc;
}
