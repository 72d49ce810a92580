// pfem2ParticleHandler<dim> over the pfem2gpu C ABI: a header-only C++ class
// with the reference handler's methods (src/pfem2particlehandler.h:18-71),
// for C++ hosts that call the particle path without deal.II types.
//
// The reference's handler reads its inputs from the solver it is attached to
// (setPfem2Solver, src/pfem2particlehandler.h:35): the mesh and DoF handler,
// the four particle parameters (src/pfem2particlehandler.cpp:34-39,47) and
// the nodal fields getSolutionV / getOldSolutionV (src/pfem2fem.cpp:622-632),
// and writes the projected field back through setSolutionV (:640-645). Here
// the mesh and parameters are given at construction and the nodal fields are
// passed to the calls that use them, as std::array<std::vector<double>, dim>
// indexed by DoF (the reference's std::array<TrilinosWrappers::MPI::Vector,
// dim>), or as `dim` raw component pointers (host memory by default, device
// memory with where = PFEM2GPU_DEVICE), which a deal.II host takes from its
// Trilinos vectors without a copy. Failures throw std::runtime_error with the
// library's message (the reference throws deal.II exceptions; it has no
// status codes).
//
// The reference times its four calls in TimerOutput sections named
// "Particle seeding", "Particle velocity correction", "Particle transport" and
// "Particle fields projection" (src/pfem2particlehandler.cpp:45,114,144,176)
// and prints the summary every time step (src/pfem2solver.cpp:173). With
// enable_timer(true) the library brackets the same calls with HIP events
// under the same names; timer_sections() and print_timer_summary() report
// them.
#ifndef PFEM2GPU_HANDLER_HPP
#define PFEM2GPU_HANDLER_HPP

#include <array>
#include <cstdint>
#include <cstdio>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "pfem2gpu.h"

namespace pfem2gpu {

// Q1 mesh: vertex coordinates as [dim][n_vertices], cell -> vertex and
// cell -> DoF tables as [n_cells][2^dim] in GeometryInfo vertex order (an
// empty cell_dofs means DoF id = vertex id).
struct MeshData {
  int dim = 2;
  std::vector<double> vertices;
  std::vector<int32_t> cell_vertices;
  std::vector<int32_t> cell_dofs;
  int64_t n_dofs = 0;
};

// pfem2ParameterHandler's particle parameters (src/pfem2parameterhandler.cpp).
struct Parameters {
  unsigned particle_integration_steps = 1;
  double time_step = 0.0;
  unsigned max_particles_per_cell_part = 2;
  std::array<unsigned, 3> particles_per_cell{{1, 1, 1}};
};

template <int dim>
using NodalField = std::array<std::vector<double>, dim>;

// The reference's TimerOutput section names of the handler's calls.
inline const std::array<const char*, 4>& timer_section_names() {
  static const std::array<const char*, 4> names{{"Particle seeding", "Particle velocity correction",
                                                  "Particle transport", "Particle fields projection"}};
  return names;
}

struct TimerSection {
  int64_t calls = 0;
  double total_ms = 0.0;
};

template <int dim>
struct Particle {
  std::array<double, dim> location;
  std::array<double, dim> reference_location;
  std::array<double, dim> velocity;
  int cell;
};

template <int dim>
class pfem2ParticleHandler {
  static_assert(dim == 2 || dim == 3, "dim must be 2 or 3");

 public:
  // pfem2ParticleHandler(const FE_Q<dim>*) + setPfem2Solver
  // (src/pfem2particlehandler.cpp:10-15,29-40).
  pfem2ParticleHandler(const MeshData& mesh, const Parameters& prm) : mesh_(mesh) {
    if (mesh.dim != dim) throw std::runtime_error("pfem2ParticleHandler: mesh dimension differs from dim");
    const int64_t nvc = 1 << dim;
    pfem2gpu_mesh m{};
    m.dim = dim;
    m.n_vertices = (int64_t)mesh.vertices.size() / dim;
    m.vertices = mesh.vertices.data();
    m.n_cells = (int64_t)mesh.cell_vertices.size() / nvc;
    m.cell_vertices = mesh.cell_vertices.data();
    m.cell_dofs = mesh.cell_dofs.empty() ? nullptr : mesh.cell_dofs.data();
    m.n_dofs = mesh.n_dofs;
    pfem2gpu_params p{};
    p.particle_integration_steps = prm.particle_integration_steps;
    p.time_step = prm.time_step;
    p.max_particles_per_cell_part = prm.max_particles_per_cell_part;
    for (int k = 0; k < 3; ++k) p.particles_per_cell[k] = k < dim ? prm.particles_per_cell[k] : 1;
    check(pfem2gpu_create(&m, &p, nullptr, &h_));
    n_dofs_ = mesh.n_dofs;
  }
  ~pfem2ParticleHandler() { pfem2gpu_destroy(h_); }
  pfem2ParticleHandler(const pfem2ParticleHandler&) = delete;
  pfem2ParticleHandler& operator=(const pfem2ParticleHandler&) = delete;

  // initialize_maps() (src/pfem2particlehandler.cpp:319-324)
  void initialize_maps() { check(pfem2gpu_initialize_maps(h_)); }

  // seed_particles() (src/pfem2particlehandler.cpp:42-109). It needs only
  // the mesh, so it may come before initialize_maps() as in
  // src/pfem2solver.cpp:154-155.
  void seed_particles(const NodalField<dim>& V) {
    flush_inserts();
    check(pfem2gpu_seed_particles(h_, ptrs(V), PFEM2GPU_HOST));
  }
  void seed_particles(const std::array<const double*, dim>& V, int where = PFEM2GPU_HOST) {
    flush_inserts();
    check(pfem2gpu_seed_particles(h_, raw(V), where));
  }

  // insert_particle() (src/pfem2particlehandler.cpp:338-344): buffered and
  // handed over with the next call that needs the particles.
  void insert_particle(const std::array<double, dim>& location, int cell,
                       const std::array<double, dim>& velocity = {}) {
    pending_cells_.push_back(cell);
    for (int k = 0; k < dim; ++k) {
      pending_x_[k].push_back(location[k]);
      pending_v_[k].push_back(velocity[k]);
    }
  }

  // remove_particle() (src/pfem2particlehandler.cpp:332-336): the particle
  // at position `index` of particles(); the others keep their order.
  void remove_particle(std::size_t index) { remove_particles(std::vector<int64_t>{(int64_t)index}); }
  void remove_particles(const std::vector<int64_t>& index) {
    flush_inserts();
    check(pfem2gpu_remove_particles(h_, (int64_t)index.size(), index.data()));
  }

  // clear_particles() (src/pfem2particlehandler.cpp:326-330)
  void clear_particles() {
    pending_cells_.clear();
    for (int k = 0; k < dim; ++k) {
      pending_x_[k].clear();
      pending_v_[k].clear();
    }
    check(pfem2gpu_clear_particles(h_));
  }

  // correct_particle_velocity() (src/pfem2particlehandler.cpp:111-139)
  void correct_particle_velocity(const NodalField<dim>& V, const NodalField<dim>& V_old) {
    flush_inserts();
    const double* a[3] = {nullptr, nullptr, nullptr};
    const double* b[3] = {nullptr, nullptr, nullptr};
    for (int k = 0; k < dim; ++k) {
      a[k] = V[k].data();
      b[k] = V_old[k].data();
    }
    check_sizes(V);
    check_sizes(V_old);
    check(pfem2gpu_correct_particle_velocity(h_, a, b, PFEM2GPU_HOST));
  }
  void correct_particle_velocity(const std::array<const double*, dim>& V,
                                 const std::array<const double*, dim>& V_old, int where = PFEM2GPU_HOST) {
    flush_inserts();
    const double* a[3] = {nullptr, nullptr, nullptr};
    const double* b[3] = {nullptr, nullptr, nullptr};
    for (int k = 0; k < dim; ++k) {
      a[k] = V[k];
      b[k] = V_old[k];
    }
    check(pfem2gpu_correct_particle_velocity(h_, a, b, where));
  }

  // move_particles() (src/pfem2particlehandler.cpp:141-171)
  void move_particles(const NodalField<dim>& V) {
    flush_inserts();
    check(pfem2gpu_move_particles(h_, ptrs(V), PFEM2GPU_HOST));
  }
  void move_particles(const std::array<const double*, dim>& V, int where = PFEM2GPU_HOST) {
    flush_inserts();
    check(pfem2gpu_move_particles(h_, raw(V), where));
  }

  // project_particle_fields() (src/pfem2particlehandler.cpp:173-215): the
  // projected field is returned instead of written with setSolutionV.
  NodalField<dim> project_particle_fields() {
    flush_inserts();
    NodalField<dim> out;
    double* o[3] = {nullptr, nullptr, nullptr};
    for (int k = 0; k < dim; ++k) {
      out[k].assign((size_t)n_dofs_, 0.0);
      o[k] = out[k].data();
    }
    check(pfem2gpu_project_particle_fields(h_, o, PFEM2GPU_HOST));
    return out;
  }
  // ... into caller-owned component arrays of n_dofs entries each
  void project_particle_fields(const std::array<double*, dim>& out, int where = PFEM2GPU_HOST) {
    flush_inserts();
    double* o[3] = {nullptr, nullptr, nullptr};
    for (int k = 0; k < dim; ++k) o[k] = out[k];
    check(pfem2gpu_project_particle_fields(h_, o, where));
  }

  // The reference's timer sections (HIP event time of everything each call
  // enqueues). Turning the timer on or off clears the totals.
  void enable_timer(bool on) { check(pfem2gpu_set_timing(h_, on ? 1 : 0)); }
  std::map<std::string, TimerSection> timer_sections() {
    const int cap = 128;
    const char* names[cap];
    int64_t calls[cap];
    double ms[cap];
    int n = 0;
    check(pfem2gpu_timings(h_, cap, names, calls, ms, &n));
    std::map<std::string, TimerSection> all;
    for (int i = 0; i < n && i < cap; ++i) all[names[i]] = TimerSection{calls[i], ms[i]};
    std::map<std::string, TimerSection> out;
    for (const char* s : timer_section_names())
      if (all.count(s)) out[s] = all[s];
    return out;
  }
  // A TimerOutput-style table of the sections (wall time of the device work)
  void print_timer_summary(std::FILE* f = stdout) {
    const auto sec = timer_sections();
    std::fprintf(f, "+---------------------------------+-----------+--------------+\n");
    std::fprintf(f, "| Section                         | no. calls | device time  |\n");
    std::fprintf(f, "+---------------------------------+-----------+--------------+\n");
    for (const auto& e : sec)
      std::fprintf(f, "| %-31s | %9lld | %10.4f ms |\n", e.first.c_str(), (long long)e.second.calls, e.second.total_ms);
    std::fprintf(f, "+---------------------------------+-----------+--------------+\n");
  }

  // output_particle_solution() (src/pfem2particlehandler.cpp:217-317)
  void output_particle_solution(int timestep_number, const std::string& directory = ".", int rank = 0,
                                int n_ranks = 1) {
    flush_inserts();
    check(pfem2gpu_output_particle_solution(h_, timestep_number, rank, n_ranks, rank, directory.c_str()));
  }

  // particles.size()
  std::size_t n_particles() {
    flush_inserts();
    int64_t n = 0;
    check(pfem2gpu_particle_count(h_, &n));
    return (std::size_t)n;
  }

  // The particles in the handle's order (the reference iterates its vector
  // with begin()/end(), src/pfem2particlehandler.h:24-25).
  std::vector<Particle<dim>> particles() {
    const std::size_t n = n_particles();
    std::vector<double> x(dim * n), xi(dim * n), v(dim * n);
    std::vector<int32_t> cell(n);
    if (n) check(pfem2gpu_download(h_, x.data(), xi.data(), v.data(), cell.data()));
    std::vector<Particle<dim>> out(n);
    for (std::size_t p = 0; p < n; ++p) {
      for (int k = 0; k < dim; ++k) {
        out[p].location[k] = x[k * n + p];
        out[p].reference_location[k] = xi[k * n + p];
        out[p].velocity[k] = v[k * n + p];
      }
      out[p].cell = cell[p];
    }
    return out;
  }

  // Records rejected by the last insertion (location not inside its cell).
  int64_t rejected_insertions() const { return rejected_; }

  pfem2gpu_handle* handle() { return h_; }

 private:
  static void check(int rc) {
    if (rc != PFEM2GPU_OK) throw std::runtime_error(pfem2gpu_last_error());
  }
  void check_sizes(const NodalField<dim>& V) const {
    for (int k = 0; k < dim; ++k)
      if ((int64_t)V[k].size() != n_dofs_) throw std::runtime_error("nodal field size differs from n_dofs");
  }
  const double* const* ptrs(const NodalField<dim>& V) {
    check_sizes(V);
    for (int k = 0; k < 3; ++k) fptr_[k] = k < dim ? V[k].data() : nullptr;
    return fptr_;
  }
  const double* const* raw(const std::array<const double*, dim>& V) {
    for (int k = 0; k < 3; ++k) fptr_[k] = k < dim ? V[k] : nullptr;
    return fptr_;
  }
  void flush_inserts() {
    if (pending_cells_.empty()) return;
    const int64_t n = (int64_t)pending_cells_.size();
    std::vector<double> x((size_t)dim * n), v((size_t)dim * n);
    for (int k = 0; k < dim; ++k)
      for (int64_t p = 0; p < n; ++p) {
        x[k * n + p] = pending_x_[k][p];
        v[k * n + p] = pending_v_[k][p];
      }
    int64_t rej = 0;
    const int rc = pfem2gpu_insert_particles(h_, n, x.data(), pending_cells_.data(), v.data(), PFEM2GPU_HOST, &rej);
    pending_cells_.clear();
    for (int k = 0; k < dim; ++k) {
      pending_x_[k].clear();
      pending_v_[k].clear();
    }
    check(rc);
    rejected_ = rej;
  }

  MeshData mesh_;
  pfem2gpu_handle* h_ = nullptr;
  int64_t n_dofs_ = 0;
  int64_t rejected_ = 0;
  const double* fptr_[3] = {nullptr, nullptr, nullptr};
  std::vector<int32_t> pending_cells_;
  std::array<std::vector<double>, dim> pending_x_;
  std::array<std::vector<double>, dim> pending_v_;
};

}  // namespace pfem2gpu

#endif  // PFEM2GPU_HANDLER_HPP
