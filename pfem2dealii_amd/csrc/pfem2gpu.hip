// Host side of the particle path: device buffers, step orchestration and the
// C ABI declared in include/pfem2gpu.h.
//
// Layout in HBM (per handle):
//   particles   SoA FP64 x[dim], v[dim] + int32 cell, sorted by cell; the
//               order inside a cell is (part bucket, canonical key);
//               cell_start[n_cells+1] delimits the cell segments and its last
//               entry is the particle count.
//   staging T   the same SoA x, v arrays (+ cell for insertion): the move
//               passes scatter into it, the settle kernel compacts back.
//   mesh        vertex coordinates SoA, cell->vertex and cell->dof tables
//               (16/32 B per cell, one vector load), box flags, the
//               vertex->cells CSR with unit centre directions, and the
//               dof->(cell, vertex) CSR used by the node quotient.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/pfem2gpu.h"
#include "pfem2gpu_adjoint.h"
#include "pfem2gpu_fem.h"
#include "pfem2gpu_output.h"
#include "pfem2gpu_wire.h"

using namespace pfem2gpu;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(expr)                                                                         \
  do {                                                                                       \
    hipError_t e_ = (expr);                                                                  \
    if (e_ != hipSuccess)                                                                    \
      return fail(PFEM2GPU_ERR_HIP, std::string(#expr ": ") + hipGetErrorString(e_));         \
  } while (0)

struct Timed {
  const char* name;
  hipEvent_t a, b;
};

}  // namespace

struct pfem2gpu_handle {
  int dim = 2, nvc = 4;
  int nc = 0, nv = 0, ndofs = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  Prm prm_host{};
  int refill_max = 0;  // upper bound of particles added by one redistribution

  // mesh
  double* X[3] = {nullptr, nullptr, nullptr};
  int* cv = nullptr;
  int* cd = nullptr;
  unsigned char* box = nullptr;
  int* v2c_off = nullptr;
  int* v2c = nullptr;
  double* v2c_dir = nullptr;
  int* d2c_off = nullptr;
  int* d2c = nullptr;
  int* bpos = nullptr;  // [nc*nvc] place of each (cell, vertex) in the d2c lists
  Prm* prm = nullptr;
  int* cell_inst = nullptr;          // [nc] instance of each cell (batches only)
  bool all_box = false;              // every cell is an axis-aligned box
  int ninst = 1;
  std::vector<Prm> prm_inst;         // per-instance parameters (host copy)
  std::vector<long> inst_cell_off;   // [ninst+1]
  std::vector<long> inst_dof_off;    // [ninst+1]
  double* geom = nullptr;  // [nc][2*dim] box corners (X_0[0] = NaN for other cells)
  int2* cvo = nullptr;     // [nc][nvc] (v2c offset, count) per cell vertex
  int4* cvc = nullptr;     // [nc][2^dim][2^dim / 4]: the cells around each cell vertex
  int* ptab = nullptr;     // [pmax] part decode table
  double* Vc = nullptr;    // [nc][nvc][dim] advecting field, cell-major
  int pshift = 0;
  bool maps_ready = false;

  // particles
  long cap = 0;
  double* px[3] = {nullptr, nullptr, nullptr};
  double* pv[3] = {nullptr, nullptr, nullptr};
  int* pcell = nullptr;
  double* tx[3] = {nullptr, nullptr, nullptr};
  double* tv[3] = {nullptr, nullptr, nullptr};
  int* tcell = nullptr;
  int* cell_start = nullptr;  // [nc+1], rewritten in place by the settle kernel
  int* count = nullptr;       // [3*nchunks+2]: bucket counts | spill counts | spill fill | spill total | ticket
  int* sstart = nullptr;      // [nchunks+1] spill offsets per chunk
  unsigned long long* lb = nullptr;  // [nchunks] look-back status words of the settle kernel
  int* d_err = nullptr;       // device-side error flags (look-back timeout)
  int nchunks = 0, cshift = 4;
  int capc = 0;               // bucket capacity per chunk
  long tsize = 0;             // staging slots: nchunks*capc + 2*cap
  int* scan_blk = nullptr;
  int scan_nb = 0;
  double* contrib = nullptr;  // [nc*nvc][dim+1] per (cell, vertex) sums in d2c order (place bpos)
  double* dVc = nullptr;  // [nc][nvc][dim] pending correction V - V_old, cell-major
  bool pending_corr = false;
  double* stageF[2][3] = {{nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}};
  unsigned long long* d_rejected = nullptr;
  int* h_pinned = nullptr;  // [0..1] synchronous reads, [2..3] adjoint loss, [4..5] count and error after a
                            // settle, [6..9] the no-allocation pack's duplicate flag and total
  int* scratch = nullptr;   // this device's no-allocation scratch (shared by every handle on the device)
  hipEvent_t settle_ev = nullptr;  // recorded after [4..5] are copied, following each eager settle
  bool settle_ev_valid = false;    // [4] is the current count once settle_ev completes
  long n_ub = 0;  // host-side upper bound of the particle count
  int cpw = 16, pmax = 1, nmax = 512, nslots = 0, settle_smem = 0;  // cpw = 1 << cshift
  unsigned char* dirty = nullptr;  // per settle chunk: sums left to k_contrib_dirty
  bool contrib_valid = false;      // per (cell, vertex) sums match the particles

  // walk queues of the split move (two, used in turn by the walk rounds)
  double* qx[2][3] = {{nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}};
  double* qv[2][3] = {{nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}};
  int* qcell[2] = {nullptr, nullptr};
  int* qs[2] = {nullptr, nullptr};
  int* qsrc[2] = {nullptr, nullptr};
  long qcap = 0;
  int* qn = nullptr;  // [qn_len] queue lengths of the walk rounds (max n_sub + 1 used)
  int qn_len = 0;

  // FEM assembly (built by the first assembly or sparsity call)
  bool fem_ready = false;
  std::vector<int64_t> fem_rowptr;  // [ndofs+1] CSR rows of the DoFs sharing a cell (host)
  std::vector<int32_t> fem_col;     // [nnz] ascending columns per row (host)
  long* fem_soff = nullptr;         // [nnz+1] contributions of each entry (runs in fem_A)
  int* fem_apos = nullptr;          // [nc*nvc*nvc] place of cell-matrix entry (c, i, j) in fem_A
  double* fem_A = nullptr;          // cell-matrix entries in contribution order (entry, then cell); planes
  int fem_A_planes = 0;             // matrices fem_A holds (1, or dim for the velocity prediction)
  double* fem_b = nullptr;          // [dim][nc*nvc] cell right-hand sides in DoF -> (cell, vertex) order
  double* fem_vals = nullptr;       // [fem_vals_planes][nnz] staging of host matrix outputs
  int fem_vals_planes = 0;
  double* fem_rhs[3] = {nullptr, nullptr, nullptr};  // [ndofs] staging of host right-hand sides
  double* fem_P = nullptr;          // [ndofs] staging of a host pressure input
  unsigned char* fem_faces = nullptr;  // [nc] staging of host face flags

  // device buffers kept between adjoint calls (pointer, bytes)
  std::vector<std::pair<void*, size_t>> adj_pool;

  // identity recording for the adjoint (null when off)
  int* rec_src = nullptr;  // where the next move writes each slot's source particle
  DPath rec_path = {nullptr, {nullptr, nullptr, nullptr}, 0};  // where the next move records its paths
  int* tsrc = nullptr;     // [tsize] staging source indices
  int* gord = nullptr;     // [tsize] grouping scratch of settle chunks that exceed the LDS stage

  // tuning (set by pfem2gpu_set_tuning; chosen by the autotuner)
  int walk_grid = 4096;     // workgroups of a walk round (multiple of kQStripes)
  int settle_grid = 2048;   // workgroups of the settle kernel
  int move_grid = 0;        // workgroups of the in-cell move pass (0: one particle per thread)
  int move_block = kBlock;  // threads per workgroup of the in-cell move pass (64, 128, 256 or 512)
  int walk_block = kBlock;  // threads per workgroup of a walk round (64, 128, 256 or 512)
  // graph-captured harness step
  bool capturing = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  const double* graph_V[3] = {nullptr, nullptr, nullptr};
  double* graph_Vp[3] = {nullptr, nullptr, nullptr};

  // timing
  bool timing = false;
  std::vector<Timed> pending;
  std::vector<hipEvent_t> pool;
  std::map<std::string, std::pair<long, double>> totals;
  std::vector<std::string> names_store;
};

namespace {

constexpr int kSettleLds = 52 * 1024;  // LDS budget of a settle workgroup (three per CU)

int grid_for(long n) {
  long g = (n + kBlock - 1) / kBlock;
  if (g < 1) g = 1;
  if (g > 16384) g = 16384;
  return (int)g;
}

hipEvent_t take_event(pfem2gpu_handle* h) {
  if (!h->pool.empty()) {
    hipEvent_t e = h->pool.back();
    h->pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

void t_begin(pfem2gpu_handle* h, const char* name) {
  if (!h->timing) return;
  Timed t{name, take_event(h), take_event(h)};
  if (t.a && t.b) {
    (void)hipEventRecord(t.a, h->stream);
    h->pending.push_back(t);
  }
}

void t_end(pfem2gpu_handle* h) {
  if (!h->timing || h->pending.empty()) return;
  (void)hipEventRecord(h->pending.back().b, h->stream);
}

int t_collect(pfem2gpu_handle* h) {
  HIPCHK(hipStreamSynchronize(h->stream));
  for (auto& t : h->pending) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, t.a, t.b) == hipSuccess) {
      auto& e = h->totals[t.name];
      e.first += 1;
      e.second += ms;
    }
    h->pool.push_back(t.a);
    h->pool.push_back(t.b);
  }
  h->pending.clear();
  return PFEM2GPU_OK;
}

// One of the reference's TimerOutput sections around a handler call
// (src/pfem2particlehandler.cpp:45,114,144,176): with timing on, a pair of
// events on the handle's stream brackets everything the call enqueues.
struct Section {
  pfem2gpu_handle* h;
  long idx = -1;
  Section(pfem2gpu_handle* hh, const char* name) : h(hh) {
    if (!h || !h->timing) return;
    Timed t{name, take_event(h), take_event(h)};
    if (t.a && t.b && hipEventRecord(t.a, h->stream) == hipSuccess) {
      h->pending.push_back(t);
      idx = (long)h->pending.size() - 1;
    }
  }
  ~Section() {
    if (idx >= 0) (void)hipEventRecord(h->pending[idx].b, h->stream);
  }
};

#define LAUNCH_B(h, name, kernel, grid, block, smem, ...)                         \
  do {                                                                            \
    t_begin(h, name);                                                             \
    hipLaunchKernelGGL(kernel, dim3(grid), dim3(block), smem, h->stream, __VA_ARGS__); \
    t_end(h);                                                                     \
    HIPCHK(hipGetLastError());                                                    \
  } while (0)
#define LAUNCH(h, name, kernel, grid, smem, ...) LAUNCH_B(h, name, kernel, grid, kBlock, smem, __VA_ARGS__)

template <typename T>
int dalloc(T** p, size_t n) {
  *p = nullptr;
  if (n == 0) n = 1;
  if (hipMalloc(reinterpret_cast<void**>(p), n * sizeof(T)) != hipSuccess)
    return fail(PFEM2GPU_ERR_MEMORY, "hipMalloc of " + std::to_string(n * sizeof(T)) + " bytes failed");
  return PFEM2GPU_OK;
}

template <typename T>
void dfree(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

#define CHK(expr)                      \
  do {                                 \
    int rc_ = (expr);                  \
    if (rc_ != PFEM2GPU_OK) return rc_; \
  } while (0)

DMesh dmesh(pfem2gpu_handle* h) {
  DMesh m;
  for (int k = 0; k < 3; ++k) m.X[k] = h->X[k];
  m.cv = h->cv;
  m.cd = h->cd;
  m.box = h->box;
  m.v2c_off = h->v2c_off;
  m.v2c = h->v2c;
  m.v2c_dir = h->v2c_dir;
  m.cell_inst = h->cell_inst;
  m.prm = h->prm;
  m.geom = h->geom;
  m.cvo = h->cvo;
  m.cvc = h->cvc;
  m.ptab = h->ptab;
  m.bpos = h->bpos;
  return m;
}

DParts dparts(pfem2gpu_handle* h) {
  DParts p;
  for (int k = 0; k < 3; ++k) {
    p.x[k] = h->px[k];
    p.v[k] = h->pv[k];
  }
  p.cell = h->pcell;
  p.src = h->rec_src;
  return p;
}

DTmp dtmp(pfem2gpu_handle* h) {
  DTmp t;
  for (int k = 0; k < 3; ++k) {
    t.x[k] = h->tx[k];
    t.v[k] = h->tv[k];
  }
  t.cell = h->tcell;
  t.src = h->rec_src ? h->tsrc : nullptr;
  return t;
}

void drop_graph(pfem2gpu_handle* h) {
  if (h->graph_exec) (void)hipGraphExecDestroy(h->graph_exec);
  if (h->graph) (void)hipGraphDestroy(h->graph);
  h->graph_exec = nullptr;
  h->graph = nullptr;
}

// Workgroups of the in-cell move pass over n particles (the tuned grid, or
// one particle per thread), and the walk-queue stripe size for it: the
// particles of the ceil(grid / kQStripes) workgroups that write to one stripe.
long move_cell_grid(const pfem2gpu_handle* h, long n) {
  const long b = h->move_block;
  long g = std::max(1L, std::min((n + b - 1) / b, 1L << 20));
  if (h->move_grid > 0) g = std::min(g, (long)h->move_grid);
  return g;
}

// particles one walk-queue stripe may receive: those of the in-cell move's
// workgroups that write to it (workgroup b writes to stripe b % kQStripes)
long walk_qreg(const pfem2gpu_handle* h, long n) {
  const long b = h->move_block;
  const long gc = move_cell_grid(h, n);
  const long per_block = b * ((n + gc * b - 1) / (gc * b));
  return ((gc + kQStripes - 1) / kQStripes) * per_block;
}

// The move kernels are instantiated for the workgroup sizes the tuning
// allows; bs selects one.
#define MOVE_LAUNCH(h, name, KERNEL, D, bs, grid, ...)                                  \
  do {                                                                                \
    switch (bs) {                                                                     \
      case 64: LAUNCH_B(h, name, (KERNEL<D, 64>), grid, 64, 0, __VA_ARGS__); break;    \
      case 128: LAUNCH_B(h, name, (KERNEL<D, 128>), grid, 128, 0, __VA_ARGS__); break; \
      case 512: LAUNCH_B(h, name, (KERNEL<D, 512>), grid, 512, 0, __VA_ARGS__); break; \
      default: LAUNCH_B(h, name, (KERNEL<D, 256>), grid, 256, 0, __VA_ARGS__); break;  \
    }                                                                                 \
  } while (0)

void free_queues(pfem2gpu_handle* h) {
  for (int b = 0; b < 2; ++b) {
    for (int k = 0; k < 3; ++k) {
      dfree(h->qx[b][k]);
      dfree(h->qv[b][k]);
    }
    dfree(h->qcell[b]);
    dfree(h->qs[b]);
    dfree(h->qsrc[b]);
  }
  h->qcap = 0;
}

// Walk queues for `need` particles (every particle may leave its cell in
// the first substep).
int grow_queues(pfem2gpu_handle* h, long need) {
  if (need <= h->qcap) return PFEM2GPU_OK;
  HIPCHK(hipStreamSynchronize(h->stream));
  free_queues(h);
  for (int b = 0; b < 2; ++b) {
    for (int k = 0; k < h->dim; ++k) {
      CHK(dalloc(&h->qx[b][k], need));
      CHK(dalloc(&h->qv[b][k], need));
    }
    CHK(dalloc(&h->qcell[b], need));
    CHK(dalloc(&h->qs[b], need));
    CHK(dalloc(&h->qsrc[b], need));
  }
  h->qcap = need;
  return PFEM2GPU_OK;
}

DWalk dwalk(pfem2gpu_handle* h, int b) {
  DWalk q;
  for (int k = 0; k < 3; ++k) {
    q.x[k] = h->qx[b][k];
    q.v[k] = h->qv[b][k];
  }
  q.cell = h->qcell[b];
  q.s = h->qs[b];
  q.src = h->qsrc[b];
  q.err = h->d_err;
  return q;
}

void free_particles(pfem2gpu_handle* h) {
  for (int k = 0; k < 3; ++k) {
    dfree(h->px[k]);
    dfree(h->pv[k]);
    dfree(h->tx[k]);
    dfree(h->tv[k]);
  }
  dfree(h->pcell);
  dfree(h->tcell);
  dfree(h->tsrc);
  dfree(h->gord);
  free_queues(h);
  h->cap = 0;
  h->tsize = 0;
}

// Grows the particle arrays to hold at least `need` particles, keeping the
// first `keep` current ones, and the staging arrays to nchunks*capc bucket
// slots plus two spill regions of the particle capacity.
int grow(pfem2gpu_handle* h, long need, long keep, int capc) {
  long ncap = h->cap;
  if (need > h->cap) {
    ncap = std::max(need, 2 * h->cap);
    ncap = std::max(ncap, 1024L);
    if (ncap > 0x7fffffffL) return fail(PFEM2GPU_ERR_MEMORY, "particle capacity exceeds int32 indexing");
  }
  capc = std::max(capc, h->capc);
  const long tsize = (long)h->nchunks * capc + 2 * ncap;
  if (ncap != h->cap || tsize != h->tsize) drop_graph(h);  // a captured step refers to the old buffers
  if (ncap != h->cap) {
    double *nx[3] = {nullptr, nullptr, nullptr}, *nvv[3] = {nullptr, nullptr, nullptr};
    int* ncell = nullptr;
    for (int k = 0; k < h->dim; ++k) {
      CHK(dalloc(&nx[k], ncap));
      CHK(dalloc(&nvv[k], ncap));
    }
    CHK(dalloc(&ncell, ncap));
    if (keep > 0) {
      for (int k = 0; k < h->dim; ++k) {
        HIPCHK(hipMemcpyAsync(nx[k], h->px[k], keep * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
        HIPCHK(hipMemcpyAsync(nvv[k], h->pv[k], keep * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
      }
      HIPCHK(hipMemcpyAsync(ncell, h->pcell, keep * sizeof(int), hipMemcpyDeviceToDevice, h->stream));
    }
    HIPCHK(hipStreamSynchronize(h->stream));
    for (int k = 0; k < 3; ++k) {
      dfree(h->px[k]);
      dfree(h->pv[k]);
    }
    dfree(h->pcell);
    for (int k = 0; k < h->dim; ++k) {
      h->px[k] = nx[k];
      h->pv[k] = nvv[k];
    }
    h->pcell = ncell;
    h->cap = ncap;
  }
  CHK(grow_queues(h, walk_qreg(h, h->cap) * kQStripes));
  if (tsize != h->tsize) {
    HIPCHK(hipStreamSynchronize(h->stream));
    for (int k = 0; k < 3; ++k) {
      dfree(h->tx[k]);
      dfree(h->tv[k]);
    }
    dfree(h->tcell);
    for (int k = 0; k < h->dim; ++k) {
      CHK(dalloc(&h->tx[k], tsize));
      CHK(dalloc(&h->tv[k], tsize));
    }
    CHK(dalloc(&h->tcell, tsize));
    dfree(h->gord);
    CHK(dalloc(&h->gord, tsize));
    if (h->tsrc) {
      dfree(h->tsrc);
      CHK(dalloc(&h->tsrc, tsize));
    }
    h->tsize = tsize;
    h->capc = capc;
  }
  return PFEM2GPU_OK;
}

// bucket capacity for about `n` particles: 2.5x the mean chunk population
int capc_for(const pfem2gpu_handle* h, long n) {
  long c = (5 * n) / (2 * (long)h->nchunks) + 64;
  c = std::max(c, 256L);
  c = (c + 63) / 64 * 64;
  return (int)std::min(c, 1L << 20);
}

// The failure a set of device error flags stands for (PFEM2GPU_OK if none).
int device_error(int flags) {
  if (flags & kErrLookback)
    return fail(PFEM2GPU_ERR_HIP, "redistribution look-back wait timed out; particle layout is invalid");
  if (flags & kErrQueue)
    return fail(PFEM2GPU_ERR_STATE, "a walk-queue stripe overflowed; particles were dropped");
  if (flags & kErrRange)
    return fail(PFEM2GPU_ERR_ARG, "an adjoint term was not finite or exceeded 2^46 in magnitude");
  return PFEM2GPU_OK;
}

// Synchronises and reports device-side failures (a look-back wait that
// timed out, or a walk queue that overflowed, leaves a wrong layout).
int check_device(pfem2gpu_handle* h) {
  HIPCHK(hipMemcpyAsync(h->h_pinned + 1, h->d_err, sizeof(int), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  return device_error(h->h_pinned[1]);
}

int read_count(pfem2gpu_handle* h, long* n) {
  HIPCHK(hipMemcpyAsync(h->h_pinned, h->cell_start + h->nc, sizeof(int), hipMemcpyDeviceToHost, h->stream));
  CHK(check_device(h));
  *n = *h->h_pinned;
  h->n_ub = *n;
  return PFEM2GPU_OK;
}

// Makes room for `extra` particles beyond the current count (particle
// arrays) and for re-binning the current plus `rebin_extra` particles
// (staging buckets). When the host-side bound says the capacity might not
// suffice, the count after the last eager settle (copied to pinned memory
// behind an event) replaces the bound: waiting for that event leaves the work
// queued after the settle running. Only if that count does not fit either is
// the stream drained and the capacity grown.
int reserve(pfem2gpu_handle* h, long extra, long rebin_extra) {
  if (h->capturing) return PFEM2GPU_OK;  // capacity was secured before the capture began
  auto fits = [&](long nb) { return nb + extra <= h->cap && capc_for(h, nb + rebin_extra) <= h->capc; };
  if (fits(h->n_ub)) return PFEM2GPU_OK;
  if (h->settle_ev_valid) {
    HIPCHK(hipEventSynchronize(h->settle_ev));
    CHK(device_error(h->h_pinned[5]));
    h->n_ub = h->h_pinned[4];
    if (fits(h->n_ub)) return PFEM2GPU_OK;
  }
  long n = 0;
  CHK(read_count(h, &n));
  const int capc = capc_for(h, n + rebin_extra);
  if (n + extra <= h->cap && capc <= h->capc) return PFEM2GPU_OK;
  return grow(h, n + extra <= h->cap ? h->cap : 2 * (n + extra), n, capc);
}

int scan(pfem2gpu_handle* h, const int* in, int n, int* out) {
  const int nb = (n + kScanChunk - 1) / kScanChunk;
  LAUNCH(h, "scan", k_scan_reduce, nb, 0, in, n, h->scan_blk);
  LAUNCH(h, "scan", k_scan_blocks, 1, 0, h->scan_blk, nb, out + n);
  LAUNCH(h, "scan", k_scan_down, nb, 0, in, n, h->scan_blk, out);
  return PFEM2GPU_OK;
}

// Field pointers for a call: device pointers pass through, host arrays are
// staged into slot `slot` (synchronous copy).
int field_in(pfem2gpu_handle* h, const double* const* F, int where, int slot, DField* out) {
  if (!F) return fail(PFEM2GPU_ERR_ARG, "null field");
  for (int k = 0; k < 3; ++k) out->c[k] = nullptr;
  for (int k = 0; k < h->dim; ++k) {
    if (!F[k]) return fail(PFEM2GPU_ERR_ARG, "null field component");
    if (where == PFEM2GPU_DEVICE) {
      out->c[k] = F[k];
    } else if (where == PFEM2GPU_HOST) {
      HIPCHK(hipMemcpyAsync(h->stageF[slot][k], F[k], h->ndofs * sizeof(double), hipMemcpyHostToDevice, h->stream));
      out->c[k] = h->stageF[slot][k];
    } else {
      return fail(PFEM2GPU_ERR_ARG, "where must be PFEM2GPU_HOST or PFEM2GPU_DEVICE");
    }
  }
  if (where == PFEM2GPU_HOST) HIPCHK(hipStreamSynchronize(h->stream));
  return PFEM2GPU_OK;
}


template <int D>
int flush_correction(pfem2gpu_handle* h) {
  if (!h->pending_corr) return PFEM2GPU_OK;
  LAUNCH(h, "k_correct", k_correct<D>, grid_for(h->n_ub), 0, dmesh(h), h->cell_start + h->nc, dparts(h), h->dVc);
  h->pending_corr = false;
  h->contrib_valid = false;
  return PFEM2GPU_OK;
}

int flush(pfem2gpu_handle* h) { return h->dim == 2 ? flush_correction<2>(h) : flush_correction<3>(h); }

// move (or re-bin) into chunk buckets -> bin the spill -> settle(count) ->
// scan -> settle(write). With advect = 0 this re-bins the first n_cur + n_add
// particles without moving them and without the per-part cap or refills.
template <int D>
int rebin_or_move(pfem2gpu_handle* h, int advect, int n_add, DField Vf) {
  const DMesh m = dmesh(h);
  const DParts P = dparts(h);
  const DTmp T = dtmp(h);
  const long nmax_p = h->n_ub + n_add;
  const int* nptr = h->cell_start + h->nc;
  const int nch = h->nchunks;
  int* bcount = h->count;
  int* scount = h->count + nch;
  int* sfill = h->count + 2 * nch;
  int* spill_n = h->count + 3 * nch;
  const long spill_base = (long)nch * h->capc;
  const long spill2_base = spill_base + h->cap;
  const int use_dv = (advect && h->pending_corr) ? 1 : 0;
  if (!advect) CHK(flush_correction<D>(h));
  LAUNCH(h, "k_fill", k_zero3, grid_for(5L * nch + 2 + (long)h->qn_len * kQStripes), 0, h->count, 3L * nch + 2,
         reinterpret_cast<int*>(h->lb), 2L * nch, h->qn, (long)h->qn_len * kQStripes);
  if (advect)
    LAUNCH(h, "k_cellfield", k_cellfield<D>, grid_for((long)h->nc * h->nvc), 0, h->nc, h->cd, Vf, h->Vc);
  int rounds = 0;
  if (advect)
    for (const Prm& q : h->prm_inst) rounds = std::max(rounds, q.n_sub);
  // a captured sequence is replayed at later counts: its grid and stripe size
  // cover every count the capacity allows (the pre-replay reserve keeps the
  // count at the start of a step within cap - refill_max)
  const long nsize = h->capturing ? std::max(nmax_p, h->cap) : nmax_p;
  const long gc = move_cell_grid(h, nsize);
  const long qreg = walk_qreg(h, nsize);
  if (qreg * kQStripes > h->qcap) {
    if (h->capturing) return fail(PFEM2GPU_ERR_STATE, "walk queues too small during graph capture");
    CHK(grow_queues(h, qreg * kQStripes));
  }
  MOVE_LAUNCH(h, advect ? "k_move" : "k_rebin", k_move_cell, D, h->move_block, gc, advect, m, h->Vc, nptr, n_add, P,
              h->dVc, use_dv, h->cshift, h->capc, bcount, spill_base, spill_n, T, dwalk(h, 0), h->qn, qreg,
              h->rec_path);
  const int gw = kQStripes * std::max(1, h->walk_grid / kQStripes);
  for (int r = 0; r < rounds; ++r)
    MOVE_LAUNCH(h, "k_walk", k_move_walk, D, h->walk_block, gw, m, h->Vc, dwalk(h, r & 1),
                h->qn + (long)r * kQStripes, qreg, h->cshift, h->capc, bcount, spill_base, spill_n, T,
                dwalk(h, (r + 1) & 1), h->qn + (long)(r + 1) * kQStripes, h->rec_path);
  if (use_dv) h->pending_corr = false;
  LAUNCH(h, "k_spill", k_spill_count, grid_for(nmax_p), 0, spill_n, spill_base, h->tcell, h->cshift, scount);
  CHK(scan(h, scount, nch, h->sstart));
  LAUNCH(h, "k_spill", k_spill_scatter<D>, grid_for(nmax_p), 0, spill_n, spill_base, spill2_base, h->cshift,
         h->sstart, sfill, T);
  const int gs = std::min(nch, h->settle_grid);
  const int cap_on = advect, refill_on = advect;
  LAUNCH_B(h, "k_settle", k_settle<D>, gs, kSettleBlock, h->settle_smem, cap_on, refill_on, m, h->nc, h->cshift, h->pshift, h->nmax,
         h->nslots, h->capc, bcount, h->sstart, spill2_base, T, h->count + 3 * nch + 1, h->lb, h->d_err, h->cell_start,
         P, h->Vc, h->gord, h->contrib, h->dirty);
  h->n_ub = nmax_p + (advect ? h->refill_max : 0);
  h->contrib_valid = true;
  h->settle_ev_valid = false;
  if (!h->capturing) {  // the count after this settle, for reserve()
    HIPCHK(hipMemcpyAsync(h->h_pinned + 4, h->cell_start + h->nc, sizeof(int), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipMemcpyAsync(h->h_pinned + 5, h->d_err, sizeof(int), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipEventRecord(h->settle_ev, h->stream));
    h->settle_ev_valid = true;
  }
  return PFEM2GPU_OK;
}

int do_rebin(pfem2gpu_handle* h, int advect, int n_add, DField Vf) {
  return h->dim == 2 ? rebin_or_move<2>(h, advect, n_add, Vf) : rebin_or_move<3>(h, advect, n_add, Vf);
}

// Stages n particles (device pointers) and appends + re-bins them.
int append_staged(pfem2gpu_handle* h, long n, const double* const* sx, const double* const* sv, const int* scell,
                  int64_t* n_rejected) {
  HIPCHK(hipMemsetAsync(h->d_rejected, 0, sizeof(unsigned long long), h->stream));
  const DMesh m = dmesh(h);
  const double* z = nullptr;
  const double* svv[3] = {sv ? sv[0] : z, sv ? sv[1] : z, (sv && h->dim > 2) ? sv[2] : z};
  const double* sxx[3] = {sx[0], sx[1], h->dim > 2 ? sx[2] : z};
  if (h->dim == 2) {
    LAUNCH(h, "k_append", k_append<2>, grid_for(n), 0, m, h->nc, h->cell_start + h->nc, (int)n, sxx[0], sxx[1],
           sxx[2], svv[0], svv[1], svv[2], scell, dparts(h), h->d_rejected);
  } else {
    LAUNCH(h, "k_append", k_append<3>, grid_for(n), 0, m, h->nc, h->cell_start + h->nc, (int)n, sxx[0], sxx[1],
           sxx[2], svv[0], svv[1], svv[2], scell, dparts(h), h->d_rejected);
  }
  DField none;
  for (int k = 0; k < 3; ++k) none.c[k] = nullptr;
  CHK(do_rebin(h, 0, (int)n, none));
  if (n_rejected) {
    unsigned long long r = 0;
    HIPCHK(hipMemcpyAsync(&r, h->d_rejected, sizeof(r), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    *n_rejected = (int64_t)r;
  }
  return PFEM2GPU_OK;
}

}  // namespace

// ------------------------------------------------------ no-allocation scratch
// Device scratch of pfem2gpu_pack_particles_noalloc: a fixed array in the
// library's static device memory, one copy per GPU, held until the process
// ends and shared by every call on that GPU (the last call's contents stay;
// calls that overlap on one GPU race). Host code never passes the array's
// name to a kernel or a copy (on the host it only stands for the symbol): it
// uses the per-device address from hipGetSymbolAddress, taken once per device
// when the first handle is created there with that device current.
constexpr long kNoAllocScratchInts = 1L << 23;  // 32 MiB
static __device__ __align__(256) int scratch[kNoAllocScratchInts];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
int* g_scratch_ptr[kMaxDevices] = {};

// The current device's scratch address (looked up on its first use).
int scratch_for_current_device(int** out) {
  int dev = -1;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(PFEM2GPU_ERR_HIP, "device ordinal out of range for the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = static_cast<int*>(p);
  }
  *out = g_scratch_ptr[dev];
  return PFEM2GPU_OK;
}

}  // namespace

// =================================================================== C ABI

static int validate_mesh(const pfem2gpu_mesh* mesh, const pfem2gpu_params* params) {
  if (mesh->dim != 2 && mesh->dim != 3) return fail(PFEM2GPU_ERR_ARG, "dim must be 2 or 3");
  if (mesh->n_cells <= 0 || mesh->n_vertices <= 0 || mesh->n_dofs <= 0 || !mesh->vertices || !mesh->cell_vertices)
    return fail(PFEM2GPU_ERR_ARG, "empty mesh");
  if (mesh->n_cells > 0x7fffffff / 8 || mesh->n_vertices > 0x7fffffff || mesh->n_dofs > 0x7fffffff)
    return fail(PFEM2GPU_ERR_ARG, "mesh too large for int32 indexing");
  const int dim = mesh->dim, nvc = 1 << dim;
  for (int k = 0; k < dim; ++k)
    if (params->particles_per_cell[k] == 0 || params->particles_per_cell[k] > 64)
      return fail(PFEM2GPU_ERR_ARG, "particles_per_cell must be in [1, 64]");
  const int32_t* cdofs = mesh->cell_dofs ? mesh->cell_dofs : mesh->cell_vertices;
  for (long i = 0; i < mesh->n_cells * nvc; ++i) {
    if (mesh->cell_vertices[i] < 0 || mesh->cell_vertices[i] >= mesh->n_vertices)
      return fail(PFEM2GPU_ERR_ARG, "cell_vertices entry out of range");
    if (cdofs[i] < 0 || cdofs[i] >= mesh->n_dofs) return fail(PFEM2GPU_ERR_ARG, "cell_dofs entry out of range");
  }
  return PFEM2GPU_OK;
}

// parameters of one instance (src/pfem2particlehandler.cpp:34-39,47,362-364)
static Prm make_prm(int dim, const pfem2gpu_params* params) {
  Prm p{};
  p.n_sub = params->particle_integration_steps == 0 ? 1 : (int)params->particle_integration_steps;
  p.dt_sub = params->time_step / (double)p.n_sub;
  p.maxpp = (int)params->max_particles_per_cell_part;
  p.P = 1;
  p.Pvalid = 1;
  for (int k = 0; k < 3; ++k) {
    p.q[k] = k < dim ? (int)params->particles_per_cell[k] : 1;
    p.Q[k] = k < dim ? p.q[k] + 1 : 1;
    p.h[k] = 1.0 / (double)(unsigned)p.q[k];
  }
  p.qpow2 = 1;
  for (int k = 0; k < 3; ++k) {
    p.qf[k] = (double)p.q[k];
    p.Qm[k] = k < dim ? (unsigned)((0x100000000ULL + (unsigned)p.Q[k] - 1) / (unsigned)p.Q[k]) : 0u;
  }
  for (int k = 0; k < dim; ++k) {
    p.P *= p.Q[k];
    p.Pvalid *= p.q[k];
    if (p.q[k] & (p.q[k] - 1)) p.qpow2 = 0;
  }
  return p;
}

// Builds a handle over `mesh` (a concatenation of instances for a batch) with
// one parameter set per instance and the instance of every cell.
static int create_impl(const pfem2gpu_mesh* mesh, const std::vector<Prm>& prms, const std::vector<int>& cell_inst,
                       const std::vector<long>& cell_off, const std::vector<long>& dof_off, void* hip_stream,
                       pfem2gpu_handle** out) {
  const int dim = mesh->dim, nvc = 1 << dim;
  const int32_t* cdofs = mesh->cell_dofs ? mesh->cell_dofs : mesh->cell_vertices;
  pfem2gpu_handle* h = new pfem2gpu_handle();
  h->dim = dim;
  h->nvc = nvc;
  h->nc = (int)mesh->n_cells;
  h->nv = (int)mesh->n_vertices;
  h->ndofs = (int)mesh->n_dofs;
  if (hip_stream) {
    h->stream = (hipStream_t)hip_stream;
  } else {
    const hipError_t e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
      delete h;
      return fail(PFEM2GPU_ERR_HIP, std::string("hipStreamCreate failed: ") + hipGetErrorString(e));
    }
    h->own_stream = true;
  }
  auto bail = [&](int rc) {
    pfem2gpu_destroy(h);
    return rc;
  };

  h->ninst = (int)prms.size();
  h->prm_inst = prms;
  h->prm_host = prms[0];
  h->inst_cell_off = cell_off;
  h->inst_dof_off = dof_off;
  int Pmax = 1, Pvmax = 1;
  h->refill_max = 0;
  for (int c = 0; c < h->nc; ++c) h->refill_max += prms[cell_inst.empty() ? 0 : cell_inst[c]].Pvalid;
  for (const Prm& q : prms) {
    Pmax = std::max(Pmax, q.P);
    Pvmax = std::max(Pvmax, q.Pvalid);
  }
  Prm p = h->prm_host;
  p.P = Pmax;
  p.Pvalid = Pvmax;  // sizing only
  h->pshift = 0;
  while ((1 << h->pshift) < p.P) ++h->pshift;
  h->pmax = 1 << h->pshift;
  // settle stage: 16 cells per workgroup (one chunk); as many final slots
  // (survivors and refills of a chunk) as fit in kSettleLds bytes of LDS
  // (three workgroups per CU), and at most as many incoming particles as the
  // workgroup holds in registers
  h->cshift = kSettleShift;  // at most 8: the settle keeps chunk-local cell indices in bytes
  h->cpw = 1 << h->cshift;
  h->nchunks = (h->nc + h->cpw - 1) >> h->cshift;
  h->nslots = 64;
  for (int ns = 64; ns <= 8192; ns += 64) {
    if (settle_layout(dim, h->cpw, h->pmax, ns).bytes > kSettleLds) break;
    h->nslots = ns;
  }
  h->nmax = std::min(h->nslots, kSettleBlock * (dim == 2 ? SettlePre<2>::n : SettlePre<3>::n));
  h->settle_smem = settle_layout(dim, h->cpw, h->pmax, h->nslots).bytes;
  if (h->settle_smem > 64 * 1024) {
    delete h;
    return fail(PFEM2GPU_ERR_ARG, "particles_per_cell too large for the redistribution stage");
  }

  int rc;
  // mesh upload
  for (int k = 0; k < dim; ++k) {
    if ((rc = dalloc(&h->X[k], h->nv))) return bail(rc);
    if (hipMemcpy(h->X[k], mesh->vertices + (size_t)k * h->nv, h->nv * sizeof(double), hipMemcpyHostToDevice))
      return bail(fail(PFEM2GPU_ERR_HIP, "upload vertices"));
  }
  if ((rc = dalloc(&h->cv, (size_t)h->nc * nvc)) || (rc = dalloc(&h->cd, (size_t)h->nc * nvc)) ||
      (rc = dalloc(&h->box, h->nc)))
    return bail(rc);
  if (hipMemcpy(h->cv, mesh->cell_vertices, (size_t)h->nc * nvc * 4, hipMemcpyHostToDevice) ||
      hipMemcpy(h->cd, cdofs, (size_t)h->nc * nvc * 4, hipMemcpyHostToDevice))
    return bail(fail(PFEM2GPU_ERR_HIP, "upload cells"));
  // axis-aligned box flag: vertex i has coordinate k of vertex 2^k if bit k
  // of i is set, else that of vertex 0
  {
    std::vector<unsigned char> box(h->nc);
    const double* V = mesh->vertices;
    for (int c = 0; c < h->nc; ++c) {
      const int32_t* cvr = mesh->cell_vertices + (size_t)c * nvc;
      bool ok = true;
      for (int i = 0; i < nvc && ok; ++i)
        for (int k = 0; k < dim; ++k) {
          const double want = ((i >> k) & 1) ? V[(size_t)k * h->nv + cvr[1 << k]] : V[(size_t)k * h->nv + cvr[0]];
          if (!(V[(size_t)k * h->nv + cvr[i]] == want)) ok = false;
        }
      box[c] = ok ? 1 : 0;
    }
    if (hipMemcpy(h->box, box.data(), h->nc, hipMemcpyHostToDevice)) return bail(fail(PFEM2GPU_ERR_HIP, "box"));
  }
  // box corners per cell (vertex 0 and vertex 2^dim-1) and the correctly
  // rounded reciprocal of each edge length; NaN in X_0[0] marks other cells
  {
    std::vector<double> geom((size_t)h->nc * 3 * dim);
    std::vector<unsigned char> boxh(h->nc);
    if (hipMemcpy(boxh.data(), h->box, h->nc, hipMemcpyDeviceToHost)) return bail(fail(PFEM2GPU_ERR_HIP, "box"));
    const double* V = mesh->vertices;
    for (int c = 0; c < h->nc; ++c) {
      const int32_t* cvr = mesh->cell_vertices + (size_t)c * nvc;
      double* g = &geom[(size_t)c * 3 * dim];
      for (int k = 0; k < dim; ++k) {
        g[k] = V[(size_t)k * h->nv + cvr[0]];
        g[dim + k] = V[(size_t)k * h->nv + cvr[nvc - 1]];
        g[2 * dim + k] = 1.0 / (g[dim + k] - g[k]);
      }
      if (!boxh[c]) g[0] = std::nan("");
    }
    h->all_box = std::all_of(boxh.begin(), boxh.end(), [](unsigned char b) { return b != 0; });
    if ((rc = dalloc(&h->geom, geom.size())) || (rc = dalloc(&h->Vc, (size_t)h->nc * nvc * dim)) ||
        (rc = dalloc(&h->dVc, (size_t)h->nc * nvc * dim)))
      return bail(rc);
    if (hipMemcpy(h->geom, geom.data(), geom.size() * 8, hipMemcpyHostToDevice))
      return bail(fail(PFEM2GPU_ERR_HIP, "upload geometry"));
  }
  // part decode table per instance: refillable part -> INT_MIN | pi_0 | pi_1 << 8 | pi_2 << 16
  {
    std::vector<int> tab((size_t)h->ninst * h->pmax, 0);
    for (int in = 0; in < h->ninst; ++in) {
      const Prm& q = prms[in];
      for (int pp = 0; pp < q.P; ++pp) {
        int rem = pp, code = 0;
        bool valid = true;
        for (int d = dim - 1; d >= 0; --d) {
          const int pk = rem % q.Q[d];
          rem /= q.Q[d];
          valid = valid && pk < q.q[d];
          code |= pk << (8 * d);
        }
        tab[(size_t)in * h->pmax + pp] = valid ? (int)(0x80000000u | (unsigned)code) : 0;
      }
    }
    if ((rc = dalloc(&h->ptab, tab.size()))) return bail(rc);
    if (hipMemcpy(h->ptab, tab.data(), tab.size() * 4, hipMemcpyHostToDevice))
      return bail(fail(PFEM2GPU_ERR_HIP, "upload part table"));
  }
  if (!cell_inst.empty()) {
    if ((rc = dalloc(&h->cell_inst, cell_inst.size()))) return bail(rc);
    if (hipMemcpy(h->cell_inst, cell_inst.data(), cell_inst.size() * 4, hipMemcpyHostToDevice))
      return bail(fail(PFEM2GPU_ERR_HIP, "upload cell instances"));
  }
  // dof -> (cell*nvc + i), ascending cell
  {
    std::vector<int> off(h->ndofs + 1, 0), lst((size_t)h->nc * nvc);
    for (long e = 0; e < (long)h->nc * nvc; ++e) off[cdofs[e] + 1]++;
    for (int j = 0; j < h->ndofs; ++j) off[j + 1] += off[j];
    std::vector<int> pos(off.begin(), off.end() - 1);
    for (long e = 0; e < (long)h->nc * nvc; ++e) lst[pos[cdofs[e]]++] = (int)e;
    std::vector<int> bp(lst.size());
    for (size_t e = 0; e < lst.size(); ++e) bp[lst[e]] = (int)e;
    if ((rc = dalloc(&h->d2c_off, h->ndofs + 1)) || (rc = dalloc(&h->d2c, lst.size())) ||
        (rc = dalloc(&h->bpos, bp.size())))
      return bail(rc);
    if (hipMemcpy(h->d2c_off, off.data(), off.size() * 4, hipMemcpyHostToDevice) ||
        hipMemcpy(h->d2c, lst.data(), lst.size() * 4, hipMemcpyHostToDevice) ||
        hipMemcpy(h->bpos, bp.data(), bp.size() * 4, hipMemcpyHostToDevice))
      return bail(fail(PFEM2GPU_ERR_HIP, "upload dof map"));
  }
  if ((rc = dalloc(&h->prm, prms.size()))) return bail(rc);
  if (hipMemcpy(h->prm, prms.data(), prms.size() * sizeof(Prm), hipMemcpyHostToDevice))
    return bail(fail(PFEM2GPU_ERR_HIP, "prm"));

  // per-cell work arrays
  h->scan_nb = (h->nc + kScanChunk - 1) / kScanChunk;
  if ((rc = dalloc(&h->cell_start, h->nc + 1)) ||
      (rc = dalloc(&h->count, 3 * (size_t)h->nchunks + 2)) || (rc = dalloc(&h->sstart, h->nchunks + 1)) ||
      (rc = dalloc(&h->lb, h->nchunks)) || (rc = dalloc(&h->d_err, 1)) ||
      (rc = dalloc(&h->scan_blk, h->scan_nb + 1)) ||
      (rc = dalloc(&h->contrib, (size_t)h->nc * nvc * (dim + 1))) || (rc = dalloc(&h->d_rejected, 1)) ||
      (rc = dalloc(&h->dirty, h->nchunks)))
    return bail(rc);
  for (const Prm& q : prms) h->qn_len = std::max(h->qn_len, q.n_sub + 2);
  if ((rc = dalloc(&h->qn, (size_t)h->qn_len * kQStripes))) return bail(rc);
  for (int k = 0; k < dim; ++k)
    if ((rc = dalloc(&h->stageF[0][k], h->ndofs)) ||
        (rc = dalloc(&h->stageF[1][k], h->ndofs)))
      return bail(rc);
  if (hipHostMalloc(reinterpret_cast<void**>(&h->h_pinned), sizeof(int) * 16, 0) != hipSuccess)
    return bail(fail(PFEM2GPU_ERR_MEMORY, "hipHostMalloc"));
  if (hipEventCreateWithFlags(&h->settle_ev, hipEventDisableTiming) != hipSuccess)
    return bail(fail(PFEM2GPU_ERR_HIP, "hipEventCreate"));
  if (hipMemsetAsync(h->d_err, 0, sizeof(int), h->stream) != hipSuccess ||
      hipMemsetAsync(h->cell_start, 0, (h->nc + 1) * sizeof(int), h->stream) != hipSuccess ||
      hipStreamSynchronize(h->stream) != hipSuccess)
    return bail(fail(PFEM2GPU_ERR_HIP, "init"));
  if ((rc = grow(h, 1024, 0, 256))) return bail(rc);
  if ((rc = scratch_for_current_device(&h->scratch))) return bail(rc);
  *out = h;
  return PFEM2GPU_OK;
}

extern "C" {

const char* pfem2gpu_last_error(void) { return g_err.c_str(); }

int pfem2gpu_create(const pfem2gpu_mesh* mesh, const pfem2gpu_params* params, void* hip_stream,
                    pfem2gpu_handle** out) {
  if (!mesh || !params || !out) return fail(PFEM2GPU_ERR_ARG, "null argument");
  *out = nullptr;
  CHK(validate_mesh(mesh, params));
  std::vector<Prm> prms(1, make_prm(mesh->dim, params));
  return create_impl(mesh, prms, {}, {0, mesh->n_cells}, {0, mesh->n_dofs}, hip_stream, out);
}

int pfem2gpu_batch_create(int32_t n_instances, const pfem2gpu_mesh* meshes, const pfem2gpu_params* params,
                          void* hip_stream, pfem2gpu_handle** out) {
  if (!meshes || !params || !out || n_instances <= 0) return fail(PFEM2GPU_ERR_ARG, "null argument");
  *out = nullptr;
  const int dim = meshes[0].dim, nvc = 1 << dim;
  long nv = 0, nc = 0, nd = 0;
  std::vector<long> coff(1, 0), doff(1, 0), voff(1, 0);
  for (int i = 0; i < n_instances; ++i) {
    CHK(validate_mesh(&meshes[i], &params[i]));
    if (meshes[i].dim != dim) return fail(PFEM2GPU_ERR_ARG, "all instances must have the same dim");
    nv += meshes[i].n_vertices;
    nc += meshes[i].n_cells;
    nd += meshes[i].n_dofs;
    voff.push_back(nv);
    coff.push_back(nc);
    doff.push_back(nd);
  }
  if (nc > 0x7fffffff / 8 || nv > 0x7fffffff || nd > 0x7fffffff)
    return fail(PFEM2GPU_ERR_ARG, "batch too large for int32 indexing");
  std::vector<double> X((size_t)dim * nv);
  std::vector<int32_t> cv((size_t)nc * nvc), cd((size_t)nc * nvc);
  std::vector<int> inst(nc);
  std::vector<Prm> prms;
  for (int i = 0; i < n_instances; ++i) {
    const pfem2gpu_mesh& m = meshes[i];
    for (int k = 0; k < dim; ++k)
      std::copy(m.vertices + (size_t)k * m.n_vertices, m.vertices + (size_t)(k + 1) * m.n_vertices,
                X.begin() + (size_t)k * nv + voff[i]);
    const int32_t* cdi = m.cell_dofs ? m.cell_dofs : m.cell_vertices;
    for (long e = 0; e < m.n_cells * nvc; ++e) {
      cv[(size_t)coff[i] * nvc + e] = m.cell_vertices[e] + (int32_t)voff[i];
      cd[(size_t)coff[i] * nvc + e] = cdi[e] + (int32_t)doff[i];
    }
    for (long c = 0; c < m.n_cells; ++c) inst[coff[i] + c] = i;
    prms.push_back(make_prm(dim, &params[i]));
  }
  pfem2gpu_mesh all{dim, nv, X.data(), nc, cv.data(), cd.data(), nd};
  return create_impl(&all, prms, inst, coff, doff, hip_stream, out);
}

int pfem2gpu_batch_layout(pfem2gpu_handle* h, int64_t* cell_offsets, int64_t* dof_offsets, int32_t* n_instances) {
  if (!h) return fail(PFEM2GPU_ERR_ARG, "null handle");
  if (n_instances) *n_instances = h->ninst;
  for (int i = 0; i <= h->ninst; ++i) {
    if (cell_offsets) cell_offsets[i] = h->inst_cell_off[i];
    if (dof_offsets) dof_offsets[i] = h->inst_dof_off[i];
  }
  return PFEM2GPU_OK;
}

int pfem2gpu_destroy(pfem2gpu_handle* h) {
  if (!h) return PFEM2GPU_OK;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  free_particles(h);
  for (auto& b : h->adj_pool) (void)hipFree(b.first);
  h->adj_pool.clear();
  for (int k = 0; k < 3; ++k) {
    dfree(h->X[k]);
    dfree(h->stageF[0][k]);
    dfree(h->stageF[1][k]);
  }
  dfree(h->cv);
  dfree(h->cd);
  dfree(h->box);
  dfree(h->v2c_off);
  dfree(h->v2c);
  dfree(h->v2c_dir);
  dfree(h->d2c_off);
  dfree(h->d2c);
  dfree(h->bpos);
  dfree(h->prm);
  dfree(h->cell_inst);
  dfree(h->geom);
  dfree(h->cvo);
  dfree(h->cvc);
  dfree(h->ptab);
  dfree(h->Vc);
  dfree(h->cell_start);
  dfree(h->count);
  dfree(h->sstart);
  dfree(h->lb);
  dfree(h->d_err);
  dfree(h->dVc);
  dfree(h->qn);
  dfree(h->scan_blk);
  dfree(h->contrib);
  dfree(h->d_rejected);
  dfree(h->dirty);
  dfree(h->fem_soff);
  dfree(h->fem_apos);
  dfree(h->fem_A);
  dfree(h->fem_b);
  dfree(h->fem_vals);
  for (int k = 0; k < 3; ++k) dfree(h->fem_rhs[k]);
  dfree(h->fem_P);
  dfree(h->fem_faces);
  if (h->h_pinned) (void)hipHostFree(h->h_pinned);
  if (h->settle_ev) (void)hipEventDestroy(h->settle_ev);
  for (auto& t : h->pending) {
    (void)hipEventDestroy(t.a);
    (void)hipEventDestroy(t.b);
  }
  for (auto e : h->pool) (void)hipEventDestroy(e);
  if (h->graph_exec) (void)hipGraphExecDestroy(h->graph_exec);
  if (h->graph) (void)hipGraphDestroy(h->graph);
  if (h->own_stream && h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return PFEM2GPU_OK;
}

// GridTools::vertex_to_cell_map + vertex_to_cell_centers_directions, built on
// the device (k_v2c_*). Centre = vertex sum * (1/n_vertices); direction
// normalised by multiplying with 1/norm (deal.II Tensor division).
int pfem2gpu_initialize_maps(pfem2gpu_handle* h) {
  if (!h) return fail(PFEM2GPU_ERR_ARG, "null handle");
  const int dim = h->dim, nvc = h->nvc, nv = h->nv;
  const long ne = (long)h->nc * nvc;
  const DMesh m = dmesh(h);
  // scratch: counts, offsets, fill cursors, the unsorted lists, scan blocks
  int *cnt = nullptr, *off = nullptr, *fill = nullptr, *lst = nullptr, *cnt2 = nullptr, *blk = nullptr;
  int* off2 = nullptr;
  const int nb = (nv + kScanChunk - 1) / kScanChunk;
  auto release = [&]() {
    dfree(cnt);
    dfree(off);
    dfree(fill);
    dfree(lst);
    dfree(cnt2);
    dfree(blk);
    dfree(off2);
  };
  int rc = PFEM2GPU_OK;
  if ((rc = dalloc(&cnt, nv + 1)) || (rc = dalloc(&off, nv + 1)) || (rc = dalloc(&fill, nv + 1)) ||
      (rc = dalloc(&lst, ne)) || (rc = dalloc(&cnt2, nv + 1)) || (rc = dalloc(&blk, nb + 1)) ||
      (rc = dalloc(&off2, nv + 1))) {
    release();
    return rc;
  }
  auto scan_into = [&](const int* in, int* out) -> int {  // exclusive; out[nv] = total
    LAUNCH(h, "scan", k_scan_reduce, nb, 0, in, nv, blk);
    LAUNCH(h, "scan", k_scan_blocks, 1, 0, blk, nb, out + nv);
    LAUNCH(h, "scan", k_scan_down, nb, 0, in, nv, blk, out);
    return PFEM2GPU_OK;
  };
  auto build = [&]() -> int {
    LAUNCH(h, "k_fill", k_zero3, grid_for(2L * (nv + 1)), 0, cnt, (long)nv + 1, fill, (long)nv + 1, nullptr, 0L);
    LAUNCH(h, "k_maps", k_v2c_count, grid_for(ne), 0, h->cv, ne, cnt);
    CHK(scan_into(cnt, off));
    LAUNCH(h, "k_maps", k_v2c_fill, grid_for(ne), 0, h->cv, ne, nvc, off, fill, lst);
    LAUNCH(h, "k_maps", k_v2c_sort, grid_for(nv), 0, nv, off, lst, cnt2);
    CHK(scan_into(cnt2, off2));
    int total = 0;
    HIPCHK(hipMemcpyAsync(&total, off2 + nv, sizeof(int), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    dfree(h->v2c_off);
    dfree(h->v2c);
    dfree(h->v2c_dir);
    dfree(h->cvo);
    dfree(h->cvc);
    CHK(dalloc(&h->v2c, total));
    CHK(dalloc(&h->v2c_dir, (size_t)total * dim));
    CHK(dalloc(&h->cvo, ne));
    CHK(dalloc(&h->cvc, (size_t)ne * (nvc / 4)));
    const double inv_n = 1.0 / (double)nvc;
    if (dim == 2)
      LAUNCH(h, "k_maps", k_v2c_emit<2>, grid_for(nv), 0, m, nv, off, lst, off2, h->v2c, h->v2c_dir, inv_n);
    else
      LAUNCH(h, "k_maps", k_v2c_emit<3>, grid_for(nv), 0, m, nv, off, lst, off2, h->v2c, h->v2c_dir, inv_n);
    LAUNCH(h, "k_maps", k_cell_vertex_maps, grid_for(ne), 0, ne, nvc, h->cv, off2, h->v2c, h->cvo, h->cvc);
    HIPCHK(hipStreamSynchronize(h->stream));
    h->v2c_off = off2;  // kept: the per-vertex offsets
    off2 = nullptr;
    return PFEM2GPU_OK;
  };
  rc = build();
  release();
  if (rc != PFEM2GPU_OK) return rc;
  h->maps_ready = true;
  return PFEM2GPU_OK;
}

int pfem2gpu_vertex_to_cells(pfem2gpu_handle* h, int32_t* offsets, int32_t* cells, double* dirs) {
  if (!h || !offsets) return fail(PFEM2GPU_ERR_ARG, "null argument");
  if (!h->maps_ready) return fail(PFEM2GPU_ERR_STATE, "vertex_to_cells before initialize_maps");
  HIPCHK(hipMemcpyAsync(offsets, h->v2c_off, (h->nv + 1) * sizeof(int), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  const int total = offsets[h->nv];
  if (cells && total > 0)
    HIPCHK(hipMemcpyAsync(cells, h->v2c, (size_t)total * sizeof(int), hipMemcpyDeviceToHost, h->stream));
  if (dirs && total > 0)
    HIPCHK(hipMemcpyAsync(dirs, h->v2c_dir, (size_t)total * h->dim * sizeof(double), hipMemcpyDeviceToHost,
                          h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  return PFEM2GPU_OK;
}

int pfem2gpu_insert_particles(pfem2gpu_handle* h, int64_t n, const double* x, const int32_t* cell, const double* v,
                              int where, int64_t* n_rejected) {
  if (!h) return fail(PFEM2GPU_ERR_ARG, "null handle");
  if (n_rejected) *n_rejected = 0;
  if (n < 0 || (n > 0 && (!x || !cell))) return fail(PFEM2GPU_ERR_ARG, "bad particle arrays");
  if (n == 0) return PFEM2GPU_OK;
  if (n > 0x7fffffffL) return fail(PFEM2GPU_ERR_ARG, "too many particles");
  if (where != PFEM2GPU_HOST && where != PFEM2GPU_DEVICE) return fail(PFEM2GPU_ERR_ARG, "bad where");
  CHK(flush(h));
  CHK(reserve(h, n, n));
  const int dim = h->dim;
  const double* sx[3] = {nullptr, nullptr, nullptr};
  const double* sv[3] = {nullptr, nullptr, nullptr};
  const int* sc = nullptr;
  if (where == PFEM2GPU_DEVICE) {
    for (int k = 0; k < dim; ++k) {
      sx[k] = x + (size_t)k * n;
      if (v) sv[k] = v + (size_t)k * n;
    }
    sc = cell;
  } else {
    // stage host arrays in the T buffers (free between steps)
    for (int k = 0; k < dim; ++k) {
      HIPCHK(hipMemcpyAsync(h->tx[k], x + (size_t)k * n, n * sizeof(double), hipMemcpyHostToDevice, h->stream));
      sx[k] = h->tx[k];
      if (v) {
        HIPCHK(hipMemcpyAsync(h->tv[k], v + (size_t)k * n, n * sizeof(double), hipMemcpyHostToDevice, h->stream));
        sv[k] = h->tv[k];
      }
    }
    HIPCHK(hipMemcpyAsync(h->tcell, cell, n * sizeof(int), hipMemcpyHostToDevice, h->stream));
    sc = h->tcell;
    HIPCHK(hipStreamSynchronize(h->stream));
  }
  // the staged T arrays are consumed by k_append before the re-bin reuses T
  return append_staged(h, n, sx, v ? sv : nullptr, sc, n_rejected);
}

int pfem2gpu_seed_particles(pfem2gpu_handle* h, const double* const* V, int where) {
  if (!h) return fail(PFEM2GPU_ERR_ARG, "null handle");
  Section sec(h, "Particle seeding");
  DField F;
  CHK(field_in(h, V, where, 0, &F));
  CHK(flush(h));
  for (const Prm& q : h->prm_inst)
    for (int k = 0; k < h->dim; ++k)
      if (q.q[k] != h->prm_host.q[k])
        return fail(PFEM2GPU_ERR_ARG, "seed_particles on a batch needs the same particles_per_cell in every instance");
  const long n = (long)h->nc * h->prm_host.Pvalid;
  if (n > 0x7fffffffL) return fail(PFEM2GPU_ERR_ARG, "too many particles");
  CHK(reserve(h, n, n));
  const DMesh m = dmesh(h);
  if (h->dim == 2)
    LAUNCH(h, "k_seed", k_seed<2>, grid_for(n), 0, m, h->nc, h->prm_host.Pvalid, F, h->tx[0], h->tx[1], h->tx[2],
           h->tv[0], h->tv[1], h->tv[2], h->tcell);
  else
    LAUNCH(h, "k_seed", k_seed<3>, grid_for(n), 0, m, h->nc, h->prm_host.Pvalid, F, h->tx[0], h->tx[1], h->tx[2],
           h->tv[0], h->tv[1], h->tv[2], h->tcell);
  const double* sx[3] = {h->tx[0], h->tx[1], h->tx[2]};
  const double* sv[3] = {h->tv[0], h->tv[1], h->tv[2]};
  return append_staged(h, n, sx, sv, h->tcell, nullptr);
}

int pfem2gpu_clear_particles(pfem2gpu_handle* h) {
  if (!h) return fail(PFEM2GPU_ERR_ARG, "null handle");
  drop_graph(h);
  HIPCHK(hipMemsetAsync(h->cell_start, 0, (h->nc + 1) * sizeof(int), h->stream));
  h->pending_corr = false;
  h->contrib_valid = false;
  h->n_ub = 0;
  h->settle_ev_valid = false;
  return PFEM2GPU_OK;
}

int pfem2gpu_reserve_particles(pfem2gpu_handle* h, int64_t n) {
  if (!h) return fail(PFEM2GPU_ERR_ARG, "null handle");
  if (n < 0 || n > 0x7fffffffL) return fail(PFEM2GPU_ERR_ARG, "n out of range");
  long cur = 0;
  CHK(read_count(h, &cur));
  const long need = std::max<long>(n, cur) + h->refill_max;
  const int capc = capc_for(h, std::max<long>(n, cur));
  if (need <= h->cap && capc <= h->capc) return PFEM2GPU_OK;
  return grow(h, std::max(need, h->cap), cur, capc);
}

int pfem2gpu_set_velocity_from_field(pfem2gpu_handle* h, const double* const* V, int where) {
  if (!h) return fail(PFEM2GPU_ERR_ARG, "null handle");
  DField F;
  CHK(field_in(h, V, where, 0, &F));
  h->pending_corr = false;  // overwritten anyway
  h->contrib_valid = false;
  if (h->dim == 2)
    LAUNCH(h, "k_setvel", k_setvel<2>, grid_for(h->n_ub), 0, dmesh(h), h->cell_start + h->nc, dparts(h), F);
  else
    LAUNCH(h, "k_setvel", k_setvel<3>, grid_for(h->n_ub), 0, dmesh(h), h->cell_start + h->nc, dparts(h), F);
  return PFEM2GPU_OK;
}

int pfem2gpu_correct_particle_velocity(pfem2gpu_handle* h, const double* const* V, const double* const* V_old,
                                       int where) {
  if (!h) return fail(PFEM2GPU_ERR_ARG, "null handle");
  Section sec(h, "Particle velocity correction");
  CHK(flush(h));
  DField A, B;
  CHK(field_in(h, V, where, 0, &A));
  CHK(field_in(h, V_old, where, 1, &B));
  if (h->dim == 2)
    LAUNCH(h, "k_sub", k_subc<2>, grid_for((long)h->nc * 4), 0, h->nc, h->cd, A, B, h->dVc);
  else
    LAUNCH(h, "k_sub", k_subc<3>, grid_for((long)h->nc * 8), 0, h->nc, h->cd, A, B, h->dVc);
  h->pending_corr = true;
  return PFEM2GPU_OK;
}

int pfem2gpu_move_particles(pfem2gpu_handle* h, const double* const* V, int where) {
  if (!h) return fail(PFEM2GPU_ERR_ARG, "null handle");
  if (!h->maps_ready) return fail(PFEM2GPU_ERR_STATE, "initialize_maps() must be called before move_particles()");
  Section sec(h, "Particle transport");
  DField F;
  CHK(field_in(h, V, where, 0, &F));
  CHK(reserve(h, h->refill_max, 0));
  return do_rebin(h, 1, 0, F);
}

int pfem2gpu_project_particle_fields(pfem2gpu_handle* h, double* const* V_out, int where) {
  if (!h) return fail(PFEM2GPU_ERR_ARG, "null handle");
  if (!V_out) return fail(PFEM2GPU_ERR_ARG, "null output");
  if (where != PFEM2GPU_HOST && where != PFEM2GPU_DEVICE) return fail(PFEM2GPU_ERR_ARG, "bad where");
  Section sec(h, "Particle fields projection");
  CHK(flush(h));
  const long ne = (long)h->nc * h->nvc;
  double* out[3] = {nullptr, nullptr, nullptr};
  for (int k = 0; k < h->dim; ++k) out[k] = (where == PFEM2GPU_DEVICE) ? V_out[k] : h->stageF[1][k];
  const int nchunks = h->nchunks;
  if (h->dim == 2) {
    if (h->contrib_valid)
      LAUNCH(h, "k_contrib_dirty", k_contrib_dirty<2>, std::min((nchunks + kBlock - 1) / kBlock, 4096), 0, dmesh(h), h->nc, h->cpw,
             h->dirty, h->cell_start, dparts(h), h->contrib);
    else
      LAUNCH(h, "k_contrib", k_contrib<2>, grid_for(ne), 0, dmesh(h), h->nc, h->cell_start, dparts(h), h->contrib);
    LAUNCH(h, "k_node", k_node<2>, grid_for(h->ndofs), 0, h->ndofs, h->d2c_off, h->contrib, out[0], out[1],
           out[2]);
  } else {
    if (h->contrib_valid)
      LAUNCH(h, "k_contrib_dirty", k_contrib_dirty<3>, std::min((nchunks + kBlock - 1) / kBlock, 4096), 0, dmesh(h), h->nc, h->cpw,
             h->dirty, h->cell_start, dparts(h), h->contrib);
    else
      LAUNCH(h, "k_contrib", k_contrib<3>, grid_for(ne), 0, dmesh(h), h->nc, h->cell_start, dparts(h), h->contrib);
    LAUNCH(h, "k_node", k_node<3>, grid_for(h->ndofs), 0, h->ndofs, h->d2c_off, h->contrib, out[0], out[1],
           out[2]);
  }
  h->contrib_valid = true;
  (void)ne;
  if (where == PFEM2GPU_HOST) {
    for (int k = 0; k < h->dim; ++k)
      HIPCHK(hipMemcpyAsync(V_out[k], out[k], h->ndofs * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
  }
  return PFEM2GPU_OK;
}

int pfem2gpu_particle_count(pfem2gpu_handle* h, int64_t* n) {
  if (!h || !n) return fail(PFEM2GPU_ERR_ARG, "null argument");
  long c = 0;
  CHK(read_count(h, &c));
  *n = c;
  return PFEM2GPU_OK;
}

int pfem2gpu_particle_count_async(pfem2gpu_handle* h, int32_t* dst) {
  if (!h || !dst) return fail(PFEM2GPU_ERR_ARG, "null argument");
  HIPCHK(hipMemcpyAsync(dst, h->cell_start + h->nc, sizeof(int), hipMemcpyDeviceToDevice, h->stream));
  return PFEM2GPU_OK;
}

int pfem2gpu_download(pfem2gpu_handle* h, double* x, double* xi, double* v, int32_t* cell) {
  if (!h) return fail(PFEM2GPU_ERR_ARG, "null handle");
  CHK(flush(h));
  long n = 0;
  CHK(read_count(h, &n));
  if (n == 0) return PFEM2GPU_OK;
  for (int k = 0; k < h->dim; ++k) {
    if (x) HIPCHK(hipMemcpyAsync(x + (size_t)k * n, h->px[k], n * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    if (v) HIPCHK(hipMemcpyAsync(v + (size_t)k * n, h->pv[k], n * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  }
  if (cell) HIPCHK(hipMemcpyAsync(cell, h->pcell, n * sizeof(int), hipMemcpyDeviceToHost, h->stream));
  if (xi) {
    if (h->dim == 2)
      LAUNCH(h, "k_xi", k_xi<2>, grid_for(n), 0, dmesh(h), h->cell_start + h->nc, dparts(h), h->tx[0], h->tx[1],
             h->tx[2]);
    else
      LAUNCH(h, "k_xi", k_xi<3>, grid_for(n), 0, dmesh(h), h->cell_start + h->nc, dparts(h), h->tx[0], h->tx[1],
             h->tx[2]);
    for (int k = 0; k < h->dim; ++k)
      HIPCHK(hipMemcpyAsync(xi + (size_t)k * n, h->tx[k], n * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  }
  HIPCHK(hipStreamSynchronize(h->stream));
  if (cell)  // the stored words also carry the placed flag
    for (long i = 0; i < n; ++i) cell[i] &= kCellMask;
  return PFEM2GPU_OK;
}

int pfem2gpu_write_particles_vtu(const char* path, int32_t dim, int64_t n, const double* x, const double* v,
                                 int32_t subdomain) {
  if (!path || dim < 1 || dim > 3 || n < 0 || (n > 0 && (!x || !v)))
    return fail(PFEM2GPU_ERR_ARG, "write_particles_vtu: bad arguments");
  if (n > 0xffffffffLL) return fail(PFEM2GPU_ERR_ARG, "write_particles_vtu: more than 2^32-1 particles");
  if (!pfem2gpu_out::write_vtu(path, dim, n, x, v, subdomain))
    return fail(PFEM2GPU_ERR_IO, std::string("cannot write ") + path);
  return PFEM2GPU_OK;
}

int pfem2gpu_write_particles_pvtu(const char* path, int32_t timestep_number, int32_t n_ranks) {
  if (!path || n_ranks < 1) return fail(PFEM2GPU_ERR_ARG, "write_particles_pvtu: bad arguments");
  if (!pfem2gpu_out::write_pvtu(path, timestep_number, n_ranks))
    return fail(PFEM2GPU_ERR_IO, std::string("cannot write ") + path);
  return PFEM2GPU_OK;
}

int pfem2gpu_output_particle_solution(pfem2gpu_handle* h, int32_t timestep_number, int32_t rank, int32_t n_ranks,
                                      int32_t subdomain, const char* directory) {
  if (!h || !directory) return fail(PFEM2GPU_ERR_ARG, "null argument");
  if (rank < 0 || n_ranks < 1 || rank >= n_ranks) return fail(PFEM2GPU_ERR_ARG, "rank out of range");
  int64_t n = 0;
  CHK(pfem2gpu_particle_count(h, &n));
  std::vector<double> x((size_t)h->dim * n), v((size_t)h->dim * n);
  CHK(pfem2gpu_download(h, x.data(), nullptr, v.data(), nullptr));
  std::string dir(directory);
  if (!dir.empty() && dir.back() != '/') dir += '/';
  const std::string step = pfem2gpu_out::padded(timestep_number, 2);
  CHK(pfem2gpu_write_particles_vtu((dir + "particles-" + step + "." + pfem2gpu_out::padded(rank, 3) + ".vtu").c_str(),
                                   h->dim, n, x.data(), v.data(), subdomain));
  if (rank == 0) CHK(pfem2gpu_write_particles_pvtu((dir + "particles-" + step + ".pvtu").c_str(), timestep_number, n_ranks));
  return PFEM2GPU_OK;
}

int pfem2gpu_cell_counts(pfem2gpu_handle* h, int32_t* counts) {
  if (!h || !counts) return fail(PFEM2GPU_ERR_ARG, "null argument");
  std::vector<int> cs(h->nc + 1);
  HIPCHK(hipMemcpyAsync(cs.data(), h->cell_start, (h->nc + 1) * sizeof(int), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  for (int c = 0; c < h->nc; ++c) counts[c] = cs[c + 1] - cs[c];
  return PFEM2GPU_OK;
}

static int step_eager(pfem2gpu_handle* h, const double* const* V, double* const* Vp) {
  CHK(pfem2gpu_move_particles(h, V, PFEM2GPU_DEVICE));
  CHK(pfem2gpu_project_particle_fields(h, Vp, PFEM2GPU_DEVICE));
  CHK(pfem2gpu_correct_particle_velocity(h, V, Vp, PFEM2GPU_DEVICE));
  return PFEM2GPU_OK;
}

int pfem2gpu_step(pfem2gpu_handle* h, const double* const* V, double* const* V_proj, int use_graph) {
  if (!h || !V || !V_proj) return fail(PFEM2GPU_ERR_ARG, "null argument");
  for (int k = 0; k < h->dim; ++k)
    if (!V[k] || !V_proj[k]) return fail(PFEM2GPU_ERR_ARG, "null field component");
  if (!h->maps_ready) return fail(PFEM2GPU_ERR_STATE, "initialize_maps() must be called before pfem2gpu_step()");
  // the captured sequence starts from a pending correction; the first step
  // of a run (none pending) and timed runs go eagerly
  if (!use_graph || !h->pending_corr || h->timing) return step_eager(h, V, V_proj);
  bool same = h->graph_exec != nullptr;
  for (int k = 0; k < h->dim; ++k) same = same && h->graph_V[k] == V[k] && h->graph_Vp[k] == V_proj[k];
  // capacity for this replay (a grow moves buffers and forces a new capture)
  if (h->n_ub + h->refill_max > h->cap || capc_for(h, h->n_ub) > h->capc) {
    double* before = h->px[0];
    const long tbefore = h->tsize;
    CHK(reserve(h, h->refill_max, 0));
    if (h->px[0] != before || h->tsize != tbefore) same = false;
  }
  if (!same) {
    drop_graph(h);
    CHK(grow_queues(h, walk_qreg(h, h->cap) * kQStripes));  // the captured stripe size (see rebin_or_move)
    const long n_ub0 = h->n_ub;
    h->capturing = true;
    HIPCHK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    const int rc = step_eager(h, V, V_proj);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(h->stream, &g);
    h->capturing = false;
    if (rc != PFEM2GPU_OK) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) return fail(PFEM2GPU_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    h->graph = g;
    HIPCHK(hipGraphInstantiate(&h->graph_exec, h->graph, nullptr, nullptr, 0));
    h->n_ub = n_ub0;  // the capture only recorded the step
    for (int k = 0; k < 3; ++k) {
      h->graph_V[k] = k < h->dim ? V[k] : nullptr;
      h->graph_Vp[k] = k < h->dim ? V_proj[k] : nullptr;
    }
  }
  HIPCHK(hipGraphLaunch(h->graph_exec, h->stream));
  h->settle_ev_valid = false;  // the graph's settle records no count
  h->n_ub += h->refill_max;
  h->pending_corr = true;
  h->contrib_valid = false;
  return PFEM2GPU_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ adjoint

namespace {

// Device buffers of one adjoint call, taken from the handle's pool of
// buffers of earlier calls (best fit) and returned to it at the end, so that
// repeated calls allocate nothing once the pool holds what they need.
struct DevBuf {
  pfem2gpu_handle* h;
  std::vector<std::pair<void*, size_t>> taken;
  explicit DevBuf(pfem2gpu_handle* hh) : h(hh) {}
  ~DevBuf() {
    for (auto& t : taken) h->adj_pool.push_back(t);
  }
  template <typename T>
  int get(T** p, size_t n) {
    const size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
    int best = -1;
    for (int i = 0; i < (int)h->adj_pool.size(); ++i)
      if (h->adj_pool[i].second >= bytes && (best < 0 || h->adj_pool[i].second < h->adj_pool[best].second)) best = i;
    if (best >= 0) {
      *p = static_cast<T*>(h->adj_pool[best].first);
      taken.push_back(h->adj_pool[best]);
      h->adj_pool.erase(h->adj_pool.begin() + best);
      return PFEM2GPU_OK;
    }
    CHK(dalloc(p, std::max<size_t>(n, 1)));
    taken.push_back({static_cast<void*>(*p), bytes});
    return PFEM2GPU_OK;
  }
};

struct Snapshot {
  long n = 0;
  double* x[3] = {nullptr, nullptr, nullptr};
  double* v[3] = {nullptr, nullptr, nullptr};
  int* cell = nullptr;
  int* cs = nullptr;
};

// Copies the current particles (count read on the device) and cell_start;
// s.n is filled in by the caller once the counts are read back.
int snap_save(pfem2gpu_handle* h, DevBuf& mem, Snapshot& s) {
  const size_t n = std::max<long>(h->cap, 1);
  for (int k = 0; k < 3; ++k) {
    s.x[k] = nullptr;
    s.v[k] = nullptr;
  }
  for (int k = 0; k < h->dim; ++k) {
    CHK(mem.get(&s.x[k], n));
    CHK(mem.get(&s.v[k], n));
  }
  CHK(mem.get(&s.cell, n));
  CHK(mem.get(&s.cs, h->nc + 1));
  LAUNCH(h, "k_copy", k_copy_state, grid_for(h->n_ub), 0, h->cell_start + h->nc, h->px[0], h->px[1],
         h->dim > 2 ? h->px[2] : nullptr, h->pv[0], h->pv[1], h->dim > 2 ? h->pv[2] : nullptr, h->pcell, s.x[0],
         s.x[1], s.x[2], s.v[0], s.v[1], s.v[2], s.cell);
  HIPCHK(hipMemcpyAsync(s.cs, h->cell_start, (h->nc + 1) * 4, hipMemcpyDeviceToDevice, h->stream));
  return PFEM2GPU_OK;
}

int snap_restore(pfem2gpu_handle* h, const Snapshot& s) {
  for (int k = 0; k < h->dim; ++k) {
    HIPCHK(hipMemcpyAsync(h->px[k], s.x[k], s.n * 8, hipMemcpyDeviceToDevice, h->stream));
    HIPCHK(hipMemcpyAsync(h->pv[k], s.v[k], s.n * 8, hipMemcpyDeviceToDevice, h->stream));
  }
  HIPCHK(hipMemcpyAsync(h->pcell, s.cell, s.n * 4, hipMemcpyDeviceToDevice, h->stream));
  HIPCHK(hipMemcpyAsync(h->cell_start, s.cs, (h->nc + 1) * 4, hipMemcpyDeviceToDevice, h->stream));
  h->n_ub = s.n;
  h->settle_ev_valid = false;
  h->pending_corr = false;
  h->contrib_valid = false;
  return PFEM2GPU_OK;
}

// Replay of a recorded move (the moves of a checkpoint segment recomputed
// in the reverse sweep): the move kernels run as in the forward (same
// inputs, same path records) but store each remaining particle at its source
// index in the staging arrays; k_adj_replay then builds the post-step state
// from the source map and end cells the forward recorded, so that the
// redistribution (spill, scan, settle) is not run again. `n_in` is the count
// at the start of step s (dcnt[s - 1] on the device); with `n_out` = 0 the
// post-step state is not built (the segment's last move).
template <int D>
int replay_move(pfem2gpu_handle* h, const int* dcnt, int s, long n_in, long n_out, const int* src, const int* ecell) {
  const DMesh m = dmesh(h);
  const DParts P = dparts(h);
  DTmp M = dtmp(h);
  M.src = nullptr;
  int rounds = 0;
  for (const Prm& q : h->prm_inst) rounds = std::max(rounds, q.n_sub);
  const long gc = move_cell_grid(h, n_in);
  const long qreg = walk_qreg(h, n_in);
  if (qreg * kQStripes > h->qcap) CHK(grow_queues(h, qreg * kQStripes));
  if (n_in > h->tsize) return fail(PFEM2GPU_ERR_STATE, "staging arrays smaller than the particle count");
  LAUNCH(h, "k_fill", k_fill_int, grid_for((long)h->qn_len * kQStripes), 0, h->qn, (long)h->qn_len * kQStripes, 0);
  MOVE_LAUNCH(h, "k_move", k_move_cell, D, h->move_block, gc, 1, m, h->Vc, dcnt + s - 1, 0, P, h->dVc, 0, h->cshift,
              -1, nullptr, 0L, nullptr, M, dwalk(h, 0), h->qn, qreg, h->rec_path);
  const int gw = kQStripes * std::max(1, h->walk_grid / kQStripes);
  for (int r = 0; r < rounds; ++r)
    MOVE_LAUNCH(h, "k_walk", k_move_walk, D, h->walk_block, gw, m, h->Vc, dwalk(h, r & 1),
                h->qn + (long)r * kQStripes, qreg, h->cshift, -1, nullptr, 0L, nullptr, M, dwalk(h, (r + 1) & 1),
                h->qn + (long)(r + 1) * kQStripes, h->rec_path);
  if (n_out > 0) {
    LAUNCH(h, "k_adj_replay", k_adj_replay<D>, grid_for(n_out), 0, m, h->Vc, h->pshift, (int)n_out, src, ecell, M, P,
           h->cell_start + h->nc);
    h->n_ub = n_out;
    h->settle_ev_valid = false;
  }
  return PFEM2GPU_OK;
}

template <int D>
int adjoint_impl(pfem2gpu_handle* h, int n_steps, int K, const double* const* Vu, const double* const* Tu,
                 double* const* gv0, double* const* gV, double* loss) {
  DevBuf mem(h);
  const int nc = h->nc, nd = h->ndofs;
  DField V, Tg;
  for (int k = 0; k < 3; ++k) {
    V.c[k] = k < D ? Vu[k] : nullptr;
    Tg.c[k] = k < D ? Tu[k] : nullptr;
  }
  CHK(flush(h));
  drop_graph(h);
  if (!h->tsrc) CHK(dalloc(&h->tsrc, h->tsize));
  // forward: checkpoints every K steps, identity maps and end cells of every step
  const int nseg = (n_steps + K - 1) / K;
  std::vector<Snapshot> ck(nseg);
  std::vector<int*> srcs(n_steps + 1, nullptr), ecell(n_steps + 1, nullptr);
  std::vector<long> cnt(n_steps + 1, 0);
  int max_sub = 1;
  for (const Prm& q : h->prm_inst) max_sub = std::max(max_sub, q.n_sub);
  // paths of one segment's moves (the last segment's are recorded by the
  // forward itself, the others by their recomputation); the capacity is
  // secured for a whole segment first, so that it bounds every count the
  // paths see
  const int last0 = (nseg - 1) * K + 1;
  std::vector<DPath> paths(K);
  int* dcnt = nullptr;  // particle count after each step (device)
  CHK(mem.get(&dcnt, n_steps + 1));
  HIPCHK(hipMemcpyAsync(dcnt, h->cell_start + h->nc, 4, hipMemcpyDeviceToDevice, h->stream));
  for (int s = 1; s <= n_steps; ++s) {
    if ((s - 1) % K == 0) CHK(snap_save(h, mem, ck[(s - 1) / K]));
    if (s == last0) {
      CHK(reserve(h, (long)(K + 1) * h->refill_max, 0));
      for (int j = 0; j < K; ++j) {
        paths[j].stride = h->cap;
        CHK(mem.get(&paths[j].c, (size_t)max_sub * h->cap));
        for (int k = 0; k < 3; ++k) paths[j].xi[k] = nullptr;
        for (int k = 0; k < D; ++k) CHK(mem.get(&paths[j].xi[k], (size_t)max_sub * h->cap));
      }
    }
    CHK(reserve(h, h->refill_max, 0));
    if (!h->tsrc || h->tsize == 0) return fail(PFEM2GPU_ERR_STATE, "staging arrays missing");
    CHK(mem.get(&srcs[s], h->cap));
    h->rec_src = srcs[s];
    if (s >= last0) h->rec_path = paths[s - last0];
    const int rc = rebin_or_move<D>(h, 1, 0, V);
    h->rec_src = nullptr;
    h->rec_path = DPath{nullptr, {nullptr, nullptr, nullptr}, 0};
    CHK(rc);
    HIPCHK(hipMemcpyAsync(dcnt + s, h->cell_start + h->nc, 4, hipMemcpyDeviceToDevice, h->stream));
    CHK(mem.get(&ecell[s], h->cap));
    LAUNCH(h, "k_copy", k_copy_state, grid_for(h->n_ub), 0, h->cell_start + h->nc, nullptr, nullptr, nullptr,
           nullptr, nullptr, nullptr, h->pcell, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, ecell[s]);
  }
  Snapshot fin;
  CHK(snap_save(h, mem, fin));
  {  // the counts of every step, read back once
    std::vector<int> hc(n_steps + 1, 0);
    HIPCHK(hipMemcpyAsync(hc.data(), dcnt, (n_steps + 1) * sizeof(int), hipMemcpyDeviceToHost, h->stream));
    CHK(check_device(h));
    for (int s = 0; s <= n_steps; ++s) cnt[s] = hc[s];
    for (int seg = 0; seg < nseg; ++seg) ck[seg].n = cnt[seg * K];
    fin.n = cnt[n_steps];
    h->n_ub = cnt[n_steps];
  }
  // final projection, loss and the projection adjoint
  double *Pn[3] = {nullptr, nullptr, nullptr}, *W = nullptr, *dloss = nullptr;
  unsigned long long* floss = nullptr;  // the loss as a fixed-point sum
  for (int k = 0; k < D; ++k) CHK(mem.get(&Pn[k], nd));
  CHK(mem.get(&W, nd));
  CHK(mem.get(&dloss, 1));
  CHK(mem.get(&floss, 2));
  const DMesh m = dmesh(h);
  LAUNCH(h, "k_contrib_dirty", k_contrib_dirty<D>, std::min((h->nchunks + kBlock - 1) / kBlock, 4096), 0, m, nc, h->cpw, h->dirty,
         h->cell_start, dparts(h), h->contrib);
  LAUNCH(h, "k_node_w", k_node_w<D>, grid_for(nd), 0, nd, h->d2c_off, h->contrib, Pn[0], Pn[1], Pn[2], W);
  LAUNCH(h, "k_fill", k_fill_u64, 1, 0, floss, 2L);
  DField Pf;
  for (int k = 0; k < 3; ++k) Pf.c[k] = Pn[k];
  LAUNCH(h, "k_loss", k_loss<D>, std::min(grid_for(nd), 1024), 0, nd, Pf, Tg, floss, h->d_err);
  LAUNCH(h, "k_loss", k_fix_to_double, 1, 0, 1L, floss, dloss);
  long nmax = 1;
  for (long c : cnt) nmax = std::max(nmax, c);
  double *xa[3] = {nullptr, nullptr, nullptr}, *va[3] = {nullptr, nullptr, nullptr};
  double *xb[3] = {nullptr, nullptr, nullptr}, *vb[3] = {nullptr, nullptr, nullptr};
  for (int k = 0; k < D; ++k) {
    CHK(mem.get(&xa[k], nmax));
    CHK(mem.get(&va[k], nmax));
    CHK(mem.get(&xb[k], nmax));
    CHK(mem.get(&vb[k], nmax));
  }
  LAUNCH(h, "k_adj_init", k_adj_init<D>, grid_for(cnt[n_steps]), 0, m, h->cell_start + nc, dparts(h), Pf, W, Tg,
         xa[0], xa[1], xa[2], va[0], va[1], va[2]);
  for (int k = 0; k < D; ++k) LAUNCH(h, "k_fill", k_fill_double, grid_for(nd), 0, gV[k], (long)nd);
  unsigned long long* gcv = nullptr;  // dL/dV per (cell, vertex) in fixed point, reduced to the DoFs at the end
  CHK(mem.get(&gcv, 2 * (size_t)nc * h->nvc * D));
  LAUNCH(h, "k_fill", k_fill_u64, grid_for(2L * nc * h->nvc * D), 0, gcv, 2L * nc * h->nvc * D);
  // backward, one checkpoint segment at a time (latest first): recompute the
  // segment's moves from its checkpoint (all but the last segment, whose
  // paths the forward recorded), recording every particle's path, then
  // reverse its steps
  for (int seg = nseg - 1; seg >= 0; --seg) {
    const int s0 = seg * K + 1, s1 = std::min(n_steps, seg * K + K);
    if (seg < nseg - 1) CHK(snap_restore(h, ck[seg]));
    for (int s = s0; s <= s1 && seg < nseg - 1; ++s) {
      h->rec_path = paths[s - s0];
      const int rc = replay_move<D>(h, dcnt, s, cnt[s - 1], s < s1 ? cnt[s] : 0, srcs[s], ecell[s]);
      h->rec_path = DPath{nullptr, {nullptr, nullptr, nullptr}, 0};
      CHK(rc);
    }
    for (int s = s1; s >= s0; --s) {
      LAUNCH(h, "k_fill", k_fill_adj, grid_for(cnt[s - 1]), 0, cnt[s - 1], xb[0], xb[1], xb[2], vb[0], vb[1], vb[2]);
      LAUNCH(h, "k_adj_unmap", k_adj_unmap<D>, grid_for(cnt[s]), 0, m, h->pshift, (int)cnt[s], srcs[s], ecell[s],
             xa[0], xa[1], xa[2], va[0], va[1], va[2], xb[0], xb[1], xb[2], vb[0], vb[1], vb[2], gcv, h->d_err);
      // (the path held for five substeps when that covers every instance,
      // the workloads' count; else for kAdjMaxSub)
      if (max_sub <= 5)
        LAUNCH(h, "k_adj_move", (k_adj_move<D, 5>), grid_for(cnt[s - 1]), 0, m, h->Vc, (int)cnt[s - 1], max_sub,
               paths[s - s0], xb[0], xb[1], xb[2], gcv, h->d_err);
      else
        LAUNCH(h, "k_adj_move", (k_adj_move<D, kAdjMaxSub>), grid_for(cnt[s - 1]), 0, m, h->Vc, (int)cnt[s - 1],
               max_sub, paths[s - s0], xb[0], xb[1], xb[2], gcv, h->d_err);
      for (int k = 0; k < D; ++k) {
        std::swap(xa[k], xb[k]);
        std::swap(va[k], vb[k]);
      }
    }
  }
  LAUNCH(h, "k_adj_gather", k_adj_gather<D>, grid_for(nd), 0, nd, h->d2c_off, h->d2c, gcv, gV[0], gV[1], gV[2]);
  for (int k = 0; k < D; ++k)
    HIPCHK(hipMemcpyAsync(gv0[k], va[k], cnt[0] * 8, hipMemcpyDeviceToDevice, h->stream));
  HIPCHK(hipMemcpyAsync(h->h_pinned + 2, dloss, 8, hipMemcpyDeviceToHost, h->stream));
  CHK(check_device(h));
  std::memcpy(loss, h->h_pinned + 2, 8);
  CHK(snap_restore(h, fin));
  CHK(check_device(h));
  return PFEM2GPU_OK;
}

}  // namespace

extern "C" {

int pfem2gpu_adjoint(pfem2gpu_handle* h, int32_t n_steps, int32_t checkpoint_every, const double* const* V,
                     const double* const* target, double* const* grad_v0, double* const* grad_V, double* loss) {
  if (!h || !V || !target || !grad_v0 || !grad_V || !loss) return fail(PFEM2GPU_ERR_ARG, "null argument");
  if (n_steps < 1 || checkpoint_every < 1) return fail(PFEM2GPU_ERR_ARG, "n_steps and checkpoint_every must be >= 1");
  if (!h->maps_ready) return fail(PFEM2GPU_ERR_STATE, "initialize_maps() must be called before the adjoint");
  if (!h->all_box) return fail(PFEM2GPU_ERR_ARG, "the adjoint supports axis-aligned box cells only");
  for (const Prm& q : h->prm_inst)
    if (q.n_sub > kAdjMaxSub) return fail(PFEM2GPU_ERR_ARG, "the adjoint supports at most 8 substeps");
  return h->dim == 2 ? adjoint_impl<2>(h, n_steps, checkpoint_every, V, target, grad_v0, grad_V, loss)
                     : adjoint_impl<3>(h, n_steps, checkpoint_every, V, target, grad_v0, grad_V, loss);
}

int pfem2gpu_move_particles_traced(pfem2gpu_handle* h, const double* const* V, int where, int32_t* src,
                                   int64_t src_capacity, int64_t* n_out) {
  if (!h || !src || !n_out) return fail(PFEM2GPU_ERR_ARG, "null argument");
  if (!h->maps_ready) return fail(PFEM2GPU_ERR_STATE, "initialize_maps() must be called before move_particles()");
  DField F;
  CHK(field_in(h, V, where, 0, &F));
  CHK(reserve(h, h->refill_max, 0));
  drop_graph(h);
  if (!h->tsrc) CHK(dalloc(&h->tsrc, h->tsize));
  int* d = nullptr;
  CHK(dalloc(&d, h->cap));
  h->rec_src = d;
  int rc = do_rebin(h, 1, 0, F);
  h->rec_src = nullptr;
  long n = 0;
  if (rc == PFEM2GPU_OK) rc = read_count(h, &n);
  if (rc == PFEM2GPU_OK && n > src_capacity) rc = fail(PFEM2GPU_ERR_ARG, "src capacity too small");
  if (rc == PFEM2GPU_OK && hipMemcpy(src, d, n * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)
    rc = fail(PFEM2GPU_ERR_HIP, "copy of the identity map failed");
  (void)hipFree(d);
  *n_out = n;
  return rc;
}

int pfem2gpu_set_tuning(pfem2gpu_handle* h, const pfem2gpu_tuning* t) {
  if (!h || !t) return fail(PFEM2GPU_ERR_ARG, "null argument");
  if (t->walk_grid < kQStripes || t->walk_grid > 64 * kQStripes || (t->walk_grid % kQStripes) != 0)
    return fail(PFEM2GPU_ERR_ARG, "walk_grid must be a multiple of 1024 in [1024, 65536]");
  if (t->settle_grid < 1 || t->settle_grid > 65536) return fail(PFEM2GPU_ERR_ARG, "settle_grid must be in [1, 65536]");
  if (t->move_grid != 0 && (t->move_grid < 256 || t->move_grid > (1 << 20)))
    return fail(PFEM2GPU_ERR_ARG, "move_grid must be 0 or in [256, 1048576]");
  for (int b : {t->move_block, t->walk_block})
    if (b != 64 && b != 128 && b != 256 && b != 512)
      return fail(PFEM2GPU_ERR_ARG, "move_block and walk_block must be 64, 128, 256 or 512");
  if (h->capturing) return fail(PFEM2GPU_ERR_STATE, "tuning changed during graph capture");
  // the walk queues are sized for the in-cell pass's stripes (walk_qreg)
  h->move_block = t->move_block;
  h->walk_block = t->walk_block;
  h->walk_grid = t->walk_grid;
  h->settle_grid = t->settle_grid;
  h->move_grid = t->move_grid;
  drop_graph(h);
  return PFEM2GPU_OK;
}

int pfem2gpu_get_tuning(pfem2gpu_handle* h, pfem2gpu_tuning* t) {
  if (!h || !t) return fail(PFEM2GPU_ERR_ARG, "null argument");
  t->walk_grid = h->walk_grid;
  t->settle_grid = h->settle_grid;
  t->move_grid = h->move_grid;
  t->move_block = h->move_block;
  t->walk_block = h->walk_block;
  return PFEM2GPU_OK;
}

void* pfem2gpu_stream(pfem2gpu_handle* h) { return h ? (void*)h->stream : nullptr; }

int pfem2gpu_synchronize(pfem2gpu_handle* h) {
  if (!h) return fail(PFEM2GPU_ERR_ARG, "null handle");
  return check_device(h);
}

int pfem2gpu_set_timing(pfem2gpu_handle* h, int enable) {
  if (!h) return fail(PFEM2GPU_ERR_ARG, "null handle");
  CHK(t_collect(h));
  h->totals.clear();
  h->timing = enable != 0;
  return PFEM2GPU_OK;
}

int pfem2gpu_timings(pfem2gpu_handle* h, int cap, const char** names, int64_t* launches, double* total_ms,
                     int* n_out) {
  if (!h || !n_out) return fail(PFEM2GPU_ERR_ARG, "null argument");
  CHK(t_collect(h));
  h->names_store.clear();
  for (auto& e : h->totals) h->names_store.push_back(e.first);
  int i = 0;
  for (auto& e : h->totals) {
    if (i >= cap) break;
    if (names) names[i] = h->names_store[i].c_str();
    if (launches) launches[i] = e.second.first;
    if (total_ms) total_ms[i] = e.second.second;
    ++i;
  }
  *n_out = (int)h->totals.size();
  return PFEM2GPU_OK;
}

}  // extern "C"

namespace {

template <int D>
int pack_impl(pfem2gpu_handle* h, int32_t n_dest, const int32_t* dest_offsets, const int32_t* cells, void* out,
              int64_t capacity, int where, int64_t* bytes, int remove) {
  const long L = dest_offsets[n_dest];
  const long rec = WireLayout<D>::kSize;
  for (int i = 0; i < n_dest; ++i) bytes[i] = 0;
  if (L == 0) return PFEM2GPU_OK;
  DevBuf mem(h);
  int *dcells = nullptr, *cnt = nullptr, *offs = nullptr;
  CHK(mem.get(&dcells, L));
  CHK(mem.get(&cnt, L));
  CHK(mem.get(&offs, L + 1));
  HIPCHK(hipMemcpyAsync(dcells, cells, L * sizeof(int), hipMemcpyHostToDevice, h->stream));
  LAUNCH(h, "k_wire", k_wire_count, grid_for(L), 0, (int)L, dcells, h->cell_start, cnt);
  CHK(scan(h, cnt, (int)L, offs));
  std::vector<int> ho(L + 1);
  HIPCHK(hipMemcpyAsync(ho.data(), offs, (L + 1) * sizeof(int), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  for (int i = 0; i < n_dest; ++i) bytes[i] = (int64_t)(ho[dest_offsets[i + 1]] - ho[dest_offsets[i]]) * rec;
  const long total = ho[L];
  if (total * rec > capacity)
    return fail(PFEM2GPU_ERR_ARG, "pack_particles: output holds " + std::to_string(capacity) + " bytes, " +
                                      std::to_string(total * rec) + " needed");
  if (total > 0) {
    unsigned long long* dst = nullptr;
    if (where == PFEM2GPU_DEVICE) dst = static_cast<unsigned long long*>(out);
    else CHK(mem.get(&dst, (size_t)total * WireLayout<D>::kWords));
    const long groups = (L + kWireCells - 1) / kWireCells;
    LAUNCH_B(h, "k_wire_pack", k_wire_pack<D>, (int)std::min<long>(groups, 65536), kWireTile, 0, dmesh(h), (int)L,
             dcells, offs, h->cell_start, dparts(h), dst);
    if (where == PFEM2GPU_HOST)
      HIPCHK(hipMemcpyAsync(out, dst, (size_t)total * rec, hipMemcpyDeviceToHost, h->stream));
  }
  if (remove && total > 0) {
    unsigned char* flag = nullptr;
    int *kcnt = nullptr, *ncs = nullptr;
    CHK(mem.get(&flag, h->nc));
    CHK(mem.get(&kcnt, h->nc));
    CHK(mem.get(&ncs, h->nc + 1));
    LAUNCH(h, "k_fill", k_fill_bytes, grid_for(h->nc), 0, flag, (long)h->nc);
    LAUNCH(h, "k_wire", k_wire_flag, grid_for(L), 0, (int)L, dcells, flag);
    LAUNCH(h, "k_wire", k_wire_keep, grid_for(h->nc), 0, h->nc, h->cell_start, flag, kcnt);
    CHK(scan(h, kcnt, h->nc, ncs));
    const DTmp T = dtmp(h);
    LAUNCH(h, "k_wire", k_wire_compact<D>, grid_for(h->n_ub), 0, h->nc, h->cell_start, ncs, flag, dparts(h), T);
    LAUNCH(h, "k_copy", k_copy_state, grid_for(h->n_ub), 0, ncs + h->nc, T.x[0], T.x[1], D > 2 ? T.x[2] : nullptr,
           T.v[0], T.v[1], D > 2 ? T.v[2] : nullptr, T.cell, h->px[0], h->px[1], D > 2 ? h->px[2] : nullptr,
           h->pv[0], h->pv[1], D > 2 ? h->pv[2] : nullptr, h->pcell);
    HIPCHK(hipMemcpyAsync(h->cell_start, ncs, (h->nc + 1) * sizeof(int), hipMemcpyDeviceToDevice, h->stream));
    h->contrib_valid = false;
    h->settle_ev_valid = false;
  }
  HIPCHK(hipStreamSynchronize(h->stream));  // the pool buffers are reused after return
  return PFEM2GPU_OK;
}

template <int D>
int unpack_impl(pfem2gpu_handle* h, const void* data, int64_t n_bytes, int where, int64_t* n_rejected) {
  const long n = n_bytes / WireLayout<D>::kSize;
  if (n == 0) return PFEM2GPU_OK;
  CHK(reserve(h, n, n));
  DevBuf mem(h);
  const unsigned long long* src = nullptr;
  if (where == PFEM2GPU_DEVICE) {
    src = static_cast<const unsigned long long*>(data);
  } else {
    unsigned long long* st = nullptr;
    CHK(mem.get(&st, (size_t)n * WireLayout<D>::kWords));
    HIPCHK(hipMemcpyAsync(st, data, (size_t)n_bytes, hipMemcpyHostToDevice, h->stream));
    src = st;
  }
  const long tiles = (n + kWireTile - 1) / kWireTile;
  LAUNCH_B(h, "k_wire_unpack", k_wire_unpack<D>, (int)std::min<long>(tiles, 65536), kWireTile, 0, src, n, h->tx[0], h->tx[1],
           h->tx[2], h->tv[0], h->tv[1], h->tv[2], h->tcell);
  const double* sx[3] = {h->tx[0], h->tx[1], h->tx[2]};
  const double* sv[3] = {h->tv[0], h->tv[1], h->tv[2]};
  CHK(append_staged(h, n, sx, sv, h->tcell, n_rejected));
  HIPCHK(hipStreamSynchronize(h->stream));  // the staging buffer returns to the pool
  return PFEM2GPU_OK;
}

}  // namespace

extern "C" {

int64_t pfem2gpu_particle_record_size(int32_t dim) {
  return dim == 2 ? WireLayout<2>::kSize : dim == 3 ? WireLayout<3>::kSize : 0;
}


int pfem2gpu_pack_particles(pfem2gpu_handle* h, int32_t n_dest, const int32_t* dest_offsets, const int32_t* cells,
                            void* out, int64_t capacity, int where, int64_t* bytes, int remove) {
  if (!h || !dest_offsets || !bytes) return fail(PFEM2GPU_ERR_ARG, "pack_particles: null argument");
  if (n_dest < 0) return fail(PFEM2GPU_ERR_ARG, "pack_particles: n_dest < 0");
  if (where != PFEM2GPU_HOST && where != PFEM2GPU_DEVICE) return fail(PFEM2GPU_ERR_ARG, "bad where");
  if (dest_offsets[0] != 0) return fail(PFEM2GPU_ERR_ARG, "pack_particles: dest_offsets[0] must be 0");
  for (int i = 0; i < n_dest; ++i)
    if (dest_offsets[i + 1] < dest_offsets[i])
      return fail(PFEM2GPU_ERR_ARG, "pack_particles: dest_offsets must be non-decreasing");
  const long L = dest_offsets[n_dest];
  if (L > 0 && !cells) return fail(PFEM2GPU_ERR_ARG, "pack_particles: null cell list");
  if (L > h->nc) return fail(PFEM2GPU_ERR_ARG, "pack_particles: more listed cells than the mesh has");
  std::vector<char> seen(h->nc, 0);
  for (long i = 0; i < L; ++i) {
    const int c = cells[i];
    if (c < 0 || c >= h->nc) return fail(PFEM2GPU_ERR_ARG, "pack_particles: cell id out of range");
    if (seen[c]) return fail(PFEM2GPU_ERR_ARG, "pack_particles: cell " + std::to_string(c) + " listed twice");
    seen[c] = 1;
  }
  if (capacity < 0 || (capacity > 0 && !out)) return fail(PFEM2GPU_ERR_ARG, "pack_particles: bad output buffer");
  if (where == PFEM2GPU_DEVICE && (reinterpret_cast<uintptr_t>(out) & 7))
    return fail(PFEM2GPU_ERR_ARG, "pack_particles: device output must be 8-byte aligned");
  CHK(flush(h));
  return h->dim == 2 ? pack_impl<2>(h, n_dest, dest_offsets, cells, out, capacity, where, bytes, remove)
                     : pack_impl<3>(h, n_dest, dest_offsets, cells, out, capacity, where, bytes, remove);
}

int pfem2gpu_unpack_particles(pfem2gpu_handle* h, const void* data, int64_t n_bytes, int where,
                              int64_t* n_rejected) {
  if (!h) return fail(PFEM2GPU_ERR_ARG, "null handle");
  if (n_rejected) *n_rejected = 0;
  if (where != PFEM2GPU_HOST && where != PFEM2GPU_DEVICE) return fail(PFEM2GPU_ERR_ARG, "bad where");
  const int64_t rec = pfem2gpu_particle_record_size(h->dim);
  if (n_bytes < 0 || n_bytes % rec != 0)
    return fail(PFEM2GPU_ERR_ARG, "unpack_particles: " + std::to_string(n_bytes) +
                                      " bytes is not a whole number of " + std::to_string(rec) + "-byte records");
  if (n_bytes > 0 && !data) return fail(PFEM2GPU_ERR_ARG, "unpack_particles: null data");
  if (n_bytes / rec > 0x7fffffffL) return fail(PFEM2GPU_ERR_ARG, "unpack_particles: too many records");
  if (where == PFEM2GPU_DEVICE && (reinterpret_cast<uintptr_t>(data) & 7))
    return fail(PFEM2GPU_ERR_ARG, "unpack_particles: device data must be 8-byte aligned");
  CHK(flush(h));
  return h->dim == 2 ? unpack_impl<2>(h, data, n_bytes, where, n_rejected)
                     : unpack_impl<3>(h, data, n_bytes, where, n_rejected);
}

}  // extern "C"

// ------------------------------------------------------ no-allocation pack

namespace {

#define NA_LAUNCH(kernel, grid, block, ...)                                              \
  do {                                                                                   \
    hipLaunchKernelGGL(kernel, dim3(grid), dim3(block), 0, h->stream, __VA_ARGS__);      \
    HIPCHK(hipGetLastError());                                                           \
  } while (0)

// pack_impl's steps with every per-call buffer carved from the device's
// scratch array (offsets in ints, each piece 256-byte aligned) and launches
// without timing events; validation of duplicates runs on the device.
template <int D>
int pack_noalloc_impl(pfem2gpu_handle* h, int32_t n_dest, const int32_t* dest_offsets, const int32_t* cells,
                      void* out, int64_t capacity, int64_t* bytes, int remove) {
  const long L = dest_offsets[n_dest];
  const long rec = WireLayout<D>::kSize;
  const int nc = h->nc;
  long o = 0;
  auto take = [&](long n) {
    const long at = o;
    o += (std::max(n, 1L) + 63) & ~63L;
    return at;
  };
  const long i_ctl = take(4), i_doff = take(n_dest + 1), i_bytes = take(2L * (n_dest + 1));
  const long i_cells = take(L), i_cnt = take(L), i_offs = take(L + 1), i_flag = take((nc + 3) / 4);
  const long i_kcnt = remove ? take(nc) : 0, i_ncs = remove ? take(nc + 1L) : 0;
  if (o > kNoAllocScratchInts)
    return fail(PFEM2GPU_ERR_ARG, "pack_particles_noalloc: this call needs " + std::to_string(o * 4) +
                                      " bytes of scratch; the limit is " + std::to_string(kNoAllocScratchInts * 4));
  if (L == 0) return PFEM2GPU_OK;
  int* S = h->scratch;
  int* dcells = S + i_cells;
  int* cnt = S + i_cnt;
  int* offs = S + i_offs;
  unsigned char* flag = reinterpret_cast<unsigned char*>(S + i_flag);
  long long* dbytes = reinterpret_cast<long long*>(S + i_bytes);
  HIPCHK(hipMemcpyAsync(dcells, cells, L * sizeof(int), hipMemcpyHostToDevice, h->stream));
  HIPCHK(hipMemcpyAsync(S + i_doff, dest_offsets, (n_dest + 1L) * sizeof(int), hipMemcpyHostToDevice, h->stream));
  // the array keeps the previous call's contents: clear the flags and the
  // duplicate word (a fill kernel, so that the sequence can be captured)
  NA_LAUNCH(k_fill_int, 1, kBlock, S + i_ctl, 4L, 0);
  NA_LAUNCH(k_fill_int, grid_for((nc + 3) / 4), kBlock, S + i_flag, (long)((nc + 3) / 4), 0);
  NA_LAUNCH(k_wire_flag_unique, grid_for(L), kBlock, (int)L, dcells, reinterpret_cast<unsigned int*>(flag), S + i_ctl);
  NA_LAUNCH(k_wire_count, grid_for(L), kBlock, (int)L, dcells, h->cell_start, cnt);
  const int nb = (int)((L + kScanChunk - 1) / kScanChunk);
  NA_LAUNCH(k_scan_reduce, nb, kBlock, cnt, (int)L, h->scan_blk);
  NA_LAUNCH(k_scan_blocks, 1, kBlock, h->scan_blk, nb, offs + L);
  NA_LAUNCH(k_scan_down, nb, kBlock, cnt, (int)L, h->scan_blk, offs);
  NA_LAUNCH(k_wire_dest_bytes, grid_for(n_dest + 1L), kBlock, n_dest, S + i_doff, offs, rec, dbytes);
  if (n_dest > 0)
    HIPCHK(hipMemcpyAsync(bytes, dbytes, n_dest * sizeof(int64_t), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipMemcpyAsync(h->h_pinned + 8, dbytes + n_dest, sizeof(int64_t), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipMemcpyAsync(h->h_pinned + 6, S + i_ctl, sizeof(int), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  if (h->h_pinned[6] != 0)
    return fail(PFEM2GPU_ERR_ARG, "pack_particles_noalloc: cell " + std::to_string(h->h_pinned[6] - 1) +
                                      " listed twice");
  int64_t total_bytes = 0;
  std::memcpy(&total_bytes, h->h_pinned + 8, sizeof(int64_t));
  if (total_bytes > capacity)
    return fail(PFEM2GPU_ERR_ARG, "pack_particles_noalloc: output holds " + std::to_string(capacity) + " bytes, " +
                                      std::to_string(total_bytes) + " needed");
  if (total_bytes == 0) return PFEM2GPU_OK;
  const long groups = (L + kWireCells - 1) / kWireCells;
  NA_LAUNCH(k_wire_pack<D>, (int)std::min<long>(groups, 65536), kWireTile, dmesh(h), (int)L, dcells, offs,
            h->cell_start, dparts(h), static_cast<unsigned long long*>(out));
  if (remove) {
    int* kcnt = S + i_kcnt;
    int* ncs = S + i_ncs;
    const int nb2 = (nc + kScanChunk - 1) / kScanChunk;
    NA_LAUNCH(k_wire_keep, grid_for(nc), kBlock, nc, h->cell_start, flag, kcnt);
    NA_LAUNCH(k_scan_reduce, nb2, kBlock, kcnt, nc, h->scan_blk);
    NA_LAUNCH(k_scan_blocks, 1, kBlock, h->scan_blk, nb2, ncs + nc);
    NA_LAUNCH(k_scan_down, nb2, kBlock, kcnt, nc, h->scan_blk, ncs);
    const DTmp T = dtmp(h);
    NA_LAUNCH(k_wire_compact<D>, grid_for(h->n_ub), kBlock, nc, h->cell_start, ncs, flag, dparts(h), T);
    NA_LAUNCH(k_copy_state, grid_for(h->n_ub), kBlock, ncs + nc, T.x[0], T.x[1], D > 2 ? T.x[2] : nullptr, T.v[0],
              T.v[1], D > 2 ? T.v[2] : nullptr, T.cell, h->px[0], h->px[1], D > 2 ? h->px[2] : nullptr, h->pv[0],
              h->pv[1], D > 2 ? h->pv[2] : nullptr, h->pcell);
    HIPCHK(hipMemcpyAsync(h->cell_start, ncs, (nc + 1L) * sizeof(int), hipMemcpyDeviceToDevice, h->stream));
    h->contrib_valid = false;
    h->settle_ev_valid = false;
  }
  HIPCHK(hipStreamSynchronize(h->stream));  // the scratch is shared with the next call on this device
  return PFEM2GPU_OK;
}

#undef NA_LAUNCH

}  // namespace

extern "C" {

int pfem2gpu_pack_particles_noalloc(pfem2gpu_handle* h, int32_t n_dest, const int32_t* dest_offsets,
                                    const int32_t* cells, void* out, int64_t capacity, int64_t* bytes, int remove) {
  if (!h || !dest_offsets || !bytes) return fail(PFEM2GPU_ERR_ARG, "pack_particles_noalloc: null argument");
  if (n_dest < 0) return fail(PFEM2GPU_ERR_ARG, "pack_particles_noalloc: n_dest < 0");
  if (dest_offsets[0] != 0) return fail(PFEM2GPU_ERR_ARG, "pack_particles_noalloc: dest_offsets[0] must be 0");
  for (int i = 0; i < n_dest; ++i)
    if (dest_offsets[i + 1] < dest_offsets[i])
      return fail(PFEM2GPU_ERR_ARG, "pack_particles_noalloc: dest_offsets must be non-decreasing");
  for (int i = 0; i < n_dest; ++i) bytes[i] = 0;
  const long L = dest_offsets[n_dest];
  if (L > 0 && !cells) return fail(PFEM2GPU_ERR_ARG, "pack_particles_noalloc: null cell list");
  if (L > h->nc) return fail(PFEM2GPU_ERR_ARG, "pack_particles_noalloc: more listed cells than the mesh has");
  for (long i = 0; i < L; ++i)
    if (cells[i] < 0 || cells[i] >= h->nc) return fail(PFEM2GPU_ERR_ARG, "pack_particles_noalloc: cell id out of range");
  if (capacity < 0 || (capacity > 0 && !out)) return fail(PFEM2GPU_ERR_ARG, "pack_particles_noalloc: bad output buffer");
  if (reinterpret_cast<uintptr_t>(out) & 7)
    return fail(PFEM2GPU_ERR_ARG, "pack_particles_noalloc: device output must be 8-byte aligned");
  if (!h->scratch) return fail(PFEM2GPU_ERR_STATE, "pack_particles_noalloc: no scratch address for this device");
  CHK(flush(h));
  return h->dim == 2 ? pack_noalloc_impl<2>(h, n_dest, dest_offsets, cells, out, capacity, bytes, remove)
                     : pack_noalloc_impl<3>(h, n_dest, dest_offsets, cells, out, capacity, bytes, remove);
}

}  // extern "C"

// ------------------------------------------------------ particle removal

namespace {

template <int D>
int remove_impl(pfem2gpu_handle* h, long n_rm, const int64_t* index, long n) {
  DevBuf mem(h);
  long long* drm = nullptr;
  int *keep = nullptr, *pos = nullptr, *blk = nullptr, *cs_new = nullptr;
  const int nb = (int)((n + kScanChunk - 1) / kScanChunk);
  CHK(mem.get(&drm, n_rm));
  CHK(mem.get(&keep, n));
  CHK(mem.get(&pos, n + 1));
  CHK(mem.get(&blk, nb + 1));
  CHK(mem.get(&cs_new, h->nc + 1L));
  HIPCHK(hipMemcpyAsync(drm, index, n_rm * sizeof(long long), hipMemcpyHostToDevice, h->stream));
  const int* nptr = h->cell_start + h->nc;
  LAUNCH(h, "k_remove", k_keep_flags, grid_for(n), 0, nptr, keep);
  LAUNCH(h, "k_remove", k_drop_listed, grid_for(n_rm), 0, n_rm, drm, keep);
  LAUNCH(h, "scan", k_scan_reduce, nb, 0, keep, (int)n, blk);
  LAUNCH(h, "scan", k_scan_blocks, 1, 0, blk, nb, pos + n);
  LAUNCH(h, "scan", k_scan_down, nb, 0, keep, (int)n, blk, pos);
  const DTmp T = dtmp(h);
  LAUNCH(h, "k_remove", k_compact_kept<D>, grid_for(std::max(n, (long)h->nc + 1)), 0, nptr, h->nc, keep, pos,
         h->cell_start, cs_new, dparts(h), T);
  LAUNCH(h, "k_copy", k_copy_state, grid_for(n), 0, cs_new + h->nc, T.x[0], T.x[1], D > 2 ? T.x[2] : nullptr, T.v[0],
         T.v[1], D > 2 ? T.v[2] : nullptr, T.cell, h->px[0], h->px[1], D > 2 ? h->px[2] : nullptr, h->pv[0], h->pv[1],
         D > 2 ? h->pv[2] : nullptr, h->pcell);
  HIPCHK(hipMemcpyAsync(h->cell_start, cs_new, (h->nc + 1L) * sizeof(int), hipMemcpyDeviceToDevice, h->stream));
  h->contrib_valid = false;
  h->settle_ev_valid = false;
  HIPCHK(hipStreamSynchronize(h->stream));  // the pool buffers are reused after return
  long c = 0;
  return read_count(h, &c);
}

}  // namespace

extern "C" {

int pfem2gpu_remove_particles(pfem2gpu_handle* h, int64_t n, const int64_t* index) {
  if (!h) return fail(PFEM2GPU_ERR_ARG, "null handle");
  if (n < 0 || (n > 0 && !index)) return fail(PFEM2GPU_ERR_ARG, "remove_particles: bad index array");
  if (n == 0) return PFEM2GPU_OK;
  CHK(flush(h));
  long count = 0;
  CHK(read_count(h, &count));
  for (int64_t i = 0; i < n; ++i)
    if (index[i] < 0 || index[i] >= count)
      return fail(PFEM2GPU_ERR_ARG, "remove_particles: index " + std::to_string(index[i]) + " out of range [0, " +
                                        std::to_string(count) + ")");
  return h->dim == 2 ? remove_impl<2>(h, n, index, count) : remove_impl<3>(h, n, index, count);
}

}  // extern "C"

// ------------------------------------------------------------- FEM assembly

namespace {

// The CSR pattern of the DoFs sharing a cell (DoFTools::make_sparsity_pattern
// without constraints, src/pfem2fem.cpp:78-95) and, for every entry, the
// offsets of its cell-matrix contributions in ascending cell order (the
// order in which distribute_local_to_global adds them over the cells).
int fem_prepare(pfem2gpu_handle* h) {
  if (h->fem_ready) return PFEM2GPU_OK;
  const int nvc = h->nvc, nc = h->nc, nd = h->ndofs;
  if ((long)nc * nvc * nvc > INT32_MAX) return fail(PFEM2GPU_ERR_ARG, "mesh too large for FEM assembly");
  std::vector<int> cd((size_t)nc * nvc);
  HIPCHK(hipMemcpy(cd.data(), h->cd, cd.size() * 4, hipMemcpyDeviceToHost));
  std::vector<int> off(nd + 1, 0), lst(cd.size());
  for (size_t e = 0; e < cd.size(); ++e) off[cd[e] + 1]++;
  for (int j = 0; j < nd; ++j) off[j + 1] += off[j];
  {
    std::vector<int> pos(off.begin(), off.end() - 1);
    for (size_t e = 0; e < cd.size(); ++e) lst[pos[cd[e]]++] = (int)e;
  }
  std::vector<int64_t> rp(nd + 1, 0);
  std::vector<int32_t> col;
  col.reserve((size_t)nd * (nvc == 4 ? 9 : 27));
  std::vector<int32_t> row;
  for (int r = 0; r < nd; ++r) {
    row.clear();
    for (int e = off[r]; e < off[r + 1]; ++e) {
      const int c = lst[e] / nvc;
      for (int j = 0; j < nvc; ++j) row.push_back(cd[(size_t)c * nvc + j]);
    }
    std::sort(row.begin(), row.end());
    row.erase(std::unique(row.begin(), row.end()), row.end());
    col.insert(col.end(), row.begin(), row.end());
    rp[r + 1] = (int64_t)col.size();
  }
  const long nnz = (long)col.size();
  std::vector<long> soff(nnz + 1, 0);
  std::vector<int> src((size_t)nc * nvc * nvc);
  // contribution (p, (c*nvc + i)*nvc + j) for r = cd[c][i], in the order
  // rows ascending, cells ascending: a stable counting sort by p (the
  // position of p inside its row is kept between the two passes)
  std::vector<uint16_t> cp;
  cp.reserve(src.size());
  for (int r = 0; r < nd; ++r) {
    if (rp[r + 1] - rp[r] > 65535) return fail(PFEM2GPU_ERR_ARG, "FEM matrix row too long");
    const int32_t* b = col.data() + rp[r];
    const int32_t* e = col.data() + rp[r + 1];
    for (int q = off[r]; q < off[r + 1]; ++q) {
      const int c = lst[q] / nvc;
      for (int j = 0; j < nvc; ++j) {
        const uint16_t k = (uint16_t)(std::lower_bound(b, e, cd[(size_t)c * nvc + j]) - b);
        cp.push_back(k);
        soff[rp[r] + k + 1]++;
      }
    }
  }
  for (long p = 0; p < nnz; ++p) soff[p + 1] += soff[p];
  {
    std::vector<long> pos(soff.begin(), soff.end() - 1);
    size_t t = 0;
    for (int r = 0; r < nd; ++r)
      for (int q = off[r]; q < off[r + 1]; ++q)
        for (int j = 0; j < nvc; ++j) src[pos[rp[r] + cp[t++]]++] = lst[q] * nvc + j;
  }
  // the places of the cell-matrix entries in contribution order (the
  // inverse list); right-hand side entries use the d2c places (bpos)
  std::vector<int> apos(src.size());
  for (size_t e = 0; e < src.size(); ++e) apos[src[e]] = (int)e;
  int rc;
  if ((rc = dalloc(&h->fem_soff, soff.size())) || (rc = dalloc(&h->fem_apos, apos.size())) ||
      (rc = dalloc(&h->fem_A, (size_t)nc * nvc * nvc)) ||
      (rc = dalloc(&h->fem_b, (size_t)nc * nvc * h->dim)))
    return rc;
  HIPCHK(hipMemcpy(h->fem_soff, soff.data(), soff.size() * sizeof(long), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(h->fem_apos, apos.data(), apos.size() * 4, hipMemcpyHostToDevice));
  h->fem_A_planes = 1;
  h->fem_rowptr.swap(rp);
  h->fem_col.swap(col);
  h->fem_ready = true;
  return PFEM2GPU_OK;
}

int fem_grid(long nc, int nvc) {
  const long cpb = kFemBlock / nvc;
  long g = (nc + cpb - 1) / cpb;
  return (int)std::max(1L, std::min(g, 65536L));
}

// Host-side staging of the optional host outputs and inputs.
int fem_stage(pfem2gpu_handle* h, int vals, int nrhs, bool P, bool faces) {
  int rc;
  if (vals > h->fem_vals_planes) {
    dfree(h->fem_vals);
    h->fem_vals_planes = 0;
    if ((rc = dalloc(&h->fem_vals, (size_t)vals * h->fem_col.size()))) return rc;
    h->fem_vals_planes = vals;
  }
  for (int k = 0; k < nrhs; ++k)
    if (!h->fem_rhs[k] && (rc = dalloc(&h->fem_rhs[k], h->ndofs))) return rc;
  if (P && !h->fem_P && (rc = dalloc(&h->fem_P, h->ndofs))) return rc;
  if (faces && !h->fem_faces && (rc = dalloc(&h->fem_faces, h->nc))) return rc;
  return PFEM2GPU_OK;
}

template <int D>
int fem_assemble(pfem2gpu_handle* h, bool pressure, const double* const* Vp, const double* P,
                 const unsigned char* faces, double c0, double c1, double* matrix, double* const* rhs, int where) {
  const bool host = where == PFEM2GPU_HOST;
  const int nrhs = pressure ? 1 : D;
  CHK(fem_prepare(h));
  CHK(fem_stage(h, (host && matrix) ? 1 : 0, host ? nrhs : 0, host && P, host && faces));
  DField Vf;
  CHK(field_in(h, Vp, where, 0, &Vf));
  const double* Pd = P;
  const unsigned char* fd = faces;
  if (host) {
    if (P) HIPCHK(hipMemcpyAsync(h->fem_P, P, h->ndofs * sizeof(double), hipMemcpyHostToDevice, h->stream));
    if (faces) HIPCHK(hipMemcpyAsync(h->fem_faces, faces, h->nc, hipMemcpyHostToDevice, h->stream));
    Pd = P ? h->fem_P : nullptr;
    fd = faces ? h->fem_faces : nullptr;
  }
  const int grid = fem_grid(h->nc, h->nvc);
  if (pressure) {
    LAUNCH_B(h, "k_fem_pressure", (k_fem_cells<D, true>), grid, kFemBlock, 0, dmesh(h), h->nc, Vf, nullptr, c0, c1,
             h->fem_apos, h->bpos, h->fem_A, h->fem_b);
    if (fd)
      LAUNCH(h, "k_fem_faces", k_fem_faces<D>, grid_for(h->nc), 0, dmesh(h), h->nc, Vf, c0, fd, h->bpos,
             h->fem_b);
  } else {
    LAUNCH_B(h, "k_fem_vcorr", (k_fem_cells<D, false>), grid, kFemBlock, 0, dmesh(h), h->nc, Vf, Pd, c0, c1,
             h->fem_apos, h->bpos, h->fem_A, h->fem_b);
  }
  const long nnz = (long)h->fem_col.size();
  if (matrix) {
    double* vd = host ? h->fem_vals : matrix;
    LAUNCH(h, pressure ? "k_fem_gather_matrix" : "k_fem_gather_mass", k_fem_segsum<long>, grid_for(nnz), 0, nnz,
           h->fem_soff, 1, 0L, h->fem_A, vd, nullptr, nullptr);
  }
  double* rd[3] = {nullptr, nullptr, nullptr};
  for (int k = 0; k < nrhs; ++k) rd[k] = host ? h->fem_rhs[k] : rhs[k];
  LAUNCH(h, pressure ? "k_fem_gather_rhs" : "k_fem_gather_rhs_v", k_fem_segsum<int>, grid_for(h->ndofs), 0,
         (long)h->ndofs, h->d2c_off, nrhs, (long)h->nc * h->nvc, h->fem_b, rd[0], rd[1], rd[2]);
  if (host) {
    if (matrix) HIPCHK(hipMemcpyAsync(matrix, h->fem_vals, nnz * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    for (int k = 0; k < nrhs; ++k)
      HIPCHK(hipMemcpyAsync(rhs[k], rd[k], h->ndofs * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
  }
  return PFEM2GPU_OK;
}

// assemble_velocity_prediction: one matrix and one right-hand side per
// component.
template <int D>
int fem_predict(pfem2gpu_handle* h, const double* const* Vo, const double* const* U, const unsigned char* faces,
                double rho, double mu, double dt, double* const* matrix, double* const* rhs, int where) {
  const bool host = where == PFEM2GPU_HOST;
  CHK(fem_prepare(h));
  const long nnz = (long)h->fem_col.size();
  const long aplane = (long)h->nc * h->nvc * h->nvc;
  if (h->fem_A_planes < D) {
    dfree(h->fem_A);
    h->fem_A_planes = 0;
    int rc;
    if ((rc = dalloc(&h->fem_A, (size_t)D * aplane))) return rc;
    h->fem_A_planes = D;
  }
  CHK(fem_stage(h, (host && matrix) ? D : 0, host ? D : 0, false, host && faces));
  DField Vf, Uf;
  CHK(field_in(h, Vo, where, 0, &Vf));
  CHK(field_in(h, U, where, 1, &Uf));
  const unsigned char* fd = faces;
  if (host && faces) {
    HIPCHK(hipMemcpyAsync(h->fem_faces, faces, h->nc, hipMemcpyHostToDevice, h->stream));
    fd = h->fem_faces;
  }
  LAUNCH_B(h, "k_fem_predict", k_fem_predict<D>, fem_grid(h->nc, h->nvc), kFemBlock, 0, dmesh(h), h->nc, Vf, Uf, rho,
           mu, dt, h->fem_apos, h->bpos, h->fem_A, h->fem_b);
  if (fd)
    LAUNCH(h, "k_fem_predict_faces", k_fem_predict_faces<D>, grid_for(h->nc), 0, dmesh(h), h->nc, Uf, mu, dt, fd,
           h->fem_apos, h->bpos, h->fem_A, h->fem_b);
  double* md[3] = {nullptr, nullptr, nullptr};
  double* rd[3] = {nullptr, nullptr, nullptr};
  for (int k = 0; k < D; ++k) {
    if (matrix) md[k] = host ? h->fem_vals + (size_t)k * nnz : matrix[k];
    rd[k] = host ? h->fem_rhs[k] : rhs[k];
  }
  if (matrix)
    LAUNCH(h, "k_fem_gather_pred", k_fem_segsum<long>, grid_for(nnz), 0, nnz, h->fem_soff, D, aplane, h->fem_A,
           md[0], md[1], md[2]);
  LAUNCH(h, "k_fem_gather_rhs_p", k_fem_segsum<int>, grid_for(h->ndofs), 0, (long)h->ndofs, h->d2c_off, D,
         (long)h->nc * h->nvc, h->fem_b, rd[0], rd[1], rd[2]);
  if (host) {
    for (int k = 0; k < D; ++k) {
      if (matrix)
        HIPCHK(hipMemcpyAsync(matrix[k], md[k], nnz * sizeof(double), hipMemcpyDeviceToHost, h->stream));
      HIPCHK(hipMemcpyAsync(rhs[k], rd[k], h->ndofs * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    }
    HIPCHK(hipStreamSynchronize(h->stream));
  }
  return PFEM2GPU_OK;
}

int fem_check(pfem2gpu_handle* h, const double* const* V, const void* rhs, int where) {
  if (!h) return fail(PFEM2GPU_ERR_ARG, "null handle");
  if (!V || !rhs) return fail(PFEM2GPU_ERR_ARG, "null field or right-hand side");
  if (where != PFEM2GPU_HOST && where != PFEM2GPU_DEVICE) return fail(PFEM2GPU_ERR_ARG, "bad where");
  return PFEM2GPU_OK;
}

}  // namespace

extern "C" {

int pfem2gpu_fem_sparsity(pfem2gpu_handle* h, int64_t* nnz, int64_t* row_ptr, int32_t* cols) {
  if (!h) return fail(PFEM2GPU_ERR_ARG, "null handle");
  CHK(fem_prepare(h));
  if (nnz) *nnz = (int64_t)h->fem_col.size();
  if (row_ptr) std::memcpy(row_ptr, h->fem_rowptr.data(), h->fem_rowptr.size() * sizeof(int64_t));
  if (cols) std::memcpy(cols, h->fem_col.data(), h->fem_col.size() * sizeof(int32_t));
  return PFEM2GPU_OK;
}

int pfem2gpu_assemble_pressure_equation(pfem2gpu_handle* h, const double* const* V_pred,
                                        const uint8_t* boundary_faces, double rho, double time_step, double* matrix,
                                        double* rhs, int where) {
  CHK(fem_check(h, V_pred, rhs, where));
  const double coeff = rho / time_step;
  double* r[1] = {rhs};
  return h->dim == 2 ? fem_assemble<2>(h, true, V_pred, nullptr, boundary_faces, coeff, 0.0, matrix, r, where)
                     : fem_assemble<3>(h, true, V_pred, nullptr, boundary_faces, coeff, 0.0, matrix, r, where);
}

int pfem2gpu_assemble_velocity_prediction(pfem2gpu_handle* h, const double* const* V_old, const double* const* V,
                                          const uint8_t* boundary_faces, double rho, double mu, double time_step,
                                          double* const* matrix, double* const* rhs, int where) {
  CHK(fem_check(h, V_old, rhs, where));
  if (!V) return fail(PFEM2GPU_ERR_ARG, "null velocity");
  for (int k = 0; k < h->dim; ++k)
    if (!rhs[k] || (matrix && !matrix[k])) return fail(PFEM2GPU_ERR_ARG, "null output component");
  return h->dim == 2 ? fem_predict<2>(h, V_old, V, boundary_faces, rho, mu, time_step, matrix, rhs, where)
                     : fem_predict<3>(h, V_old, V, boundary_faces, rho, mu, time_step, matrix, rhs, where);
}

int pfem2gpu_assemble_velocity_correction(pfem2gpu_handle* h, const double* const* V_pred, const double* P,
                                          double rho, double time_step, double* matrix, double* const* rhs,
                                          int where) {
  CHK(fem_check(h, V_pred, rhs, where));
  if (!P) return fail(PFEM2GPU_ERR_ARG, "null pressure");
  for (int k = 0; k < h->dim; ++k)
    if (!rhs[k]) return fail(PFEM2GPU_ERR_ARG, "null right-hand side component");
  return h->dim == 2 ? fem_assemble<2>(h, false, V_pred, P, nullptr, rho, time_step, matrix, rhs, where)
                     : fem_assemble<3>(h, false, V_pred, P, nullptr, rho, time_step, matrix, rhs, where);
}

}  // extern "C"
