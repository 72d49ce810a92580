// Particle records of the reference's MPI wire format and the kernels that
// write and read them.
//
// A record is what pfem2Particle::write_data serialises
// (src/pfem2particle.cpp:221-256) and the reading constructor parses back
// (:17-44), fields back to back with no padding:
//   u32 id, i32 tria_position, i32 cell_dofs[2^D],
//   f64 location[D], f64 reference_location[D], f64 velocity[D],
//   f64 velocity_ext[D]
// = 88 bytes in 2D and 136 in 3D (serialized_size_in_bytes, :213-219). Every
// double sits at a multiple of 8 bytes, so a record is a whole number of
// 64-bit words and the kernels move it as such (little-endian: the id is the
// low half of word 0, tria_position the high half).
#pragma once

#include "pfem2gpu_kernels.h"

namespace pfem2gpu {

template <int D>
struct WireLayout {
  static constexpr int NV = 1 << D;
  static constexpr int kDofs = 8;                 // byte offsets
  static constexpr int kLoc = kDofs + 4 * NV;
  static constexpr int kRef = kLoc + 8 * D;
  static constexpr int kVel = kRef + 8 * D;
  static constexpr int kExt = kVel + 8 * D;
  static constexpr int kSize = kExt + 8 * D;
  static constexpr int kWords = kSize / 8;
};
static_assert(WireLayout<2>::kSize == 88 && WireLayout<3>::kSize == 136, "record sizes of the reference");
static_assert(WireLayout<2>::kLoc % 8 == 0 && WireLayout<3>::kLoc % 8 == 0, "doubles at 8-byte offsets");

constexpr int kWireTile = 256;  // records per workgroup tile (one per thread)

__device__ __forceinline__ unsigned long long wire_pair(unsigned int lo, unsigned int hi) {
  return (unsigned long long)lo | ((unsigned long long)hi << 32);
}

// Particles in each listed cell: cnt[i] = count of cells[i].
__global__ void k_wire_count(int L, const int* __restrict__ cells, const int* __restrict__ cs, int* __restrict__ cnt) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < L; i += gridDim.x * blockDim.x) {
    const int c = cells[i];
    cnt[i] = cs[c + 1] - cs[c];
  }
}

// Records of the listed cells [i0, i0 + kWireCells) are contiguous in the
// output, offs[i0] .. offs[i0 + kWireCells) (offs: exclusive scan of the listed
// counts), so a workgroup takes one such group: the group's offsets and cell
// ids go to LDS, and each thread finds its record's cell there (a search over
// at most kWireCells entries), builds the record in LDS, and the tile is
// written out with contiguous 64-bit stores. The id written is the particle's
// slot in the handle (the reference never reads ids in arithmetic, and its
// own are uninitialised counters); velocity_ext is written as zero (the
// reference overwrites it at the start of every substep,
// src/pfem2particlehandler.cpp:152-161, before anything reads it).
constexpr int kWireCells = 32;

template <int D>
__global__ __launch_bounds__(kWireTile) void k_wire_pack(DMesh m, int L, const int* __restrict__ cells,
                                                        const int* __restrict__ offs, const int* __restrict__ cs,
                                                        DParts P, unsigned long long* __restrict__ out) {
  using WL = WireLayout<D>;
  constexpr int W = WL::kWords;
  constexpr int NV = WL::NV;
  __shared__ unsigned long long s_rec[kWireTile * W];
  __shared__ int s_off[kWireCells + 1];
  __shared__ int s_cell[kWireCells];
  const int ngroups = (L + kWireCells - 1) / kWireCells;
  for (int g = blockIdx.x; g < ngroups; g += gridDim.x) {
    const int i0 = g * kWireCells;
    const int nci = min(kWireCells, L - i0);
    if ((int)threadIdx.x <= nci) s_off[threadIdx.x] = offs[i0 + threadIdx.x];
    if ((int)threadIdx.x < nci) s_cell[threadIdx.x] = cells[i0 + threadIdx.x];
    __syncthreads();
    const long base = s_off[0];
    const int total = s_off[nci] - s_off[0];
    for (int j0 = 0; j0 < total; j0 += kWireTile) {
      const int cnt = min(kWireTile, total - j0);
      if ((int)threadIdx.x < cnt) {
        const int j = s_off[0] + j0 + threadIdx.x;
        int lo = 0, hi = nci;  // largest lo with s_off[lo] <= j
        while (hi - lo > 1) {
          const int mid = (lo + hi) >> 1;
          if (s_off[mid] <= j) lo = mid;
          else hi = mid;
        }
        const int c = s_cell[lo];
        const int p = cs[c] + (j - s_off[lo]);
        double x[D], xi[D];
#pragma unroll
        for (int k = 0; k < D; ++k) x[k] = P.x[k][p];
        recover_xi_c<D>(m, m.prm[inst_of(m, c)], c, placed_of(P.cell[p]), x, xi);
        unsigned long long* w = s_rec + threadIdx.x * W;
        w[0] = wire_pair((unsigned int)p, (unsigned int)c);
#pragma unroll
        for (int q = 0; q < NV / 2; ++q)
          w[WL::kDofs / 8 + q] = wire_pair((unsigned int)m.cd[(size_t)c * NV + 2 * q],
                                           (unsigned int)m.cd[(size_t)c * NV + 2 * q + 1]);
#pragma unroll
        for (int k = 0; k < D; ++k) {
          w[WL::kLoc / 8 + k] = (unsigned long long)__double_as_longlong(x[k]);
          w[WL::kRef / 8 + k] = (unsigned long long)__double_as_longlong(xi[k]);
          w[WL::kVel / 8 + k] = (unsigned long long)__double_as_longlong(P.v[k][p]);
          w[WL::kExt / 8 + k] = 0ULL;
        }
      }
      __syncthreads();
      unsigned long long* dst = out + (base + j0) * W;
      for (int i = threadIdx.x; i < cnt * W; i += kWireTile) dst[i] = s_rec[i];
      __syncthreads();
    }
    __syncthreads();  // s_off and s_cell are rewritten for the next group
  }
}

// Parses n records into the staging arrays (position, velocity, cell =
// tria_position); the tile is read with contiguous 64-bit loads first.
// cell_dofs, reference_location and velocity_ext are not used: the
// receiving handle takes DoFs from its mesh and recomputes the reference
// location from (cell, position) with its own inverse map.
template <int D>
__global__ __launch_bounds__(kWireTile) void k_wire_unpack(const unsigned long long* __restrict__ in, long n,
                                                          double* tx0, double* tx1, double* tx2, double* tv0,
                                                          double* tv1, double* tv2, int* __restrict__ tcell) {
  using WL = WireLayout<D>;
  constexpr int W = WL::kWords;
  __shared__ unsigned long long s_rec[kWireTile * W];
  double* tx[3] = {tx0, tx1, tx2};
  double* tv[3] = {tv0, tv1, tv2};
  const long ntiles = (n + kWireTile - 1) / kWireTile;
  for (long tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const long r0 = tile * kWireTile;
    const int cnt = (int)min((long)kWireTile, n - r0);
    const unsigned long long* src = in + r0 * W;
    for (int i = threadIdx.x; i < cnt * W; i += kWireTile) s_rec[i] = src[i];
    __syncthreads();
    if ((int)threadIdx.x < cnt) {
      const long r = r0 + threadIdx.x;
      const unsigned long long* w = s_rec + threadIdx.x * W;
      tcell[r] = (int)(unsigned int)(w[0] >> 32);
#pragma unroll
      for (int k = 0; k < D; ++k) {
        tx[k][r] = __longlong_as_double((long long)w[WL::kLoc / 8 + k]);
        tv[k][r] = __longlong_as_double((long long)w[WL::kVel / 8 + k]);
      }
    }
    __syncthreads();
  }
}

__global__ void k_fill_bytes(unsigned char* __restrict__ a, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) a[i] = 0;
}

// Removal of the particles of listed cells (they left for another
// subdomain, src/pfem2particlehandler.cpp:484-486): a flag per listed cell,
// the kept counts per cell, then each kept particle moves to its cell's new
// start + its rank in the cell (the canonical order is kept).
__global__ void k_wire_flag(int L, const int* __restrict__ cells, unsigned char* __restrict__ flag) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < L; i += gridDim.x * blockDim.x) flag[cells[i]] = 1;
}

// The flag of every listed cell, set with a word-wide atomic OR on the byte
// array (4-byte aligned); a cell whose byte was already set is listed twice
// and is reported through *dup (cell + 1).
__global__ void k_wire_flag_unique(int L, const int* __restrict__ cells, unsigned int* __restrict__ flag_words,
                                   int* __restrict__ dup) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < L; i += gridDim.x * blockDim.x) {
    const int c = cells[i];
    const unsigned int bit = 1u << (8 * (c & 3));
    const unsigned int old = atomicOr(&flag_words[c >> 2], bit);
    if (old & bit) atomicExch(dup, c + 1);
  }
}

// Byte counts per destination from the record offsets of the listed cells:
// out[i] = (offs[doff[i + 1]] - offs[doff[i]]) * rec, out[n_dest] = offs[L] * rec.
__global__ void k_wire_dest_bytes(int n_dest, const int* __restrict__ doff, const int* __restrict__ offs, long rec,
                                  long long* __restrict__ out) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i <= n_dest; i += gridDim.x * blockDim.x)
    out[i] = i < n_dest ? (long long)(offs[doff[i + 1]] - offs[doff[i]]) * rec : (long long)offs[doff[n_dest]] * rec;
}

__global__ void k_wire_keep(int nc, const int* __restrict__ cs, const unsigned char* __restrict__ flag,
                            int* __restrict__ cnt) {
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < nc; c += gridDim.x * blockDim.x)
    cnt[c] = flag[c] ? 0 : cs[c + 1] - cs[c];
}

template <int D>
__global__ __launch_bounds__(kBlock) void k_wire_compact(int nc, const int* __restrict__ cs,
                                                        const int* __restrict__ ncs,
                                                        const unsigned char* __restrict__ flag, DParts P, DTmp T) {
  const int n = cs[nc];
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
    const int w = P.cell[p];
    const int c = cell_of(w);
    if (flag[c]) continue;
    const int q = ncs[c] + (p - cs[c]);
#pragma unroll
    for (int k = 0; k < D; ++k) {
      T.x[k][q] = P.x[k][p];
      T.v[k][q] = P.v[k][p];
    }
    T.cell[q] = w;
  }
}

}  // namespace pfem2gpu
