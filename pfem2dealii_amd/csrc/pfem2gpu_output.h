// Particle output in the reference's VTK format (host code, no device use).
//
// output_particle_solution (src/pfem2particlehandler.cpp:217-317) writes one
// ASCII UnstructuredGrid piece per rank with every particle as a vertex cell
// and, on rank 0, the parallel index file that lists the pieces. The text is
// reproduced byte for byte: positions and velocities print their first
// component through the stream's default double formatting (six significant
// digits, printf %g) and the others through std::to_string (printf %f), the
// remaining components of a 2D particle as the literal 0.0; the subdomain is
// an unsigned integer; file names zero-pad the step to two digits and the
// rank to three (deal.II Utilities::int_to_string pads and never truncates).
// Lines are built with snprintf into a buffer written in large blocks.
#pragma once

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>

namespace pfem2gpu_out {

// Output file with a large write buffer; every append is checked at close.
class Out {
 public:
  explicit Out(const char* path) : f_(std::fopen(path, "wb")) { buf_.reserve(kChunk + 256); }
  ~Out() {
    if (f_) std::fclose(f_);
  }
  bool ok() const { return f_ != nullptr && !failed_; }
  void put(const char* s, size_t n) {
    buf_.append(s, n);
    if (buf_.size() >= kChunk) flush();
  }
  void put(const char* s) { put(s, std::strlen(s)); }
  void put(const std::string& s) { put(s.data(), s.size()); }
  // printf-style append of one short field
  template <typename... A>
  void fmt(const char* f, A... a) {
    char tmp[128];
    const int n = std::snprintf(tmp, sizeof(tmp), f, a...);
    if (n < 0 || n >= (int)sizeof(tmp)) {
      failed_ = true;
      return;
    }
    put(tmp, (size_t)n);
  }
  bool close() {
    flush();
    if (!f_) return false;
    const bool good = std::fclose(f_) == 0 && !failed_;
    f_ = nullptr;
    return good;
  }

 private:
  static constexpr size_t kChunk = 1 << 22;
  void flush() {
    if (!f_ || buf_.empty()) return;
    if (std::fwrite(buf_.data(), 1, buf_.size(), f_) != buf_.size()) failed_ = true;
    buf_.clear();
  }
  std::FILE* f_;
  std::string buf_;
  bool failed_ = false;
};

// Utilities::int_to_string(value, digits): zero-padded to at least `digits`
inline std::string padded(int value, int digits) {
  std::string s = std::to_string(value);
  const size_t at = (!s.empty() && s[0] == '-') ? 1 : 0;
  if ((int)s.size() < digits) s.insert(at, std::string(digits - s.size(), '0'));
  return s;
}

// one "x y z" line of a 3-component Float32 array (src/pfem2particlehandler.cpp:231-236,269-273)
inline void vec3_line(Out& o, int dim, const double* a, int64_t n, int64_t i) {
  o.fmt("          %g ", a[i]);
  if (dim > 1)
    o.fmt("%f ", a[n + i]);
  else
    o.put("0.0 ");
  if (dim > 2)
    o.fmt("%f\n", a[2 * n + i]);
  else
    o.put("0.0\n");
}

// One rank's piece (x, v: [dim][n] host arrays). Returns false if the file
// cannot be written.
inline bool write_vtu(const char* path, int dim, int64_t n, const double* x, const double* v, int32_t subdomain) {
  Out o(path);
  if (!o.ok()) return false;
  const unsigned sub = (unsigned)subdomain;  // types::subdomain_id
  o.put("<?xml version=\"1.0\" ?> \n");
  o.put("<VTKFile type=\"UnstructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\">\n");
  o.put("  <UnstructuredGrid>\n");
  o.fmt("    <Piece NumberOfPoints=\"%lld\" NumberOfCells=\"%lld\">\n", (long long)n, (long long)n);
  o.put("      <Points>\n");
  o.put("        <DataArray type=\"Float32\" NumberOfComponents=\"3\" Format=\"ascii\">\n");
  for (int64_t i = 0; i < n; ++i) vec3_line(o, dim, x, n, i);
  o.put("        </DataArray>\n");
  o.put("      </Points>\n");
  o.put("      <Cells>\n");
  o.put("        <DataArray type=\"Int32\" Name=\"connectivity\" Format=\"ascii\">\n");
  for (int64_t i = 0; i < n; ++i) o.fmt("          %u\n", (unsigned)i);
  o.put("        </DataArray>\n");
  o.put("        <DataArray type=\"Int32\" Name=\"offsets\" Format=\"ascii\">\n");
  o.put("        ");
  for (int64_t i = 0; i < n; ++i) o.fmt("  %u", (unsigned)(i + 1));
  o.put("\n");
  o.put("        </DataArray>\n");
  o.put("        <DataArray type=\"Int32\" Name=\"types\" Format=\"ascii\">\n");
  o.put("        ");
  for (int64_t i = 0; i < n; ++i) o.put("  1");
  o.put("\n");
  o.put("        </DataArray>\n");
  o.put("      </Cells>\n");
  o.put("      <PointData Scalars=\"scalars\">\n");
  o.put("        <DataArray type=\"Float32\" Name=\"velocity\" NumberOfComponents=\"3\" Format=\"ascii\">\n");
  for (int64_t i = 0; i < n; ++i) vec3_line(o, dim, v, n, i);
  o.put("        </DataArray>\n");
  o.put("        <DataArray type=\"Float32\" Name=\"subdomain\" Format=\"ascii\">\n");
  o.put("        ");
  for (int64_t i = 0; i < n; ++i) o.fmt("  %u", sub);
  o.put("\n");
  o.put("        </DataArray>\n");
  o.put("      </PointData>\n");
  o.put("    </Piece>\n");
  o.put("  </UnstructuredGrid>\n");
  o.put("</VTKFile>\n");
  return o.close();
}

// The index file of n_ranks pieces of step timestep_number.
inline bool write_pvtu(const char* path, int32_t timestep_number, int32_t n_ranks) {
  Out o(path);
  if (!o.ok()) return false;
  const std::string step = padded(timestep_number, 2);
  o.put("<?xml version=\"1.0\" ?> \n");
  o.put("<VTKFile type=\"PUnstructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\">\n");
  o.put("  <PUnstructuredGrid GhostLevel=\"0\">\n");
  o.put("    <PPointData Scalars=\"scalars\">\n");
  o.put("      <PDataArray type=\"Float32\" Name=\"velocity\" NumberOfComponents=\"3\" format=\"ascii\"/>\n");
  o.put("      <PDataArray type=\"Float32\" Name=\"subdomain\" format=\"ascii\"/>\n");
  o.put("    </PPointData>\n");
  o.put("    <PPoints>\n");
  o.put("      <PDataArray type=\"Float32\" NumberOfComponents=\"3\"/>\n");
  o.put("    </PPoints>\n");
  for (int r = 0; r < n_ranks; ++r) o.put("    <Piece Source=\"particles-" + step + "." + padded(r, 3) + ".vtu\"/>");
  o.put("  </PUnstructuredGrid>\n");
  o.put("</VTKFile>\n");
  return o.close();
}

}  // namespace pfem2gpu_out
