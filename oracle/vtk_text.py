"""CPU restatement of the reference's particle output text. Test infrastructure.

output_particle_solution (src/pfem2particlehandler.cpp:217-317) writes, per
rank, an ASCII VTK UnstructuredGrid piece with one vertex cell per particle,
and on rank 0 the PUnstructuredGrid index of the pieces. The arithmetic-free
parts are fixed strings; the numbers follow three C++ library rules:
  - `ostream << double` with default flags and precision 6 prints what
    printf("%.6g") prints (the C++ standard defines num_put through printf
    conversions): the first component of positions and velocities (:231, :270);
  - std::to_string(double) prints printf("%f"): the other components (:232-233,
    :271-272); a missing dimension prints the literal 0.0;
  - deal.II Utilities::int_to_string(v, digits) zero-pads to at least `digits`
    characters (file names, :219, :312).
Python's %-formatting follows C's printf for finite values; a NaN with the
sign bit set is printed "-nan" by glibc, which the helpers below reproduce.
"""
import math

import numpy as np


def _signed_nan(a, text):
    if math.isnan(a):
        return "-nan" if math.copysign(1.0, a) < 0 else "nan"
    return text


def stream_double(a):
    """`std::ostream << a` with default formatting (printf %.6g)."""
    a = float(a)
    return _signed_nan(a, "%g" % a)


def to_string(a):
    """std::to_string(double a) (printf %f)."""
    a = float(a)
    return _signed_nan(a, "%f" % a)


def int_to_string(value, digits):
    """deal.II Utilities::int_to_string(value, digits)."""
    s = str(int(value))
    at = 1 if s.startswith("-") else 0
    if len(s) < digits:
        s = s[:at] + "0" * (digits - len(s)) + s[at:]
    return s


def _vec3_lines(dim, a):
    n = a.shape[1]
    out = []
    for i in range(n):
        y = to_string(a[1, i]) if dim > 1 else "0.0"
        z = to_string(a[2, i]) if dim > 2 else "0.0"
        out.append("          " + stream_double(a[0, i]) + " " + y + " " + z + "\n")
    return out


def vtu_text(x, v, subdomain=0):
    """The piece of src/pfem2particlehandler.cpp:219-294 for positions x and
    velocities v ([dim][n], in output order)."""
    x = np.asarray(x, dtype=np.float64)
    v = np.asarray(v, dtype=np.float64)
    dim, n = x.shape
    sub = str(int(subdomain) & 0xFFFFFFFF)  # types::subdomain_id is unsigned
    t = ["<?xml version=\"1.0\" ?> \n",
         "<VTKFile type=\"UnstructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\">\n",
         "  <UnstructuredGrid>\n",
         "    <Piece NumberOfPoints=\"%d\" NumberOfCells=\"%d\">\n" % (n, n),
         "      <Points>\n",
         "        <DataArray type=\"Float32\" NumberOfComponents=\"3\" Format=\"ascii\">\n"]
    t += _vec3_lines(dim, x)
    t += ["        </DataArray>\n",
          "      </Points>\n",
          "      <Cells>\n",
          "        <DataArray type=\"Int32\" Name=\"connectivity\" Format=\"ascii\">\n"]
    t += ["          %d\n" % i for i in range(n)]
    t += ["        </DataArray>\n",
          "        <DataArray type=\"Int32\" Name=\"offsets\" Format=\"ascii\">\n",
          "        " + "".join("  %d" % (i + 1) for i in range(n)) + "\n",
          "        </DataArray>\n",
          "        <DataArray type=\"Int32\" Name=\"types\" Format=\"ascii\">\n",
          "        " + "  1" * n + "\n",
          "        </DataArray>\n",
          "      </Cells>\n",
          "      <PointData Scalars=\"scalars\">\n",
          "        <DataArray type=\"Float32\" Name=\"velocity\" NumberOfComponents=\"3\" Format=\"ascii\">\n"]
    t += _vec3_lines(dim, v)
    t += ["        </DataArray>\n",
          "        <DataArray type=\"Float32\" Name=\"subdomain\" Format=\"ascii\">\n",
          "        " + ("  " + sub) * n + "\n",
          "        </DataArray>\n",
          "      </PointData>\n",
          "    </Piece>\n",
          "  </UnstructuredGrid>\n",
          "</VTKFile>\n"]
    return "".join(t)


def pvtu_text(timestep_number, n_ranks):
    """The index file of src/pfem2particlehandler.cpp:296-316 (the pieces are
    written without line breaks between them, :312)."""
    step = int_to_string(timestep_number, 2)
    t = ["<?xml version=\"1.0\" ?> \n",
         "<VTKFile type=\"PUnstructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\">\n",
         "  <PUnstructuredGrid GhostLevel=\"0\">\n",
         "    <PPointData Scalars=\"scalars\">\n",
         "      <PDataArray type=\"Float32\" Name=\"velocity\" NumberOfComponents=\"3\" format=\"ascii\"/>\n",
         "      <PDataArray type=\"Float32\" Name=\"subdomain\" format=\"ascii\"/>\n",
         "    </PPointData>\n",
         "    <PPoints>\n",
         "      <PDataArray type=\"Float32\" NumberOfComponents=\"3\"/>\n",
         "    </PPoints>\n"]
    t += ["    <Piece Source=\"particles-%s.%s.vtu\"/>" % (step, int_to_string(r, 3)) for r in range(n_ranks)]
    t += ["  </PUnstructuredGrid>\n", "</VTKFile>\n"]
    return "".join(t)


def file_names(timestep_number, rank):
    """(piece, index) file names of src/pfem2particlehandler.cpp:219 and :297."""
    step = int_to_string(timestep_number, 2)
    return "particles-%s.%s.vtu" % (step, int_to_string(rank, 3)), "particles-%s.pvtu" % step
