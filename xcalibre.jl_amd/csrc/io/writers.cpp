// Result writers (include/xcal_io.h), host C++.
//
// File layouts follow the reference's writers byte for byte:
//   2D legacy ASCII  -- src/IOFormats/VTK/VTK_writer.jl:6-124
//   3D XML polyhedra -- src/IOFormats/VTK/VTK_writer_3D.jl:16-224
//   OpenFOAM case    -- src/IOFormats/OpenFOAM/OpenFOAM_writer.jl:1-359
// Each file is formatted into one buffer and written with one fwrite.
#include "xcal_io.h"

#include "xcal.h"

#include <charconv>
#include <filesystem>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

struct xcalw_writer {
  int dim = 0;
  bool foam = false;
  int64_t n_cells = 0;
  std::string header;  // 2D: everything up to and including "CELL_DATA n\n"
  std::string footer;  // 3D only
  std::string dir;                  // OpenFOAM: the case directory
  std::vector<std::string> patches; // OpenFOAM: patch names
};

namespace {

thread_local std::string g_err;

struct Fail : std::runtime_error {
  using std::runtime_error::runtime_error;
};

void need(bool ok, const std::string& msg) {
  if (!ok) throw Fail(msg);
}

// Julia's print(::Float64): shortest round-trip digits (Ryu, as Julia's
// Base.Ryu.writeshortest), plain when the decimal exponent of the first
// digit is in [-4, 5], else d.ddd e exp with no '+' and no zero padding.
int fmt_double(double x, char* out) {
  char* o = out;
  if (std::isnan(x)) {
    std::memcpy(out, "NaN", 4);
    return 3;
  }
  if (std::isinf(x)) {
    const char* s = x > 0 ? "Inf" : "-Inf";
    std::strcpy(out, s);
    return (int)std::strlen(s);
  }
  if (x == 0.0) {
    const char* s = std::signbit(x) ? "-0.0" : "0.0";
    std::strcpy(out, s);
    return (int)std::strlen(s);
  }
  char tmp[40];
  auto r = std::to_chars(tmp, tmp + sizeof tmp, x, std::chars_format::scientific);
  *r.ptr = '\0';
  const char* p = tmp;
  if (*p == '-') {
    *o++ = '-';
    ++p;
  }
  char dig[24];
  int nd = 0;
  for (; *p && *p != 'e'; ++p)
    if (*p != '.') dig[nd++] = *p;
  const int e = std::atoi(p + 1);  // value = d.ddd * 10^e
  const int pt = e + 1;            // digits before the decimal point
  if (-4 < pt && pt <= 6) {
    if (pt <= 0) {
      *o++ = '0';
      *o++ = '.';
      for (int i = 0; i < -pt; ++i) *o++ = '0';
      for (int i = 0; i < nd; ++i) *o++ = dig[i];
    } else if (pt >= nd) {
      for (int i = 0; i < nd; ++i) *o++ = dig[i];
      for (int i = nd; i < pt; ++i) *o++ = '0';
      *o++ = '.';
      *o++ = '0';
    } else {
      for (int i = 0; i < pt; ++i) *o++ = dig[i];
      *o++ = '.';
      for (int i = pt; i < nd; ++i) *o++ = dig[i];
    }
  } else {
    *o++ = dig[0];
    *o++ = '.';
    if (nd == 1) *o++ = '0';
    for (int i = 1; i < nd; ++i) *o++ = dig[i];
    *o++ = 'e';
    auto q = std::to_chars(o, o + 8, e);
    o = q.ptr;
  }
  *o = '\0';
  return (int)(o - out);
}

struct Buf {
  std::string s;
  Buf& d(double x) {
    char b[40];
    s.append(b, (size_t)fmt_double(x, b));
    return *this;
  }
  Buf& i(int64_t v) {
    char b[24];
    auto r = std::to_chars(b, b + sizeof b, v);
    s.append(b, (size_t)(r.ptr - b));
    return *this;
  }
  Buf& t(const char* x) {
    s.append(x);
    return *this;
  }
  Buf& t(const std::string& x) {
    s.append(x);
    return *this;
  }
};

template <class F>
int guarded(F&& f) {
  try {
    f();
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  } catch (...) {
    g_err = "unknown error";
    return -1;
  }
}

std::string result_name(const xcalw_writer* w, int64_t iteration, double time,
                        const char* suffix) {
  // `iteration == time` compares the Int with the Float64 numerically
  std::string name = (double)iteration == time ? "iteration_" : "time_";
  name += std::to_string(iteration);
  if (suffix) name += suffix;
  name += w->dim == 2 ? ".vtk" : ".vtu";
  return name;
}

void check_cells(int64_t n_nodes, int64_t n_cells, const int64_t* cnr, const int64_t* cn) {
  for (int64_t c = 0; c < n_cells; ++c) {
    need(cnr[2 * c] >= 1 && cnr[2 * c + 1] >= cnr[2 * c] - 1, "bad cell_nodes_range");
    for (int64_t k = cnr[2 * c]; k <= cnr[2 * c + 1]; ++k)
      need(cn[k - 1] >= 1 && cn[k - 1] <= n_nodes, "cell node index out of range");
  }
}

// VTK_writer.jl:26-77: header lines, POINTS, CELLS (" " before every node id,
// so a quad reads "4  0 1 2 3"), CELL_TYPES (3 nodes: 5, 4 nodes: 9, more: 7)
// and the CELL_DATA line.
void build_2d(xcalw_writer* w, int64_t n_nodes, const double* X, int64_t n_cells,
              const int64_t* cnr, const int64_t* cn) {
  check_cells(n_nodes, n_cells, cnr, cn);
  Buf b;
  b.s.reserve((size_t)(n_nodes * 60 + n_cells * 40 + 256));
  b.t("# vtk DataFile Version 3.0\n").t("XCALibre.jl simulation data\n").t("ASCII\n");
  b.t("DATASET UNSTRUCTURED_GRID\n");
  b.t("POINTS ").i(n_nodes).t(" double\n");
  for (int64_t i = 0; i < n_nodes; ++i)
    b.d(X[3 * i]).t(" ").d(X[3 * i + 1]).t(" ").d(X[3 * i + 2]).t("\n");
  int64_t sum = 0;
  for (int64_t c = 0; c < n_cells; ++c) sum += cnr[2 * c + 1] - cnr[2 * c] + 1;
  b.t("CELLS ").i(n_cells).t(" ").i(sum + n_cells).t("\n");
  for (int64_t c = 0; c < n_cells; ++c) {
    b.i(cnr[2 * c + 1] - cnr[2 * c] + 1).t(" ");
    for (int64_t k = cnr[2 * c]; k <= cnr[2 * c + 1]; ++k) b.t(" ").i(cn[k - 1] - 1);
    b.t("\n");
  }
  b.t("CELL_TYPES ").i(n_cells).t("\n");
  for (int64_t c = 0; c < n_cells; ++c) {
    const int64_t nn = cnr[2 * c + 1] - cnr[2 * c] + 1;
    need(nn >= 3, "a 2D cell with fewer than 3 nodes has no VTK cell type");
    b.t(nn == 3 ? "5\n" : nn == 4 ? "9\n" : "7\n");
  }
  b.t("CELL_DATA ").i(n_cells).t("\n");
  w->header = std::move(b.s);
}

struct V3 {
  double x, y, z;
};
V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }

// VTK_writer_3D.jl:16-146.
void build_3d(xcalw_writer* w, int64_t n_nodes, const double* X, int64_t n_cells,
              const int64_t* cnr, const int64_t* cn, const double* cc, int64_t n_faces,
              const int64_t* fnr, const int64_t* fn, const int64_t* fo, const double* fc) {
  check_cells(n_nodes, n_cells, cnr, cn);
  for (int64_t f = 0; f < n_faces; ++f) {
    need(fo[2 * f] >= 1 && fo[2 * f] <= n_cells && fo[2 * f + 1] >= 1 &&
             fo[2 * f + 1] <= n_cells, "face owner out of range");
    need(fnr[2 * f] >= 1 && fnr[2 * f + 1] - fnr[2 * f] + 1 >= 3, "a face needs 3 nodes");
    for (int64_t k = fnr[2 * f]; k <= fnr[2 * f + 1]; ++k)
      need(fn[k - 1] >= 1 && fn[k - 1] <= n_nodes, "face node index out of range");
  }
  auto node = [&](int64_t id) { return V3{X[3 * (id - 1)], X[3 * (id - 1) + 1], X[3 * (id - 1) + 2]}; };
  Buf b;
  b.s.reserve((size_t)(n_nodes * 66 + n_cells * 200 + n_faces * 40 + 1024));
  b.t("<?xml version=\"1.0\"?>\n");
  b.t(" <VTKFile type=\"UnstructuredGrid\" version=\"1.0\" byte_order=\"LittleEndian\">\n");
  b.t("  <UnstructuredGrid>\n");
  b.t("   <Piece NumberOfPoints=\"").i(n_nodes).t("\" NumberOfCells=\"").i(n_cells).t("\">\n");
  b.t("    <Points>\n");
  b.t("     <DataArray type=\"Float64\" NumberOfComponents=\"3\" format=\"ascii\">\n");
  for (int64_t i = 0; i < n_nodes; ++i)
    b.t("      ").d(X[3 * i]).t(" ").d(X[3 * i + 1]).t(" ").d(X[3 * i + 2]).t("\n");
  b.t("     </DataArray>\n").t("    </Points>\n").t("    <Cells>\n");
  b.t("     <DataArray type=\"Int64\" Name=\"connectivity\" format=\"ascii\">\n");
  for (int64_t c = 0; c < n_cells; ++c) {
    b.t("      ");
    for (int64_t k = cnr[2 * c]; k <= cnr[2 * c + 1]; ++k) {
      if (k > cnr[2 * c]) b.t(" ");
      b.i(cn[k - 1] - 1);
    }
    b.t("\n");
  }
  b.t("     </DataArray>\n");
  b.t("     <DataArray type=\"Int64\" Name=\"offsets\" format=\"ascii\">\n");
  int64_t counter = 0;
  for (int64_t c = 0; c < n_cells; ++c) {
    counter += cnr[2 * c + 1] - cnr[2 * c] + 1;
    b.t("      ").i(counter).t("\n");
  }
  b.t("     </DataArray>\n");
  b.t("     <DataArray type=\"Int64\" Name=\"faces\" format=\"ascii\">\n");
  // every face of every cell, boundary faces included, in ascending face
  // order: a CSR of the owner lists (VTK_writer_3D.jl:76-87)
  std::vector<int64_t> start(n_cells + 1, 0), list;
  for (int64_t f = 0; f < n_faces; ++f) {
    ++start[fo[2 * f]];
    if (fo[2 * f + 1] != fo[2 * f]) ++start[fo[2 * f + 1]];
  }
  for (int64_t c = 0; c < n_cells; ++c) start[c + 1] += start[c];
  list.resize(start[n_cells]);
  {
    std::vector<int64_t> fill(start.begin(), start.end() - 1);
    for (int64_t f = 0; f < n_faces; ++f) {
      list[fill[fo[2 * f] - 1]++] = f;
      if (fo[2 * f + 1] != fo[2 * f]) list[fill[fo[2 * f + 1] - 1]++] = f;
    }
  }
  for (int64_t c = 0; c < n_cells; ++c) {
    b.t("\t").i(start[c + 1] - start[c]).t("\n");
    const V3 C{cc[3 * c], cc[3 * c + 1], cc[3 * c + 2]};
    for (int64_t q = start[c]; q < start[c + 1]; ++q) {
      const int64_t f = list[q];
      const int64_t k0 = fnr[2 * f], k1 = fnr[2 * f + 1];
      // normal of the first three nodes, (n2 - n1) x (n3 - n1) / norm,
      // against the face centre seen from the cell centre
      // (VTK_writer_3D.jl:91-113)
      const V3 n1 = node(fn[k0 - 1]), n2 = node(fn[k0]), n3 = node(fn[k0 + 1]);
      const V3 a = sub(n2, n1), d = sub(n3, n1);
      V3 nrm{a.y * d.z - a.z * d.y, a.z * d.x - a.x * d.z, a.x * d.y - a.y * d.x};
      const double len = std::sqrt(nrm.x * nrm.x + nrm.y * nrm.y + nrm.z * nrm.z);
      nrm = {nrm.x / len, nrm.y / len, nrm.z / len};
      const V3 dfc = sub(V3{fc[3 * f], fc[3 * f + 1], fc[3 * f + 2]}, C);
      const bool flip = dfc.x * nrm.x + dfc.y * nrm.y + dfc.z * nrm.z < 0.0;
      b.t("\t").i(k1 - k0 + 1).t(" ");
      for (int64_t j = 0; j <= k1 - k0; ++j) {
        if (j) b.t(" ");
        b.i(fn[(flip ? k1 - j : k0 + j) - 1] - 1);
      }
      b.t("\n");
    }
  }
  b.t("     </DataArray>\n");
  b.t("     <DataArray type=\"Int64\" Name=\"faceoffsets\" format=\"ascii\">\n");
  int64_t x = 0;
  for (int64_t c = 0; c < n_cells; ++c) {
    int64_t total = 0;
    for (int64_t q = start[c]; q < start[c + 1]; ++q)
      total += fnr[2 * list[q] + 1] - fnr[2 * list[q]] + 1;
    x += 1 + (start[c + 1] - start[c]) + total;
    b.t("     ").i(x).t("\n");
  }
  b.t("     </DataArray>\n");
  b.t("     <DataArray type=\"Int64\" Name=\"types\" format=\"ascii\">\n");
  for (int64_t c = 0; c < n_cells; ++c) b.t("      42\n");
  b.t("     </DataArray>\n").t("    </Cells>\n").t("    <CellData>\n");
  w->header = std::move(b.s);
  w->footer = "    </CellData>\n   </Piece>\n  </UnstructuredGrid>\n </VTKFile>\n";
}

// the CELL_DATA / CellData blocks (VTK_writer.jl:79-116,
// VTK_writer_3D.jl:167-218): one value line per cell, three lines per cell
// for a tensor
void write_field(Buf& b, int dim, int64_t n, const xcalw_field& f) {
  need(f.label && f.values, "field without label or values");
  need(f.ncomp == 1 || f.ncomp == 3 || f.ncomp == 9,
       "Input data should be a ScalarField, VectorField or TensorField e.g. (\"U\", U)");
  const double* v = f.values;
  if (dim == 2) {
    if (f.ncomp == 1)
      b.t("SCALARS ").t(f.label).t(" double 1\n").t("LOOKUP_TABLE CellColors\n");
    else
      b.t(f.ncomp == 3 ? "VECTORS " : "TENSORS ").t(f.label).t(" double\n");
  } else {
    b.t("     <DataArray type=\"Float32\" Name=\"").t(f.label).t("\" format=\"ascii\"");
    if (f.ncomp == 3) b.t(" NumberOfComponents=\"3\"");
    if (f.ncomp == 9) b.t(" NumberOfComponents=\"9\"");
    b.t(">\n");
  }
  for (int64_t i = 0; i < n; ++i) {
    if (f.ncomp == 1) {
      b.d(v[i]).t("\n");
    } else {
      for (int r = 0; r < f.ncomp / 3; ++r) {
        const int c0 = 3 * r;
        b.d(v[c0 * n + i]).t(" ").d(v[(c0 + 1) * n + i]).t(" ").d(v[(c0 + 2) * n + i]).t("\n");
      }
    }
  }
  if (dim == 3) b.t("     </DataArray>\n");
}


std::string join_path(const std::string& dir, const std::string& rel) {
  return dir.empty() ? rel : dir + "/" + rel;
}

void write_file(const std::string& path, const std::string& text) {
  FILE* fp = std::fopen(path.c_str(), "wb");
  need(fp != nullptr, "cannot open " + path + " for writing");
  bool ok = std::fwrite(text.data(), 1, text.size(), fp) == text.size();
  ok = (std::fclose(fp) == 0) && ok;
  need(ok, "write to " + path + " failed");
}

void make_dirs(const std::string& path) {
  std::error_code ec;
  std::filesystem::create_directories(path, ec);
  need(!ec, "cannot create directory " + path + ": " + ec.message());
}

// The FoamFile header the reference prints with println of a triple-quoted
// string: the block, then a blank line.
std::string foam_header(const char* cls, const std::string& location, const char* object,
                        const std::string& note = "") {
  std::string h = "FoamFile\n{\n    version     2.0;\n    format      ascii;\n    class       ";
  h += cls;
  h += ";\n";
  if (!note.empty()) h += "    note        \"" + note + "\";\n";
  h += "    location    " + location + ";\n    object      " + object + ";\n}\n\n";
  return h;
}

// OpenFOAM_writer.jl:1-185
void build_foam(xcalw_writer* w, const std::string& dir, int64_t n_nodes, const double* X,
                int64_t n_cells, int64_t n_faces, int64_t n_bf, const int64_t* fnr,
                const int64_t* fn, const int64_t* fo, int32_t n_patches,
                const char* const* names, const int64_t* pr) {
  need(n_bf >= 0 && n_bf <= n_faces, "bad boundary face count");
  for (int64_t f = 0; f < n_faces; ++f) {
    need(fnr[2 * f] >= 1 && fnr[2 * f + 1] >= fnr[2 * f], "bad face_nodes_range");
    need(fo[2 * f] >= 1 && fo[2 * f] <= n_cells && fo[2 * f + 1] >= 1 &&
             fo[2 * f + 1] <= n_cells, "face owner out of range");
  }
  for (int32_t i = 0; i < n_patches; ++i)
    need(names[i] && pr[2 * i] >= 1 && pr[2 * i + 1] <= n_bf, "bad patch");
  w->dir = dir;
  for (int32_t i = 0; i < n_patches; ++i) w->patches.emplace_back(names[i]);
  if (!dir.empty()) make_dirs(dir);
  // touch("XCALibre.foam")
  {
    const std::string t = join_path(dir, "XCALibre.foam");
    FILE* fp = std::fopen(t.c_str(), "ab");
    need(fp != nullptr, "cannot create " + t);
    std::fclose(fp);
  }
  const std::string pm = join_path(dir, "constant/polyMesh");
  make_dirs(pm);
  const int64_t n_if = n_faces - n_bf;
  const std::string loc = "\"constant/polyMesh\"";
  {
    Buf b;
    b.s.reserve((size_t)n_nodes * 40 + 512);
    b.t(foam_header("vectorField", loc, "points")).i(n_nodes).t("\n(\n");
    char line[128];
    for (int64_t i = 0; i < n_nodes; ++i) {
      const int k = std::snprintf(line, sizeof line, "(%g %g %g)\n", X[3 * i], X[3 * i + 1],
                                  X[3 * i + 2]);
      b.s.append(line, (size_t)k);
    }
    b.t(")\n");
    write_file(pm + "/points", b.s);
  }
  // internal faces first, then the boundary faces (OpenFOAM numbering)
  auto face_order = [&](auto&& emit) {
    for (int64_t f = n_bf; f < n_faces; ++f) emit(f);
    for (int64_t f = 0; f < n_bf; ++f) emit(f);
  };
  {
    Buf b;
    b.s.reserve((size_t)n_faces * 40 + 512);
    b.t(foam_header("faceList", loc, "faces")).i(n_faces).t("\n(\n");
    face_order([&](int64_t f) {
      b.i(fnr[2 * f + 1] - fnr[2 * f] + 1).t("(");
      for (int64_t k = fnr[2 * f]; k <= fnr[2 * f + 1]; ++k) b.i(fn[k - 1] - 1).t(" ");
      b.t(")\n");
    });
    b.t(")\n");
    write_file(pm + "/faces", b.s);
  }
  const std::string note = "nPoints: " + std::to_string(n_nodes) + " nCells: " +
                           std::to_string(n_cells) + " nFaces: " + std::to_string(n_faces) +
                           " nInternalFaces: " + std::to_string(n_if);
  {
    Buf b;
    b.t(foam_header("labelList", loc, "owner", note)).i(n_faces).t("\n(\n");
    face_order([&](int64_t f) { b.i(fo[2 * f] - 1).t("\n"); });
    b.t(")\n");
    write_file(pm + "/owner", b.s);
  }
  {
    Buf b;
    b.t(foam_header("labelList", loc, "neighbour", note)).i(n_if).t("\n(\n");
    for (int64_t f = n_bf; f < n_faces; ++f) b.i(fo[2 * f + 1] - 1).t("\n");
    b.t(")\n");
    write_file(pm + "/neighbour", b.s);
  }
  {
    Buf b;
    b.t(foam_header("polyBoundaryMesh", loc, "boundary")).i(n_patches).t("\n(\n");
    for (int32_t i = 0; i < n_patches; ++i) {
      b.t(names[i]).t("\n{\n    type            patch;\n    nFaces          ");
      b.i(pr[2 * i + 1] - pr[2 * i] + 1).t(";\n    startFace       ");
      b.i(pr[2 * i] + n_if - 1).t(";\n}\n");
    }
    b.t(")\n");
    write_file(pm + "/boundary", b.s);
  }
}

std::string foam_time_dir(int64_t iteration, double time) {
  if ((double)iteration == time) return std::to_string(iteration);
  char b[64];
  std::snprintf(b, sizeof b, "%.8f", time);
  return b;
}

// _foam_boundary_entry (OpenFOAM_writer.jl:277-359)
void foam_bc_entry(Buf& b, const xcalw_foam_bc& bc, int ncomp) {
  const bool fixed = bc.kind == XCAL_BC_DIRICHLET || (bc.kind == XCAL_BC_WALL && ncomp == 3);
  if (bc.kind == XCAL_BC_EMPTY) {
    b.t("\t{\n\t\ttype empty;\n\t}\n");
  } else if (fixed) {
    b.t("\t{\n\t\ttype fixedValue;\n\t\tvalue uniform ");
    if (ncomp == 1)
      b.d(bc.value[0]);
    else
      b.t("(").d(bc.value[0]).t(" ").d(bc.value[1]).t(" ").d(bc.value[2]).t(")");
    b.t(";\n\t}\n");
  } else {
    b.t("\t{\n\t\ttype zeroGradient;\n\t}\n");
  }
}

// OpenFOAM_writer.jl:191-273
void write_foam_field(const xcalw_writer* w, const std::string& tdir, int64_t iteration,
                      const xcalw_foam_field& f) {
  need(f.label && f.values, "field without label or values");
  need(f.ncomp == 1 || f.ncomp == 3,
       "Input data should be a ScalarField or VectorField e.g. (\"U\", U)");
  need(f.n_bcs >= 0 && (f.n_bcs == 0 || f.bcs), "bad boundary conditions");
  const int64_t n = w->n_cells;
  Buf b;
  b.s.reserve((size_t)n * 26 * f.ncomp + 1024);
  b.t(foam_header(f.ncomp == 1 ? "volScalarField" : "volVectorField",
                  "\"" + std::to_string(iteration) + "\"", f.label));
  b.t(f.ncomp == 1 ? "internalField   nonuniform List<scalar>\n"
                   : "internalField   nonuniform List<vector>\n");
  b.i(n).t("\n(\n");
  const double* v = f.values;
  for (int64_t i = 0; i < n; ++i) {
    if (f.ncomp == 1)
      b.d(v[i]).t("\n");
    else
      b.t("(").d(v[i]).t(" ").d(v[n + i]).t(" ").d(v[2 * n + i]).t(")\n");
  }
  b.t(");\n").t("boundaryField\n").t("{\n");
  for (int32_t k = 0; k < f.n_bcs; ++k) {
    const xcalw_foam_bc& bc = f.bcs[k];
    need(bc.patch >= 0 && bc.patch < (int32_t)w->patches.size(), "bad patch index");
    b.t("\t").t(w->patches[bc.patch]).t("\n");
    foam_bc_entry(b, bc, f.ncomp);
    b.t("\n");
  }
  b.t("}\n");
  write_file(tdir + "/" + f.label, b.s);
}

}  // namespace

extern "C" {

const char* xcalw_last_error(void) { return g_err.c_str(); }

int32_t xcalw_format_double(double x, char* buf) { return fmt_double(x, buf); }

int xcalw_vtk2d_init(int64_t n_nodes, const double* node_coords, int64_t n_cells,
                     const int64_t* cell_nodes_range, const int64_t* cell_nodes,
                     xcalw_writer** out) {
  return guarded([&] {
    need(out && node_coords && cell_nodes_range && cell_nodes && n_nodes >= 0 && n_cells >= 0,
         "bad arguments");
    auto* w = new xcalw_writer;
    try {
      w->dim = 2;
      w->n_cells = n_cells;
      build_2d(w, n_nodes, node_coords, n_cells, cell_nodes_range, cell_nodes);
    } catch (...) {
      delete w;
      throw;
    }
    *out = w;
  });
}

int xcalw_vtu3d_init(int64_t n_nodes, const double* node_coords, int64_t n_cells,
                     const int64_t* cell_nodes_range, const int64_t* cell_nodes,
                     const double* cell_centre, int64_t n_faces,
                     const int64_t* face_nodes_range, const int64_t* face_nodes,
                     const int64_t* face_owner, const double* face_centre,
                     xcalw_writer** out) {
  return guarded([&] {
    need(out && node_coords && cell_nodes_range && cell_nodes && cell_centre &&
             face_nodes_range && face_nodes && face_owner && face_centre && n_nodes >= 0 &&
             n_cells >= 0 && n_faces >= 0,
         "bad arguments");
    auto* w = new xcalw_writer;
    try {
      w->dim = 3;
      w->n_cells = n_cells;
      build_3d(w, n_nodes, node_coords, n_cells, cell_nodes_range, cell_nodes, cell_centre,
               n_faces, face_nodes_range, face_nodes, face_owner, face_centre);
    } catch (...) {
      delete w;
      throw;
    }
    *out = w;
  });
}

int32_t xcalw_result_name(const xcalw_writer* w, int64_t iteration, double time,
                          const char* suffix, char* buf, int32_t cap) {
  if (!w || !buf) return -1;
  const std::string name = result_name(w, iteration, time, suffix);
  if ((int64_t)name.size() + 1 > cap) return -1;
  std::memcpy(buf, name.c_str(), name.size() + 1);
  return (int32_t)name.size();
}

int xcalw_write_results(const xcalw_writer* w, int64_t iteration, double time,
                        const char* suffix, const char* directory, int32_t n_fields,
                        const xcalw_field* fields) {
  return guarded([&] {
    need(w && !w->foam && n_fields >= 0 && (n_fields == 0 || fields), "bad arguments");
    std::string path = result_name(w, iteration, time, suffix);
    if (directory && *directory) path = std::string(directory) + "/" + path;
    Buf b;
    size_t per = 0;
    for (int32_t i = 0; i < n_fields; ++i) per += (size_t)fields[i].ncomp * 26;
    b.s.reserve((size_t)w->n_cells * per + 256 * (size_t)(n_fields + 1));
    for (int32_t i = 0; i < n_fields; ++i) write_field(b, w->dim, w->n_cells, fields[i]);
    FILE* fp = std::fopen(path.c_str(), "wb");
    need(fp != nullptr, "cannot open " + path + " for writing");
    bool ok = std::fwrite(w->header.data(), 1, w->header.size(), fp) == w->header.size();
    ok = ok && std::fwrite(b.s.data(), 1, b.s.size(), fp) == b.s.size();
    ok = ok && std::fwrite(w->footer.data(), 1, w->footer.size(), fp) == w->footer.size();
    ok = (std::fclose(fp) == 0) && ok;
    need(ok, "write to " + path + " failed");
  });
}

int xcalw_foam_init(const char* directory, int32_t dim, int64_t n_nodes,
                    const double* node_coords, int64_t n_cells, int64_t n_faces,
                    int64_t n_bfaces, const int64_t* face_nodes_range,
                    const int64_t* face_nodes, const int64_t* face_owner, int32_t n_patches,
                    const char* const* patch_names, const int64_t* patch_range,
                    xcalw_writer** out) {
  return guarded([&] {
    need(dim == 3,
         "The OpenFOAM format can only be used for 3D simulations. Use `output=VTK()` "
         "instead.");
    need(out && node_coords && face_nodes_range && face_nodes && face_owner && n_nodes >= 0 &&
             n_cells >= 0 && n_faces >= 0 && n_patches >= 0 &&
             (n_patches == 0 || (patch_names && patch_range)),
         "bad arguments");
    auto* w = new xcalw_writer;
    try {
      w->dim = 3;
      w->foam = true;
      w->n_cells = n_cells;
      build_foam(w, directory ? directory : "", n_nodes, node_coords, n_cells, n_faces,
                 n_bfaces, face_nodes_range, face_nodes, face_owner, n_patches, patch_names,
                 patch_range);
    } catch (...) {
      delete w;
      throw;
    }
    *out = w;
  });
}

int32_t xcalw_foam_time_dir(int64_t iteration, double time, char* buf, int32_t cap) {
  if (!buf) return -1;
  const std::string d = foam_time_dir(iteration, time);
  if ((int64_t)d.size() + 1 > cap) return -1;
  std::memcpy(buf, d.c_str(), d.size() + 1);
  return (int32_t)d.size();
}

int xcalw_foam_write_results(const xcalw_writer* w, int64_t iteration, double time,
                             int32_t n_fields, const xcalw_foam_field* fields) {
  return guarded([&] {
    need(w && w->foam && n_fields >= 0 && (n_fields == 0 || fields), "bad arguments");
    const std::string tdir = join_path(w->dir, foam_time_dir(iteration, time));
    make_dirs(tdir);
    for (int32_t i = 0; i < n_fields; ++i) write_foam_field(w, tdir, iteration, fields[i]);
  });
}

void xcalw_free(xcalw_writer* w) { delete w; }

}  // extern "C"
