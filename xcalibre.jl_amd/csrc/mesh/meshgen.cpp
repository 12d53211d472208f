// Host mesh ingest and generation for the MI355X finite-volume path.
//
// Every builder produces the reference's Mesh2/Mesh3 array layout
// (src/Mesh/Mesh_0_types.jl:135-185): 1-based Int64 indices, Float64
// geometry, boundary faces first.  Floating-point expressions are written in
// the same evaluation order as the Julia code they restate (StaticArrays dot
// products left to right, a*b/c as (a*b)/c) so that a mesh built here is the
// mesh the reference would build from the same input.  Vector norms use
// sqrt(x*x+y*y+z*z).
#include "xcal_mesh.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <limits>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

struct V3 {
  double x, y, z;
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
inline V3 operator*(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline V3 operator*(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 operator/(V3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline double norm(V3 a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
inline V3 cross(V3 a, V3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
// Julia eps(x): spacing from |x| to the next larger float.
inline double jeps(double x) {
  double a = std::fabs(x);
  return std::nextafter(a, std::numeric_limits<double>::infinity()) - a;
}

}  // namespace

// Result container, reference layout.
struct xcalm_mesh {
  int32_t dim = 3;
  int64_t n_cells = 0, n_faces = 0, n_bfaces = 0, n_nodes = 0;
  std::vector<double> cell_centre, cell_volume;
  std::vector<int64_t> cell_faces_range, cell_nodes_range, cell_nodes;
  std::vector<int64_t> cell_faces, cell_neighbours, cell_nsign;
  std::vector<int64_t> face_nodes_range, face_nodes, face_owner;
  std::vector<double> face_centre, face_normal, face_e, face_area, face_delta,
      face_weight;
  std::vector<int64_t> boundary_cellsID;
  std::vector<std::string> patch_names;
  std::vector<int64_t> patch_range;
  std::vector<double> node_coords;
};

namespace {

// ---------------------------------------------------------------------------
// Mesh.weight_delta_e (src/Mesh/Mesh_1_functions.jl:20-51)
// ---------------------------------------------------------------------------
void weight_delta_e_internal(V3 C1F1, V3 C2F1, V3 C1C2, V3 normal, double& weight,
                             double& delta, V3& e) {
  double projection = dot(C1C2, normal);
  if (std::isfinite(projection) && std::fabs(projection) > jeps(projection)) {
    double wi = dot(C1F1, normal) / projection;
    weight = 1.0 - wi;
  } else {
    double d1 = norm(C1F1), d2 = norm(C2F1);
    double dsum = d1 + d2;
    weight = dsum > 0.0 ? d2 / dsum : 0.5;
  }
  delta = norm(C1C2);
  if (delta > 0.0) {
    e = C1C2 / delta;
  } else {
    delta = norm(C1F1) + norm(C2F1);
    e = normal;
  }
  delta = std::max(delta, jeps(1.0));
  weight = std::min(std::max(weight, 0.0), 1.0);
}

void weight_delta_e_boundary(V3 C1F1, V3 normal, double& weight, double& delta,
                             V3& e) {
  weight = 1.0;
  delta = norm(C1F1);
  e = delta > 0.0 ? C1F1 / delta : normal;
  delta = std::max(delta, jeps(1.0));
}

// ---------------------------------------------------------------------------
// FOAM path
// ---------------------------------------------------------------------------
struct FoamData {
  std::vector<V3> points;
  std::vector<int64_t> face_nodes;   // 1-based node ids, foam face order
  std::vector<int64_t> fn_first;     // 1-based first index per foam face
  std::vector<int64_t> fn_last;      // 1-based last index per foam face
  std::vector<int64_t> face_owner;   // 1-based, n_faces
  std::vector<int64_t> face_neighbour;  // 1-based, n_faces (boundary: owner)
  int64_t n_cells = 0, n_faces = 0, n_ifaces = 0, n_bfaces = 0;
  std::vector<std::string> bnames;
  std::vector<int64_t> bnfaces, bstart;  // startFace 1-based
};

// assign_faces! (src/FoamMesh/FoamMesh_1_read.jl:24-41)
void foam_assign_faces(FoamData& fd, const std::vector<int64_t>& neighbours) {
  fd.n_faces = (int64_t)fd.face_owner.size();
  fd.n_ifaces = (int64_t)neighbours.size();
  fd.n_bfaces = fd.n_faces - fd.n_ifaces;
  int64_t mo = 0, mn = 0;
  for (auto v : fd.face_owner) mo = std::max(mo, v);
  for (auto v : neighbours) mn = std::max(mn, v);
  fd.n_cells = std::max(mo, mn);
  fd.face_neighbour.assign(fd.n_faces, 0);
  for (int64_t fi = 0; fi < fd.n_ifaces; ++fi) fd.face_neighbour[fi] = neighbours[fi];
  for (int64_t fi = fd.n_ifaces; fi < fd.n_faces; ++fi)
    fd.face_neighbour[fi] = fd.face_owner[fi];
}

void foam_build(FoamData& fd, xcalm_mesh& m) {
  const int64_t nc = fd.n_cells, nf = fd.n_faces, nif = fd.n_ifaces,
                nbf = fd.n_bfaces;
  if (nc <= 0 || nf <= 0) throw std::runtime_error("empty polyMesh");
  for (int64_t fi = 0; fi < nf; ++fi) {
    if (fd.face_owner[fi] < 1 || fd.face_owner[fi] > nc ||
        fd.face_neighbour[fi] < 1 || fd.face_neighbour[fi] > nc)
      throw std::runtime_error("owner/neighbour index out of range");
  }
  const int64_t np = (int64_t)fd.points.size();
  m.dim = 3;
  m.n_cells = nc;
  m.n_faces = nf;
  m.n_bfaces = nbf;
  m.n_nodes = np;

  // connect_cell_faces (FoamMesh_2_connect.jl:24-69)
  std::vector<int64_t> count(nc, 0), offset(nc, 0);
  for (int64_t fi = 0; fi < nif; ++fi) {
    count[fd.face_owner[fi] - 1] += 1;
    count[fd.face_neighbour[fi] - 1] += 1;
  }
  m.cell_faces_range.assign(2 * nc, 0);
  for (int64_t c = 0; c < nc; ++c) {
    offset[c] = c == 0 ? 0 : offset[c - 1] + count[c - 1];
    m.cell_faces_range[2 * c] = offset[c] + 1;
    m.cell_faces_range[2 * c + 1] = offset[c] + count[c];
  }
  m.cell_faces.assign(2 * nif, 0);
  m.cell_nsign.assign(2 * nif, 0);
  m.cell_neighbours.assign(2 * nif, 0);
  {
    std::vector<int64_t> cursor(offset);
    for (int64_t fi = 0; fi < nif; ++fi) {
      int64_t o = fd.face_owner[fi], n = fd.face_neighbour[fi];
      int64_t fID = fi + 1 + nbf;
      int64_t k = ++cursor[o - 1];
      m.cell_faces[k - 1] = fID;
      m.cell_nsign[k - 1] = 1;
      m.cell_neighbours[k - 1] = n;
      k = ++cursor[n - 1];
      m.cell_faces[k - 1] = fID;
      m.cell_nsign[k - 1] = -1;
      m.cell_neighbours[k - 1] = o;
    }
  }

  // connect_cell_nodes (FoamMesh_2_connect.jl:71-142)
  {
    std::vector<int64_t> rcount(nc, 0), roff(nc, 0);
    for (int64_t fi = 0; fi < nf; ++fi) {
      int64_t nn = fd.fn_last[fi] - fd.fn_first[fi] + 1;
      rcount[fd.face_owner[fi] - 1] += nn;
      rcount[fd.face_neighbour[fi] - 1] += nn;
    }
    for (int64_t c = 0; c < nc; ++c) roff[c] = c == 0 ? 0 : roff[c - 1] + rcount[c - 1];
    std::vector<int64_t> raw(roff[nc - 1] + rcount[nc - 1]);
    std::vector<int64_t> rcur(roff);
    for (int64_t fi = 0; fi < nf; ++fi) {
      int64_t o = fd.face_owner[fi], nb = fd.face_neighbour[fi];
      for (int64_t k = fd.fn_first[fi]; k <= fd.fn_last[fi]; ++k)
        raw[(++rcur[o - 1]) - 1] = fd.face_nodes[k - 1];
      for (int64_t k = fd.fn_first[fi]; k <= fd.fn_last[fi]; ++k)
        raw[(++rcur[nb - 1]) - 1] = fd.face_nodes[k - 1];
    }
    std::vector<int64_t> seen(np, 0), dcount(nc, 0), doff(nc, 0);
    for (int64_t c = 0; c < nc; ++c)
      for (int64_t k = roff[c]; k < roff[c] + rcount[c]; ++k) {
        int64_t nid = raw[k];
        if (seen[nid - 1] != c + 1) {
          dcount[c] += 1;
          seen[nid - 1] = c + 1;
        }
      }
    m.cell_nodes_range.assign(2 * nc, 0);
    for (int64_t c = 0; c < nc; ++c) {
      doff[c] = c == 0 ? 0 : doff[c - 1] + dcount[c - 1];
      m.cell_nodes_range[2 * c] = doff[c] + 1;
      m.cell_nodes_range[2 * c + 1] = doff[c] + dcount[c];
    }
    m.cell_nodes.assign(doff[nc - 1] + dcount[nc - 1], 0);
    std::fill(seen.begin(), seen.end(), 0);
    std::vector<int64_t> dcur(doff);
    for (int64_t c = 0; c < nc; ++c)
      for (int64_t k = roff[c]; k < roff[c] + rcount[c]; ++k) {
        int64_t nid = raw[k];
        if (seen[nid - 1] != c + 1) {
          m.cell_nodes[(++dcur[c]) - 1] = nid;
          seen[nid - 1] = c + 1;
        }
      }
  }

  // connect_face_nodes (FoamMesh_2_connect.jl:144-195): boundary faces first
  m.face_nodes.assign(fd.face_nodes.size(), 0);
  m.face_nodes_range.assign(2 * nf, 0);
  {
    int64_t nodei = 0, start = 1, fID = 0;
    auto emit = [&](int64_t src) {
      int64_t nn = fd.fn_last[src] - fd.fn_first[src] + 1;
      m.face_nodes_range[2 * fID] = start;
      m.face_nodes_range[2 * fID + 1] = start + nn - 1;
      start += nn;
      for (int64_t k = fd.fn_first[src]; k <= fd.fn_last[src]; ++k)
        m.face_nodes[nodei++] = fd.face_nodes[k - 1];
      ++fID;
    };
    for (int64_t b = nif; b < nf; ++b) emit(b);
    for (int64_t b = 0; b < nif; ++b) emit(b);
  }

  // generate_faces (FoamMesh_3_generate.jl:46-64)
  m.face_owner.assign(2 * nf, 0);
  for (int64_t fID = 1; fID <= nf; ++fID) {
    int64_t src = fID <= nbf ? (nif + fID) : (fID - nbf);
    m.face_owner[2 * (fID - 1)] = fd.face_owner[src - 1];
    m.face_owner[2 * (fID - 1) + 1] = fd.face_neighbour[src - 1];
  }

  // generate_boundaries (FoamMesh_3_generate.jl:66-81)
  m.patch_names = fd.bnames;
  m.patch_range.assign(2 * fd.bnames.size(), 0);
  {
    int64_t start = 1;
    for (size_t bi = 0; bi < fd.bnames.size(); ++bi) {
      int64_t end = start + fd.bnfaces[bi] - 1;
      m.patch_range[2 * bi] = start;
      m.patch_range[2 * bi + 1] = end;
      start += fd.bnfaces[bi];
    }
  }
  // generate_boundary_cells (FoamMesh_3_generate.jl:32-44)
  m.boundary_cellsID.assign(nbf, 0);
  for (size_t bi = 0; bi < fd.bnames.size(); ++bi)
    for (int64_t fID = m.patch_range[2 * bi]; fID <= m.patch_range[2 * bi + 1]; ++fID)
      m.boundary_cellsID[fID - 1] = m.face_owner[2 * (fID - 1)];

  m.node_coords.resize(3 * np);
  for (int64_t i = 0; i < np; ++i) {
    m.node_coords[3 * i] = fd.points[i].x;
    m.node_coords[3 * i + 1] = fd.points[i].y;
    m.node_coords[3 * i + 2] = fd.points[i].z;
  }

  // compute_geometry! (FoamMesh_4_geometry.jl:1-158)
  std::vector<V3> ccentre(nc), fcentre(nf), fnormal(nf), fe(nf);
  std::vector<double> farea(nf), fdelta(nf), fweight(nf);
  for (int64_t c = 0; c < nc; ++c) {
    V3 s{0, 0, 0};
    int64_t a = m.cell_nodes_range[2 * c], b = m.cell_nodes_range[2 * c + 1];
    for (int64_t k = a; k <= b; ++k) s = s + fd.points[m.cell_nodes[k - 1] - 1];
    ccentre[c] = s / (double)(b - a + 1);
  }
  for (int64_t f = 0; f < nf; ++f) {
    V3 s{0, 0, 0};
    int64_t a = m.face_nodes_range[2 * f], b = m.face_nodes_range[2 * f + 1];
    for (int64_t k = a; k <= b; ++k) s = s + fd.points[m.face_nodes[k - 1] - 1];
    fcentre[f] = s / (double)(b - a + 1);
  }
  // _face_geometry: fan triangles about the node-average apex
  auto face_geometry = [&](int64_t f, V3 apex, V3& normal, double& mag, V3& centre) {
    int64_t a = m.face_nodes_range[2 * f], b = m.face_nodes_range[2 * f + 1];
    int64_t n = b - a + 1;
    V3 area_vec{0, 0, 0}, centre_sum{0, 0, 0};
    double sum_area = 0.0;
    for (int64_t i = 0; i < n; ++i) {
      int64_t j = i == n - 1 ? 0 : i + 1;
      V3 p1 = fd.points[m.face_nodes[a - 1 + i] - 1];
      V3 p2 = fd.points[m.face_nodes[a - 1 + j] - 1];
      V3 tri_vec = cross(p1 - apex, p2 - apex) / 2.0;
      double tri_area = norm(tri_vec);
      if (!(tri_area > 0.0)) continue;
      area_vec = area_vec + tri_vec;
      centre_sum = centre_sum + (tri_area * ((apex + p1) + p2)) / 3.0;
      sum_area += tri_area;
    }
    mag = norm(area_vec);
    normal = mag > 0.0 ? area_vec / mag : V3{0, 0, 0};
    centre = sum_area > 0.0 ? centre_sum / sum_area : apex;
  };
  for (int64_t f = 0; f < nbf; ++f) {
    V3 normal, centre;
    double area;
    face_geometry(f, fcentre[f], normal, area, centre);
    V3 C1 = ccentre[m.face_owner[2 * f] - 1];
    V3 C1F1 = centre - C1;
    if (dot(C1F1, normal) < 0.0) normal = -normal;
    double w, d;
    V3 e;
    weight_delta_e_boundary(C1F1, normal, w, d, e);
    fcentre[f] = centre;
    fnormal[f] = normal;
    farea[f] = area;
    fdelta[f] = d;
    fe[f] = e;
    fweight[f] = w;
  }
  for (int64_t f = nbf; f < nf; ++f) {
    V3 normal, centre;
    double area;
    face_geometry(f, fcentre[f], normal, area, centre);
    V3 C1 = ccentre[m.face_owner[2 * f] - 1];
    V3 C2 = ccentre[m.face_owner[2 * f + 1] - 1];
    V3 C1F1 = centre - C1, C2F1 = centre - C2, C1C2 = C2 - C1;
    if (dot(C1C2, normal) < 0.0) normal = -normal;
    double w, d;
    V3 e;
    weight_delta_e_internal(C1F1, C2F1, C1C2, normal, w, d, e);
    fcentre[f] = centre;
    fnormal[f] = normal;
    farea[f] = area;
    fdelta[f] = d;
    fe[f] = e;
    fweight[f] = w;
  }
  // calculate_cell_volumes!
  const double oneThird = 1.0 / 3.0;
  std::vector<double> vol(nc, 0.0), maxa(nc, 0.0);
  for (int64_t c = 0; c < nc; ++c) {
    for (int64_t k = m.cell_faces_range[2 * c]; k <= m.cell_faces_range[2 * c + 1]; ++k) {
      int64_t f = m.cell_faces[k - 1] - 1;
      double ns = (double)m.cell_nsign[k - 1];
      V3 xf = fcentre[f] - ccentre[c];
      vol[c] += ((oneThird * dot(xf, fnormal[f])) * ns) * farea[f];
      maxa[c] = std::max(maxa[c], farea[f]);
    }
  }
  for (size_t bi = 0; bi < fd.bnames.size(); ++bi)
    for (int64_t fID = m.patch_range[2 * bi]; fID <= m.patch_range[2 * bi + 1]; ++fID) {
      int64_t c = m.boundary_cellsID[fID - 1] - 1, f = fID - 1;
      V3 xf = fcentre[f] - ccentre[c];
      vol[c] += (oneThird * dot(xf, fnormal[f])) * farea[f];
      maxa[c] = std::max(maxa[c], farea[f]);
    }
  int64_t fixed = 0;
  for (int64_t c = 0; c < nc; ++c) {
    double V = vol[c];
    if (!(std::isfinite(V) && V > 0.0)) {
      double est = std::pow(maxa[c], 1.5) * 1e-3;
      V = std::max(std::isfinite(V) ? std::fabs(V) : 0.0, est);
      ++fixed;
    }
    vol[c] = V;
  }
  if (fixed > 0)
    std::fprintf(stderr, "compute_geometry!: %lld cell(s) had non-positive volume\n",
                 (long long)fixed);

  m.cell_centre.resize(3 * nc);
  m.cell_volume = vol;
  for (int64_t c = 0; c < nc; ++c) {
    m.cell_centre[3 * c] = ccentre[c].x;
    m.cell_centre[3 * c + 1] = ccentre[c].y;
    m.cell_centre[3 * c + 2] = ccentre[c].z;
  }
  m.face_centre.resize(3 * nf);
  m.face_normal.resize(3 * nf);
  m.face_e.resize(3 * nf);
  for (int64_t f = 0; f < nf; ++f) {
    m.face_centre[3 * f] = fcentre[f].x;
    m.face_centre[3 * f + 1] = fcentre[f].y;
    m.face_centre[3 * f + 2] = fcentre[f].z;
    m.face_normal[3 * f] = fnormal[f].x;
    m.face_normal[3 * f + 1] = fnormal[f].y;
    m.face_normal[3 * f + 2] = fnormal[f].z;
    m.face_e[3 * f] = fe[f].x;
    m.face_e[3 * f + 1] = fe[f].y;
    m.face_e[3 * f + 2] = fe[f].z;
  }
  m.face_area = farea;
  m.face_delta = fdelta;
  m.face_weight = fweight;
}

// ---------------------------------------------------------------------------
// PRNG for the generated hex box: SplitMix64 seeding xoshiro256**.
// ---------------------------------------------------------------------------
struct Xoshiro {
  uint64_t s[4];
  static uint64_t splitmix(uint64_t& x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  explicit Xoshiro(uint64_t seed) {
    uint64_t x = seed;
    for (auto& v : s) v = splitmix(x);
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    uint64_t r = rotl(s[1] * 5, 7) * 9;
    uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  double u01() { return (double)(next() >> 11) * 0x1.0p-53; }
};

// ---------------------------------------------------------------------------
// BlockMesher2D restatement
// ---------------------------------------------------------------------------
struct BEdge {
  std::vector<int64_t> nodesID;  // 1-based
  int64_t ncells;
  bool boundary;
};
struct BBlock {
  int64_t edgesID[4];
  int64_t nx, ny;
  std::vector<int64_t> nodesID;     // (nx+1) x (ny+1), column-major
  std::vector<int64_t> elementsID;  // nx x ny
  std::vector<int64_t> facesNS;     // nx x (ny+1)
  std::vector<int64_t> facesEW;     // (nx+1) x ny
  int64_t& node(int64_t i, int64_t j) { return nodesID[(i - 1) + (nx + 1) * (j - 1)]; }
  int64_t& elem(int64_t i, int64_t j) { return elementsID[(i - 1) + nx * (j - 1)]; }
  int64_t& ns(int64_t i, int64_t j) { return facesNS[(i - 1) + nx * (j - 1)]; }
  int64_t& ew(int64_t i, int64_t j) { return facesEW[(i - 1) + (nx + 1) * (j - 1)]; }
};
struct OldFace {
  int64_t n1 = 0, n2 = 0;
  int64_t o1 = 0, o2 = 0;
  V3 centre{0, 0, 0}, normal{0, 0, 0}, e{0, 0, 0};
  double area = 0, delta = 0, weight = 0;
};
struct OldCell {
  std::vector<int64_t> nodesID, facesID, neighbours, nsign;
  V3 centre{0, 0, 0};
  double volume = 0;
};
struct OldBoundary {
  std::string name;
  std::vector<int64_t> facesID, cellsID;
};

V3 geometric_centre(const std::vector<V3>& nodes, const int64_t* ids, int64_t n) {
  V3 s{0, 0, 0};
  for (int64_t k = 0; k < n; ++k) s = s + nodes[ids[k] - 1];
  return s / (double)n;
}

// UNV2.update_mesh_format (src/UNV2/UNV2_4_update_format.jl:2-171)
void update_mesh_format(const std::vector<V3>& nodes, const std::vector<OldCell>& cells,
                        const std::vector<OldFace>& faces,
                        const std::vector<OldBoundary>& bnds, xcalm_mesh& m) {
  const int64_t nc = (int64_t)cells.size(), nf = (int64_t)faces.size();
  m.dim = 2;
  m.n_cells = nc;
  m.n_faces = nf;
  m.n_nodes = (int64_t)nodes.size();
  m.patch_names.clear();
  m.patch_range.clear();
  for (auto& b : bnds) {
    if (b.facesID.empty()) throw std::runtime_error("empty boundary patch " + b.name);
    m.patch_names.push_back(b.name);
    m.patch_range.push_back(*std::min_element(b.facesID.begin(), b.facesID.end()));
    m.patch_range.push_back(*std::max_element(b.facesID.begin(), b.facesID.end()));
  }
  int64_t nbf = *std::max_element(bnds.back().facesID.begin(), bnds.back().facesID.end());
  m.n_bfaces = nbf;
  m.boundary_cellsID.assign(nbf, 0);
  for (auto& b : bnds)
    for (size_t i = 0; i < b.facesID.size(); ++i) m.boundary_cellsID[b.facesID[i] - 1] = b.cellsID[i];
  m.node_coords.resize(3 * nodes.size());
  for (size_t i = 0; i < nodes.size(); ++i) {
    m.node_coords[3 * i] = nodes[i].x;
    m.node_coords[3 * i + 1] = nodes[i].y;
    m.node_coords[3 * i + 2] = nodes[i].z;
  }
  m.cell_nodes.clear();
  m.cell_faces.clear();
  m.cell_neighbours.clear();
  m.cell_nsign.clear();
  m.cell_nodes_range.assign(2 * nc, 0);
  m.cell_faces_range.assign(2 * nc, 0);
  m.cell_centre.resize(3 * nc);
  m.cell_volume.resize(nc);
  int64_t ni = 1, fi = 1;
  for (int64_t i = 0; i < nc; ++i) {
    const OldCell& c = cells[i];
    m.cell_nodes_range[2 * i] = ni;
    m.cell_nodes_range[2 * i + 1] = ni + (int64_t)c.nodesID.size() - 1;
    for (auto v : c.nodesID) {
      m.cell_nodes.push_back(v);
      ++ni;
    }
    m.cell_faces_range[2 * i] = fi;
    m.cell_faces_range[2 * i + 1] = fi + (int64_t)c.facesID.size() - 1;
    for (size_t j = 0; j < c.facesID.size(); ++j) {
      m.cell_faces.push_back(c.facesID[j]);
      m.cell_neighbours.push_back(c.neighbours[j]);
      m.cell_nsign.push_back(c.nsign[j]);
      ++fi;
    }
    m.cell_centre[3 * i] = c.centre.x;
    m.cell_centre[3 * i + 1] = c.centre.y;
    m.cell_centre[3 * i + 2] = c.centre.z;
    m.cell_volume[i] = c.volume;
  }
  m.face_nodes.clear();
  m.face_nodes_range.assign(2 * nf, 0);
  m.face_owner.assign(2 * nf, 0);
  m.face_centre.resize(3 * nf);
  m.face_normal.resize(3 * nf);
  m.face_e.resize(3 * nf);
  m.face_area.resize(nf);
  m.face_delta.resize(nf);
  m.face_weight.resize(nf);
  ni = 1;
  for (int64_t i = 0; i < nf; ++i) {
    const OldFace& f = faces[i];
    m.face_nodes_range[2 * i] = ni;
    m.face_nodes_range[2 * i + 1] = ni + 1;
    m.face_nodes.push_back(f.n1);
    m.face_nodes.push_back(f.n2);
    ni += 2;
    m.face_owner[2 * i] = f.o1;
    m.face_owner[2 * i + 1] = f.o2;
    m.face_centre[3 * i] = f.centre.x;
    m.face_centre[3 * i + 1] = f.centre.y;
    m.face_centre[3 * i + 2] = f.centre.z;
    m.face_normal[3 * i] = f.normal.x;
    m.face_normal[3 * i + 1] = f.normal.y;
    m.face_normal[3 * i + 2] = f.normal.z;
    m.face_e[3 * i] = f.e.x;
    m.face_e[3 * i + 1] = f.e.y;
    m.face_e[3 * i + 2] = f.e.z;
    m.face_area[i] = f.area;
    m.face_delta[i] = f.delta;
    m.face_weight[i] = f.weight;
  }
}

void block2d_build(std::vector<V3> points, const std::vector<std::array<int64_t, 3>>& edef,
                   const std::vector<std::array<int64_t, 4>>& bdef,
                   const std::vector<std::string>& pnames,
                   const std::vector<std::vector<int64_t>>& pedges, xcalm_mesh& m) {
  // line! for each edge in list order (2_builder.jl:287-299, 302-309)
  std::vector<BEdge> edges;
  for (auto& ed : edef) {
    int64_t p1 = ed[0], p2 = ed[1], ncells = ed[2];
    if (p1 < 1 || p2 < 1 || p1 > (int64_t)points.size() || p2 > (int64_t)points.size() ||
        ncells < 1)
      throw std::runtime_error("bad edge definition");
    BEdge e;
    e.nodesID.assign(ncells + 1, 0);
    e.nodesID[0] = p1;
    e.nodesID[ncells] = p2;
    V3 a = points[p1 - 1], b = points[p2 - 1];
    V3 d = b - a;
    double dmag = norm(d);
    V3 normal = d / dmag;
    for (int64_t j = 2; j <= ncells; ++j) {
      double dx = (dmag / (double)ncells) * (double)(j - 1);
      points.push_back(a + dx * normal);
      e.nodesID[j - 1] = (int64_t)points.size();
    }
    e.ncells = ncells;
    e.boundary = false;
    edges.push_back(e);
  }
  // quad (2_builder.jl:337-351)
  std::vector<BBlock> blocks;
  for (auto& bd : bdef) {
    BBlock B;
    for (int k = 0; k < 4; ++k) {
      if (bd[k] < 1 || bd[k] > (int64_t)edges.size()) throw std::runtime_error("bad block edge");
      B.edgesID[k] = bd[k];
    }
    B.nx = edges[bd[0] - 1].ncells;
    B.ny = edges[bd[2] - 1].ncells;
    if (edges[bd[1] - 1].ncells != B.nx || edges[bd[3] - 1].ncells != B.ny)
      throw std::runtime_error("block edges have inconsistent cell counts");
    B.nodesID.assign((B.nx + 1) * (B.ny + 1), 0);
    for (int64_t i = 1; i <= B.nx + 1; ++i) {
      B.node(i, 1) = edges[bd[0] - 1].nodesID[i - 1];
      B.node(i, B.ny + 1) = edges[bd[1] - 1].nodesID[i - 1];
    }
    for (int64_t j = 1; j <= B.ny + 1; ++j) {
      B.node(1, j) = edges[bd[2] - 1].nodesID[j - 1];
      B.node(B.nx + 1, j) = edges[bd[3] - 1].nodesID[j - 1];
    }
    B.elementsID.assign(B.nx * B.ny, 0);
    B.facesNS.assign(B.nx * (B.ny + 1), 0);
    B.facesEW.assign((B.nx + 1) * B.ny, 0);
    blocks.push_back(B);
  }
  // preallocate_mesh (2_builder.jl:240-256)
  int64_t inner = 0, ncell_tot = 0;
  for (auto& B : blocks) {
    inner += (B.nx - 1) * (B.ny - 1);
    ncell_tot += B.nx * B.ny;
  }
  std::vector<V3> nodes(points.size() + inner, V3{0, 0, 0});
  for (size_t i = 0; i < points.size(); ++i) nodes[i] = points[i];
  for (auto& pe : pedges)
    for (auto id : pe) {
      if (id < 1 || id > (int64_t)edges.size()) throw std::runtime_error("bad patch edge");
      edges[id - 1].boundary = true;
    }
  int64_t nfaces = 0;
  for (auto& B : blocks) nfaces += B.nx * (B.ny + 1) + (B.nx + 1) * B.ny;
  for (auto& e : edges)
    if (!e.boundary) nfaces -= e.ncells;
  std::vector<OldFace> faces(nfaces);
  std::vector<OldCell> cells(ncell_tot);
  std::vector<OldBoundary> bnds;
  for (size_t p = 0; p < pnames.size(); ++p) {
    OldBoundary b;
    b.name = pnames[p];
    int64_t nc = 0;
    for (auto id : pedges[p]) nc += edges[id - 1].ncells;
    b.facesID.assign(nc, 0);
    b.cellsID.assign(nc, 0);
    bnds.push_back(b);
  }
  // generate_inner_points! (2_builder.jl:219-238)
  int64_t pc = (int64_t)points.size();
  for (auto& B : blocks) {
    BEdge& ey1 = edges[B.edgesID[2] - 1];
    BEdge& ey2 = edges[B.edgesID[3] - 1];
    for (int64_t yi = 2; yi <= B.ny; ++yi) {
      V3 a = points[ey1.nodesID[yi - 1] - 1], b = points[ey2.nodesID[yi - 1] - 1];
      V3 d = b - a;
      double dmag = norm(d);
      V3 normal = d / dmag;
      for (int64_t xi = 2; xi <= B.nx; ++xi) {
        ++pc;
        double dx = (dmag / (double)B.nx) * (double)(xi - 1);
        nodes[pc - 1] = a + dx * normal;
        B.node(xi, yi) = pc;
      }
    }
  }
  // generate_elements! (2_builder.jl:193-217)
  int64_t celli = 0;
  for (auto& B : blocks)
    for (int64_t yi = 1; yi <= B.ny; ++yi)
      for (int64_t xi = 1; xi <= B.nx; ++xi) {
        ++celli;
        int64_t nl[4] = {B.node(xi, yi), B.node(xi + 1, yi), B.node(xi + 1, yi + 1),
                         B.node(xi, yi + 1)};
        B.elem(xi, yi) = celli;
        cells[celli - 1].nodesID.assign(nl, nl + 4);
        cells[celli - 1].centre = geometric_centre(nodes, nl, 4);
      }
  auto index_to_block_edge = [](BBlock& B, int64_t edgei, int64_t pointi, int64_t facei) {
    if (edgei == 1) B.ns(pointi, 1) = facei;
    else if (edgei == 2) B.ns(pointi, B.ny + 1) = facei;
    else if (edgei == 3) B.ew(1, pointi) = facei;
    else if (edgei == 4) B.ew(B.nx + 1, pointi) = facei;
  };
  auto set_face = [&](int64_t facei, int64_t p1, int64_t p2) {
    if (facei < 1 || facei > nfaces) throw std::runtime_error("face count mismatch");
    OldFace& f = faces[facei - 1];
    f.n1 = p1;
    f.n2 = p2;
    int64_t ids[2] = {p1, p2};
    f.centre = geometric_centre(nodes, ids, 2);
  };
  // generate_boundary_faces! (2_builder.jl:144-171)
  int64_t facei = 1;
  for (size_t p = 0; p < pnames.size(); ++p) {
    int64_t i = 1;
    for (auto edgeID : pedges[p]) {
      BBlock* blk = nullptr;
      int64_t edgei = 0;
      for (auto& B : blocks) {
        for (int k = 0; k < 4; ++k)
          if (B.edgesID[k] == edgeID) {
            blk = &B;
            edgei = k + 1;
            break;
          }
        if (blk) break;
      }
      if (!blk) throw std::runtime_error("patch edge not in any block");
      auto& nid = edges[edgeID - 1].nodesID;
      for (size_t pi = 1; pi < nid.size(); ++pi) {
        set_face(facei, nid[pi - 1], nid[pi]);
        index_to_block_edge(*blk, edgei, (int64_t)pi, facei);
        bnds[p].facesID[i - 1] = facei;
        ++facei;
        ++i;
      }
    }
  }
  // generate_interface_faces! (2_builder.jl:80-107)
  std::vector<BBlock*> pairB(2, nullptr);
  std::vector<int64_t> pairE(2, 0);
  auto find_pair = [&](int64_t edgeID) {
    int c = 0;
    for (auto& B : blocks)
      for (int k = 0; k < 4; ++k)
        if (B.edgesID[k] == edgeID && c < 2) {
          pairB[c] = &B;
          pairE[c] = k + 1;
          ++c;
        }
    return c;
  };
  for (size_t ei = 0; ei < edges.size(); ++ei) {
    if (edges[ei].boundary) continue;
    if (find_pair((int64_t)ei + 1) != 2) throw std::runtime_error("interface edge not shared by two blocks");
    auto& nid = edges[ei].nodesID;
    for (size_t pi = 1; pi < nid.size(); ++pi) {
      set_face(facei, nid[pi - 1], nid[pi]);
      for (int k = 0; k < 2; ++k) index_to_block_edge(*pairB[k], pairE[k], (int64_t)pi, facei);
      ++facei;
    }
  }
  // generate_internal_faces! (2_builder.jl:13-43)
  for (auto& B : blocks) {
    for (int64_t yi = 1; yi <= B.ny - 1; ++yi)
      for (int64_t xi = 1; xi <= B.nx; ++xi) {
        set_face(facei, B.node(xi, yi + 1), B.node(xi + 1, yi + 1));
        B.ns(xi, yi + 1) = facei;
        ++facei;
      }
    for (int64_t xi = 1; xi <= B.nx - 1; ++xi)
      for (int64_t yi = 1; yi <= B.ny; ++yi) {
        set_face(facei, B.node(xi + 1, yi), B.node(xi + 1, yi + 1));
        B.ew(xi + 1, yi) = facei;
        ++facei;
      }
  }
  // connect! (3_connectivity.jl)
  auto setown = [&](int64_t fID, int64_t a, int64_t b) {
    faces[fID - 1].o1 = a;
    faces[fID - 1].o2 = b;
  };
  for (auto& B : blocks) {
    int64_t accNS[2][2] = {{1, 1}, {B.ny + 1, B.ny}};
    for (auto& acc : accNS)
      for (int64_t row = 1; row <= B.nx; ++row) {
        int64_t e = B.elem(row, acc[1]);
        setown(B.ns(row, acc[0]), e, e);
      }
    int64_t accEW[2][2] = {{1, 1}, {B.nx + 1, B.nx}};
    for (auto& acc : accEW)
      for (int64_t col = 1; col <= B.ny; ++col) {
        int64_t e = B.elem(acc[1], col);
        setown(B.ew(acc[0], col), e, e);
      }
  }
  for (auto& b : bnds)
    for (size_t fi2 = 0; fi2 < b.facesID.size(); ++fi2) b.cellsID[fi2] = faces[b.facesID[fi2] - 1].o1;
  for (size_t ei = 0; ei < edges.size(); ++ei) {
    if (edges[ei].boundary) continue;
    find_pair((int64_t)ei + 1);
    BBlock& b1 = *pairB[0];
    BBlock& b2 = *pairB[1];
    if (pairE[0] == 3 && pairE[1] == 4)
      for (int64_t j = 1; j <= b1.ny; ++j) setown(b1.ew(1, j), b1.elem(1, j), b2.elem(b2.nx, j));
    if (pairE[0] == 4 && pairE[1] == 3)
      for (int64_t j = 1; j <= b1.ny; ++j) setown(b1.ew(b1.nx + 1, j), b1.elem(b1.nx, j), b2.elem(1, j));
    if (pairE[0] == 1 && pairE[1] == 2)
      for (int64_t i = 1; i <= b1.nx; ++i) setown(b1.ns(i, 1), b1.elem(i, 1), b2.elem(i, b2.ny));
    if (pairE[0] == 2 && pairE[1] == 1)
      for (int64_t i = 1; i <= b1.nx; ++i) setown(b1.ns(i, b1.ny + 1), b1.elem(i, b1.ny), b2.elem(i, 1));
  }
  for (auto& B : blocks) {
    for (int64_t j = 2; j <= B.ny; ++j)
      for (int64_t i = 1; i <= B.nx; ++i) setown(B.ns(i, j), B.elem(i, j - 1), B.elem(i, j));
    for (int64_t j = 1; j <= B.ny; ++j)
      for (int64_t i = 2; i <= B.nx; ++i) setown(B.ew(i, j), B.elem(i - 1, j), B.elem(i, j));
  }
  int64_t nbf = 0;
  for (auto& b : bnds) nbf += (int64_t)b.facesID.size();
  for (auto& B : blocks)
    for (int64_t j = 1; j <= B.ny; ++j)
      for (int64_t i = 1; i <= B.nx; ++i) {
        int64_t cid = B.elem(i, j);
        int64_t senw[4] = {B.ns(i, j), B.ew(i + 1, j), B.ns(i, j + 1), B.ew(i, j)};
        for (auto id : senw)
          if (id > nbf) cells[cid - 1].facesID.push_back(id);
      }
  for (int64_t c = 0; c < ncell_tot; ++c)
    for (auto fID : cells[c].facesID) {
      const OldFace& f = faces[fID - 1];
      cells[c].neighbours.push_back(f.o1 != c + 1 ? f.o1 : f.o2);
    }
  // geometry! (4_geometry.jl:21-156)
  const V3 k{0.0, 0.0, 1.0};
  for (int64_t fidx = nbf + 1; fidx <= nfaces; ++fidx) {
    OldFace& f = faces[fidx - 1];
    V3 tangent = nodes[f.n2 - 1] - nodes[f.n1 - 1];
    double area = norm(tangent);
    V3 c1 = cells[f.o1 - 1].centre, c2 = cells[f.o2 - 1].centre, cf = f.centre;
    V3 d_f2 = c2 - cf, d_12 = c2 - c1;
    V3 normal = cross(tangent / area, k);
    if (dot(d_12, normal) < 0.0) normal = -1.0 * normal;
    double delta = norm(d_12);
    f.area = area;
    f.normal = normal;
    f.delta = delta;
    f.e = d_12 / delta;
    f.weight = norm(d_f2) / std::fabs(dot(d_12, normal));
  }
  for (auto& b : bnds)
    for (auto fID : b.facesID) {
      OldFace& f = faces[fID - 1];
      V3 tangent = nodes[f.n2 - 1] - nodes[f.n1 - 1];
      double area = norm(tangent);
      V3 normal = cross(tangent / area, k);
      V3 d_cf = f.centre - cells[f.o1 - 1].centre;
      if (dot(d_cf, normal) < 0.0) normal = -1.0 * normal;
      double delta = norm(d_cf);
      f.area = area;
      f.normal = normal;
      f.delta = delta;
      f.e = d_cf / delta;
      f.weight = 1.0;
    }
  for (auto& c : cells) {
    double volume = 0.0;
    for (auto fID : c.facesID) {
      const OldFace& f = faces[fID - 1];
      V3 d_cf = f.centre - c.centre;
      int64_t s = dot(d_cf, f.normal) > 0.0 ? 1 : -1;
      c.nsign.push_back(s);
      volume += (dot(d_cf, f.normal) * (double)s) * f.area;
      c.volume = 0.5 * volume;
    }
  }
  for (auto& b : bnds)
    for (size_t i = 0; i < b.cellsID.size(); ++i) {
      OldCell& c = cells[b.cellsID[i] - 1];
      const OldFace& f = faces[b.facesID[i] - 1];
      V3 d_cf = f.centre - c.centre;
      c.volume = c.volume + (0.5 * dot(d_cf, f.normal)) * f.area;
    }
  update_mesh_format(nodes, cells, faces, bnds, m);
}

// ---------------------------------------------------------------------------
// UNV2 reader restatement (src/UNV2/UNV2_1_loader.jl, UNV2_3_builder.jl,
// UNV2_2_geometry.jl)
// ---------------------------------------------------------------------------
std::vector<std::string> split_ws(const std::string& s) {
  std::vector<std::string> out;
  std::istringstream is(s);
  std::string t;
  while (is >> t) out.push_back(t);
  return out;
}
bool parse_int(const std::string& s, int64_t& v) {
  if (s.empty()) return false;
  char* end = nullptr;
  long long r = std::strtoll(s.c_str(), &end, 10);
  if (*end != '\0') return false;
  v = r;
  return true;
}

void unv2_build(const std::string& path, double scale, xcalm_mesh& m) {
  std::ifstream in(path);
  if (!in) throw std::runtime_error("cannot open " + path);
  struct Elem {
    int64_t index, nv;
    std::vector<int64_t> v;
  };
  struct BL {
    std::string name;
    std::vector<int64_t> elements;
  };
  std::vector<V3> pts;
  std::vector<Elem> elems;
  std::vector<BL> bls;
  bool block = false, d2411 = false, d2412 = false, d2467 = false;
  int64_t index = 0, vcount = 0;
  std::string line;
  while (std::getline(in, line)) {
    auto sl = split_ws(line);
    if (sl.empty()) continue;
    if (!block && sl[0] == "-1") {
      block = true;
      continue;
    }
    if (block && sl[0] == "-1") {
      block = d2411 = d2412 = d2467 = false;
      continue;
    }
    if (block && !d2411 && !d2412 && !d2467) {
      if (sl[0] == "2411") { d2411 = true; continue; }
      if (sl[0] == "2412") { d2412 = true; continue; }
      if (sl[0] == "2467") { d2467 = true; continue; }
    }
    if (d2411) {
      if (sl[0].find('.') != std::string::npos && sl.size() >= 3) {
        pts.push_back(V3{std::stod(sl[0]), std::stod(sl[1]), std::stod(sl[2])});
        continue;
      }
    }
    if (d2412) {
      int64_t first;
      if (!parse_int(sl[0], first)) continue;
      if (first == 0) continue;
      int64_t last;
      if (sl.size() == 6 && parse_int(sl.back(), last) && (last == 2 || last == 3 || last == 4)) {
        vcount = last;
        index = first;
        continue;
      }
      if (sl.size() == 2 || sl.size() == 3 || sl.size() == 4) {
        Elem e;
        e.index = index;
        e.nv = vcount;
        for (auto& t : sl) {
          int64_t v;
          if (!parse_int(t, v)) throw std::runtime_error("bad UNV element line");
          e.v.push_back(v);
        }
        elems.push_back(e);
        continue;
      }
    }
    if (d2467) {
      int64_t a, b;
      if (parse_int(sl[0], a) && sl.size() > 1 && parse_int(sl[1], b) && b == 0) {
        bls.push_back(BL{});
        continue;
      }
      if (!parse_int(sl[0], a)) {
        if (bls.empty()) throw std::runtime_error("UNV group name before header");
        bls.back().name = sl[0];
        continue;
      }
      if (bls.empty()) continue;
      int64_t v;
      if (sl.size() > 1 && parse_int(sl[1], v)) bls.back().elements.push_back(v);
      if (sl.size() >= 6 && parse_int(sl[5], v)) bls.back().elements.push_back(v);
    }
  }
  if (scale != 1.0)
    for (auto& p : pts) p = p * scale;
  int64_t bfaces = 0;
  for (auto& b : bls) bfaces += (int64_t)b.elements.size();
  size_t first = 0;
  while (first < elems.size() && elems[first].nv <= 2) ++first;
  if (first == elems.size()) throw std::runtime_error("no 2D elements in UNV file");
  // generate_nodes
  std::vector<std::vector<int64_t>> ncells(pts.size());
  int64_t cellID = 0;
  for (size_t i = first; i < elems.size(); ++i) {
    ++cellID;
    for (auto nid : elems[i].v) ncells.at(nid - 1).push_back(cellID);
  }
  // generate_faces + unique!
  std::vector<std::pair<int64_t, int64_t>> flist;
  std::map<std::pair<int64_t, int64_t>, int64_t> seen;
  auto push_face = [&](int64_t a, int64_t b) {
    if (a == b) throw std::runtime_error("Boundary elements are inconsistent: possible mesh corruption");
    auto key = a < b ? std::make_pair(a, b) : std::make_pair(b, a);
    if (seen.count(key)) return;
    seen[key] = (int64_t)flist.size() + 1;
    flist.push_back(key);
  };
  for (auto& b : bls)
    for (auto ei : b.elements) {
      const Elem& e = elems.at(ei - 1);
      push_face(e.v.at(0), e.v.at(1));
    }
  for (size_t i = first; i < elems.size(); ++i) {
    auto& v = elems[i].v;
    for (size_t vi = 0; vi < v.size(); ++vi) push_face(v[vi], v[(vi + 1) % v.size()]);
  }
  const int64_t nf = (int64_t)flist.size();
  std::vector<OldFace> faces(nf);
  std::vector<OldCell> cells(elems.size() - first);
  for (size_t i = first; i < elems.size(); ++i) cells[i - first].nodesID = elems[i].v;
  // face_cell_connectivity!
  for (int64_t f = 0; f < nf; ++f) {
    faces[f].n1 = flist[f].first;
    faces[f].n2 = flist[f].second;
    int64_t own[2] = {0, 0};
    int cnt = 0;
    for (auto c1 : ncells[faces[f].n1 - 1])
      for (auto c2 : ncells[faces[f].n2 - 1])
        if (c1 == c2) {
          if (cnt < 2) own[cnt] = c1;
          ++cnt;
        }
    if (own[1] != 0) {
      faces[f].o1 = own[0];
      faces[f].o2 = own[1];
      cells[own[0] - 1].facesID.push_back(f + 1);
      cells[own[1] - 1].facesID.push_back(f + 1);
      cells[own[0] - 1].neighbours.push_back(own[1]);
      cells[own[1] - 1].neighbours.push_back(own[0]);
    } else {
      faces[f].o1 = faces[f].o2 = own[0];
    }
  }
  // generate_boundaries + boundary_connectivity!
  std::vector<OldBoundary> bnds;
  for (auto& b : bls) {
    OldBoundary ob;
    ob.name = b.name;
    for (auto ei : b.elements) {
      auto v = elems.at(ei - 1).v;
      std::sort(v.begin(), v.end());
      for (int64_t f = 0; f < bfaces && f < nf; ++f)
        if (faces[f].n1 == v[0] && faces[f].n2 == v[1]) {
          ob.facesID.push_back(f + 1);
          ob.cellsID.push_back(faces[f].o1);
        }
    }
    bnds.push_back(ob);
  }
  // process_geometry!: node-average centres, then geometry!
  for (auto& c : cells) c.centre = geometric_centre(pts, c.nodesID.data(), (int64_t)c.nodesID.size());
  for (auto& f : faces) {
    int64_t ids[2] = {f.n1, f.n2};
    f.centre = geometric_centre(pts, ids, 2);
  }
  int64_t nbf = 0;
  for (auto& b : bnds) nbf += (int64_t)b.facesID.size();
  const V3 k{0.0, 0.0, 1.0};
  for (int64_t fidx = nbf + 1; fidx <= nf; ++fidx) {
    OldFace& f = faces[fidx - 1];
    V3 tangent = pts[f.n2 - 1] - pts[f.n1 - 1];
    double area = norm(tangent);
    V3 F1 = f.centre, C1 = cells[f.o1 - 1].centre, C2 = cells[f.o2 - 1].centre;
    V3 C1F1 = F1 - C1, C2F1 = F1 - C2, C1C2 = C2 - C1;
    V3 normal = cross(tangent / area, k);
    if (dot(C1C2, normal) < 0.0) normal = -normal;
    double w, d;
    V3 e;
    weight_delta_e_internal(C1F1, C2F1, C1C2, normal, w, d, e);
    f.area = area;
    f.normal = normal;
    f.delta = d;
    f.e = e;
    f.weight = w;
  }
  for (auto& b : bnds)
    for (auto fID : b.facesID) {
      OldFace& f = faces[fID - 1];
      V3 tangent = pts[f.n2 - 1] - pts[f.n1 - 1];
      double area = norm(tangent);
      V3 normal = cross(tangent / area, k);
      V3 C1F1 = f.centre - cells[f.o1 - 1].centre;
      if (dot(C1F1, normal) < 0.0) normal = -normal;
      double w, d;
      V3 e;
      weight_delta_e_boundary(C1F1, normal, w, d, e);
      f.area = area;
      f.normal = normal;
      f.delta = d;
      f.e = e;
      f.weight = w;
    }
  // cell_properties! (area-weighted centroid over all faces, then volume)
  std::vector<std::vector<int64_t>> allf(cells.size());
  for (size_t c = 0; c < cells.size(); ++c) allf[c] = cells[c].facesID;
  for (auto& b : bnds)
    for (size_t i = 0; i < b.cellsID.size(); ++i) allf[b.cellsID[i] - 1].push_back(b.facesID[i]);
  for (size_t c = 0; c < cells.size(); ++c) {
    OldCell& cell = cells[c];
    double surf = 0.0;
    V3 sumc{0, 0, 0};
    for (auto fID : allf[c]) {
      const OldFace& f = faces[fID - 1];
      sumc = sumc + f.centre * f.area;
      surf += f.area;
    }
    V3 tc = sumc / surf;
    double volume = 0.0;
    cell.nsign.clear();
    for (auto fID : allf[c]) {
      const OldFace& f = faces[fID - 1];
      V3 d_cf = f.centre - tc;
      int64_t s = dot(d_cf, f.normal) > 0.0 ? 1 : -1;
      if (std::find(cell.facesID.begin(), cell.facesID.end(), fID) != cell.facesID.end())
        cell.nsign.push_back(s);
      volume += (dot(d_cf, f.normal) * (double)s) * f.area;
    }
    cell.centre = tc;
    cell.volume = 0.5 * volume;
  }
  update_mesh_format(pts, cells, faces, bnds, m);
}

// ---------------------------------------------------------------------------
// ASCII polyMesh reader (FoamMesh_1_read.jl)
// ---------------------------------------------------------------------------
std::string slurp(const std::string& p) {
  std::ifstream in(p, std::ios::binary);
  if (!in) throw std::runtime_error("cannot open " + p);
  std::ostringstream ss;
  ss << in.rdbuf();
  return ss.str();
}
// position just after the first line that is a bare integer; returns the count
size_t foam_count_line(const std::string& s, int64_t& count) {
  size_t pos = 0;
  while (pos < s.size()) {
    size_t e = s.find('\n', pos);
    if (e == std::string::npos) e = s.size();
    std::string ln = s.substr(pos, e - pos);
    while (!ln.empty() && (ln.back() == '\r' || ln.back() == ' ')) ln.pop_back();
    int64_t v;
    if (parse_int(ln, v)) {
      count = v;
      return e + 1;
    }
    pos = e + 1;
  }
  throw std::runtime_error("no count line in polyMesh file");
}
int64_t next_uint(const std::string& s, size_t& pos) {
  while (pos < s.size() && (s[pos] < '0' || s[pos] > '9')) ++pos;
  int64_t v = 0;
  while (pos < s.size() && s[pos] >= '0' && s[pos] <= '9') v = 10 * v + (s[pos++] - '0');
  return v;
}

void read_foam(const std::string& dir, double scale, FoamData& fd) {
  {
    std::string s = slurp(dir + "/points");
    int64_t n;
    size_t pos = foam_count_line(s, n);
    fd.points.resize(n);
    for (int64_t i = 0; i < n; ++i) {
      pos = s.find('(', pos);
      if (pos == std::string::npos) throw std::runtime_error("truncated points file");
      ++pos;
      double c[3];
      for (int k = 0; k < 3; ++k) {
        while (pos < s.size() && (s[pos] == '(' || s[pos] == ')' || s[pos] == ' ' ||
                                  s[pos] == '\t' || s[pos] == '\n' || s[pos] == '\r'))
          ++pos;
        size_t lo = pos;
        while (pos < s.size() && s[pos] != '(' && s[pos] != ')' && s[pos] != ' ' &&
               s[pos] != '\t' && s[pos] != '\n' && s[pos] != '\r')
          ++pos;
        c[k] = scale * std::stod(s.substr(lo, pos - lo));
      }
      fd.points[i] = V3{c[0], c[1], c[2]};
    }
  }
  {
    std::string s = slurp(dir + "/faces");
    int64_t n;
    size_t pos = foam_count_line(s, n);
    fd.fn_first.resize(n);
    fd.fn_last.resize(n);
    int64_t start = 1;
    for (int64_t f = 0; f < n; ++f) {
      int64_t nn = next_uint(s, pos);
      for (int64_t i = 0; i < nn; ++i) fd.face_nodes.push_back(next_uint(s, pos) + 1);
      fd.fn_first[f] = start;
      fd.fn_last[f] = start + nn - 1;
      start += nn;
    }
  }
  std::vector<int64_t> neigh;
  auto read_labels = [&](const std::string& p, std::vector<int64_t>& out) {
    std::string s = slurp(p);
    int64_t n;
    size_t pos = foam_count_line(s, n);
    out.resize(n);
    for (int64_t i = 0; i < n; ++i) out[i] = next_uint(s, pos) + 1;
  };
  read_labels(dir + "/neighbour", neigh);
  read_labels(dir + "/owner", fd.face_owner);
  {
    std::string s = slurp(dir + "/boundary");
    int64_t nb;
    size_t pos = foam_count_line(s, nb);
    std::istringstream is(s.substr(pos));
    std::string ln;
    int cur = -1;
    while (std::getline(is, ln)) {
      if (ln == ")") break;
      std::string t = ln;
      for (char& ch : t)
        if (ch == ';' || ch == '{' || ch == '}') ch = ' ';
      auto sl = split_ws(t);
      if (sl.size() == 1 && sl[0] != "(") {
        fd.bnames.push_back(sl[0]);
        fd.bnfaces.push_back(0);
        fd.bstart.push_back(0);
        cur = (int)fd.bnames.size() - 1;
        continue;
      }
      if (cur >= 0 && sl.size() == 2 && sl[0] == "nFaces") fd.bnfaces[cur] = std::stoll(sl[1]);
      if (cur >= 0 && sl.size() == 2 && sl[0] == "startFace") fd.bstart[cur] = std::stoll(sl[1]) + 1;
    }
    if ((int64_t)fd.bnames.size() != nb) throw std::runtime_error("boundary file patch count mismatch");
  }
  foam_assign_faces(fd, neigh);
}

// Generated hex box emitted as polyMesh data.
void box3d_build(int64_t nx, int64_t ny, int64_t nz, double Lx, double Ly, double Lz,
                 double jitter, uint64_t seed_jitter, uint64_t seed_perm, FoamData& fd) {
  if (nx < 1 || ny < 1 || nz < 1) throw std::runtime_error("box dimensions must be >= 1");
  const int64_t px = nx + 1, py = ny + 1, pz = nz + 1;
  auto nid = [&](int64_t i, int64_t j, int64_t k) { return i + px * (j + py * k); };
  fd.points.resize(px * py * pz);
  const double hx = Lx / (double)nx, hy = Ly / (double)ny, hz = Lz / (double)nz;
  Xoshiro rj(seed_jitter);
  for (int64_t k = 0; k < pz; ++k)
    for (int64_t j = 0; j < py; ++j)
      for (int64_t i = 0; i < px; ++i) {
        double ux = 2.0 * rj.u01() - 1.0, uy = 2.0 * rj.u01() - 1.0, uz = 2.0 * rj.u01() - 1.0;
        double x = (Lx * (double)i) / (double)nx;
        double y = (Ly * (double)j) / (double)ny;
        double z = (Lz * (double)k) / (double)nz;
        if (jitter != 0.0) {
          if (i > 0 && i < nx) x += ux * jitter * hx;
          if (j > 0 && j < ny) y += uy * jitter * hy;
          if (k > 0 && k < nz) z += uz * jitter * hz;
        }
        fd.points[nid(i, j, k)] = V3{x, y, z};
      }
  const int64_t nc = nx * ny * nz;
  std::vector<int64_t> perm(nc);
  for (int64_t c = 0; c < nc; ++c) perm[c] = c;
  if (seed_perm != 0) {
    Xoshiro rp(seed_perm);
    for (int64_t i = nc - 1; i >= 1; --i) {
      int64_t j = (int64_t)(rp.next() % (uint64_t)(i + 1));
      std::swap(perm[i], perm[j]);
    }
  }
  auto cid = [&](int64_t i, int64_t j, int64_t k) { return perm[i + nx * (j + ny * k)]; };
  struct F {
    int64_t o, n;
    int64_t v[4];
  };
  std::vector<F> internal;
  internal.reserve((nx - 1) * ny * nz + nx * (ny - 1) * nz + nx * ny * (nz - 1));
  // quad whose right-hand normal points +axis; reversed when owner is the + side
  auto add_internal = [&](int64_t cm, int64_t cp, int64_t a, int64_t b, int64_t c, int64_t d) {
    F f;
    if (cm < cp) {
      f.o = cm; f.n = cp;
      f.v[0] = a; f.v[1] = b; f.v[2] = c; f.v[3] = d;
    } else {
      f.o = cp; f.n = cm;
      f.v[0] = a; f.v[1] = d; f.v[2] = c; f.v[3] = b;
    }
    internal.push_back(f);
  };
  for (int64_t k = 0; k < nz; ++k)
    for (int64_t j = 0; j < ny; ++j)
      for (int64_t i = 1; i < nx; ++i)
        add_internal(cid(i - 1, j, k), cid(i, j, k), nid(i, j, k), nid(i, j + 1, k),
                     nid(i, j + 1, k + 1), nid(i, j, k + 1));
  for (int64_t k = 0; k < nz; ++k)
    for (int64_t j = 1; j < ny; ++j)
      for (int64_t i = 0; i < nx; ++i)
        add_internal(cid(i, j - 1, k), cid(i, j, k), nid(i, j, k), nid(i, j, k + 1),
                     nid(i + 1, j, k + 1), nid(i + 1, j, k));
  for (int64_t k = 1; k < nz; ++k)
    for (int64_t j = 0; j < ny; ++j)
      for (int64_t i = 0; i < nx; ++i)
        add_internal(cid(i, j, k - 1), cid(i, j, k), nid(i, j, k), nid(i + 1, j, k),
                     nid(i + 1, j + 1, k), nid(i, j + 1, k));
  std::sort(internal.begin(), internal.end(), [](const F& a, const F& b) {
    return a.o != b.o ? a.o < b.o : a.n < b.n;
  });
  std::vector<F> bnd;
  // inlet x=0 (outward -x), outlet x=Lx (+x)
  for (int64_t k = 0; k < nz; ++k)
    for (int64_t j = 0; j < ny; ++j)
      bnd.push_back(F{cid(0, j, k), -1, {nid(0, j, k), nid(0, j, k + 1), nid(0, j + 1, k + 1), nid(0, j + 1, k)}});
  for (int64_t k = 0; k < nz; ++k)
    for (int64_t j = 0; j < ny; ++j)
      bnd.push_back(F{cid(nx - 1, j, k), -1, {nid(nx, j, k), nid(nx, j + 1, k), nid(nx, j + 1, k + 1), nid(nx, j, k + 1)}});
  // walls: y=0, y=Ly, z=0, z=Lz
  for (int64_t k = 0; k < nz; ++k)
    for (int64_t i = 0; i < nx; ++i)
      bnd.push_back(F{cid(i, 0, k), -1, {nid(i, 0, k), nid(i + 1, 0, k), nid(i + 1, 0, k + 1), nid(i, 0, k + 1)}});
  for (int64_t k = 0; k < nz; ++k)
    for (int64_t i = 0; i < nx; ++i)
      bnd.push_back(F{cid(i, ny - 1, k), -1, {nid(i, ny, k), nid(i, ny, k + 1), nid(i + 1, ny, k + 1), nid(i + 1, ny, k)}});
  for (int64_t j = 0; j < ny; ++j)
    for (int64_t i = 0; i < nx; ++i)
      bnd.push_back(F{cid(i, j, 0), -1, {nid(i, j, 0), nid(i, j + 1, 0), nid(i + 1, j + 1, 0), nid(i + 1, j, 0)}});
  for (int64_t j = 0; j < ny; ++j)
    for (int64_t i = 0; i < nx; ++i)
      bnd.push_back(F{cid(i, j, nz - 1), -1, {nid(i, j, nz), nid(i + 1, j, nz), nid(i + 1, j + 1, nz), nid(i, j + 1, nz)}});
  const int64_t nif = (int64_t)internal.size(), nbf = (int64_t)bnd.size();
  fd.face_nodes.reserve(4 * (nif + nbf));
  fd.fn_first.reserve(nif + nbf);
  fd.fn_last.reserve(nif + nbf);
  fd.face_owner.reserve(nif + nbf);
  std::vector<int64_t> neigh(nif);
  int64_t start = 1;
  auto emit = [&](const F& f) {
    for (int q = 0; q < 4; ++q) fd.face_nodes.push_back(f.v[q] + 1);
    fd.fn_first.push_back(start);
    fd.fn_last.push_back(start + 3);
    start += 4;
    fd.face_owner.push_back(f.o + 1);
  };
  for (int64_t f = 0; f < nif; ++f) {
    emit(internal[f]);
    neigh[f] = internal[f].n + 1;
  }
  for (auto& f : bnd) emit(f);
  fd.bnames = {"inlet", "outlet", "walls"};
  fd.bnfaces = {ny * nz, ny * nz, 2 * nx * nz + 2 * nx * ny};
  fd.bstart = {nif + 1, nif + 1 + ny * nz, nif + 1 + 2 * ny * nz};
  foam_assign_faces(fd, neigh);
}

template <class Fn>
int guard(Fn fn) {
  try {
    fn();
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  } catch (...) {
    g_err = "unknown error";
    return -1;
  }
}

}  // namespace

extern "C" {

const char* xcalm_last_error(void) { return g_err.c_str(); }

int xcalm_foam_from_arrays(int64_t n_points, const double* points, int64_t n_faces,
                           const int64_t* face_nodes_ptr, const int64_t* face_nodes,
                           const int64_t* owner, int64_t n_ifaces, const int64_t* neighbour,
                           int32_t n_patches, const char* const* patch_names,
                           const int64_t* patch_start, const int64_t* patch_nfaces,
                           xcalm_mesh** out) {
  return guard([&] {
    FoamData fd;
    fd.points.resize(n_points);
    for (int64_t i = 0; i < n_points; ++i)
      fd.points[i] = V3{points[3 * i], points[3 * i + 1], points[3 * i + 2]};
    for (int64_t f = 0; f < n_faces; ++f) {
      fd.fn_first.push_back(face_nodes_ptr[f] + 1);
      fd.fn_last.push_back(face_nodes_ptr[f + 1]);
    }
    for (int64_t k = 0; k < face_nodes_ptr[n_faces]; ++k) fd.face_nodes.push_back(face_nodes[k] + 1);
    for (int64_t f = 0; f < n_faces; ++f) fd.face_owner.push_back(owner[f] + 1);
    std::vector<int64_t> neigh(n_ifaces);
    for (int64_t f = 0; f < n_ifaces; ++f) neigh[f] = neighbour[f] + 1;
    for (int32_t p = 0; p < n_patches; ++p) {
      fd.bnames.push_back(patch_names[p]);
      fd.bnfaces.push_back(patch_nfaces[p]);
      fd.bstart.push_back(patch_start[p] + 1);
    }
    foam_assign_faces(fd, neigh);
    auto* m = new xcalm_mesh();
    try {
      foam_build(fd, *m);
    } catch (...) {
      delete m;
      throw;
    }
    *out = m;
  });
}

int xcalm_read_foam(const char* dir, double scale, xcalm_mesh** out) {
  return guard([&] {
    FoamData fd;
    read_foam(dir, scale, fd);
    auto* m = new xcalm_mesh();
    try {
      foam_build(fd, *m);
    } catch (...) {
      delete m;
      throw;
    }
    *out = m;
  });
}

int xcalm_box3d(int64_t nx, int64_t ny, int64_t nz, double Lx, double Ly, double Lz,
                double jitter, uint64_t seed_jitter, uint64_t seed_perm, xcalm_mesh** out) {
  return guard([&] {
    FoamData fd;
    box3d_build(nx, ny, nz, Lx, Ly, Lz, jitter, seed_jitter, seed_perm, fd);
    auto* m = new xcalm_mesh();
    try {
      foam_build(fd, *m);
    } catch (...) {
      delete m;
      throw;
    }
    *out = m;
  });
}

int xcalm_block2d(int64_t n_points, const double* points, int64_t n_edges, const int64_t* edges,
                  int64_t n_blocks, const int64_t* blocks, int32_t n_patches,
                  const char* const* patch_names, const int64_t* patch_edge_ptr,
                  const int64_t* patch_edges, xcalm_mesh** out) {
  return guard([&] {
    std::vector<V3> pts(n_points);
    for (int64_t i = 0; i < n_points; ++i)
      pts[i] = V3{points[3 * i], points[3 * i + 1], points[3 * i + 2]};
    std::vector<std::array<int64_t, 3>> ed(n_edges);
    for (int64_t i = 0; i < n_edges; ++i) ed[i] = {edges[3 * i], edges[3 * i + 1], edges[3 * i + 2]};
    std::vector<std::array<int64_t, 4>> bd(n_blocks);
    for (int64_t i = 0; i < n_blocks; ++i)
      bd[i] = {blocks[4 * i], blocks[4 * i + 1], blocks[4 * i + 2], blocks[4 * i + 3]};
    std::vector<std::string> pn;
    std::vector<std::vector<int64_t>> pe;
    for (int32_t p = 0; p < n_patches; ++p) {
      pn.push_back(patch_names[p]);
      pe.emplace_back(patch_edges + patch_edge_ptr[p], patch_edges + patch_edge_ptr[p + 1]);
    }
    auto* m = new xcalm_mesh();
    try {
      block2d_build(pts, ed, bd, pn, pe, *m);
    } catch (...) {
      delete m;
      throw;
    }
    *out = m;
  });
}

int xcalm_read_unv2(const char* path, double scale, xcalm_mesh** out) {
  return guard([&] {
    auto* m = new xcalm_mesh();
    try {
      unv2_build(path, scale, *m);
    } catch (...) {
      delete m;
      throw;
    }
    *out = m;
  });
}

int xcalm_array(const xcalm_mesh* m, const char* name, const void** ptr, int64_t* len,
                int32_t* is_float) {
  return guard([&] {
    std::string n(name);
#define XI(NAME)                      \
  if (n == #NAME) {                   \
    *ptr = m->NAME.data();            \
    *len = (int64_t)m->NAME.size();   \
    *is_float = 0;                    \
    return;                           \
  }
#define XF(NAME)                      \
  if (n == #NAME) {                   \
    *ptr = m->NAME.data();            \
    *len = (int64_t)m->NAME.size();   \
    *is_float = 1;                    \
    return;                           \
  }
    XF(cell_centre) XF(cell_volume) XI(cell_faces_range) XI(cell_nodes_range) XI(cell_nodes)
    XI(cell_faces) XI(cell_neighbours) XI(cell_nsign) XI(face_nodes_range) XI(face_nodes)
    XI(face_owner) XF(face_centre) XF(face_normal) XF(face_e) XF(face_area) XF(face_delta)
    XF(face_weight) XI(boundary_cellsID) XI(patch_range) XF(node_coords)
#undef XI
#undef XF
    throw std::runtime_error("unknown mesh array " + n);
  });
}

int xcalm_sizes(const xcalm_mesh* m, int32_t* dim, int64_t* n_cells, int64_t* n_faces,
                int64_t* n_bfaces, int64_t* n_nodes, int32_t* n_patches) {
  *dim = m->dim;
  *n_cells = m->n_cells;
  *n_faces = m->n_faces;
  *n_bfaces = m->n_bfaces;
  *n_nodes = m->n_nodes;
  *n_patches = (int32_t)m->patch_names.size();
  return 0;
}

const char* xcalm_patch_name(const xcalm_mesh* m, int32_t i) {
  if (i < 0 || i >= (int32_t)m->patch_names.size()) return nullptr;
  return m->patch_names[i].c_str();
}

void xcalm_free(xcalm_mesh* m) { delete m; }

}  // extern "C"
