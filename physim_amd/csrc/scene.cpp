// scene.cpp -- scene ingestion in the reference's own formats: the XML
// elements Serializer_XML reads for a volume FEM model, its boundary
// conditions, the optimization problem and the solver options, and TetGen
// volume meshes (.node / .ele), the format Serializer_Mesh::ReadVolume_TET
// declares (Serializer_Mesh.h:46-52) but the reference never implements.
//
// A scene document is a <scene> root holding, under the reference's element
// and attribute names:
//   <simulable> with the children LoadSimulable_VolumeFEM requires
//     (Serializer_XML.cpp:1199-1325): meshInit, discretization,
//     shapeFunction, materialParam, materialModel, quadrature;
//   <boundary> with <bc> children (LoadBoundary, :1507-1647);
//   <problem> (LoadOptimProblem, :1327-1354);
//   <solver> (LoadOptimSolver, :1356-1505).
// Reading follows the reference's semantics where it works and corrects it
// where it does not (each case is stated where it is handled):
//   - meshInit FILEVOLUME reads the `path` attribute (the reference checks
//     `path` but reads `mesh`, :928-939; `mesh` is accepted too);
//   - meshInit GRID builds the 5-tet hex grid of CreateTetMesh_Grid
//     (MeshUtils.cpp:927-991), which the reference's TET4 reader leaves as a
//     TODO (:959);
//   - meshRest, position, scale and rotation as LoadMesh_TetVolume
//     (x' = Rz Ry Rx S x + t, Euler angles in degrees,
//     GeometryUtils.cpp:79-114).
// Supported: TET4 / TET_LIN / TET_1, CONH and STVK (3 in 3), parameter sets
// by YOUNG + POISSON + DENSITY or LAME1 + LAME2 + DENSITY
// (ParameterSet.cpp:54-70), CONSTANT and VARIABLE distributions, Gravity and
// hard FixPosition conditions with AllNodes / IndexNodes / BoxNodes /
// BoundaryNodes selections, Static problems, the Newton / Steepest / LBFGS
// steps and the EigenCG / EigenLDLT linear solvers with the Simple line
// search.  Anything else fails with PGN_ERR_INVALID and a message naming it.
#include <algorithm>
#include <array>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "pgn_internal.hpp"

namespace pgn {
int api_guard_set(const char* msg, int code);
}

using namespace pgn;

struct pgn_scene {
  int n = 0, m = 0;
  std::vector<double> X0;        // 3n rest positions
  std::vector<double> X;         // 3n initial positions (meshInit)
  std::vector<int> T;            // 4m
  std::vector<uint8_t> kind;     // m
  std::vector<double> l1, l2, rho;
  std::vector<uint8_t> fixed;    // 3n
  bool has_g = false;
  double g[3] = {0.0, 0.0, 0.0};
  int g_stages = 1;
  pgn_newton_opts opts{};
  int dynamic = 0;
  double dt = 0.0;
  int grid[3] = {0, 0, 0};  // cells per axis of a GRID mesh, 0 otherwise
};

namespace {

// ---- a minimal XML reader (elements, attributes, text, comments) ----------
struct XNode {
  std::string name, text;
  std::map<std::string, std::string> attr;
  std::vector<std::unique_ptr<XNode>> kids;
  const XNode* child(const char* n) const {
    for (auto& k : kids)
      if (k->name == n) return k.get();
    return nullptr;
  }
  std::vector<const XNode*> children(const char* n) const {
    std::vector<const XNode*> v;
    for (auto& k : kids)
      if (k->name == n) v.push_back(k.get());
    return v;
  }
  const std::string* at(const char* a) const {
    auto it = attr.find(a);
    return it == attr.end() ? nullptr : &it->second;
  }
};

std::string xml_unescape(const std::string& s) {
  std::string o;
  for (size_t i = 0; i < s.size(); ++i) {
    if (s[i] != '&') {
      o += s[i];
      continue;
    }
    const size_t j = s.find(';', i);
    PGN_CHECK(j != std::string::npos, "XML: unterminated entity");
    const std::string e = s.substr(i + 1, j - i - 1);
    if (e == "lt") o += '<';
    else if (e == "gt") o += '>';
    else if (e == "amp") o += '&';
    else if (e == "quot") o += '"';
    else if (e == "apos") o += '\'';
    else throw Error(PGN_ERR_INVALID, "XML: unknown entity &" + e + ";");
    i = j;
  }
  return o;
}

class XmlReader {
 public:
  explicit XmlReader(const std::string& s) : s_(s) {}
  std::unique_ptr<XNode> document() {
    skip_misc();
    auto root = element();
    skip_misc();
    PGN_CHECK(p_ >= s_.size(), "XML: content after the root element");
    return root;
  }

 private:
  const std::string& s_;
  size_t p_ = 0;
  [[noreturn]] void fail(const char* what) {
    throw Error(PGN_ERR_INVALID, std::string("XML: ") + what + " at offset " + std::to_string(p_));
  }
  bool starts(const char* t) const { return s_.compare(p_, std::strlen(t), t) == 0; }
  void ws() {
    while (p_ < s_.size() && std::isspace((unsigned char)s_[p_])) ++p_;
  }
  void skip_misc() {  // whitespace, declarations, comments, doctype
    for (;;) {
      ws();
      if (starts("<?")) {
        const size_t e = s_.find("?>", p_);
        if (e == std::string::npos) fail("unterminated declaration");
        p_ = e + 2;
      } else if (starts("<!--")) {
        const size_t e = s_.find("-->", p_);
        if (e == std::string::npos) fail("unterminated comment");
        p_ = e + 3;
      } else if (starts("<!")) {
        const size_t e = s_.find('>', p_);
        if (e == std::string::npos) fail("unterminated declaration");
        p_ = e + 1;
      } else {
        return;
      }
    }
  }
  std::string name() {
    const size_t b = p_;
    while (p_ < s_.size() && (std::isalnum((unsigned char)s_[p_]) || s_[p_] == '_' ||
                              s_[p_] == '-' || s_[p_] == '.' || s_[p_] == ':'))
      ++p_;
    if (p_ == b) fail("expected a name");
    return s_.substr(b, p_ - b);
  }
  std::unique_ptr<XNode> element() {
    if (!starts("<")) fail("expected an element");
    ++p_;
    auto n = std::make_unique<XNode>();
    n->name = name();
    for (;;) {
      ws();
      if (starts("/>")) {
        p_ += 2;
        return n;
      }
      if (starts(">")) {
        ++p_;
        break;
      }
      const std::string a = name();
      ws();
      if (!starts("=")) fail("expected = after an attribute name");
      ++p_;
      ws();
      if (p_ >= s_.size() || (s_[p_] != '"' && s_[p_] != '\'')) fail("expected a quoted value");
      const char q = s_[p_++];
      const size_t e = s_.find(q, p_);
      if (e == std::string::npos) fail("unterminated attribute value");
      n->attr[a] = xml_unescape(s_.substr(p_, e - p_));
      p_ = e + 1;
    }
    for (;;) {  // content
      if (p_ >= s_.size()) fail("unterminated element");
      if (starts("<!--")) {
        const size_t e = s_.find("-->", p_);
        if (e == std::string::npos) fail("unterminated comment");
        p_ = e + 3;
      } else if (starts("</")) {
        p_ += 2;
        if (name() != n->name) fail("mismatched closing tag");
        ws();
        if (!starts(">")) fail("expected >");
        ++p_;
        return n;
      } else if (starts("<")) {
        n->kids.push_back(element());
      } else {
        const size_t e = s_.find('<', p_);
        if (e == std::string::npos) fail("unterminated element");
        n->text += xml_unescape(s_.substr(p_, e - p_));
        p_ = e;
      }
    }
  }
};

std::string read_file(const std::string& path) {
  std::ifstream f(path, std::ios::binary);
  PGN_CHECK(f.good(), "cannot open " + path);
  std::stringstream ss;
  ss << f.rdbuf();
  return ss.str();
}

std::string dir_of(const std::string& path) {
  const size_t k = path.find_last_of('/');
  return k == std::string::npos ? std::string() : path.substr(0, k + 1);
}

// IOUtils::stringToVector_CSV: comma-separated numbers
std::vector<double> csv(const std::string& s) {
  std::vector<double> v;
  std::string tok;
  std::stringstream ss(s);
  while (std::getline(ss, tok, ',')) {
    const char* b = tok.c_str();
    while (*b && std::isspace((unsigned char)*b)) ++b;
    if (!*b) continue;
    char* e = nullptr;
    const double x = std::strtod(b, &e);
    PGN_CHECK(e != b, "not a number: '" + tok + "'");
    v.push_back(x);
  }
  return v;
}

std::array<double, 3> vec3(const XNode* n, const char* what) {
  PGN_CHECK(n, std::string("missing <") + what + ">");
  const auto v = csv(n->text);
  PGN_CHECK(v.size() == 3, std::string("<") + what + "> needs 3 comma-separated values");
  return {v[0], v[1], v[2]};
}

const XNode* need(const XNode* n, const char* c) {
  const XNode* k = n->child(c);
  PGN_CHECK(k, "<" + n->name + "> needs a <" + c + "> child");
  return k;
}

std::string need_attr(const XNode* n, const char* a) {
  const std::string* v = n->at(a);
  PGN_CHECK(v, "<" + n->name + "> needs the attribute " + a);
  return *v;
}

bool as_bool(const std::string& s) { return s == "true" || s == "1" || s == "True"; }

}  // namespace

namespace pgn {

// TetGen volume mesh: <base>.node ("<#points> <dim> <#attr> <#markers>",
// then "<index> x y z ...") and <base>.ele ("<#tets> <nodes per tet>
// <#attr>", then "<index> n1 n2 n3 n4 ..."); '#' starts a comment; indices
// start at 0 or 1, whichever the .node file uses.  Tets whose corners give
// a negative volume get their last two corners swapped, since the element
// needs det(Dm) > 0 at rest.
void read_tetgen(const std::string& base_in, std::vector<double>& V, std::vector<int>& T) {
  std::string base = base_in;
  for (const char* ext : {".node", ".ele"})
    if (base.size() > std::strlen(ext) &&
        base.compare(base.size() - std::strlen(ext), std::strlen(ext), ext) == 0)
      base.resize(base.size() - std::strlen(ext));
  auto tokens = [](const std::string& text) {
    std::vector<std::vector<std::string>> lines;
    std::stringstream ss(text);
    std::string line;
    while (std::getline(ss, line)) {
      const size_t c = line.find('#');
      if (c != std::string::npos) line.resize(c);
      std::stringstream ls(line);
      std::vector<std::string> t;
      std::string w;
      while (ls >> w) t.push_back(w);
      if (!t.empty()) lines.push_back(t);
    }
    return lines;
  };
  const auto nl = tokens(read_file(base + ".node"));
  PGN_CHECK(!nl.empty() && nl[0].size() >= 2, base + ".node: missing header");
  const long np = std::stol(nl[0][0]);
  PGN_CHECK(std::stoi(nl[0][1]) == 3, base + ".node: points must be 3D");
  PGN_CHECK(np >= 1 && (long)nl.size() >= np + 1, base + ".node: fewer points than declared");
  const long first = std::stol(nl[1][0]);
  PGN_CHECK(first == 0 || first == 1, base + ".node: indices must start at 0 or 1");
  V.assign(3 * np, 0.0);
  for (long i = 0; i < np; ++i) {
    const auto& t = nl[1 + i];
    PGN_CHECK(t.size() >= 4 && std::stol(t[0]) == first + i,
              base + ".node: point " + std::to_string(first + i) + " malformed or out of order");
    for (int c = 0; c < 3; ++c) V[3 * i + c] = std::stod(t[1 + c]);
  }
  const auto el = tokens(read_file(base + ".ele"));
  PGN_CHECK(!el.empty() && el[0].size() >= 2, base + ".ele: missing header");
  const long nt = std::stol(el[0][0]);
  const int per = std::stoi(el[0][1]);
  PGN_CHECK(per == 4, base + ".ele: only 4-node tets (Tet4) are supported");
  PGN_CHECK(nt >= 0 && (long)el.size() >= nt + 1, base + ".ele: fewer tets than declared");
  T.assign(4 * nt, 0);
  for (long e = 0; e < nt; ++e) {
    const auto& t = el[1 + e];
    PGN_CHECK(t.size() >= 5, base + ".ele: tet " + std::to_string(e) + " malformed");
    int q[4];
    for (int a = 0; a < 4; ++a) {
      const long id = std::stol(t[1 + a]) - first;
      PGN_CHECK(id >= 0 && id < np, base + ".ele: node id out of range");
      q[a] = (int)id;
    }
    double d[3][3];
    for (int a = 1; a < 4; ++a)
      for (int c = 0; c < 3; ++c) d[a - 1][c] = V[3L * q[a] + c] - V[3L * q[0] + c];
    const double det = d[0][0] * (d[1][1] * d[2][2] - d[1][2] * d[2][1]) -
                       d[0][1] * (d[1][0] * d[2][2] - d[1][2] * d[2][0]) +
                       d[0][2] * (d[1][0] * d[2][1] - d[1][1] * d[2][0]);
    PGN_CHECK(det != 0.0, base + ".ele: degenerate tet " + std::to_string(e));
    if (det < 0.0) std::swap(q[2], q[3]);
    for (int a = 0; a < 4; ++a) T[4 * e + a] = q[a];
  }
}

// CreateTetMesh_Grid (MeshUtils.cpp:927-991): nodes x-fastest on the
// regular grid of `size` cut into dims cells, 5 tets per hex cell
void tet_grid(const double size[3], const int dims[3], std::vector<double>& V,
              std::vector<int>& T) {
  const int nx = dims[0] + 1, ny = dims[1] + 1, nz = dims[2] + 1;
  V.resize(3L * nx * ny * nz);
  long c = 0;
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i, ++c) {
        V[3 * c] = i * (size[0] / dims[0]);
        V[3 * c + 1] = j * (size[1] / dims[1]);
        V[3 * c + 2] = k * (size[2] / dims[2]);
      }
  T.clear();
  T.reserve(20L * dims[0] * dims[1] * dims[2]);
  const int line = nx, slab = nx * ny;
  for (int k = 0; k < dims[2]; ++k)
    for (int j = 0; j < dims[1]; ++j)
      for (int i = 0; i < dims[0]; ++i) {
        const int o = slab * k + line * j + i;
        // hex corners: bit 0 = +x, bit 1 = +y, bit 2 = +z
        const int h[8] = {o, o + 1, o + line, o + line + 1, o + slab, o + slab + 1,
                          o + slab + line, o + slab + line + 1};
        const int split[5][4] = {{0, 1, 2, 4}, {6, 7, 4, 2}, {5, 4, 7, 1}, {3, 2, 1, 7},
                                 {1, 2, 4, 7}};
        for (auto& t : split)
          for (int a = 0; a < 4; ++a) T.push_back(h[t[a]]);
      }
}

}  // namespace pgn

namespace {

// LoadMesh_TetVolume (Serializer_XML.cpp:878-983)
void load_tet_volume(const XNode* e, const std::string& dir, std::vector<double>& V,
                     std::vector<int>& T, int* grid) {
  const std::string type = need_attr(e, "type");
  if (type == "FILEVOLUME") {
    const std::string* p = e->at("path");
    if (!p) p = e->at("mesh");
    PGN_CHECK(p, "<" + e->name + " type=\"FILEVOLUME\"> needs the attribute path");
    const std::string path = (!p->empty() && (*p)[0] == '/') ? *p : dir + *p;
    read_tetgen(path, V, T);
  } else if (type == "GRID" || type == "Grid") {
    const auto d = vec3(e->child("dims"), "dims");
    const auto z = vec3(e->child("size"), "size");
    const int dims[3] = {(int)d[0], (int)d[1], (int)d[2]};
    PGN_CHECK(dims[0] >= 1 && dims[1] >= 1 && dims[2] >= 1, "<dims> must be >= 1");
    const double size[3] = {z[0], z[1], z[2]};
    tet_grid(size, dims, V, T);
    if (grid)
      for (int c = 0; c < 3; ++c) grid[c] = dims[c];
  } else {
    throw Error(PGN_ERR_INVALID, "meshInit type " + type +
                                     " is not supported (FILEVOLUME or GRID); FILESURFACE "
                                     "needs a tetrahedralizer");
  }
  // position, scale, rotation: x' = R S x + t (R = Rz Ry Rx, degrees)
  std::array<double, 3> t{0, 0, 0}, sc{1, 1, 1}, r{0, 0, 0};
  if (e->child("position")) t = vec3(e->child("position"), "position");
  if (e->child("scale")) sc = vec3(e->child("scale"), "scale");
  if (e->child("rotation")) r = vec3(e->child("rotation"), "rotation");
  const double deg = std::acos(-1.0) / 180.0;
  const double a = r[0] * deg, b = r[1] * deg, c = r[2] * deg;
  const double Rx[3][3] = {{1, 0, 0}, {0, std::cos(a), -std::sin(a)}, {0, std::sin(a), std::cos(a)}};
  const double Ry[3][3] = {{std::cos(b), 0, std::sin(b)}, {0, 1, 0}, {-std::sin(b), 0, std::cos(b)}};
  const double Rz[3][3] = {{std::cos(c), -std::sin(c), 0}, {std::sin(c), std::cos(c), 0}, {0, 0, 1}};
  double Ryx[3][3], R[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      Ryx[i][j] = 0.0;
      for (int k = 0; k < 3; ++k) Ryx[i][j] += Ry[i][k] * Rx[k][j];
    }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      R[i][j] = 0.0;
      for (int k = 0; k < 3; ++k) R[i][j] += Rz[i][k] * Ryx[k][j];
    }
  const bool identity = r == std::array<double, 3>{0, 0, 0} &&
                        sc == std::array<double, 3>{1, 1, 1} && t == std::array<double, 3>{0, 0, 0};
  if (identity) return;
  for (size_t q = 0; q + 2 < V.size(); q += 3) {
    const double s[3] = {sc[0] * V[q], sc[1] * V[q + 1], sc[2] * V[q + 2]};
    for (int i = 0; i < 3; ++i) V[q + i] = R[i][0] * s[0] + R[i][1] * s[1] + R[i][2] * s[2] + t[i];
  }
}

// LoadDistribution (Serializer_XML.cpp:618-733): CONSTANT (one object,
// the child of the same name) or VARIABLE (a catalogue and an optional
// <distribution> of per-element catalogue indices, identity by default)
template <class Obj, class F>
std::vector<Obj> load_distribution(const XNode* e, int m, F&& load_one) {
  const std::string type = need_attr(e, "type");
  const auto items = e->children(e->name.c_str());
  PGN_CHECK(!items.empty(), "<" + e->name + "> needs <" + e->name + "> children");
  if (type == "CONSTANT") return std::vector<Obj>(m, load_one(items[0]));
  PGN_CHECK(type == "VARIABLE", "<" + e->name + "> distribution " + type +
                                    " is not supported (CONSTANT or VARIABLE)");
  std::vector<Obj> cat;
  for (auto* it : items) cat.push_back(load_one(it));
  std::vector<Obj> out(m);
  if (const XNode* d = e->child("distribution")) {
    const auto idx = csv(d->text);
    PGN_CHECK((long)idx.size() == m, "<distribution> needs one index per element");
    for (int k = 0; k < m; ++k) {
      const int i = (int)idx[k];
      PGN_CHECK(i >= 0 && i < (int)cat.size(), "<distribution> index out of range");
      out[k] = cat[i];
    }
  } else {
    PGN_CHECK((int)cat.size() == m, "a VARIABLE distribution without <distribution> needs one "
                                    "entry per element");
    out = cat;
  }
  return out;
}

struct Params {
  double l1 = 0, l2 = 0, rho = 0;
};

// LoadDistributed<ParameterSet> (Serializer_XML.cpp:577-616) and
// ParameterSet::InitLinearFromYoungPoisson (ParameterSet.cpp:54-70)
Params load_params(const XNode* e) {
  std::map<std::string, double> p;
  for (auto* q : e->children("parameter"))
    p[need_attr(q, "name")] = std::strtod(need_attr(q, "value").c_str(), nullptr);
  Params r;
  PGN_CHECK(p.count("DENSITY"), "a parameter set needs DENSITY");
  r.rho = p["DENSITY"];
  if (p.count("YOUNG") && p.count("POISSON")) {
    const double E = p["YOUNG"], nu = p["POISSON"];
    r.l1 = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
    r.l2 = E / (2.0 * (1.0 + nu));
  } else if (p.count("LAME1") && p.count("LAME2")) {
    r.l1 = p["LAME1"];
    r.l2 = p["LAME2"];
  } else {
    throw Error(PGN_ERR_INVALID, "a parameter set needs YOUNG and POISSON or LAME1 and LAME2");
  }
  return r;
}

// LoadDistributed<IMaterialModel> (Serializer_XML.cpp:528-575), 3 in 3
uint8_t load_material(const XNode* e) {
  const std::string t = need_attr(e, "type");
  PGN_CHECK(need_attr(e, "dimSpace") == "3" && need_attr(e, "dimBasis") == "3",
            "volume materials are dimSpace 3, dimBasis 3");
  if (t == "CONH") return PGN_CONH;
  if (t == "STVK") return PGN_STVK;
  throw Error(PGN_ERR_INVALID, "material model " + t + " is not supported (CONH or STVK)");
}

// the single-object loaders the scene needs to see but that change nothing
// here: the linear Tet4 shape function and the one-point rule
std::string load_name(const XNode* e) { return need_attr(e, "type"); }

// boundary nodes: on a GRID mesh the nodes of the box's faces (its surface
// is the grid's, create3DGridIndex_Tri3; the 5-tet split does not conform
// across cells, so face counting would see interior faces), otherwise the
// nodes of faces that belong to one tet only
std::vector<uint8_t> boundary_nodes(int n, const std::vector<int>& T, const int grid[3]) {
  if (grid[0] > 0) {
    std::vector<uint8_t> on(n, 0);
    const int nx = grid[0] + 1, ny = grid[1] + 1, nz = grid[2] + 1;
    for (int k = 0; k < nz; ++k)
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i)
          on[(k * ny + j) * nx + i] = (i == 0 || j == 0 || k == 0 || i == nx - 1 ||
                                       j == ny - 1 || k == nz - 1) ? 1 : 0;
    return on;
  }
  std::map<std::array<int, 3>, int> faces;
  const int fl[4][3] = {{0, 1, 2}, {0, 1, 3}, {0, 2, 3}, {1, 2, 3}};
  for (size_t e = 0; e + 3 < T.size(); e += 4)
    for (auto& f : fl) {
      std::array<int, 3> k{T[e + f[0]], T[e + f[1]], T[e + f[2]]};
      std::sort(k.begin(), k.end());
      faces[k]++;
    }
  std::vector<uint8_t> on(n, 0);
  for (auto& kv : faces)
    if (kv.second == 1)
      for (int v : kv.first) on[v] = 1;
  return on;
}

// LoadGeometryFilter (Serializer_XML.cpp:1758-1868), node selections
std::vector<uint8_t> select_nodes(const XNode* e, const pgn_scene& S) {
  const std::string t = need_attr(e, "type");
  std::vector<uint8_t> on(S.n, 0);
  if (t == "AllNodes") {
    std::fill(on.begin(), on.end(), 1);
  } else if (t == "BoundaryNodes") {
    on = boundary_nodes(S.n, S.T, S.grid);
  } else if (t == "IndexNodes") {
    for (double d : csv(need(e, "indices")->text)) {
      const long i = (long)d;
      PGN_CHECK(i >= 0 && i < S.n, "<indices>: node index out of range");
      on[i] = 1;
    }
  } else if (t == "BoxNodes") {
    const auto lo = vec3(e->child("minCorner"), "minCorner");
    const auto hi = vec3(e->child("maxCorner"), "maxCorner");
    const std::string f = need_attr(e, "field");
    PGN_CHECK(f == "Pos0" || f == "PosX", "BoxNodes field must be Pos0 or PosX");
    const std::vector<double>& P = f == "Pos0" ? S.X0 : S.X;
    for (int i = 0; i < S.n; ++i) {
      bool in = true;
      for (int c = 0; c < 3; ++c) in &= P[3L * i + c] >= lo[c] && P[3L * i + c] <= hi[c];
      on[i] = in ? 1 : 0;
    }
  } else {
    throw Error(PGN_ERR_INVALID, "selection " + t +
                                     " is not supported (AllNodes, BoundaryNodes, IndexNodes, "
                                     "BoxNodes)");
  }
  return on;
}

void load_scene(const std::string& path, pgn_scene& S) {
  const std::string text = read_file(path);
  auto root = XmlReader(text).document();
  PGN_CHECK(root->name == "scene", "the root element must be <scene>");
  const std::string dir = dir_of(path);
  // ---- the model (LoadSimulable_VolumeFEM) ----
  const XNode* sim = need(root.get(), "simulable");
  for (const char* c : {"meshInit", "discretization", "shapeFunction", "materialParam",
                        "materialModel", "quadrature"})
    need(sim, c);
  const std::string dis = load_name(need(sim, "discretization"));
  PGN_CHECK(dis == "TET4", "discretization " + dis + " is not supported (TET4)");
  std::vector<double> V;
  load_tet_volume(need(sim, "meshInit"), dir, V, S.T, S.grid);
  S.n = (int)(V.size() / 3);
  S.m = (int)(S.T.size() / 4);
  PGN_CHECK(S.n > 0 && S.m > 0, "empty mesh");
  S.X = V;
  if (const XNode* rest = sim->child("meshRest")) {
    std::vector<int> T0;
    load_tet_volume(rest, dir, S.X0, T0, nullptr);
    PGN_CHECK(T0 == S.T && S.X0.size() == V.size(), "meshRest must match meshInit's elements");
  } else {
    S.X0 = V;
  }
  const auto sf = load_distribution<std::string>(need(sim, "shapeFunction"), S.m, load_name);
  const auto qd = load_distribution<std::string>(need(sim, "quadrature"), S.m, load_name);
  for (int e = 0; e < S.m; ++e) {
    PGN_CHECK(sf[e] == "TET_LIN", "shape function " + sf[e] + " is not supported (TET_LIN)");
    PGN_CHECK(qd[e] == "TET_1", "quadrature " + qd[e] + " is not supported (TET_1)");
  }
  S.kind = load_distribution<uint8_t>(need(sim, "materialModel"), S.m, load_material);
  const auto pr = load_distribution<Params>(need(sim, "materialParam"), S.m, load_params);
  S.l1.resize(S.m);
  S.l2.resize(S.m);
  S.rho.resize(S.m);
  for (int e = 0; e < S.m; ++e) {
    S.l1[e] = pr[e].l1;
    S.l2[e] = pr[e].l2;
    S.rho[e] = pr[e].rho;
  }
  // ---- boundary conditions (LoadBoundary) ----
  S.fixed.assign(3L * S.n, 0);
  if (const XNode* bnd = root->child("boundary")) {
    for (auto* bc : bnd->children("bc")) {
      const std::string t = need_attr(bc, "type");
      const std::string* ns = bc->at("numStep");
      if (t == "Gravity") {
        // BC_Gravity without an animator: the same value at every stage;
        // the stage count only decides when the problem is fully loaded
        // (BCondition.h:128-155); default (0, -9.8, 0) (:1537)
        PGN_CHECK(!S.has_g, "one Gravity condition per scene");
        std::array<double, 3> g{0.0, -9.8, 0.0};
        if (bc->child("gravity")) g = vec3(bc->child("gravity"), "gravity");
        S.has_g = true;
        for (int c = 0; c < 3; ++c) S.g[c] = g[c];
        S.g_stages = ns ? std::atoi(ns->c_str()) : 1;
        PGN_CHECK(S.g_stages >= 1, "numStep must be >= 1");
      } else if (t == "FixPosition") {
        PGN_CHECK(!as_bool(need_attr(bc, "soft")),
                  "soft FixPosition conditions are not supported (soft=\"false\")");
        PGN_CHECK(!bc->child("animation"), "animated FixPosition conditions are not supported");
        const auto on = select_nodes(need(bc, "selection"), S);
        for (int i = 0; i < S.n; ++i)
          if (on[i]) S.fixed[3L * i] = S.fixed[3L * i + 1] = S.fixed[3L * i + 2] = 1;
      } else if (t == "AlignFace" || t == "Force") {
        // constructed as no-ops by the reference (:1595-1597, :1631-1633)
      } else {
        throw Error(PGN_ERR_INVALID, "boundary condition " + t +
                                         " is not supported (Gravity, FixPosition)");
      }
    }
  }
  // ---- problem (LoadOptimProblem) and solver (LoadOptimSolver) ----
  if (const XNode* pb = root->child("problem")) {
    const std::string t = need_attr(pb, "type");
    if (t == "Dynamic") {
      S.dynamic = 1;
      S.dt = std::strtod(need_attr(pb, "step").c_str(), nullptr);
      PGN_CHECK(S.dt > 0.0, "a Dynamic problem needs step > 0");
    } else {
      PGN_CHECK(t == "Static", "problem type " + t + " is not supported (Static, Dynamic)");
    }
  }
  pgn_newton_default_opts(&S.opts);
  if (const XNode* so = root->child("solver")) {
    if (auto* v = so->at("maxIters")) S.opts.numMaxIters = std::atoi(v->c_str());
    if (auto* v = so->at("maxError")) S.opts.tolMaxError = std::strtod(v->c_str(), nullptr);
    if (auto* v = so->at("qpStep")) {
      if (*v == "Newton") S.opts.qpSolverType = PGN_QP_NEWTON;
      else if (*v == "Steepest") S.opts.qpSolverType = PGN_QP_STEEPEST;
      else if (*v == "LBFGS") S.opts.qpSolverType = PGN_QP_LBFGS;
      else throw Error(PGN_ERR_INVALID, "qpStep " + *v + " is not supported (Newton, Steepest, LBFGS)");
    }
    if (const XNode* li = so->child("linear")) {
      if (auto* v = li->at("type")) {
        if (*v == "EigenCG") S.opts.lsSolverType = PGN_LS_EIGENCG;
        else if (*v == "EigenLDLT") S.opts.lsSolverType = PGN_LS_EIGENLDLT;
        else throw Error(PGN_ERR_INVALID, "linear solver " + *v + " is not supported (EigenCG, EigenLDLT)");
      }
      if (auto* v = li->at("maxIters")) S.opts.lsNumMaxIters = std::atoi(v->c_str());
      if (auto* v = li->at("maxError")) S.opts.lsTolMaxError = std::strtod(v->c_str(), nullptr);
    }
    if (const XNode* ls = so->child("lsearch")) {
      if (auto* v = ls->at("type"))
        PGN_CHECK(*v == "Simple", "line search " + *v + " is not supported (Simple)");
      if (auto* v = ls->at("iters")) S.opts.lsearch_numTry = std::atoi(v->c_str());
      if (auto* v = ls->at("alpha")) S.opts.lSearch_alpha = std::strtod(v->c_str(), nullptr);
    }
  }
}

template <class F>
int sc_guard(F&& f) {
  try {
    f();
    return PGN_OK;
  } catch (const Error& e) {
    return api_guard_set(e.what(), e.code);
  } catch (const std::bad_alloc&) {
    return api_guard_set("host allocation failed", PGN_ERR_NOMEM);
  } catch (const std::exception& e) {
    return api_guard_set(e.what(), PGN_ERR_INVALID);
  }
}

}  // namespace

extern "C" {

int pgn_tetgen_sizes(const char* base, int* n_nodes, int* n_tets) {
  return sc_guard([&] {
    PGN_CHECK(base && n_nodes && n_tets, "null argument");
    std::vector<double> V;
    std::vector<int> T;
    read_tetgen(base, V, T);
    *n_nodes = (int)(V.size() / 3);
    *n_tets = (int)(T.size() / 4);
  });
}

int pgn_tetgen_read(const char* base, double* X, int* T) {
  return sc_guard([&] {
    PGN_CHECK(base && X && T, "null argument");
    std::vector<double> V;
    std::vector<int> E;
    read_tetgen(base, V, E);
    std::copy(V.begin(), V.end(), X);
    std::copy(E.begin(), E.end(), T);
  });
}

int pgn_scene_load(const char* path, pgn_scene** out) {
  return sc_guard([&] {
    PGN_CHECK(path && out, "null argument");
    *out = nullptr;
    auto S = std::make_unique<pgn_scene>();
    load_scene(path, *S);
    *out = S.release();
  });
}

void pgn_scene_destroy(pgn_scene* scene) { delete scene; }

int pgn_scene_sizes(const pgn_scene* S, int* n_nodes, int* n_tets) {
  return sc_guard([&] {
    PGN_CHECK(S && n_nodes && n_tets, "null argument");
    *n_nodes = S->n;
    *n_tets = S->m;
  });
}

int pgn_scene_mesh(const pgn_scene* S, double* X0, double* X, int* T, unsigned char* material,
                   double* lame1, double* lame2, double* density) {
  return sc_guard([&] {
    PGN_CHECK(S, "null argument");
    if (X0) std::copy(S->X0.begin(), S->X0.end(), X0);
    if (X) std::copy(S->X.begin(), S->X.end(), X);
    if (T) std::copy(S->T.begin(), S->T.end(), T);
    if (material) std::copy(S->kind.begin(), S->kind.end(), material);
    if (lame1) std::copy(S->l1.begin(), S->l1.end(), lame1);
    if (lame2) std::copy(S->l2.begin(), S->l2.end(), lame2);
    if (density) std::copy(S->rho.begin(), S->rho.end(), density);
  });
}

int pgn_scene_boundary(const pgn_scene* S, unsigned char* fixed, double g_ini[3],
                       double g_end[3], int* stages, int* has_gravity) {
  return sc_guard([&] {
    PGN_CHECK(S, "null argument");
    if (fixed) std::copy(S->fixed.begin(), S->fixed.end(), fixed);
    for (int c = 0; c < 3; ++c) {
      if (g_ini) g_ini[c] = S->g[c];
      if (g_end) g_end[c] = S->g[c];
    }
    if (stages) *stages = S->g_stages;
    if (has_gravity) *has_gravity = S->has_g ? 1 : 0;
  });
}

int pgn_scene_solver(const pgn_scene* S, pgn_newton_opts* opts, int* dynamic, double* dt) {
  return sc_guard([&] {
    PGN_CHECK(S, "null argument");
    if (opts) *opts = S->opts;
    if (dynamic) *dynamic = S->dynamic;
    if (dt) *dt = S->dt;
  });
}

int pgn_scene_create_model(pgn_ctx* ctx, const pgn_scene* S, int precision, pgn_model** out) {
  return sc_guard([&] {
    PGN_CHECK(ctx && S && out, "null argument");
    *out = nullptr;
    pgn_model* M = nullptr;
    int rc = pgn_model_create_tet(ctx, S->n, S->X0.data(), S->m, S->T.data(), S->kind.data(),
                                  S->l1.data(), S->l2.data(), S->rho.data(), precision, &M);
    if (rc != PGN_OK) throw Error(rc, pgn_last_error());
    auto drop = [&](int code) {
      const std::string msg = pgn_last_error();
      pgn_model_destroy(M);
      throw Error(code, msg);
    };
    if ((rc = pgn_model_set_fixed(M, S->fixed.data())) != PGN_OK) drop(rc);
    if (S->has_g && (rc = pgn_model_set_gravity(M, S->g, S->g, S->g_stages)) != PGN_OK) drop(rc);
    if (S->X != S->X0 && (rc = pgn_model_set_x(M, S->X.data())) != PGN_OK) drop(rc);
    *out = M;
  });
}

}  // extern "C"
