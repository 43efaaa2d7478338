// mesh.cpp -- Gmsh 2.2 reader, periodic face pairing and global flux-point connectivity.
//
// Output contract (kept bit-exact with the reference, src/geometry.cpp:1257-1744 and
// src/solver.cpp:312-437): fpt2gfpt[fpt][ele] and slot[fpt][ele] as the reference has
// them after orient_fpts.  The construction is this project's own:
//  * a single pass over element faces with a hash map keyed by sorted node ids
//    (the reference uses std::map/std::set with a separate counting pass);
//  * periodic partners found through hashed face centroids instead of the reference's
//    all-pairs scan (src/geometry.cpp:1159-1255), with the reference's acceptance test
//    and its tie rule (first match in sorted-node order);
//  * the orientation step written as a closed form: the new number of a flux point is
//    the rank, among all slot-0 element flux points ordered by (fpt, ele), of the
//    slot-0 flux point whose in-face position equals the point's fuzzy-sorted rank.
//    That is what the reference's fuzzysort + address sort + index inversion computes.
#include <algorithm>
#include <array>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

#include "common.hpp"
#include "host.hpp"

namespace zfr {

namespace {

struct KeyHash {
  size_t operator()(const FaceKey &k) const {
    uint64_t h = 1469598103934665603ull;
    for (unsigned v : k) { h ^= v; h *= 1099511628211ull; }
    return (size_t)h;
  }
};

// Minimal tokenizer over the whole file.
struct Tok {
  std::string buf;
  size_t pos = 0;
  explicit Tok(const std::string &file) {
    std::ifstream f(file, std::ios::binary);
    if (!f.is_open()) raise("could not open mesh file " + file);
    std::stringstream ss;
    ss << f.rdbuf();
    buf = ss.str();
  }
  bool seek(const char *tag) {
    const size_t p = buf.find(tag);
    if (p == std::string::npos) return false;
    pos = buf.find('\n', p);
    pos = (pos == std::string::npos) ? buf.size() : pos + 1;
    return true;
  }
  void skip_ws() { while (pos < buf.size() && (unsigned char)buf[pos] <= ' ') pos++; }
  long integer() {
    skip_ws();
    char *end;
    const long v = std::strtol(buf.c_str() + pos, &end, 10);
    if (end == buf.c_str() + pos) raise("mesh file: expected an integer");
    pos = end - buf.c_str();
    return v;
  }
  double real() {
    skip_ws();
    char *end;
    const double v = std::strtod(buf.c_str() + pos, &end);
    if (end == buf.c_str() + pos) raise("mesh file: expected a number");
    pos = end - buf.c_str();
    return v;
  }
  std::string line() {
    const size_t e = buf.find('\n', pos);
    std::string s = buf.substr(pos, (e == std::string::npos ? buf.size() : e) - pos);
    pos = (e == std::string::npos) ? buf.size() : e + 1;
    return s;
  }
};

FaceKey make_key(const unsigned *nodes, int n) {
  FaceKey k = {UINT_MAX, UINT_MAX, UINT_MAX, UINT_MAX};
  for (int i = 0; i < n; i++) k[i] = nodes[i];
  std::sort(k.begin(), k.begin() + n);
  return k;
}

int key_len(const FaceKey &k) {
  int n = 0;
  while (n < 4 && k[n] != UINT_MAX) n++;
  return n;
}

// Element face node tables (src/geometry.cpp:218-229, 247-265).
const std::vector<std::vector<int>> &face_nodes(int nDims) {
  static const std::vector<std::vector<int>> quad = {{0, 1}, {1, 2}, {2, 3}, {3, 0}};
  static const std::vector<std::vector<int>> hex = {{0, 1, 2, 3}, {5, 4, 7, 6}, {0, 3, 7, 4},
                                                    {2, 1, 5, 6}, {1, 0, 4, 5}, {3, 2, 6, 7}};
  return nDims == 2 ? quad : hex;
}

std::string lower(std::string s) {
  std::transform(s.begin(), s.end(), s.begin(), ::tolower);
  return s;
}

}  // namespace

// Gmsh 2.2 ASCII (src/geometry.cpp:307-958), restricted to linear quads / hexes and the
// periodic / fluid physical groups.
Mesh read_gmsh(const std::string &file, const zefr_input_t &in) {
  Mesh m;
  m.nDims = (int)in.nDims;
  if (m.nDims != 2 && m.nDims != 3) raise("nDims must be 2 or 3");
  Tok t(file);
  // physical names -> boundary condition through the deck's mesh_bound map
  if (!t.seek("$PhysicalNames")) raise("Meshfile missing $PhysicalNames tag");
  const long nPhys = t.integer();
  t.line();
  std::map<long, int> phys;  // gmsh id -> 0 fluid, 1 periodic
  for (long i = 0; i < nPhys; i++) {
    std::stringstream ss(t.line());
    int bdim; long bid; std::string name;
    ss >> bdim >> bid >> name;
    name.erase(std::remove(name.begin(), name.end(), '"'), name.end());
    name = lower(name);
    std::string bc = name;
    if (in.nMeshBounds > 0) {
      bool found = false;
      for (int k = 0; k < in.nMeshBounds; k++)
        if (name == in.meshBoundName[k]) { bc = in.meshBoundBC[k]; found = true; }
      if (!found) raise("Unrecognized mesh boundary: \"" + name + "\"");
    }
    if (bc == "fluid") {
      if (bdim != m.nDims) raise("nDims in mesh file does not match input-specified nDims.");
      phys[bid] = 0;
    } else if (bc == "periodic") {
      phys[bid] = 1;
    } else {
      raise("boundary condition '" + bc + "' is not supported: the residual path handles periodic meshes");
    }
  }
  if (!t.seek("$Nodes")) raise("Meshfile missing $Nodes tag");
  m.nNodes = (int)t.integer();
  m.coord.assign((size_t)m.nNodes * m.nDims, 0.0);
  for (int n = 0; n < m.nNodes; n++) {
    t.integer();
    double xyz[3];
    for (int d = 0; d < 3; d++) xyz[d] = t.real();
    for (int d = 0; d < m.nDims; d++) m.coord[(size_t)n * m.nDims + d] = xyz[d];
  }
  if (!t.seek("$Elements")) raise("Meshfile missing $Elements tag");
  const long nTot = t.integer();
  m.etype = (m.nDims == 2) ? QUAD_T : HEX_T;
  m.nNodesPerEle = (m.nDims == 2) ? 4 : 8;
  const long eleType = (m.nDims == 2) ? 3 : 5, bndType = (m.nDims == 2) ? 1 : 3;
  const int nBndNodes = (m.nDims == 2) ? 2 : 4;
  for (long n = 0; n < nTot; n++) {
    t.integer();
    const long type = t.integer();
    const long nTags = t.integer();
    long tag0 = -1;
    for (long k = 0; k < nTags; k++) {
      const long v = t.integer();
      if (k == 0) tag0 = v;
    }
    if (type == eleType) {
      for (int k = 0; k < m.nNodesPerEle; k++) m.ele2nodes.push_back((int)t.integer() - 1);
    } else if (type == bndType) {
      unsigned nd[4];
      for (int k = 0; k < nBndNodes; k++) nd[k] = (unsigned)(t.integer() - 1);
      auto it = phys.find(tag0);
      if (it == phys.end()) raise("boundary face with an unknown physical tag");
      if (it->second != 1) raise("boundary face tagged as fluid");
      m.per_faces.push_back(make_key(nd, nBndNodes));
    } else if (type == 2 && m.nDims == 2) {
      // The reference's triangle operators (src/elements.cpp:555-744) leave oppDiv without
      // its volume derivative for every non-tensor element (it is filled only for quads and
      // hexes, :591-621, and becomes -M3 M2 at :728-730): its triangle residual is the
      // interface correction alone, so there is no triangle result to reproduce.
      raise("triangular elements (Gmsh type 2) are not supported: the reference solver's "
            "triangle residual has no volume divergence term (src/elements.cpp:591-730)");
    } else {
      raise("Gmsh element type " + std::to_string(type) +
            " is not supported (linear quads / hexes and their boundary faces)");
    }
    t.line();
  }
  m.nEles = (int)(m.ele2nodes.size() / m.nNodesPerEle);
  if (m.nEles == 0) raise("mesh has no elements");
  for (int v : m.ele2nodes)
    if (v < 0 || v >= m.nNodes) raise("element references a node outside $Nodes");
  // duplicates would make the boundary map ambiguous; the reference keeps the last one
  std::sort(m.per_faces.begin(), m.per_faces.end());
  m.per_faces.erase(std::unique(m.per_faces.begin(), m.per_faces.end()), m.per_faces.end());
  couple_periodic(m);
  return m;
}

// Pair every periodic face with the face whose centroid matches it (|dc| < 1e-6) in all
// directions in which the face is not planar (src/geometry.cpp:1159-1255).  Faces are
// bucketed by their centroid with the normal direction left out; per_faces is sorted, so
// the smallest accepted candidate index is the reference's first match.
void couple_periodic(Mesh &m) {
  const int D = m.nDims;
  const size_t nP = m.per_faces.size();
  m.per_pair.assign(nP, -1);
  if (nP == 0) return;
  const double tol = 1.e-6, cell = 2.e-6;
  std::vector<double> cen(nP * D, 0.0);
  std::vector<int> planar_dim(nP, -1), n_planar(nP, 0);
  for (size_t i = 0; i < nP; i++) {
    const FaceKey &k = m.per_faces[i];
    const int n = key_len(k);
    for (int d = 0; d < D; d++) {
      double c = 0.0;
      for (int a = 0; a < n; a++) c += m.coord[(size_t)k[a] * D + d];
      cen[i * D + d] = c / n;
      bool onPlane = true;
      const double c0 = m.coord[(size_t)k[0] * D + d];
      for (int a = 1; a < n; a++)
        if (std::abs(c0 - m.coord[(size_t)k[a] * D + d]) > tol) onPlane = false;
      if (onPlane) { n_planar[i]++; planar_dim[i] = d; }
    }
  }
  // the reference's acceptance test for (face1 = i, face2 = j)
  auto accepts = [&](size_t i, size_t j) {
    if (i == j || key_len(m.per_faces[i]) != key_len(m.per_faces[j])) return false;
    const FaceKey &k = m.per_faces[i];
    const int n = key_len(k);
    int count = 0;
    for (int d = 0; d < D; d++) {
      if (std::abs(cen[i * D + d] - cen[j * D + d]) < tol) {
        bool onPlane = true;
        const double c0 = m.coord[(size_t)k[0] * D + d];
        for (int a = 1; a < n; a++)
          if (std::abs(c0 - m.coord[(size_t)k[a] * D + d]) > tol) onPlane = false;
        if (!onPlane) count++;
      }
    }
    return count == D - 1;
  };
  // one bucket index per left-out direction
  auto bucket_key = [&](size_t i, int skip, const long *shift) {
    uint64_t h = 1469598103934665603ull ^ (uint64_t)skip;
    int q = 0;
    for (int d = 0; d < D; d++) {
      if (d == skip) continue;
      const long b = (long)std::floor(cen[i * D + d] / cell) + (shift ? shift[q] : 0);
      h ^= (uint64_t)b; h *= 1099511628211ull;
      q++;
    }
    return h;
  };
  std::vector<std::unordered_multimap<uint64_t, size_t>> index(D);
  for (int skip = 0; skip < D; skip++) {
    index[skip].reserve(nP * 2);
    for (size_t j = 0; j < nP; j++) index[skip].emplace(bucket_key(j, skip, nullptr), j);
  }
  for (size_t i = 0; i < nP; i++) {
    long best = -1;
    if (n_planar[i] == 1) {
      const int skip = planar_dim[i];
      const int nq = D - 1;
      long shift[2] = {0, 0};
      const int ncomb = (nq == 1) ? 3 : 9;
      for (int c = 0; c < ncomb; c++) {
        shift[0] = c % 3 - 1;
        shift[1] = (c / 3) % 3 - 1;
        auto range = index[skip].equal_range(bucket_key(i, skip, shift));
        for (auto it = range.first; it != range.second; ++it) {
          const size_t j = it->second;
          if (accepts(i, j) && (best < 0 || (long)j < best)) best = (long)j;
        }
      }
    } else {
      for (size_t j = 0; j < nP; j++)
        if (accepts(i, j)) { best = (long)j; break; }
    }
    if (best < 0) raise("Unpaired periodic face detected. Check your mesh.");
    m.per_pair[i] = (int)best;
  }
}

namespace {

// Fuzzy lexicographic order of n points (include/funcs.hpp:254-276): order by the first
// coordinate, treat runs whose values stay within tol of the run's first value as ties
// and order those by the next coordinate.  rank[i] is the position of point i.
void fuzzy_rank(const double *pts, int n, int D, int *rank, double tol = 1e-6) {
  int order[256];
  for (int i = 0; i < n; i++) order[i] = i;
  struct Frame { int lo, hi, dim; };
  Frame stack[512];
  int sp = 0;
  stack[sp++] = {0, n, 0};
  while (sp > 0) {
    const Frame fr = stack[--sp];
    const int dim = fr.dim;
    std::sort(order + fr.lo, order + fr.hi,
              [&](int a, int b) { return pts[a * D + dim] < pts[b * D + dim]; });
    if (dim + 1 >= D) continue;
    int g = fr.lo;
    for (int j = fr.lo + 1; j <= fr.hi; j++) {
      if (j == fr.hi || pts[order[j] * D + dim] - pts[order[g] * D + dim] >= tol) {
        if (j - g > 1) stack[sp++] = {g, j, dim + 1};
        g = j;
      }
    }
  }
  for (int r = 0; r < n; r++) rank[order[r]] = r;
}

}  // namespace

Connectivity build_connectivity(const Mesh &m, const RefElement &re,
                                const std::function<void(int, int, double *)> &fpt_coord) {
  const int D = m.nDims, E = m.nEles, nF = re.nFpts, n = re.nFptsPerFace;
  const auto &fn = face_nodes(D);
  const int nFaceEle = (int)fn.size();
  Connectivity c;
  c.nFptsPerFace = n;
  std::unordered_map<FaceKey, int, KeyHash> per_index;
  per_index.reserve(m.per_faces.size() * 2);
  for (size_t i = 0; i < m.per_faces.size(); i++) per_index.emplace(m.per_faces[i], (int)i);

  // first / second owner of every face, in the reference's traversal order
  struct FaceRec { int eL, lfL, eR, lfR; };
  std::vector<FaceRec> faces;
  faces.reserve((size_t)E * nFaceEle / 2 + 16);
  std::unordered_map<FaceKey, int, KeyHash> seen;
  seen.reserve((size_t)E * nFaceEle);
  std::vector<int> ele_face((size_t)E * nFaceEle, -1);
  std::vector<signed char> ele_side((size_t)E * nFaceEle, -1);
  for (int e = 0; e < E; e++) {
    for (int lf = 0; lf < nFaceEle; lf++) {
      unsigned nd[4];
      for (size_t a = 0; a < fn[lf].size(); a++) nd[a] = (unsigned)m.ele2nodes[(size_t)e * m.nNodesPerEle + fn[lf][a]];
      FaceKey key = make_key(nd, (int)fn[lf].size());
      auto pit = per_index.find(key);
      if (pit != per_index.end()) {
        const FaceKey &pk = m.per_faces[m.per_pair[pit->second]];
        if (seen.count(pk)) key = pk;
      }
      auto it = seen.find(key);
      if (it == seen.end()) {
        seen.emplace(key, (int)faces.size());
        ele_face[(size_t)e * nFaceEle + lf] = (int)faces.size();
        ele_side[(size_t)e * nFaceEle + lf] = 0;
        faces.push_back({e, lf, -1, -1});
      } else {
        FaceRec &r = faces[it->second];
        if (r.eR >= 0) raise("a face is shared by more than two elements");
        r.eR = e; r.lfR = lf;
        ele_face[(size_t)e * nFaceEle + lf] = it->second;
        ele_side[(size_t)e * nFaceEle + lf] = 1;
      }
    }
  }
  for (const auto &r : faces)
    if (r.eR < 0) raise("mesh has a boundary face that is not periodic; only periodic meshes are supported");
  c.nFaces = (int)faces.size();
  c.nGfpts = c.nGfpts_int = c.nFaces * n;

  // slot of every element fpt and the rank of slot-0 fpts in (fpt, ele) order
  c.slot.assign((size_t)nF * E, 0);
  for (int e = 0; e < E; e++)
    for (int lf = 0; lf < nFaceEle; lf++)
      for (int i = 0; i < n; i++) c.slot[(size_t)(lf * n + i) * E + e] = ele_side[(size_t)e * nFaceEle + lf];
  std::vector<int> rank0((size_t)nF * E, -1);
  {
    int r = 0;
    for (size_t k = 0; k < (size_t)nF * E; k++)
      if (c.slot[k] == 0) rank0[k] = r++;
  }
  c.fpt2gfpt.assign((size_t)nF * E, -1);
  const long nFc = (long)faces.size();
#pragma omp parallel
  {
    std::vector<double> pl((size_t)n * D), pr((size_t)n * D);
    std::vector<int> rl(n), rr(n);
#pragma omp for schedule(static)
    for (long k = 0; k < nFc; k++) {
      const FaceRec &r = faces[k];
      const int fL = r.lfL * n, fR = r.lfR * n;
      for (int i = 0; i < n; i++) {
        fpt_coord(r.eL, fL + i, &pl[(size_t)i * D]);
        fpt_coord(r.eR, fR + (n - 1 - i), &pr[(size_t)i * D]);  // second owner sees the face reversed
      }
      fuzzy_rank(pl.data(), n, D, rl.data());
      fuzzy_rank(pr.data(), n, D, rr.data());
      for (int i = 0; i < n; i++) {
        c.fpt2gfpt[(size_t)(fL + i) * E + r.eL] = rank0[(size_t)(fL + rl[i]) * E + r.eL];
        c.fpt2gfpt[(size_t)(fR + (n - 1 - i)) * E + r.eR] = rank0[(size_t)(fL + rr[i]) * E + r.eL];
      }
    }
  }
  c.gf_ele.assign((size_t)2 * c.nGfpts, -1);
  c.gf_fpt.assign((size_t)2 * c.nGfpts, -1);
  for (int e = 0; e < E; e++)
    for (int f = 0; f < nF; f++) {
      const int g = c.fpt2gfpt[(size_t)f * E + e], s = c.slot[(size_t)f * E + e];
      if (g < 0 || g >= c.nGfpts) raise("internal: flux point left unnumbered");
      if (c.gf_ele[(size_t)s * c.nGfpts + g] != -1) raise("internal: flux point numbered twice");
      c.gf_ele[(size_t)s * c.nGfpts + g] = e;
      c.gf_fpt[(size_t)s * c.nGfpts + g] = f;
    }
  return c;
}

}  // namespace zfr
