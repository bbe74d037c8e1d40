// Host-side mesh: synthetic box generator, GMSH reader/writer, cell renumbering and the
// polyFvGrid build (face pairs, volumes, normals).
//
// The pair semantics follow the reference's getNeibPolyFvGrid
// (src/libfosolvers/polyGrid/polyFvGrid.f90:60-198): a cell face first matches facet
// elements (one boundary pair per matching facet), otherwise the one other cell holding the
// same nodes (one internal pair, m = lower cell); area, unit normal and centre come from the
// face node order of cell m (getINF, polyGrid.f90:211-278, geometry.f90:17-58), so the
// normal points out of m for positively oriented cells.  Matching is done here by sorting
// face node keys (O(F log F)) instead of scanning node-to-element lists.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <numeric>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/fosolvers_amd.h"
#include "common.h"

namespace {

using fos::xalloc;

int n_faces(int s) {
  switch (s) {
    case FOS_TRI: case FOS_QUAD: case FOS_TRI6: return 1;
    case FOS_TET: case FOS_TET10: return 4;
    case FOS_HEX: return 6;
    default: return 0;
  }
}
// node count of a shape (0: a shape the grid build skips)
int n_nodes(int s) {
  switch (s) {
    case FOS_TRI: return 3;
    case FOS_QUAD: return 4;
    case FOS_TRI6: return 6;
    case FOS_TET: return 4;
    case FOS_HEX: return 8;
    case FOS_TET10: return 10;
    default: return 0;
  }
}
bool is_cell(int s) { return s == FOS_TET || s == FOS_HEX || s == FOS_TET10; }
bool is_facet(int s) { return s == FOS_TRI || s == FOS_QUAD || s == FOS_TRI6; }

// local (0-based) node numbers of face f of shape s; returns count
int face_nodes(int s, int f, int *ind) {
  static const int tet[4][3] = {{0, 2, 1}, {0, 1, 3}, {0, 3, 2}, {1, 2, 3}};
  static const int hex[6][4] = {{1, 2, 6, 5}, {0, 4, 7, 3}, {2, 3, 7, 6}, {0, 1, 5, 4}, {4, 5, 6, 7}, {0, 3, 2, 1}};
  static const int tet10[4][6] = {{0, 2, 1, 6, 5, 4}, {0, 1, 3, 4, 9, 7}, {0, 3, 2, 7, 8, 6}, {1, 2, 3, 5, 8, 9}};
  switch (s) {
    case FOS_TRI: for (int k = 0; k < 3; k++) ind[k] = k; return 3;
    case FOS_QUAD: for (int k = 0; k < 4; k++) ind[k] = k; return 4;
    case FOS_TRI6: for (int k = 0; k < 6; k++) ind[k] = k; return 6;
    case FOS_TET: for (int k = 0; k < 3; k++) ind[k] = tet[f][k]; return 3;
    case FOS_HEX: for (int k = 0; k < 4; k++) ind[k] = hex[f][k]; return 4;
    case FOS_TET10: for (int k = 0; k < 6; k++) ind[k] = tet10[f][k]; return 6;
    default: return 0;
  }
}

struct V3 {
  double x, y, z;
};
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// triangle area from edges p2-p1, p3-p1 (geometry.f90:17-27)
double tri_area(V3 a, V3 b, V3 c) {
  V3 n = cross(sub(b, a), sub(c, a));
  return std::sqrt(dot(n, n)) / 2.0;
}
// unit normal (p2-p1) x (p3-p2) (geometry.f90:38-49)
V3 tri_normal(V3 a, V3 b, V3 c) {
  V3 n = cross(sub(b, a), sub(c, b));
  double l = std::sqrt(dot(n, n));
  return {n.x / l, n.y / l, n.z / l};
}
// |det(b-a, c-a, d-a)|/6 -- the same volume as the cofactor form of geometry.f90:61-82
double tet_volume(V3 a, V3 b, V3 c, V3 d) {
  return std::fabs(dot(sub(b, a), cross(sub(c, a), sub(d, a)))) / 6.0;
}
double tet_signed6(V3 a, V3 b, V3 c, V3 d) { return dot(sub(b, a), cross(sub(c, a), sub(d, a))); }

inline uint64_t splitmix(uint64_t &s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
inline double unif(uint64_t &s) { return (splitmix(s) >> 11) * (1.0 / 9007199254740992.0); }

void alloc_poly(fos_poly_grid *g, int nN, int nE, int maxNNE) {
  g->nN = nN;
  g->nE = nE;
  g->maxNNE = maxNNE;
  g->pN = xalloc<double>(3 * (size_t)nN);
  g->sE = xalloc<int>(nE);
  g->nNE = xalloc<int>(nE);
  g->iNE = xalloc<int>((size_t)maxNNE * nE);
  g->gid = xalloc<int>(nE);
  std::fill(g->iNE, g->iNE + (size_t)maxNNE * nE, -1);
}

// cell-to-cell adjacency through shared faces, cells indexed in element order among cells
void cell_adjacency(const fos_poly_grid *g, std::vector<int> &cellOf, std::vector<int> &elemOf,
                    std::vector<int> &xadj, std::vector<int> &adj) {
  cellOf.assign(g->nE, -1);
  elemOf.clear();
  for (int e = 0; e < g->nE; e++)
    if (is_cell(g->sE[e])) { cellOf[e] = (int)elemOf.size(); elemOf.push_back(e); }
  const int nc = (int)elemOf.size();
  struct FK { std::array<int, 6> k; int cell; };
  std::vector<FK> keys;
  keys.reserve((size_t)nc * 4);
  for (int c = 0; c < nc; c++) {
    int e = elemOf[c], s = g->sE[e];
    for (int f = 0; f < n_faces(s); f++) {
      int ind[6], cnt = face_nodes(s, f, ind);
      FK k;
      k.k.fill(-1);
      for (int q = 0; q < cnt; q++) k.k[q] = g->iNE[(size_t)g->maxNNE * e + ind[q]];
      std::sort(k.k.begin(), k.k.begin() + cnt);
      k.cell = c;
      keys.push_back(k);
    }
  }
  // order the records by (nodes, cell): a counting sort on the smallest node, then a sort inside each
  // node's bucket (a few dozen records), which gives the order of one global sort at a fraction of
  // its cost
  auto less = [](const FK &a, const FK &b) { return a.k != b.k ? a.k < b.k : a.cell < b.cell; };
  {
    std::vector<size_t> off((size_t)g->nN + 1, 0);
    for (const FK &f : keys) off[(size_t)f.k[0] + 1]++;
    for (int n = 0; n < g->nN; n++) off[n + 1] += off[n];
    std::vector<FK> byNode(keys.size());
    std::vector<size_t> pos(off.begin(), off.end() - 1);
    for (const FK &f : keys) byNode[pos[f.k[0]]++] = f;
    for (int n = 0; n < g->nN; n++) std::sort(byNode.begin() + off[n], byNode.begin() + off[n + 1], less);
    keys.swap(byNode);
  }
  // cells sharing a face, both directions, as rows of ascending neighbours
  xadj.assign(nc + 1, 0);
  for (size_t i = 0; i + 1 < keys.size(); i++)
    if (keys[i].k == keys[i + 1].k && keys[i].cell != keys[i + 1].cell) {
      xadj[keys[i].cell + 1]++;
      xadj[keys[i + 1].cell + 1]++;
    }
  for (int c = 0; c < nc; c++) xadj[c + 1] += xadj[c];
  adj.resize(xadj[nc]);
  std::vector<int> fill(xadj.begin(), xadj.end() - 1);
  for (size_t i = 0; i + 1 < keys.size(); i++)
    if (keys[i].k == keys[i + 1].k && keys[i].cell != keys[i + 1].cell) {
      adj[fill[keys[i].cell]++] = keys[i + 1].cell;
      adj[fill[keys[i + 1].cell]++] = keys[i].cell;
    }
  for (int c = 0; c < nc; c++) std::sort(adj.begin() + xadj[c], adj.begin() + xadj[c + 1]);
}

// reverse Cuthill-McKee over all connected components
std::vector<int> rcm_order(int n, const std::vector<int> &xadj, const std::vector<int> &adj) {
  std::vector<int> deg(n), order, level(n);
  for (int i = 0; i < n; i++) deg[i] = xadj[i + 1] - xadj[i];
  std::vector<char> seen(n, 0);
  order.reserve(n);
  auto bfs_last_level = [&](int root, std::vector<int> &lvl) {
    // returns a node of the deepest level with minimum degree; lvl = depth per node
    std::vector<int> q{root};
    std::vector<char> vis(0);
    lvl[root] = 0;
    std::vector<int> touched{root};
    std::vector<char> mark(n, 0);
    mark[root] = 1;
    size_t h = 0;
    int maxl = 0;
    while (h < q.size()) {
      int u = q[h++];
      for (int k = xadj[u]; k < xadj[u + 1]; k++) {
        int w = adj[k];
        if (!mark[w]) { mark[w] = 1; lvl[w] = lvl[u] + 1; maxl = std::max(maxl, lvl[w]); q.push_back(w); }
      }
    }
    int best = root;
    for (int u : q)
      if (lvl[u] == maxl && (best == root || deg[u] < deg[best])) best = u;
    return std::make_pair(best, maxl);
  };
  for (int start = 0; start < n; start++) {
    if (seen[start]) continue;
    // pseudo-peripheral root: repeat BFS from the far end while the depth grows
    int root = start;
    auto r = bfs_last_level(root, level);
    for (int it = 0; it < 4; it++) {
      auto r2 = bfs_last_level(r.first, level);
      if (r2.second <= r.second) break;
      root = r.first;
      r = r2;
    }
    root = r.first;
    size_t h = order.size();
    order.push_back(root);
    seen[root] = 1;
    std::vector<int> nb;
    while (h < order.size()) {
      int u = order[h++];
      nb.clear();
      for (int k = xadj[u]; k < xadj[u + 1]; k++)
        if (!seen[adj[k]]) { nb.push_back(adj[k]); seen[adj[k]] = 1; }
      std::sort(nb.begin(), nb.end(), [&](int a, int b) { return deg[a] != deg[b] ? deg[a] < deg[b] : a < b; });
      order.insert(order.end(), nb.begin(), nb.end());
    }
  }
  std::reverse(order.begin(), order.end());
  return order;
}

// recursive coordinate bisection into chunks of at most cb cells; keeps input order inside
void rcb(std::vector<int> &idx, size_t lo, size_t hi, const std::vector<V3> &ctr, int cb) {
  size_t n = hi - lo;
  if (n <= (size_t)cb) {
    std::sort(idx.begin() + lo, idx.begin() + hi);
    return;
  }
  V3 mn{1e300, 1e300, 1e300}, mx{-1e300, -1e300, -1e300};
  for (size_t i = lo; i < hi; i++) {
    V3 p = ctr[idx[i]];
    mn = {std::min(mn.x, p.x), std::min(mn.y, p.y), std::min(mn.z, p.z)};
    mx = {std::max(mx.x, p.x), std::max(mx.y, p.y), std::max(mx.z, p.z)};
  }
  double ex = mx.x - mn.x, ey = mx.y - mn.y, ez = mx.z - mn.z;
  int ax = ex >= ey && ex >= ez ? 0 : (ey >= ez ? 1 : 2);
  size_t nchunk = (n + cb - 1) / cb;
  size_t left = (size_t)cb * ((nchunk + 1) / 2);
  auto key = [&](int i) { V3 p = ctr[i]; return ax == 0 ? p.x : ax == 1 ? p.y : p.z; };
  std::nth_element(idx.begin() + lo, idx.begin() + lo + left, idx.begin() + hi, [&](int a, int b) {
    double ka = key(a), kb = key(b);
    return ka != kb ? ka < kb : a < b;
  });
  rcb(idx, lo, lo + left, ctr, cb);
  rcb(idx, lo + left, hi, ctr, cb);
}

bool parse_gmsh(std::istream &in, fos_poly_grid *out) {
  // GMSH.f90:4-157 semantics
  std::string line;
  int version = 0;
  std::vector<double> P;
  int nNodes = 0;
  struct El { int t, np, g; std::array<int, 27> c; };
  std::vector<El> els;
  auto trans = [](int t, int &np) -> int {
    switch (t) {
      case 2: np = 3; return FOS_TRI;
      case 3: np = 4; return FOS_QUAD;
      case 9: np = 6; return FOS_TRI6;
      case 4: np = 4; return FOS_TET;
      case 5: np = 8; return FOS_HEX;
      case 11: np = 10; return FOS_TET10;
      default: np = 0; return 0;
    }
  };
  bool haveEl = false;
  while (std::getline(in, line)) {
    if (line.compare(0, 11, "$MeshFormat") == 0) {
      std::getline(in, line);
      version = (int)std::floor(std::stod(line));
    } else if (line.compare(0, 6, "$Nodes") == 0 && version == 2) {
      std::getline(in, line);
      nNodes = std::stoi(line);
      P.assign(3 * (size_t)nNodes, 0.0);
      for (int i = 0; i < nNodes; i++) {
        std::getline(in, line);
        std::istringstream ss(line);
        long j;
        ss >> j >> P[3 * i] >> P[3 * i + 1] >> P[3 * i + 2];  // node ids assumed 1..n in order (GMSH.f90:46)
      }
    } else if (line.compare(0, 6, "$Nodes") == 0 && version == 4) {
      std::getline(in, line);
      std::istringstream ss(line);
      int nEnt;
      ss >> nEnt >> nNodes;
      P.assign(3 * (size_t)nNodes, 0.0);
      for (int b = 0; b < nEnt; b++) {
        std::getline(in, line);
        std::istringstream sb(line);
        int d, tag, par, cnt;
        sb >> d >> tag >> par >> cnt;
        std::vector<long> ids(cnt);
        for (int i = 0; i < cnt; i++) { std::getline(in, line); ids[i] = std::stol(line); }
        for (int i = 0; i < cnt; i++) {
          std::getline(in, line);
          std::istringstream sc(line);
          long id = ids[i] - 1;
          if (id < 0 || id >= nNodes) return false;
          sc >> P[3 * id] >> P[3 * id + 1] >> P[3 * id + 2];
        }
      }
    } else if (line.compare(0, 9, "$Elements") == 0 && version == 2) {
      std::getline(in, line);
      int m = std::stoi(line);
      els.resize(m);
      for (int i = 0; i < m; i++) {
        std::getline(in, line);
        std::istringstream ss(line);
        long j;
        int t, l;
        ss >> j >> t >> l;
        std::vector<int> tags(l);
        for (int q = 0; q < l; q++) ss >> tags[q];
        El e{};
        e.t = trans(t, e.np);
        e.g = l >= 2 ? tags[1] : 0;  // second tag: elementary entity (GMSH.f90:87-88)
        for (int q = 0; q < e.np; q++) { ss >> e.c[q]; e.c[q]--; }
        els[i] = e;
      }
      haveEl = true;
      break;
    } else if (line.compare(0, 9, "$Elements") == 0 && version == 4) {
      std::getline(in, line);
      std::istringstream ss(line);
      int nEnt, m;
      ss >> nEnt >> m;
      els.assign(m, El{});
      for (int b = 0; b < nEnt; b++) {
        std::getline(in, line);
        std::istringstream sb(line);
        int d, tag, t, cnt;
        sb >> d >> tag >> t >> cnt;
        int np;
        int s = trans(t, np);
        for (int i = 0; i < cnt; i++) {
          std::getline(in, line);
          if (!s) continue;
          std::istringstream se(line);
          long j;
          se >> j;
          if (j < 1 || j > m) return false;
          El &e = els[j - 1];
          e.t = s;
          e.np = np;
          e.g = tag;  // entity tag (GMSH.f90:67)
          for (int q = 0; q < np; q++) { se >> e.c[q]; e.c[q]--; }
        }
      }
      haveEl = true;
      break;
    }
  }
  if (!haveEl || P.empty()) {
    fos::set_error("readGMSH: nodes or elements not found");
    return false;
  }
  int nE = 0, maxNNE = 0;
  for (auto &e : els)
    if (e.np > 0) { nE++; maxNNE = std::max(maxNNE, e.np); }
  alloc_poly(out, nNodes, nE, maxNNE);
  std::memcpy(out->pN, P.data(), sizeof(double) * P.size());
  int k = 0;
  for (auto &e : els) {
    if (e.np <= 0) continue;
    out->sE[k] = e.t;
    out->nNE[k] = e.np;
    out->gid[k] = e.g;
    for (int q = 0; q < e.np; q++) out->iNE[(size_t)maxNNE * k + q] = e.c[q];
    k++;
  }
  return true;
}

}  // namespace

namespace fos {
int validate_poly_grid(const fos_poly_grid *g, const char *who) {
  if (!g) { set_error("%s: null grid", who); return -1; }
  if (g->nN < 0 || g->nE < 0 || g->maxNNE < 0) {
    set_error("%s: negative size (nN %d, nE %d, maxNNE %d)", who, g->nN, g->nE, g->maxNNE);
    return -1;
  }
  if (g->nE == 0) return 0;
  if (g->maxNNE < 1 || !g->sE || !g->nNE || !g->iNE || !g->gid || (g->nN > 0 && !g->pN)) {
    set_error("%s: element arrays missing or maxNNE %d < 1", who, g->maxNNE);
    return -1;
  }
  for (int e = 0; e < g->nE; e++) {
    const int s = g->sE[e], k = g->nNE[e], want = n_nodes(s);
    if (want ? k != want : (k < 0 || k > g->maxNNE)) {
      set_error("%s: element %d of shape %d has %d nodes (expected %d)", who, e, s, k, want ? want : g->maxNNE);
      return -1;
    }
    if (k > g->maxNNE) {
      set_error("%s: element %d has %d nodes, more than maxNNE %d", who, e, k, g->maxNNE);
      return -1;
    }
    for (int q = 0; q < k; q++) {
      const int n = g->iNE[(size_t)g->maxNNE * e + q];
      if (n < 0 || n >= g->nN) {
        set_error("%s: element %d names node %d outside [0, %d)", who, e, n, g->nN);
        return -1;
      }
    }
  }
  return 0;
}
}  // namespace fos

extern "C" {

void fos_poly_grid_free(fos_poly_grid *g) {
  if (!g) return;
  std::free(g->pN); std::free(g->sE); std::free(g->nNE); std::free(g->iNE); std::free(g->gid);
  std::memset(g, 0, sizeof(*g));
}

void fos_fv_grid_free(fos_fv_grid *g) {
  if (!g) return;
  std::free(g->pN); std::free(g->sE); std::free(g->nNE); std::free(g->iNE); std::free(g->gid);
  std::free(g->p); std::free(g->v); std::free(g->iEP); std::free(g->aP); std::free(g->normP);
  std::free(g->pP); std::free(g->neib);
  std::memset(g, 0, sizeof(*g));
}

int fos_grid_box(int nx, int ny, int nz, double lx, double ly, double lz, double jitter,
                 uint64_t seed, const int gids[7], fos_poly_grid *out) {
  if (nx < 1 || ny < 1 || nz < 1 || !out || !gids) { fos::set_error("fos_grid_box: bad arguments"); return -1; }
  if (!(jitter >= 0.0 && jitter < 0.5)) { fos::set_error("fos_grid_box: jitter must be in [0,0.5)"); return -1; }
  const long long NX = nx + 1, NY = ny + 1, NZ = nz + 1;
  const long long nN = NX * NY * NZ, nHex = (long long)nx * ny * nz;
  const long long nTri = 4LL * ((long long)ny * nz + (long long)nx * nz + (long long)nx * ny);
  if (nN > 0x7fffffffLL || 6 * nHex + nTri > 0x7fffffffLL) { fos::set_error("fos_grid_box: mesh too large"); return -1; }
  alloc_poly(out, (int)nN, (int)(6 * nHex + nTri), 4);
  const double hx = lx / nx, hy = ly / ny, hz = lz / nz;
  auto nid = [&](long long i, long long j, long long k) { return (int)(i + NX * (j + NY * k)); };
  for (long long k = 0; k < NZ; k++)
    for (long long j = 0; j < NY; j++)
      for (long long i = 0; i < NX; i++) {
        uint64_t s = seed ^ (0xD1B54A32D192ED03ull * (uint64_t)(nid(i, j, k) + 1));
        double dx = (2.0 * unif(s) - 1.0) * jitter, dy = (2.0 * unif(s) - 1.0) * jitter, dz = (2.0 * unif(s) - 1.0) * jitter;
        if (i == 0 || i == nx) dx = 0.0;
        if (j == 0 || j == ny) dy = 0.0;
        if (k == 0 || k == nz) dz = 0.0;
        double *P = out->pN + 3 * (size_t)nid(i, j, k);
        P[0] = (i + dx) * hx;
        P[1] = (j + dy) * hy;
        P[2] = (k + dz) * hz;
        if (i == nx) P[0] = lx;
        if (j == ny) P[1] = ly;
        if (k == nz) P[2] = lz;
      }
  // 6 Kuhn tets per hex: v0=(0,0,0), v1=v0+e_a, v2=v1+e_b, v3=(1,1,1) for each axis order
  static const int perms[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
  long long e = 0;
  auto P = [&](int n) { const double *q = out->pN + 3 * (size_t)n; return V3{q[0], q[1], q[2]}; };
  int bad = 0;
  for (long long k = 0; k < nz; k++)
    for (long long j = 0; j < ny; j++)
      for (long long i = 0; i < nx; i++) {
        for (int t = 0; t < 6; t++) {
          int o[3] = {0, 0, 0};
          int v[4];
          v[0] = nid(i, j, k);
          o[perms[t][0]] = 1;
          v[1] = nid(i + o[0], j + o[1], k + o[2]);
          o[perms[t][1]] = 1;
          v[2] = nid(i + o[0], j + o[1], k + o[2]);
          v[3] = nid(i + 1, j + 1, k + 1);
          if (tet_signed6(P(v[0]), P(v[1]), P(v[2]), P(v[3])) < 0.0) std::swap(v[1], v[2]);
          if (!(tet_signed6(P(v[0]), P(v[1]), P(v[2]), P(v[3])) > 0.0)) bad++;
          out->sE[e] = FOS_TET;
          out->nNE[e] = 4;
          out->gid[e] = gids[6];
          for (int q = 0; q < 4; q++) out->iNE[4 * e + q] = v[q];
          e++;
        }
      }
  // boundary triangles, diagonal from the face's (low,low) corner to its (high,high) corner
  auto quad = [&](int a, int b, int c, int d, int g) {  // a=(0,0) b=(1,0) c=(1,1) d=(0,1)
    int tri[2][3] = {{a, b, c}, {a, c, d}};
    for (int t = 0; t < 2; t++) {
      out->sE[e] = FOS_TRI;
      out->nNE[e] = 3;
      out->gid[e] = g;
      for (int q = 0; q < 3; q++) out->iNE[4 * e + q] = tri[t][q];
      e++;
    }
  };
  for (int side = 0; side < 2; side++) {
    long long i = side ? nx : 0;
    for (long long k = 0; k < nz; k++)
      for (long long j = 0; j < ny; j++)
        quad(nid(i, j, k), nid(i, j + 1, k), nid(i, j + 1, k + 1), nid(i, j, k + 1), gids[side]);
  }
  for (int side = 0; side < 2; side++) {
    long long j = side ? ny : 0;
    for (long long k = 0; k < nz; k++)
      for (long long i = 0; i < nx; i++)
        quad(nid(i, j, k), nid(i + 1, j, k), nid(i + 1, j, k + 1), nid(i, j, k + 1), gids[2 + side]);
  }
  for (int side = 0; side < 2; side++) {
    long long k = side ? nz : 0;
    for (long long j = 0; j < ny; j++)
      for (long long i = 0; i < nx; i++)
        quad(nid(i, j, k), nid(i + 1, j, k), nid(i + 1, j + 1, k), nid(i, j + 1, k), gids[4 + side]);
  }
  if (bad) {
    fos::set_error("fos_grid_box: %d tets are inverted or degenerate after jitter", bad);
    fos_poly_grid_free(out);
    return -2;
  }
  return 0;
}

int fos_grid_read_gmsh(const char *path, fos_poly_grid *out) {
  std::ifstream in(path);
  if (!in) { fos::set_error("readGMSH: cannot open %s", path); return -1; }
  std::memset(out, 0, sizeof(*out));
  if (!parse_gmsh(in, out)) return -1;
  if (fos::validate_poly_grid(out, "readGMSH")) {
    fos_poly_grid_free(out);
    return -1;
  }
  return 0;
}

int fos_grid_write_gmsh(const char *path, const fos_poly_grid *g) {
  FILE *f = std::fopen(path, "w");
  if (!f) { fos::set_error("writeGMSH: cannot open %s", path); return -1; }
  std::fprintf(f, "$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$Nodes\n%d\n", g->nN);
  for (int i = 0; i < g->nN; i++)
    std::fprintf(f, "%d %.17g %.17g %.17g\n", i + 1, g->pN[3 * i], g->pN[3 * i + 1], g->pN[3 * i + 2]);
  std::fprintf(f, "$EndNodes\n$Elements\n%d\n", g->nE);
  for (int e = 0; e < g->nE; e++) {
    int t = 0;
    switch (g->sE[e]) {
      case FOS_TRI: t = 2; break;
      case FOS_QUAD: t = 3; break;
      case FOS_TRI6: t = 9; break;
      case FOS_TET: t = 4; break;
      case FOS_HEX: t = 5; break;
      case FOS_TET10: t = 11; break;
    }
    std::fprintf(f, "%d %d 2 %d %d", e + 1, t, g->gid[e], g->gid[e]);
    for (int q = 0; q < g->nNE[e]; q++) std::fprintf(f, " %d", g->iNE[(size_t)g->maxNNE * e + q] + 1);
    std::fprintf(f, "\n");
  }
  std::fprintf(f, "$EndElements\n");
  std::fclose(f);
  return 0;
}

int fos_grid_renumber(fos_poly_grid *g, int method, int *perm) {
  if (fos::validate_poly_grid(g, "fos_grid_renumber")) return -1;
  std::vector<int> cellOf, elemOf, xadj, adj;
  if (method == FOS_RENUMBER_RCM) {
    cell_adjacency(g, cellOf, elemOf, xadj, adj);
  } else {  // the bisection needs only the cells, not their adjacency
    for (int e = 0; e < g->nE; e++)
      if (is_cell(g->sE[e])) elemOf.push_back(e);
  }
  const int nc = (int)elemOf.size();
  std::vector<int> order;
  if (method == FOS_RENUMBER_RCM) {
    order = rcm_order(nc, xadj, adj);
  } else if (method == FOS_RENUMBER_CHUNKS) {
    std::vector<V3> ctr(nc);
    for (int c = 0; c < nc; c++) {
      int e = elemOf[c];
      V3 s{0, 0, 0};
      for (int q = 0; q < g->nNE[e]; q++) {
        const double *P = g->pN + 3 * (size_t)g->iNE[(size_t)g->maxNNE * e + q];
        s.x += P[0]; s.y += P[1]; s.z += P[2];
      }
      ctr[c] = {s.x / g->nNE[e], s.y / g->nNE[e], s.z / g->nNE[e]};
    }
    order.resize(nc);
    std::iota(order.begin(), order.end(), 0);
    rcb(order, 0, nc, ctr, FOS_CHUNK_CELLS);
  } else {
    fos::set_error("fos_grid_renumber: unknown method %d", method);
    return -1;
  }
  // rewrite the cell slots (in element order) with cells in the new order
  std::vector<int> sE(nc), nNE(nc), gid(nc), iNE((size_t)nc * g->maxNNE);
  for (int c = 0; c < nc; c++) {
    int e = elemOf[order[c]];
    sE[c] = g->sE[e];
    nNE[c] = g->nNE[e];
    gid[c] = g->gid[e];
    std::memcpy(&iNE[(size_t)c * g->maxNNE], g->iNE + (size_t)g->maxNNE * e, sizeof(int) * g->maxNNE);
  }
  for (int c = 0; c < nc; c++) {
    int e = elemOf[c];
    g->sE[e] = sE[c];
    g->nNE[e] = nNE[c];
    g->gid[e] = gid[c];
    std::memcpy(g->iNE + (size_t)g->maxNNE * e, &iNE[(size_t)c * g->maxNNE], sizeof(int) * g->maxNNE);
  }
  if (perm) std::memcpy(perm, order.data(), sizeof(int) * nc);
  return 0;
}

int fos_grid_sort_pairs_reference(fos_fv_grid *g) {
  if (!g || g->nP < 0 || (g->nP > 0 && (!g->iEP || !g->aP || !g->normP || !g->pP || !g->iNE || !g->nNE))) {
    fos::set_error("fos_grid_sort_pairs_reference: incomplete grid");
    return -1;
  }
  const int nP = g->nP, M = g->maxNNE;
  // getNeib's discovery order (polyFvGrid.f90:103-183): cells ascending, their faces ascending, the
  // facets on a face by index, then the cell neighbour.  The face of m a pair crosses is the one
  // whose nodes all belong to the other element n.
  std::vector<long long> disc(nP);
  for (int i = 0; i < nP; i++) {
    const int m = g->iEP[2 * i], n = g->iEP[2 * i + 1];
    if (m < 0 || m >= g->nC || n < 0 || n >= g->nE) {
      fos::set_error("fos_grid_sort_pairs_reference: pair %d names elements (%d, %d) outside the grid", i, m, n);
      return -1;
    }
    const int *nn = g->iNE + (size_t)M * n;
    int face = -1;
    for (int f = 0; f < n_faces(g->sE[m]) && face < 0; f++) {
      int ind[6];
      const int cnt = face_nodes(g->sE[m], f, ind);
      bool all = cnt > 0;
      for (int q = 0; q < cnt && all; q++) {
        const int node = g->iNE[(size_t)M * m + ind[q]];
        all = std::find(nn, nn + g->nNE[n], node) != nn + g->nNE[n];
      }
      if (all) face = f;
    }
    if (face < 0) {
      fos::set_error("fos_grid_sort_pairs_reference: pair %d: no face of %d lies on element %d", i, m, n);
      return -1;
    }
    disc[i] = ((long long)m * 8 + face) * ((long long)g->nE + 1) + n;
  }
  std::vector<int> order(nP);
  std::iota(order.begin(), order.end(), 0);
  std::sort(order.begin(), order.end(), [&](int x, int y) { return disc[x] < disc[y]; });
  // sortPolyFvGrid (polyFvGrid.f90:201-227): key max(m, n), sorted by sort.f90's quicksort (middle
  // pivot, Hoare scans, swaps carried into the permutation).  The sort is not stable, so its tie
  // order is part of the reference's pair order: the same partition steps give the same order.
  std::vector<int> key(nP), perm(order);
  for (int i = 0; i < nP; i++) key[i] = std::max(g->iEP[2 * order[i]], g->iEP[2 * order[i] + 1]);
  std::vector<std::pair<int, int>> todo;
  if (nP > 1) todo.push_back({0, nP - 1});
  while (!todo.empty()) {  // the sub-ranges are disjoint: any processing order gives the same result
    const int first = todo.back().first, last = todo.back().second;
    todo.pop_back();
    const int pivot = key[(first + last) / 2];
    int i = first, j = last;
    for (;;) {
      while (key[i] < pivot) i++;
      while (pivot < key[j]) j--;
      if (i >= j) break;
      std::swap(key[i], key[j]);
      std::swap(perm[i], perm[j]);
      i++;
      j--;
    }
    if (first < i - 1) todo.push_back({first, i - 1});
    if (j + 1 < last) todo.push_back({j + 1, last});
  }
  std::vector<int> iEP(2 * (size_t)nP);
  std::vector<double> aP(nP), nrm(3 * (size_t)nP), pP(3 * (size_t)nP);
  for (int i = 0; i < nP; i++) {
    const int s = perm[i];
    iEP[2 * i] = g->iEP[2 * s];
    iEP[2 * i + 1] = g->iEP[2 * s + 1];
    aP[i] = g->aP[s];
    for (int d = 0; d < 3; d++) {
      nrm[3 * i + d] = g->normP[3 * (size_t)s + d];
      pP[3 * i + d] = g->pP[3 * (size_t)s + d];
    }
  }
  std::copy(iEP.begin(), iEP.end(), g->iEP);
  std::copy(aP.begin(), aP.end(), g->aP);
  std::copy(nrm.begin(), nrm.end(), g->normP);
  std::copy(pP.begin(), pP.end(), g->pP);
  return 0;
}

int fos_grid_up(const fos_poly_grid *in, fos_fv_grid *out) {
  if (!out) { fos::set_error("fos_grid_up: null output"); return -1; }
  std::memset(out, 0, sizeof(*out));
  if (fos::validate_poly_grid(in, "fos_grid_up")) return -1;
  const int nE = in->nE, M = in->maxNNE;
  // element sort: cells, facets, others, each stable (polyGrid.f90:88-136)
  std::vector<int> ord;
  ord.reserve(nE);
  for (int pass = 0; pass < 3; pass++)
    for (int e = 0; e < nE; e++) {
      int s = in->sE[e];
      bool take = pass == 0 ? is_cell(s) : pass == 1 ? is_facet(s) : !(is_cell(s) || is_facet(s));
      if (take) ord.push_back(e);
    }
  out->nN = in->nN;
  out->nE = nE;
  out->maxNNE = M;
  out->pN = xalloc<double>(3 * (size_t)in->nN);
  std::memcpy(out->pN, in->pN, sizeof(double) * 3 * in->nN);
  out->sE = xalloc<int>(nE);
  out->nNE = xalloc<int>(nE);
  out->iNE = xalloc<int>((size_t)M * nE);
  out->gid = xalloc<int>(nE);
  out->p = xalloc<double>(3 * (size_t)nE);
  int nC = 0, nF = 0, maxNF = 0;
  for (int k = 0; k < nE; k++) {
    int e = ord[k];
    out->sE[k] = in->sE[e];
    out->nNE[k] = in->nNE[e];
    out->gid[k] = in->gid[e];
    std::memcpy(out->iNE + (size_t)M * k, in->iNE + (size_t)M * e, sizeof(int) * M);
    double s[3] = {0, 0, 0};
    for (int q = 0; q < in->nNE[e]; q++)
      for (int d = 0; d < 3; d++) s[d] += in->pN[3 * (size_t)in->iNE[(size_t)M * e + q] + d];
    for (int d = 0; d < 3; d++) out->p[3 * (size_t)k + d] = s[d] / in->nNE[e];
    if (is_cell(in->sE[e])) nC++;
    else if (is_facet(in->sE[e])) nF++;
    maxNF = std::max(maxNF, n_faces(in->sE[e]));
  }
  out->nC = nC;
  out->nF = nF;
  out->maxNF = maxNF;
  auto node = [&](int e, int q) {
    const double *P = out->pN + 3 * (size_t)out->iNE[(size_t)M * e + q];
    return V3{P[0], P[1], P[2]};
  };
  out->v = xalloc<double>(nC);
  for (int c = 0; c < nC; c++) {
    int s = out->sE[c];
    if (s == FOS_TET || s == FOS_TET10) {
      out->v[c] = tet_volume(node(c, 0), node(c, 1), node(c, 2), node(c, 3));
    } else if (s == FOS_HEX) {
      static const int sub6[6][4] = {{0, 1, 2, 5}, {0, 2, 3, 5}, {0, 3, 4, 5}, {3, 4, 5, 6}, {2, 3, 5, 6}, {3, 4, 6, 7}};
      double v = 0;
      for (auto &t : sub6) v += tet_volume(node(c, t[0]), node(c, t[1]), node(c, t[2]), node(c, t[3]));
      out->v[c] = v;  // the hex split of geometry.f90:85-91
    } else {
      out->v[c] = 0.0;
    }
  }
  // face keys of every element face
  struct FK { std::array<int, 6> k; int elem; int face; };
  std::vector<FK> keys;
  keys.reserve((size_t)nC * 4 + nF);
  for (int e = 0; e < nE; e++) {
    int s = out->sE[e];
    if (!is_cell(s) && !is_facet(s)) continue;
    for (int f = 0; f < n_faces(s); f++) {
      int ind[6], cnt = face_nodes(s, f, ind);
      FK k;
      k.k.fill(-1);
      for (int q = 0; q < cnt; q++) k.k[q] = out->iNE[(size_t)M * e + ind[q]];
      std::sort(k.k.begin(), k.k.begin() + cnt);
      k.elem = e;
      k.face = f;
      keys.push_back(k);
    }
  }
  std::sort(keys.begin(), keys.end(), [](const FK &a, const FK &b) {
    if (a.k != b.k) return a.k < b.k;
    return a.elem != b.elem ? a.elem < b.elem : a.face < b.face;
  });
  out->neib = xalloc<int>((size_t)std::max(maxNF, 1) * nC);
  std::fill(out->neib, out->neib + (size_t)maxNF * nC, -1);
  struct Pair { int m, n, face; };
  std::vector<Pair> pairs;
  pairs.reserve((size_t)2 * nC + nF);
  int nPi = 0, nPb = 0;
  for (size_t a = 0; a < keys.size();) {
    size_t b = a;
    while (b < keys.size() && keys[b].k == keys[a].k) b++;
    std::vector<int> cells, facets;
    for (size_t q = a; q < b; q++) (keys[q].elem < nC ? cells : facets).push_back((int)q);
    if (!facets.empty()) {
      for (int ci : cells) {
        for (int fi : facets) {
          pairs.push_back({keys[ci].elem, keys[fi].elem, keys[ci].face});
          out->neib[(size_t)maxNF * keys[ci].elem + keys[ci].face] = keys[fi].elem;
          nPb++;
        }
      }
    } else if (cells.size() == 2) {
      const FK &A = keys[cells[0]], &B = keys[cells[1]];  // A.elem < B.elem
      pairs.push_back({A.elem, B.elem, A.face});
      out->neib[(size_t)maxNF * A.elem + A.face] = B.elem;
      out->neib[(size_t)maxNF * B.elem + B.face] = A.elem;
      nPi++;
    } else if (cells.size() > 2) {
      fos::set_error("fos_grid_up: non-manifold face shared by %zu cells", cells.size());
      fos_fv_grid_free(out);
      return -1;
    }
    a = b;
  }
  // pair order: ascending max(m,n), ties by (m, face) -- boundary pairs land at the tail
  std::stable_sort(pairs.begin(), pairs.end(), [](const Pair &x, const Pair &y) {
    int kx = std::max(x.m, x.n), ky = std::max(y.m, y.n);
    if (kx != ky) return kx < ky;
    return x.m != y.m ? x.m < y.m : x.face < y.face;
  });
  const int nP = (int)pairs.size();
  out->nP = nP;
  out->nPi = nPi;
  out->nPb = nPb;
  out->iEP = xalloc<int>(2 * (size_t)nP);
  out->aP = xalloc<double>(nP);
  out->normP = xalloc<double>(3 * (size_t)nP);
  out->pP = xalloc<double>(3 * (size_t)nP);
  for (int i = 0; i < nP; i++) {
    const Pair &pr = pairs[i];
    out->iEP[2 * i] = pr.m;
    out->iEP[2 * i + 1] = pr.n;
    int ind[6], cnt = face_nodes(out->sE[pr.m], pr.face, ind);
    V3 q[6];
    for (int k = 0; k < cnt; k++) q[k] = node(pr.m, ind[k]);
    double a = 0.0;
    V3 n{0, 0, 0}, c{0, 0, 0};
    if (cnt == 3) {
      a = tri_area(q[0], q[1], q[2]);
      n = tri_normal(q[0], q[1], q[2]);
      c = {(q[0].x + q[1].x + q[2].x) / 3.0, (q[0].y + q[1].y + q[2].y) / 3.0, (q[0].z + q[1].z + q[2].z) / 3.0};
    } else if (cnt == 4) {
      a = tri_area(q[0], q[1], q[2]) + tri_area(q[2], q[3], q[0]);
      V3 n1 = tri_normal(q[0], q[1], q[2]), n2 = tri_normal(q[2], q[3], q[0]);
      n = {n1.x + n2.x, n1.y + n2.y, n1.z + n2.z};
      double l = std::sqrt(dot(n, n));
      n = {n.x / l, n.y / l, n.z / l};
      c = {(q[0].x + q[1].x + q[2].x + q[3].x) / 4.0, (q[0].y + q[1].y + q[2].y + q[3].y) / 4.0,
           (q[0].z + q[1].z + q[2].z + q[3].z) / 4.0};
    }  // other node counts: zero area, normal and centre (polyFvGrid.f90:135-138)
    out->aP[i] = a;
    out->normP[3 * i] = n.x; out->normP[3 * i + 1] = n.y; out->normP[3 * i + 2] = n.z;
    out->pP[3 * i] = c.x; out->pP[3 * i + 1] = c.y; out->pP[3 * i + 2] = c.z;
  }
  return 0;
}

}  // extern "C"
