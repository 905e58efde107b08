// Host geometry of the wind-tunnel workloads: triangle-soup generators (icosphere, finned rocket,
// extruded NACA 00tt section, box), mesh statistics with the projected frontal area, and the flat BVH that
// the GPU voxelizer traverses.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <tuple>
#include <vector>

#include "../../include/aerolbm.h"
#include "d3q19.h"
#include "geometry.h"

namespace {

struct V3 {
  double x, y, z;
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator*(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline V3 norm(V3 a) {
  const double l = std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
  return {a.x / l, a.y / l, a.z / l};
}

struct Soup {
  std::vector<float> v;
  void tri(V3 a, V3 b, V3 c) {
    const V3 p[3] = {a, b, c};
    for (int i = 0; i < 3; ++i) {
      v.push_back((float)p[i].x);
      v.push_back((float)p[i].y);
      v.push_back((float)p[i].z);
    }
  }
  void quad(V3 a, V3 b, V3 c, V3 d) {  // a b c d counter-clockwise seen from outside
    tri(a, b, c);
    tri(a, c, d);
  }
  int32_t n() const { return (int32_t)(v.size() / 9); }
};

int emit(const Soup& s, float* out, int32_t cap, int32_t* ntri) {
  if (!ntri) return AERO_E_ARG;
  *ntri = s.n();
  if (!out) return AERO_OK;
  if (cap < s.n()) return AERO_E_ARG;
  std::memcpy(out, s.v.data(), s.v.size() * 4);
  return AERO_OK;
}

double jitter_factor(uint64_t seed, int k, double amp) {
  return 1.0 + amp * aero::noise_unit(seed, 0x524f434bull, k);  // stream "ROCK"
}

}  // namespace

extern "C" {

int aero_mesh_icosphere(int32_t level, double radius, const double centre[3], float* out, int32_t cap, int32_t* ntri) {
  if (level < 0 || level > 8 || !(radius > 0) || !centre) return AERO_E_ARG;
  const double t = (1.0 + std::sqrt(5.0)) / 2.0;
  std::vector<V3> vs = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
                        {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
  for (auto& p : vs) p = norm(p);
  std::vector<std::array<V3, 3>> faces;
  const int F[20][3] = {{0, 11, 5}, {0, 5, 1},  {0, 1, 7},   {0, 7, 10}, {0, 10, 11}, {1, 5, 9}, {5, 11, 4},
                        {11, 10, 2}, {10, 7, 6}, {7, 1, 8},  {3, 9, 4},  {3, 4, 2},   {3, 2, 6}, {3, 6, 8},
                        {3, 8, 9},  {4, 9, 5},  {2, 4, 11}, {6, 2, 10}, {8, 6, 7},   {9, 8, 1}};
  for (auto& f : F) faces.push_back({vs[f[0]], vs[f[1]], vs[f[2]]});
  for (int l = 0; l < level; ++l) {
    std::vector<std::array<V3, 3>> nf;
    nf.reserve(faces.size() * 4);
    for (auto& f : faces) {
      const V3 ab = norm(0.5 * (f[0] + f[1])), bc = norm(0.5 * (f[1] + f[2])), ca = norm(0.5 * (f[2] + f[0]));
      nf.push_back({f[0], ab, ca});
      nf.push_back({f[1], bc, ab});
      nf.push_back({f[2], ca, bc});
      nf.push_back({ab, bc, ca});
    }
    faces.swap(nf);
  }
  const V3 c{centre[0], centre[1], centre[2]};
  Soup s;
  for (auto& f : faces) s.tri(radius * f[0] + c, radius * f[1] + c, radius * f[2] + c);
  return emit(s, out, cap, ntri);
}

int aero_mesh_rocket(const double tip[3], double body_len, double diam, double nose_len, double jitter, uint64_t seed,
                     int32_t segments, float* out, int32_t cap, int32_t* ntri) {
  if (!tip || !(body_len > 0) || !(diam > 0) || !(nose_len > 0) || jitter < 0 || jitter > 0.5 || segments < 8)
    return AERO_E_ARG;
  const double Lb = body_len * jitter_factor(seed, 0, jitter);
  const double R = 0.5 * diam * jitter_factor(seed, 1, jitter);
  const double Ln = nose_len * jitter_factor(seed, 2, jitter);
  // Fin planform in units of the body diameter, each dimension jittered.
  const double D = 2.0 * R;
  const double root = 1.35 * D * jitter_factor(seed, 3, jitter);
  const double tipc = 0.65 * D * jitter_factor(seed, 4, jitter);
  const double span = 0.85 * D * jitter_factor(seed, 5, jitter);
  const double sweep = 0.60 * D * jitter_factor(seed, 6, jitter);
  const double thick = std::max(0.08 * D, 1e-3);
  const V3 T{tip[0], tip[1], tip[2]};
  const int nr = std::max(8, segments * 3 / 8);   // nose rings
  const int br = std::max(2, segments * 15 / 64);  // body rings
  const double rho = (R * R + Ln * Ln) / (2.0 * R);  // tangent-ogive radius
  auto nose_r = [&](double x) {
    const double d = Ln - x;
    const double v = rho * rho - d * d;
    return std::max(0.0, std::sqrt(std::max(0.0, v)) + R - rho);
  };
  std::vector<double> xs, rs;
  for (int i = 0; i <= nr; ++i) {
    const double s = (double)i / nr;
    const double x = Ln * (1.0 - std::cos(0.5 * M_PI * s));  // finer rings near the tip
    xs.push_back(x);
    rs.push_back(i == nr ? R : nose_r(x));
  }
  for (int i = 1; i <= br; ++i) {
    xs.push_back(Ln + Lb * i / br);
    rs.push_back(R);
  }
  auto ring = [&](int i, int j) {
    const double a = 2.0 * M_PI * j / segments;
    return T + V3{xs[i], rs[i] * std::cos(a), rs[i] * std::sin(a)};
  };
  Soup s;
  // tip fan
  for (int j = 0; j < segments; ++j) s.tri(T, ring(1, j + 1), ring(1, j));
  for (size_t i = 1; i + 1 < xs.size(); ++i)
    for (int j = 0; j < segments; ++j) s.quad(ring(i, j), ring(i, j + 1), ring(i + 1, j + 1), ring(i + 1, j));
  // flat base
  const size_t last = xs.size() - 1;
  const V3 bc = T + V3{xs[last], 0, 0};
  for (int j = 0; j < segments; ++j) s.tri(bc, ring(last, j), ring(last, j + 1));
  // four trapezoidal fins, trailing edge of the root at the base, closed prisms touching the body
  const double xe = xs[last];
  for (int k = 0; k < 4; ++k) {
    const double a = 0.5 * M_PI * k;
    const V3 er{0, std::cos(a), std::sin(a)}, et{0, -std::sin(a), std::cos(a)};
    auto P = [&](double x, double r, double t) { return T + V3{x, 0, 0} + r * er + t * et; };
    const double r0 = R, r1 = R + span;
    const double xr0 = xe - root, xr1 = xe;                 // root chord
    const double xt0 = xr0 + sweep, xt1 = xt0 + tipc;       // tip chord
    const double h = 0.5 * thick;
    const V3 A0 = P(xr0, r0, -h), B0 = P(xr1, r0, -h), C0 = P(xt1, r1, -h), D0 = P(xt0, r1, -h);
    const V3 A1 = P(xr0, r0, h), B1 = P(xr1, r0, h), C1 = P(xt1, r1, h), D1 = P(xt0, r1, h);
    s.quad(A0, D0, C0, B0);  // -t face
    s.quad(A1, B1, C1, D1);  // +t face
    s.quad(A0, B0, B1, A1);  // root
    s.quad(D0, D1, C1, C0);  // tip
    s.quad(A0, A1, D1, D0);  // leading edge
    s.quad(B0, C0, C1, B1);  // trailing edge
  }
  return emit(s, out, cap, ntri);
}

int aero_mesh_naca(double t, double chord, const double le[3], double alpha_deg, double zmin, double zmax, int32_t npts,
                   float* out, int32_t cap, int32_t* ntri) {
  if (!(t > 0 && t < 0.5) || !(chord > 0) || !le || !(zmax > zmin) || npts < 8) return AERO_E_ARG;
  // closed-trailing-edge thickness form of the symmetric 4-digit section
  auto yt = [&](double xi) {
    return 5.0 * t * chord *
           (0.2969 * std::sqrt(xi) - 0.1260 * xi - 0.3516 * xi * xi + 0.2843 * xi * xi * xi - 0.1036 * xi * xi * xi * xi);
  };
  std::vector<double> px, py;  // closed loop: TE -> upper -> LE -> lower -> TE (counter-clockwise)
  for (int i = 0; i <= npts; ++i) {
    const double xi = 0.5 * (1.0 + std::cos(M_PI * i / npts));  // 1 -> 0
    px.push_back(xi * chord);
    py.push_back(i == 0 ? 0.0 : yt(xi));
  }
  for (int i = npts - 1; i > 0; --i) {
    const double xi = 0.5 * (1.0 + std::cos(M_PI * i / npts));  // 0 -> 1
    px.push_back(xi * chord);
    py.push_back(-yt(xi));
  }
  const double a = alpha_deg * M_PI / 180.0, ca = std::cos(a), sa = std::sin(a);
  const double xp = 0.25 * chord;
  std::vector<V3> pts;
  for (size_t i = 0; i < px.size(); ++i) {
    const double dx = px[i] - xp, dy = py[i];
    pts.push_back({le[0] + xp + dx * ca + dy * sa, le[1] - dx * sa + dy * ca, 0.0});
  }
  const size_t n = pts.size();
  Soup s;
  auto at = [&](size_t i, double z) { return V3{pts[i % n].x, pts[i % n].y, z}; };
  for (size_t i = 0; i < n; ++i) s.quad(at(i, zmin), at(i + 1, zmin), at(i + 1, zmax), at(i, zmax));
  V3 cen{0, 0, 0};
  for (auto& p : pts) cen = cen + (1.0 / n) * p;
  for (size_t i = 0; i < n; ++i) {
    s.tri(V3{cen.x, cen.y, zmax}, at(i, zmax), at(i + 1, zmax));
    s.tri(V3{cen.x, cen.y, zmin}, at(i + 1, zmin), at(i, zmin));
  }
  return emit(s, out, cap, ntri);
}

int aero_mesh_box(const double lo[3], const double hi[3], float* out, int32_t cap, int32_t* ntri) {
  if (!lo || !hi || !(hi[0] > lo[0] && hi[1] > lo[1] && hi[2] > lo[2])) return AERO_E_ARG;
  auto P = [&](int i, int j, int k) { return V3{i ? hi[0] : lo[0], j ? hi[1] : lo[1], k ? hi[2] : lo[2]}; };
  Soup s;
  s.quad(P(0, 0, 0), P(0, 0, 1), P(0, 1, 1), P(0, 1, 0));  // -x
  s.quad(P(1, 0, 0), P(1, 1, 0), P(1, 1, 1), P(1, 0, 1));  // +x
  s.quad(P(0, 0, 0), P(1, 0, 0), P(1, 0, 1), P(0, 0, 1));  // -y
  s.quad(P(0, 1, 0), P(0, 1, 1), P(1, 1, 1), P(1, 1, 0));  // +y
  s.quad(P(0, 0, 0), P(0, 1, 0), P(1, 1, 0), P(1, 0, 0));  // -z
  s.quad(P(0, 0, 1), P(1, 0, 1), P(1, 1, 1), P(0, 1, 1));  // +z
  return emit(s, out, cap, ntri);
}

int aero_mesh_stats(const float* tri, int32_t ntri, const double* world, double* out) {
  if (!tri || ntri <= 0 || !out) return AERO_E_ARG;
  std::vector<double> w;
  aero::world_positions(tri, ntri, world, w);
  double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
  double vol6 = 0;
  for (int32_t t = 0; t < ntri; ++t) {
    const double* a = &w[(size_t)t * 9];
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 3; ++k) {
        mn[k] = aero::js_min(mn[k], a[j * 3 + k]);
        mx[k] = aero::js_max(mx[k], a[j * 3 + k]);
      }
    const double *p0 = a, *p1 = a + 3, *p2 = a + 6;
    vol6 += p0[0] * (p1[1] * p2[2] - p1[2] * p2[1]) + p0[1] * (p1[2] * p2[0] - p1[0] * p2[2]) +
            p0[2] * (p1[0] * p2[1] - p1[1] * p2[0]);
  }
  // open edges on vertices quantized to 1e-5 (stlLoader.ts:160-208), on the untransformed positions. The
  // reference keys vertices by the strings of Math.round(v * 1e5): equal numbers give equal keys, -0 and +0
  // included, so the key here is the bit pattern of the rounded double with -0 and NaN made canonical. The
  // reference skips the check above 100000 triangles and then reports no open edge.
  std::map<std::tuple<int64_t, int64_t, int64_t, int64_t, int64_t, int64_t>, int> edges;
  auto q = [](float v) {
    double r = aero::js_round((double)v * 1e5);
    if (r == 0) r = 0.0;
    if (r != r) r = NAN;
    int64_t b;
    std::memcpy(&b, &r, 8);
    return b;
  };
  auto key = [&](int32_t t, int j) {
    const float* p = tri + (size_t)t * 9 + j * 3;
    return std::make_tuple(q(p[0]), q(p[1]), q(p[2]));
  };
  for (int32_t t = 0; t < ntri && ntri <= 100000; ++t)
    for (int j = 0; j < 3; ++j) {
      auto a = key(t, j), b = key(t, (j + 1) % 3);
      if (a == b) continue;
      if (b < a) std::swap(a, b);
      edges[std::tuple_cat(a, b)]++;
    }
  long long open = 0;
  for (auto& e : edges)
    if (e.second != 2) ++open;
  out[0] = aero::frontal_area(w.data(), ntri, 512);
  out[1] = aero::js_max(0.0, mx[0] - mn[0]);
  out[2] = aero::js_max(0.0, mx[1] - mn[1]);
  out[3] = aero::js_max(0.0, mx[2] - mn[2]);
  out[4] = std::fabs(vol6) / 6.0;
  out[5] = (double)open;
  return AERO_OK;
}

}  // extern "C"

namespace aero {

// THREE.Vector3.applyMatrix4 on every vertex (column-major elements), in double.
void world_positions(const float* tri, int32_t ntri, const double* e, std::vector<double>& out) {
  out.resize((size_t)ntri * 9);
  for (size_t i = 0; i < (size_t)ntri * 3; ++i) {
    double x = tri[i * 3], y = tri[i * 3 + 1], z = tri[i * 3 + 2];
    if (e) {
      const double w = 1.0 / (e[3] * x + e[7] * y + e[11] * z + e[15]);
      const double X = (e[0] * x + e[4] * y + e[8] * z + e[12]) * w;
      const double Y = (e[1] * x + e[5] * y + e[9] * z + e[13]) * w;
      const double Z = (e[2] * x + e[6] * y + e[10] * z + e[14]) * w;
      x = X; y = Y; z = Z;
    }
    out[i * 3] = x; out[i * 3 + 1] = y; out[i * 3 + 2] = z;
  }
}

// Projected frontal area on the y-z plane: gridRes^2 cells of 4x4 sub-pixels over the projected bounding
// box, edge-function coverage (frontalArea.ts:22-179; the same raster so A_ref matches the reference).
double frontal_area(const double* w, int32_t ntri, int gridRes) {
  const size_t nv = (size_t)ntri * 3;
  double minY = INFINITY, maxY = -INFINITY, minZ = INFINITY, maxZ = -INFINITY;
  for (size_t i = 0; i < nv; ++i) {
    minY = std::min(minY, w[i * 3 + 1]); maxY = std::max(maxY, w[i * 3 + 1]);
    minZ = std::min(minZ, w[i * 3 + 2]); maxZ = std::max(maxZ, w[i * 3 + 2]);
  }
  const double spanY = maxY - minY, spanZ = maxZ - minZ;
  if (spanY <= 1e-9 || spanZ <= 1e-9) return 0.0;
  const int ny = gridRes, nz = gridRes;
  std::vector<uint16_t> grid((size_t)ny * nz, 0);
  const double icy = ny / spanY, icz = nz / spanZ;
  const double sub[4] = {0.125, 0.375, 0.625, 0.875};
  for (int32_t t = 0; t < ntri; ++t) {
    const double* p = w + (size_t)t * 9;
    const double z0 = (p[2] - minZ) * icz, y0 = (p[1] - minY) * icy;
    const double z1 = (p[5] - minZ) * icz, y1 = (p[4] - minY) * icy;
    const double z2 = (p[8] - minZ) * icz, y2 = (p[7] - minY) * icy;
    const double area2 = (z1 - z0) * (y2 - y0) - (y1 - y0) * (z2 - z0);
    if (std::fabs(area2) < 1e-11) continue;
    double az = z0, ay = y0, bz = z1, by = y1, cz = z2, cy = y2;
    if (area2 < 0) { bz = z2; by = y2; cz = z1; cy = y1; }
    const int zlo = std::max(0, (int)std::floor(std::min(az, std::min(bz, cz))));
    const int zhi = std::min(nz - 1, (int)std::floor(std::max(az, std::max(bz, cz))));
    const int ylo = std::max(0, (int)std::floor(std::min(ay, std::min(by, cy))));
    const int yhi = std::min(ny - 1, (int)std::floor(std::max(ay, std::max(by, cy))));
    const double e01z = bz - az, e01y = by - ay, e12z = cz - bz, e12y = cy - by, e20z = az - cz, e20y = ay - cy;
    for (int iy = ylo; iy <= yhi; ++iy)
      for (int iz = zlo; iz <= zhi; ++iz) {
        uint16_t& m = grid[(size_t)iy * nz + iz];
        if (m == 0xffff) continue;
        unsigned mask = m, bit = 1;
        for (int sy = 0; sy < 4; ++sy)
          for (int sz = 0; sz < 4; ++sz, bit <<= 1) {
            if (mask & bit) continue;
            const double py = iy + sub[sy], pz = iz + sub[sz];
            if (e01z * (py - ay) - e01y * (pz - az) >= -1e-9 && e12z * (py - by) - e12y * (pz - bz) >= -1e-9 &&
                e20z * (py - cy) - e20y * (pz - cz) >= -1e-9)
              mask |= bit;
          }
        m = (uint16_t)mask;
      }
  }
  long long covered = 0;
  for (uint16_t m : grid) covered += __builtin_popcount(m);
  return ((double)covered / ((double)ny * nz * 16)) * (spanY * spanZ);
}

// Flat BVH over world-space triangles: median split of the centroid bounds along the largest extent,
// leaves of at most 4 triangles, children of a node stored consecutively. The construction matches the
// reference's (bvh.ts:141-253) so that the GPU traversal prunes exactly the same nodes.
void build_bvh(const double* w, int32_t ntri, HostBVH& b) {
  b.ntri = ntri;
  b.tri.assign((size_t)ntri * 9, 0.0);
  b.idx.resize(ntri);
  std::vector<double> cen((size_t)ntri * 3), bb((size_t)ntri * 6);
  for (int k = 0; k < 3; ++k) { b.bmin[k] = INFINITY; b.bmax[k] = -INFINITY; }
  for (int32_t t = 0; t < ntri; ++t) {
    b.idx[t] = t;
    const double* v = w + (size_t)t * 9;
    for (int k = 0; k < 3; ++k) {
      b.tri[(size_t)t * 9 + k] = v[k];
      b.tri[(size_t)t * 9 + 3 + k] = v[3 + k] - v[k];
      b.tri[(size_t)t * 9 + 6 + k] = v[6 + k] - v[k];
      cen[(size_t)t * 3 + k] = (v[k] + v[3 + k] + v[6 + k]) / 3.0;
      const double lo = std::min(v[k], std::min(v[3 + k], v[6 + k]));
      const double hi = std::max(v[k], std::max(v[3 + k], v[6 + k]));
      bb[(size_t)t * 6 + k] = lo;
      bb[(size_t)t * 6 + 3 + k] = hi;
      if (lo < b.bmin[k]) b.bmin[k] = lo;
      if (hi > b.bmax[k]) b.bmax[k] = hi;
    }
  }
  const int maxNodes = std::max(1, ntri * 2);
  b.node.assign((size_t)maxNodes * 6, 0.0);
  b.lf.assign(maxNodes, 0);
  b.cnt.assign(maxNodes, 0);
  b.depth = 0;
  if (ntri == 0) return;
  int next = 1;
  struct Item { int node, first, count, depth; };
  std::vector<Item> work{{0, 0, ntri, 1}};
  // depth-first in the same order as a recursive build (left subtree first)
  while (!work.empty()) {
    Item it = work.back();
    work.pop_back();
    b.depth = std::max(b.depth, it.depth);
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    double clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int i = it.first; i < it.first + it.count; ++i) {
      const int t = b.idx[i];
      for (int k = 0; k < 3; ++k) {
        lo[k] = std::min(lo[k], bb[(size_t)t * 6 + k]);
        hi[k] = std::max(hi[k], bb[(size_t)t * 6 + 3 + k]);
        clo[k] = std::min(clo[k], cen[(size_t)t * 3 + k]);
        chi[k] = std::max(chi[k], cen[(size_t)t * 3 + k]);
      }
    }
    for (int k = 0; k < 3; ++k) { b.node[(size_t)it.node * 6 + k] = lo[k]; b.node[(size_t)it.node * 6 + 3 + k] = hi[k]; }
    if (it.count <= 4) { b.lf[it.node] = it.first; b.cnt[it.node] = it.count; continue; }
    const double ext[3] = {chi[0] - clo[0], chi[1] - clo[1], chi[2] - clo[2]};
    int axis = 0;
    double split = 0.5 * (clo[0] + chi[0]), best = ext[0];
    if (ext[1] > best) { axis = 1; best = ext[1]; split = 0.5 * (clo[1] + chi[1]); }
    if (ext[2] > best) { axis = 2; best = ext[2]; split = 0.5 * (clo[2] + chi[2]); }
    if (best <= 1e-12) { b.lf[it.node] = it.first; b.cnt[it.node] = it.count; continue; }
    int i = it.first, j = it.first + it.count - 1;
    while (i <= j) {
      if (cen[(size_t)b.idx[i] * 3 + axis] < split) ++i;
      else { std::swap(b.idx[i], b.idx[j]); --j; }
    }
    int left = i - it.first;
    if (left == 0 || left == it.count) left = it.count >> 1;
    const int child = next;
    next += 2;
    b.lf[it.node] = child;
    b.cnt[it.node] = 0;
    work.push_back({child + 1, it.first + left, it.count - left, it.depth + 1});
    work.push_back({child, it.first, left, it.depth + 1});
  }
  b.nnodes = next;
}

}  // namespace aero
