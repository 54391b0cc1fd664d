// Procedural scenes for tests and benchmarks (the reference loads .bsp maps;
// these replace that input with seeded axis-aligned rooms).  Triangles are
// emitted exactly as Environment.AddQuad / AddAxisAlignedRectangularSolid do
// (raytracer/environment.go:71-117): a quad (v1,v2,v3,v4) becomes (v1,v2,v3)
// with id and (v1,v3,v4) with id+1; a box is six quads in the order far,
// near, left, right, top, bottom.  Leaf patches are the faces chopped into
// equal cells no larger than `chop` (rad/patches/subdivide.go chops to the
// same bound).
#include <algorithm>
#include <cmath>

#include "vr_internal.h"

namespace vr {
namespace {

struct Rng {  // SplitMix64
  uint64_t s;
  explicit Rng(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uni() { return (double)(next() >> 11) * 0x1p-53; }            // [0,1)
  double uni(double a, double b) { return a + (b - a) * uni(); }
  int64_t irange(int64_t a, int64_t b) { return a + (int64_t)(next() % (uint64_t)(b - a + 1)); }
};

struct V3 {
  double x, y, z;
};
static V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
static V3 mul(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
static double len(V3 a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }

struct Box {
  double x0, y0, z0, x1, y1, z1;
};

class Builder {
 public:
  explicit Builder(uint64_t seed) : rng(seed) {}
  GenScene g;
  Rng rng;

  void tri(int32_t id, V3 a, V3 b, V3 c) {
    const V3 v[3] = {a, b, c};
    for (const V3 &p : v) {
      g.v9.push_back((float)p.x);
      g.v9.push_back((float)p.y);
      g.v9.push_back((float)p.z);
    }
    g.ids.push_back(id);
    g.flags.push_back(0);
  }
  // Environment.AddQuad, raytracer/environment.go:71-75
  void quad(int32_t id, V3 v1, V3 v2, V3 v3, V3 v4) {
    tri(id, v1, v2, v3);
    tri(id + 1, v1, v3, v4);
  }
  // Environment.AddAxisAlignedRectangularSolid, raytracer/environment.go:77-117
  void solid(int32_t id, V3 mn, V3 mx) {
    quad(id, {mn.x, mx.y, mx.z}, {mx.x, mx.y, mx.z}, {mx.x, mn.y, mx.z}, {mn.x, mn.y, mx.z});
    quad(id, {mn.x, mx.y, mn.z}, {mx.x, mx.y, mn.z}, {mx.x, mn.y, mn.z}, {mn.x, mn.y, mn.z});
    quad(id, {mn.x, mx.y, mx.z}, {mn.x, mx.y, mn.z}, {mn.x, mn.y, mn.z}, {mn.x, mn.y, mx.z});
    quad(id, {mx.x, mx.y, mx.z}, {mx.x, mx.y, mn.z}, {mx.x, mn.y, mn.z}, {mx.x, mn.y, mx.z});
    quad(id, {mn.x, mx.y, mx.z}, {mx.x, mx.y, mx.z}, {mx.x, mx.y, mn.z}, {mn.x, mx.y, mn.z});
    quad(id, {mn.x, mn.y, mx.z}, {mx.x, mn.y, mx.z}, {mx.x, mn.y, mn.z}, {mn.x, mn.y, mn.z});
  }
  void face(V3 P, V3 U, V3 V, V3 n) {
    const double f[12] = {P.x, P.y, P.z, U.x, U.y, U.z, V.x, V.y, V.z, n.x, n.y, n.z};
    g.faces.insert(g.faces.end(), f, f + 12);
  }
  void patch(V3 o, V3 n, double area) {
    g.po.push_back((float)o.x);
    g.po.push_back((float)o.y);
    g.po.push_back((float)o.z);
    g.pn.push_back((float)n.x);
    g.pn.push_back((float)n.y);
    g.pn.push_back((float)n.z);
    g.parea.push_back((float)area);
    for (int c = 0; c < 3; ++c) g.prefl.push_back((float)rng.uni(0.2, 0.8));
  }
  // Rectangle P + a U + b V (a, b in [0,1]) with unit normal n: tessellated
  // into quads of at most `tess`, chopped into patches of at most `chop`.
  // skip(center) drops patches whose centre is hidden.
  template <class Skip>
  void rect(int32_t id, V3 P, V3 U, V3 V, V3 n, double tess, double chop, bool patches,
            Skip skip) {
    double lu = len(U), lv = len(V);
    if (lu <= 0 || lv <= 0) return;
    int nu = std::max(1, (int)std::ceil(lu / tess - 1e-9)), nv = std::max(1, (int)std::ceil(lv / tess - 1e-9));
    for (int b = 0; b < nv; ++b)
      for (int a = 0; a < nu; ++a) {
        V3 c00 = add(P, add(mul(U, (double)a / nu), mul(V, (double)b / nv)));
        V3 c10 = add(P, add(mul(U, (double)(a + 1) / nu), mul(V, (double)b / nv)));
        V3 c11 = add(P, add(mul(U, (double)(a + 1) / nu), mul(V, (double)(b + 1) / nv)));
        V3 c01 = add(P, add(mul(U, (double)a / nu), mul(V, (double)(b + 1) / nv)));
        quad(id, c00, c10, c11, c01);
      }
    if (!patches) return;
    face(P, U, V, n);
    int mu = std::max(1, (int)std::ceil(lu / chop - 1e-9)), mv = std::max(1, (int)std::ceil(lv / chop - 1e-9));
    double area = (lu / mu) * (lv / mv);
    for (int b = 0; b < mv; ++b)
      for (int a = 0; a < mu; ++a) {
        V3 o = add(P, add(mul(U, (a + 0.5) / mu), mul(V, (b + 0.5) / mv)));
        if (!skip(o)) patch(o, n, area);
      }
  }
  void rect(int32_t id, V3 P, V3 U, V3 V, V3 n, double tess, double chop, bool patches = true) {
    rect(id, P, U, V, n, tess, chop, patches, [](V3) { return false; });
  }
};

// Patches of a rectangle whose triangles were emitted elsewhere.
struct PatchOnly {
  Builder &b;
  void rect(V3 P, V3 U, V3 V, V3 n, double chop) {
    b.face(P, U, V, n);
    double lu = len(U), lv = len(V);
    int mu = std::max(1, (int)std::ceil(lu / chop - 1e-9)), mv = std::max(1, (int)std::ceil(lv / chop - 1e-9));
    double area = (lu / mu) * (lv / mv);
    for (int y = 0; y < mv; ++y)
      for (int x = 0; x < mu; ++x) b.patch(add(P, add(mul(U, (x + 0.5) / mu), mul(V, (y + 0.5) / mv))), n, area);
  }
};

static void pillar_patches(Builder &B, const Box &b, double ceil_z, double chop) {
  B.solid(VR_TRACE_ID_OPAQUE, {b.x0, b.y0, b.z0}, {b.x1, b.y1, b.z1});
  PatchOnly P{B};
  const double h = b.z1 - b.z0, w = b.x1 - b.x0, d = b.y1 - b.y0;
  P.rect({b.x0, b.y0, b.z0}, {0, d, 0}, {0, 0, h}, {-1, 0, 0}, chop);
  P.rect({b.x1, b.y0, b.z0}, {0, d, 0}, {0, 0, h}, {1, 0, 0}, chop);
  P.rect({b.x0, b.y0, b.z0}, {w, 0, 0}, {0, 0, h}, {0, -1, 0}, chop);
  P.rect({b.x0, b.y1, b.z0}, {w, 0, 0}, {0, 0, h}, {0, 1, 0}, chop);
  if (b.z1 < ceil_z) P.rect({b.x0, b.y0, b.z1}, {w, 0, 0}, {0, d, 0}, {0, 0, 1}, chop);
}

static bool inside_xy(const std::vector<Box> &bs, double x, double y, double m) {
  for (const Box &b : bs)
    if (x > b.x0 - m && x < b.x1 + m && y > b.y0 - m && y < b.y1 + m) return true;
  return false;
}

static bool overlaps(const std::vector<Box> &bs, const Box &c, double m) {
  for (const Box &b : bs)
    if (c.x0 < b.x1 + m && c.x1 > b.x0 - m && c.y0 < b.y1 + m && c.y1 > b.y0 - m) return true;
  return false;
}

static vr_light point_light(double x, double y, double z, float intensity) {
  vr_light L;
  std::memset(&L, 0, sizeof L);
  L.type = VR_EMIT_POINT;
  L.origin[0] = (float)x;
  L.origin[1] = (float)y;
  L.origin[2] = (float)z;
  // intensity 10000 = unit intensity rescaled by c + 100 l + 100^2 q with
  // q = 1 (SetLightFalloffParams, rad/lightmap/lights.go:335-339)
  L.intensity[0] = L.intensity[1] = L.intensity[2] = intensity;
  L.quadratic_attn = 1.0f;
  L.start_fade = 0.0f;
  L.end_fade = -1.0f;
  L.cap_distance = 1.0e22f;
  return L;
}

}  // namespace

// Configuration 1: closed room 1024 x 1024 x 256 with seeded pillars.
int gen_room(uint64_t seed, int n_pillars, double chop, double tess, GenScene *out) {
  Builder B(seed);
  const double X = 1024, Y = 1024, Z = 256;
  std::vector<Box> pil;
  for (int k = 0, tries = 0; k < n_pillars && tries < 100000; ++tries) {
    double sx = (double)B.rng.irange(16, 128), sy = (double)B.rng.irange(16, 128);
    double x0 = (double)B.rng.irange(0, (int64_t)(X - sx)), y0 = (double)B.rng.irange(0, (int64_t)(Y - sy));
    Box b{x0, y0, 0, x0 + sx, y0 + sy, Z};
    if (overlaps(pil, b, 16)) continue;
    pil.push_back(b);
    ++k;
  }
  auto hidden = [&](V3 o) { return inside_xy(pil, o.x, o.y, 0); };
  const int32_t op = VR_TRACE_ID_OPAQUE;
  B.rect(op, {0, 0, 0}, {X, 0, 0}, {0, Y, 0}, {0, 0, 1}, tess, chop, true, hidden);   // floor
  B.rect(op, {0, 0, Z}, {X, 0, 0}, {0, Y, 0}, {0, 0, -1}, tess, chop, true, hidden);  // ceiling
  auto wall_hidden = [&](V3 o) { return inside_xy(pil, o.x, o.y, 0.5); };
  B.rect(op, {0, 0, 0}, {0, Y, 0}, {0, 0, Z}, {1, 0, 0}, tess, chop, true, wall_hidden);
  B.rect(op, {X, 0, 0}, {0, Y, 0}, {0, 0, Z}, {-1, 0, 0}, tess, chop, true, wall_hidden);
  B.rect(op, {0, 0, 0}, {X, 0, 0}, {0, 0, Z}, {0, 1, 0}, tess, chop, true, wall_hidden);
  B.rect(op, {0, Y, 0}, {X, 0, 0}, {0, 0, Z}, {0, -1, 0}, tess, chop, true, wall_hidden);
  for (const Box &b : pil) pillar_patches(B, b, Z, chop);
  // one point light clear of the pillars
  double lx = 512, ly = 512, lz = 128;
  for (int tries = 0; tries < 100000; ++tries) {
    lx = B.rng.uni(64, X - 64);
    ly = B.rng.uni(64, Y - 64);
    lz = B.rng.uni(64, Z - 64);
    if (!inside_xy(pil, lx, ly, 8)) break;
  }
  B.g.lights.push_back(point_light(lx, ly, lz, 10000.0f));
  B.g.grid_x = {0.0, X};
  B.g.grid_y = {0.0, Y};
  *out = std::move(B.g);
  return VR_OK;
}

// Configuration 2/3: grid of connected rooms.
int gen_level(uint64_t seed, int rx, int ry, double chop, double tess, int n_lights, int sun,
              GenScene *out) {
  Builder B(seed);
  const double T = 16, half = T / 2, DW = 128, DH = 192;
  std::vector<double> W(rx), D(ry), Xs(rx + 1, 0), Ys(ry + 1, 0);
  for (int c = 0; c < rx; ++c) W[c] = 16.0 * (double)B.rng.irange(16, 128);  // 256..2048
  for (int r = 0; r < ry; ++r) D[r] = 16.0 * (double)B.rng.irange(16, 128);
  for (int c = 0; c < rx; ++c) Xs[c + 1] = Xs[c] + W[c];
  for (int r = 0; r < ry; ++r) Ys[r + 1] = Ys[r] + D[r];
  std::vector<double> H((size_t)rx * ry);
  for (auto &h : H) h = 16.0 * (double)B.rng.irange(16, 32);  // 256..512
  // door positions: east doors (c,c+1) per row, north doors (r,r+1) per column
  std::vector<double> de((size_t)rx * ry, -1), dn((size_t)rx * ry, -1);
  for (int r = 0; r < ry; ++r)
    for (int c = 0; c + 1 < rx; ++c) {
      double y0 = Ys[r] + half + 32, y1 = Ys[r + 1] - half - 32 - DW;
      de[(size_t)r * rx + c] = std::floor(B.rng.uni(y0, std::max(y0, y1)));
    }
  for (int r = 0; r + 1 < ry; ++r)
    for (int c = 0; c < rx; ++c) {
      double x0 = Xs[c] + half + 32, x1 = Xs[c + 1] - half - 32 - DW;
      dn[(size_t)r * rx + c] = std::floor(B.rng.uni(x0, std::max(x0, x1)));
    }
  const int32_t op = VR_TRACE_ID_OPAQUE;
  std::vector<std::vector<Box>> obst((size_t)rx * ry);
  std::vector<Box> interiors((size_t)rx * ry);
  for (int r = 0; r < ry; ++r)
    for (int c = 0; c < rx; ++c) {
      const size_t k = (size_t)r * rx + c;
      const double x0 = Xs[c] + half, x1 = Xs[c + 1] - half, y0 = Ys[r] + half, y1 = Ys[r + 1] - half;
      const double h = H[k];
      interiors[k] = Box{x0, y0, 0, x1, y1, h};
      // obstacles: pillars and at most one ramp, away from the walls
      std::vector<Box> &ob = obst[k];
      int np = (int)B.rng.irange(0, 4);
      for (int p = 0, tries = 0; p < np && tries < 200; ++tries) {
        double sx = (double)B.rng.irange(16, 128), sy = (double)B.rng.irange(16, 128);
        if (x1 - x0 < sx + 160 || y1 - y0 < sy + 160) break;
        double px = std::floor(B.rng.uni(x0 + 80, x1 - 80 - sx)), py = std::floor(B.rng.uni(y0 + 80, y1 - 80 - sy));
        double top = B.rng.uni() < 0.5 ? h : 16.0 * (double)B.rng.irange(4, (int64_t)(h / 16) - 2);
        Box b{px, py, 0, px + sx, py + sy, top};
        if (overlaps(ob, b, 32)) continue;
        ob.push_back(b);
        ++p;
      }
      std::vector<Box> pil = ob;
      bool ramp = B.rng.uni() < 0.5 && (x1 - x0) > 400 && (y1 - y0) > 300;
      Box rb{0, 0, 0, 0, 0, 0};
      if (ramp) {
        double L = 16.0 * (double)B.rng.irange(8, 16), Wd = 16.0 * (double)B.rng.irange(4, 8);
        double rh = 16.0 * (double)B.rng.irange(2, 6);
        double px = std::floor(B.rng.uni(x0 + 80, x1 - 80 - L)), py = std::floor(B.rng.uni(y0 + 80, y1 - 80 - Wd));
        rb = Box{px, py, 0, px + L, py + Wd, rh};
        if (overlaps(ob, rb, 32)) ramp = false;
        else ob.push_back(rb);
      }
      bool sky = B.rng.uni() < 0.3 && (x1 - x0) > 400 && (y1 - y0) > 400;
      double sx0 = 0, sy0 = 0, ss = 192;
      if (sky) {
        sx0 = std::floor(B.rng.uni(x0 + 64, x1 - 64 - ss));
        sy0 = std::floor(B.rng.uni(y0 + 64, y1 - 64 - ss));
      }
      auto floor_hidden = [&](V3 o) { return inside_xy(ob, o.x, o.y, 0); };
      auto ceil_hidden = [&](V3 o) {
        for (const Box &b : pil)
          if (b.z1 >= h && o.x > b.x0 && o.x < b.x1 && o.y > b.y0 && o.y < b.y1) return true;
        return false;
      };
      // floor
      B.rect(op, {x0, y0, 0}, {x1 - x0, 0, 0}, {0, y1 - y0, 0}, {0, 0, 1}, tess, chop, true, floor_hidden);
      // ceiling (with a skylight hole: four rectangles around it)
      if (!sky) {
        B.rect(op, {x0, y0, h}, {x1 - x0, 0, 0}, {0, y1 - y0, 0}, {0, 0, -1}, tess, chop, true, ceil_hidden);
      } else {
        B.rect(op, {x0, y0, h}, {sx0 - x0, 0, 0}, {0, y1 - y0, 0}, {0, 0, -1}, tess, chop, true, ceil_hidden);
        B.rect(op, {sx0 + ss, y0, h}, {x1 - sx0 - ss, 0, 0}, {0, y1 - y0, 0}, {0, 0, -1}, tess, chop, true, ceil_hidden);
        B.rect(op, {sx0, y0, h}, {ss, 0, 0}, {0, sy0 - y0, 0}, {0, 0, -1}, tess, chop, true, ceil_hidden);
        B.rect(op, {sx0, sy0 + ss, h}, {ss, 0, 0}, {0, y1 - sy0 - ss, 0}, {0, 0, -1}, tess, chop, true, ceil_hidden);
        // shaft walls (16 high) and the sky quad above them
        B.rect(op, {sx0, sy0, h}, {0, ss, 0}, {0, 0, T}, {1, 0, 0}, tess, chop);
        B.rect(op, {sx0 + ss, sy0, h}, {0, ss, 0}, {0, 0, T}, {-1, 0, 0}, tess, chop);
        B.rect(op, {sx0, sy0, h}, {ss, 0, 0}, {0, 0, T}, {0, 1, 0}, tess, chop);
        B.rect(op, {sx0, sy0 + ss, h}, {ss, 0, 0}, {0, 0, T}, {0, -1, 0}, tess, chop);
        B.quad(VR_TRACE_ID_SKY, {sx0, sy0, h + T}, {sx0 + ss, sy0, h + T}, {sx0 + ss, sy0 + ss, h + T},
               {sx0, sy0 + ss, h + T});
      }
      // walls; a wall with a door is three rectangles (two sides + lintel)
      struct WallSide {
        bool along_x;  // wall runs along x (south/north) or along y (west/east)
        double fixed;  // the constant coordinate of the face
        double a0, a1; // extent along the wall
        V3 n;
        double door;   // door start along the wall, or -1
      };
      double dW = (c > 0) ? de[(size_t)r * rx + c - 1] : -1, dE = (c + 1 < rx) ? de[k] : -1;
      double dS = (r > 0) ? dn[(size_t)(r - 1) * rx + c] : -1, dN = (r + 1 < ry) ? dn[k] : -1;
      WallSide sides[4] = {{false, x0, y0, y1, {1, 0, 0}, dW},
                           {false, x1, y0, y1, {-1, 0, 0}, dE},
                           {true, y0, x0, x1, {0, 1, 0}, dS},
                           {true, y1, x0, x1, {0, -1, 0}, dN}};
      for (const WallSide &s : sides) {
        auto P = [&](double a, double z) -> V3 { return s.along_x ? V3{a, s.fixed, z} : V3{s.fixed, a, z}; };
        auto A = [&](double da) -> V3 { return s.along_x ? V3{da, 0, 0} : V3{0, da, 0}; };
        if (s.door < 0) {
          B.rect(op, P(s.a0, 0), A(s.a1 - s.a0), {0, 0, h}, s.n, tess, chop);
        } else {
          double d0 = s.door, d1 = s.door + DW;
          B.rect(op, P(s.a0, 0), A(d0 - s.a0), {0, 0, h}, s.n, tess, chop);
          B.rect(op, P(d1, 0), A(s.a1 - d1), {0, 0, h}, s.n, tess, chop);
          B.rect(op, P(d0, DH), A(DW), {0, 0, h - DH}, s.n, tess, chop);
        }
      }
      for (const Box &b : pil) pillar_patches(B, b, h, chop);
      if (ramp) {
        // wedge rising along +x from z=0 at rb.x0 to rb.z1 at rb.x1
        const double L = rb.x1 - rb.x0, Wd = rb.y1 - rb.y0, rh = rb.z1;
        V3 a{rb.x0, rb.y0, 0}, b{rb.x1, rb.y0, rh}, cc{rb.x1, rb.y1, rh}, d{rb.x0, rb.y1, 0};
        V3 U{L, 0, rh}, V{0, Wd, 0};
        double ln = std::sqrt(L * L + rh * rh);
        V3 nrm{-rh / ln, 0, L / ln};
        B.rect(op, a, U, V, nrm, tess, chop);                                            // slope
        B.rect(op, {rb.x1, rb.y0, 0}, {0, Wd, 0}, {0, 0, rh}, {1, 0, 0}, tess, chop);    // back
        B.tri(op, {rb.x0, rb.y0, 0}, {rb.x1, rb.y0, 0}, {rb.x1, rb.y0, rh});              // sides
        B.tri(op, {rb.x0, rb.y1, 0}, {rb.x1, rb.y1, 0}, {rb.x1, rb.y1, rh});
        B.patch({rb.x0 + 2 * L / 3, rb.y0, rh / 3}, {0, -1, 0}, 0.5 * L * rh);
        B.patch({rb.x0 + 2 * L / 3, rb.y1, rh / 3}, {0, 1, 0}, 0.5 * L * rh);
        (void)b; (void)cc; (void)d;
      }
    }
  // door tunnels through the 16-unit walls
  for (int r = 0; r < ry; ++r)
    for (int c = 0; c + 1 < rx; ++c) {
      double y0 = de[(size_t)r * rx + c], xw = Xs[c + 1];
      V3 P{xw - half, y0, 0};
      B.rect(op, P, {T, 0, 0}, {0, 0, DH}, {0, 1, 0}, tess, chop);                       // jamb
      B.rect(op, {xw - half, y0 + DW, 0}, {T, 0, 0}, {0, 0, DH}, {0, -1, 0}, tess, chop);
      B.rect(op, {xw - half, y0, DH}, {T, 0, 0}, {0, DW, 0}, {0, 0, -1}, tess, chop);     // lintel
      B.rect(op, P, {T, 0, 0}, {0, DW, 0}, {0, 0, 1}, tess, chop);                        // sill
    }
  for (int r = 0; r + 1 < ry; ++r)
    for (int c = 0; c < rx; ++c) {
      double x0 = dn[(size_t)r * rx + c], yw = Ys[r + 1];
      V3 P{x0, yw - half, 0};
      B.rect(op, P, {0, T, 0}, {0, 0, DH}, {1, 0, 0}, tess, chop);
      B.rect(op, {x0 + DW, yw - half, 0}, {0, T, 0}, {0, 0, DH}, {-1, 0, 0}, tess, chop);
      B.rect(op, {x0, yw - half, DH}, {DW, 0, 0}, {0, T, 0}, {0, 0, -1}, tess, chop);
      B.rect(op, P, {DW, 0, 0}, {0, T, 0}, {0, 0, 1}, tess, chop);
    }
  // point lights in distinct rooms (cycling when there are more lights)
  std::vector<size_t> order((size_t)rx * ry);
  for (size_t k = 0; k < order.size(); ++k) order[k] = k;
  for (size_t k = order.size(); k > 1; --k) std::swap(order[k - 1], order[(size_t)(B.rng.next() % k)]);
  for (int l = 0; l < n_lights; ++l) {
    size_t k = order[(size_t)l % order.size()];
    const Box &in = interiors[k];
    double lx = 0.5 * (in.x0 + in.x1), ly = 0.5 * (in.y0 + in.y1), lz = 0.5 * in.z1;
    for (int tries = 0; tries < 1000; ++tries) {
      lx = B.rng.uni(in.x0 + 32, in.x1 - 32);
      ly = B.rng.uni(in.y0 + 32, in.y1 - 32);
      lz = B.rng.uni(48, in.z1 - 32);
      bool bad = false;
      for (const Box &b : obst[k])
        if (lx > b.x0 - 16 && lx < b.x1 + 16 && ly > b.y0 - 16 && ly < b.y1 + 16 && lz < b.z1 + 16) bad = true;
      if (!bad) break;
    }
    B.g.lights.push_back(point_light(lx, ly, lz, 10000.0f));
  }
  if (sun) {
    vr_light L;
    std::memset(&L, 0, sizeof L);
    L.type = VR_EMIT_SKYLIGHT;
    double ux = B.rng.uni(-0.6, 0.6), uy = B.rng.uni(-0.6, 0.6), n = std::sqrt(ux * ux + uy * uy + 1.0);
    L.normal[0] = (float)(ux / n);
    L.normal[1] = (float)(uy / n);
    L.normal[2] = (float)(-1.0 / n);
    L.intensity[0] = 1.0f;
    L.intensity[1] = 0.95f;
    L.intensity[2] = 0.85f;
    L.end_fade = -1.0f;
    L.cap_distance = 1.0e22f;
    B.g.lights.push_back(L);
  }
  B.g.grid_x = Xs;
  B.g.grid_y = Ys;
  *out = std::move(B.g);
  return VR_OK;
}

}  // namespace vr
