// Delaunay triangulation for the synthetic NACA 0012 meshes: incremental
// Bowyer-Watson insertion with exact predicates.  Host code only (built with
// g++ into libmeshgen.so, loaded by gpueularsolver2d_amd/meshgen.py), so the
// mesh generator needs no third-party triangulator.
//
// The caller's coordinates must lie on the grid of multiples of 2^-40 (the
// generator rounds its points to it) with |x|, |y| < 32.  The predicates
// then run on the exact integers x * 2^40 (< 2^48 in magnitude, the far
// vertices included):
//   orient2d: products of 49-bit differences, exact in __int128;
//   incircle: lifted norms (< 2^99) times 2x2 minors (< 2^99), summed exactly
//             in 256-bit two's complement.
// With exact predicates the cavity of every insertion (the triangles whose
// circumcircle strictly contains the new point) is star-shaped from it, and
// the result is a Delaunay triangulation (the unique one for points with no
// four on a circle).
//
// Points are inserted in Hilbert order, each located by a walk from the last
// triangle made.  Three far vertices enclose the points; the triangles that
// use them are dropped at the end, which leaves the triangulation of the
// points' convex hull when the hull edges are Delaunay edges of the points
// and the far vertices together (the generator keeps every boundary edge's
// diametral circle empty, which makes it so).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

namespace {

using i128 = __int128;
using u128 = unsigned __int128;
using i64 = int64_t;

constexpr double kScale = 1099511627776.0;  // 2^40
constexpr double kLimit = 32.0;             // |coordinate| bound of the caller's points
// the enclosing triangle (-F, -F), (F, -F), (0, F): it holds the square
// |x|, |y| < kLimit (its half-width at y = kLimit is (F - kLimit) / 2)
constexpr double kFar = 250.0;

struct U256 {
  uint64_t w[4];  // little-endian limbs
};

U256 mul_u128(u128 a, u128 b) {
  const uint64_t a0 = (uint64_t)a, a1 = (uint64_t)(a >> 64);
  const uint64_t b0 = (uint64_t)b, b1 = (uint64_t)(b >> 64);
  const u128 p00 = (u128)a0 * b0, p01 = (u128)a0 * b1, p10 = (u128)a1 * b0, p11 = (u128)a1 * b1;
  U256 r{};
  r.w[0] = (uint64_t)p00;
  u128 mid = (p00 >> 64) + (uint64_t)p01 + (uint64_t)p10;
  r.w[1] = (uint64_t)mid;
  u128 hi = (mid >> 64) + (p01 >> 64) + (p10 >> 64) + (uint64_t)p11;
  r.w[2] = (uint64_t)hi;
  r.w[3] = (uint64_t)((hi >> 64) + (p11 >> 64));
  return r;
}

U256 neg(U256 a) {
  U256 r{};
  uint64_t carry = 1;
  for (int i = 0; i < 4; i++) {
    const u128 s = (u128)(~a.w[i]) + carry;
    r.w[i] = (uint64_t)s;
    carry = (uint64_t)(s >> 64);
  }
  return r;
}

U256 add(U256 a, U256 b) {
  U256 r{};
  uint64_t carry = 0;
  for (int i = 0; i < 4; i++) {
    const u128 s = (u128)a.w[i] + b.w[i] + carry;
    r.w[i] = (uint64_t)s;
    carry = (uint64_t)(s >> 64);
  }
  return r;
}

// a * b for signed operands whose product fits 255 bits, in two's complement
U256 mul_signed(i128 a, i128 b) {
  const bool negative = (a < 0) != (b < 0);
  const u128 ua = a < 0 ? (u128)(-a) : (u128)a, ub = b < 0 ? (u128)(-b) : (u128)b;
  const U256 m = mul_u128(ua, ub);
  return negative ? neg(m) : m;
}

int sign(U256 a) {
  if (a.w[3] >> 63) return -1;
  return (a.w[0] | a.w[1] | a.w[2] | a.w[3]) ? 1 : 0;
}

struct P {
  i64 x, y;
};

// > 0: c left of a->b (a, b, c counter-clockwise)
int orient(const P& a, const P& b, const P& c) {
  const i128 d = (i128)(b.x - a.x) * (c.y - a.y) - (i128)(b.y - a.y) * (c.x - a.x);
  return d > 0 ? 1 : (d < 0 ? -1 : 0);
}

// > 0: d strictly inside the circumcircle of the counter-clockwise a, b, c
int incircle(const P& a, const P& b, const P& c, const P& d) {
  const i128 adx = a.x - d.x, ady = a.y - d.y, bdx = b.x - d.x, bdy = b.y - d.y;
  const i128 cdx = c.x - d.x, cdy = c.y - d.y;
  const i128 al = adx * adx + ady * ady, bl = bdx * bdx + bdy * bdy, cl = cdx * cdx + cdy * cdy;
  const i128 m1 = bdx * cdy - bdy * cdx, m2 = cdx * ady - cdy * adx, m3 = adx * bdy - ady * bdx;
  return sign(add(add(mul_signed(al, m1), mul_signed(bl, m2)), mul_signed(cl, m3)));
}

struct Tri {
  int v[3];  // counter-clockwise
  int n[3];  // neighbour across the edge opposite v[i], -1 outside
  bool alive;
};

// Index of (x, y) along a Hilbert curve over the n x n grid (n a power of 2)
uint64_t hilbert(i64 n, i64 x, i64 y) {
  uint64_t d = 0;
  for (i64 s = n / 2; s > 0; s /= 2) {
    const i64 rx = (x & s) > 0, ry = (y & s) > 0;
    d += (uint64_t)s * (uint64_t)s * (uint64_t)((3 * rx) ^ ry);
    if (ry == 0) {
      if (rx == 1) {
        x = n - 1 - x;
        y = n - 1 - y;
      }
      std::swap(x, y);
    }
  }
  return d;
}

class Delaunay {
 public:
  explicit Delaunay(std::vector<P> pts) : p_(std::move(pts)) {}

  // 0 on success, 2 for a repeated point
  int run() {
    const int n = (int)p_.size();
    const i64 F = (i64)(kFar * kScale);
    p_.push_back({-F, -F});
    p_.push_back({F, -F});
    p_.push_back({0, F});
    t_.push_back(Tri{{n, n + 1, n + 2}, {-1, -1, -1}, true});
    // Hilbert order of the points
    i64 x0 = p_[0].x, x1 = x0, y0 = p_[0].y, y1 = y0;
    for (int i = 0; i < n; i++) {
      x0 = std::min(x0, p_[i].x);
      x1 = std::max(x1, p_[i].x);
      y0 = std::min(y0, p_[i].y);
      y1 = std::max(y1, p_[i].y);
    }
    const double sx = (double)std::max<i64>(x1 - x0, 1), sy = (double)std::max<i64>(y1 - y0, 1);
    std::vector<std::pair<uint64_t, int>> order(n);
    for (int i = 0; i < n; i++) {
      const i64 hx = (i64)((double)(p_[i].x - x0) / sx * 65535.0);
      const i64 hy = (i64)((double)(p_[i].y - y0) / sy * 65535.0);
      order[i] = {hilbert(65536, hx, hy), i};
    }
    std::sort(order.begin(), order.end());
    int last = 0;
    for (auto& o : order) {
      const int r = insert(o.second, last);
      if (r < 0) return 2;
      last = r;
    }
    return 0;
  }

  // the triangles that use no enclosing vertex
  void result(std::vector<int>& out) const {
    const int n = (int)p_.size() - 3;
    for (const Tri& t : t_)
      if (t.alive && t.v[0] < n && t.v[1] < n && t.v[2] < n)
        out.insert(out.end(), {t.v[0], t.v[1], t.v[2]});
  }

 private:
  // a triangle containing point q (inside or on its boundary), by a
  // visibility walk from triangle `start`
  // (every point lies inside the enclosing triangle, so a crossing never
  // leaves the triangulation); the first edge tested rotates to avoid cycles
  int locate(const P& q, int start) const {
    int t = start;
    int rot = 0;
    for (;;) {
      const Tri& T = t_[t];
      int next = -1;
      for (int k = 0; k < 3 && next < 0; k++) {
        const int i = (k + rot) % 3;
        if (T.n[i] >= 0 && orient(p_[T.v[(i + 1) % 3]], p_[T.v[(i + 2) % 3]], q) < 0) next = T.n[i];
      }
      if (next < 0) return t;
      t = next;
      rot = (rot + 1) % 3;
    }
  }

  int new_tri(int a, int b, int c) {
    Tri T{{a, b, c}, {-1, -1, -1}, true};
    if (!free_.empty()) {
      const int i = free_.back();
      free_.pop_back();
      t_[i] = T;
      return i;
    }
    t_.push_back(T);
    return (int)t_.size() - 1;
  }

  // inserts point i; returns a triangle made, or -1 if i repeats a point
  int insert(int i, int start) {
    const P& q = p_[i];
    const int t0 = locate(q, start);
    for (int k = 0; k < 3; k++) {
      const P& v = p_[t_[t0].v[k]];
      if (v.x == q.x && v.y == q.y) return -1;
    }
    // cavity: triangles whose circumcircle strictly contains q, grown from
    // the containing triangle
    cavity_.clear();
    stack_.clear();
    stack_.push_back(t0);
    mark_.resize(t_.size(), 0);
    mark_[t0] = 1;
    while (!stack_.empty()) {
      const int t = stack_.back();
      stack_.pop_back();
      cavity_.push_back(t);
      for (int k = 0; k < 3; k++) {
        const int m = t_[t].n[k];
        if (m < 0 || mark_[m]) continue;
        const Tri& M = t_[m];
        if (incircle(p_[M.v[0]], p_[M.v[1]], p_[M.v[2]], q) > 0) {
          mark_[m] = 1;
          stack_.push_back(m);
        }
      }
    }
    // the cavity's boundary edges (a -> b counter-clockwise), each with the
    // triangle outside it
    edges_.clear();
    for (int t : cavity_)
      for (int k = 0; k < 3; k++) {
        const int m = t_[t].n[k];
        if (m >= 0 && mark_[m]) continue;
        edges_.push_back({t_[t].v[(k + 1) % 3], t_[t].v[(k + 2) % 3], m});
      }
    for (int t : cavity_) {
      mark_[t] = 0;
      t_[t].alive = false;
      free_.push_back(t);
    }
    // one triangle (a, b, q) per boundary edge
    made_.clear();
    for (const Edge& e : edges_) {
      const int t = new_tri(e.a, e.b, i);
      if ((int)mark_.size() < (int)t_.size()) mark_.resize(t_.size(), 0);
      Tri& T = t_[t];
      T.n[2] = e.out;  // across (a, b)
      if (e.out >= 0) {
        Tri& O = t_[e.out];
        for (int k = 0; k < 3; k++)
          if (O.v[(k + 1) % 3] == e.b && O.v[(k + 2) % 3] == e.a) O.n[k] = t;
      }
      made_.push_back(t);
    }
    // neighbours among the new triangles: (a, b, q) and (b, c, q) share q-b
    start_of_.clear();
    for (int t : made_) start_of_.push_back({t_[t].v[0], t});
    std::sort(start_of_.begin(), start_of_.end());
    for (int t : made_) {
      const int b = t_[t].v[1];
      const auto it = std::lower_bound(start_of_.begin(), start_of_.end(), std::make_pair(b, -1));
      const int u = it->second;  // (b, c, q)
      t_[t].n[0] = u;            // across (b, q), opposite a
      t_[u].n[1] = t;            // across (q, b), opposite c
    }
    return made_.back();
  }

  struct Edge {
    int a, b, out;
  };
  std::vector<P> p_;
  std::vector<Tri> t_;
  std::vector<int> free_, cavity_, stack_, made_;
  std::vector<char> mark_;
  std::vector<Edge> edges_;
  std::vector<std::pair<int, int>> start_of_;
};

}  // namespace

extern "C" {

// Delaunay triangulation of n points (see the header comment for the grid
// the coordinates must lie on).  Writes at most max_tri counter-clockwise
// triangles (0-based point ids) to tri and their count to *ntri.
// Returns 0, 1 for coordinates off the grid or out of range, 2 for a
// repeated point, 3 when max_tri is too small.
int fvm_meshgen_delaunay(int n, const double* x, const double* y, int* tri, int max_tri, int* ntri) {
  if (n < 3 || !x || !y || !tri || !ntri) return 1;
  std::vector<P> pts(n);
  for (int i = 0; i < n; i++) {
    if (!(std::fabs(x[i]) < kLimit) || !(std::fabs(y[i]) < kLimit)) return 1;
    const double sx = x[i] * kScale, sy = y[i] * kScale;
    if (sx != std::nearbyint(sx) || sy != std::nearbyint(sy)) return 1;
    pts[i] = {(i64)sx, (i64)sy};
  }
  Delaunay d(std::move(pts));
  const int r = d.run();
  if (r) return r;
  std::vector<int> out;
  d.result(out);
  *ntri = (int)(out.size() / 3);
  if (*ntri > max_tri) return 3;
  std::copy(out.begin(), out.end(), tri);
  return 0;
}

}  // extern "C"
