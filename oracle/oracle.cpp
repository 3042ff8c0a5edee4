// CPU oracle for the nonlinear magnetostatic solve path.
//
// TEST INFRASTRUCTURE ONLY: an FP64 CPU restatement of the reference's host
// algorithms, used as the checker for the MI355X library.  Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the
// product library never links, loads or calls it.
//
// Each function names the reference file:line it restates (paths relative to
// the reference's gpu_solver/src/).  Accumulation orders follow the reference
// so that the assembled matrices agree bit-for-bit with the reference host
// assembler compiled from its own sources.

#include "oracle.h"

#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstring>
#include <limits>
#include <map>
#include <thread>
#include <utility>
#include <vector>

namespace
{

constexpr double kPi = 3.141592653589793238462643383279502884;
constexpr double kTwoPi = 6.283185307179586476925286766559005768;
constexpr double kMu0 = 4.0e-7 * kPi;

struct Curve {
  std::vector<double> b, h, slope;
  double cold = 0.0;
  int passes = 0;
};

// physics/nonlinear_model.cpp:13-113.  FEMM DC natural spline: dense
// tridiagonal-shaped system solved by Gaussian elimination with partial
// pivoting, then 3-point smoothing repeated until every Hermite segment is
// free of interior derivative roots.
bool BuildSpline(Curve *c)
{
  const int n = static_cast<int>(c->b.size());
  if (n < 2 || c->h.size() != c->b.size() || c->b[0] != 0.0 || c->h[0] != 0.0)
    return false;
  if (!(c->cold > 0.0))
    c->cold = c->h[1] / c->b[1];
  for (int i = 1; i < n; ++i)
    if (!(c->b[i] > c->b[i - 1]) || c->h[i] < 0.0 || !std::isfinite(c->b[i]) ||
        !std::isfinite(c->h[i]))
      return false;
  std::vector<double> m(static_cast<size_t>(n) * n, 0.0), r(n, 0.0);
  const double l0 = c->b[1] - c->b[0];
  m[0] = 4.0 / l0;
  m[1] = 2.0 / l0;
  r[0] = 6.0 * (c->h[1] - c->h[0]) / (l0 * l0);
  const double ln = c->b[n - 1] - c->b[n - 2];
  m[(n - 1) * n + n - 1] = 4.0 / ln;
  m[(n - 1) * n + n - 2] = 2.0 / ln;
  r[n - 1] = 6.0 * (c->h[n - 1] - c->h[n - 2]) / (ln * ln);
  for (int i = 1; i < n - 1; ++i) {
    const double lf = c->b[i] - c->b[i - 1], rt = c->b[i + 1] - c->b[i];
    m[i * n + i - 1] = 2.0 / lf;
    m[i * n + i] = 4.0 * (lf + rt) / (lf * rt);
    m[i * n + i + 1] = 2.0 / rt;
    r[i] = 6.0 * (c->h[i] - c->h[i - 1]) / (lf * lf) + 6.0 * (c->h[i + 1] - c->h[i]) / (rt * rt);
  }
  for (int col = 0; col < n; ++col) {
    int piv = col;
    for (int row = col + 1; row < n; ++row)
      if (std::abs(m[row * n + col]) > std::abs(m[piv * n + col]))
        piv = row;
    if (!(std::abs(m[piv * n + col]) > 0.0))
      return false;
    for (int k = col; k < n; ++k)
      std::swap(m[col * n + k], m[piv * n + k]);
    std::swap(r[col], r[piv]);
    for (int row = col + 1; row < n; ++row) {
      const double f = m[row * n + col] / m[col * n + col];
      for (int k = col; k < n; ++k)
        m[row * n + k] -= f * m[col * n + k];
      r[row] -= f * r[col];
    }
  }
  c->slope.assign(n, 0.0);
  for (int row = n - 1; row >= 0; --row) {
    double v = r[row];
    for (int k = row + 1; k < n; ++k)
      v -= m[row * n + k] * c->slope[k];
    c->slope[row] = v / m[row * n + row];
    if (!std::isfinite(c->slope[row]))
      return false;
  }
  bool monotone = true;
  for (int i = 1; i < n && monotone; ++i) {
    const double L = c->b[i] - c->b[i - 1];
    const double d0 = c->slope[i - 1], d1 = c->slope[i], h0 = c->h[i - 1], h1 = c->h[i];
    const double c0 = d0;
    const double c1 = -2.0 * (2.0 * d0 * L + d1 * L + 3.0 * h0 - 3.0 * h1) / (L * L);
    const double c2 = 3.0 * (d0 * L + d1 * L + 2.0 * h0 - 2.0 * h1) / (L * L * L);
    double r0 = -1.0, r1 = -1.0;
    const double disc = c1 * c1 - 4.0 * c0 * c2;
    if (c2 == 0.0) {
      if (c1 != 0.0)
        r0 = -c0 / c1;
    } else if (disc > 0.0) {
      const double s = std::sqrt(disc);
      r0 = -(c1 + s) / (2.0 * c2);
      r1 = (-c1 + s) / (2.0 * c2);
    }
    if ((r0 >= 0.0 && r0 <= L) || (r1 >= 0.0 && r1 <= L))
      monotone = false;
  }
  if (monotone)
    return true;
  if (++c->passes > 1024)
    return false;
  std::vector<double> nb = c->b, nh = c->h;
  for (int i = 1; i < n - 1; ++i) {
    nb[i] = (c->b[i - 1] + c->b[i] + c->b[i + 1]) / 3.0;
    nh[i] = (c->h[i - 1] + c->h[i] + c->h[i + 1]) / 3.0;
  }
  c->b = std::move(nb);
  c->h = std::move(nh);
  c->slope.clear();
  return BuildSpline(c);
}

struct BhEval {
  double h = 0.0, dh = 0.0, nu = 0.0, dnu = 0.0;
};

// physics/nonlinear_model.cpp:115-148.  Hermite segment found by a linear scan;
// linear extrapolation past the last knot; B = 0 gives nu = slope[0], nu' = 0.
BhEval EvalBh(const double *b, const double *h, const double *s, int n, double field)
{
  BhEval e;
  const double x = std::abs(field);
  if (n <= 0 || x == 0.0) {
    e.dh = n > 0 ? s[0] : std::numeric_limits<double>::quiet_NaN();
    e.nu = e.dh;
    return e;
  }
  if (x > b[n - 1]) {
    e.h = h[n - 1] + s[n - 1] * (x - b[n - 1]);
    e.dh = s[n - 1];
  } else {
    for (int i = 0; i < n - 1; ++i) {
      if (x >= b[i] && x <= b[i + 1]) {
        const double L = b[i + 1] - b[i];
        const double z = (x - b[i]) / L, z2 = z * z;
        e.h = (1.0 - 3.0 * z2 + 2.0 * z2 * z) * h[i] + z * (1.0 - 2.0 * z + z2) * L * s[i] +
              z2 * (3.0 - 2.0 * z) * h[i + 1] + z2 * (z - 1.0) * L * s[i + 1];
        e.dh = 6.0 * z * (z - 1.0) * h[i] / L + (1.0 - 4.0 * z + 3.0 * z2) * s[i] +
               6.0 * z * (1.0 - z) * h[i + 1] / L + z * (3.0 * z - 2.0) * s[i + 1];
        break;
      }
    }
  }
  e.nu = e.h / x;
  e.dnu = 0.5 * (e.dh / (x * x) - e.h / (x * x * x));
  return e;
}

struct Node {
  double x, y;
};

struct Model {
  int problem = FEMM_PLANAR;
  std::vector<Node> nodes;
  std::vector<std::array<int32_t, 3>> tri;
  std::vector<int32_t> tri_mat;
  std::vector<femm_material_t> mat;
  std::vector<Curve> curves;
  std::vector<int32_t> dir_nodes;
  std::vector<double> dir_values;
  std::vector<femm_node_constraint_t> cons;
  int circuits = 1;
  double depth = 0.0;
  bool curves_ok = true;
  // Native air-gap elements (AirGapElement, nonlinear_model.h:63-81).
  struct AirGap {
    bool antiperiodic = false;
    int cells = 0;
    double cx = 0.0, cy = 0.0, ri = 0.0, ro = 0.0, arc_deg = 0.0, ci = 0.0, co = 0.0;
    std::vector<femm_age_point_t> points;   // [cells + 1]
  };
  std::vector<AirGap> ages;
};

bool LoadModel(const femm_model_t *d, Model *m)
{
  if (d == nullptr || d->node_count < 0 || d->triangle_count < 0 || d->material_count < 0 ||
      d->curve_count < 0 || d->dirichlet_count < 0 || d->constraint_count < 0)
    return false;
  m->problem = d->problem_type;
  m->nodes.resize(d->node_count);
  for (int i = 0; i < d->node_count; ++i)
    m->nodes[i] = {d->node_xy[2 * i], d->node_xy[2 * i + 1]};
  m->tri.resize(d->triangle_count);
  m->tri_mat.resize(d->triangle_count);
  for (int e = 0; e < d->triangle_count; ++e) {
    for (int k = 0; k < 3; ++k)
      m->tri[e][k] = d->triangle_nodes[3 * e + k];
    m->tri_mat[e] = d->triangle_material[e];
  }
  m->mat.assign(d->materials, d->materials + d->material_count);
  m->curves.resize(d->curve_count);
  for (int c = 0; c < d->curve_count; ++c) {
    const int lo = d->curve_offsets[c], hi = d->curve_offsets[c + 1];
    m->curves[c].b.assign(d->curve_b + lo, d->curve_b + hi);
    m->curves[c].h.assign(d->curve_h + lo, d->curve_h + hi);
    if (!BuildSpline(&m->curves[c]))
      m->curves_ok = false;
  }
  m->dir_nodes.assign(d->dirichlet_nodes, d->dirichlet_nodes + d->dirichlet_count);
  m->dir_values.assign(d->dirichlet_values, d->dirichlet_values + d->dirichlet_count);
  m->cons.assign(d->constraints, d->constraints + d->constraint_count);
  m->circuits = d->circuit_count;
  m->depth = d->depth_m;
  if (d->air_gap_count < 0 || (d->air_gap_count > 0 && d->air_gaps == nullptr))
    return false;
  m->ages.resize(d->air_gap_count);
  for (int g = 0; g < d->air_gap_count; ++g) {
    const femm_air_gap_t &src = d->air_gaps[g];
    Model::AirGap &a = m->ages[g];
    a.antiperiodic = src.antiperiodic != 0;
    a.cells = src.cells;
    a.cx = src.center_xy[0];
    a.cy = src.center_xy[1];
    a.ri = src.inner_radius_m;
    a.ro = src.outer_radius_m;
    a.arc_deg = src.arc_length_deg;
    a.ci = src.inner_shift;
    a.co = src.outer_shift;
    if (src.cells >= 0 && src.points != nullptr)
      a.points.assign(src.points, src.points + src.cells + 1);
  }
  return true;
}

// signed_node_constraints.h:33-106.
struct DofMap {
  std::vector<int32_t> root;
  std::vector<int8_t> sign;
  std::vector<char> root_fixed;
  std::vector<double> root_value;
  std::vector<int32_t> root_to_free;
  std::vector<int32_t> free_roots;
};

bool BuildDofMap(const Model &m, DofMap *out)
{
  const size_t n = m.nodes.size();
  if (n == 0 || m.dir_nodes.size() != m.dir_values.size())
    return false;
  std::vector<std::vector<std::pair<int32_t, int8_t>>> adj(n);
  for (const auto &c : m.cons) {
    if (c.node_a < 0 || c.node_b < 0 || static_cast<size_t>(c.node_a) >= n ||
        static_cast<size_t>(c.node_b) >= n || (c.sign != 1 && c.sign != -1))
      return false;
    if (c.node_a == c.node_b)
      return false;
    adj[c.node_a].push_back({c.node_b, static_cast<int8_t>(c.sign)});
    adj[c.node_b].push_back({c.node_a, static_cast<int8_t>(c.sign)});
  }
  DofMap d;
  d.root.assign(n, -1);
  d.sign.assign(n, 0);
  for (size_t s = 0; s < n; ++s) {
    if (d.root[s] >= 0)
      continue;
    d.root[s] = static_cast<int32_t>(s);
    d.sign[s] = 1;
    std::vector<int32_t> stack{static_cast<int32_t>(s)};
    while (!stack.empty()) {
      const int32_t v = stack.back();
      stack.pop_back();
      for (const auto &e : adj[v]) {
        const int8_t ns = static_cast<int8_t>(d.sign[v] * e.second);
        if (d.root[e.first] < 0) {
          d.root[e.first] = static_cast<int32_t>(s);
          d.sign[e.first] = ns;
          stack.push_back(e.first);
        } else if (d.root[e.first] != static_cast<int32_t>(s) || d.sign[e.first] != ns) {
          return false;
        }
      }
    }
  }
  d.root_fixed.assign(n, 0);
  d.root_value.assign(n, 0.0);
  for (size_t i = 0; i < m.dir_nodes.size(); ++i) {
    const int32_t v = m.dir_nodes[i];
    if (v < 0 || static_cast<size_t>(v) >= n || !std::isfinite(m.dir_values[i]))
      return false;
    const int32_t r = d.root[v];
    const double rv = static_cast<double>(d.sign[v]) * m.dir_values[i];
    if (d.root_fixed[r] && d.root_value[r] != rv)
      return false;
    d.root_fixed[r] = 1;
    d.root_value[r] = rv;
  }
  d.root_to_free.assign(n, -1);
  for (size_t v = 0; v < n; ++v)
    if (d.root[v] == static_cast<int32_t>(v) && !d.root_fixed[v]) {
      d.root_to_free[v] = static_cast<int32_t>(d.free_roots.size());
      d.free_roots.push_back(static_cast<int32_t>(v));
    }
  *out = std::move(d);
  return true;
}

// physics/nonlinear_model.cpp:207-221.
struct Terms {
  double area = 0.0, bx[3] = {}, by[3] = {};
};

bool ElementTerms(const Node p[3], Terms *t)
{
  const double det = (p[1].x - p[0].x) * (p[2].y - p[0].y) - (p[2].x - p[0].x) * (p[1].y - p[0].y);
  if (!(det > 0.0) || !std::isfinite(det))
    return false;
  t->area = 0.5 * det;
  const double dx[3] = {p[1].y - p[2].y, p[2].y - p[0].y, p[0].y - p[1].y};
  const double dy[3] = {p[2].x - p[1].x, p[0].x - p[2].x, p[1].x - p[0].x};
  for (int i = 0; i < 3; ++i) {
    t->bx[i] = dy[i] / det;
    t->by[i] = -dx[i] / det;
  }
  return true;
}

// axisymmetric_p1.h:26-77.  Seven-point Dunavant rule; Br = -dN/dz,
// Bz = dN/dr + N/r.
struct AxPoint {
  double weight_area = 0.0, radius = 0.0, shape[3] = {}, br[3] = {}, bz[3] = {};
};

bool AxElement(const double r[3], const double z[3], AxPoint pts[7], double *area)
{
  const double det = (r[1] - r[0]) * (z[2] - z[0]) - (r[2] - r[0]) * (z[1] - z[0]);
  if (!(det > 0.0) || !std::isfinite(det))
    return false;
  const double dndr[3] = {(z[1] - z[2]) / det, (z[2] - z[0]) / det, (z[0] - z[1]) / det};
  const double dndz[3] = {(r[2] - r[1]) / det, (r[0] - r[2]) / det, (r[1] - r[0]) / det};
  static const double bary[7][3] = {
      {1.0 / 3.0, 1.0 / 3.0, 1.0 / 3.0},
      {0.059715871789770, 0.470142064105115, 0.470142064105115},
      {0.470142064105115, 0.059715871789770, 0.470142064105115},
      {0.470142064105115, 0.470142064105115, 0.059715871789770},
      {0.797426985353087, 0.101286507323456, 0.101286507323456},
      {0.101286507323456, 0.797426985353087, 0.101286507323456},
      {0.101286507323456, 0.101286507323456, 0.797426985353087},
  };
  static const double w[7] = {0.225,
                              0.132394152788506,
                              0.132394152788506,
                              0.132394152788506,
                              0.125939180544827,
                              0.125939180544827,
                              0.125939180544827};
  *area = 0.5 * det;
  for (int q = 0; q < 7; ++q) {
    AxPoint &o = pts[q];
    o = AxPoint{};
    o.weight_area = *area * w[q];
    for (int k = 0; k < 3; ++k) {
      o.shape[k] = bary[q][k];
      o.radius += o.shape[k] * r[k];
    }
    if (!(o.radius > 0.0) || !std::isfinite(o.radius))
      return false;
    for (int k = 0; k < 3; ++k) {
      o.br[k] = -dndz[k];
      o.bz[k] = dndr[k] + o.shape[k] / o.radius;
    }
  }
  return true;
}

// physics/nonlinear_model.cpp:223-319 (AGE checks omitted: this oracle has no AGE).
// ---------------------------------------------------------------------------------------
// Native air-gap element.  The cell matrix of BuildNativeFemmAirGapMatrix
// (native_airgap_matrix.cpp:5-22; its text is fkn/prob1big.cpp:158-213), restated from
// the element it is: a cell is the rectangle u, v in [0, 1] (arc, radius) whose potential
// is bilinear between four virtual corners; each corner value is the cubic Hermite
// interpolant of the ring nodes around it with central-difference slopes (Catmull-Rom), the
// ring's nodes sitting at u = j - 2 + shift.  The stiffness K int (dA/du)^2 + Ki int
// (dA/dv)^2 over the rectangle is taken with the mean of the 2 x 2 Gauss rule and the corner
// (trapezoid) rule.  The product library computes the same matrix in closed form.

constexpr double kAgeNu0 = 1.0 / (4.0e-7 * 3.141592653589793238462643383279502884);

// Weights of p0..p3 in the Hermite cubic on [p1, p2] at parameter t, slopes (p2 - p0) / 2
// and (p3 - p1) / 2.
void HermiteCentral(double t, double w[4])
{
  const double h00 = 2 * t * t * t - 3 * t * t + 1, h10 = t * t * t - 2 * t * t + t;
  const double h01 = -2 * t * t * t + 3 * t * t, h11 = t * t * t - t * t;
  w[0] = -0.5 * h10;
  w[1] = h00 - 0.5 * h11;
  w[2] = h01 + 0.5 * h10;
  w[3] = 0.5 * h11;
}

bool AgeMatrix(double K, double Ki, double ci, double co, double out[10][10])
{
  if (!(K > 0.0) || !(Ki > 0.0) || !std::isfinite(K) || !std::isfinite(Ki) ||
      !std::isfinite(ci) || !std::isfinite(co))
    return false;
  double T[4][10] = {};
  double wi[4], wo[4];
  HermiteCentral(1.0 - ci, wi);
  HermiteCentral(1.0 - co, wo);
  for (int q = 0; q < 4; ++q) {
    T[0][q] = wi[q];
    T[1][q + 1] = wi[q];
    T[2][q + 5] = wo[q];
    T[3][q + 6] = wo[q];
  }
  // Bilinear shape gradients of corner c = (cu, cv) at (u, v).
  auto du = [](int c, double v) { return ((c & 1) ? 1.0 : -1.0) * ((c >> 1) ? v : 1.0 - v); };
  auto dv = [](int c, double u) { return ((c >> 1) ? 1.0 : -1.0) * ((c & 1) ? u : 1.0 - u); };
  const double g = 0.5 / std::sqrt(3.0);
  const double gauss[2] = {0.5 - g, 0.5 + g};
  double S[4][4] = {};
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) {
      double gq = 0.0, tq = 0.0;
      for (int x = 0; x < 2; ++x)
        for (int y = 0; y < 2; ++y) {
          gq += 0.25 * (K * du(a, gauss[y]) * du(b, gauss[y]) + Ki * dv(a, gauss[x]) * dv(b, gauss[x]));
          tq += 0.25 * (K * du(a, y) * du(b, y) + Ki * dv(a, x) * dv(b, x));
        }
      S[a][b] = 0.5 * (gq + tq);
    }
  for (int i = 0; i < 10; ++i)
    for (int j = 0; j < 10; ++j) {
      double acc = 0.0;
      for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 4; ++b)
          acc += T[a][i] * S[a][b] * T[b][j];
      out[i][j] = acc;
    }
  return true;
}

// K, Ki and the relative shifts of nonlinear_model.cpp:573-587.
void AgeParameters(const Model::AirGap &g, double *K, double *Ki, double *ci, double *co)
{
  const double dt = (kPi / 180.0) * g.arc_deg / static_cast<double>(g.cells);
  *K = 2.0 * (g.ro - g.ri) / (dt * (g.ro + g.ri));
  *Ki = 1.0 / *K;
  *ci = g.ci;
  *co = g.co;
  if (*ci > *co) {
    *ci -= *co;
    *co = 0.0;
  } else {
    *ci = 1.0 - *co + *ci;
    *co = 1.0;
  }
}

// BuildAirGapStencil (postprocess/airgap_postprocess.cpp:10-38).
void AgeStencil(const Model::AirGap &g, int k, int32_t node[10], double weight[10])
{
  const int n = g.cells;
  const int prev = k == 0 ? n - 1 : k - 1, next = k + 1, next2 = k + 2 > n ? 1 : k + 2;
  const femm_age_point_t &P = g.points[prev], &C = g.points[k], &X = g.points[next],
                         &Y = g.points[next2];
  const int32_t nodes[10] = {P.node[0], C.node[0], C.node[1], X.node[1], Y.node[1],
                             P.node[2], C.node[2], C.node[3], X.node[3], Y.node[3]};
  const double weights[10] = {P.weight[0], C.weight[0], C.weight[1], X.weight[1], Y.weight[1],
                              P.weight[2], C.weight[2], C.weight[3], X.weight[3], Y.weight[3]};
  for (int q = 0; q < 10; ++q) {
    node[q] = nodes[q];
    weight[q] = weights[q];
  }
  if (g.antiperiodic && k == 0) {
    weight[0] = -weight[0];
    weight[5] = -weight[5];
  }
  if (g.antiperiodic && k + 1 == n) {
    weight[4] = -weight[4];
    weight[9] = -weight[9];
  }
}

int Validate(const Model &m)
{
  if (m.nodes.empty() || m.tri.empty() || m.mat.empty() ||
      (m.problem == FEMM_PLANAR && (!(m.depth > 0.0) || !std::isfinite(m.depth))) ||
      m.circuits < 0 || m.dir_nodes.empty() || m.dir_nodes.size() != m.dir_values.size())
    return FEMM_INVALID_ARGUMENT;
  std::vector<char> bnd(m.nodes.size(), 0);
  for (size_t i = 0; i < m.dir_nodes.size(); ++i) {
    const int32_t v = m.dir_nodes[i];
    if (v < 0 || static_cast<size_t>(v) >= m.nodes.size() || bnd[v] ||
        !std::isfinite(m.dir_values[i]))
      return FEMM_BOUNDARY_INVALID;
    bnd[v] = 1;
  }
  if (m.dir_nodes.size() == m.nodes.size())
    return FEMM_BOUNDARY_INVALID;
  for (const Node &p : m.nodes)
    if (!std::isfinite(p.x) || !std::isfinite(p.y) || (m.problem == FEMM_AXISYMMETRIC && p.x < 0.0))
      return FEMM_MESH_INVALID;
  if (!m.curves_ok)
    return FEMM_INVALID_MATERIAL;
  for (const femm_material_t &mt : m.mat) {
    if (!(mt.reluctivity_zero > 0.0) || mt.alpha < 0.0 || !std::isfinite(mt.reluctivity_zero) ||
        !std::isfinite(mt.alpha) || !std::isfinite(mt.source_j_per_a) || !std::isfinite(mt.hc) ||
        !std::isfinite(mt.magnetization_deg))
      return FEMM_INVALID_MATERIAL;
    if (mt.circuit < -1 || mt.circuit >= m.circuits || (mt.source_j_per_a != 0.0 && mt.circuit < 0))
      return FEMM_INVALID_MATERIAL;
    if (mt.bh_curve < -1 ||
        (mt.bh_curve >= 0 && static_cast<size_t>(mt.bh_curve) >= m.curves.size()))
      return FEMM_INVALID_MATERIAL;
    if (mt.bh_curve >= 0) {
      const Curve &c = m.curves[mt.bh_curve];
      if (c.b.size() < 2 || c.slope.size() != c.b.size() || !(c.cold > 0.0))
        return FEMM_INVALID_MATERIAL;
    }
  }
  for (size_t e = 0; e < m.tri.size(); ++e) {
    if (m.tri_mat[e] < 0 || static_cast<size_t>(m.tri_mat[e]) >= m.mat.size())
      return FEMM_INVALID_MATERIAL;
    Node p[3];
    for (int k = 0; k < 3; ++k) {
      if (m.tri[e][k] < 0 || static_cast<size_t>(m.tri[e][k]) >= m.nodes.size())
        return FEMM_MESH_INVALID;
      p[k] = m.nodes[m.tri[e][k]];
    }
    Terms t;
    if (!ElementTerms(p, &t))
      return FEMM_MESH_INVALID;
  }
  // nonlinear_model.cpp:281-297.
  for (const Model::AirGap &g : m.ages) {
    if (m.problem != FEMM_PLANAR)
      return FEMM_UNSUPPORTED_FEATURE;
    if (g.cells < 2 || g.points.size() != static_cast<size_t>(g.cells) + 1 || !(g.ri > 0.0) ||
        !(g.ro > g.ri) || !(g.arc_deg > 0.0) || !std::isfinite(g.cx) || !std::isfinite(g.cy) ||
        !std::isfinite(g.ci) || !std::isfinite(g.co))
      return FEMM_MESH_INVALID;
    for (const femm_age_point_t &pt : g.points)
      for (int q = 0; q < 4; ++q)
        if (pt.node[q] < 0 || static_cast<size_t>(pt.node[q]) >= m.nodes.size() ||
            !std::isfinite(pt.weight[q]))
          return FEMM_MESH_INVALID;
    const double dt = g.arc_deg / static_cast<double>(g.cells);
    if (!(dt > 0.0) || !std::isfinite(dt))
      return FEMM_MESH_INVALID;
  }
  DofMap d;
  if (!BuildDofMap(m, &d) || d.free_roots.empty())
    return FEMM_BOUNDARY_INVALID;
  if (m.problem == FEMM_AXISYMMETRIC) {
    double scale = 1e-3;
    for (const Node &p : m.nodes)
      scale = std::max(scale, p.x);
    const double tol = 1e-12 * scale;
    for (size_t v = 0; v < m.nodes.size(); ++v) {
      const int32_t r = d.root[v];
      if (m.nodes[v].x <= tol && (!d.root_fixed[r] || d.root_value[r] != 0.0))
        return FEMM_BOUNDARY_INVALID;
    }
  }
  return FEMM_OK;
}

// physics/nonlinear_model.cpp:428-451.
bool MaterialNu(const Model &m, const femm_material_t &mt, double bmag, bool newton, double *nu,
                double *dnu)
{
  BhEval bh;
  if (mt.bh_curve >= 0) {
    const Curve &c = m.curves[mt.bh_curve];
    bh = EvalBh(c.b.data(), c.h.data(), c.slope.data(), static_cast<int>(c.b.size()), bmag);
  }
  *nu = mt.bh_curve >= 0 ? bh.nu : mt.reluctivity_zero * (1.0 + mt.alpha * bmag * bmag);
  *dnu = newton ? (mt.bh_curve >= 0 ? bh.dnu : mt.reluctivity_zero * mt.alpha) : 0.0;
  if (!newton && mt.bh_curve >= 0)
    *nu = m.curves[mt.bh_curve].cold;
  return *nu > 0.0 && std::isfinite(*nu) && std::isfinite(*dnu);
}

using Rows = std::vector<std::map<int32_t, double>>;

bool AddEntry(Rows *rows, int32_t r, int32_t c, double v)
{
  if (r < 0 || c < 0 || static_cast<size_t>(r) >= rows->size() ||
      static_cast<size_t>(c) >= rows->size() || !std::isfinite(v))
    return false;
  double &acc = (*rows)[r][c];
  acc += v;
  return std::isfinite(acc);
}

struct Assembly {
  std::vector<int32_t> rows, cols, free_nodes, node_to_free;
  std::vector<double> values, diag, rhs_per_amp, rhs_offset, boundary;
  std::vector<int8_t> node_sign;
  // Per-ampere coil load of each circuit on every node (source_load_per_circuit,
  // physics/nonlinear_model.cpp:471-546).
  std::vector<std::vector<double>> source_per_circuit;
};

// physics/nonlinear_model.cpp:321-394.  Row sign times column sign; fixed
// columns move into rhs_offset; exact cancellation slots are kept.
int Reduce(const Model &m, const Rows &full, const std::vector<double> &full_rhs, Assembly *a)
{
  DofMap d;
  if (!BuildDofMap(m, &d) || d.free_roots.empty())
    return FEMM_BOUNDARY_INVALID;
  const size_t nf = d.free_roots.size(), nn = m.nodes.size();
  Rows red(nf);
  std::vector<double> rrhs(nf, 0.0), roff(nf, 0.0);
  a->boundary.assign(nn, 0.0);
  a->node_to_free.assign(nn, -1);
  a->node_sign = d.sign;
  for (size_t v = 0; v < nn; ++v) {
    const int32_t r = d.root[v];
    a->node_to_free[v] = d.root_to_free[r];
    if (d.root_fixed[r])
      a->boundary[v] = static_cast<double>(d.sign[v]) * d.root_value[r];
  }
  for (size_t v = 0; v < nn; ++v) {
    const int32_t row = a->node_to_free[v];
    if (row < 0)
      continue;
    const double rs = static_cast<double>(d.sign[v]);
    rrhs[row] += rs * full_rhs[v];
    for (const auto &e : full[v]) {
      const int32_t col = a->node_to_free[e.first];
      if (col < 0) {
        roff[row] -= rs * e.second * a->boundary[e.first];
      } else if (!AddEntry(&red, row, col, rs * static_cast<double>(d.sign[e.first]) * e.second)) {
        return FEMM_NUMERICAL_NONFINITE;
      }
    }
  }
  for (const auto &row : red)
    for (const auto &e : row)
      if (!std::isfinite(e.second))
        return FEMM_NUMERICAL_NONFINITE;
  a->free_nodes = d.free_roots;
  a->rows.assign(nf + 1, 0);
  a->cols.clear();
  a->values.clear();
  a->diag.assign(nf, 0.0);
  a->rhs_per_amp = std::move(rrhs);
  a->rhs_offset = std::move(roff);
  for (size_t r = 0; r < nf; ++r) {
    a->rows[r] = static_cast<int32_t>(a->values.size());
    for (const auto &e : red[r]) {
      a->cols.push_back(e.first);
      a->values.push_back(e.second);
      if (e.first == static_cast<int32_t>(r))
        a->diag[r] = e.second;
    }
    if (!(a->diag[r] > 0.0) || !std::isfinite(a->diag[r]) || !std::isfinite(a->rhs_per_amp[r]) ||
        !std::isfinite(a->rhs_offset[r]))
      return FEMM_ASSEMBLY_FAILED;
  }
  a->rows[nf] = static_cast<int32_t>(a->values.size());
  return FEMM_OK;
}

// physics/nonlinear_model.cpp:453-627 (planar and r-z; no AGE).  The Newton
// system is for the new A: matrix K + C, rhs = sources + PM + C a_old.
int Assemble(const Model &m, const std::vector<double> &a_old, const std::vector<double> &cur,
             bool newton, Assembly *out)
{
  if (a_old.size() != m.nodes.size() || cur.size() != static_cast<size_t>(m.circuits))
    return FEMM_INVALID_ARGUMENT;
  for (double c : cur)
    if (!std::isfinite(c))
      return FEMM_INVALID_ARGUMENT;
  const int v = Validate(m);
  if (v != FEMM_OK)
    return v;
  const size_t nn = m.nodes.size();
  Rows jac(nn);
  std::vector<double> rhs(nn, 0.0);
  std::vector<std::vector<double>> source(static_cast<size_t>(m.circuits),
                                          std::vector<double>(nn, 0.0));
  for (size_t e = 0; e < m.tri.size(); ++e) {
    const auto &t = m.tri[e];
    const femm_material_t &mt = m.mat[m.tri_mat[e]];
    const Node p[3] = {m.nodes[t[0]], m.nodes[t[1]], m.nodes[t[2]]};
    const double th = mt.magnetization_deg * kPi / 180.0;
    const double hcx = mt.hc * std::cos(th), hcy = mt.hc * std::sin(th);
    if (m.problem == FEMM_PLANAR) {
      Terms tm;
      if (!ElementTerms(p, &tm))
        return FEMM_MESH_INVALID;
      double bx = 0.0, by = 0.0;
      for (int i = 0; i < 3; ++i) {
        bx += a_old[t[i]] * tm.bx[i];
        by += a_old[t[i]] * tm.by[i];
      }
      double nu = 0.0, dnu = 0.0;
      if (!MaterialNu(m, mt, std::hypot(bx, by), newton, &nu, &dnu))
        return FEMM_NUMERICAL_NONFINITE;
      for (int i = 0; i < 3; ++i) {
        const double coil = tm.area * mt.source_j_per_a / 3.0;
        const double pm = tm.area * (hcx * tm.bx[i] + hcy * tm.by[i]);
        if (mt.circuit >= 0) {
          source[mt.circuit][t[i]] += coil;
          rhs[t[i]] += cur[mt.circuit] * coil;
        }
        rhs[t[i]] += pm;
        const double bi = bx * tm.bx[i] + by * tm.by[i];
        for (int j = 0; j < 3; ++j) {
          const double bj = bx * tm.bx[j] + by * tm.by[j];
          const double k = tm.area * nu * (tm.bx[i] * tm.bx[j] + tm.by[i] * tm.by[j]);
          const double c = newton ? 2.0 * tm.area * dnu * bi * bj : 0.0;
          if (!AddEntry(&jac, t[i], t[j], k + c))
            return FEMM_NUMERICAL_NONFINITE;
          if (newton)
            rhs[t[i]] += c * a_old[t[j]];
        }
      }
    } else {
      const double r[3] = {p[0].x, p[1].x, p[2].x}, z[3] = {p[0].y, p[1].y, p[2].y};
      AxPoint pts[7];
      double area = 0.0;
      if (!AxElement(r, z, pts, &area))
        return FEMM_MESH_INVALID;
      for (const AxPoint &q : pts) {
        double br = 0.0, bz = 0.0;
        for (int k = 0; k < 3; ++k) {
          br += a_old[t[k]] * q.br[k];
          bz += a_old[t[k]] * q.bz[k];
        }
        double nu = 0.0, dnu = 0.0;
        if (!MaterialNu(m, mt, std::hypot(br, bz), newton, &nu, &dnu))
          return FEMM_NUMERICAL_NONFINITE;
        const double vw = kTwoPi * q.radius * q.weight_area;
        for (int i = 0; i < 3; ++i) {
          const double coil = vw * mt.source_j_per_a * q.shape[i];
          const double pm = vw * (hcx * q.br[i] + hcy * q.bz[i]);
          if (mt.circuit >= 0) {
            source[mt.circuit][t[i]] += coil;
            rhs[t[i]] += cur[mt.circuit] * coil;
          }
          rhs[t[i]] += pm;
          const double bi = br * q.br[i] + bz * q.bz[i];
          for (int j = 0; j < 3; ++j) {
            const double bj = br * q.br[j] + bz * q.bz[j];
            const double k = vw * nu * (q.br[i] * q.br[j] + q.bz[i] * q.bz[j]);
            const double c = newton ? 2.0 * vw * dnu * bi * bj : 0.0;
            if (!AddEntry(&jac, t[i], t[j], k + c))
              return FEMM_NUMERICAL_NONFINITE;
            if (newton)
              rhs[t[i]] += c * a_old[t[j]];
          }
        }
      }
    }
  }
  // Air-gap elements after every triangle (nonlinear_model.cpp:569-622).
  for (const Model::AirGap &g : m.ages) {
    double K, Ki, ci, co, mat[10][10];
    AgeParameters(g, &K, &Ki, &ci, &co);
    if (!AgeMatrix(K, Ki, ci, co, mat))
      return FEMM_MESH_INVALID;
    for (int k = 0; k < g.cells; ++k) {
      int32_t node[10];
      double weight[10];
      AgeStencil(g, k, node, weight);
      for (int i = 0; i < 10; ++i)
        for (int j = 0; j < 10; ++j)
          if (!AddEntry(&jac, node[i], node[j], mat[i][j] * weight[i] * weight[j] * kAgeNu0))
            return FEMM_NUMERICAL_NONFINITE;
    }
  }
  for (const auto &row : jac)
    for (const auto &e : row)
      if (!std::isfinite(e.second))
        return FEMM_NUMERICAL_NONFINITE;
  const int st = Reduce(m, jac, rhs, out);
  out->source_per_circuit = std::move(source);
  return st;
}

std::vector<double> Expand(const Assembly &a, const std::vector<double> &free_solution)
{
  std::vector<double> x = a.boundary;
  for (size_t v = 0; v < x.size(); ++v) {
    const int32_t r = a.node_to_free[v];
    if (r >= 0 && static_cast<size_t>(r) < free_solution.size())
      x[v] = static_cast<double>(a.node_sign[v]) * free_solution[r];
  }
  return x;
}

struct Symbolic {
  Assembly sym;
  std::vector<int32_t> diag_slots, mat_offsets, rhs_offsets;
  std::vector<std::array<int32_t, 4>> mat_entries;
  std::vector<std::array<int32_t, 3>> rhs_entries;
  std::vector<double> pm_rhs, circuit_rhs;
  std::vector<double> age_values, age_rhs;
};

// physics/nonlinear_topology.cpp:12-268 (planar and r-z; no AGE).  Per free
// row, columns are the sorted union of triangle couplings among free roots;
// contributors are emitted in element / i / j order.
int PrepareSymbolic(const Model &m, Symbolic *s)
{
  if (Validate(m) != FEMM_OK)
    return FEMM_INVALID_ARGUMENT;
  DofMap d;
  if (!BuildDofMap(m, &d))
    return FEMM_BOUNDARY_INVALID;
  const size_t nn = m.nodes.size();
  std::vector<int32_t> n2f(nn, -1);
  std::vector<double> bval(nn, 0.0);
  for (size_t v = 0; v < nn; ++v) {
    const int32_t r = d.root[v];
    n2f[v] = d.root_to_free[r];
    if (n2f[v] < 0)
      bval[v] = static_cast<double>(d.sign[v]) * d.root_value[r];
  }
  const size_t nf = d.free_roots.size();
  std::vector<std::map<int32_t, int32_t>> st(nf);
  for (const auto &t : m.tri)
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        if (n2f[t[i]] >= 0 && n2f[t[j]] >= 0)
          st[n2f[t[i]]][n2f[t[j]]] = -1;
  for (const Model::AirGap &g : m.ages)
    for (int k = 0; k < g.cells; ++k) {
      int32_t node[10];
      double weight[10];
      AgeStencil(g, k, node, weight);
      for (int i = 0; i < 10; ++i)
        for (int j = 0; j < 10; ++j)
          if (n2f[node[i]] >= 0 && n2f[node[j]] >= 0)
            st[n2f[node[i]]][n2f[node[j]]] = -1;
    }
  Assembly &a = s->sym;
  a.free_nodes = d.free_roots;
  a.boundary = bval;
  a.node_to_free = n2f;
  a.node_sign = d.sign;
  a.rows.assign(nf + 1, 0);
  s->diag_slots.assign(nf, -1);
  for (size_t r = 0; r < nf; ++r) {
    a.rows[r] = static_cast<int32_t>(a.cols.size());
    for (auto &e : st[r]) {
      const int32_t slot = static_cast<int32_t>(a.cols.size());
      e.second = slot;
      a.cols.push_back(e.first);
      if (e.first == static_cast<int32_t>(r))
        s->diag_slots[r] = slot;
    }
    if (s->diag_slots[r] < 0)
      return FEMM_ASSEMBLY_FAILED;
  }
  a.rows[nf] = static_cast<int32_t>(a.cols.size());
  const size_t nnz = a.cols.size();
  std::vector<std::vector<std::array<int32_t, 4>>> contrib(nnz);
  std::vector<std::vector<std::array<int32_t, 3>>> rcontrib(nf);
  s->pm_rhs.assign(nf, 0.0);
  s->circuit_rhs.assign(static_cast<size_t>(m.circuits) * nf, 0.0);
  for (size_t e = 0; e < m.tri.size(); ++e) {
    const auto &t = m.tri[e];
    const femm_material_t &mt = m.mat[m.tri_mat[e]];
    const Node p[3] = {m.nodes[t[0]], m.nodes[t[1]], m.nodes[t[2]]};
    const double th = mt.magnetization_deg * kPi / 180.0;
    const double hcx = mt.hc * std::cos(th), hcy = mt.hc * std::sin(th);
    int nq = 1;
    double w[7] = {}, gx[7][3] = {}, gy[7][3] = {}, sh[7][3] = {};
    if (m.problem == FEMM_PLANAR) {
      Terms tm;
      if (!ElementTerms(p, &tm))
        return FEMM_MESH_INVALID;
      w[0] = tm.area;
      for (int k = 0; k < 3; ++k) {
        gx[0][k] = tm.bx[k];
        gy[0][k] = tm.by[k];
        sh[0][k] = 1.0 / 3.0;
      }
    } else {
      const double r[3] = {p[0].x, p[1].x, p[2].x}, z[3] = {p[0].y, p[1].y, p[2].y};
      AxPoint pts[7];
      double area = 0.0;
      if (!AxElement(r, z, pts, &area))
        return FEMM_MESH_INVALID;
      nq = 7;
      for (int q = 0; q < 7; ++q) {
        w[q] = kTwoPi * pts[q].radius * pts[q].weight_area;
        for (int k = 0; k < 3; ++k) {
          gx[q][k] = pts[q].br[k];
          gy[q][k] = pts[q].bz[k];
          sh[q][k] = pts[q].shape[k];
        }
      }
    }
    for (int i = 0; i < 3; ++i) {
      double pm = 0.0, coil = 0.0;
      for (int q = 0; q < nq; ++q) {
        pm += w[q] * (hcx * gx[q][i] + hcy * gy[q][i]);
        coil += w[q] * mt.source_j_per_a * sh[q][i];
      }
      const int32_t row = n2f[t[i]];
      if (row >= 0) {
        rcontrib[row].push_back({static_cast<int32_t>(e), i, d.sign[t[i]]});
        s->pm_rhs[row] += static_cast<double>(d.sign[t[i]]) * pm;
        if (mt.circuit >= 0)
          s->circuit_rhs[static_cast<size_t>(mt.circuit) * nf + row] +=
              static_cast<double>(d.sign[t[i]]) * coil;
      }
      for (int j = 0; j < 3; ++j)
        if (n2f[t[i]] >= 0 && n2f[t[j]] >= 0)
          contrib[st[n2f[t[i]]][n2f[t[j]]]].push_back(
              {static_cast<int32_t>(e), i, j, d.sign[t[i]] * d.sign[t[j]]});
    }
  }
  // Fixed AGE terms per slot and fixed-node loads (nonlinear_topology.cpp:180-232).
  s->age_values.assign(nnz, 0.0);
  s->age_rhs.assign(nf, 0.0);
  for (const Model::AirGap &g : m.ages) {
    double K, Ki, ci, co, mat[10][10];
    AgeParameters(g, &K, &Ki, &ci, &co);
    if (!AgeMatrix(K, Ki, ci, co, mat))
      return FEMM_MESH_INVALID;
    for (int k = 0; k < g.cells; ++k) {
      int32_t node[10];
      double weight[10];
      AgeStencil(g, k, node, weight);
      for (int i = 0; i < 10; ++i)
        for (int j = 0; j < 10; ++j) {
          const double value = mat[i][j] * weight[i] * weight[j] * kAgeNu0;
          const int32_t row = n2f[node[i]], col = n2f[node[j]];
          if (row >= 0 && col >= 0)
            s->age_values[st[row][col]] += static_cast<double>(d.sign[node[i]] * d.sign[node[j]]) * value;
          else if (row >= 0)
            s->age_rhs[row] -= static_cast<double>(d.sign[node[i]]) * value * bval[node[j]];
        }
    }
  }
  s->mat_offsets.assign(nnz + 1, 0);
  for (size_t k = 0; k < nnz; ++k) {
    s->mat_offsets[k] = static_cast<int32_t>(s->mat_entries.size());
    s->mat_entries.insert(s->mat_entries.end(), contrib[k].begin(), contrib[k].end());
  }
  s->mat_offsets[nnz] = static_cast<int32_t>(s->mat_entries.size());
  s->rhs_offsets.assign(nf + 1, 0);
  for (size_t r = 0; r < nf; ++r) {
    s->rhs_offsets[r] = static_cast<int32_t>(s->rhs_entries.size());
    s->rhs_entries.insert(s->rhs_entries.end(), rcontrib[r].begin(), rcontrib[r].end());
  }
  s->rhs_offsets[nf] = static_cast<int32_t>(s->rhs_entries.size());
  return FEMM_OK;
}

// Block-Jacobi preconditioner (PcgPreconditioner::kBlockJacobi, solvers/options.h:8-10):
// BuildBlockInverse (solvers/block_jacobi_preconditioner.cu:73-168).  Block b holds the
// rows [b bs, b bs + width), width = min(bs, n - b bs).  Its principal submatrix, scaled by
// sqrt(diag) on both sides, must have finite reciprocal diagonals >= DBL_MIN and be
// symmetric to 1e-12 (relative to max(1, |a|, |b|)); it is Cholesky-factored and each
// column j of the inverse solved by forward / back substitution, then unscaled.
// planes[j n + b bs + k] = entry k of column j; fallback[b] = 1 marks an unusable block,
// whose rows use the point Jacobi inverse instead (pcg_device.cuh:56-72).
void BuildBlockJacobi(int n, const int32_t *rp, const int32_t *ci, const double *val,
                      const double *dg, int bs, std::vector<double> *planes,
                      std::vector<char> *fallback)
{
  const int blocks = (n + bs - 1) / bs;
  planes->assign(static_cast<size_t>(bs) * n, 0.0);
  fallback->assign(blocks, 0);
  std::vector<double> f(bs * bs), scale(bs), x(bs);
  for (int b = 0; b < blocks; ++b) {
    const int first = b * bs, width = std::min(bs, n - first);
    std::fill(f.begin(), f.end(), 0.0);
    for (int k = 0; k < bs; ++k)
      scale[k] = k < width ? std::sqrt(dg[first + k]) : 1.0;
    for (int k = 0; k < width; ++k)
      for (int e = rp[first + k]; e < rp[first + k + 1]; ++e) {
        const int c = ci[e] - first;
        if (c >= 0 && c < width)
          f[k * bs + c] += val[e] / scale[k] / scale[c];
      }
    bool bad = false;
    for (int r = 0; r < width; ++r) {
      const double reciprocal = 1.0 / dg[first + r];
      if (!std::isfinite(reciprocal) || reciprocal < std::numeric_limits<double>::min())
        bad = true;
      for (int c = 0; c <= r; ++c) {
        const double u = f[r * bs + c], v = f[c * bs + r];
        if (!std::isfinite(u) || !std::isfinite(v) ||
            std::fabs(u - v) > 1e-12 * std::fmax(1.0, std::fmax(std::fabs(u), std::fabs(v))))
          bad = true;
      }
    }
    for (int c = 0; c < width && !bad; ++c) {
      double pivot = f[c * bs + c];
      for (int k = 0; k < c; ++k)
        pivot -= f[c * bs + k] * f[c * bs + k];
      if (!(pivot > 0.0) || !std::isfinite(pivot)) {
        bad = true;
        break;
      }
      f[c * bs + c] = std::sqrt(pivot);
      for (int r = c + 1; r < width; ++r) {
        double v = f[r * bs + c];
        for (int k = 0; k < c; ++k)
          v -= f[r * bs + k] * f[c * bs + k];
        f[r * bs + c] = v / f[c * bs + c];
      }
    }
    // Column j of the inverse (all columns, then the block is kept or dropped whole).
    std::vector<double> cols(static_cast<size_t>(bs) * bs, 0.0);
    for (int j = 0; j < width && !bad; ++j) {
      for (int r = 0; r < width; ++r) {
        double v = r == j ? 1.0 : 0.0;
        for (int k = 0; k < r; ++k)
          v -= f[r * bs + k] * x[k];
        x[r] = v / f[r * bs + r];
      }
      for (int r = width - 1; r >= 0; --r) {
        double v = x[r];
        for (int k = r + 1; k < width; ++k)
          v -= f[k * bs + r] * x[k];
        x[r] = v / f[r * bs + r];
      }
      for (int r = 0; r < width; ++r) {
        x[r] = x[r] / scale[r] / scale[j];
        if (!std::isfinite(x[r]))
          bad = true;
        cols[j * bs + r] = x[r];
      }
    }
    (*fallback)[b] = bad ? 1 : 0;
    if (!bad)
      for (int j = 0; j < width; ++j)
        for (int k = 0; k < width; ++k)
          (*planes)[static_cast<size_t>(j) * n + first + k] = cols[j * bs + k];
  }
}

// z = M^-1 r: point Jacobi (bs <= 1) or the block form of BuildBlockJacobi
// (PcgPreconditionRow, solvers/pcg_device.cuh:56-72; sum over the block's columns in order).
void Precondition(int n, int bs, const std::vector<double> &planes,
                  const std::vector<char> &fallback, const std::vector<double> &inv,
                  const std::vector<double> &r, std::vector<double> *z)
{
  for (int i = 0; i < n; ++i) {
    if (bs <= 1 || fallback[i / bs]) {
      (*z)[i] = r[i] * inv[i];
      continue;
    }
    const int first = i / bs * bs, width = std::min(bs, n - first);
    double v = 0.0;
    for (int j = 0; j < width; ++j)
      v += planes[static_cast<size_t>(j) * n + i] * r[first + j];
    (*z)[i] = v;
  }
}

// Block size of the PCG preconditioner for the Newton drivers below (the reference
// selects it process-wide, solvers/options.h:9 and app/cli.cpp:38-47); 1 = point Jacobi.
std::atomic<int> g_block_size{1};

// solvers/pcg_kernels.cu:10-344, restated with sequential sums.  Jacobi (or, with
// block_size > 1, block-Jacobi) preconditioner; x0 = 0; rhs scaled by max|b|;
// acceptance only on the recomputed true residual; restart otherwise.
void Pcg(int n, const int32_t *rp, const int32_t *ci, const double *val, const double *dg,
         const double *b, double tol, int maxit, double *x, double info[3], int block_size = -1)
{
  const int bs = block_size < 0 ? g_block_size.load() : std::max(block_size, 1);
  info[0] = FEMM_INTERNAL_ERROR;
  info[1] = 0;
  info[2] = std::numeric_limits<double>::infinity();
  if (n <= 0 || maxit < 0 || !(tol > 0.0) || !(tol < 1.0) || !std::isfinite(tol)) {
    info[0] = FEMM_INVALID_ARGUMENT;
    return;
  }
  for (int i = 0; i < n; ++i)
    if (!std::isfinite(b[i]) || !std::isfinite(dg[i]) || !(dg[i] > 0.0)) {
      info[0] = (!std::isfinite(b[i]) || !std::isfinite(dg[i])) ? FEMM_NUMERICAL_NONFINITE
                                                                : FEMM_LINEAR_SOLVE_BREAKDOWN;
      return;
    }
  std::vector<double> r(n), z(n), p(n), ap(n), inv(n);
  for (int i = 0; i < n; ++i) {
    x[i] = 0.0;
    inv[i] = 1.0 / dg[i];
  }
  std::vector<double> planes;
  std::vector<char> fallback;
  if (bs > 1)
    BuildBlockJacobi(n, rp, ci, val, dg, bs, &planes, &fallback);
  double scale = 0.0;
  for (int i = 0; i < n; ++i)
    scale = std::fmax(scale, std::fabs(b[i]));
  if (scale == 0.0) {
    info[0] = FEMM_OK;
    info[2] = 0.0;
    return;
  }
  auto spmv = [&](const double *v, int row) {
    double s = 0.0;
    for (int k = rp[row]; k < rp[row + 1]; ++k)
      s += val[k] * v[ci[k]];
    return s;
  };
  double rr = 0.0, rho = 0.0;
  for (int i = 0; i < n; ++i)
    r[i] = b[i] / scale;
  Precondition(n, bs, planes, fallback, inv, r, &z);
  for (int i = 0; i < n; ++i)
    p[i] = z[i];
  for (int i = 0; i < n; ++i) {
    rr += r[i] * r[i];
    rho += r[i] * z[i];
  }
  if (!std::isfinite(rr) || !std::isfinite(rho)) {
    info[0] = FEMM_NUMERICAL_NONFINITE;
    return;
  }
  info[2] = std::sqrt(rr);
  const double threshold = tol * std::sqrt(rr);
  for (int it = 0; it < maxit; ++it) {
    double pap = 0.0;
    for (int i = 0; i < n; ++i) {
      ap[i] = spmv(p.data(), i);
      pap += p[i] * ap[i];
    }
    if (!std::isfinite(pap) || !std::isfinite(rho)) {
      info[0] = FEMM_NUMERICAL_NONFINITE;
      return;
    }
    if (!(pap > 0.0) || !(rho > 0.0)) {
      info[0] = FEMM_LINEAR_SOLVE_BREAKDOWN;
      return;
    }
    const double alpha = rho / pap;
    double rsq = 0.0, rz = 0.0;
    for (int i = 0; i < n; ++i) {
      x[i] += alpha * p[i];
      r[i] -= alpha * ap[i];
      rsq += r[i] * r[i];
    }
    Precondition(n, bs, planes, fallback, inv, r, &z);
    for (int i = 0; i < n; ++i)
      rz += r[i] * z[i];
    info[1] = it + 1;
    if (!std::isfinite(rsq)) {
      info[0] = FEMM_NUMERICAL_NONFINITE;
      return;
    }
    info[2] = std::sqrt(rsq);
    if (info[2] <= threshold) {
      double tr = 0.0;
      for (int i = 0; i < n; ++i) {
        r[i] = b[i] / scale - spmv(x, i);
        tr += r[i] * r[i];
      }
      if (!std::isfinite(tr)) {
        info[0] = FEMM_NUMERICAL_NONFINITE;
        return;
      }
      info[2] = std::sqrt(tr);
      if (info[2] <= threshold) {
        for (int i = 0; i < n; ++i) {
          x[i] *= scale;
          if (!std::isfinite(x[i])) {
            info[0] = FEMM_NUMERICAL_NONFINITE;
            return;
          }
        }
        info[2] *= scale;
        info[0] = std::isfinite(info[2]) ? FEMM_OK : FEMM_NUMERICAL_NONFINITE;
        return;
      }
      double restart = 0.0;
      Precondition(n, bs, planes, fallback, inv, r, &z);
      for (int i = 0; i < n; ++i) {
        p[i] = z[i];
        restart += r[i] * z[i];
      }
      if (!std::isfinite(restart) || !(restart > 0.0)) {
        info[0] = FEMM_LINEAR_SOLVE_BREAKDOWN;
        return;
      }
      rho = restart;
      continue;
    }
    if (!std::isfinite(rz) || !(rz > 0.0)) {
      info[0] = FEMM_LINEAR_SOLVE_BREAKDOWN;
      return;
    }
    const double beta = rz / rho;
    for (int i = 0; i < n; ++i)
      p[i] = z[i] + beta * p[i];
    rho = rz;
  }
  double fr = 0.0;
  for (int i = 0; i < n; ++i) {
    r[i] = b[i] / scale - spmv(x, i);
    fr += r[i] * r[i];
  }
  if (!std::isfinite(fr)) {
    info[0] = FEMM_NUMERICAL_NONFINITE;
    return;
  }
  for (int i = 0; i < n; ++i) {
    x[i] *= scale;
    if (!std::isfinite(x[i])) {
      info[0] = FEMM_NUMERICAL_NONFINITE;
      return;
    }
  }
  info[2] = std::sqrt(fr) * scale;
  info[0] = std::isfinite(info[2]) ? FEMM_LINEAR_SOLVE_NOT_CONVERGED : FEMM_NUMERICAL_NONFINITE;
}

// physics/nonlinear_model.cpp:166-193.
double LinearTolerance(const femm_newton_options_t &o, double prev)
{
  const double exact = o.linear_relative_tolerance;
  if (!o.adaptive_linear_tolerance || !std::isfinite(prev) || prev <= 100.0 * o.relative_tolerance ||
      !(exact > 0.0) || !std::isfinite(exact))
    return exact;
  return std::max(exact, 1e-8 * std::min(1.0, prev));
}

bool StepConverged(const femm_newton_options_t &o, int it, double step, double tol)
{
  return it > 0 && step < 100.0 * o.relative_tolerance && tol <= o.linear_relative_tolerance;
}

bool HasNonlinear(const Model &m)
{
  for (const auto &mt : m.mat)
    if (mt.bh_curve >= 0 || mt.alpha != 0.0)
      return true;
  return false;
}

// Equation norm of nonlinear_solver.cu:549-558 (hypot accumulation).
double EquationNorm(const Assembly &a, const std::vector<double> &x)
{
  double norm = 0.0;
  for (size_t r = 0; r < a.free_nodes.size(); ++r) {
    double v = -a.rhs_per_amp[r] - a.rhs_offset[r];
    for (int k = a.rows[r]; k < a.rows[r + 1]; ++k)
      v += a.values[k] * x[a.free_nodes[a.cols[k]]];
    norm = std::hypot(norm, v);
  }
  return norm;
}

// Initial state with prescribed signed roots expanded (nonlinear_solver.cu:545-569).
bool StartState(const Model &m, std::vector<double> *state)
{
  DofMap d;
  if (!BuildDofMap(m, &d))
    return false;
  state->assign(m.nodes.size(), 0.0);
  for (size_t i = 0; i < m.dir_nodes.size(); ++i)
    (*state)[m.dir_nodes[i]] = m.dir_values[i];
  for (size_t v = 0; v < m.nodes.size(); ++v) {
    const int32_t r = d.root[v];
    if (d.root_fixed[r])
      (*state)[v] = static_cast<double>(d.sign[v]) * d.root_value[r];
  }
  return true;
}

double EquationResidual(const Model &m, const std::vector<double> &a,
                        const std::vector<double> &cur)
{
  std::vector<double> start;
  if (!StartState(m, &start))
    return std::numeric_limits<double>::infinity();
  Assembly s0, s1;
  if (Assemble(m, start, cur, true, &s0) != FEMM_OK || Assemble(m, a, cur, true, &s1) != FEMM_OK)
    return std::numeric_limits<double>::infinity();
  const double scale = std::max(EquationNorm(s0, start), std::numeric_limits<double>::min());
  return EquationNorm(s1, a) / scale;
}

// Linear tolerance of a safeguarded Newton step.  forcing = 0: the options' exact tolerance
// (the reference's SolveSafeguarded).  forcing > 0: inexact Newton with a forcing term
// proportional to the current relative equation residual (Eisenstat & Walker, "Choosing the
// forcing terms in an inexact Newton method", SIAM J. Sci. Comput. 17, 1996), kept within
// [exact, 1e-3].
double ForcingTolerance(const femm_newton_options_t &o, double forcing, double rel_residual)
{
  const double exact = o.linear_relative_tolerance;
  if (!(forcing > 0.0))
    return exact;
  return std::fmax(exact, std::fmin(1e-3, forcing * rel_residual));
}

// nonlinear_solver.cu:534-659: host-assembled Newton with Armijo backtracking
// on the equation residual.
// With `warm`, the iteration starts from the warm state (free nodes taken from it, the
// prescribed values kept) while the residual scale stays the norm at the plain start:
// the warm-start hook of Solve (nonlinear_solver.cu:667-698) combined with this loop.
// With forcing > 0 the linear solves are inexact (ForcingTolerance); a step that meets both
// gates on an inexact solve is solved again (same linearization) to the exact tolerance,
// which counts as a Newton iteration, and a step that meets both gates on an exact solve is
// tried at alpha = 1: the item is accepted at that candidate when the candidate's own
// equation residual passes, else at the current state.
int Safeguarded(const Model &m, const std::vector<double> &cur, const femm_newton_options_t &o,
                std::vector<double> *state_out, oracle_result_t *res, const double *warm = nullptr,
                double forcing = 0.0)
{
  std::vector<double> state;
  if (!StartState(m, &state))
    return FEMM_BOUNDARY_INVALID;
  Assembly as;
  int st = Assemble(m, state, cur, true, &as);
  if (st != FEMM_OK)
    return st;
  double norm = EquationNorm(as, state);
  const double scale = std::max(norm, std::numeric_limits<double>::min());
  if (warm) {
    for (size_t v = 0; v < state.size(); ++v)
      if (as.node_to_free[v] >= 0)
        state[v] = warm[v];
    st = Assemble(m, state, cur, true, &as);
    if (st != FEMM_OK)
      return st;
    norm = EquationNorm(as, state);
  }
  const double exact = o.linear_relative_tolerance;
  int status = FEMM_NONLINEAR_SOLVE_NOT_CONVERGED;
  bool exact_next = false;
  for (int it = 0; it < o.max_newton_iterations; ++it) {
    if (!std::isfinite(norm)) {
      status = FEMM_NUMERICAL_NONFINITE;
      break;
    }
    const size_t nf = as.free_nodes.size();
    std::vector<double> rhs(nf), sol(nf);
    for (size_t r = 0; r < nf; ++r)
      rhs[r] = as.rhs_per_amp[r] + as.rhs_offset[r];
    const double tol = exact_next ? exact : ForcingTolerance(o, forcing, norm / scale);
    exact_next = false;
    double info[3];
    Pcg(static_cast<int>(nf), as.rows.data(), as.cols.data(), as.values.data(), as.diag.data(),
        rhs.data(), tol, o.max_linear_iterations, sol.data(), info);
    res->linear_iterations += static_cast<int32_t>(info[1]);
    if (static_cast<int>(info[0]) != FEMM_OK) {
      status = static_cast<int>(info[0]);
      break;
    }
    const std::vector<double> cand = Expand(as, sol);
    double sn = 0.0, cn = 0.0;
    for (size_t v = 0; v < state.size(); ++v) {
      sn = std::hypot(sn, cand[v] - state[v]);
      cn = std::hypot(cn, cand[v]);
    }
    const double step = sn / std::max(cn, std::numeric_limits<double>::min());
    res->newton_iterations += 1;
    res->relative_step = step;
    if (!std::isfinite(step)) {
      status = FEMM_NUMERICAL_NONFINITE;
      break;
    }
    const bool gates = step < 100.0 * o.relative_tolerance && norm / scale <= o.relative_tolerance;
    if (gates && tol > exact) {
      exact_next = true;
      continue;
    }
    if (gates && !(forcing > 0.0)) {
      status = FEMM_OK;
      break;
    }
    bool accepted = false;
    double alpha = 1.0;
    for (int bt = 0; bt < 20; ++bt, alpha *= 0.5) {
      std::vector<double> trial(state.size());
      for (size_t v = 0; v < state.size(); ++v)
        trial[v] = state[v] + alpha * (cand[v] - state[v]);
      Assembly ta;
      st = Assemble(m, trial, cur, true, &ta);
      if (st != FEMM_OK) {
        if (!gates)
          status = st;
        break;
      }
      const double tn = EquationNorm(ta, trial);
      const bool take = gates ? std::isfinite(tn) && tn / scale <= o.relative_tolerance
                              : std::isfinite(tn) && (tn <= (1.0 - 1e-4 * alpha) * norm ||
                                                      tn / scale <= o.relative_tolerance);
      if (take) {
        state = std::move(trial);
        as = std::move(ta);
        norm = tn;
        accepted = true;
      }
      if (take || gates)
        break;
    }
    if (gates) {
      status = FEMM_OK;
      break;
    }
    if (!accepted)
      break;
  }
  *state_out = std::move(state);
  res->equation_residual = norm / scale;
  return status;
}

// nonlinear_solver.cu:16-83.
int Finalize(const Model &m, const std::vector<double> &a, double *bx, double *by, double *flux)
{
  std::vector<double> fl(static_cast<size_t>(m.circuits), 0.0);
  for (size_t e = 0; e < m.tri.size(); ++e) {
    const auto &t = m.tri[e];
    const femm_material_t &mt = m.mat[m.tri_mat[e]];
    const Node p[3] = {m.nodes[t[0]], m.nodes[t[1]], m.nodes[t[2]]};
    double ex = 0.0, ey = 0.0;
    if (m.problem == FEMM_PLANAR) {
      Terms tm;
      if (!ElementTerms(p, &tm))
        return FEMM_MESH_INVALID;
      for (int i = 0; i < 3; ++i) {
        ex += a[t[i]] * tm.bx[i];
        ey += a[t[i]] * tm.by[i];
      }
      if (mt.circuit >= 0)
        fl[mt.circuit] += tm.area * mt.source_j_per_a / 3.0 * (a[t[0]] + a[t[1]] + a[t[2]]);
    } else {
      const double r[3] = {p[0].x, p[1].x, p[2].x}, z[3] = {p[0].y, p[1].y, p[2].y};
      AxPoint pts[7];
      double area = 0.0;
      if (!AxElement(r, z, pts, &area))
        return FEMM_MESH_INVALID;
      for (int k = 0; k < 3; ++k) {
        ex += a[t[k]] * pts[0].br[k];
        ey += a[t[k]] * pts[0].bz[k];
      }
      if (mt.circuit >= 0) {
        double f = 0.0;
        for (const AxPoint &q : pts) {
          double av = 0.0;
          for (int k = 0; k < 3; ++k)
            av += q.shape[k] * a[t[k]];
          f += kTwoPi * q.radius * q.weight_area * mt.source_j_per_a * av;
        }
        fl[mt.circuit] += f;
      }
    }
    if (bx)
      bx[e] = ex;
    if (by)
      by[e] = ey;
  }
  int status = FEMM_OK;
  for (int c = 0; c < m.circuits; ++c) {
    if (m.problem == FEMM_PLANAR)
      fl[c] *= m.depth;
    if (!std::isfinite(fl[c]))
      status = FEMM_NUMERICAL_NONFINITE;
    if (flux)
      flux[c] = fl[c];
  }
  return status;
}

// nonlinear_solver.cu:667-769 (single-item Newton), then the recovery of
// :515-528 for an exhausted item, then the equation residual at the result.
int NewtonSolve(const Model &m, const std::vector<double> &cur, femm_newton_options_t o,
                const double *warm, std::vector<double> *out, oracle_result_t *res)
{
  *res = oracle_result_t{};
  res->status = FEMM_INVALID_ARGUMENT;
  res->relative_step = std::numeric_limits<double>::infinity();
  res->equation_residual = std::numeric_limits<double>::infinity();
  const int v = Validate(m);
  if (v != FEMM_OK)
    return res->status = v;
  if (cur.size() != static_cast<size_t>(m.circuits) || !(o.relative_tolerance > 0.0) ||
      !std::isfinite(o.relative_tolerance) || o.max_newton_iterations < 0)
    return res->status;
  for (double c : cur)
    if (!std::isfinite(c))
      return res->status;
  o.adaptive_linear_tolerance = o.adaptive_linear_tolerance && HasNonlinear(m);
  std::vector<double> a;
  if (!StartState(m, &a))
    return res->status = FEMM_BOUNDARY_INVALID;
  if (warm) {
    for (size_t i = 0; i < a.size(); ++i)
      a[i] = warm[i];
    for (size_t i = 0; i < m.dir_nodes.size(); ++i)
      a[m.dir_nodes[i]] = m.dir_values[i];
  }
  double omega = 1.0, prev = std::numeric_limits<double>::infinity();
  int status = FEMM_NONLINEAR_SOLVE_NOT_CONVERGED;
  for (int it = 0; it < o.max_newton_iterations; ++it) {
    Assembly as;
    const bool newton = warm != nullptr || it > 0;
    const int st = Assemble(m, a, cur, newton, &as);
    if (st != FEMM_OK) {
      status = st;
      break;
    }
    const size_t nf = as.free_nodes.size();
    std::vector<double> rhs(nf), sol(nf);
    for (size_t r = 0; r < nf; ++r)
      rhs[r] = as.rhs_per_amp[r] + as.rhs_offset[r];
    const double tol = LinearTolerance(o, prev);
    double info[3];
    Pcg(static_cast<int>(nf), as.rows.data(), as.cols.data(), as.values.data(), as.diag.data(),
        rhs.data(), tol, o.max_linear_iterations, sol.data(), info);
    res->linear_iterations += static_cast<int32_t>(info[1]);
    if (static_cast<int>(info[0]) != FEMM_OK) {
      status = static_cast<int>(info[0]);
      break;
    }
    const std::vector<double> cand = Expand(as, sol);
    double dsq = 0.0, csq = 0.0;
    for (size_t i = 0; i < cand.size(); ++i) {
      const double dd = cand[i] - a[i];
      dsq += dd * dd;
      csq += cand[i] * cand[i];
    }
    const double step =
        std::sqrt(dsq) / std::max(std::sqrt(csq), std::numeric_limits<double>::min());
    if (!std::isfinite(step)) {
      status = FEMM_NUMERICAL_NONFINITE;
      break;
    }
    if (it > 5) {
      if (step > prev && omega > 0.125)
        omega *= 0.5;
      else
        omega += 0.1 * (1.0 - omega);
    }
    for (size_t i = 0; i < a.size(); ++i)
      a[i] += omega * (cand[i] - a[i]);
    res->newton_iterations = it + 1;
    res->relative_step = step;
    if (StepConverged(o, it, step, tol)) {
      if (!o.require_equation_residual) {
        status = FEMM_OK;
        break;
      }
      // Both gates of nonlinear_solver.cu:623-627: the equation residual at the
      // returned state must also pass before the item is accepted.
      const double eq = EquationResidual(m, a, cur);
      res->equation_residual = eq;
      if (eq <= o.relative_tolerance) {
        status = FEMM_OK;
        break;
      }
    }
    prev = step;
  }
  if (status == FEMM_NONLINEAR_SOLVE_NOT_CONVERGED && o.recover_nonconvergence &&
      o.max_newton_iterations > 0) {
    const int original = res->newton_iterations;
    const int linear = res->linear_iterations;
    oracle_result_t rec{};
    std::vector<double> state;
    status = Safeguarded(m, cur, o, &state, &rec);
    a = std::move(state);
    res->newton_iterations = original + rec.newton_iterations;
    res->linear_iterations = linear + rec.linear_iterations;
    res->relative_step = rec.relative_step;
  }
  res->equation_residual = EquationResidual(m, a, cur);
  *out = std::move(a);
  return res->status = status;
}

// Weighted-stress mask (weighted_stress.cu:24-183), plan
// (field_postprocess.cpp:221-339), loads and probes (:341-403).
bool Selected(const Model &m, int mat_index, int sel_label, int sel_group)
{
  return sel_group >= 0 ? m.mat[mat_index].group == sel_group : mat_index == sel_label;
}

std::pair<int32_t, int32_t> Edge(int32_t a, int32_t b) { return {std::min(a, b), std::max(a, b)}; }

bool Barycentric(const Node p[3], double x, double y, double l[3])
{
  const double det = (p[1].x - p[0].x) * (p[2].y - p[0].y) - (p[2].x - p[0].x) * (p[1].y - p[0].y);
  if (!(det > 0.0) || !std::isfinite(det))
    return false;
  l[1] = ((x - p[0].x) * (p[2].y - p[0].y) - (p[2].x - p[0].x) * (y - p[0].y)) / det;
  l[2] = ((p[1].x - p[0].x) * (y - p[0].y) - (x - p[0].x) * (p[1].y - p[0].y)) / det;
  l[0] = 1.0 - l[1] - l[2];
  return std::isfinite(l[0]) && std::isfinite(l[1]) && std::isfinite(l[2]);
}

int WeightedStressMask(const Model &m, int sl, int al, int sg, int ag, std::vector<double> *mask)
{
  const size_t nn = m.nodes.size();
  if ((sg < 0 && (sl < 0 || static_cast<size_t>(sl) >= m.mat.size())) ||
      (ag < 0 && (al < 0 || static_cast<size_t>(al) >= m.mat.size())) || (sg >= 0 && sg == ag) ||
      (sg < 0 && ag < 0 && sl == al))
    return FEMM_INVALID_ARGUMENT;
  const int v = Validate(m);
  if (v != FEMM_OK)
    return v;
  std::map<std::pair<int32_t, int32_t>, std::vector<int32_t>> edges;
  bool seen = false;
  for (size_t e = 0; e < m.tri.size(); ++e) {
    seen = seen || Selected(m, m.tri_mat[e], sl, sg);
    for (int k = 0; k < 3; ++k)
      edges[Edge(m.tri[e][k], m.tri[e][(k + 1) % 3])].push_back(static_cast<int32_t>(e));
  }
  if (!seen)
    return FEMM_INVALID_MATERIAL;
  for (const auto &ed : edges) {
    bool sel = false, other = false;
    for (int32_t e : ed.second) {
      const bool s = Selected(m, m.tri_mat[e], sl, sg), a = Selected(m, m.tri_mat[e], al, ag);
      sel = sel || s;
      other = other || (!s && !a);
    }
    if (sel && other)
      return FEMM_INVALID_MATERIAL;
  }
  const double nan = std::numeric_limits<double>::quiet_NaN();
  std::vector<double> fixed(nn, nan);
  for (int32_t b : m.dir_nodes)
    fixed[b] = 0.0;
  for (size_t e = 0; e < m.tri.size(); ++e) {
    if (Selected(m, m.tri_mat[e], sl, sg)) {
      for (int32_t v2 : m.tri[e]) {
        if (std::isfinite(fixed[v2]) && fixed[v2] != 1.0)
          return FEMM_BOUNDARY_INVALID;
        fixed[v2] = 1.0;
      }
    } else if (!Selected(m, m.tri_mat[e], al, ag)) {
      for (int32_t v2 : m.tri[e])
        if (!std::isfinite(fixed[v2]))
          fixed[v2] = 0.0;
    }
  }
  Rows k(nn);
  for (const auto &t : m.tri) {
    const Node p[3] = {m.nodes[t[0]], m.nodes[t[1]], m.nodes[t[2]]};
    Terms tm;
    if (!ElementTerms(p, &tm))
      return FEMM_MESH_INVALID;
    const double wgt = std::sqrt(tm.area);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        if (!AddEntry(&k, t[i], t[j], wgt * tm.area * (tm.bx[i] * tm.bx[j] + tm.by[i] * tm.by[j])))
          return FEMM_NUMERICAL_NONFINITE;
  }
  bool all_fixed = true;
  for (double f : fixed)
    all_fixed = all_fixed && std::isfinite(f);
  if (all_fixed) {
    mask->assign(nn, 0.0);
    for (size_t i = 0; i < nn; ++i)
      (*mask)[i] = fixed[i] > 0.5 ? 1.0 : 0.0;
    return FEMM_OK;
  }
  Model mm = m;
  for (auto &c : mm.cons)
    c.sign = 1;
  mm.dir_nodes.clear();
  mm.dir_values.clear();
  for (size_t i = 0; i < nn; ++i)
    if (std::isfinite(fixed[i])) {
      mm.dir_nodes.push_back(static_cast<int32_t>(i));
      mm.dir_values.push_back(fixed[i]);
    }
  Assembly as;
  const int st = Reduce(mm, k, std::vector<double>(nn, 0.0), &as);
  if (st != FEMM_OK)
    return st;
  const size_t nf = as.free_nodes.size();
  std::vector<double> rhs(nf), sol(nf);
  for (size_t r = 0; r < nf; ++r)
    rhs[r] = as.rhs_per_amp[r] + as.rhs_offset[r];
  double info[3];
  Pcg(static_cast<int>(nf), as.rows.data(), as.cols.data(), as.values.data(), as.diag.data(),
      rhs.data(), 1e-12, 256, sol.data(), info);
  if (static_cast<int>(info[0]) != FEMM_OK)
    return static_cast<int>(info[0]);
  *mask = Expand(as, sol);
  for (double &x : *mask)
    x = x > 0.5 ? 1.0 : 0.0;
  return FEMM_OK;
}

// field_postprocess.cpp:34-190: FEMM GetNodalB subset.
bool SmoothedB(const Model &m, const std::vector<double> &a, const double *bx, const double *by,
               const std::vector<std::vector<int32_t>> &inc, int32_t el, double nbx[3],
               double nby[3])
{
  const auto &tri = m.tri[el];
  for (int loc = 0; loc < 3; ++loc) {
    const int32_t node = tri[loc];
    bool homo = true;
    for (int32_t nb : inc[node])
      homo = homo && m.tri_mat[el] == m.tri_mat[nb];
    if (homo) {
      double ws = 0.0;
      nbx[loc] = 0.0;
      nby[loc] = 0.0;
      for (int32_t nb : inc[node]) {
        const auto &c = m.tri[nb];
        const double cx = (m.nodes[c[0]].x + m.nodes[c[1]].x + m.nodes[c[2]].x) / 3.0;
        const double cy = (m.nodes[c[0]].y + m.nodes[c[1]].y + m.nodes[c[2]].y) / 3.0;
        const double w = 1.0 / std::hypot(m.nodes[node].x - cx, m.nodes[node].y - cy);
        if (!(w > 0.0) || !std::isfinite(w))
          return false;
        ws += w;
        nbx[loc] += w * bx[nb];
        nby[loc] += w * by[nb];
      }
      nbx[loc] /= ws;
      nby[loc] /= ws;
      continue;
    }
    auto local_index = [&](int32_t cand) {
      for (int i = 0; i < 3; ++i)
        if (m.tri[cand][i] == node)
          return i;
      return -1;
    };
    auto next_across = [&](int32_t cur, int32_t other) {
      for (int32_t cand : inc[node]) {
        if (cand == cur)
          continue;
        for (int i = 0; i < 3; ++i)
          if (m.tri[cand][i] == other)
            return cand;
      }
      return static_cast<int32_t>(-1);
    };
    double ws = 0.0, v1x = 0.0, v1y = 0.0, v2x = 0.0, v2y = 0.0;
    nbx[loc] = 0.0;
    nby[loc] = 0.0;
    bool special = false;
    auto add_interface = [&](int32_t src, int32_t other) {
      const double dx = m.nodes[other].x - m.nodes[node].x, dy = m.nodes[other].y - m.nodes[node].y;
      const double len = std::hypot(dx, dy);
      if (!(len > 0.0))
        return false;
      const double tx = dx / len, ty = dy / len;
      const double bt = bx[src] * tx + by[src] * ty;
      const double bn = (a[other] - a[node]) / len;
      const double w = 0.5 / len;
      nbx[loc] += w * (tx * bt + ty * bn);
      nby[loc] += w * (ty * bt - tx * bn);
      ws += w;
      return true;
    };
    for (int pass = 0; pass < 2 && !special; ++pass) {
      int32_t cur = el;
      for (size_t scan = 0; scan < inc[node].size(); ++scan) {
        const int idx = local_index(cur);
        if (idx < 0)
          return false;
        const int32_t other = m.tri[cur][(idx + (pass == 0 ? 2 : 1)) % 3];
        const int32_t nx = next_across(cur, other);
        if (nx < 0) {
          nbx[loc] = bx[cur];
          nby[loc] = by[cur];
          special = true;
          break;
        }
        if (m.tri_mat[el] != m.tri_mat[nx]) {
          if (!add_interface(cur, other))
            return false;
          const double len = std::hypot(m.nodes[other].x - m.nodes[node].x,
                                        m.nodes[other].y - m.nodes[node].y);
          const double ux = (m.nodes[other].x - m.nodes[node].x) / len;
          const double uy = (m.nodes[other].y - m.nodes[node].y) / len;
          if (pass == 0) {
            v1x = ux;
            v1y = uy;
          } else {
            v2x = ux;
            v2y = uy;
          }
          break;
        }
        cur = nx;
      }
    }
    if (!special) {
      if (!(ws > 0.0))
        return false;
      nbx[loc] /= ws;
      nby[loc] /= ws;
      const bool simple = std::hypot(v1x, v1y) < 0.9 || std::hypot(v2x, v2y) < 0.9 ||
                          (-v1x * v2x - v1y * v2y) > 0.985;
      if (!simple) {
        double mag = 0.0;
        for (int32_t nb : inc[node])
          if (m.tri_mat[nb] == m.tri_mat[el])
            mag = std::max(mag, std::hypot(bx[nb], by[nb]));
        const double own = std::hypot(bx[el], by[el]);
        if (own > 0.0) {
          nbx[loc] = mag * bx[el] / own;
          nby[loc] = mag * by[el] / own;
        } else {
          nbx[loc] = 0.0;
          nby[loc] = 0.0;
        }
      }
    }
  }
  return true;
}

// ComputeAirGapElementPostprocess (postprocess/airgap_postprocess.cpp:41-160).  A cell's
// field at its centre is the gradient of the bilinear interpolant of its four virtual
// corners (Catmull-Rom values of the rings at the element's own shifts): Br across the arc
// over R dt, Bt = -dA/dr across the gap.  Sums run over elements and cells in order.
int AgePostprocess(const Model &m, const std::vector<double> &a, const double *angles, int count,
                   double out[3], double *radial)
{
  if (m.ages.empty() || a.size() != m.nodes.size())
    return FEMM_INVALID_ARGUMENT;
  constexpr double kMu0 = 4.0e-7 * 3.141592653589793238462643383279502884;
  out[0] = out[1] = out[2] = 0.0;
  std::vector<double> first_br;
  for (size_t g = 0; g < m.ages.size(); ++g) {
    const Model::AirGap &age = m.ages[g];
    const int n = age.cells;
    const double dt = (kPi / 180.0) * age.arc_deg / static_cast<double>(n);
    const double radius = 0.5 * (age.ri + age.ro), dr = age.ro - age.ri;
    if (!(dt > 0.0) || !(radius > 0.0) || !(dr > 0.0) || n < 2)
      return FEMM_MESH_INVALID;
    double wi[4], wo[4];
    HermiteCentral(1.0 - age.ci, wi);
    HermiteCentral(1.0 - age.co, wo);
    for (int k = 0; k < n; ++k) {
      int32_t node[10];
      double weight[10], v[10];
      AgeStencil(age, k, node, weight);
      for (int q = 0; q < 10; ++q)
        v[q] = a[node[q]] * weight[q];
      double corner[4] = {0.0, 0.0, 0.0, 0.0};   // inner u=0, inner u=1, outer u=0, outer u=1
      for (int q = 0; q < 4; ++q) {
        corner[0] += wi[q] * v[q];
        corner[1] += wi[q] * v[q + 1];
        corner[2] += wo[q] * v[q + 5];
        corner[3] += wo[q] * v[q + 6];
      }
      const double br = ((corner[1] - corner[0]) + (corner[3] - corner[2])) / (2.0 * dt * radius);
      const double bt = ((corner[0] + corner[1]) - (corner[2] + corner[3])) / (2.0 * dr);
      if (!std::isfinite(br) || !std::isfinite(bt))
        return FEMM_NUMERICAL_NONFINITE;
      const double theta = (static_cast<double>(k) + 0.5) * dt;
      const double normal = br * br - bt * bt, shear = 2.0 * br * bt;
      const double scale = m.depth * radius * dt / (2.0 * kMu0);
      out[0] += scale * (normal * std::cos(theta) - shear * std::sin(theta));
      out[1] += scale * (normal * std::sin(theta) + shear * std::cos(theta));
      out[2] += m.depth * radius * radius * dt * br * bt / kMu0;
      if (g == 0)
        first_br.push_back(br);
    }
  }
  const Model::AirGap &age = m.ages.front();
  const int n = age.cells;
  const double dt = (kPi / 180.0) * age.arc_deg / static_cast<double>(n);
  const size_t harmonics = age.antiperiodic ? static_cast<size_t>(n + 1) / 2
                                            : static_cast<size_t>(n) / 2 + 1;
  const int32_t base =
      static_cast<int32_t>(std::lround((age.antiperiodic ? 180.0 : 360.0) / age.arc_deg));
  std::vector<double> bc(harmonics, 0.0), bs(harmonics, 0.0);
  std::vector<int32_t> order(harmonics);
  for (size_t h = 0; h < harmonics; ++h) {
    order[h] = age.antiperiodic ? base * static_cast<int32_t>(2 * h + 1)
                                : base * static_cast<int32_t>(h);
    for (int k = 0; k < n; ++k) {
      const double phase = (static_cast<double>(k) + 0.5) * dt * static_cast<double>(order[h]);
      bc[h] += first_br[k] * std::cos(phase);
      bs[h] += first_br[k] * std::sin(phase);
    }
    const bool edge = order[h] == 0 || (!age.antiperiodic && h + 1 == harmonics && n % 2 == 0);
    const double norm = edge ? static_cast<double>(n) : static_cast<double>(n) / 2.0;
    bc[h] /= norm;
    bs[h] /= norm;
  }
  for (int p = 0; p < count; ++p) {
    if (!std::isfinite(angles[p]))
      return FEMM_INVALID_ARGUMENT;
    double value = 0.0;
    for (size_t h = 0; h < harmonics; ++h) {
      const double phase = static_cast<double>(order[h]) * angles[p];
      value += bc[h] * std::cos(phase) + bs[h] * std::sin(phase);
    }
    radial[p] = value;
  }
  if (!std::isfinite(out[0]) || !std::isfinite(out[1]) || !std::isfinite(out[2]))
    return FEMM_NUMERICAL_NONFINITE;
  return FEMM_OK;
}

}  // namespace

struct oracle_assembly {
  Assembly a;
};

struct oracle_symbolic {
  Symbolic s;
};

// NonlinearP1Solver::SolveTangentMultiRhs (solvers/nonlinear_solver.cu:85-172): the Newton
// operator at the base state, one right-hand side per direction built circuit by circuit and
// node by node (rhs[row] += dir_c * sign * source_c[node]), Jacobi PCG from x0 = 0, the
// derivative expanded with signs and zero on prescribed nodes
// (ExpandAssemblyDerivative, physics/nonlinear_model.cpp:413-426), then FinalizeResult.
int Tangent(const Model &m, const std::vector<double> &a, const std::vector<double> &cur,
            int directions, const double *dirs, const femm_newton_options_t &o, double *da,
            double *dflux, oracle_result_t *res)
{
  for (int d = 0; d < directions; ++d) {
    res[d] = oracle_result_t{};
    res[d].status = FEMM_INVALID_ARGUMENT;
  }
  if (!(o.linear_relative_tolerance > 0.0) || o.max_linear_iterations < 0)
    return FEMM_OK;
  for (int i = 0; i < directions * m.circuits; ++i)
    if (!std::isfinite(dirs[i]))
      return FEMM_OK;
  Assembly tg;
  const int st = Assemble(m, a, cur, true, &tg);
  if (st != FEMM_OK) {
    for (int d = 0; d < directions; ++d)
      res[d].status = st;
    return FEMM_OK;
  }
  const size_t nf = tg.free_nodes.size(), nn = m.nodes.size();
  for (int d = 0; d < directions; ++d) {
    std::vector<double> rhs(nf, 0.0), sol(nf, 0.0);
    for (int c = 0; c < m.circuits; ++c) {
      const double scale = dirs[static_cast<size_t>(d) * m.circuits + c];
      if (scale == 0.0)
        continue;
      for (size_t v = 0; v < nn; ++v) {
        const int32_t row = tg.node_to_free[v];
        if (row >= 0)
          rhs[row] += scale * static_cast<double>(tg.node_sign[v]) * tg.source_per_circuit[c][v];
      }
    }
    double info[3];
    Pcg(static_cast<int>(nf), tg.rows.data(), tg.cols.data(), tg.values.data(), tg.diag.data(),
        rhs.data(), o.linear_relative_tolerance, o.max_linear_iterations, sol.data(), info, 1);
    res[d].status = static_cast<int>(info[0]);
    res[d].linear_iterations = static_cast<int32_t>(info[1]);
    std::vector<double> expanded(nn, 0.0);
    if (res[d].status == FEMM_OK)
      for (size_t v = 0; v < nn; ++v) {
        const int32_t row = tg.node_to_free[v];
        if (row >= 0)
          expanded[v] = static_cast<double>(tg.node_sign[v]) * sol[row];
      }
    std::copy(expanded.begin(), expanded.end(), da + static_cast<size_t>(d) * nn);
    std::vector<double> bx(m.tri.size()), by(m.tri.size()), fl(std::max(m.circuits, 1));
    if (Finalize(m, expanded, bx.data(), by.data(), fl.data()) == FEMM_OK)
      std::copy(fl.begin(), fl.begin() + m.circuits, dflux + static_cast<size_t>(d) * m.circuits);
  }
  return FEMM_OK;
}

extern "C" {

int oracle_bh_spline(const double *b, const double *h, int n, double *out_b, double *out_h,
                     double *out_slope, double *cold_secant, int *smoothing_passes)
{
  Curve c;
  c.b.assign(b, b + n);
  c.h.assign(h, h + n);
  const bool ok = BuildSpline(&c);
  if (ok) {
    std::copy(c.b.begin(), c.b.end(), out_b);
    std::copy(c.h.begin(), c.h.end(), out_h);
    std::copy(c.slope.begin(), c.slope.end(), out_slope);
  }
  *cold_secant = c.cold;
  *smoothing_passes = c.passes;
  return ok ? 1 : 0;
}

void oracle_bh_eval(const double *b, const double *h, const double *slope, int n, double field_b,
                    double out[4])
{
  const BhEval e = EvalBh(b, h, slope, n, field_b);
  out[0] = e.h;
  out[1] = e.dh;
  out[2] = e.nu;
  out[3] = e.dnu;
}

int oracle_validate(const femm_model_t *model)
{
  Model m;
  if (!LoadModel(model, &m))
    return FEMM_INVALID_ARGUMENT;
  return Validate(m);
}

int oracle_assemble(const femm_model_t *model, const double *a_old, const double *currents,
                    int use_newton, oracle_assembly **out)
{
  *out = nullptr;
  Model m;
  if (!LoadModel(model, &m))
    return FEMM_INVALID_ARGUMENT;
  auto *h = new oracle_assembly;
  const int st = Assemble(m, std::vector<double>(a_old, a_old + m.nodes.size()),
                          std::vector<double>(currents, currents + m.circuits), use_newton != 0,
                          &h->a);
  if (st != FEMM_OK) {
    delete h;
    return st;
  }
  *out = h;
  return FEMM_OK;
}

void oracle_assembly_sizes(const oracle_assembly *h, int32_t *free_count, int32_t *nnz)
{
  *free_count = static_cast<int32_t>(h->a.free_nodes.size());
  *nnz = static_cast<int32_t>(h->a.values.size());
}

void oracle_assembly_copy(const oracle_assembly *h, int32_t *row_offsets, int32_t *column_indices,
                          double *values, double *diagonal, double *rhs_per_amp,
                          double *rhs_offset, int32_t *free_nodes, int32_t *node_to_free,
                          int8_t *node_sign, double *boundary_values)
{
  const Assembly &a = h->a;
  std::copy(a.rows.begin(), a.rows.end(), row_offsets);
  std::copy(a.cols.begin(), a.cols.end(), column_indices);
  std::copy(a.values.begin(), a.values.end(), values);
  std::copy(a.diag.begin(), a.diag.end(), diagonal);
  std::copy(a.rhs_per_amp.begin(), a.rhs_per_amp.end(), rhs_per_amp);
  std::copy(a.rhs_offset.begin(), a.rhs_offset.end(), rhs_offset);
  std::copy(a.free_nodes.begin(), a.free_nodes.end(), free_nodes);
  std::copy(a.node_to_free.begin(), a.node_to_free.end(), node_to_free);
  std::copy(a.node_sign.begin(), a.node_sign.end(), node_sign);
  std::copy(a.boundary.begin(), a.boundary.end(), boundary_values);
}

void oracle_assembly_destroy(oracle_assembly *h) { delete h; }

int oracle_prepare_symbolic(const femm_model_t *model, oracle_symbolic **out)
{
  *out = nullptr;
  Model m;
  if (!LoadModel(model, &m))
    return FEMM_INVALID_ARGUMENT;
  auto *h = new oracle_symbolic;
  const int st = PrepareSymbolic(m, &h->s);
  if (st != FEMM_OK) {
    delete h;
    return st;
  }
  *out = h;
  return FEMM_OK;
}

void oracle_symbolic_sizes(const oracle_symbolic *h, int32_t *free_count, int32_t *nnz,
                           int32_t *matrix_entries, int32_t *rhs_entries)
{
  *free_count = static_cast<int32_t>(h->s.sym.free_nodes.size());
  *nnz = static_cast<int32_t>(h->s.sym.cols.size());
  *matrix_entries = static_cast<int32_t>(h->s.mat_entries.size());
  *rhs_entries = static_cast<int32_t>(h->s.rhs_entries.size());
}

void oracle_symbolic_copy(const oracle_symbolic *h, int32_t *row_offsets, int32_t *column_indices,
                          int32_t *diagonal_slots, int32_t *matrix_offsets, int32_t *matrix_entries,
                          int32_t *rhs_offsets, int32_t *rhs_entries, double *pm_rhs,
                          double *circuit_rhs, int32_t *free_nodes, int32_t *node_to_free,
                          int8_t *node_sign, double *boundary_values)
{
  const Symbolic &s = h->s;
  std::copy(s.sym.rows.begin(), s.sym.rows.end(), row_offsets);
  std::copy(s.sym.cols.begin(), s.sym.cols.end(), column_indices);
  std::copy(s.diag_slots.begin(), s.diag_slots.end(), diagonal_slots);
  std::copy(s.mat_offsets.begin(), s.mat_offsets.end(), matrix_offsets);
  for (size_t k = 0; k < s.mat_entries.size(); ++k)
    for (int j = 0; j < 4; ++j)
      matrix_entries[4 * k + j] = s.mat_entries[k][j];
  std::copy(s.rhs_offsets.begin(), s.rhs_offsets.end(), rhs_offsets);
  for (size_t k = 0; k < s.rhs_entries.size(); ++k)
    for (int j = 0; j < 3; ++j)
      rhs_entries[3 * k + j] = s.rhs_entries[k][j];
  std::copy(s.pm_rhs.begin(), s.pm_rhs.end(), pm_rhs);
  std::copy(s.circuit_rhs.begin(), s.circuit_rhs.end(), circuit_rhs);
  std::copy(s.sym.free_nodes.begin(), s.sym.free_nodes.end(), free_nodes);
  std::copy(s.sym.node_to_free.begin(), s.sym.node_to_free.end(), node_to_free);
  std::copy(s.sym.node_sign.begin(), s.sym.node_sign.end(), node_sign);
  std::copy(s.sym.boundary.begin(), s.sym.boundary.end(), boundary_values);
}

void oracle_symbolic_destroy(oracle_symbolic *h) { delete h; }

void oracle_symbolic_age(const oracle_symbolic *h, double *age_values, double *age_rhs)
{
  if (age_values)
    std::copy(h->s.age_values.begin(), h->s.age_values.end(), age_values);
  if (age_rhs)
    std::copy(h->s.age_rhs.begin(), h->s.age_rhs.end(), age_rhs);
}

int oracle_age_postprocess(const femm_model_t *model, const double *a, const double *angles_rad,
                           int angle_count, double out[3], double *radial_b)
{
  Model m;
  if (!LoadModel(model, &m))
    return FEMM_INVALID_ARGUMENT;
  return AgePostprocess(m, std::vector<double>(a, a + m.nodes.size()), angles_rad, angle_count,
                        out, radial_b);
}

int oracle_age_matrix(double K, double Ki, double ci, double co, double out[100])
{
  double m[10][10];
  if (!AgeMatrix(K, Ki, ci, co, m))
    return 0;
  for (int i = 0; i < 10; ++i)
    for (int j = 0; j < 10; ++j)
      out[10 * i + j] = m[i][j];
  return 1;
}

void oracle_pcg(int n, const int32_t *row_offsets, const int32_t *column_indices,
                const double *values, const double *diagonal, const double *rhs,
                double relative_tolerance, int max_iterations, double *x, double info[3])
{
  Pcg(n, row_offsets, column_indices, values, diagonal, rhs, relative_tolerance, max_iterations, x,
      info, 1);
}

void oracle_pcg_block(int n, const int32_t *row_offsets, const int32_t *column_indices,
                      const double *values, const double *diagonal, const double *rhs,
                      double relative_tolerance, int max_iterations, int block_size, double *x,
                      double info[3])
{
  Pcg(n, row_offsets, column_indices, values, diagonal, rhs, relative_tolerance, max_iterations, x,
      info, block_size);
}

void oracle_block_jacobi(int n, const int32_t *row_offsets, const int32_t *column_indices,
                         const double *values, const double *diagonal, int block_size,
                         double *planes, int8_t *fallback)
{
  std::vector<double> pl;
  std::vector<char> fb;
  BuildBlockJacobi(n, row_offsets, column_indices, values, diagonal, block_size, &pl, &fb);
  std::copy(pl.begin(), pl.end(), planes);
  for (size_t b = 0; b < fb.size(); ++b)
    fallback[b] = static_cast<int8_t>(fb[b]);
}

void oracle_set_block_size(int block_size) { g_block_size = std::max(block_size, 1); }

int oracle_tangent_multi_rhs(const femm_model_t *model, const double *base_a,
                             const double *base_currents, int directions,
                             const double *direction_currents,
                             const femm_newton_options_t *options, double *da, double *dflux,
                             oracle_result_t *results)
{
  Model m;
  if (!LoadModel(model, &m) || directions <= 0)
    return FEMM_INVALID_ARGUMENT;
  return Tangent(m, std::vector<double>(base_a, base_a + m.nodes.size()),
                 std::vector<double>(base_currents, base_currents + m.circuits), directions,
                 direction_currents, *options, da, dflux, results);
}

int oracle_newton_solve(const femm_model_t *model, const double *currents,
                        const femm_newton_options_t *options, const double *warm_start, double *a,
                        double *bx, double *by, double *flux, oracle_result_t *result)
{
  Model m;
  if (!LoadModel(model, &m))
    return FEMM_INVALID_ARGUMENT;
  std::vector<double> state;
  const int st = NewtonSolve(m, std::vector<double>(currents, currents + m.circuits), *options,
                             warm_start, &state, result);
  if (state.size() == m.nodes.size()) {
    std::copy(state.begin(), state.end(), a);
    const int fs = Finalize(m, state, bx, by, flux);
    if (st == FEMM_OK && fs != FEMM_OK)
      return result->status = fs;
  }
  return st;
}

int oracle_safeguarded_solve_forcing(const femm_model_t *model, const double *currents,
                                     const femm_newton_options_t *options, double forcing,
                                     const double *warm_start, double *a,
                                     oracle_result_t *result)
{
  *result = oracle_result_t{};
  result->status = FEMM_INVALID_ARGUMENT;
  result->relative_step = std::numeric_limits<double>::infinity();
  result->equation_residual = std::numeric_limits<double>::infinity();
  Model m;
  if (!LoadModel(model, &m))
    return result->status;
  const int v = Validate(m);
  if (v != FEMM_OK)
    return result->status = v;
  const femm_newton_options_t &o = *options;
  std::vector<double> cur(currents, currents + m.circuits);
  if (!(o.relative_tolerance > 0.0) || !std::isfinite(o.relative_tolerance) ||
      o.max_newton_iterations < 0)
    return result->status;
  for (double c : cur)
    if (!std::isfinite(c))
      return result->status;
  std::vector<double> state;
  const int st = Safeguarded(m, cur, o, &state, result, warm_start, forcing);
  if (state.size() == m.nodes.size())
    std::copy(state.begin(), state.end(), a);
  return result->status = st;
}

int oracle_safeguarded_solve(const femm_model_t *model, const double *currents,
                             const femm_newton_options_t *options, const double *warm_start,
                             double *a, oracle_result_t *result)
{
  return oracle_safeguarded_solve_forcing(model, currents, options, 0.0, warm_start, a, result);
}

int oracle_newton_batch(const femm_model_t *model, int items, const double *currents,
                        const femm_newton_options_t *options, int threads, double *a,
                        double *flux, oracle_result_t *results)
{
  Model m;
  if (!LoadModel(model, &m))
    return FEMM_INVALID_ARGUMENT;
  const size_t nn = m.nodes.size(), nc = static_cast<size_t>(m.circuits);
  std::atomic<int> next{0};
  auto work = [&]() {
    for (;;) {
      const int item = next.fetch_add(1);
      if (item >= items)
        return;
      std::vector<double> state;
      NewtonSolve(m, std::vector<double>(currents + item * nc, currents + (item + 1) * nc),
                  *options, nullptr, &state, &results[item]);
      if (state.size() == nn) {
        std::copy(state.begin(), state.end(), a + item * nn);
        Finalize(m, state, nullptr, nullptr, flux + item * nc);
      }
    }
  };
  std::vector<std::thread> pool;
  for (int t = 1; t < std::max(1, threads); ++t)
    pool.emplace_back(work);
  work();
  for (auto &t : pool)
    t.join();
  return FEMM_OK;
}

double oracle_equation_residual(const femm_model_t *model, const double *a, const double *currents)
{
  Model m;
  if (!LoadModel(model, &m))
    return std::numeric_limits<double>::infinity();
  return EquationResidual(m, std::vector<double>(a, a + m.nodes.size()),
                          std::vector<double>(currents, currents + m.circuits));
}

int oracle_finalize(const femm_model_t *model, const double *a, double *bx, double *by,
                    double *flux)
{
  Model m;
  if (!LoadModel(model, &m))
    return FEMM_INVALID_ARGUMENT;
  return Finalize(m, std::vector<double>(a, a + m.nodes.size()), bx, by, flux);
}

int oracle_weighted_stress(const femm_model_t *model, const double *a, const double *bx,
                           const double *by, int32_t selected_label, int32_t air_label,
                           int32_t selected_group, int32_t air_group, double radius_m,
                           const double *angles_rad, int angle_count, double out[3],
                           double *radial_b, double *mask_out)
{
  Model m;
  if (!LoadModel(model, &m))
    return FEMM_INVALID_ARGUMENT;
  std::vector<double> mask;
  int st = WeightedStressMask(m, selected_label, air_label, selected_group, air_group, &mask);
  if (st != FEMM_OK)
    return st;
  if (mask_out)
    std::copy(mask.begin(), mask.end(), mask_out);
  const size_t nt = m.tri.size();
  std::vector<std::vector<int32_t>> inc(m.nodes.size());
  double fx = 0.0, fy = 0.0, tq = 0.0;
  for (size_t e = 0; e < nt; ++e) {
    const auto &t = m.tri[e];
    for (int k = 0; k < 3; ++k)
      inc[t[k]].push_back(static_cast<int32_t>(e));
    const Node p[3] = {m.nodes[t[0]], m.nodes[t[1]], m.nodes[t[2]]};
    Terms tm;
    if (!ElementTerms(p, &tm))
      return FEMM_MESH_INVALID;
    double hx = 0.0, hy = 0.0;
    for (int k = 0; k < 3; ++k) {
      hx += mask[t[k]] * tm.by[k];
      hy -= mask[t[k]] * tm.bx[k];
    }
    const double w = tm.area * m.depth;
    const double cx = (p[0].x + p[1].x + p[2].x) / 3.0, cy = (p[0].y + p[1].y + p[2].y) / 3.0;
    const double ex = bx[e], ey = by[e];
    const double fxd = ((ex * ex - ey * ey) * hx + 2.0 * ex * ey * hy) / (2.0 * kMu0);
    const double fyd = (2.0 * ex * ey * hx + (ey * ey - ex * ex) * hy) / (2.0 * kMu0);
    fx += w * fxd;
    fy += w * fyd;
    tq += w * (cx * fyd - cy * fxd);
  }
  out[0] = fx;
  out[1] = fy;
  out[2] = tq;
  const std::vector<double> av(a, a + m.nodes.size());
  for (int s = 0; s < angle_count; ++s) {
    const double c = std::cos(angles_rad[s]), sn = std::sin(angles_rad[s]);
    const double x = radius_m * c, y = radius_m * sn;
    int32_t found = -1;
    double lam[3] = {};
    for (size_t e = 0; e < nt && found < 0; ++e) {
      const auto &t = m.tri[e];
      const Node p[3] = {m.nodes[t[0]], m.nodes[t[1]], m.nodes[t[2]]};
      double l[3];
      if (Barycentric(p, x, y, l) && l[0] >= -1e-12 && l[1] >= -1e-12 && l[2] >= -1e-12) {
        found = static_cast<int32_t>(e);
        std::copy(l, l + 3, lam);
      }
    }
    if (found < 0 || !Selected(m, m.tri_mat[found], air_label, air_group))
      return FEMM_INVALID_ARGUMENT;
    double nbx[3], nby[3];
    if (!SmoothedB(m, av, bx, by, inc, found, nbx, nby))
      return FEMM_MESH_INVALID;
    double px = 0.0, py = 0.0;
    for (int k = 0; k < 3; ++k) {
      px += lam[k] * nbx[k];
      py += lam[k] * nby[k];
    }
    radial_b[s] = px * c + py * sn;
  }
  if (!std::isfinite(fx) || !std::isfinite(fy) || !std::isfinite(tq))
    return FEMM_NUMERICAL_NONFINITE;
  return FEMM_OK;
}

}  // extern "C"
