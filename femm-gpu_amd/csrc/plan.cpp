// Host plan construction for the MI355X solve path.  See plan.hpp.

#include "plan.hpp"

#include "age.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

namespace femm
{
namespace
{

constexpr double kPi = 3.141592653589793238462643383279502884;
constexpr double kTwoPi = 6.283185307179586476925286766559005768;

// Seven-point Dunavant rule of the reference's r-z element (axisymmetric_p1.h:45-58).
constexpr double kBary[7][3] = {
    {1.0 / 3.0, 1.0 / 3.0, 1.0 / 3.0},
    {0.059715871789770, 0.470142064105115, 0.470142064105115},
    {0.470142064105115, 0.059715871789770, 0.470142064105115},
    {0.470142064105115, 0.470142064105115, 0.059715871789770},
    {0.797426985353087, 0.101286507323456, 0.101286507323456},
    {0.101286507323456, 0.797426985353087, 0.101286507323456},
    {0.101286507323456, 0.101286507323456, 0.797426985353087},
};
constexpr double kWeight[7] = {0.225,
                               0.132394152788506, 0.132394152788506, 0.132394152788506,
                               0.125939180544827, 0.125939180544827, 0.125939180544827};

// Curl basis of one P1 triangle at its quadrature points.  Planar: one point,
// b_x = dN/dy, b_y = -dN/dx, weight = area (nonlinear_model.cpp:207-221).  r-z:
// Br = -dN/dz, Bz = dN/dr + N/r, weight = 2 pi r w area (axisymmetric_p1.h:26-77).
struct ElementBasis {
  int q = 1;
  double area = 0.0;
  double w[kMaxQuadrature] = {}, gx[kMaxQuadrature][3] = {}, gy[kMaxQuadrature][3] = {},
         shape[kMaxQuadrature][3] = {};
};

bool Basis(const double x[3], const double y[3], bool planar, ElementBasis *e)
{
  const double det = (x[1] - x[0]) * (y[2] - y[0]) - (x[2] - x[0]) * (y[1] - y[0]);
  if (!(det > 0.0) || !std::isfinite(det))
    return false;
  e->area = 0.5 * det;
  if (planar) {
    const double dx[3] = {y[1] - y[2], y[2] - y[0], y[0] - y[1]};
    const double dy[3] = {x[2] - x[1], x[0] - x[2], x[1] - x[0]};
    e->q = 1;
    e->w[0] = e->area;
    for (int k = 0; k < 3; ++k) {
      e->gx[0][k] = dy[k] / det;
      e->gy[0][k] = -dx[k] / det;
      e->shape[0][k] = 1.0 / 3.0;
    }
    return true;
  }
  const double dndr[3] = {(y[1] - y[2]) / det, (y[2] - y[0]) / det, (y[0] - y[1]) / det};
  const double dndz[3] = {(x[2] - x[1]) / det, (x[0] - x[2]) / det, (x[1] - x[0]) / det};
  e->q = 7;
  for (int p = 0; p < 7; ++p) {
    double r = 0.0;
    for (int k = 0; k < 3; ++k) {
      e->shape[p][k] = kBary[p][k];
      r += kBary[p][k] * x[k];
    }
    if (!(r > 0.0) || !std::isfinite(r))
      return false;
    e->w[p] = kTwoPi * r * (e->area * kWeight[p]);
    for (int k = 0; k < 3; ++k) {
      e->gx[p][k] = -dndz[k];
      e->gy[p][k] = dndr[k] + kBary[p][k] / r;
    }
  }
  return true;
}

// Signed DOF map as a weighted union-find.  The reference's DFS
// (signed_node_constraints.h:33-106) roots every component at its smallest
// node; a union by smallest index reproduces that root, and the parity
// relative to the root is path-independent when the constraints are consistent.
struct SignedUnionFind {
  std::vector<int32_t> parent;
  std::vector<int8_t> parity;  // sign of node relative to parent
  explicit SignedUnionFind(size_t n) : parent(n), parity(n, 1)
  {
    std::iota(parent.begin(), parent.end(), 0);
  }
  std::pair<int32_t, int8_t> Find(int32_t v)
  {
    int8_t sign = 1;
    int32_t r = v;
    while (parent[r] != r) {
      sign = static_cast<int8_t>(sign * parity[r]);
      r = parent[r];
    }
    // path compression with parity fix-up
    int8_t s = sign;
    while (parent[v] != v) {
      const int32_t next = parent[v];
      const int8_t p = parity[v];
      parent[v] = r;
      parity[v] = s;
      s = static_cast<int8_t>(s * p);
      v = next;
    }
    return {r, sign};
  }
  // A(b) = sign * A(a).  Returns false on a contradictory cycle.
  bool Unite(int32_t a, int32_t b, int8_t sign)
  {
    auto [ra, sa] = Find(a);
    auto [rb, sb] = Find(b);
    if (ra == rb)
      return static_cast<int8_t>(sa * sb) == sign;
    // A(b) = sb A(rb), A(a) = sa A(ra) -> A(rb) = sb * sign * sa * A(ra)
    const int8_t rel = static_cast<int8_t>(sb * sign * sa);
    if (ra < rb) {
      parent[rb] = ra;
      parity[rb] = rel;
    } else {
      parent[ra] = rb;
      parity[ra] = rel;
    }
    return true;
  }
};

}  // namespace

bool BuildSpline(Spline *s)
{
  const int n = static_cast<int>(s->b.size());
  if (n < 2 || s->h.size() != s->b.size() || s->b[0] != 0.0 || s->h[0] != 0.0)
    return false;
  if (!(s->cold > 0.0))
    s->cold = s->h[1] / s->b[1];
  for (int i = 1; i < n; ++i)
    if (!(s->b[i] > s->b[i - 1]) || s->h[i] < 0.0 || !std::isfinite(s->b[i]) ||
        !std::isfinite(s->h[i]))
      return false;
  // Tridiagonal natural-spline system; the matrix is diagonally dominant, so
  // elimination without row exchanges performs the same arithmetic as the
  // reference's dense partial-pivoting elimination.
  std::vector<double> lo(n, 0.0), di(n, 0.0), up(n, 0.0), r(n, 0.0);
  const double l0 = s->b[1] - s->b[0], ln = s->b[n - 1] - s->b[n - 2];
  di[0] = 4.0 / l0;
  up[0] = 2.0 / l0;
  r[0] = 6.0 * (s->h[1] - s->h[0]) / (l0 * l0);
  di[n - 1] = 4.0 / ln;
  lo[n - 1] = 2.0 / ln;
  r[n - 1] = 6.0 * (s->h[n - 1] - s->h[n - 2]) / (ln * ln);
  for (int i = 1; i < n - 1; ++i) {
    const double a = s->b[i] - s->b[i - 1], c = s->b[i + 1] - s->b[i];
    lo[i] = 2.0 / a;
    di[i] = 4.0 * (a + c) / (a * c);
    up[i] = 2.0 / c;
    r[i] = 6.0 * (s->h[i] - s->h[i - 1]) / (a * a) + 6.0 * (s->h[i + 1] - s->h[i]) / (c * c);
  }
  for (int k = 0; k + 1 < n; ++k) {
    if (!(std::abs(di[k]) >= std::abs(lo[k + 1])) || !(std::abs(di[k]) > 0.0))
      return false;   // would need a row exchange: outside the FEMM table domain
    const double f = lo[k + 1] / di[k];
    di[k + 1] -= f * up[k];
    r[k + 1] -= f * r[k];
  }
  if (!(std::abs(di[n - 1]) > 0.0))
    return false;
  s->slope.assign(n, 0.0);
  for (int k = n - 1; k >= 0; --k) {
    double v = r[k];
    if (k + 1 < n)
      v -= up[k] * s->slope[k + 1];
    s->slope[k] = v / di[k];
    if (!std::isfinite(s->slope[k]))
      return false;
  }
  // Interior derivative roots of each Hermite segment trigger FEMM's 3-point smoothing.
  for (int i = 1; i < n; ++i) {
    const double L = s->b[i] - s->b[i - 1];
    const double d0 = s->slope[i - 1], d1 = s->slope[i], h0 = s->h[i - 1], h1 = s->h[i];
    const double c1 = -2.0 * (2.0 * d0 * L + d1 * L + 3.0 * h0 - 3.0 * h1) / (L * L);
    const double c2 = 3.0 * (d0 * L + d1 * L + 2.0 * h0 - 2.0 * h1) / (L * L * L);
    double r0 = -1.0, r1 = -1.0;
    const double disc = c1 * c1 - 4.0 * d0 * c2;
    if (c2 == 0.0) {
      if (c1 != 0.0)
        r0 = -d0 / c1;
    } else if (disc > 0.0) {
      const double q = std::sqrt(disc);
      r0 = -(c1 + q) / (2.0 * c2);
      r1 = (-c1 + q) / (2.0 * c2);
    }
    if ((r0 >= 0.0 && r0 <= L) || (r1 >= 0.0 && r1 <= L)) {
      if (++s->passes > 1024)
        return false;
      std::vector<double> nb = s->b, nh = s->h;
      for (int k = 1; k < n - 1; ++k) {
        nb[k] = (s->b[k - 1] + s->b[k] + s->b[k + 1]) / 3.0;
        nh[k] = (s->h[k - 1] + s->h[k] + s->h[k + 1]) / 3.0;
      }
      s->b.swap(nb);
      s->h.swap(nh);
      return BuildSpline(s);
    }
  }
  return true;
}

bool SameTopology(const femm_model_t *a, const femm_model_t *b)
{
  if (a == nullptr || b == nullptr)
    return false;
  if (a->problem_type != b->problem_type || a->node_count != b->node_count ||
      a->triangle_count != b->triangle_count || a->material_count != b->material_count ||
      a->curve_count != b->curve_count || a->dirichlet_count != b->dirichlet_count ||
      a->constraint_count != b->constraint_count || a->circuit_count != b->circuit_count ||
      !(a->depth_m == b->depth_m) || a->node_count <= 0 || a->triangle_count <= 0)
    return false;
  auto same = [](const void *x, const void *y, size_t bytes) {
    return bytes == 0 || (x != nullptr && y != nullptr && std::memcmp(x, y, bytes) == 0);
  };
  const size_t N = static_cast<size_t>(a->node_count), T = static_cast<size_t>(a->triangle_count);
  if (a->air_gap_count != b->air_gap_count || a->air_gap_count < 0)
    return false;
  for (int32_t g = 0; g < a->air_gap_count; ++g) {
    const femm_air_gap_t &x = a->air_gaps[g], &y = b->air_gaps[g];
    if (x.antiperiodic != y.antiperiodic || x.cells != y.cells ||
        !(x.center_xy[0] == y.center_xy[0]) || !(x.center_xy[1] == y.center_xy[1]) ||
        !(x.inner_radius_m == y.inner_radius_m) || !(x.outer_radius_m == y.outer_radius_m) ||
        !(x.arc_length_deg == y.arc_length_deg) || !(x.inner_shift == y.inner_shift) ||
        !(x.outer_shift == y.outer_shift) || x.cells < 0 ||
        !same(x.points, y.points, sizeof(femm_age_point_t) * (static_cast<size_t>(x.cells) + 1)))
      return false;
  }
  return same(a->node_xy, b->node_xy, sizeof(double) * 2 * N) &&
         same(a->triangle_nodes, b->triangle_nodes, sizeof(int32_t) * 3 * T) &&
         same(a->curve_offsets, b->curve_offsets,
              a->curve_count > 0 ? sizeof(int32_t) * (a->curve_count + 1) : 0) &&
         same(a->dirichlet_nodes, b->dirichlet_nodes, sizeof(int32_t) * a->dirichlet_count) &&
         same(a->dirichlet_values, b->dirichlet_values, sizeof(double) * a->dirichlet_count) &&
         same(a->constraints, b->constraints,
              sizeof(femm_node_constraint_t) * a->constraint_count);
}

// Fixed AGE stiffness per CSR slot and fixed-node load per row, accumulated element by
// element, cell by cell, i then j (nonlinear_topology.cpp:180-232); also keeps a copy of the
// elements for the postprocess.
static int BuildAirGapTerms(const femm_model_t *m, HostPlan *p)
{
  p->age_vals.assign(p->nnz, 0.0);
  p->age_rhs.assign(p->n, 0.0);
  auto slot_of = [p](int32_t row, int32_t col) {
    const int32_t *lo = p->cols.data() + p->row_offsets[row];
    const int32_t *hi = p->cols.data() + p->row_offsets[row + 1];
    return static_cast<int32_t>(std::lower_bound(lo, hi, col) - p->cols.data());
  };
  size_t total_points = 0;
  for (int32_t g = 0; g < m->air_gap_count; ++g)
    total_points += static_cast<size_t>(m->air_gaps[g].cells) + 1;
  p->ages.assign(m->air_gaps, m->air_gaps + m->air_gap_count);
  p->age_points.clear();
  p->age_points.reserve(total_points);
  for (int32_t g = 0; g < m->air_gap_count; ++g) {
    const femm_air_gap_t &age = m->air_gaps[g];
    p->age_points.insert(p->age_points.end(), age.points, age.points + age.cells + 1);
    double K, Ki, ci, co, matrix[10][10];
    AgeAssemblyParameters(age, &K, &Ki, &ci, &co);
    if (!AgeCellMatrix(K, Ki, ci, co, matrix))
      return FEMM_MESH_INVALID;
    for (int32_t k = 0; k < age.cells; ++k) {
      int32_t node[10];
      double weight[10];
      AgeStencil(age, k, node, weight);
      for (int i = 0; i < 10; ++i) {
        const int32_t row = p->node_to_free[node[i]];
        if (row < 0)
          continue;
        const double si = static_cast<double>(p->node_sign[node[i]]);
        for (int j = 0; j < 10; ++j) {
          const double value = matrix[i][j] * weight[i] * weight[j] * kAgeReluctivity;
          const int32_t col = p->node_to_free[node[j]];
          if (col >= 0)
            p->age_vals[slot_of(row, col)] +=
                static_cast<double>(p->node_sign[node[i]] * p->node_sign[node[j]]) * value;
          else
            p->age_rhs[row] -= si * value * p->boundary[node[j]];
        }
      }
    }
  }
  size_t at = 0;
  for (femm_air_gap_t &age : p->ages) {
    age.points = p->age_points.data() + at;
    at += static_cast<size_t>(age.cells) + 1;
  }
  return FEMM_OK;
}

int BuildHostPlan(const femm_model_t *m, HostPlan *p, std::string *error, const HostPlan *topology,
                  bool validate_only)
{
  auto fail = [error](int code, const char *why) {
    if (error)
      *error = why;
    return code;
  };
  if (m == nullptr || p == nullptr)
    return fail(FEMM_INVALID_ARGUMENT, "null model");
  *p = HostPlan{};
  const bool planar = m->problem_type == FEMM_PLANAR;
  if (m->problem_type != FEMM_PLANAR && m->problem_type != FEMM_AXISYMMETRIC)
    return fail(FEMM_UNSUPPORTED_FEATURE, "problem type");
  // Validation order and codes follow ValidateNonlinearModel (nonlinear_model.cpp:223-319).
  if (m->node_count <= 0 || m->triangle_count <= 0 || m->material_count <= 0 ||
      (planar && (!(m->depth_m > 0.0) || !std::isfinite(m->depth_m))) || m->circuit_count < 0 ||
      m->dirichlet_count <= 0)
    return fail(FEMM_INVALID_ARGUMENT, "empty model, bad depth or no Dirichlet nodes");
  const int32_t N = m->node_count, T = m->triangle_count, C = m->circuit_count;
  std::vector<char> fixed_literal(N, 0);
  for (int i = 0; i < m->dirichlet_count; ++i) {
    const int32_t v = m->dirichlet_nodes[i];
    if (v < 0 || v >= N || fixed_literal[v] || !std::isfinite(m->dirichlet_values[i]))
      return fail(FEMM_BOUNDARY_INVALID, "bad Dirichlet node");
    fixed_literal[v] = 1;
  }
  if (m->dirichlet_count == N)
    return fail(FEMM_BOUNDARY_INVALID, "every node fixed");
  for (int32_t v = 0; v < N; ++v) {
    const double x = m->node_xy[2 * v], y = m->node_xy[2 * v + 1];
    if (!std::isfinite(x) || !std::isfinite(y) || (!planar && x < 0.0))
      return fail(FEMM_MESH_INVALID, "bad node coordinate");
  }
  std::vector<Spline> splines(m->curve_count);
  for (int c = 0; c < m->curve_count; ++c) {
    const int lo = m->curve_offsets[c], hi = m->curve_offsets[c + 1];
    splines[c].b.assign(m->curve_b + lo, m->curve_b + hi);
    splines[c].h.assign(m->curve_h + lo, m->curve_h + hi);
    if (!BuildSpline(&splines[c]))
      return fail(FEMM_INVALID_MATERIAL, "invalid B-H curve");
  }
  for (int k = 0; k < m->material_count; ++k) {
    const femm_material_t &mt = m->materials[k];
    if (!(mt.reluctivity_zero > 0.0) || mt.alpha < 0.0 || !std::isfinite(mt.reluctivity_zero) ||
        !std::isfinite(mt.alpha) || !std::isfinite(mt.source_j_per_a) || !std::isfinite(mt.hc) ||
        !std::isfinite(mt.magnetization_deg))
      return fail(FEMM_INVALID_MATERIAL, "material value");
    if (mt.circuit < -1 || mt.circuit >= C || (mt.source_j_per_a != 0.0 && mt.circuit < 0))
      return fail(FEMM_INVALID_MATERIAL, "material circuit");
    if (mt.bh_curve < -1 || mt.bh_curve >= m->curve_count)
      return fail(FEMM_INVALID_MATERIAL, "material curve");
  }
  std::vector<ElementBasis> basis(validate_only ? 0 : T);   // a check keeps no geometry
  ElementBasis scratch_basis;
  for (int32_t e = 0; e < T; ++e) {
    const int32_t mat = m->triangle_material[e];
    if (mat < 0 || mat >= m->material_count)
      return fail(FEMM_INVALID_MATERIAL, "triangle material");
    double x[3], y[3];
    for (int k = 0; k < 3; ++k) {
      const int32_t v = m->triangle_nodes[3 * e + k];
      if (v < 0 || v >= N)
        return fail(FEMM_MESH_INVALID, "triangle node");
      x[k] = m->node_xy[2 * v];
      y[k] = m->node_xy[2 * v + 1];
    }
    // The planar determinant test of the reference applies to both problem types.
    const double det = (x[1] - x[0]) * (y[2] - y[0]) - (x[2] - x[0]) * (y[1] - y[0]);
    if (!(det > 0.0) || !std::isfinite(det))
      return fail(FEMM_MESH_INVALID, "triangle orientation");
    if (!Basis(x, y, true, validate_only ? &scratch_basis : &basis[e]))
      return fail(FEMM_MESH_INVALID, "triangle basis");
  }
  const int age_status = ValidateAirGaps(m);
  if (age_status != FEMM_OK)
    return fail(age_status, "invalid air-gap element");
  if (topology != nullptr) {
    // Variant of an existing plan: same constraints and Dirichlet data, same DOF map.
    p->node_sign = topology->node_sign;
    p->node_to_free = topology->node_to_free;
    p->free_nodes = topology->free_nodes;
    p->boundary = topology->boundary;
    p->any_boundary_value = topology->any_boundary_value;
  } else {
    // Signed DOF map.
    SignedUnionFind uf(N);
    for (int k = 0; k < m->constraint_count; ++k) {
      const femm_node_constraint_t &c = m->constraints[k];
      if (c.node_a < 0 || c.node_b < 0 || c.node_a >= N || c.node_b >= N ||
          (c.sign != 1 && c.sign != -1) || c.node_a == c.node_b)
        return fail(FEMM_BOUNDARY_INVALID, "invalid node constraint");
      if (!uf.Unite(c.node_a, c.node_b, static_cast<int8_t>(c.sign)))
        return fail(FEMM_BOUNDARY_INVALID, "contradictory constraint cycle");
    }
    std::vector<int32_t> root(N);
    p->node_sign.assign(N, 1);
    for (int32_t v = 0; v < N; ++v) {
      auto rs = uf.Find(v);
      root[v] = rs.first;
      p->node_sign[v] = rs.second;
    }
    std::vector<char> root_fixed(N, 0);
    std::vector<double> root_value(N, 0.0);
    for (int i = 0; i < m->dirichlet_count; ++i) {
      const int32_t v = m->dirichlet_nodes[i];
      const double rv = static_cast<double>(p->node_sign[v]) * m->dirichlet_values[i];
      if (root_fixed[root[v]] && root_value[root[v]] != rv)
        return fail(FEMM_BOUNDARY_INVALID, "inconsistent fixed value");
      root_fixed[root[v]] = 1;
      root_value[root[v]] = rv;
    }
    std::vector<int32_t> root_to_free(N, -1);
    for (int32_t v = 0; v < N; ++v)
      if (root[v] == v && !root_fixed[v]) {
        root_to_free[v] = static_cast<int32_t>(p->free_nodes.size());
        p->free_nodes.push_back(v);
      }
    if (p->free_nodes.empty())
      return fail(FEMM_BOUNDARY_INVALID, "no free DOF");
    p->node_to_free.assign(N, -1);
    p->boundary.assign(N, 0.0);
    for (int32_t v = 0; v < N; ++v) {
      p->node_to_free[v] = root_to_free[root[v]];
      if (p->node_to_free[v] < 0) {
        p->boundary[v] = static_cast<double>(p->node_sign[v]) * root_value[root[v]];
        p->any_boundary_value = p->any_boundary_value || p->boundary[v] != 0.0;
      }
    }
    if (!planar) {
      double scale = 1e-3;
      for (int32_t v = 0; v < N; ++v)
        scale = std::max(scale, m->node_xy[2 * v]);
      for (int32_t v = 0; v < N; ++v)
        if (m->node_xy[2 * v] <= 1e-12 * scale &&
            (p->node_to_free[v] >= 0 || p->boundary[v] != 0.0))
          return fail(FEMM_BOUNDARY_INVALID, "axis node not fixed to zero");
    }
  }
  const int32_t n = static_cast<int32_t>(p->free_nodes.size());
  p->N = N;
  p->T = T;
  p->n = n;
  p->C = C;
  if (validate_only) {
    // The one check of the geometry build below: the r-z element basis.
    for (int32_t e = 0; !planar && e < T; ++e) {
      double x[3], y[3];
      for (int k = 0; k < 3; ++k) {
        x[k] = m->node_xy[2 * m->triangle_nodes[3 * e + k]];
        y[k] = m->node_xy[2 * m->triangle_nodes[3 * e + k] + 1];
      }
      if (!Basis(x, y, false, &scratch_basis))
        return fail(FEMM_MESH_INVALID, "r-z basis");
    }
    return FEMM_OK;
  }
  p->planar = planar ? 1 : 0;
  p->Q = planar ? 1 : 7;
  p->depth = planar ? m->depth_m : 1.0;
  p->dirichlet_nodes.assign(m->dirichlet_nodes, m->dirichlet_nodes + m->dirichlet_count);

  // Materials and curves.
  p->mat_nu0.resize(m->material_count);
  p->mat_alpha.resize(m->material_count);
  p->mat_cold.resize(m->material_count);
  p->mat_curve_off.resize(m->material_count);
  p->mat_curve_cnt.resize(m->material_count);
  std::vector<int32_t> curve_offset(m->curve_count);
  for (int c = 0; c < m->curve_count; ++c) {
    curve_offset[c] = static_cast<int32_t>(p->curve_b.size());
    p->curve_b.insert(p->curve_b.end(), splines[c].b.begin(), splines[c].b.end());
    p->curve_h.insert(p->curve_h.end(), splines[c].h.begin(), splines[c].h.end());
    p->curve_s.insert(p->curve_s.end(), splines[c].slope.begin(), splines[c].slope.end());
  }
  for (int k = 0; k < m->material_count; ++k) {
    const femm_material_t &mt = m->materials[k];
    p->mat_nu0[k] = mt.reluctivity_zero;
    p->mat_alpha[k] = mt.alpha;
    p->mat_curve_off[k] = mt.bh_curve >= 0 ? curve_offset[mt.bh_curve] : -1;
    p->mat_curve_cnt[k] = mt.bh_curve >= 0 ? static_cast<int32_t>(splines[mt.bh_curve].b.size()) : 0;
    p->mat_cold[k] = mt.bh_curve >= 0 ? splines[mt.bh_curve].cold : 0.0;
    p->any_nonlinear = p->any_nonlinear || mt.bh_curve >= 0 || mt.alpha != 0.0;
  }
  if (p->curve_b.empty()) {
    p->curve_b.push_back(0.0);
    p->curve_h.push_back(0.0);
    p->curve_s.push_back(0.0);
  }

  // Element geometry, sources and flux weights.
  const int Q = p->Q;
  p->tri_nodes.assign(m->triangle_nodes, m->triangle_nodes + 3 * static_cast<size_t>(T));
  p->tri_mat.assign(m->triangle_material, m->triangle_material + T);
  p->tri_w.assign(static_cast<size_t>(T) * Q, 0.0);
  p->tri_gx.assign(static_cast<size_t>(T) * Q * 3, 0.0);
  p->tri_gy.assign(static_cast<size_t>(T) * Q * 3, 0.0);
  p->tri_flux.assign(static_cast<size_t>(T) * 3, 0.0);
  p->tri_circuit.assign(T, -1);
  p->pm_rhs.assign(n, 0.0);
  p->circuit_rhs.assign(static_cast<size_t>(std::max(C, 1)) * n, 0.0);
  for (int32_t e = 0; e < T; ++e) {
    const int32_t *tn = m->triangle_nodes + 3 * e;
    ElementBasis eb;
    if (!planar) {
      double x[3], y[3];
      for (int k = 0; k < 3; ++k) {
        x[k] = m->node_xy[2 * tn[k]];
        y[k] = m->node_xy[2 * tn[k] + 1];
      }
      if (!Basis(x, y, false, &eb))
        return fail(FEMM_MESH_INVALID, "r-z basis");
    } else {
      eb = basis[e];
    }
    for (int q = 0; q < Q; ++q) {
      p->tri_w[static_cast<size_t>(e) * Q + q] = eb.w[q];
      for (int k = 0; k < 3; ++k) {
        p->tri_gx[(static_cast<size_t>(e) * Q + q) * 3 + k] = eb.gx[q][k];
        p->tri_gy[(static_cast<size_t>(e) * Q + q) * 3 + k] = eb.gy[q][k];
      }
    }
    const femm_material_t &mt = m->materials[p->tri_mat[e]];
    const double th = mt.magnetization_deg * kPi / 180.0;
    const double hcx = mt.hc * std::cos(th), hcy = mt.hc * std::sin(th);
    p->tri_circuit[e] = mt.circuit;
    for (int i = 0; i < 3; ++i) {
      double pm = 0.0, coil = 0.0;
      for (int q = 0; q < Q; ++q) {
        pm += eb.w[q] * (hcx * eb.gx[q][i] + hcy * eb.gy[q][i]);
        coil += eb.w[q] * mt.source_j_per_a * eb.shape[q][i];
      }
      if (mt.circuit >= 0)
        p->tri_flux[3 * static_cast<size_t>(e) + i] = coil;
      const int32_t row = p->node_to_free[tn[i]];
      if (row >= 0) {
        const double s = static_cast<double>(p->node_sign[tn[i]]);
        p->pm_rhs[row] += s * pm;
        if (mt.circuit >= 0)
          p->circuit_rhs[static_cast<size_t>(mt.circuit) * n + row] += s * coil;
      }
    }
  }

  if (topology != nullptr) {
    // A variant shares the base plan's air-gap elements and their fixed-node load.
    for (size_t r = 0; r < topology->age_rhs.size(); ++r)
      p->pm_rhs[r] = p->pm_rhs[r] + topology->age_rhs[r];
    return FEMM_OK;
  }

  // Symbolic CSR: sorted unique (row, col) among free roots of every triangle
  // coupling (the rule of nonlinear_topology.cpp:37-85), built by a stable sort of
  // the element/i/j emission stream so each slot's contributor list keeps that order.
  std::vector<uint64_t> keys;
  std::vector<int32_t> packed;
  keys.reserve(static_cast<size_t>(T) * 9);
  packed.reserve(static_cast<size_t>(T) * 9);
  std::vector<uint64_t> rkeys;
  std::vector<int32_t> rpacked;
  rkeys.reserve(static_cast<size_t>(T) * 3);
  rpacked.reserve(static_cast<size_t>(T) * 3);
  for (int32_t e = 0; e < T; ++e) {
    const int32_t *tn = m->triangle_nodes + 3 * e;
    for (int i = 0; i < 3; ++i) {
      const int32_t ri = p->node_to_free[tn[i]];
      if (ri < 0)
        continue;
      rkeys.push_back(static_cast<uint64_t>(ri));
      rpacked.push_back((e << 3) | (i << 1) | (p->node_sign[tn[i]] < 0 ? 1 : 0));
      for (int j = 0; j < 3; ++j) {
        const int32_t cj = p->node_to_free[tn[j]];
        if (cj < 0)
          continue;
        keys.push_back((static_cast<uint64_t>(ri) << 32) | static_cast<uint32_t>(cj));
        const int neg = p->node_sign[tn[i]] * p->node_sign[tn[j]] < 0 ? 1 : 0;
        packed.push_back((e << 5) | (i << 3) | (j << 1) | neg);
      }
    }
  }
  // Air-gap element couplings join the pattern (nonlinear_topology.cpp:46-59) but carry no
  // triangle contributor (marker -1).
  for (int32_t g = 0; g < m->air_gap_count; ++g)
    for (int32_t k = 0; k < m->air_gaps[g].cells; ++k) {
      int32_t node[10];
      double weight[10];
      AgeStencil(m->air_gaps[g], k, node, weight);
      for (int i = 0; i < 10; ++i) {
        const int32_t ri = p->node_to_free[node[i]];
        if (ri < 0)
          continue;
        for (int j = 0; j < 10; ++j) {
          const int32_t cj = p->node_to_free[node[j]];
          if (cj < 0)
            continue;
          keys.push_back((static_cast<uint64_t>(ri) << 32) | static_cast<uint32_t>(cj));
          packed.push_back(-1);
        }
      }
    }
  std::vector<int32_t> order(keys.size());
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(),
                   [&keys](int32_t a, int32_t b) { return keys[a] < keys[b]; });
  p->row_offsets.assign(n + 1, 0);
  p->contrib.clear();
  p->contrib.reserve(static_cast<size_t>(T) * 9);
  p->slot_off.clear();
  uint64_t last = ~0ull;
  for (size_t k = 0; k < order.size(); ++k) {
    const uint64_t key = keys[order[k]];
    if (key != last) {
      p->slot_off.push_back(static_cast<int32_t>(p->contrib.size()));
      p->cols.push_back(static_cast<int32_t>(key & 0xffffffffu));
      p->row_offsets[(key >> 32) + 1]++;
      last = key;
    }
    if (packed[order[k]] >= 0)
      p->contrib.push_back(packed[order[k]]);
  }
  p->slot_off.push_back(static_cast<int32_t>(p->contrib.size()));
  for (int32_t r = 0; r < n; ++r)
    p->row_offsets[r + 1] += p->row_offsets[r];
  p->nnz = static_cast<int32_t>(p->cols.size());
  p->diag_slot.assign(n, -1);
  for (int32_t r = 0; r < n; ++r) {
    for (int32_t k = p->row_offsets[r]; k < p->row_offsets[r + 1]; ++k)
      if (p->cols[k] == r)
        p->diag_slot[r] = k;
    if (p->diag_slot[r] < 0)
      return fail(FEMM_ASSEMBLY_FAILED, "missing diagonal");
  }
  if (m->air_gap_count > 0) {
    const int age_built = BuildAirGapTerms(m, p);
    if (age_built != FEMM_OK)
      return fail(age_built, "air-gap element matrix");
    // The rhs gather starts from pm_rhs[row] + age_rhs[row] (nonlinear_assembly.cu:155).
    for (int32_t r = 0; r < n; ++r)
      p->pm_rhs[r] = p->pm_rhs[r] + p->age_rhs[r];
  }
  std::vector<int32_t> rorder(rkeys.size());
  std::iota(rorder.begin(), rorder.end(), 0);
  std::stable_sort(rorder.begin(), rorder.end(),
                   [&rkeys](int32_t a, int32_t b) { return rkeys[a] < rkeys[b]; });
  p->row_off.assign(n + 1, 0);
  p->rcontrib.resize(rorder.size());
  for (size_t k = 0; k < rorder.size(); ++k) {
    p->rcontrib[k] = rpacked[rorder[k]];
    p->row_off[rkeys[rorder[k]] + 1]++;
  }
  for (int32_t r = 0; r < n; ++r)
    p->row_off[r + 1] += p->row_off[r];

  // SELL-64: slice s holds rows 64s..64s+63 column-interleaved; width = longest row;
  // padding entries point at the row itself with a zero value.
  p->nslices = (n + kSliceRows - 1) / kSliceRows;
  p->slice_ptr.assign(p->nslices + 1, 0);
  p->slice_width.assign(p->nslices, 0);
  for (int32_t s = 0; s < p->nslices; ++s) {
    int32_t w = 0;
    for (int32_t r = s * kSliceRows; r < std::min(n, (s + 1) * kSliceRows); ++r)
      w = std::max(w, p->row_offsets[r + 1] - p->row_offsets[r]);
    p->slice_width[s] = w;
    p->slice_ptr[s + 1] = p->slice_ptr[s] + w * kSliceRows;
  }
  p->sell_size = p->slice_ptr[p->nslices];
  p->sell_cols.assign(p->sell_size, 0);
  p->slot_to_sell.assign(p->nnz, 0);
  p->diag_sell.assign(n, 0);
  for (int32_t s = 0; s < p->nslices; ++s)
    for (int32_t lane = 0; lane < kSliceRows; ++lane) {
      const int32_t r = s * kSliceRows + lane;
      const int32_t len = r < n ? p->row_offsets[r + 1] - p->row_offsets[r] : 0;
      for (int32_t k = 0; k < p->slice_width[s]; ++k) {
        const int32_t pos = p->slice_ptr[s] + k * kSliceRows + lane;
        if (k < len) {
          const int32_t slot = p->row_offsets[r] + k;
          p->sell_cols[pos] = p->cols[slot];
          p->slot_to_sell[slot] = pos;
          if (slot == p->diag_slot[r])
            p->diag_sell[r] = pos;
        } else {
          p->sell_cols[pos] = r < n ? r : 0;
        }
      }
    }
  return FEMM_OK;
}

}  // namespace femm
