// FEMM's native air-gap element on the host (see age.hpp for the derivation).

#include "age.hpp"

#include <cmath>

namespace femm
{

namespace
{
constexpr double kPi = 3.141592653589793238462643383279502884;
}  // namespace

void CatmullRomWeights(double t, double w[4])
{
  const double t2 = t * t, t3 = t2 * t;
  w[0] = 0.5 * (-t + 2.0 * t2 - t3);
  w[1] = 0.5 * (2.0 - 5.0 * t2 + 3.0 * t3);
  w[2] = 0.5 * (t + 4.0 * t2 - 3.0 * t3);
  w[3] = 0.5 * (t3 - t2);
}

bool AgeCellMatrix(double K, double Ki, double ci, double co, double matrix[10][10])
{
  if (!(K > 0.0) || !(Ki > 0.0) || !std::isfinite(K) || !std::isfinite(Ki) ||
      !std::isfinite(ci) || !std::isfinite(co))
    return false;
  // Corner interpolation T[corner][node]; corners (u, v) = (0,0), (1,0), (0,1), (1,1).
  double T[4][10] = {};
  double wi[4], wo[4];
  CatmullRomWeights(1.0 - ci, wi);
  CatmullRomWeights(1.0 - co, wo);
  for (int q = 0; q < 4; ++q) {
    T[0][q] = wi[q];       // inner ring, u = 0: nodes 0..3
    T[1][q + 1] = wi[q];   // inner ring, u = 1: nodes 1..4
    T[2][q + 5] = wo[q];   // outer ring, u = 0: nodes 5..8
    T[3][q + 6] = wo[q];   // outer ring, u = 1: nodes 6..9
  }
  // Rectangle stiffness S = K A(u) M(v) + Ki M(u) A(v).
  const double A[2][2] = {{1.0, -1.0}, {-1.0, 1.0}};
  const double M[2][2] = {{5.0 / 12.0, 1.0 / 12.0}, {1.0 / 12.0, 5.0 / 12.0}};
  double S[4][4];
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) {
      const int ua = a & 1, va = a >> 1, ub = b & 1, vb = b >> 1;
      S[a][b] = K * A[ua][ub] * M[va][vb] + Ki * M[ua][ub] * A[va][vb];
    }
  double ST[4][10];
  for (int a = 0; a < 4; ++a)
    for (int j = 0; j < 10; ++j) {
      double s = 0.0;
      for (int b = 0; b < 4; ++b)
        s += S[a][b] * T[b][j];
      ST[a][j] = s;
    }
  for (int i = 0; i < 10; ++i)
    for (int j = i; j < 10; ++j) {
      double s = 0.0;
      for (int a = 0; a < 4; ++a)
        s += T[a][i] * ST[a][j];
      matrix[i][j] = s;
      matrix[j][i] = s;
    }
  return true;
}

void AgeAssemblyParameters(const femm_air_gap_t &age, double *K, double *Ki, double *ci,
                           double *co)
{
  const double dt = (kPi / 180.0) * age.arc_length_deg / static_cast<double>(age.cells);
  *K = 2.0 * (age.outer_radius_m - age.inner_radius_m) /
       (dt * (age.outer_radius_m + age.inner_radius_m));
  *Ki = 1.0 / *K;
  double i = age.inner_shift, o = age.outer_shift;
  if (i > o) {
    i -= o;
    o = 0.0;
  } else {
    i = 1.0 - o + i;
    o = 1.0;
  }
  *ci = i;
  *co = o;
}

bool AgeStencil(const femm_air_gap_t &age, int32_t k, int32_t node[10], double weight[10])
{
  const int32_t cells = age.cells;
  if (cells < 2 || k < 0 || k >= cells || age.points == nullptr)
    return false;
  const femm_age_point_t &before = age.points[k == 0 ? cells - 1 : k - 1];
  const femm_age_point_t &here = age.points[k];
  const femm_age_point_t &after = age.points[k + 1];
  const femm_age_point_t &after2 = age.points[k + 2 > cells ? 1 : k + 2];
  const femm_age_point_t *rec[5] = {&before, &here, &here, &after, &after2};
  // inner ring: records' (node 0, node 0, node 1, node 1, node 1); outer: (2, 2, 3, 3, 3)
  const int slot[5] = {0, 0, 1, 1, 1};
  for (int q = 0; q < 5; ++q) {
    node[q] = rec[q]->node[slot[q]];
    weight[q] = rec[q]->weight[slot[q]];
    node[q + 5] = rec[q]->node[slot[q] + 2];
    weight[q + 5] = rec[q]->weight[slot[q] + 2];
  }
  if (age.antiperiodic && k == 0) {
    weight[0] = -weight[0];
    weight[5] = -weight[5];
  }
  if (age.antiperiodic && k + 1 == cells) {
    weight[4] = -weight[4];
    weight[9] = -weight[9];
  }
  return true;
}

int ValidateAirGaps(const femm_model_t *m)
{
  if (m->air_gap_count < 0 || (m->air_gap_count > 0 && m->air_gaps == nullptr))
    return FEMM_INVALID_ARGUMENT;
  for (int32_t g = 0; g < m->air_gap_count; ++g) {
    const femm_air_gap_t &age = m->air_gaps[g];
    if (m->problem_type != FEMM_PLANAR)
      return FEMM_UNSUPPORTED_FEATURE;
    if (age.cells < 2 || age.points == nullptr || !(age.inner_radius_m > 0.0) ||
        !(age.outer_radius_m > age.inner_radius_m) || !(age.arc_length_deg > 0.0) ||
        !std::isfinite(age.center_xy[0]) || !std::isfinite(age.center_xy[1]) ||
        !std::isfinite(age.inner_shift) || !std::isfinite(age.outer_shift))
      return FEMM_MESH_INVALID;
    for (int32_t k = 0; k <= age.cells; ++k)
      for (int q = 0; q < 4; ++q) {
        const int32_t v = age.points[k].node[q];
        if (v < 0 || v >= m->node_count || !std::isfinite(age.points[k].weight[q]))
          return FEMM_MESH_INVALID;
      }
    const double dt = age.arc_length_deg / static_cast<double>(age.cells);
    if (!(dt > 0.0) || !std::isfinite(dt))
      return FEMM_MESH_INVALID;
  }
  return FEMM_OK;
}

}  // namespace femm
