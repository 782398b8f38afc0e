// d3q15.h -- D3Q15 lattice and the reference's mode-space collision, as
// constexpr tables and fully unrolled per-cell routines shared by host and
// device code.
//
// Model (otbrown/LAMBReX):
//   velocities             include/d3q15_bgk.h:14-22 (population order kept,
//                          so f_i arrays compare one to one)
//   mode matrix M          src/AmrSim.cpp:1037-1054
//   inverse MI             src/AmrSim.cpp:1056-1073
//   DELTA = I/15           src/AmrSim.cpp:1033-1035
//   collide                src/AmrSim.cpp:28-104
//   equilibrium            src/AmrSim.cpp:873-927
//   hydro (rho, u)         src/AmrSim.cpp:957-971
//
// The tables are not typed in: M is built from the velocity set and the
// mode definitions with exact rationals, and MI from the weighted
// orthogonality of the basis, MI[p][m] = w_p M[m][p] / sum_i w_i M[m][i]^2,
// each entry rounded once from its exact value (which is what the
// reference's a./b. literals are).
//
// Bit-exactness: the FP64 routines perform the reference's operations in the
// reference's order.  The only rewrites are the exact ones: terms with a
// zero coefficient are skipped (a running sum that starts at +0.0 can never
// be -0.0, so adding a signed zero never changes it for finite data), and
// multiplications by +1 / -1 become an add / subtract.  Build device code
// with -ffp-contract=off so no multiply-add is fused.
#pragma once

#include <cstdint>
#include <utility>

#if defined(__HIPCC__)
#define LBX_HD __host__ __device__ __forceinline__
#else
#define LBX_HD inline
#endif

namespace lbx {

constexpr int NV = 15;
constexpr int ND = 3;

constexpr int kC[NV][3] = {
    {0, 0, 0},
    {1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1},
    {1, 1, 1}, {1, 1, -1}, {1, -1, 1}, {1, -1, -1},
    {-1, 1, 1}, {-1, 1, -1}, {-1, -1, 1}, {-1, -1, -1},
};

// ---------------------------------------------------------------- rationals
struct Rat {
  long long n, d;
};
constexpr long long cgcd(long long a, long long b) {
  a = a < 0 ? -a : a;
  b = b < 0 ? -b : b;
  while (b) {
    long long t = a % b;
    a = b;
    b = t;
  }
  return a;
}
constexpr Rat rat(long long n, long long d) {
  if (d < 0) { n = -n; d = -d; }
  long long g = cgcd(n, d);
  if (g == 0) g = 1;
  return n == 0 ? Rat{0, 1} : Rat{n / g, d / g};
}
constexpr Rat operator+(Rat a, Rat b) { return rat(a.n * b.d + b.n * a.d, a.d * b.d); }
constexpr Rat operator*(Rat a, Rat b) { return rat(a.n * b.n, a.d * b.d); }
constexpr Rat operator/(Rat a, Rat b) { return rat(a.n * b.d, a.d * b.n); }
constexpr double to_double(Rat r) { return double(r.n) / double(r.d); }

constexpr int csq(int p) { return kC[p][0] * kC[p][0] + kC[p][1] * kC[p][1] + kC[p][2] * kC[p][2]; }

// Lattice weights: 2/9 rest, 1/9 axis, 1/72 diagonal.
constexpr Rat weight_rat(int p) { return csq(p) == 0 ? rat(2, 9) : csq(p) == 1 ? rat(1, 9) : rat(1, 72); }

// Mode basis: rho; j_x, j_y, j_z; c_a c_b - delta_ab/3 for ab in
// xx, xy, xz, yy, yz, zz; ghosts g, c_x g, c_y g, c_z g, c_x c_y c_z with
// g = 1 on axis links and -2 on rest and diagonal links.
constexpr Rat mode_rat(int m, int p) {
  constexpr int pa[6] = {0, 0, 0, 1, 1, 2};
  constexpr int pb[6] = {0, 1, 2, 1, 2, 2};
  const int g = csq(p) == 1 ? 1 : -2;
  if (m == 0) return rat(1, 1);
  if (m <= 3) return rat(kC[p][m - 1], 1);
  if (m <= 9) {
    const int a = pa[m - 4], b = pb[m - 4];
    return rat(3 * kC[p][a] * kC[p][b] - (a == b ? 1 : 0), 3);
  }
  if (m == 10) return rat(g, 1);
  if (m <= 13) return rat(kC[p][m - 11] * g, 1);
  return rat(kC[p][0] * kC[p][1] * kC[p][2], 1);
}

constexpr Rat mode_norm(int m) {
  Rat s = rat(0, 1);
  for (int i = 0; i < NV; ++i) s = s + weight_rat(i) * mode_rat(m, i) * mode_rat(m, i);
  return s;
}

constexpr Rat inv_rat(int p, int m) { return weight_rat(p) * mode_rat(m, p) / mode_norm(m); }

struct Tables {
  double M[NV][NV];
  double MI[NV][NV];
  double W[NV];
};
constexpr Tables make_tables() {
  Tables t{};
  for (int m = 0; m < NV; ++m)
    for (int p = 0; p < NV; ++p) {
      t.M[m][p] = to_double(mode_rat(m, p));
      t.MI[p][m] = to_double(inv_rat(p, m));
    }
  for (int p = 0; p < NV; ++p) t.W[p] = to_double(weight_rat(p));
  return t;
}
constexpr Tables kT = make_tables();

// Opposite population: c_opp = -c.
constexpr int opposite(int p) {
  for (int q = 0; q < NV; ++q)
    if (kC[q][0] == -kC[p][0] && kC[q][1] == -kC[p][1] && kC[q][2] == -kC[p][2]) return q;
  return -1;
}

// ------------------------------------------------ exact FP64 sparse algebra
// s (+)= f * coef with the reference's rounding, zero terms skipped.
template <typename T, int COEF_ROW, int COEF_COL, bool INVERSE>
LBX_HD T acc_term(T s, T v) {
  constexpr double c = INVERSE ? kT.MI[COEF_ROW][COEF_COL] : kT.M[COEF_ROW][COEF_COL];
  if constexpr (c == 0.0) {
    return s;
  } else if constexpr (c == 1.0) {
    return s + v;
  } else if constexpr (c == -1.0) {
    return s - v;
  } else {
    return s + v * T(c);
  }
}

// mode[m] = sum_p f[p] * M[m][p], p ascending, starting from +0.0.
template <int m, int... P>
LBX_HD double mode_of(const double* f, std::integer_sequence<int, P...>) {
  double s = 0.0;
  ((s = acc_term<double, m, P, false>(s, f[P])), ...);
  return s;
}

// f[p] = sum_m mode[m] * MI[p][m] over m = 0..9 (modes 10..14 are zero).
template <int p, int... Mi>
LBX_HD double pop_of(const double* mode, std::integer_sequence<int, Mi...>) {
  double s = 0.0;
  ((s = acc_term<double, p, Mi, true>(s, mode[Mi])), ...);
  return s;
}

template <int... Mi>
LBX_HD void modes_0_9(const double* f, double* mode, std::integer_sequence<int, Mi...>) {
  ((mode[Mi] = mode_of<Mi>(f, std::make_integer_sequence<int, NV>{})), ...);
}

template <int... P>
LBX_HD void project_back(const double* mode, double* f, std::integer_sequence<int, P...>) {
  ((f[P] = pop_of<P>(mode, std::make_integer_sequence<int, 10>{})), ...);
}

// One reference collision in place (src/AmrSim.cpp:28-104), bit-exact.
LBX_HD void collide_f64(double* f, double omega_s, double omega_b) {
  constexpr double DELTA_D = 1.0 / 15;  // DELTA[a][a]; off-diagonal entries are 0
  double mode[10];
  modes_0_9(f, mode, std::make_integer_sequence<int, 10>{});
  const double density = mode[0];
  const double v0 = mode[1] / density;
  double usq = 0.0;
  usq = usq + v0 * v0;
  const double v1 = mode[2] / density;
  usq = usq + v1 * v1;
  const double v2 = mode[3] / density;
  usq = usq + v2 * v2;

  double s00 = mode[4], s01 = mode[5], s02 = mode[6];
  double s11 = mode[7], s12 = mode[8], s22 = mode[9];
  double TrS = 0.0;
  TrS = TrS + s00;
  TrS = TrS + s11;
  TrS = TrS + s22;
  const double third_tr = TrS / 3.0;
  s00 = s00 - third_tr;
  s11 = s11 - third_tr;
  s22 = s22 - third_tr;
  TrS = TrS - omega_b * (TrS - density * usq);
  const double third_relaxed = TrS / 3.0;
  // Off-diagonal: usq * DELTA = +0, and x - (+0) == x exactly.
  s00 = s00 - omega_s * (s00 - density * (v0 * v0 - usq * DELTA_D));
  s01 = s01 - omega_s * (s01 - density * (v0 * v1));
  s02 = s02 - omega_s * (s02 - density * (v0 * v2));
  s00 = s00 + third_relaxed;
  s11 = s11 - omega_s * (s11 - density * (v1 * v1 - usq * DELTA_D));
  s12 = s12 - omega_s * (s12 - density * (v1 * v2));
  s11 = s11 + third_relaxed;
  s22 = s22 - omega_s * (s22 - density * (v2 * v2 - usq * DELTA_D));
  s22 = s22 + third_relaxed;
  mode[4] = s00;
  mode[5] = s01;
  mode[6] = s02;
  mode[7] = s11;
  mode[8] = s12;
  mode[9] = s22;
  project_back(mode, f, std::make_integer_sequence<int, NV>{});
}

// rho = mode 0, u_a = mode_{a+1} / mode_0 (src/AmrSim.cpp:957-971), bit-exact.
LBX_HD void hydro_f64(const double* f, double& rho, double& u0, double& u1, double& u2) {
  const double m0 = mode_of<0>(f, std::make_integer_sequence<int, NV>{});
  const double m1 = mode_of<1>(f, std::make_integer_sequence<int, NV>{});
  const double m2 = mode_of<2>(f, std::make_integer_sequence<int, NV>{});
  const double m3 = mode_of<3>(f, std::make_integer_sequence<int, NV>{});
  rho = m0;
  u0 = m1 / m0;
  u1 = m2 / m0;
  u2 = m3 / m0;
}

// Equilibrium (src/AmrSim.cpp:873-927), reference operation order.
LBX_HD void equilibrium_f64(double rho, double ux, double uy, double uz, double* f) {
  constexpr double CS2 = 1.0 / 3.0;
  const double rw0 = rho * 2.0 / 9.0;
  const double rw1 = rho / 9.0;
  const double rw2 = rho / 72.0;
  const double u2x = ux * ux, u2y = uy * uy, u2z = uz * uz;
  const double ucx = ux / CS2, ucy = uy / CS2, ucz = uz / CS2;
  const double q2x = u2x / (2.0 * CS2 * CS2);
  const double q2y = u2y / (2.0 * CS2 * CS2);
  const double q2z = u2z / (2.0 * CS2 * CS2);
  const double uv = ucx * ucy, vw = ucy * ucz, uw = ucx * ucz;
  const double mod_sq = (u2x + u2y + u2z) / (2.0 * CS2);
  const double mod_sq_2 = (u2x + u2y + u2z) * (1 - CS2) / (2.0 * CS2 * CS2);
  f[0] = rw0 * (1.0 - mod_sq);
  f[1] = rw1 * (1.0 - mod_sq + ucx + q2x);
  f[2] = rw1 * (1.0 - mod_sq - ucx + q2x);
  f[3] = rw1 * (1.0 - mod_sq + ucy + q2y);
  f[4] = rw1 * (1.0 - mod_sq - ucy + q2y);
  f[5] = rw1 * (1.0 - mod_sq + ucz + q2z);
  f[6] = rw1 * (1.0 - mod_sq - ucz + q2z);
  // Diagonals: signs follow c_x, c_y, c_z, c_x c_y, c_y c_z, c_x c_z.
  f[7] = rw2 * (1.0 + ucx + ucy + ucz + uv + vw + uw + mod_sq_2);
  f[8] = rw2 * (1.0 + ucx + ucy - ucz + uv - vw - uw + mod_sq_2);
  f[9] = rw2 * (1.0 + ucx - ucy + ucz - uv - vw + uw + mod_sq_2);
  f[10] = rw2 * (1.0 + ucx - ucy - ucz - uv + vw - uw + mod_sq_2);
  f[11] = rw2 * (1.0 - ucx + ucy + ucz - uv + vw - uw + mod_sq_2);
  f[12] = rw2 * (1.0 - ucx + ucy - ucz - uv - vw + uw + mod_sq_2);
  f[13] = rw2 * (1.0 - ucx - ucy + ucz + uv - vw - uw + mod_sq_2);
  f[14] = rw2 * (1.0 - ucx - ucy - ucz + uv + vw + uw + mod_sq_2);
}

// ------------------------------------------- FP32 shifted-storage collision
// Storage holds d_p = f_p - w_p rho0.  Because sum_i w_i M[m][i] = delta_m0
// and MI[p][0] = w_p, only mode 0 sees the shift: rho = rho0 + sum_p d_p, the
// other modes are the same linear forms of d, and the collided
// d'_p = w_p (rho - rho0) + sum_{m=1..9} mode'_m MI[p][m].
template <int m, int... P>
LBX_HD float mode_of_f32(const float* d, std::integer_sequence<int, P...>) {
  float s = 0.0f;
  ((s = acc_term<float, m, P, false>(s, d[P])), ...);
  return s;
}
template <int p, int... Mi>
LBX_HD float pop_of_f32(const float* mode, std::integer_sequence<int, Mi...>) {
  // mode[0] carries drho, whose coefficient is MI[p][0] = w_p.
  float s = 0.0f;
  ((s = acc_term<float, p, Mi, true>(s, mode[Mi])), ...);
  return s;
}
template <int... Mi>
LBX_HD void modes_0_9_f32(const float* d, float* mode, std::integer_sequence<int, Mi...>) {
  ((mode[Mi] = mode_of_f32<Mi>(d, std::make_integer_sequence<int, NV>{})), ...);
}
template <int... P>
LBX_HD void project_back_f32(const float* mode, float* d, std::integer_sequence<int, P...>) {
  ((d[P] = pop_of_f32<P>(mode, std::make_integer_sequence<int, 10>{})), ...);
}

LBX_HD void collide_f32_shifted(float* d, float rho0, float omega_s, float omega_b) {
  float mode[10];
  modes_0_9_f32(d, mode, std::make_integer_sequence<int, 10>{});
  const float drho = mode[0];
  const float rho = rho0 + drho;
  const float inv_rho = 1.0f / rho;
  const float v0 = mode[1] * inv_rho, v1 = mode[2] * inv_rho, v2 = mode[3] * inv_rho;
  const float usq = v0 * v0 + v1 * v1 + v2 * v2;
  float s00 = mode[4], s01 = mode[5], s02 = mode[6];
  float s11 = mode[7], s12 = mode[8], s22 = mode[9];
  float TrS = s00 + s11 + s22;
  const float third_tr = TrS * (1.0f / 3.0f);
  s00 -= third_tr;
  s11 -= third_tr;
  s22 -= third_tr;
  TrS -= omega_b * (TrS - rho * usq);
  const float third_relaxed = TrS * (1.0f / 3.0f);
  const float ud = usq * (1.0f / 15.0f);
  s00 = s00 - omega_s * (s00 - rho * (v0 * v0 - ud)) + third_relaxed;
  s11 = s11 - omega_s * (s11 - rho * (v1 * v1 - ud)) + third_relaxed;
  s22 = s22 - omega_s * (s22 - rho * (v2 * v2 - ud)) + third_relaxed;
  s01 = s01 - omega_s * (s01 - rho * (v0 * v1));
  s02 = s02 - omega_s * (s02 - rho * (v0 * v2));
  s12 = s12 - omega_s * (s12 - rho * (v1 * v2));
  mode[4] = s00;
  mode[5] = s01;
  mode[6] = s02;
  mode[7] = s11;
  mode[8] = s12;
  mode[9] = s22;
  project_back_f32(mode, d, std::make_integer_sequence<int, NV>{});
}

}  // namespace lbx
