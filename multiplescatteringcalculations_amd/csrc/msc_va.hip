// Volume-averaged nearest-neighbour 9x9 propagator blocks (host, FP64).
//
// Replaces inter_voxel_propagator_9x9 (cubic_scattering/inter_voxel_propagator.py:
// 1295-1355) as called by _build_slab_kernels(volume_averaged=True)
// (slab_scattering.py:316-329) for the 26 offsets R in {-1,0,1}^3 \ {0}:
//   G = <G_ij>: 1/(4 pi mu)[d_ij Phi - eta_s d2Psi_ij] + the w^2, w^4, w^6 series
//   C_i,(mn) = Voigt-symmetrised dG_imn,  H = C^T
//   S = Voigt contraction of P_ijkl(w) = sum_n w^2n P^(n),
//       P^(n) = -1/(2 mu_n)[d_ik A_jl + d_jk A_il - 2 eta_n B_ijkl]
// from the unit-cube potential derivatives tabulated for the face, edge and
// corner directions (msc_va_tables.hpp), mapped to R by a signed permutation.
// As in the reference, the values are not scaled with the cube size, the
// corner fourth-derivative tensors are filled from the index groups
// {0012, 0102, 0122} (the {0112} class stays zero), and the signed permutation
// is the first one a row-by-row greedy match finds.
#include <cmath>
#include <cstring>

#include "msc_setup.hpp"
#include "msc_va_tables.hpp"

namespace msc {

namespace {

using T2 = double[3][3];
using T3 = double[3][3][3];
using T4 = double[3][3][3][3];

constexpr double kPi = 3.14159265358979323846264338327950288;

int ntype(const int R[3]) {
  const int s = std::abs(R[0]) + std::abs(R[1]) + std::abs(R[2]);
  return s - 1;  // 1 face, 2 edge, 3 corner -> 0, 1, 2 (R is a nearest neighbour)
}

void a_matrix(int t, const double v[3], T2 A) {
  std::memset(A, 0, sizeof(T2));
  if (t == 0) {
    A[0][0] = v[0];
    A[1][1] = A[2][2] = v[1];
  } else if (t == 1) {
    A[0][0] = A[1][1] = v[0];
    A[2][2] = v[1];
    A[0][1] = A[1][0] = v[2];
  } else {
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) A[i][j] = i == j ? v[0] : v[2];
  }
}

// index groups that carry each tabulated d4W value (every permutation of each
// listed index quadruple receives it), terminated by -1
struct Group {
  int n;
  int idx[6][4];
};
const Group kBGroups[3][4] = {
    {{1, {{0, 0, 0, 0}}},
     {2, {{0, 0, 1, 1}, {0, 0, 2, 2}}},
     {2, {{1, 1, 1, 1}, {2, 2, 2, 2}}},
     {1, {{1, 1, 2, 2}}}},
    {{2, {{0, 0, 0, 0}, {1, 1, 1, 1}}},
     {1, {{0, 0, 1, 1}}},
     {2, {{0, 0, 2, 2}, {1, 1, 2, 2}}},
     {1, {{2, 2, 2, 2}}}},
    {{3, {{0, 0, 0, 0}, {1, 1, 1, 1}, {2, 2, 2, 2}}},
     {3, {{0, 0, 1, 1}, {0, 0, 2, 2}, {1, 1, 2, 2}}},
     {6, {{0, 0, 0, 1}, {0, 0, 0, 2}, {0, 1, 1, 1}, {1, 1, 1, 2}, {0, 2, 2, 2}, {1, 2, 2, 2}}},
     {3, {{0, 0, 1, 2}, {0, 1, 0, 2}, {0, 1, 2, 2}}}},
};
// edge has six values: the last two groups
const Group kBEdgeExtra[2] = {{2, {{0, 0, 0, 1}, {0, 1, 1, 1}}}, {1, {{0, 1, 2, 2}}}};

void set_all_perms4(T4 B, const int q[4], double v) {
  int p[4] = {0, 1, 2, 3};
  // all 24 orderings of the positions
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b)
      for (int c = 0; c < 4; ++c)
        for (int d = 0; d < 4; ++d) {
          if (a == b || a == c || a == d || b == c || b == d || c == d) continue;
          B[q[a]][q[b]][q[c]][q[d]] = v;
        }
  (void)p;
}

void b_tensor(int t, int order, double scale, T4 B) {
  std::memset(B, 0, sizeof(T4));
  const int ng = t == 1 ? 6 : 4;
  for (int g = 0; g < ng; ++g) {
    const Group& G = (t == 1 && g >= 4) ? kBEdgeExtra[g - 4] : kBGroups[t][g];
    const double v = va::kB[t][order][g] * scale;
    for (int m = 0; m < G.n; ++m) set_all_perms4(B, G.idx[m], v);
  }
}

void d3_tensor(int t, int order, T3 D) {
  std::memset(D, 0, sizeof(T3));
  auto put = [&](int i, int j, int k, double v) {
    const int q[3] = {i, j, k};
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b)
        for (int c = 0; c < 3; ++c)
          if (a != b && a != c && b != c) D[q[a]][q[b]][q[c]] = v;
  };
  const double* v = va::kD3[t][order];
  if (t == 0) {
    put(0, 0, 0, v[0]);
    put(0, 1, 1, v[1]);
    put(0, 2, 2, v[1]);
  } else if (t == 1) {
    put(0, 0, 0, v[0]);
    put(1, 1, 1, v[0]);
    put(0, 0, 1, v[1]);
    put(0, 1, 1, v[1]);
    put(0, 2, 2, v[2]);
    put(1, 2, 2, v[2]);
  } else {
    for (int i = 0; i < 3; ++i) put(i, i, i, v[0]);
    put(0, 0, 1, v[1]);
    put(0, 0, 2, v[1]);
    put(0, 1, 1, v[1]);
    put(0, 2, 2, v[1]);
    put(1, 1, 2, v[1]);
    put(1, 2, 2, v[1]);
    put(0, 1, 2, v[2]);
  }
}

void dw_vector(int t, int order, double w[3]) {
  const double v = va::kDW[t][order][0];
  w[0] = v;
  w[1] = t >= 1 ? v : 0.0;
  w[2] = t == 2 ? v : 0.0;
}

// normalisations of the tabulated values per order (A: orders 1..4, B: 1..3)
double norm_a(int n) {
  switch (n) {
    case 1: return 1.0 / (8.0 * kPi);
    case 2: return -1.0 / (96.0 * kPi);
    case 3: return 1.0 / (2880.0 * kPi);
    default: return -1.0 / (161280.0 * kPi);
  }
}
double norm_b(int n) {
  switch (n) {
    case 1: return 1.0 / (96.0 * kPi);
    case 2: return -1.0 / (2880.0 * kPi);
    default: return 1.0 / (161280.0 * kPi);
  }
}

// P_ijkl += c * (-1/(2 mu)) [d_ik A_jl + d_jk A_il - 2 eta B_ijkl]
void add_p(T4 P, const T2 A, const T4 B, double mu, double eta, double c) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 3; ++k)
        for (int l = 0; l < 3; ++l) {
          double v = -2.0 * eta * B[i][j][k][l];
          if (i == k) v += A[j][l];
          if (j == k) v += A[i][l];
          P[i][j][k][l] += c * (-1.0 / (2.0 * mu)) * v;
        }
}

// signed permutation Q with Q s = R, s the canonical direction of R's type
void oh_map(const int R[3], int t, double Q[3][3]) {
  const double s[3] = {1.0, t >= 1 ? 1.0 : 0.0, t == 2 ? 1.0 : 0.0};
  std::memset(Q, 0, sizeof(double) * 9);
  bool used[3] = {false, false, false};
  for (int i = 0; i < 3; ++i) {
    const double ti = (double)R[i];
    for (int j = 0; j < 3; ++j) {
      if (!used[j] && std::fabs(std::fabs(ti) - s[j]) < 1e-10 && s[j] > 1e-10) {
        Q[i][j] = (ti > 0 ? 1.0 : -1.0);
        used[j] = true;
        break;
      }
      if (!used[j] && s[j] < 1e-10 && std::fabs(ti) < 1e-10) {
        Q[i][j] = 1.0;
        used[j] = true;
        break;
      }
    }
  }
}

constexpr int kVP[6][2] = {{0, 0}, {1, 1}, {2, 2}, {1, 2}, {0, 2}, {0, 1}};

}  // namespace

// 26 blocks, row-major 9x9 each, in the order R = (dz, dx, dy) with
// dz, dx, dy = -1..1 lexicographic and (0,0,0) skipped.
void va_blocks(double alpha, double beta, double rho, double omega, int n_orders,
               double* out26x81) {
  const double mu = rho * beta * beta;
  const double nu = (alpha * alpha - 2.0 * beta * beta) / (2.0 * (alpha * alpha - beta * beta));
  const int nord = n_orders < 0 ? 0 : (n_orders > 3 ? 3 : n_orders);
  int slot = 0;
  for (int dz = -1; dz <= 1; ++dz)
    for (int dx = -1; dx <= 1; ++dx)
      for (int dy = -1; dy <= 1; ++dy) {
        if (dz == 0 && dx == 0 && dy == 0) continue;
        const int R[3] = {dz, dx, dy};
        const int t = ntype(R);
        double Q[3][3];
        oh_map(R, t, Q);
        // strain propagator, canonical direction
        T4 P, B;
        T2 A;
        std::memset(P, 0, sizeof(T4));
        a_matrix(t, va::kA[t][0], A);
        b_tensor(t, 0, 1.0, B);
        add_p(P, A, B, mu, 1.0 / (2.0 * (1.0 - nu)), 1.0);
        for (int n = 1; n <= nord; ++n) {
          double an[3];
          for (int q = 0; q < 3; ++q) an[q] = va::kA[t][n][q] * norm_a(n);
          a_matrix(t, an, A);
          b_tensor(t, n, norm_b(n), B);
          const double mun = rho * std::pow(beta, 2 * n + 2);
          const double etan = 1.0 - std::pow(beta / alpha, 2 * n + 2);
          add_p(P, A, B, mun, etan, std::pow(omega, 2 * n));
        }
        // G and dG, canonical direction
        const double eta_s = 1.0 / (4.0 * (1.0 - nu));
        T2 G, d2;
        T3 dG, D3;
        double dw[3];
        a_matrix(t, va::kA[t][1], d2);
        const double phi = 0.5 * (d2[0][0] + d2[1][1] + d2[2][2]);
        d3_tensor(t, 0, D3);
        dw_vector(t, 0, dw);
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j) {
            G[i][j] = ((i == j ? phi : 0.0) - eta_s * d2[i][j]) / (4.0 * kPi * mu);
            for (int k = 0; k < 3; ++k)
              dG[i][j][k] = ((i == j ? dw[k] : 0.0) - eta_s * D3[i][j][k]) / (4.0 * kPi * mu);
          }
        // w^2n terms: coefficient, Laplacian divisor (12, 30, 56), eta_n
        const double cs = beta, cp = alpha;
        const double coef[3] = {-(omega * omega) / (8.0 * kPi * rho * std::pow(cs, 4)),
                                std::pow(omega, 4) / (96.0 * kPi * rho * std::pow(cs, 6)),
                                -std::pow(omega, 6) / (2880.0 * kPi * rho * std::pow(cs, 8))};
        const double lap[3] = {12.0, 30.0, 56.0};
        for (int n = 1; n <= nord; ++n) {
          const double etan = 1.0 - std::pow(cs / cp, 2 * n + 2);
          a_matrix(t, va::kA[t][n + 1], d2);
          const double tr = d2[0][0] + d2[1][1] + d2[2][2];
          d3_tensor(t, n, D3);
          dw_vector(t, n, dw);
          for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
              G[i][j] += coef[n - 1] * ((i == j ? tr / lap[n - 1] : 0.0) -
                                        etan / lap[n - 1] * d2[i][j]);
              if (omega != 0.0)
                for (int k = 0; k < 3; ++k)
                  dG[i][j][k] += coef[n - 1] * ((i == j ? dw[k] : 0.0) -
                                                etan / lap[n - 1] * D3[i][j][k]);
            }
        }
        // rotate to R: X'_{i...} = Q_ia ... X_{a...}
        T2 Gr;
        T3 dGr;
        T4 Pr;
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j) {
            double g = 0.0;
            for (int a = 0; a < 3; ++a)
              for (int b = 0; b < 3; ++b) g += Q[i][a] * Q[j][b] * G[a][b];
            Gr[i][j] = g;
            for (int k = 0; k < 3; ++k) {
              double v = 0.0;
              for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b)
                  for (int c = 0; c < 3; ++c) v += Q[i][a] * Q[j][b] * Q[k][c] * dG[a][b][c];
              dGr[i][j][k] = v;
              for (int l = 0; l < 3; ++l) {
                double w = 0.0;
                for (int a = 0; a < 3; ++a)
                  for (int b = 0; b < 3; ++b)
                    for (int c = 0; c < 3; ++c)
                      for (int d = 0; d < 3; ++d)
                        w += Q[i][a] * Q[j][b] * Q[k][c] * Q[l][d] * P[a][b][c][d];
                Pr[i][j][k][l] = w;
              }
            }
          }
        double* blk = out26x81 + (size_t)slot * 81;
        for (int q = 0; q < 81; ++q) blk[q] = 0.0;
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j) blk[i * 9 + j] = Gr[i][j];
        for (int al = 0; al < 6; ++al) {
          const int m = kVP[al][0], n = kVP[al][1];
          for (int i = 0; i < 3; ++i) {
            const double c = m == n ? dGr[i][m][m] : 0.5 * (dGr[i][m][n] + dGr[i][n][m]);
            blk[i * 9 + 3 + al] = c;        // C
            blk[(3 + al) * 9 + i] = c;      // H = C^T
          }
          const int p = kVP[al][0], q = kVP[al][1];
          const double mult = p != q ? 2.0 : 1.0;
          for (int be = 0; be < 6; ++be) {
            const int mm = kVP[be][0], nn = kVP[be][1];
            double v = mult * Pr[p][q][mm][nn];
            if (mm != nn) v += mult * Pr[p][q][nn][mm];
            blk[(3 + al) * 9 + 3 + be] = v * (be >= 3 ? 0.5 : 1.0);  // S
          }
        }
        ++slot;
      }
}

}  // namespace msc
