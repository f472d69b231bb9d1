// FP64 Kupradze propagator on the device: the 45 independent entries of the
// 9x9 block [[G, C], [H, S]] at one offset (resonance_tmatrix.py:162-428), and
// their expansion to the full 9x9 block.  Shared by the G0 table, the
// volume-averaged correction and the resonance T-matrix assembly.
#pragma once
#include <hip/hip_runtime.h>

namespace msc {

struct Rad {
  double2 phi, psi, phi1, psi1, phi2, psi2;
};

__device__ __forceinline__ double2 zc() { return make_double2(0.0, 0.0); }
__device__ __forceinline__ double2 cmd(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 csc(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
__device__ __forceinline__ double2 cad(double2 a, double2 b) {
  return make_double2(a.x + b.x, a.y + b.y);
}

// c e^{ikr} r^p and its first two r-derivatives accumulated into f, f1, f2
// (resonance_tmatrix.py:224-271).
__device__ __forceinline__ void rterm(double2 c, double2 e, double k, double r, int p, double2& f,
                                      double2& f1, double2& f2) {
  const double rp = pow(r, (double)p);
  const double2 ce = cmd(c, e);
  const double2 ikr = make_double2(0.0, k * r);
  f = cad(f, csc(ce, rp));
  // (ikr + p)
  f1 = cad(f1, csc(cmd(ce, make_double2((double)p, k * r)), rp / r));
  // (ikr)^2 + 2p ikr + p(p-1) = (-(kr)^2 + p(p-1)) + i 2p kr
  const double2 q = make_double2(-(k * r) * (k * r) + (double)(p * (p - 1)), 2.0 * p * ikr.y);
  f2 = cad(f2, csc(cmd(ce, q), rp / (r * r)));
}

__device__ inline Rad radial(double r, double kP, double kS) {
  double sP, cP, sS, cS;
  sincos(kP * r, &sP, &cP);
  sincos(kS * r, &sS, &cS);
  const double2 eP = make_double2(cP, sP), eS = make_double2(cS, sS);
  Rad o;
  o.phi = o.phi1 = o.phi2 = zc();
  o.psi = o.psi1 = o.psi2 = zc();
  rterm(make_double2(kS * kS, 0.0), eS, kS, r, -1, o.phi, o.phi1, o.phi2);
  rterm(make_double2(0.0, kS), eS, kS, r, -2, o.phi, o.phi1, o.phi2);
  rterm(make_double2(-1.0, 0.0), eS, kS, r, -3, o.phi, o.phi1, o.phi2);
  rterm(make_double2(0.0, -kP), eP, kP, r, -2, o.phi, o.phi1, o.phi2);
  rterm(make_double2(1.0, 0.0), eP, kP, r, -3, o.phi, o.phi1, o.phi2);
  rterm(make_double2(-kS * kS, 0.0), eS, kS, r, -1, o.psi, o.psi1, o.psi2);
  rterm(make_double2(0.0, -3.0 * kS), eS, kS, r, -2, o.psi, o.psi1, o.psi2);
  rterm(make_double2(3.0, 0.0), eS, kS, r, -3, o.psi, o.psi1, o.psi2);
  rterm(make_double2(kP * kP, 0.0), eP, kP, r, -1, o.psi, o.psi1, o.psi2);
  rterm(make_double2(0.0, 3.0 * kP), eP, kP, r, -2, o.psi, o.psi1, o.psi2);
  rterm(make_double2(-3.0, 0.0), eP, kP, r, -3, o.psi, o.psi1, o.psi2);
  return o;
}

__device__ __forceinline__ double dl(int i, int j) { return i == j ? 1.0 : 0.0; }

// The 45 independent entries of [[G, C], [H, S]] (without the common factor
// P = 1/(4 pi rho w^2)), added into acc:
//   G_ij        (i <= j), 6 entries
//   C[i][al]    3 x 6 (engineering halving of off-diagonal al)
//   S'[al][be]  = S[al][be] / W_be, al <= be, 21 entries
__device__ inline void add_prop45(double rz, double rx, double ry, double kP, double kS, double pre,
                           double2* acc) {
  const double r = sqrt(rz * rz + rx * rx + ry * ry);
  const double g[3] = {rz / r, rx / r, ry / r};
  const Rad R = radial(r, kP, kS);
  const double2 c1 = R.phi1;
  const double2 c2 = cad(R.psi1, csc(R.psi, -2.0 / r));
  const double2 c3 = csc(R.psi, 1.0 / r);
  const double2 t1 = csc(R.phi1, 1.0 / r);
  const double2 t2 = cad(R.phi2, csc(R.phi1, -1.0 / r));
  const double2 t3 = cad(csc(R.psi1, 1.0 / r), csc(R.psi, -2.0 / (r * r)));
  const double2 t4 = csc(R.psi, 1.0 / (r * r));
  const double2 t7 = cad(cad(R.psi2, csc(R.psi1, -5.0 / r)), csc(R.psi, 8.0 / (r * r)));

  auto Gd = [&](int i, int j, int k) -> double2 {
    double2 v = csc(c1, g[k] * dl(i, j));
    v = cad(v, csc(c2, g[i] * g[j] * g[k]));
    v = cad(v, csc(c3, dl(i, k) * g[j] + dl(j, k) * g[i]));
    return v;
  };
  auto Gdd = [&](int i, int j, int k, int l) -> double2 {
    double2 v = csc(t1, dl(i, j) * dl(k, l));
    v = cad(v, csc(t2, dl(i, j) * g[k] * g[l]));
    v = cad(v, csc(t3, g[i] * g[j] * dl(k, l) + dl(i, l) * g[j] * g[k] + dl(j, l) * g[i] * g[k] +
                           dl(i, k) * g[j] * g[l] + dl(j, k) * g[i] * g[l]));
    v = cad(v, csc(t4, dl(i, k) * dl(j, l) + dl(j, k) * dl(i, l)));
    v = cad(v, csc(t7, g[i] * g[j] * g[k] * g[l]));
    return v;
  };
  const int VP[6][2] = {{0, 0}, {1, 1}, {2, 2}, {1, 2}, {0, 2}, {0, 1}};
  int e = 0;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
#pragma unroll
    for (int j = i; j < 3; ++j) {
      const double2 v = cad(csc(R.phi, dl(i, j)), csc(R.psi, g[i] * g[j]));
      acc[e] = cad(acc[e], csc(v, pre));
      ++e;
    }
  }
#pragma unroll
  for (int i = 0; i < 3; ++i) {
#pragma unroll
    for (int al = 0; al < 6; ++al) {
      const int p = VP[al][0], q = VP[al][1];
      const double2 v = p == q ? Gd(i, p, p) : csc(cad(Gd(i, p, q), Gd(i, q, p)), 0.5);
      acc[e] = cad(acc[e], csc(v, pre));
      ++e;
    }
  }
#pragma unroll
  for (int al = 0; al < 6; ++al) {
#pragma unroll
    for (int be = al; be < 6; ++be) {
      const int p = VP[al][0], q = VP[al][1], m = VP[be][0], nn = VP[be][1];
      double2 v = Gdd(p, m, nn, q);
      if (m != nn) v = cad(v, Gdd(p, nn, m, q));
      if (p != q) {
        v = cad(v, Gdd(q, m, nn, p));
        if (m != nn) v = cad(v, Gdd(q, nn, m, p));
      }
      acc[e] = cad(acc[e], csc(v, pre));
      ++e;
    }
  }
}

// Full 9x9 block (row-major) from the 45 entries: G symmetric; C; H = W^-1 C^T
// (W = diag(1,1,1,2,2,2)); S[al][be] = S'[al][be] W_be, S' symmetric.
__device__ inline void expand45(const double2* e45, double2* K) {
  const int gi[6][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 2}};
  for (int e = 0; e < 6; ++e) {
    K[gi[e][0] * 9 + gi[e][1]] = e45[e];
    K[gi[e][1] * 9 + gi[e][0]] = e45[e];
  }
  for (int i = 0; i < 3; ++i)
    for (int al = 0; al < 6; ++al) {
      const double2 c = e45[6 + 6 * i + al];
      K[i * 9 + 3 + al] = c;
      const double f = al < 3 ? 1.0 : 2.0;
      K[(3 + al) * 9 + i] = make_double2(f * c.x, f * c.y);
    }
  int e = 24;
  for (int al = 0; al < 6; ++al)
    for (int be = al; be < 6; ++be, ++e) {
      const double2 sv = e45[e];
      const double wb = be < 3 ? 1.0 : 0.5, wa = al < 3 ? 1.0 : 0.5;
      K[(3 + al) * 9 + 3 + be] = make_double2(sv.x * wb, sv.y * wb);
      K[(3 + be) * 9 + 3 + al] = make_double2(sv.x * wa, sv.y * wa);
    }
}

}  // namespace msc
