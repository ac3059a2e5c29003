/* Two-electron repulsion integrals over contracted Cartesian Gaussian shells (oracle -- test
 * infrastructure, not part of the product).
 *
 * McMurchie-Davidson scheme, written from the published recurrences (McMurchie & Davidson,
 * J. Comput. Phys. 26, 218 (1978); Helgaker, Jorgensen, Olsen, "Molecular Electronic-Structure
 * Theory", sections 9.5 and 9.9): Hermite expansion coefficients E^{ij}_t per Cartesian direction,
 * Hermite Coulomb integrals R^n_{tuv} from the Boys function, and
 *   (ab|cd) = 2 pi^{5/2} / (p q sqrt(p + q)) sum_{tuv} E^{ab}_{tuv} sum_{t'u'v'} (-1)^{t'+u'+v'}
 *             E^{cd}_{t'u'v'} R_{t+t', u+u', v+v'}(alpha, P - Q).
 * The Boys function uses its power series for small arguments (downward recursion) and the erf
 * closed form with upward recursion for large ones (Helgaker et al. eqs. 9.8.12-9.8.15).
 *
 * Shell s has angular momentum l[s] <= 2, centre xyz[3 s ..], nprim[s] primitives starting at
 * first[s] in exps / coef (coef multiplies the primitive x^i y^j z^k exp(-a r^2) as given, for every
 * component). Components are ordered (i, j, k) with i descending, then j descending, and the basis
 * function index of component c of shell s is bf0[s] + c. The full eri[n][n][n][n] (chemist order,
 * n = total components) is filled using the 8-fold symmetry.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#define LMAX 2
#define NCOMP_MAX 6
#define RMAX (4 * LMAX)

static void comps(int l, int c[][3]) {
  int n = 0;
  for (int i = l; i >= 0; --i)
    for (int j = l - i; j >= 0; --j) {
      c[n][0] = i;
      c[n][1] = j;
      c[n][2] = l - i - j;
      ++n;
    }
}

/* F_n(T) for n = 0 .. nmax */
static void boys(int nmax, double T, double* F) {
  if (T < 1e-15) {
    for (int n = 0; n <= nmax; ++n) F[n] = 1.0 / (2 * n + 1) - T / (2 * n + 3);
    return;
  }
  const double et = exp(-T);
  if (T < 35.0) {
    /* F_m(T) = e^{-T} sum_k (2T)^k / ((2m+1)(2m+3)...(2m+2k+1)), all terms positive */
    double term = 1.0 / (2 * nmax + 1), sum = term;
    for (int k = 1; k < 400; ++k) {
      term *= 2.0 * T / (2 * nmax + 2 * k + 1);
      sum += term;
      if (term < 1e-17 * sum) break;
    }
    F[nmax] = et * sum;
    for (int n = nmax; n > 0; --n) F[n - 1] = (2.0 * T * F[n] + et) / (2 * n - 1);
  } else {
    F[0] = 0.5 * sqrt(M_PI / T) * erf(sqrt(T));
    for (int n = 0; n < nmax; ++n) F[n + 1] = ((2 * n + 1) * F[n] - et) / (2.0 * T);
  }
}

/* E^{ij}_t for i, j <= LMAX, t <= i + j: E[i][j][t] */
static void hermite_E(double a, double b, double Q, double E[LMAX + 1][LMAX + 1][2 * LMAX + 1]) {
  const double p = a + b, mu = a * b / p;
  const double XPA = -b / p * Q, XPB = a / p * Q; /* Q = A - B */
  memset(E, 0, sizeof(double) * (LMAX + 1) * (LMAX + 1) * (2 * LMAX + 1));
  E[0][0][0] = exp(-mu * Q * Q);
  for (int i = 0; i <= LMAX; ++i)
    for (int j = 0; j <= LMAX; ++j) {
      if (i == 0 && j == 0) continue;
      for (int t = 0; t <= i + j; ++t) {
        double v;
        if (i > 0) {
          v = XPA * E[i - 1][j][t];
          if (t > 0) v += E[i - 1][j][t - 1] / (2 * p);
          if (t + 1 <= i - 1 + j) v += (t + 1) * E[i - 1][j][t + 1];
        } else {
          v = XPB * E[i][j - 1][t];
          if (t > 0) v += E[i][j - 1][t - 1] / (2 * p);
          if (t + 1 <= i + j - 1) v += (t + 1) * E[i][j - 1][t + 1];
        }
        E[i][j][t] = v;
      }
    }
}

/* R[t][u][v] = R^0_{tuv}(alpha, PQ) for t + u + v <= L, through R^n with n <= L */
static void hermite_R(int L, double alpha, const double PQ[3], double R[RMAX + 1][RMAX + 1][RMAX + 1]) {
  static double Rn[RMAX + 1][RMAX + 1][RMAX + 1][RMAX + 1];
  double F[RMAX + 1];
  const double X = PQ[0], Y = PQ[1], Z = PQ[2];
  boys(L, alpha * (X * X + Y * Y + Z * Z), F);
  double m2a = 1.0;
  for (int n = 0; n <= L; ++n) {
    Rn[n][0][0][0] = m2a * F[n];
    m2a *= -2.0 * alpha;
  }
  /* R^n_{t+1,u,v} = t R^{n+1}_{t-1,u,v} + X R^{n+1}_{t,u,v}, same for u (Y) and v (Z) */
  for (int s = 1; s <= L; ++s)
    for (int n = 0; n <= L - s; ++n)
      for (int t = 0; t <= s; ++t)
        for (int u = 0; u <= s - t; ++u) {
          const int v = s - t - u;
          double r;
          if (t > 0) {
            r = X * Rn[n + 1][t - 1][u][v];
            if (t > 1) r += (t - 1) * Rn[n + 1][t - 2][u][v];
          } else if (u > 0) {
            r = Y * Rn[n + 1][t][u - 1][v];
            if (u > 1) r += (u - 1) * Rn[n + 1][t][u - 2][v];
          } else {
            r = Z * Rn[n + 1][t][u][v - 1];
            if (v > 1) r += (v - 1) * Rn[n + 1][t][u][v - 2];
          }
          Rn[n][t][u][v] = r;
        }
  for (int t = 0; t <= L; ++t)
    for (int u = 0; u <= L - t; ++u)
      for (int v = 0; v <= L - t - u; ++v) R[t][u][v] = Rn[0][t][u][v];
}

int md_eri(int ns, const int* l, const double* xyz, const int* nprim, const int* first, const double* exps,
           const double* coef, const int* bf0, int n, double* eri) {
  for (int s = 0; s < ns; ++s)
    if (l[s] < 0 || l[s] > LMAX) return -1;
  static double R[RMAX + 1][RMAX + 1][RMAX + 1];
  double Eab[3][LMAX + 1][LMAX + 1][2 * LMAX + 1], Ecd[3][LMAX + 1][LMAX + 1][2 * LMAX + 1];
  double buf[NCOMP_MAX][NCOMP_MAX][NCOMP_MAX][NCOMP_MAX];
  int ca[NCOMP_MAX][3], cb[NCOMP_MAX][3], cc[NCOMP_MAX][3], cd[NCOMP_MAX][3];
  for (int P = 0; P < ns; ++P)
    for (int Q = 0; Q <= P; ++Q)
      for (int Rs = 0; Rs < ns; ++Rs)
        for (int S = 0; S <= Rs; ++S) {
          if (P * (P + 1) / 2 + Q < Rs * (Rs + 1) / 2 + S) continue;
          const int la = l[P], lb = l[Q], lc = l[Rs], ld = l[S];
          const int na = (la + 1) * (la + 2) / 2, nb = (lb + 1) * (lb + 2) / 2;
          const int nc = (lc + 1) * (lc + 2) / 2, nd = (ld + 1) * (ld + 2) / 2;
          comps(la, ca);
          comps(lb, cb);
          comps(lc, cc);
          comps(ld, cd);
          memset(buf, 0, sizeof(buf));
          const double *A = xyz + 3 * P, *B = xyz + 3 * Q, *C = xyz + 3 * Rs, *D = xyz + 3 * S;
          const int L = la + lb + lc + ld;
          for (int ia = 0; ia < nprim[P]; ++ia)
            for (int ib = 0; ib < nprim[Q]; ++ib) {
              const double a = exps[first[P] + ia], b = exps[first[Q] + ib];
              const double p = a + b, cab = coef[first[P] + ia] * coef[first[Q] + ib];
              double Pc[3];
              for (int k = 0; k < 3; ++k) {
                Pc[k] = (a * A[k] + b * B[k]) / p;
                hermite_E(a, b, A[k] - B[k], Eab[k]);
              }
              for (int ic = 0; ic < nprim[Rs]; ++ic)
                for (int id = 0; id < nprim[S]; ++id) {
                  const double c = exps[first[Rs] + ic], d = exps[first[S] + id];
                  const double q = c + d, ccd = coef[first[Rs] + ic] * coef[first[S] + id];
                  double Qc[3], PQ[3];
                  for (int k = 0; k < 3; ++k) {
                    Qc[k] = (c * C[k] + d * D[k]) / q;
                    PQ[k] = Pc[k] - Qc[k];
                    hermite_E(c, d, C[k] - D[k], Ecd[k]);
                  }
                  const double alpha = p * q / (p + q);
                  hermite_R(L, alpha, PQ, R);
                  const double pref = cab * ccd * 2.0 * pow(M_PI, 2.5) / (p * q * sqrt(p + q));
                  for (int x = 0; x < na; ++x)
                    for (int y = 0; y < nb; ++y) {
                      const int i1 = ca[x][0], j1 = ca[x][1], k1 = ca[x][2];
                      const int i2 = cb[y][0], j2 = cb[y][1], k2 = cb[y][2];
                      for (int z = 0; z < nc; ++z)
                        for (int w = 0; w < nd; ++w) {
                          const int i3 = cc[z][0], j3 = cc[z][1], k3 = cc[z][2];
                          const int i4 = cd[w][0], j4 = cd[w][1], k4 = cd[w][2];
                          double val = 0.0;
                          for (int t = 0; t <= i1 + i2; ++t)
                            for (int u = 0; u <= j1 + j2; ++u)
                              for (int v = 0; v <= k1 + k2; ++v) {
                                const double e1 = Eab[0][i1][i2][t] * Eab[1][j1][j2][u] * Eab[2][k1][k2][v];
                                if (e1 == 0.0) continue;
                                for (int t2 = 0; t2 <= i3 + i4; ++t2)
                                  for (int u2 = 0; u2 <= j3 + j4; ++u2)
                                    for (int v2 = 0; v2 <= k3 + k4; ++v2) {
                                      const double e2 =
                                          Ecd[0][i3][i4][t2] * Ecd[1][j3][j4][u2] * Ecd[2][k3][k4][v2];
                                      const double sg = ((t2 + u2 + v2) & 1) ? -1.0 : 1.0;
                                      val += e1 * e2 * sg * R[t + t2][u + u2][v + v2];
                                    }
                              }
                          buf[x][y][z][w] += pref * val;
                        }
                    }
                }
            }
          for (int x = 0; x < na; ++x)
            for (int y = 0; y < nb; ++y)
              for (int z = 0; z < nc; ++z)
                for (int w = 0; w < nd; ++w) {
                  const int64_t i = bf0[P] + x, j = bf0[Q] + y, k = bf0[Rs] + z, m = bf0[S] + w;
                  const double val = buf[x][y][z][w];
                  const int64_t idx[8][4] = {{i, j, k, m}, {j, i, k, m}, {i, j, m, k}, {j, i, m, k},
                                             {k, m, i, j}, {m, k, i, j}, {k, m, j, i}, {m, k, j, i}};
                  for (int e = 0; e < 8; ++e)
                    eri[((idx[e][0] * n + idx[e][1]) * n + idx[e][2]) * n + idx[e][3]] = val;
                }
        }
  return 0;
}
