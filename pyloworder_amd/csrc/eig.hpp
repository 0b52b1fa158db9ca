// Eigenvalues and right eigenvectors of a small real nonsymmetric matrix, the
// counterpart of numpy.linalg.eig (LAPACK ?geev) used by the reference's
// `eigen` (pyLOM/vmmath/maths.py:161-183) for the DMD operator.
//
// Algorithm (textbook: Wilkinson and Reinsch, Handbook for Automatic
// Computation II; Golub and Van Loan, Matrix Computations, 7.4-7.6):
//   1. Householder reduction to upper Hessenberg form, H = Z^T A Z, with Z
//      accumulated;
//   2. Francis implicit double-shift QR on H down to real Schur form
//      T = Q^T H Q, accumulated into Z (deflation on negligible subdiagonal
//      entries, exceptional shifts after 10 and 30 iterations);
//   3. eigenvectors of the quasi-triangular T by back substitution (2x2
//      blocks give conjugate pairs), back-transformed by Z;
//   4. ?geev's normalisation: unit 2-norm; for a conjugate pair the vector is
//      also rotated so that its largest component is real.
// Eigenvalues come in the order of T's diagonal, pairs with the positive
// imaginary part first (as ?geev).  The order within T depends on the
// deflation sequence, so it can differ from LAPACK's for the same matrix.
//
// Plain scalar code, __host__ __device__: one GPU thread runs it (n is the
// truncated DMD rank, small).  Work arrays are the caller's.
#pragma once
#include <cmath>

#ifndef PC_EIG_HD
#if defined(__HIPCC__)
#define PC_EIG_HD __host__ __device__
#else
#define PC_EIG_HD
#endif
#endif

namespace eig {

PC_EIG_HD inline double dabs(double x) { return x < 0 ? -x : x; }
PC_EIG_HD inline double dmax(double a, double b) { return a > b ? a : b; }

// (ar + i ai) / (br + i bi) without forming |b|^2 (Smith)
PC_EIG_HD inline void cdiv(double ar, double ai, double br, double bi, double& cr, double& ci) {
  if (dabs(br) > dabs(bi)) {
    const double r = bi / br, d = br + r * bi;
    cr = (ar + r * ai) / d;
    ci = (ai - r * ar) / d;
  } else {
    const double r = br / bi, d = bi + r * br;
    cr = (r * ar + ai) / d;
    ci = (r * ai - ar) / d;
  }
}

// H (n x n, row-major, ld n) is overwritten; Z (n x n) receives the Schur
// vectors and then the (real-packed) eigenvectors; wr / wi the eigenvalues;
// ort (n) work.  Returns 0, or -1 when the QR iteration did not converge.
PC_EIG_HD inline int real_eig(double* H, double* Z, double* wr, double* wi, double* ort, int nn) {
#define Hm(i, j) H[(i) * nn + (j)]
#define Zm(i, j) Z[(i) * nn + (j)]
  const double eps = 2.220446049250313e-16;
  const int low = 0, high = nn - 1;
  // ---- 1. Hessenberg reduction
  for (int m = low + 1; m <= high - 1; m++) {
    double scale = 0.0;
    for (int i = m; i <= high; i++) scale += dabs(Hm(i, m - 1));
    if (scale != 0.0) {
      double h = 0.0;
      for (int i = high; i >= m; i--) {
        ort[i] = Hm(i, m - 1) / scale;
        h += ort[i] * ort[i];
      }
      double g = sqrt(h);
      if (ort[m] > 0) g = -g;
      h -= ort[m] * g;
      ort[m] -= g;
      for (int j = m; j < nn; j++) {
        double f = 0.0;
        for (int i = high; i >= m; i--) f += ort[i] * Hm(i, j);
        f /= h;
        for (int i = m; i <= high; i++) Hm(i, j) -= f * ort[i];
      }
      for (int i = 0; i <= high; i++) {
        double f = 0.0;
        for (int j = high; j >= m; j--) f += ort[j] * Hm(i, j);
        f /= h;
        for (int j = m; j <= high; j++) Hm(i, j) -= f * ort[j];
      }
      ort[m] = scale * ort[m];
      Hm(m, m - 1) = scale * g;
    }
  }
  for (int i = 0; i < nn; i++)
    for (int j = 0; j < nn; j++) Zm(i, j) = (i == j ? 1.0 : 0.0);
  for (int m = high - 1; m >= low + 1; m--) {
    if (Hm(m, m - 1) != 0.0) {
      for (int i = m + 1; i <= high; i++) ort[i] = Hm(i, m - 1);
      for (int j = m; j <= high; j++) {
        double g = 0.0;
        for (int i = m; i <= high; i++) g += ort[i] * Zm(i, j);
        g = (g / ort[m]) / Hm(m, m - 1);  // two divisions: no underflow of the product
        for (int i = m; i <= high; i++) Zm(i, j) += g * ort[i];
      }
    }
  }
  // ---- 2. Francis double-shift QR
  int n = nn - 1;
  double exshift = 0.0, p = 0, q = 0, r = 0, s = 0, z = 0, t, w, x, y;
  double norm = 0.0;
  for (int i = 0; i < nn; i++)
    for (int j = (i > 0 ? i - 1 : 0); j < nn; j++) norm += dabs(Hm(i, j));
  int iter = 0, total = 0;
  while (n >= low) {
    int l = n;
    while (l > low) {
      s = dabs(Hm(l - 1, l - 1)) + dabs(Hm(l, l));
      if (s == 0.0) s = norm;
      if (dabs(Hm(l, l - 1)) < eps * s) break;
      l--;
    }
    if (l == n) {  // one root
      Hm(n, n) = Hm(n, n) + exshift;
      wr[n] = Hm(n, n);
      wi[n] = 0.0;
      n--;
      iter = 0;
    } else if (l == n - 1) {  // two roots
      w = Hm(n, n - 1) * Hm(n - 1, n);
      p = (Hm(n - 1, n - 1) - Hm(n, n)) / 2.0;
      q = p * p + w;
      z = sqrt(dabs(q));
      Hm(n, n) = Hm(n, n) + exshift;
      Hm(n - 1, n - 1) = Hm(n - 1, n - 1) + exshift;
      x = Hm(n, n);
      if (q >= 0) {  // real pair: rotate the 2x2 block to upper triangular
        z = (p >= 0) ? p + z : p - z;
        wr[n - 1] = x + z;
        wr[n] = wr[n - 1];
        if (z != 0.0) wr[n] = x - w / z;
        wi[n - 1] = 0.0;
        wi[n] = 0.0;
        x = Hm(n, n - 1);
        s = dabs(x) + dabs(z);
        p = x / s;
        q = z / s;
        r = sqrt(p * p + q * q);
        p /= r;
        q /= r;
        for (int j = n - 1; j < nn; j++) {
          z = Hm(n - 1, j);
          Hm(n - 1, j) = q * z + p * Hm(n, j);
          Hm(n, j) = q * Hm(n, j) - p * z;
        }
        for (int i = 0; i <= n; i++) {
          z = Hm(i, n - 1);
          Hm(i, n - 1) = q * z + p * Hm(i, n);
          Hm(i, n) = q * Hm(i, n) - p * z;
        }
        for (int i = low; i <= high; i++) {
          z = Zm(i, n - 1);
          Zm(i, n - 1) = q * z + p * Zm(i, n);
          Zm(i, n) = q * Zm(i, n) - p * z;
        }
      } else {  // complex pair
        wr[n - 1] = x + p;
        wr[n] = x + p;
        wi[n - 1] = z;
        wi[n] = -z;
      }
      n -= 2;
      iter = 0;
    } else {  // no convergence yet: a double-shift step on rows / columns l..n
      x = Hm(n, n);
      y = 0.0;
      w = 0.0;
      if (l < n) {
        y = Hm(n - 1, n - 1);
        w = Hm(n, n - 1) * Hm(n - 1, n);
      }
      if (iter == 10) {  // exceptional shift (Wilkinson)
        exshift += x;
        for (int i = low; i <= n; i++) Hm(i, i) -= x;
        s = dabs(Hm(n, n - 1)) + dabs(Hm(n - 1, n - 2));
        x = y = 0.75 * s;
        w = -0.4375 * s * s;
      }
      if (iter == 30) {  // second exceptional shift
        s = (y - x) / 2.0;
        s = s * s + w;
        if (s > 0) {
          s = sqrt(s);
          if (y < x) s = -s;
          s = x - w / ((y - x) / 2.0 + s);
          for (int i = low; i <= n; i++) Hm(i, i) -= s;
          exshift += s;
          x = y = w = 0.964;
        }
      }
      iter++;
      if (++total > 60 * nn) return -1;
      int m = n - 2;
      while (m >= l) {  // two consecutive small subdiagonal entries
        z = Hm(m, m);
        r = x - z;
        s = y - z;
        p = (r * s - w) / Hm(m + 1, m) + Hm(m, m + 1);
        q = Hm(m + 1, m + 1) - z - r - s;
        r = Hm(m + 2, m + 1);
        s = dabs(p) + dabs(q) + dabs(r);
        p /= s;
        q /= s;
        r /= s;
        if (m == l) break;
        if (dabs(Hm(m, m - 1)) * (dabs(q) + dabs(r)) <
            eps * (dabs(p) * (dabs(Hm(m - 1, m - 1)) + dabs(z) + dabs(Hm(m + 1, m + 1)))))
          break;
        m--;
      }
      for (int i = m + 2; i <= n; i++) {
        Hm(i, i - 2) = 0.0;
        if (i > m + 2) Hm(i, i - 3) = 0.0;
      }
      for (int k = m; k <= n - 1; k++) {  // chase the bulge
        const bool notlast = (k != n - 1);
        if (k != m) {
          p = Hm(k, k - 1);
          q = Hm(k + 1, k - 1);
          r = notlast ? Hm(k + 2, k - 1) : 0.0;
          x = dabs(p) + dabs(q) + dabs(r);
          if (x == 0.0) continue;
          p /= x;
          q /= x;
          r /= x;
        }
        s = sqrt(p * p + q * q + r * r);
        if (p < 0) s = -s;
        if (s != 0) {
          if (k != m) Hm(k, k - 1) = -s * x;
          else if (l != m) Hm(k, k - 1) = -Hm(k, k - 1);
          p += s;
          x = p / s;
          y = q / s;
          z = r / s;
          q /= p;
          r /= p;
          for (int j = k; j < nn; j++) {
            p = Hm(k, j) + q * Hm(k + 1, j);
            if (notlast) {
              p += r * Hm(k + 2, j);
              Hm(k + 2, j) -= p * z;
            }
            Hm(k, j) -= p * x;
            Hm(k + 1, j) -= p * y;
          }
          const int ilim = n < k + 3 ? n : k + 3;
          for (int i = 0; i <= ilim; i++) {
            p = x * Hm(i, k) + y * Hm(i, k + 1);
            if (notlast) {
              p += z * Hm(i, k + 2);
              Hm(i, k + 2) -= p * r;
            }
            Hm(i, k) -= p;
            Hm(i, k + 1) -= p * q;
          }
          for (int i = low; i <= high; i++) {
            p = x * Zm(i, k) + y * Zm(i, k + 1);
            if (notlast) {
              p += z * Zm(i, k + 2);
              Zm(i, k + 2) -= p * r;
            }
            Zm(i, k) -= p;
            Zm(i, k + 1) -= p * q;
          }
        }
      }
    }
  }
  // ---- 3. eigenvectors of the quasi-triangular form, then back-transformed
  if (norm == 0.0) return 0;
  for (n = nn - 1; n >= 0; n--) {
    p = wr[n];
    q = wi[n];
    if (q == 0) {  // real vector
      int l = n;
      Hm(n, n) = 1.0;
      for (int i = n - 1; i >= 0; i--) {
        w = Hm(i, i) - p;
        r = 0.0;
        for (int j = l; j <= n; j++) r += Hm(i, j) * Hm(j, n);
        if (wi[i] < 0.0) {
          z = w;
          s = r;
        } else {
          l = i;
          if (wi[i] == 0.0) {
            Hm(i, n) = (w != 0.0) ? -r / w : -r / (eps * norm);
          } else {  // the 2x2 block of a conjugate pair above
            x = Hm(i, i + 1);
            y = Hm(i + 1, i);
            q = (wr[i] - p) * (wr[i] - p) + wi[i] * wi[i];
            t = (x * s - z * r) / q;
            Hm(i, n) = t;
            Hm(i + 1, n) = (dabs(x) > dabs(z)) ? (-r - w * t) / x : (-s - y * t) / z;
          }
          t = dabs(Hm(i, n));
          if ((eps * t) * t > 1)
            for (int j = i; j <= n; j++) Hm(j, n) /= t;
        }
      }
    } else if (q < 0) {  // complex vector: real part in column n-1, imaginary in column n
      int l = n - 1;
      if (dabs(Hm(n, n - 1)) > dabs(Hm(n - 1, n))) {
        Hm(n - 1, n - 1) = q / Hm(n, n - 1);
        Hm(n - 1, n) = -(Hm(n, n) - p) / Hm(n, n - 1);
      } else {
        cdiv(0.0, -Hm(n - 1, n), Hm(n - 1, n - 1) - p, q, Hm(n - 1, n - 1), Hm(n - 1, n));
      }
      Hm(n, n - 1) = 0.0;
      Hm(n, n) = 1.0;
      for (int i = n - 2; i >= 0; i--) {
        double ra = 0.0, sa = 0.0, vr, vi;
        for (int j = l; j <= n; j++) {
          ra += Hm(i, j) * Hm(j, n - 1);
          sa += Hm(i, j) * Hm(j, n);
        }
        w = Hm(i, i) - p;
        if (wi[i] < 0.0) {
          z = w;
          r = ra;
          s = sa;
        } else {
          l = i;
          if (wi[i] == 0) {
            cdiv(-ra, -sa, w, q, Hm(i, n - 1), Hm(i, n));
          } else {
            x = Hm(i, i + 1);
            y = Hm(i + 1, i);
            vr = (wr[i] - p) * (wr[i] - p) + wi[i] * wi[i] - q * q;
            vi = (wr[i] - p) * 2.0 * q;
            if (vr == 0.0 && vi == 0.0) vr = eps * norm * (dabs(w) + dabs(q) + dabs(x) + dabs(y) + dabs(z));
            cdiv(x * r - z * ra + q * sa, x * s - z * sa - q * ra, vr, vi, Hm(i, n - 1), Hm(i, n));
            if (dabs(x) > (dabs(z) + dabs(q))) {
              Hm(i + 1, n - 1) = (-ra - w * Hm(i, n - 1) + q * Hm(i, n)) / x;
              Hm(i + 1, n) = (-sa - w * Hm(i, n) - q * Hm(i, n - 1)) / x;
            } else {
              cdiv(-r - y * Hm(i, n - 1), -s - y * Hm(i, n), z, q, Hm(i + 1, n - 1), Hm(i + 1, n));
            }
          }
          t = dmax(dabs(Hm(i, n - 1)), dabs(Hm(i, n)));
          if ((eps * t) * t > 1)
            for (int j = i; j <= n; j++) {
              Hm(j, n - 1) /= t;
              Hm(j, n) /= t;
            }
        }
      }
    }
  }
  for (int j = nn - 1; j >= low; j--)
    for (int i = low; i <= high; i++) {
      z = 0.0;
      for (int k = low; k <= (j < high ? j : high); k++) z += Zm(i, k) * Hm(k, j);
      Zm(i, j) = z;
    }
#undef Hm
#undef Zm
  return 0;
}

// ?geev's normalisation of the packed real vectors in Z into complex vectors
// V (n x n, interleaved re/im, row-major): a real eigenvalue's column scaled
// to unit 2-norm; a pair (j, j+1) (wi[j] > 0) gives v_j = (z_j + i z_{j+1})
// scaled to unit norm and rotated so that its largest-modulus component is
// real with the sign of its real part, and v_{j+1} = conj(v_j).
PC_EIG_HD inline void normalise(const double* Z, const double* wi, int nn, double* V) {
  for (int j = 0; j < nn; j++) {
    if (wi[j] == 0.0) {
      double s = 0.0;
      for (int i = 0; i < nn; i++) s += Z[i * nn + j] * Z[i * nn + j];
      s = s > 0 ? 1.0 / sqrt(s) : 1.0;
      for (int i = 0; i < nn; i++) {
        V[2 * (i * nn + j)] = Z[i * nn + j] * s;
        V[2 * (i * nn + j) + 1] = 0.0;
      }
    } else if (wi[j] > 0.0 && j + 1 < nn) {
      double s = 0.0, best = -1.0;
      int kb = 0;
      for (int i = 0; i < nn; i++) {
        const double a = Z[i * nn + j], b = Z[i * nn + j + 1], m2 = a * a + b * b;
        s += m2;
        if (m2 > best) { best = m2; kb = i; }
      }
      s = s > 0 ? 1.0 / sqrt(s) : 1.0;
      const double f = Z[kb * nn + j] * s, g = Z[kb * nn + j + 1] * s;
      const double d = sqrt(f * f + g * g);
      const double rr = f < 0 ? -d : d;
      const double c = d > 0 ? f / rr : 1.0, sn = d > 0 ? g / rr : 0.0;
      for (int i = 0; i < nn; i++) {
        const double a = Z[i * nn + j] * s, b = Z[i * nn + j + 1] * s;
        const double re = c * a + sn * b, im = c * b - sn * a;
        V[2 * (i * nn + j)] = re;
        V[2 * (i * nn + j) + 1] = i == kb ? 0.0 : im;
        V[2 * (i * nn + j + 1)] = re;
        V[2 * (i * nn + j + 1) + 1] = i == kb ? 0.0 : -im;
      }
      j++;
    }
  }
}

}  // namespace eig
