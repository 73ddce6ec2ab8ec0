// Geometric primitives and sign predicates of the restricted power diagram.
//
// Shared by the HIP kernels and the host-side C++ (neighbour builder).  All
// arithmetic is IEEE binary64 with contraction disabled at build time
// (-ffp-contract=off), so the host and the GPU produce bit-identical planes
// and homogeneous vertices from the same inputs.
//
// Planes follow the reference's conventions:
//  * tet face plane = tri2plane: n = (v1-v0) x (v2-v0), d = -(v0 . n)
//    (include/common_cuda.h:248-254);
//  * bisector of seed S and neighbour B with weights w = r^2:
//    dir = S - B, d = -(dot3(S + B, dir) + (w_B - w_S)) / 2, positive side
//    towards the seed (src/rpd3d/convex_cell.cu:561-592).
//  * a cell vertex is the triple of planes (p1,p2,p3); its homogeneous point
//    X is the cofactor vector of compute_vertex_coordinates
//    (convex_cell.cu:319-351), so det4x4(p1,p2,p3,e) == dot(e, X)
//    (convex_cell.cu:437-500) and "in conflict" means det > 0.
//
// The conflict sign is decided by a dynamic filter with a rigorous forward
// error bound (gamma_10 * sum_r |e_r| * perm_r) and, when the filter cannot
// decide, by exact floating-point expansion arithmetic (two-sum / FMA
// two-product, zero-eliminating expansion sums, after Shewchuk 1997).
#pragma once
#include <stdint.h>
#include <math.h>

#if defined(__HIPCC__)
#define RPD_HD __host__ __device__ __forceinline__
#define RPD_HDI __host__ __device__ inline
#else
#define RPD_HD inline
#define RPD_HDI inline
#endif

struct rpd_d4 {
  double x, y, z, w;
};

// ----------------------------------------------------------------- planes
RPD_HD rpd_d4 rpd_tri_plane(double ax, double ay, double az, double bx,
                            double by, double bz, double cx, double cy,
                            double cz) {
  double ux = bx - ax, uy = by - ay, uz = bz - az;
  double vx = cx - ax, vy = cy - ay, vz = cz - az;
  double nx = uy * vz - uz * vy;
  double ny = uz * vx - ux * vz;
  double nz = ux * vy - uy * vx;
  rpd_d4 p;
  p.x = nx;
  p.y = ny;
  p.z = nz;
  p.w = -(ax * nx + ay * ny + az * nz);
  return p;
}

// S, B: (x, y, z, w = r^2)
RPD_HD rpd_d4 rpd_bisector(rpd_d4 S, rpd_d4 B) {
  double dx = S.x - B.x, dy = S.y - B.y, dz = S.z - B.z;
  double ax = S.x + B.x, ay = S.y + B.y, az = S.z + B.z;
  double dot = ax * dx + ay * dy + az * dz;
  dot = dot + (B.w - S.w);
  rpd_d4 p;
  p.x = dx;
  p.y = dy;
  p.z = dz;
  p.w = -dot * 0.5;
  return p;
}

// power distance |v - c|^2 - w, summed x, y, z in order
RPD_HD double rpd_power_dist(double vx, double vy, double vz, rpd_d4 s) {
  double dx = s.x - vx, dy = s.y - vy, dz = s.z - vz;
  return dx * dx + dy * dy + dz * dz - s.w;
}

// ------------------------------------------------- homogeneous vertices
// 3x3 determinant in the reference's evaluation order, plus its permanent
// (same expression with absolute values), used for the error bound.
RPD_HD void rpd_det3p(double a11, double a12, double a13, double a21,
                      double a22, double a23, double a31, double a32,
                      double a33, double& det, double& perm) {
  double m1 = a22 * a33 - a23 * a32;
  double m2 = a12 * a33 - a13 * a32;
  double m3 = a12 * a23 - a13 * a22;
  double q1 = fabs(a22 * a33) + fabs(a23 * a32);
  double q2 = fabs(a12 * a33) + fabs(a13 * a32);
  double q3 = fabs(a12 * a23) + fabs(a13 * a22);
  det = a11 * m1 - a21 * m2 + a31 * m3;
  perm = fabs(a11) * q1 + fabs(a21) * q2 + fabs(a31) * q3;
}

// Homogeneous point X of the vertex (p1,p2,p3) and permanents Q of its four
// cofactors.  X.xyz / X.w is the vertex position.
RPD_HD void rpd_vertex_homog(rpd_d4 p1, rpd_d4 p2, rpd_d4 p3, rpd_d4& X,
                             rpd_d4& Q) {
  double d, q;
  rpd_det3p(p1.w, p1.y, p1.z, p2.w, p2.y, p2.z, p3.w, p3.y, p3.z, d, q);
  X.x = -d;
  Q.x = q;
  rpd_det3p(p1.x, p1.w, p1.z, p2.x, p2.w, p2.z, p3.x, p3.w, p3.z, d, q);
  X.y = -d;
  Q.y = q;
  rpd_det3p(p1.x, p1.y, p1.w, p2.x, p2.y, p2.w, p3.x, p3.y, p3.w, d, q);
  X.z = -d;
  Q.z = q;
  rpd_det3p(p1.x, p1.y, p1.z, p2.x, p2.y, p2.z, p3.x, p3.y, p3.z, d, q);
  X.w = d;
  Q.w = q;
}

// Error constant for dot(e, X~): every monomial of det4x4 passes through at
// most 10 roundings (5 in the cofactor, 1 product and 3 sums in the dot, and
// a safety rounding), so |computed - exact| <= gamma_10 * sum|e_r|Q_r with
// gamma_10 = 10u/(1-10u) = 1.1102e-15; 1.2e-15 also covers the rounding of
// the bound itself.
#define RPD_FILTER_C 1.2e-15
#define RPD_FILTER_ABS 1e-290
#define RPD_FILTER_MAX 1e280

// Returns +1 if dot(e, X) > 0 for certain, -1 if < 0 for certain, 0 if the
// filter cannot decide (the caller then uses rpd_exact_det4_sign).
RPD_HD int rpd_filtered_sign4(rpd_d4 e, rpd_d4 X, rpd_d4 Q) {
  double s = e.x * X.x + e.y * X.y + e.z * X.z + e.w * X.w;
  double b =
      fabs(e.x) * Q.x + fabs(e.y) * Q.y + fabs(e.z) * Q.z + fabs(e.w) * Q.w;
  double eps = RPD_FILTER_C * b + RPD_FILTER_ABS;
  if (!(b < RPD_FILTER_MAX)) return 0;  // overflow or NaN: undecided
  if (s > eps) return 1;
  if (s < -eps) return -1;
  return 0;
}

// Sign of X.w (the 3x3 determinant of the normals), filtered: 5 roundings.
RPD_HD int rpd_filtered_sign_w(rpd_d4 X, rpd_d4 Q) {
  double eps = 6.0e-16 * Q.w + RPD_FILTER_ABS;
  if (!(Q.w < RPD_FILTER_MAX)) return 0;
  if (X.w > eps) return 1;
  if (X.w < -eps) return -1;
  return 0;
}

// ------------------------------------------------ plane masks (binary32)
// A bisector component rounded to binary32 for the plane masks: zero stays
// zero; magnitudes in [1e-30, 1e30] round to nearest (relative error at
// most 2^-24, no underflow or overflow); others become NaN, which
// rpd_mask_negative never takes for certain.
RPD_HD float rpd_mask_coef(double v) {
  const double a = fabs(v);
  return v == 0.0 ? 0.0f
                  : (a >= 1e-30 && a <= 1e30 ? (float)v : __builtin_nanf(""));
}

// dot(e, X~_k) < 0 for certain at all four corners k of a tet cell
// (X_k = rpd_vertex_homog of the corner's face planes, sQ_k the sum of its
// four cofactor permanents), decided from f = rpd_mask_coef(e) alone
// (|f_r - e_r| <= 5.97e-8 |f_r|): the filter of rpd_filtered_sign4 for
// dot(f, X~), its bound sum|f_r| Q_r replaced by the larger
// max|f_r| * sum Q_r (the computed product and sum short by at most 4u
// relative; products may be fused into the sums, no monomial then passing
// through more roundings), plus |dot(f - e, X~)| <= 5.97e-8 (1 + 1e-15)
// max|f_r| sum Q_r (|X~_r| <= Q_r (1 + 6e-16)); 6e-8 covers both, 1e-250
// the absolute terms.  A NaN component decides nothing.
RPD_HD bool rpd_mask_negative(double fx, double fy, double fz, double fw,
                              const rpd_d4 X[4], const double sQ[4]) {
#if defined(__clang__)
#pragma clang fp contract(fast)
#endif
  const double m = fmax(fmax(fabs(fx), fabs(fy)), fmax(fabs(fz), fabs(fw)));
  bool neg = true;
#if defined(__clang__)
#pragma unroll
#endif
  for (int k = 0; k < 4 && neg; ++k) {
    const double s = fx * X[k].x + fy * X[k].y + fz * X[k].z + fw * X[k].w;
    const double b = m * sQ[k];
    neg = b < RPD_FILTER_MAX && s < -(6e-8 * b + 1e-250);
  }
  return neg;
}

// ---------------------------------------------- expansion arithmetic
// Components are stored in increasing magnitude, non-overlapping, zero-free
// (a zero value is the one-component expansion {0}).
RPD_HD void rpd_two_sum(double a, double b, double& x, double& y) {
  x = a + b;
  double bv = x - a;
  double av = x - bv;
  double br = b - bv;
  double ar = a - av;
  y = ar + br;
}
RPD_HD void rpd_fast_two_sum(double a, double b, double& x, double& y) {
  x = a + b;
  double bv = x - a;
  y = b - bv;
}
RPD_HD void rpd_two_prod(double a, double b, double& x, double& y) {
  x = a * b;
  y = fma(a, b, -x);
}

// h = e + f   (h needs room for elen + flen components)
RPD_HDI int rpd_exp_sum(int elen, const double* e, int flen, const double* f,
                        double* h) {
  double Q, Qn, hh, enow, fnow;
  int ei = 0, fi = 0, hi = 0;
  enow = e[0];
  fnow = f[0];
  if ((fnow > enow) == (fnow > -enow)) {
    Q = enow;
    ++ei;
    enow = ei < elen ? e[ei] : 0.0;
  } else {
    Q = fnow;
    ++fi;
    fnow = fi < flen ? f[fi] : 0.0;
  }
  if (ei < elen && fi < flen) {
    if ((fnow > enow) == (fnow > -enow)) {
      rpd_fast_two_sum(enow, Q, Qn, hh);
      ++ei;
      enow = ei < elen ? e[ei] : 0.0;
    } else {
      rpd_fast_two_sum(fnow, Q, Qn, hh);
      ++fi;
      fnow = fi < flen ? f[fi] : 0.0;
    }
    Q = Qn;
    if (hh != 0.0) h[hi++] = hh;
    while (ei < elen && fi < flen) {
      if ((fnow > enow) == (fnow > -enow)) {
        rpd_two_sum(Q, enow, Qn, hh);
        ++ei;
        enow = ei < elen ? e[ei] : 0.0;
      } else {
        rpd_two_sum(Q, fnow, Qn, hh);
        ++fi;
        fnow = fi < flen ? f[fi] : 0.0;
      }
      Q = Qn;
      if (hh != 0.0) h[hi++] = hh;
    }
  }
  while (ei < elen) {
    rpd_two_sum(Q, enow, Qn, hh);
    ++ei;
    enow = ei < elen ? e[ei] : 0.0;
    Q = Qn;
    if (hh != 0.0) h[hi++] = hh;
  }
  while (fi < flen) {
    rpd_two_sum(Q, fnow, Qn, hh);
    ++fi;
    fnow = fi < flen ? f[fi] : 0.0;
    Q = Qn;
    if (hh != 0.0) h[hi++] = hh;
  }
  if (Q != 0.0 || hi == 0) h[hi++] = Q;
  return hi;
}

// h = b * e   (h needs room for 2 * elen components)
RPD_HDI int rpd_exp_scale(int elen, const double* e, double b, double* h) {
  double Q, hh, p1, p0, sum;
  int hi = 0;
  rpd_two_prod(e[0], b, Q, hh);
  if (hh != 0.0) h[hi++] = hh;
  for (int i = 1; i < elen; ++i) {
    rpd_two_prod(e[i], b, p1, p0);
    rpd_two_sum(Q, p0, sum, hh);
    if (hh != 0.0) h[hi++] = hh;
    rpd_fast_two_sum(p1, sum, Q, hh);
    if (hh != 0.0) h[hi++] = hh;
  }
  if (Q != 0.0 || hi == 0) h[hi++] = Q;
  return hi;
}

// h = a*b - c*d exactly (<= 4 components)
RPD_HDI int rpd_exp_2x2(double a, double b, double c, double d, double* h) {
  double x1, x0, y1, y0;
  rpd_two_prod(a, b, x1, x0);
  rpd_two_prod(c, d, y1, y0);
  double e[2], f[2];
  int el = 0, fl = 0;
  if (x0 != 0.0) e[el++] = x0;
  if (x1 != 0.0 || el == 0) e[el++] = x1;
  if (y0 != 0.0) f[fl++] = -y0;
  if (y1 != 0.0 || fl == 0) f[fl++] = -y1;
  return rpd_exp_sum(el, e, fl, f, h);
}

// exact 3x3 determinant a11*(a22 a33 - a23 a32) - a21*(a12 a33 - a13 a32)
//                     + a31*(a12 a23 - a13 a22)   (<= 24 components)
RPD_HDI int rpd_exp_det3(double a11, double a12, double a13, double a21,
                         double a22, double a23, double a31, double a32,
                         double a33, double* h) {
  double m[4], t1[8], t2[8], t3[8], s12[16];
  int ml, l1, l2, l3, l12;
  ml = rpd_exp_2x2(a22, a33, a23, a32, m);
  l1 = rpd_exp_scale(ml, m, a11, t1);
  ml = rpd_exp_2x2(a12, a33, a13, a32, m);
  l2 = rpd_exp_scale(ml, m, -a21, t2);
  ml = rpd_exp_2x2(a12, a23, a13, a22, m);
  l3 = rpd_exp_scale(ml, m, a31, t3);
  l12 = rpd_exp_sum(l1, t1, l2, t2, s12);
  return rpd_exp_sum(l12, s12, l3, t3, h);
}

RPD_HD int rpd_exp_sign(int len, const double* h) {
  double top = h[len - 1];
  return top > 0.0 ? 1 : (top < 0.0 ? -1 : 0);
}

// Exact sign of det4x4(p1,p2,p3,e) == dot(e, X) where X are the cofactors
// of rpd_vertex_homog.  Uses ~3.7 KB of private storage.
RPD_HDI int rpd_exact_det4_sign(rpd_d4 p1, rpd_d4 p2, rpd_d4 p3, rpd_d4 e) {
  double c[24], sc[48], acc[192], tmp[192];
  int cl, sl, al = 1, tl;
  acc[0] = 0.0;
  // X.x = -det3(w,y,z rows)
  cl = rpd_exp_det3(p1.w, p1.y, p1.z, p2.w, p2.y, p2.z, p3.w, p3.y, p3.z, c);
  sl = rpd_exp_scale(cl, c, -e.x, sc);
  tl = rpd_exp_sum(al, acc, sl, sc, tmp);
  // X.y = -det3(x,w,z rows)
  cl = rpd_exp_det3(p1.x, p1.w, p1.z, p2.x, p2.w, p2.z, p3.x, p3.w, p3.z, c);
  sl = rpd_exp_scale(cl, c, -e.y, sc);
  al = rpd_exp_sum(tl, tmp, sl, sc, acc);
  // X.z = -det3(x,y,w rows)
  cl = rpd_exp_det3(p1.x, p1.y, p1.w, p2.x, p2.y, p2.w, p3.x, p3.y, p3.w, c);
  sl = rpd_exp_scale(cl, c, -e.z, sc);
  tl = rpd_exp_sum(al, acc, sl, sc, tmp);
  // X.w = det3(x,y,z rows)
  cl = rpd_exp_det3(p1.x, p1.y, p1.z, p2.x, p2.y, p2.z, p3.x, p3.y, p3.z, c);
  sl = rpd_exp_scale(cl, c, e.w, sc);
  al = rpd_exp_sum(tl, tmp, sl, sc, acc);
  return rpd_exp_sign(al, acc);
}

// Exact sign of X.w = det3(x,y,z rows) of the three normals.
RPD_HDI int rpd_exact_w_sign(rpd_d4 p1, rpd_d4 p2, rpd_d4 p3) {
  double c[24];
  int cl =
      rpd_exp_det3(p1.x, p1.y, p1.z, p2.x, p2.y, p2.z, p3.x, p3.y, p3.z, c);
  return rpd_exp_sign(cl, c);
}
