// Reads cases from stdin, one per line: a tet's four corners (12 doubles)
// and a plane e (4 doubles), hex floats.  Builds the corners' homogeneous
// points and cofactor permanents from the tet's binary64 face planes as the
// plane-mask kernel does (faces (2,1,3), (0,2,3), (1,0,3), (0,1,2); corners
// from the face triples (1,3,2), (0,2,3), (0,3,1), (0,1,2)) and prints the
// mask decision from the binary32 entry of e (1 = dot(e, X~) < 0 at every
// corner for certain) and the four exact signs of dot(e, X~).
#include <cstdio>

#include "../../libmat_amd/csrc/common/rpd_predicates.h"

int main() {
  static const int tri[4][3] = {{2, 1, 3}, {0, 2, 3}, {1, 0, 3}, {0, 1, 2}};
  static const int cor[4][3] = {{1, 3, 2}, {0, 2, 3}, {0, 3, 1}, {0, 1, 2}};
  double a[16];
  while (true) {
    for (int k = 0; k < 16; ++k)
      if (std::scanf("%la", &a[k]) != 1) return 0;
    rpd_d4 pl[4];
    for (int f = 0; f < 4; ++f) {
      const double* p = a + 3 * tri[f][0];
      const double* q = a + 3 * tri[f][1];
      const double* r = a + 3 * tri[f][2];
      pl[f] = rpd_tri_plane(p[0], p[1], p[2], q[0], q[1], q[2], r[0], r[1],
                            r[2]);
    }
    rpd_d4 X[4], Q[4];
    double sQ[4];
    for (int k = 0; k < 4; ++k) {
      rpd_vertex_homog(pl[cor[k][0]], pl[cor[k][1]], pl[cor[k][2]], X[k],
                       Q[k]);
      sQ[k] = ((Q[k].x + Q[k].y) + Q[k].z) + Q[k].w;
    }
    const rpd_d4 e{a[12], a[13], a[14], a[15]};
    const int mask = rpd_mask_negative(rpd_mask_coef(e.x), rpd_mask_coef(e.y),
                                       rpd_mask_coef(e.z), rpd_mask_coef(e.w),
                                       X, sQ);
    int x[4];
    for (int k = 0; k < 4; ++k)
      x[k] = rpd_exact_det4_sign(pl[cor[k][0]], pl[cor[k][1]], pl[cor[k][2]],
                                 e);
    std::printf("%d %d %d %d %d\n", mask, x[0], x[1], x[2], x[3]);
  }
}
