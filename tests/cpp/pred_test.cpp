// Reads 4x4 matrices (16 doubles, row-major, hex floats) from stdin, one
// per line, and prints the product header's exact sign of det4x4 computed
// as dot(e, X) of rpd_predicates.h, and its filtered sign (0 = undecided).
// Rows of the input are the x, y, z, w components; columns the planes
// p1, p2, p3, e.
#include <cstdio>

#include "../../libmat_amd/csrc/common/rpd_predicates.h"

int main() {
  double a[16];
  while (true) {
    for (int k = 0; k < 16; ++k)
      if (std::scanf("%la", &a[k]) != 1) return 0;
    rpd_d4 p1{a[0], a[4], a[8], a[12]}, p2{a[1], a[5], a[9], a[13]},
        p3{a[2], a[6], a[10], a[14]}, e{a[3], a[7], a[11], a[15]};
    rpd_d4 X, Q;
    rpd_vertex_homog(p1, p2, p3, X, Q);
    int f = rpd_filtered_sign4(e, X, Q);
    int x = rpd_exact_det4_sign(p1, p2, p3, e);
    int w = rpd_exact_w_sign(p1, p2, p3);
    std::printf("%d %d %d\n", x, f, w);
  }
}
