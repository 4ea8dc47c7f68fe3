// The cube decompositions of the tetrahedron method (get_tetra,
// tetrahedron.f:219-271, mode 1), shared by the Hilbert-mesh weights
// (tetra.hip) and the Fermi-surface weights (drude.hip).
#pragma once
#include "common.hpp"

// the four decompositions of get_tetra (tetrahedron.f:253-268), corners
// 0..7 = offsets (c & 1, (c >> 1) & 1, c >> 2) of the cube, one row of four
// corners per tetrahedron; DEC is the index of the shortest diagonal
// (3-6, 2-7, 1-8, 4-5)
__device__ constexpr int kDecomp[4][6][4] = {
    {{0, 1, 2, 5}, {4, 6, 2, 5}, {0, 4, 2, 5}, {1, 3, 2, 5}, {3, 7, 2, 5}, {6, 7, 2, 5}},
    {{4, 5, 1, 6}, {0, 4, 1, 6}, {0, 2, 1, 6}, {7, 5, 1, 6}, {3, 2, 1, 6}, {7, 3, 1, 6}},
    {{1, 5, 0, 7}, {1, 3, 0, 7}, {2, 3, 0, 7}, {2, 6, 0, 7}, {4, 6, 0, 7}, {4, 5, 0, 7}},
    {{1, 5, 3, 4}, {0, 1, 3, 4}, {0, 2, 3, 4}, {2, 6, 3, 4}, {6, 7, 3, 4}, {5, 7, 3, 4}}};

// position of corner p in tetrahedron it of decomposition DEC, or -1
template <int DEC>
__device__ constexpr int corner_pos(int it, int p) {
  return kDecomp[DEC][it][0] == p   ? 0
         : kDecomp[DEC][it][1] == p ? 1
         : kDecomp[DEC][it][2] == p ? 2
         : kDecomp[DEC][it][3] == p ? 3
                                    : -1;
}

// get_tetra (mode 1) on the host: the decomposition around the shortest
// diagonal of the mesh cell spanned by rlat(:,i)/nkpt3(i), first on ties.
inline int tetra_choice(const int *nkpt3, const double *rlat) {
  double rl[3][3];
  for (int j = 0; j < 3; j++)
    for (int x = 0; x < 3; x++) rl[j][x] = rlat[x + 3 * j] / nkpt3[j];
  static const int sg[4][3] = {{1, -1, 1}, {-1, 1, 1}, {1, 1, 1}, {1, 1, -1}};
  double diag[4];
  for (int d = 0; d < 4; d++) {
    double s = 0.0;
    for (int x = 0; x < 3; x++) {
      const double v = sg[d][0] * rl[0][x] + sg[d][1] * rl[1][x] + sg[d][2] * rl[2][x];
      s += v * v;
    }
    diag[d] = s;
  }
  int idmin = 0;
  for (int d = 1; d < 4; d++)
    if (diag[d] < diag[idmin]) idmin = d;
  return idmin;
}
