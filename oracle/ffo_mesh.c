/*
 * ffo_mesh.c — oracle restatement of FreeFEM's structured mesh builders.
 * TEST INFRASTRUCTURE ONLY (see ffo.h): what the tests compare against, never the product.
 */
#include "ffo.h"

/* BuildCube, kind = 6 (msh3.cpp:7879-8132).  The reference selects the hex
 * split hex_decoupe[code6] by a search over its 74 candidate splits
 * (msh3.cpp:7891-8014) for the one whose six face diagonals all have code 0;
 * re-running that search over the tables simplexes[58][4] / hex_decoupe[74][7]
 * (msh3.cpp:7682-7742) yields code6 = 70, i.e. simplexes {7,15,19,27,26,11}.
 * Those six tets, in hex-local vertex numbering a+2b+4c (msh3.cpp:8039-8045),
 * are listed here in the reference's order. */
static const int kCubeSplit6[6][4] = {
    {4, 0, 6, 7}, {0, 4, 5, 7}, {1, 0, 5, 7},
    {0, 1, 3, 7}, {2, 0, 3, 7}, {0, 2, 6, 7}};
/* Tet faces, GenericElement<DataTet>::nvface (Mesh3dn.cpp:72). */
static const int kTetFace[4][3] = {{3, 2, 1}, {0, 2, 3}, {3, 1, 0}, {0, 1, 2}};

void ffo_cube_counts(int nx, int ny, int nz, int64_t out[3]) {
  out[0] = (int64_t)(nx + 1) * (ny + 1) * (nz + 1);
  out[1] = (int64_t)6 * nx * ny * nz;
  out[2] = (int64_t)4 * ((int64_t)nx * ny + (int64_t)nx * nz + (int64_t)ny * nz);
}

int ffo_cube(int nx, int ny, int nz, const int32_t labels[6], double *xyz,
             int32_t *tets, int32_t *btri, int32_t *blab) {
  if (nx < 1 || ny < 1 || nz < 1) return -1;
  const int64_t nj = nx + 1, nk = nj * (ny + 1);
  const double xd = 1. / nx, yd = 1. / ny, zd = 1. / nz;
  const int nff[6] = {3, 1, 0, 2, 4, 5}; /* msh3.cpp:8050 */
  int64_t p = 0;
  /* vertices: x0 + xd*i with x0 = 0 (msh3.cpp:8024-8044) */
  for (int k = 0; k <= nz; ++k)
    for (int j = 0; j <= ny; ++j)
      for (int i = 0; i <= nx; ++i, ++p) {
        xyz[3 * p + 0] = 0. + xd * i;
        xyz[3 * p + 1] = 0. + yd * j;
        xyz[3 * p + 2] = 0. + zd * k;
      }
  int64_t t = 0, kf = 0;
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        int64_t n8[8];
        for (int c = 0; c < 2; ++c)
          for (int b = 0; b < 2; ++b)
            for (int a = 0; a < 2; ++a)
              n8[a + 2 * b + 4 * c] = (i + a) + nj * (j + b) + nk * (k + c);
        for (int d = 0; d < 6; ++d, ++t) {
          int32_t nu[4];
          for (int q = 0; q < 4; ++q) nu[q] = (int32_t)n8[kCubeSplit6[d][q]];
          for (int q = 0; q < 4; ++q) tets[4 * t + q] = nu[q];
          /* boundary faces: a face whose 3 vertices share exactly one
           * boundary bit (msh3.cpp:8090-8110) */
          for (int f = 0; f < 4; ++f) {
            int32_t nf[3] = {nu[kTetFace[f][0]], nu[kTetFace[f][1]],
                             nu[kTetFace[f][2]]};
            int l = 63;
            for (int e = 0; e < 3; ++e) {
              int64_t v = nf[e];
              int vi = (int)(v % nj), vj = (int)((v / nj) % (ny + 1)),
                  vk = (int)(v / nk);
              int lab = 1 * (vi == 0) + 2 * (vi == nx) + 4 * (vj == 0) +
                        8 * (vj == ny) + 16 * (vk == 0) + 32 * (vk == nz);
              l &= lab;
            }
            for (int b = 0; b < 6; ++b)
              if (l == (1 << b)) {
                btri[3 * kf + 0] = nf[0];
                btri[3 * kf + 1] = nf[1];
                btri[3 * kf + 2] = nf[2];
                blab[kf] = labels[nff[b]];
                ++kf;
              }
          }
        }
      }
  int64_t cnt[3];
  ffo_cube_counts(nx, ny, nz, cnt);
  return (kf == cnt[2] && t == cnt[1]) ? 0 : -2;
}

/* Carre_ (lgmesh.cpp:1229-1384) with flags = 0 ("diag 1") and no transform. */
void ffo_square_counts(int nx, int ny, int64_t out[3]) {
  out[0] = (int64_t)(nx + 1) * (ny + 1);
  out[1] = (int64_t)2 * nx * ny;
  out[2] = (int64_t)2 * (nx + ny);
}

int ffo_square(int nx, int ny, const int32_t labels[4], double *xy,
               int32_t *tris, int32_t *bedges, int32_t *blab) {
  if (nx < 1 || ny < 1) return -1;
  const int nx1 = nx + 1, ny1 = ny + 1;
  int64_t p = 0;
  for (int j = 0; j < ny1; j++)
    for (int i = 0; i < nx1; i++, p++) {
      xy[2 * p + 0] = (double)i / nx;
      xy[2 * p + 1] = (double)j / ny;
    }
  /* "direct" holds for the unit square, so the diag-1 split is (i0,i1,i2),
   * (i0,i2,i3) (lgmesh.cpp:1300-1320). */
  int64_t t = 0;
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++) {
      int32_t i0 = i + j * nx1, i1 = i0 + 1, i2 = i1 + nx1, i3 = i2 - 1;
      tris[3 * t + 0] = i0; tris[3 * t + 1] = i1; tris[3 * t + 2] = i2; ++t;
      tris[3 * t + 0] = i0; tris[3 * t + 1] = i2; tris[3 * t + 2] = i3; ++t;
    }
  int64_t b = 0;
  for (int i = 0; i < nx; i++) { /* bottom */
    bedges[2 * b] = i; bedges[2 * b + 1] = i + 1; blab[b++] = labels[0];
  }
  for (int j = 0; j < ny; j++) { /* right */
    int32_t i1 = nx + j * nx1;
    bedges[2 * b] = i1; bedges[2 * b + 1] = i1 + nx1; blab[b++] = labels[1];
  }
  for (int i = 0; i < nx; i++) { /* up */
    int32_t i1 = i + ny * nx1;
    bedges[2 * b] = i1; bedges[2 * b + 1] = i1 + 1; blab[b++] = labels[2];
  }
  for (int j = 0; j < ny; j++) { /* left */
    int32_t i1 = j * nx1;
    bedges[2 * b] = i1; bedges[2 * b + 1] = i1 + nx1; blab[b++] = labels[3];
  }
  return 0;
}
