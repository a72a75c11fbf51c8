/* Marching cubes, CPU restatement — TEST INFRASTRUCTURE (the oracle).
 *
 * Restates the marching-cubes stage the reference runs at
 * projects/neuralangelo/utils/mesh.py:119-121 (mcubes.marching_cubes(volume, 0.))
 * on each lattice block of projects/neuralangelo/utils/mesh.py:24-49.  The
 * arithmetic lives in the un-vendored PyMCubes package (requirements.txt:23,
 * no version pin), restated here from its published algorithm:
 *   - cell (i,j,k) has corners v0..v7 in Bourke order, cube index bit m set
 *     iff value[m] <= iso;
 *   - cells are visited with i outermost, then j, then k;
 *   - per cell the vertices on the three "far" edges 6, 5, 10 are always
 *     created (in that order); edges 0,1,2,3,4,7,8,9,11 create a vertex only
 *     on the lattice boundary where no earlier cell owns them (0: j=k=0;
 *     1, 2: k=0; 3: i=k=0; 4: j=0; 7: i=0; 8: i=j=0; 9: j=0; 11: i=0), in
 *     that order, and otherwise reuse the index recorded by the owning cell;
 *   - a vertex on edge (a -> b) is x_a + (x_b - x_a) * (iso - f_a) / (f_b - f_a)
 *     in double precision, array-index coordinates;
 *   - triangles follow the case table order; faces index the vertex list.
 * Exact PyMCubes emission order cannot be checked here (the package is
 * absent); this restatement defines the canonical order the GPU kernel must
 * reproduce bit-for-bit.
 *
 * Build: oracle/build.py (gcc -O2 -shared), output under oracle/_build/.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../neuralangelo_amd/csrc/mc_tables.h"

static const int kCorner[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                  {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};

/* axis along which edge e runs: x edges 0,2,4,6; y edges 1,3,5,7; z edges 8..11 */
static int edge_axis(int e) { return e < 8 ? (e & 1) : 2; }

static double interp(double x1, double x2, double f1, double f2, double iso) {
  if (f2 == f1) return (x1 + x2) / 2;
  volatile double d = x2 - x1;
  volatile double n = iso - f1;
  volatile double q = f2 - f1;
  volatile double t = d * n;
  volatile double r = t / q;
  return r + x1;
}

/* index of the shared-edge slot for lattice edge along `axis` at lattice point (a,b,c) */
typedef struct {
  int64_t nx, ny, nz;
  int64_t* xs; /* x-edges: [nx-1][ny][nz] */
  int64_t* ys; /* y-edges: [nx][ny-1][nz] */
  int64_t* zs; /* z-edges: [nx][ny][nz-1] */
} Shared;

static int64_t* slot(Shared* s, int axis, int64_t a, int64_t b, int64_t c) {
  if (axis == 0) return &s->xs[(a * s->ny + b) * s->nz + c];
  if (axis == 1) return &s->ys[(a * (s->ny - 1) + b) * s->nz + c];
  return &s->zs[(a * s->ny + b) * (s->nz - 1) + c];
}

static int owned(int e, int64_t i, int64_t j, int64_t k) {
  switch (e) {
    case 6: case 5: case 10: return 1;
    case 0: return j == 0 && k == 0;
    case 1: case 2: return k == 0;
    case 3: return i == 0 && k == 0;
    case 4: return j == 0;
    case 7: return i == 0;
    case 8: return i == 0 && j == 0;
    case 9: return j == 0;
    case 11: return i == 0;
  }
  return 0;
}

static const int kOrder[12] = {6, 5, 10, 0, 1, 2, 3, 4, 7, 8, 9, 11};

/* Returns 0 on success.  With verts/faces NULL only counts are produced. */
int mc_oracle(const float* vol, int64_t nx, int64_t ny, int64_t nz, double iso, double* verts,
              int64_t* faces, int64_t* nv_out, int64_t* nf_out) {
  *nv_out = 0;
  *nf_out = 0;
  if (nx < 2 || ny < 2 || nz < 2) return 0;
  Shared sh;
  sh.nx = nx; sh.ny = ny; sh.nz = nz;
  sh.xs = (int64_t*)malloc(sizeof(int64_t) * (nx - 1) * ny * nz);
  sh.ys = (int64_t*)malloc(sizeof(int64_t) * nx * (ny - 1) * nz);
  sh.zs = (int64_t*)malloc(sizeof(int64_t) * nx * ny * (nz - 1));
  if (!sh.xs || !sh.ys || !sh.zs) return -1;
  int64_t nv = 0, nf = 0;
  for (int64_t i = 0; i < nx - 1; ++i)
    for (int64_t j = 0; j < ny - 1; ++j)
      for (int64_t k = 0; k < nz - 1; ++k) {
        double v[8];
        int cube = 0;
        for (int m = 0; m < 8; ++m) {
          v[m] = (double)vol[((i + kCorner[m][0]) * ny + (j + kCorner[m][1])) * nz +
                             (k + kCorner[m][2])];
          if (v[m] <= iso) cube |= 1 << m;
        }
        const int edges = kMcEdgeTable[cube];
        int64_t idx[12];
        for (int q = 0; q < 12; ++q) {
          const int e = kOrder[q];
          if (!(edges & (1 << e))) continue;
          const int a = kMcEdgeCorners[e][0], b = kMcEdgeCorners[e][1];
          const int ax = edge_axis(e);
          /* lattice coordinates of the edge's low end along its axis */
          int64_t p[3] = {i + kCorner[a][0], j + kCorner[a][1], k + kCorner[a][2]};
          int64_t pb[3] = {i + kCorner[b][0], j + kCorner[b][1], k + kCorner[b][2]};
          int64_t lo[3] = {p[0] < pb[0] ? p[0] : pb[0], p[1] < pb[1] ? p[1] : pb[1],
                           p[2] < pb[2] ? p[2] : pb[2]};
          int64_t* sl = slot(&sh, ax, lo[0], lo[1], lo[2]);
          if (owned(e, i, j, k)) {
            idx[e] = nv;
            *sl = nv;
            if (verts) {
              double* o = verts + 3 * nv;
              o[0] = (double)p[0];
              o[1] = (double)p[1];
              o[2] = (double)p[2];
              o[ax] = interp((double)p[ax], (double)pb[ax], v[a], v[b], iso);
            }
            ++nv;
          } else {
            idx[e] = *sl;
          }
        }
        const int8_t* t = kMcTriTable[cube];
        for (int q = 0; q < 16 && t[q] >= 0; q += 3) {
          if (faces) {
            faces[3 * nf] = idx[t[q]];
            faces[3 * nf + 1] = idx[t[q + 1]];
            faces[3 * nf + 2] = idx[t[q + 2]];
          }
          ++nf;
        }
      }
  free(sh.xs);
  free(sh.ys);
  free(sh.zs);
  *nv_out = nv;
  *nf_out = nf;
  return 0;
}
