/* Seeded hard-sphere configurations of equal unit spheres in a cubic periodic
 * box (benchmark and test workloads, not part of the solver).
 *
 * fcc_mc(): FCC lattice with the fewest cells per side that hold n sites at
 * the target volume fraction, random vacancies down to n spheres, then
 * `sweeps` Metropolis sweeps of hard-sphere displacement moves (cell list,
 * step size adapted towards 35% acceptance during the first half).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static __thread uint64_t rng_state; /* per thread: configurations can be built concurrently */
static inline uint64_t next_u64(void) { /* splitmix64 */
  uint64_t z = (rng_state += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
static inline double uniform(void) { return (next_u64() >> 11) * (1.0 / 9007199254740992.0); }

typedef struct {
  int nc;
  double cs;
  int* head;
  int* next;
  int* prev;
  int* cell;
} Cells;

static int cell_of(const Cells* c, const double* p) {
  int ix = (int)(p[0] / c->cs), iy = (int)(p[1] / c->cs), iz = (int)(p[2] / c->cs);
  if (ix >= c->nc) ix = c->nc - 1;
  if (iy >= c->nc) iy = c->nc - 1;
  if (iz >= c->nc) iz = c->nc - 1;
  return ix + c->nc * (iy + c->nc * iz);
}
static void cell_insert(Cells* c, int i, int k) {
  c->cell[i] = k;
  c->prev[i] = -1;
  c->next[i] = c->head[k];
  if (c->head[k] >= 0) c->prev[c->head[k]] = i;
  c->head[k] = i;
}
static void cell_remove(Cells* c, int i) {
  int k = c->cell[i];
  if (c->prev[i] >= 0) c->next[c->prev[i]] = c->next[i];
  else c->head[k] = c->next[i];
  if (c->next[i] >= 0) c->prev[c->next[i]] = c->prev[i];
}

static int overlaps(const Cells* c, const double* x, int self, const double* p, double L) {
  int k = cell_of(c, p);
  int ix = k % c->nc, iy = (k / c->nc) % c->nc, iz = k / (c->nc * c->nc);
  for (int dz = -1; dz <= 1; ++dz)
    for (int dy = -1; dy <= 1; ++dy)
      for (int dx = -1; dx <= 1; ++dx) {
        int jx = (ix + dx + c->nc) % c->nc, jy = (iy + dy + c->nc) % c->nc, jz = (iz + dz + c->nc) % c->nc;
        for (int j = c->head[jx + c->nc * (jy + c->nc * jz)]; j >= 0; j = c->next[j]) {
          if (j == self) continue;
          double d2 = 0.0;
          for (int a = 0; a < 3; ++a) {
            double d = p[a] - x[3 * j + a];
            d -= L * rint(d / L);
            d2 += d * d;
          }
          if (d2 < 4.0) return 1;
        }
      }
  return 0;
}

/* returns accepted fraction of the last sweep, or -1 on bad input */
double fcc_mc(int n, double phi, uint64_t seed, int sweeps, double* x, double* box_out) {
  if (n < 1 || !(phi > 0.0) || phi > 0.7) return -1.0;
  rng_state = seed * 0x2545F4914F6CDD1DULL + 1;
  const double L = cbrt(n * 4.0 / 3.0 * M_PI / phi);
  int m = 1;
  while (4 * m * m * m < n) ++m;
  const double a = L / m;
  if (a / sqrt(2.0) < 2.0) return -1.0; /* lattice would overlap */
  const int ns = 4 * m * m * m;
  double* sites = (double*)malloc(sizeof(double) * 3 * ns);
  const double basis[4][3] = {{0, 0, 0}, {0.5, 0.5, 0}, {0.5, 0, 0.5}, {0, 0.5, 0.5}};
  int s = 0;
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < m; ++j)
      for (int k = 0; k < m; ++k)
        for (int b = 0; b < 4; ++b) {
          sites[3 * s] = (i + basis[b][0] + 0.25) * a;
          sites[3 * s + 1] = (j + basis[b][1] + 0.25) * a;
          sites[3 * s + 2] = (k + basis[b][2] + 0.25) * a;
          ++s;
        }
  /* random subset of n sites (partial Fisher-Yates) */
  int* idx = (int*)malloc(sizeof(int) * ns);
  for (int i = 0; i < ns; ++i) idx[i] = i;
  for (int i = 0; i < n; ++i) {
    int j = i + (int)(uniform() * (ns - i));
    int t = idx[i]; idx[i] = idx[j]; idx[j] = t;
    memcpy(&x[3 * i], &sites[3 * idx[i]], 3 * sizeof(double));
  }
  free(idx);
  free(sites);
  Cells c;
  c.nc = (int)floor(L / 2.0);
  if (c.nc < 3) c.nc = 3;
  c.cs = L / c.nc;
  const int ncell = c.nc * c.nc * c.nc;
  c.head = (int*)malloc(sizeof(int) * ncell);
  c.next = (int*)malloc(sizeof(int) * n);
  c.prev = (int*)malloc(sizeof(int) * n);
  c.cell = (int*)malloc(sizeof(int) * n);
  for (int k = 0; k < ncell; ++k) c.head[k] = -1;
  for (int i = 0; i < n; ++i) cell_insert(&c, i, cell_of(&c, &x[3 * i]));
  double step = 0.2, frac = 0.0;
  for (int sw = 0; sw < sweeps; ++sw) {
    long acc = 0;
    for (int t = 0; t < n; ++t) {
      int i = (int)(uniform() * n);
      double p[3];
      for (int d = 0; d < 3; ++d) {
        p[d] = x[3 * i + d] + step * (2.0 * uniform() - 1.0);
        p[d] -= L * floor(p[d] / L);
      }
      if (!overlaps(&c, x, i, p, L)) {
        cell_remove(&c, i);
        memcpy(&x[3 * i], p, sizeof(p));
        cell_insert(&c, i, cell_of(&c, p));
        ++acc;
      }
    }
    frac = (double)acc / n;
    if (2 * sw < sweeps) step *= (frac > 0.35) ? 1.05 : 0.95;
    if (step > 0.5 * c.cs) step = 0.5 * c.cs;
  }
  free(c.head); free(c.next); free(c.prev); free(c.cell);
  box_out[0] = box_out[1] = box_out[2] = L;
  return frac;
}

/* minimum pair distance (O(n) with the cell list); for checks */
double min_distance(int n, const double* x, double L) {
  double best = 1e300;
  Cells c;
  c.nc = (int)floor(L / 2.0);
  if (c.nc < 3) c.nc = 3;
  c.cs = L / c.nc;
  const int ncell = c.nc * c.nc * c.nc;
  c.head = (int*)malloc(sizeof(int) * ncell);
  c.next = (int*)malloc(sizeof(int) * n);
  c.prev = (int*)malloc(sizeof(int) * n);
  c.cell = (int*)malloc(sizeof(int) * n);
  for (int k = 0; k < ncell; ++k) c.head[k] = -1;
  for (int i = 0; i < n; ++i) {
    double p[3];
    for (int d = 0; d < 3; ++d) p[d] = x[3 * i + d] - L * floor(x[3 * i + d] / L);
    cell_insert(&c, i, cell_of(&c, p));
  }
  for (int i = 0; i < n; ++i) {
    int k = c.cell[i];
    int ix = k % c.nc, iy = (k / c.nc) % c.nc, iz = k / (c.nc * c.nc);
    for (int dz = -1; dz <= 1; ++dz)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          int jx = (ix + dx + c.nc) % c.nc, jy = (iy + dy + c.nc) % c.nc, jz = (iz + dz + c.nc) % c.nc;
          for (int j = c.head[jx + c.nc * (jy + c.nc * jz)]; j >= 0; j = c.next[j]) {
            if (j <= i) continue;
            double d2 = 0.0;
            for (int a = 0; a < 3; ++a) {
              double d = x[3 * i + a] - x[3 * j + a];
              d -= L * rint(d / L);
              d2 += d * d;
            }
            if (d2 < best) best = d2;
          }
        }
  }
  free(c.head); free(c.next); free(c.prev); free(c.cell);
  return sqrt(best);
}

/* smallest d_ij / (a_i + a_j) over the pairs in neighbouring cells */
static double pair_ratio_min(const Cells* c, const double* x, const double* a, int n, double L) {
  double best = 1e300;
  for (int i = 0; i < n; ++i) {
    int k = c->cell[i];
    int ix = k % c->nc, iy = (k / c->nc) % c->nc, iz = k / (c->nc * c->nc);
    for (int dz = -1; dz <= 1; ++dz)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          int jx = (ix + dx + c->nc) % c->nc, jy = (iy + dy + c->nc) % c->nc, jz = (iz + dz + c->nc) % c->nc;
          for (int j = c->head[jx + c->nc * (jy + c->nc * jz)]; j >= 0; j = c->next[j]) {
            if (j <= i) continue;
            double d2 = 0.0;
            for (int q = 0; q < 3; ++q) {
              double d = x[3 * i + q] - x[3 * j + q];
              d -= L * rint(d / L);
              d2 += d * d;
            }
            double r = sqrt(d2) / (a[i] + a[j]);
            if (r < best) best = r;
          }
        }
  }
  return best;
}

static void cells_init(Cells* c, int n, double L, double amax) {
  c->nc = (int)floor(L / (2.0 * amax));
  if (c->nc < 3) c->nc = 3;
  c->cs = L / c->nc;
  const int ncell = c->nc * c->nc * c->nc;
  c->head = (int*)malloc(sizeof(int) * ncell);
  c->next = (int*)malloc(sizeof(int) * n);
  c->prev = (int*)malloc(sizeof(int) * n);
  c->cell = (int*)malloc(sizeof(int) * n);
}
static void cells_fill(Cells* c, const double* x, int n) {
  const int ncell = c->nc * c->nc * c->nc;
  for (int k = 0; k < ncell; ++k) c->head[k] = -1;
  for (int i = 0; i < n; ++i) cell_insert(c, i, cell_of(c, &x[3 * i]));
}
static void cells_free(Cells* c) { free(c->head); free(c->next); free(c->prev); free(c->cell); }

/* Polydisperse hard spheres: radii a[0..n), cubic periodic box of side L with
 * sum(4/3 pi a^3) / L^3 = phi.  Uniform random positions, then overlap
 * removal: every overlapping pair (d < a_i + a_j) is pushed apart along its
 * axis to the distance (a_i + a_j)(1 + 1e-6), half each, all pairs at once
 * (Jacobi), until no pair overlaps.  Returns the number of relaxation passes,
 * or -1 on bad input or when max_iter passes do not remove every overlap. */
double poly_pack(int n, const double* a, double phi, uint64_t seed, int max_iter, double* x, double* box_out) {
  if (n < 2 || !(phi > 0.0) || phi > 0.55) return -1.0;
  double vol = 0.0, amax = 0.0;
  for (int i = 0; i < n; ++i) {
    if (!(a[i] > 0.0)) return -1.0;
    vol += 4.0 / 3.0 * M_PI * a[i] * a[i] * a[i];
    if (a[i] > amax) amax = a[i];
  }
  const double L = cbrt(vol / phi);
  if (L < 4.0 * amax) return -1.0; /* contact distance must stay below L/2 */
  rng_state = seed * 0x2545F4914F6CDD1DULL + 1;
  for (int i = 0; i < 3 * n; ++i) x[i] = uniform() * L;
  Cells c;
  cells_init(&c, n, L, amax);
  double* dx = (double*)malloc(sizeof(double) * 3 * n);
  int it = 0;
  for (; it < max_iter; ++it) {
    cells_fill(&c, x, n);
    memset(dx, 0, sizeof(double) * 3 * n);
    int any = 0;
    for (int i = 0; i < n; ++i) {
      int k = c.cell[i];
      int ix = k % c.nc, iy = (k / c.nc) % c.nc, iz = k / (c.nc * c.nc);
      for (int ddz = -1; ddz <= 1; ++ddz)
        for (int ddy = -1; ddy <= 1; ++ddy)
          for (int ddx = -1; ddx <= 1; ++ddx) {
            int jx = (ix + ddx + c.nc) % c.nc, jy = (iy + ddy + c.nc) % c.nc, jz = (iz + ddz + c.nc) % c.nc;
            for (int j = c.head[jx + c.nc * (jy + c.nc * jz)]; j >= 0; j = c.next[j]) {
              if (j <= i) continue;
              double r[3], d2 = 0.0;
              for (int q = 0; q < 3; ++q) {
                r[q] = x[3 * i + q] - x[3 * j + q];
                r[q] -= L * rint(r[q] / L);
                d2 += r[q] * r[q];
              }
              const double contact = a[i] + a[j], target = contact * (1.0 + 1e-6);
              if (d2 >= contact * contact) continue;
              any = 1;
              double d = sqrt(d2);
              double u[3];
              if (d > 1e-12) {
                for (int q = 0; q < 3; ++q) u[q] = r[q] / d;
              } else { /* coincident centres: a random axis */
                double z = 2.0 * uniform() - 1.0, t = 2.0 * M_PI * uniform(), w = sqrt(1.0 - z * z);
                u[0] = w * cos(t); u[1] = w * sin(t); u[2] = z;
              }
              const double push = 0.5 * (target - d) * 1.0001;
              for (int q = 0; q < 3; ++q) {
                dx[3 * i + q] += push * u[q];
                dx[3 * j + q] -= push * u[q];
              }
            }
          }
    }
    if (!any) break;
    for (int i = 0; i < 3 * n; ++i) {
      x[i] += dx[i];
      x[i] -= L * floor(x[i] / L);
    }
  }
  free(dx);
  cells_free(&c);
  box_out[0] = box_out[1] = box_out[2] = L;
  return it < max_iter ? (double)it : -1.0;
}

/* smallest d_ij / (a_i + a_j) over all pairs (1e300 if none within the cell
 * range); >= 1 means no overlaps */
double poly_min_ratio(int n, const double* x, const double* a, double L) {
  if (n < 2) return 1e300;
  double amax = 0.0;
  for (int i = 0; i < n; ++i)
    if (a[i] > amax) amax = a[i];
  Cells c;
  cells_init(&c, n, L, amax);
  double* w = (double*)malloc(sizeof(double) * 3 * n);
  for (int i = 0; i < 3 * n; ++i) w[i] = x[i] - L * floor(x[i] / L);
  cells_fill(&c, w, n);
  double r = pair_ratio_min(&c, w, a, n, L);
  free(w);
  cells_free(&c);
  return r;
}
