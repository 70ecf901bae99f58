/*
 * bf_io.c -- the reference's case files, VTK frames and state checkpoints (see
 * include/biofem_io.h).  Plain C, host only.
 */
#define _POSIX_C_SOURCE 200809L
#include "../../include/biofem_io.h"

#include <errno.h>
#include <limits.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <sys/types.h>

#define BF_NO_NEIGHBOUR 1000000 /* Input's marker for "no patch node" (src/io.c:478) */
#define VEC_FILE_CLASSID 1211214 /* PETSc's class id of a Vec in a binary file */

static void list_path(char *out, size_t n, const char *dir, const char *kind, int ibi) {
  snprintf(out, n, "%s/%s%2.2d", dir, kind, ibi);  /* "%s/nlist%2.2d" (src/io.c:44) */
}

/* the first integer of a file, the rest of that line skipped */
static int first_int(FILE *f, int *v) {
  if (fscanf(f, "%d", v) != 1) return 0;
  int c;
  while ((c = fgetc(f)) != EOF && c != '\n') {
  }
  return 1;
}

int bf_io_dimension(const char *in_dir, int ibi, int *n_v, int *n_elmt, int *n_edge,
                    int *n_ghosts, int *sum_n_bnodes) {
  if (!in_dir) return BF_ERR_ARG;
  char fn[4096];
  int nv = 0, ne = 0, nedge = 0, ng = 0;
  long long sum = 0;
  list_path(fn, sizeof fn, in_dir, "nlist", ibi);
  FILE *f = fopen(fn, "r");
  if (!f) return BF_ERR_IO;
  int ok = first_int(f, &nv);
  fclose(f);
  if (!ok || nv < 0) return BF_ERR_IO;
  list_path(fn, sizeof fn, in_dir, "elist", ibi);
  f = fopen(fn, "r");
  if (!f) return BF_ERR_IO;
  ok = first_int(f, &ne);
  fclose(f);
  if (!ok || ne < 0) return BF_ERR_IO;
  list_path(fn, sizeof fn, in_dir, "blist", ibi);
  f = fopen(fn, "r");
  if (!f) return BF_ERR_IO;
  ok = first_int(f, &nedge) && first_int(f, &ng) && nedge >= 0 && ng >= 0;
  for (int i = 0; ok && i < nedge; i++) {
    int c = 0;
    ok = fscanf(f, "%d", &c) == 1 && c >= 0;
    sum += c;
    ok = ok && sum <= INT_MAX;  /* the list lengths size the caller's arrays */
  }
  fclose(f);
  if (!ok) return BF_ERR_IO;
  if (n_v) *n_v = nv;
  if (n_elmt) *n_elmt = ne;
  if (n_edge) *n_edge = nedge;
  if (n_ghosts) *n_ghosts = ng;
  if (sum_n_bnodes) *sum_n_bnodes = (int)sum;
  return BF_OK;
}

/* ---- edge table for the neighbour search ---- */
typedef struct {
  uint64_t key;  /* (lo + 1) << 32 | (hi + 1); 0 = empty */
  int id;        /* edge number: row of the incident-triangle lists */
  int last_e;    /* last triangle counted on the edge (a triangle lists an edge once) */
} EdgeSlot;

static uint64_t edge_key(int a, int b) {
  const uint32_t lo = (uint32_t)(a < b ? a : b), hi = (uint32_t)(a < b ? b : a);
  return ((uint64_t)(lo + 1u) << 32) | (uint64_t)(hi + 1u);
}

static EdgeSlot *edge_find(EdgeSlot *t, size_t cap, uint64_t key, int insert, int *n_edges) {
  size_t h = (size_t)((key * 0x9E3779B97F4A7C15ull) >> 17) & (cap - 1);
  for (;;) {
    if (t[h].key == key) return &t[h];
    if (t[h].key == 0) {
      if (!insert) return NULL;
      t[h].key = key;
      t[h].id = (*n_edges)++;
      t[h].last_e = -1;
      return &t[h];
    }
    h = (h + 1) & (cap - 1);
  }
}

/* nv4/nv5/nv6 of Input (src/io.c:476-519): for element i and each element j in increasing
 * order, mn counts i's corners found among j's; mn == 2 writes j's first corner that is
 * not one of i's to nv4 (shared corners nv2, nv3), nv5 (nv1, nv3) or nv6 (nv1, nv2), so
 * the last such j wins.  Every j with mn == 2 holds both shared corners, so it is listed
 * under that pair in an edge table of all incident triangles (ascending): the search
 * scans the pair's list from its end for the first j with mn == 2.  A triangle with a
 * repeated corner (not a valid surface triangle) gets no neighbour across that pair. */
static int neighbours(int n_elmt, const int *nv1, const int *nv2, const int *nv3, int *nv4,
                      int *nv5, int *nv6) {
  size_t cap = 16;
  while (cap < (size_t)n_elmt * 6 + 16) cap <<= 1;
  EdgeSlot *t = (EdgeSlot *)calloc(cap, sizeof(EdgeSlot));
  int *ptr = (int *)calloc((size_t)3 * n_elmt + 2, sizeof(int));
  int *lst = (int *)malloc(((size_t)3 * n_elmt + 1) * sizeof(int));
  if (!t || !ptr || !lst) {
    free(t); free(ptr); free(lst);
    return BF_ERR_NOMEM;
  }
  int n_edges = 0;
  /* pass 1: edge numbers and the number of triangles per edge */
  for (int j = 0; j < n_elmt; j++) {
    const int c[3] = {nv1[j], nv2[j], nv3[j]};
    for (int k = 0; k < 3; k++) {
      const int a = c[k], b = c[(k + 1) % 3];
      if (a == b) continue;
      EdgeSlot *s = edge_find(t, cap, edge_key(a, b), 1, &n_edges);
      if (s->last_e == j) continue;
      s->last_e = j;
      ptr[s->id + 1]++;
    }
  }
  for (int e = 0; e < n_edges; e++) ptr[e + 1] += ptr[e];
  /* pass 2: the lists, ascending in j */
  for (size_t h = 0; h < cap; h++) t[h].last_e = -1;
  {
    int *fill = (int *)malloc(((size_t)n_edges + 1) * sizeof(int));
    if (!fill) {
      free(t); free(ptr); free(lst);
      return BF_ERR_NOMEM;
    }
    memcpy(fill, ptr, (size_t)n_edges * sizeof(int));
    for (int j = 0; j < n_elmt; j++) {
      const int c[3] = {nv1[j], nv2[j], nv3[j]};
      for (int k = 0; k < 3; k++) {
        const int a = c[k], b = c[(k + 1) % 3];
        if (a == b) continue;
        EdgeSlot *s = edge_find(t, cap, edge_key(a, b), 0, &n_edges);
        if (s->last_e == j) continue;
        s->last_e = j;
        lst[fill[s->id]++] = j;
      }
    }
    free(fill);
  }
  for (int i = 0; i < n_elmt; i++) {
    const int c[3] = {nv1[i], nv2[i], nv3[i]};
    /* pair without corner k: (nv2,nv3) -> nv4, (nv1,nv3) -> nv5, (nv1,nv2) -> nv6 */
    int out[3] = {BF_NO_NEIGHBOUR, BF_NO_NEIGHBOUR, BF_NO_NEIGHBOUR};
    for (int k = 0; k < 3; k++) {
      const int a = c[(k + 1) % 3], b = c[(k + 2) % 3];
      if (a == b) continue;
      EdgeSlot *s = edge_find(t, cap, edge_key(a, b), 0, &n_edges);
      if (!s) continue;
      for (int p = ptr[s->id + 1] - 1; p >= ptr[s->id]; p--) {
        const int j = lst[p];
        const int d[3] = {nv1[j], nv2[j], nv3[j]};
        int mn = 0;
        int in_j[3];
        for (int q = 0; q < 3; q++) {
          in_j[q] = (c[q] == d[0] || c[q] == d[1] || c[q] == d[2]);
          mn += in_j[q];
        }
        /* mn == 2 with a and b shared means corner k is the one missing from j */
        if (mn != 2 || in_j[k]) continue;
        int npe = d[2];
        if (d[0] != c[0] && d[0] != c[1] && d[0] != c[2])
          npe = d[0];
        else if (d[1] != c[0] && d[1] != c[1] && d[1] != c[2])
          npe = d[1];
        out[k] = npe;
        break;
      }
    }
    if (nv4) nv4[i] = out[0];
    if (nv5) nv5[i] = out[1];
    if (nv6) nv6[i] = out[2];
  }
  free(t);
  free(ptr);
  free(lst);
  return BF_OK;
}

int bf_io_input(const char *in_dir, int ibi, const double *char_length, double *x, double *y,
                double *z, int *nv1, int *nv2, int *nv3, int *n_bnodes, int *bnodes, int *nv4,
                int *nv5, int *nv6) {
  if (!in_dir || !x || !y || !z || !nv1 || !nv2 || !nv3) return BF_ERR_ARG;
  int n_v, n_elmt, n_edge, n_ghosts, sum;
  int rc = bf_io_dimension(in_dir, ibi, &n_v, &n_elmt, &n_edge, &n_ghosts, &sum);
  if (rc != BF_OK) return rc;
  if ((n_edge > 0 && !n_bnodes) || (sum > 0 && !bnodes)) return BF_ERR_ARG;
  const double cl[3] = {char_length ? char_length[0] : 1.0, char_length ? char_length[1] : 1.0,
                        char_length ? char_length[2] : 1.0};
  char fn[4096];
  /* nlist: n_v, then "i 0 0 x y z" per node (src/io.c:382-391) */
  list_path(fn, sizeof fn, in_dir, "nlist", ibi);
  FILE *f = fopen(fn, "r");
  if (!f) return BF_ERR_IO;
  int n = 0, ok = first_int(f, &n);
  for (int i = 0; ok && i < n_v; i++) {
    int i1, i2, i3;
    double a, b, c;
    ok = fscanf(f, "%d %d %d %le %le %le", &i1, &i2, &i3, &a, &b, &c) == 6;
    if (ok) {
      x[i] = a / cl[0] + 0.0;
      y[i] = b / cl[1] + 0.0;
      z[i] = c / cl[2] + 0.0;
    }
  }
  fclose(f);
  if (!ok) return BF_ERR_IO;
  /* elist: n_elmt, then 15 integers per element, corners 12..14 (1-based) (:422-427) */
  list_path(fn, sizeof fn, in_dir, "elist", ibi);
  f = fopen(fn, "r");
  if (!f) return BF_ERR_IO;
  ok = first_int(f, &n);
  for (int e = 0; ok && e < n_elmt; e++) {
    int v[15];
    for (int k = 0; ok && k < 15; k++) ok = fscanf(f, "%d", &v[k]) == 1;
    if (ok) {
      nv1[e] = v[11] - 1;
      nv2[e] = v[12] - 1;
      nv3[e] = v[13] - 1;
      ok = nv1[e] >= 0 && nv1[e] < n_v && nv2[e] >= 0 && nv2[e] < n_v && nv3[e] >= 0 &&
           nv3[e] < n_v;
    }
  }
  fclose(f);
  if (!ok) return BF_ERR_IO;
  /* blist: n_edge, n_ghosts, n_edge list lengths, then the nodes (1-based) (:435-466) */
  list_path(fn, sizeof fn, in_dir, "blist", ibi);
  f = fopen(fn, "r");
  if (!f) return BF_ERR_IO;
  int m1, m2;
  ok = first_int(f, &m1) && first_int(f, &m2);
  for (int i = 0; ok && i < n_edge; i++) ok = fscanf(f, "%d", &n_bnodes[i]) == 1;
  for (int i = 0; ok && i < sum; i++) {
    ok = fscanf(f, "%d", &bnodes[i]) == 1;
    if (ok) {
      bnodes[i] -= 1;
      ok = bnodes[i] >= 0 && bnodes[i] < n_v;
    }
  }
  fclose(f);
  if (!ok) return BF_ERR_IO;
  if (nv4 || nv5 || nv6) return neighbours(n_elmt, nv1, nv2, nv3, nv4, nv5, nv6);
  return BF_OK;
}

int bf_io_write_lists(const char *dir, int ibi, int n_v, const double *x, const double *y,
                      const double *z, int n_elmt, const int *nv1, const int *nv2,
                      const int *nv3, int n_edge, const int *n_bnodes, const int *bnodes,
                      int n_ghosts) {
  if (!dir || n_v < 0 || n_elmt < 0 || n_edge < 0 || !x || !y || !z || !nv1 || !nv2 || !nv3 ||
      (n_edge > 0 && (!n_bnodes || !bnodes)))
    return BF_ERR_ARG;
  char fn[4096];
  list_path(fn, sizeof fn, dir, "nlist", ibi);
  FILE *f = fopen(fn, "w");
  if (!f) return BF_ERR_IO;
  fprintf(f, "%d\n", n_v);
  for (int i = 0; i < n_v; i++)
    fprintf(f, "%d\t0\t0\t%.17g %.17g %.17g\n", i + 1, x[i], y[i], z[i]);
  if (fclose(f) != 0) return BF_ERR_IO;
  list_path(fn, sizeof fn, dir, "elist", ibi);
  f = fopen(fn, "w");
  if (!f) return BF_ERR_IO;
  fprintf(f, "%d\n", n_elmt);
  for (int e = 0; e < n_elmt; e++)
    fprintf(f, "1\t1\t0\t0\t0\t0\t0\t0\t4\t0\t%d\t%d\t%d\t%d\t%d\n", e + 1, nv1[e] + 1,
            nv2[e] + 1, nv3[e] + 1, nv3[e] + 1);
  if (fclose(f) != 0) return BF_ERR_IO;
  list_path(fn, sizeof fn, dir, "blist", ibi);
  f = fopen(fn, "w");
  if (!f) return BF_ERR_IO;
  fprintf(f, "%d\n%d\n", n_edge, n_ghosts);
  for (int i = 0; i < n_edge; i++) fprintf(f, "%d\n", n_bnodes[i]);
  int k = 0;
  for (int i = 0; i < n_edge; i++) {
    for (int j = 0; j < n_bnodes[i]; j++) fprintf(f, j ? " %d" : "%d", bnodes[k++] + 1);
    fprintf(f, "\n");
  }
  if (fclose(f) != 0) return BF_ERR_IO;
  return BF_OK;
}

/* ---- VTK ---- */
static double at(const double *a, long i) { return a ? a[i] : 0.0; }

static void vec_block(FILE *f, const char *label, const double *v, int n_v) {
  fprintf(f, "VECTORS %s float\n", label);
  for (int i = 0; i < n_v; i++)
    fprintf(f, "%f %f %f\n", at(v, 3L * i), at(v, 3L * i + 1), at(v, 3L * i + 2));
}

static void scalar_block(FILE *f, const char *label, const double *v, int n) {
  fprintf(f, "SCALARS %s float\n", label);
  fprintf(f, "LOOKUP_TABLE default\n");
  for (int i = 0; i < n; i++) fprintf(f, "%f \n", at(v, i));
}

int bf_io_write_vtk(const char *out_dir, int ibi, int ti, const bf_vtk_frame *fr) {
  if (!fr || fr->n_v < 0 || fr->n_elmt < 0 || !fr->x || !fr->y || !fr->z || !fr->nv1 ||
      !fr->nv2 || !fr->nv3)
    return BF_ERR_ARG;
  const int n_v = fr->n_v, E = fr->n_elmt;
  const char *dir = (out_dir && strlen(out_dir) > 0) ? out_dir : ".";
  mkdir(dir, 0777);
  char fn[4096];
  snprintf(fn, sizeof fn, "%s/surface%2.2d_%5.5d.vtk", dir, ibi, ti);
  FILE *f = fopen(fn, "w");
  if (!f) return BF_ERR_IO;
  fprintf(f, "# vtk DataFile Version 2.0\n");
  fprintf(f, "Surface Grid\n");
  fprintf(f, "ASCII\n");
  fprintf(f, "DATASET UNSTRUCTURED_GRID\n");
  fprintf(f, "POINTS  %d float\n", n_v);
  for (int i = 0; i < n_v; i++) fprintf(f, "%f %f %f\n", fr->x[i], fr->y[i], fr->z[i]);
  fprintf(f, "CELLS %d %d\n", E, 4 * E);
  for (int e = 0; e < E; e++) fprintf(f, "%d  %d %d %d\n", 3, fr->nv1[e], fr->nv2[e], fr->nv3[e]);
  fprintf(f, "CELL_TYPES %d\n", E);
  for (int e = 0; e < E; e++) fprintf(f, "%d\n", 5);
  fprintf(f, "POINT_DATA %d\n", n_v);
  fprintf(f, "SCALARS contact integer\n");
  fprintf(f, "LOOKUP_TABLE default\n");
  for (int i = 0; i < n_v; i++) fprintf(f, "%d\n", fr->contact ? fr->contact[i] : 0);
  scalar_block(f, "m", fr->m, n_v);
  scalar_block(f, "IE", fr->IE, n_v);
  fprintf(f, "VECTORS disp float\n");
  for (int i = 0; i < n_v; i++)
    fprintf(f, "%f %f %f\n", fr->x[i] - at(fr->x0, i), fr->y[i] - at(fr->y0, i),
            fr->z[i] - at(fr->z0, i));
  vec_block(f, "Fint", fr->Fint, n_v);
  vec_block(f, "Fext", fr->Fext, n_v);
  vec_block(f, "Fdyn", fr->Fdyn, n_v);
  vec_block(f, "Fcnt", fr->Fcnt, n_v);
  vec_block(f, "u", fr->u, n_v);
  if (fr->manufactured && fr->muscle_activation) vec_block(f, "xdd", fr->xdd, n_v);
  if (fr->muscle_activation) {
    fprintf(f, "VECTORS Res float\n");
    for (int i = 0; i < n_v; i++) {
      double r[3];
      for (int d = 0; d < 3; d++) {
        const long k = 3L * i + d;
        r[d] = at(fr->Fint, k) - at(fr->Fext, k) + at(fr->Fdyn, k);
      }
      fprintf(f, "%f %f %f\n", r[0], r[1], r[2]);
    }
  }
  fprintf(f, "CELL_DATA %d\n", E);
  if (fr->lv_geom_process) vec_block(f, "nfib", fr->n_fib, E);
  fprintf(f, "SCALARS dA float\n");
  fprintf(f, "LOOKUP_TABLE default\n");
  for (int e = 0; e < E; e++) {
    double a;
    if (fr->dA) {
      a = fr->dA[e];
    } else {  /* AreaNormal (src/io.c:1646-1669) on the current coordinates */
      const int p = fr->nv1[e], q = fr->nv2[e], r = fr->nv3[e];
      const double ax = fr->x[q] - fr->x[p], ay = fr->y[q] - fr->y[p], az = fr->z[q] - fr->z[p];
      const double bx = fr->x[r] - fr->x[p], by = fr->y[r] - fr->y[p], bz = fr->z[r] - fr->z[p];
      const double cx = ay * bz - az * by, cy = az * bx - ax * bz, cz = ax * by - ay * bx;
      a = 0.5 * sqrt(cx * cx + cy * cy + cz * cz);
    }
    fprintf(f, "%f \n", a);
  }
  scalar_block(f, "dA0", fr->dA0, E);
  if (fr->muscle_activation) scalar_block(f, "gamma", fr->gamma, E);
  if (fclose(f) != 0) return BF_ERR_IO;
  return BF_OK;
}

/* ---- state checkpoints ---- */
static void put_be64(unsigned char *p, double v) {
  uint64_t u;
  memcpy(&u, &v, 8);
  for (int k = 7; k >= 0; k--) {
    p[k] = (unsigned char)(u & 0xffu);
    u >>= 8;
  }
}
static double get_be64(const unsigned char *p) {
  uint64_t u = 0;
  for (int k = 0; k < 8; k++) u = (u << 8) | p[k];
  double v;
  memcpy(&v, &u, 8);
  return v;
}
static void put_be32(unsigned char *p, uint32_t v) {
  for (int k = 3; k >= 0; k--) {
    p[k] = (unsigned char)(v & 0xffu);
    v >>= 8;
  }
}
static uint64_t get_be(const unsigned char *p, int n) {
  uint64_t u = 0;
  for (int k = 0; k < n; k++) u = (u << 8) | p[k];
  return u;
}

static const char *const kStateTag[4] = {"x", "xn", "xd", "xdd"};

static void state_path(char *out, size_t n, const char *dir, int k, int ibi, int ti) {
  snprintf(out, n, "%s/%s%1.1d_%5.5d.dat", dir, kStateTag[k], ibi, ti);  /* src/io.c:1332 */
}

int bf_io_write_state(const char *dir, int ibi, int ti, int n_total, const double *x,
                      const double *xn, const double *xd, const double *xdd, int layout) {
  if (n_total < 0 || (layout != BF_IO_STATE_RAW && layout != BF_IO_STATE_VECVIEW))
    return BF_ERR_ARG;
  const char *d = (dir && strlen(dir) > 0) ? dir : ".";
  mkdir(d, 0777);
  const double *v[4] = {x, xn, xd, xdd};
  const size_t n = 3 * (size_t)n_total;
  unsigned char *buf = (unsigned char *)malloc(8 * n + 8);
  if (!buf) return BF_ERR_NOMEM;
  int rc = BF_OK;
  for (int k = 0; k < 4 && rc == BF_OK; k++) {
    if (!v[k]) continue;
    size_t off = 0;
    if (layout == BF_IO_STATE_VECVIEW) {
      put_be32(buf, VEC_FILE_CLASSID);
      put_be32(buf + 4, (uint32_t)n);
      off = 8;
    }
    for (size_t i = 0; i < n; i++) put_be64(buf + off + 8 * i, v[k][i]);
    char fn[4096];
    state_path(fn, sizeof fn, d, k, ibi, ti);
    FILE *f = fopen(fn, "wb");
    if (!f) {
      rc = BF_ERR_IO;
      break;
    }
    if (fwrite(buf, 1, off + 8 * n, f) != off + 8 * n) rc = BF_ERR_IO;
    if (fclose(f) != 0) rc = BF_ERR_IO;
  }
  free(buf);
  return rc;
}

int bf_io_read_state(const char *dir, int ibi, int ti, int n_total, double *x, double *xn,
                     double *xd, double *xdd, int *layout) {
  if (n_total < 0) return BF_ERR_ARG;
  const char *d = (dir && strlen(dir) > 0) ? dir : ".";
  double *v[4] = {x, xn, xd, xdd};
  const size_t n = 3 * (size_t)n_total;
  unsigned char *buf = (unsigned char *)malloc(8 * n + 16);
  if (!buf) return BF_ERR_NOMEM;
  int rc = BF_OK, seen = -1;
  for (int k = 0; k < 4 && rc == BF_OK; k++) {
    if (!v[k]) continue;
    char fn[4096];
    state_path(fn, sizeof fn, d, k, ibi, ti);
    FILE *f = fopen(fn, "rb");
    if (!f) {
      rc = BF_ERR_IO;
      break;
    }
    const size_t got = fread(buf, 1, 8 * n + 16, f);
    const int longer = fgetc(f) != EOF;
    fclose(f);
    if (longer) {
      rc = BF_ERR_IO;
      break;
    }
    size_t off;
    int lay;
    if (got == 8 * n) {  /* the gathered, header-less layout */
      off = 0;
      lay = BF_IO_STATE_RAW;
    } else if (got == 8 * n + 8 && get_be(buf, 4) == VEC_FILE_CLASSID &&
               get_be(buf + 4, 4) == n) {  /* VecView, 32-bit PetscInt */
      off = 8;
      lay = BF_IO_STATE_VECVIEW;
    } else if (got == 8 * n + 16 && get_be(buf, 8) == VEC_FILE_CLASSID &&
               get_be(buf + 8, 8) == n) {  /* VecView, 64-bit PetscInt */
      off = 16;
      lay = BF_IO_STATE_VECVIEW;
    } else {
      rc = BF_ERR_IO;
      break;
    }
    for (size_t i = 0; i < n; i++) v[k][i] = get_be64(buf + off + 8 * i);
    if (seen < 0) seen = lay;
  }
  free(buf);
  if (rc == BF_OK && layout) *layout = seen;
  return rc;
}
