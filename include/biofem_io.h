/*
 * biofem_io.h -- the reference's on-disk formats either side of the force path.
 *
 * Plain C on the host (no device): the case inputs nlist/elist/blist that Dimension and
 * Input read (src/io.c:34-83, 333-537) and bin/biofem writes (bin/biofem:551-575), the VTK
 * surface frame of Output (src/io.c:540-930), and the x/xn/xd/xdd state checkpoints of
 * LocationOut / LocationIn (src/io.c:1272-1376, 1511-1640).  With these, an existing
 * BioFEM case directory loads into this library, and its states restart and render as the
 * reference's do.  All functions return BF_OK or a BF_ERR_* code (BF_ERR_IO: a file is
 * missing, cannot be written, or does not hold what its format requires).
 */
#ifndef BIOFEM_IO_H
#define BIOFEM_IO_H

#include "biofem_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Dimension (src/io.c:34-83): n_v from nlist<ibi>, n_elmt from elist<ibi>, n_edge and
 * n_ghosts from the first two lines of blist<ibi>; sum_n_bnodes (the sum of the n_edge
 * list lengths, which Input reads next) so the caller can size the boundary arrays.  Any
 * output may be NULL. */
int bf_io_dimension(const char *in_dir, int ibi, int *n_v, int *n_elmt, int *n_edge,
                    int *n_ghosts, int *sum_n_bnodes);

/* Input (src/io.c:333-537).  Coordinates divided by char_length[3] (NULL: 1, 1, 1) as the
 * reference does (x / char_length_x + 0); corners and boundary nodes converted from the
 * files' 1-based numbering to 0-based.  n_bnodes [n_edge], bnodes [sum_n_bnodes] (the
 * sizes bf_io_dimension returns).  nv4, nv5, nv6 [n_elmt] (any may be NULL): the vertex of
 * the neighbour across edge (nv2, nv3), (nv1, nv3) and (nv1, nv2), 1000000 where there is
 * none -- Input's O(E^2) search, here by an edge table, with its rule that the
 * highest-numbered neighbour wins on an edge shared by more than two triangles. */
int bf_io_input(const char *in_dir, int ibi, const double *char_length, double *x, double *y,
                double *z, int *nv1, int *nv2, int *nv3, int *n_bnodes, int *bnodes, int *nv4,
                int *nv5, int *nv6);

/* Writes nlist<ibi>, elist<ibi> and blist<ibi> in the layout bin/biofem writes and Input
 * reads (0-based arrays in, 1-based numbers in the files; coordinates with 17 significant
 * digits, where the reference's generator rounds to 6 decimals).  No reference counterpart
 * in C (the reference writes these files from its Python CLI). */
int bf_io_write_lists(const char *dir, int ibi, int n_v, const double *x, const double *y,
                      const double *z, int n_elmt, const int *nv1, const int *nv2,
                      const int *nv3, int n_edge, const int *n_bnodes, const int *bnodes,
                      int n_ghosts);

/* One VTK frame as Output writes it (src/io.c:540-930): file <out_dir>/surface<ibi:02>_
 * <ti:05>.vtk, legacy ASCII unstructured grid of the n_v real nodes and n_elmt triangles,
 * every number through printf's %f and %d in the reference's order.  NULL node or element
 * arrays are written as zeros (the reference writes its zeroed Vecs), except dA, which is
 * the current area of each triangle (AreaNormal) when NULL.  Blocks the reference writes
 * only under an option follow the same flags: xdd (manufactured && muscle_activation), Res
 * = Fint - Fext + Fdyn and gamma (muscle_activation), nfib (lv_geom_process).  The Fung
 * coefficient blocks of the nonlinear material are outside this path. */
typedef struct {
  int n_v, n_nodes, n_elmt;          /* n_nodes: length of the interleaved vectors / 3 */
  const double *x, *y, *z;           /* current coordinates [>= n_v] */
  const double *x0, *y0, *z0;        /* reference coordinates [>= n_v] (disp = x - x0) */
  const int *nv1, *nv2, *nv3;        /* 0-based corners [n_elmt] */
  const int *contact;                /* [n_v] */
  const double *m, *IE;              /* [n_v] */
  const double *Fint, *Fext, *Fdyn, *Fcnt, *u, *xdd; /* interleaved [3 n_nodes]; u = xd */
  const double *dA, *dA0;            /* [n_elmt] */
  const double *n_fib;               /* [3 n_elmt] */
  const double *gamma;               /* [n_elmt]: GammaOfTimeCurrent x gamma_scale */
  int muscle_activation, manufactured, lv_geom_process;
} bf_vtk_frame;
int bf_io_write_vtk(const char *out_dir, int ibi, int ti, const bf_vtk_frame *frame);

/* State checkpoints <dir>/{x,xn,xd,xdd}<ibi:1>_<ti:05>.dat, each 3 * n_total FP64 values
 * (n_total = n_v + n_ghosts, interleaved xyz), big-endian as PETSc binary files are.
 * BF_IO_STATE_RAW is LocationOut's process-count-invariant layout (PetscBinaryWrite of the
 * gathered buffer, no header, src/io.c:1330-1335); BF_IO_STATE_VECVIEW is its serial path
 * (VecView: int32 classid 1211214, int32 length, then the values).  The reader takes either
 * (and a VecView header with 64-bit integers), chosen by the file size; *layout (may be
 * NULL) reports which.  Any of x, xn, xd, xdd may be NULL to skip that file. */
#define BF_IO_STATE_RAW 0
#define BF_IO_STATE_VECVIEW 1
int bf_io_write_state(const char *dir, int ibi, int ti, int n_total, const double *x,
                      const double *xn, const double *xd, const double *xdd, int layout);
int bf_io_read_state(const char *dir, int ibi, int ti, int n_total, double *x, double *xn,
                     double *xd, double *xdd, int *layout);

#ifdef __cplusplus
}
#endif
#endif
