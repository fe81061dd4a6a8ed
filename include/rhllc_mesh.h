/*
 * rhllc_mesh.h -- host-side mesh and file-format tools for the RHLLC step.
 *
 * Plain C ABI (extern "C", pointers and sizes only).  Everything here runs on
 * the CPU; it needs no GPU and is built into libradrhd_mesh.so.
 *
 * Record formats are the reference's on-disk geometry records:
 *   geo_face_t  40 bytes  lib/grids/include/solvers_struct.h:87-102
 *     u32  id_l:25 | id_l_node:6 | is_regular:1
 *     i32  id_r:26 (signed; < 0 is a boundary code) | id_r_node:6
 *     f64  n[3]   outward normal of cell id_l      (offset 8)
 *     f64  S      face area                        (offset 32)
 *   geo_cell_t  56 bytes  lib/grids/include/solvers_struct.h:120-154
 *     i32  id_faces[4]                             (offset 0)
 *     f64  V                                       (offset 16)
 *     f64  center[3]                               (offset 24)
 *     u8   sign_n  bit j = 1 <=> cell is id_l of face j   (offset 48)
 *     u8   node:6 | is_bound:2                     (offset 49)
 *   flux_t      40 bytes  lib/grids/include/solvers_struct.h:51-70
 *     f64  d, v[3], p   (primitive W = rho, v, p  or conserved U = D, m, E)
 *
 * Face numbering and orientation follow ConvertBinToGeo
 * (lib/geometry/src/convert_bin_to_geo.cpp:3-48): half-faces (cell i, slot j)
 * are scanned in increasing i*4+j; the first visit of a face creates it with
 * id_l = i, the normal of (i, j) and sign bit j = 1; the partner slot gets
 * sign bit 0.  Boundary half-faces get id_r = boundary code and bump is_bound.
 */
#ifndef RHLLC_MESH_H
#define RHLLC_MESH_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RHLLC_GEO_FACE_BYTES 40
#define RHLLC_GEO_CELL_BYTES 56
#define RHLLC_FLUX_BYTES 40

/* Boundary codes, lib/global_types.h:25-31 (e_boundary_types_t). */
#define RHLLC_BOUND_FREE (-1)
#define RHLLC_BOUND_LOCK (-2)
#define RHLLC_BOUND_OUT_SOURCE (-3)
#define RHLLC_BOUND_INNER_SOURCE (-4)
#define RHLLC_BOUND_OUTER_SURFACE (-5)

/* Periodicity mask bits for rhllc_box_desc.periodic. */
#define RHLLC_PERIODIC_X 1
#define RHLLC_PERIODIC_Y 2
#define RHLLC_PERIODIC_Z 4

/* A box of nx*ny*nz hexahedra, each cut into 6 Kuhn tetrahedra. */
typedef struct rhllc_box_desc {
  int32_t nx, ny, nz;
  double origin[3];  /* lower corner */
  double length[3];  /* edge lengths of the box */
  int32_t periodic;  /* RHLLC_PERIODIC_* mask; non-periodic sides get boundary code -1 */
  double jitter;     /* vertex jitter, fraction of the hex edge (0.05 = 5%) */
  uint64_t seed;     /* jitter uses seed, the cell permutation seed + 1 */
  int32_t permute;   /* non-zero: randomly permute the cell numbering */
} rhllc_box_desc;

/* Number of cells and faces the box will have. */
int rhllc_box_sizes(const rhllc_box_desc *d, int64_t *n_cells, int64_t *n_faces);

/*
 * Build the box into caller-owned arrays:
 *   geo_cells : n_cells * 56 bytes (geo_cell_t records)
 *   geo_faces : n_faces * 40 bytes (geo_face_t records)
 * Cells are numbered after the optional permutation; faces follow the
 * ConvertBinToGeo rule above.  Returns 0 on success.
 */
int rhllc_box_build(const rhllc_box_desc *d, void *geo_cells, void *geo_faces);

/*
 * h_min as computed by utils/src/get_average_volume.cpp:31-46:
 * min over cells of (3/4) * V / (sum of the 4 face areas), the areas summed in
 * slot order.  Needs the face records to look up the areas.  The records are
 * validated first (ids in range, slots agreeing with their faces); a
 * malformed mesh returns -1 with the reason in rhllc_mesh_last_error().
 */
double rhllc_mesh_h_min(int64_t n_cells, const void *geo_cells, int64_t n_faces, const void *geo_faces);

/* Initial states of the synthetic workloads (flux_t W records, n_cells*40 B). */
#define RHLLC_INIT_SHOCK_TUBE 0  /* x<0.5: rho=10,p=13.33; else rho=1,p=1e-6; v=0 */
#define RHLLC_INIT_BLAST 1       /* rho=1; p=1000 for |x-c|<0.1, else 1; v=0     */
#define RHLLC_INIT_SHEAR 2       /* vx=0.995 tanh(y/0.01), rho 1/0.1, p=1, seeded vy */
#define RHLLC_INIT_UNIFORM 3     /* params[0..4] = rho, vx, vy, vz, p everywhere    */
int rhllc_init_state(int32_t kind, int64_t n_cells, const void *geo_cells,
                     const double *params, uint64_t seed, void *W_out);

/*
 * Seeded radiation fields for the full-form workload: per cell (FP32)
 * Ur log-uniform in [lo, hi], Fr = Ur * f * w (f ~ U[0,1], w isotropic),
 * kappa_a, kappa_s ~ U[kmin, kmax].
 */
int rhllc_init_radiation(int64_t n_cells, uint64_t seed, double ur_lo, double ur_hi,
                         double kmin, double kmax, float *Ur, float *Fr3, float *ka,
                         float *ks);

/*
 * Relabel-only locality order: a permutation new->old of the cells sorted by
 * the Hilbert index of their centres (21 bits per axis, one common scale),
 * and new->old of the faces sorted by the earlier new index of their two
 * cells, id_l for a boundary face (ties keep the old face order).  The mesh
 * itself (orientation, slots, id_l/id_r roles) is not changed.
 */
int rhllc_curve_order(int64_t n_cells, const void *geo_cells, int64_t n_faces,
                      const void *geo_faces, int32_t *cell_new_to_old,
                      int32_t *face_new_to_old);

/*
 * Legacy-format VTK files of tetrahedral unstructured grids (no VTK library).
 * Replaces files_sys::vtk::Read (lib/files_sys/include/reader_vtk.h:43-61),
 * which reads the file with VTK 9.3's vtkGenericDataObjectReader.
 * Versions 2.0-5.1, ASCII or BINARY; every cell must be a VTK_TETRA (10).
 * Points declared "float" are rounded to float, as VTK stores them.
 *
 *   rhllc_vtk_open   parses the file; *n_points, *n_cells receive the sizes
 *   rhllc_vtk_copy   points: 3*n_points doubles; tets: 4*n_cells point ids
 *                    (either may be NULL)
 *   rhllc_vtk_close  frees the parsed mesh
 */
typedef struct rhllc_vtk_mesh rhllc_vtk_mesh;
int rhllc_vtk_open(const char *path, rhllc_vtk_mesh **out, int64_t *n_points, int64_t *n_cells);
int rhllc_vtk_copy(const rhllc_vtk_mesh *m, double *points, int32_t *tets);
void rhllc_vtk_close(rhllc_vtk_mesh *m);

/* Boundary-code rules of SetTypeOfBound (lib/geometry/src/make_internal_format.cpp:12-108). */
#define RHLLC_GEO_BOUND_CONE 0 /* GEOMETRY_TYPE Cone, the reference's build (include/solvers_config.h:29):
                                  normal within 1e-7 of -x: -3; of +x: -5; otherwise -2 */
#define RHLLC_GEO_BOUND_FREE 1 /* GEOMETRY_TYPE Cube: every boundary face -1 */

/*
 * The per-cell arrays that BuildDataFromVTK writes
 * (lib/geometry/src/make_internal_format.cpp:272-336), from a tet mesh:
 *   neighbors[4N]  neighbour cell*4 + its local face, or a boundary code
 *                  (GetNeighborFace3D + SetTypeOfBound; file neighbors.bin)
 *   normals[12N]   outward unit normal of local face j at [12i + 3j]
 *                  (NormalAndSquareFace3D; file normals.bin)
 *   areas[4N]      face areas (areas.bin)
 *   volumes[N]     -(a . (b x c)) / 6, signed as in GetVolume3D (volume.bin)
 *   centers[3N]    insphere centres, CenterOfTetra3D (centers.bin)
 * Local face j of cell i is VTK's vtkTetra face j: points (0,1,3), (1,2,3),
 * (2,0,3), (0,2,1) of the cell.  Results are bitwise the reference's.
 * Fails on repeated or out-of-range point ids and on a face shared by more
 * than two cells.
 */
int rhllc_tet_geometry(int64_t n_points, const double *points, int64_t n_cells, const int32_t *tets,
                       int32_t bound_rule, int32_t *neighbors, double *normals, double *areas,
                       double *volumes, double *centers);

/*
 * ConvertBinToGeo (lib/geometry/src/convert_bin_to_geo.cpp:3-48): geometry
 * records from the arrays above, numbered by the rule at the top of this
 * header.  *n_faces receives the face count; with geo_cells and geo_faces both
 * NULL only the count is computed.  geo_cells: n_cells * 56 bytes, geo_faces:
 * *n_faces * 40 bytes.
 */
int rhllc_bin_to_geo(int64_t n_cells, const int32_t *neighbors, const double *normals, const double *areas,
                     const double *volumes, const double *centers, int64_t *n_faces, void *geo_cells,
                     void *geo_faces);

/*
 * The cooling-function table (tab_func/rcf_H.dat), as
 * files_sys::txt::ReadTableFunc (lib/files_sys/src/reader_txt.cpp:59-118)
 * reads it: "... = Nd" and "... = Nt" lines, two header lines, then Nd*Nt
 * triples (ln D, ln T, ln Lambda).  *nx = Nd, *ny = Nt; header = {min_x,
 * max_x, step_x, min_y, max_y, step_y} over the ln D and ln T axes, step =
 * second value - first; data (Nd*Nt doubles at [i*Nt + j]) is filled when not
 * NULL, so a first call with data = NULL gives the sizes.  Values are those
 * of the reference's reader bit for bit.  Fails on a missing file, a bad
 * count line (Nd or Nt below 2) or fewer than Nd*Nt triples.
 */
int rhllc_table_read(const char *path, int32_t *nx, int32_t *ny, double header[6], double *data);

/* Read/write "int32 n, then n records" files (reader_bin.h:113-135,232-252). */
int64_t rhllc_file_count(const char *path);
int rhllc_file_read_records(const char *path, size_t rec_bytes, int64_t n, void *out);
int rhllc_file_write_records(const char *path, size_t rec_bytes, int64_t n, const void *recs);

const char *rhllc_mesh_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* RHLLC_MESH_H */
